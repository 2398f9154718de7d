// Drop-in for hcorepp::api::HCore<T> (reference include/hcorepp/api/HCore.hpp:23-128).
// Gemm keeps the reference signature and forwards to hcorepp_gemm_d / hcorepp_gemm_s.
#pragma once

#include <type_traits>

#include <hcorepp/data-units/memory-handlers/MemoryHandler.hpp>
#include <hcorepp/kernels/RunContext.hpp>
#include <hcorepp/operators/concrete/Compressed.hpp>
#include <hcorepp/operators/concrete/Dense.hpp>
#include <hcorepp/operators/helpers/CompressionParameters.hpp>

namespace hcorepp {
namespace api {

template <typename T>
class HCore {
    static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value, "float or double");

public:
    static void Gemm(T aAlpha, const operators::Tile<T> &aA, const blas::Op &aAOp, const operators::Tile<T> &aB,
                     const blas::Op &aBOp, T aBeta, operators::Tile<T> &aC, const kernels::RunContext &aContext,
                     size_t &aFlops, dataunits::MemoryUnit<T> &aMemoryUnit,
                     const operators::CompressionParameters &aCompressionParameters = {1e-9},
                     bool aCholesky = false) {
        (void) aMemoryUnit;
        const hcorepp_cparams p = aCompressionParameters.ToC();
        uint64_t flops = 0;
        hcorepp_status st;
        if (std::is_same<T, double>::value)
            st = hcorepp_gemm_d(aContext.Handle(), (double) aAlpha, &aA.Desc(), (char) aAOp, &aB.Desc(), (char) aBOp,
                                (double) aBeta, &aC.Desc(), &p, aCholesky, &flops, nullptr, nullptr);
        else
            st = hcorepp_gemm_s(aContext.Handle(), (float) aAlpha, &aA.Desc(), (char) aAOp, &aB.Desc(), (char) aBOp,
                                (float) aBeta, &aC.Desc(), &p, aCholesky, &flops, nullptr, nullptr);
        kernels::ThrowIfError(st);
        aFlops += flops;
    }

    // Workspace (elements) one update needs; reference HCore<T>::CalculateMemoryPoolSize
    // (src/api/HCore.cpp:417-480), with this engine's layout.
    static size_t CalculateMemoryPoolSize(const operators::Tile<T> &aA, const operators::Tile<T> &aB,
                                          const operators::Tile<T> &aC,
                                          const operators::CompressionParameters &aParams,
                                          const kernels::RunContext &) {
        const hcorepp_cparams p = aParams.ToC();
        return hcorepp_gemm_pool_size_d(&aA.Desc(), &aB.Desc(), &aC.Desc(), &p);
    }
};

}  // namespace api
}  // namespace hcorepp
