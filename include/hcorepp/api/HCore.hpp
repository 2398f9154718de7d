// Drop-in for hcorepp::api::HCore<T> (reference include/hcorepp/api/HCore.hpp:23-128).
// Gemm, Syrk, Potrf and Trsm keep the reference signatures and forward to the C-ABI
// (hcorepp_gemm_d/_s, hcorepp_tile_syrk_d/_s, hcorepp_tile_potrf_d/_s, hcorepp_tile_trsm_d/_s).
// In the Cholesky flow (Gemm with aCholesky = true, Syrk, Trsm) compressed tiles follow the
// reference's convention for that flow: A = U V^T with V stored cols x rank.
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

    // Reference include/hcorepp/api/HCore.hpp:66-77 (src/api/HCore.cpp:484-584).  A compressed A
    // uses alpha = -1 as the reference does.
    static void Syrk(T aAlpha, const operators::Tile<T> &aA, const blas::Op &aAOp, blas::Uplo aUplo, T aBeta,
                     operators::Tile<T> &aC, const kernels::RunContext &aContext, size_t &aFlops,
                     dataunits::MemoryUnit<T> &aMemoryHandler) {
        (void) aMemoryHandler;
        hcorepp_status st;
        if (std::is_same<T, double>::value)
            st = hcorepp_tile_syrk_d(aContext.Handle(), (double) aAlpha, &aA.Desc(), (char) aAOp, (char) aUplo,
                                     (double) aBeta, &aC.Desc());
        else
            st = hcorepp_tile_syrk_s(aContext.Handle(), (float) aAlpha, &aA.Desc(), (char) aAOp, (char) aUplo,
                                     (float) aBeta, &aC.Desc());
        kernels::ThrowIfError(st);
        const size_t n = (size_t) aC.GetNumOfRows();
        aFlops += aA.isCompressed() ? 0 : n * (n + 1) * (size_t) aC.GetNumOfCols();
    }

    // Reference include/hcorepp/api/HCore.hpp:78-90 (src/api/HCore.cpp:586-622).  A tile that is
    // not positive definite throws (the reference leaves LAPACK's info unread).
    static void Potrf(operators::Tile<T> &aA, blas::Uplo aUplo, const kernels::RunContext &aContext, size_t &aFlops,
                      dataunits::MemoryUnit<T> &aMemoryHandler) {
        (void) aMemoryHandler;
        hcorepp_status st = std::is_same<T, double>::value
                                    ? hcorepp_tile_potrf_d(aContext.Handle(), (char) aUplo, &aA.Desc())
                                    : hcorepp_tile_potrf_s(aContext.Handle(), (char) aUplo, &aA.Desc());
        kernels::ThrowIfError(st);
        const size_t n = (size_t) aA.GetNumOfRows();
        aFlops += n * n * n / 3;
    }

    // Reference include/hcorepp/api/HCore.hpp:92-106 (src/api/HCore.cpp:624-647): trsm on the V
    // factor of the compressed B.
    static void Trsm(blas::Side aSide, blas::Uplo aUplo, blas::Op aTrans, blas::Diag aDiag, T aAlpha,
                     operators::Tile<T> &aA, operators::Tile<T> &aB, const kernels::RunContext &aContext,
                     size_t &aFlops, dataunits::MemoryUnit<T> &aMemoryHandler) {
        (void) aMemoryHandler;
        hcorepp_status st;
        if (std::is_same<T, double>::value)
            st = hcorepp_tile_trsm_d(aContext.Handle(), (char) aSide, (char) aUplo, (char) aTrans, (char) aDiag,
                                     (double) aAlpha, &aA.Desc(), &aB.Desc());
        else
            st = hcorepp_tile_trsm_s(aContext.Handle(), (char) aSide, (char) aUplo, (char) aTrans, (char) aDiag,
                                     (float) aAlpha, &aA.Desc(), &aB.Desc());
        kernels::ThrowIfError(st);
        const size_t order = (size_t) aA.GetNumOfRows();
        aFlops += order * order * (size_t) aB.Desc().rank;
    }

    // Workspace (elements) one update needs; reference HCore<T>::CalculateMemoryPoolSize
    // (src/api/HCore.cpp:417-480), with this engine's layout.
    static size_t CalculateMemoryPoolSize(const operators::Tile<T> &aA, const operators::Tile<T> &aB,
                                          const operators::Tile<T> &aC,
                                          const operators::CompressionParameters &aParams,
                                          const kernels::RunContext &) {
        const hcorepp_cparams p = aParams.ToC();
        return std::is_same<T, double>::value ? hcorepp_gemm_pool_size_d(&aA.Desc(), &aB.Desc(), &aC.Desc(), &p)
                                              : hcorepp_gemm_pool_size_s(&aA.Desc(), &aB.Desc(), &aC.Desc(), &p);
    }
};

}  // namespace api
}  // namespace hcorepp
