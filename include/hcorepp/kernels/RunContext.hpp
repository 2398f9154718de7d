// Drop-in for hcorepp::kernels::RunContext (reference src/kernels/cuda/RunContext.cpp:11-78):
// one HIP stream plus the device arena the updates take their workspace from.
#pragma once

#include <stdexcept>
#include <string>

#include <hcorepp_amd.h>

namespace hcorepp {
namespace kernels {

inline void ThrowIfError(hcorepp_status st) {
    if (st != HCOREPP_OK) throw std::runtime_error(std::string("hcorepp_amd: ") + hcorepp_last_error());
}

class RunContext {
public:
    explicit RunContext(int aDevice = 0, size_t aArenaBytes = 0) {
        ThrowIfError(hcorepp_ctx_create(aDevice, aArenaBytes, &mCtx));
    }
    RunContext(const RunContext &) = delete;
    RunContext &operator=(const RunContext &) = delete;
    ~RunContext() { hcorepp_ctx_destroy(mCtx); }
    void Sync() const { ThrowIfError(hcorepp_ctx_sync(mCtx)); }
    hcorepp_ctx *Handle() const { return mCtx; }
    void *GetStream() const { return hcorepp_ctx_stream(mCtx); }

private:
    hcorepp_ctx *mCtx = nullptr;
};

}  // namespace kernels
}  // namespace hcorepp
