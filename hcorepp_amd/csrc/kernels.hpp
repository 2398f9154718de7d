// Declarations shared between the kernels (kernels.hip) and the host side of
// the C-ABI (api.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"
#include "gemm.hpp"
#include "recompress.hpp"
#include "learned.hpp"

namespace hcamd {

// Every engine kernel runs 256 threads and is register-capped for two waves per SIMD
// (256 VGPRs per lane), so that two workgroups of the batched kernel share a CU: the
// latency-bound Jacobi rounds and panel reductions of one overlap the other's.  The
// device functions the kernels share inherit the cap only when every caller has it.
#define HCAMD_KERNEL __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2)))

// One temporary factor product W = op(A) * op(B) (alpha 1, beta 0), as the
// reference's temporary DenseTile GEMMs (src/api/HCore.cpp:176-264).
template <typename T>
struct GemmSpec {
    const T *a;
    const T *b;
    T *c;
    int64_t lda, ldb, ldc;
    int32_t m, n, k;
    int32_t ta, tb;
    int32_t pad;
};

// One update of a batch: up to two temporary products then the recompression.  The
// temporaries and the recompression workspace live in the slot of the workgroup that runs the
// update: the pointers named by the bits of `rel` hold byte offsets into that slot.
enum : int32_t {
    REL_G0A = 1, REL_G0B = 2, REL_G0C = 4, REL_G1A = 8, REL_G1B = 16, REL_G1C = 32,
    REL_AFIN = 64, REL_BFIN = 128, REL_WS = 256,
};
template <typename T>
struct BatchEntry {
    GemmSpec<T> gemm[2];
    int32_t ngemm;
    int32_t rel;
    RecompDesc<T> rd;
};

struct BatchCounters {
    unsigned long long next;
    unsigned long long status;
    unsigned long long max_sweeps;
    unsigned long long fallbacks;  // updates whose learned core sizing was rejected
};

int pick_pb(int64_t rows, int64_t lds_elems);
bool pb_fits(int64_t rows, int pb, int64_t lds_elems);

template <typename T>
hipError_t launch_gemm(hipStream_t s, int M, int N, int Kd, T alpha, const T *A, int64_t lda, bool ta, const T *B,
                       int64_t ldb, bool tb, T beta, T *C, int64_t ldc);
// Cores of at least MW_MIN_COLS columns run their Jacobi over many workgroups when `jflags`
// (JACOBI_MAX_SWEEPS device ints) and `host_flag` (one pinned int) are given; that path
// synchronises the stream between sweep chunks.  QRs of at least MW_MIN_QR_COLS columns run
// panel by panel with multi-workgroup trailing updates when `vpan` (max(m, n) * 16 + 1024
// elements of device scratch: the panel, then column norms) is given.
constexpr int MW_MIN_COLS = 96;
constexpr int MW_MIN_QR_COLS = 128;
template <typename T>
hipError_t launch_recomp_pre(hipStream_t s, const RecompDesc<T> &d, int64_t K, int64_t lds_bytes, int pb,
                             int32_t *jflags = nullptr, int32_t *host_flag = nullptr, T *vpan = nullptr);
template <typename T>
hipError_t launch_recomp_post(hipStream_t s, const RecompDesc<T> &d, int64_t r, int64_t lds_bytes, int pb);
hipError_t launch_set_i64(hipStream_t s, int64_t *p, int64_t v);
template <typename T>
hipError_t launch_readjust_ddc(hipStream_t s, const T *W, int64_t m, int64_t n, T *cdata);
template <typename T>
hipError_t launch_fill_identity(hipStream_t s, T *I, int64_t n);
hipError_t launch_reset_counters(hipStream_t s, BatchCounters *c);
// Batched persistent launch: workgroup g works in slot g of `slots` (slot_bytes each) and copies
// each update's workspace header to hdrs[HDR_WORDS * update] when the update ends.
template <typename T>
hipError_t launch_batched(hipStream_t s, const BatchEntry<T> *entries, const int32_t *order, int64_t n,
                          BatchCounters *ctr, int grid, int64_t lds_bytes, int pb, int mixed, char *slots,
                          int64_t slot_bytes, int64_t *hdrs);
template <typename T>
hipError_t set_kernel_lds_limits(int64_t lds_bytes);

}  // namespace hcamd
