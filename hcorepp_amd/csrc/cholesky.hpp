// Job descriptors and launchers of the TLR Cholesky tile kernels (cholesky.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hcamd {

struct CholTrsm {   // V (rank x n, ld rank) <- V L^{-T}
    double *v;
    const double *l;
    int64_t rank, n, ldl;
};
struct CholSyrk {   // A (n x n, ld lda) <- A - U (V V^T) U^T ; scratch w (rank^2), t (n rank)
    double *a;
    const double *u, *v;
    double *w, *t;
    int64_t n, rank, lda;
};
struct CholTrans {  // dst = src^T for compressed tiles
    const double *src;
    double *dst;
    int64_t m, n, rank, src_max_rank, dst_max_rank;
};

// Lower Cholesky factor in place (blocked); the strict upper part is zeroed.  info: first
// non-positive pivot (1-based) or 0.
template <typename T>
hipError_t launch_potrf(hipStream_t s, T *A, int n, int64_t lda, int *info);
hipError_t launch_trsm_v(hipStream_t s, const CholTrsm *jobs, int n);
hipError_t launch_syrk(hipStream_t s, const CholSyrk *jobs, int n);
hipError_t launch_transpose_factors(hipStream_t s, const CholTrans *jobs, int n);

}  // namespace hcamd
