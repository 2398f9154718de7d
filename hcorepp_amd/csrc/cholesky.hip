// Tile kernels of the TLR Cholesky factorisation A = L L^T (lower), in the engine's tile
// convention (compressed tile = U (m x r) V (r x n), stored [U | V], ld m and r):
//   potrf  dense diagonal tile A_kk = L_kk L_kk^T (in place, lower; the strict upper part is zeroed)
//   trsm   L_ik = A_ik L_kk^{-T}: only the V factor of a compressed A_ik changes, V <- V L_kk^{-T}
//   syrk   A_ii <- A_ii - L_ik L_ik^T = A_ii - U (V V^T) U^T for a compressed L_ik (dense A_ii)
//   transpose  compressed B = A^T: U_B = V_A^T, V_B = U_A^T (the B operand of the trailing
//              updates A_ij <- A_ij - L_ik L_jk^T, which then run as ordinary CCC recompressions)
// The trsm, syrk and transpose kernels handle a batch of independent tiles, one workgroup per
// tile; potrf is blocked over several launches.
#include "kernels.hpp"
#include "cholesky.hpp"

namespace hcamd {

constexpr int CB = 32;   // column block of the blocked potrf
constexpr int RB = 16;   // column block of the row-wise triangular solves
constexpr int ROWS_TRSM_LDS = RB * (RB + 1) + RB * NT;

// X <- X L^{-T} for X (rows x n, column k at X + k * ldx) and lower-triangular L (n x n, ld ldl),
// i.e. the solve X L^T = B by rows.  A thread owns a row: blocks of RB columns are accumulated in
// registers from the earlier (already solved) columns, with L staged through LDS in RB x RB
// pieces; the diagonal block is then solved by forward substitution in LDS (the thread's RB
// values in column-interleaved slots, so that loops stay rolled).  `lds` holds ROWS_TRSM_LDS
// doubles.  Every thread of the workgroup must call it (it synchronises).
template <typename T>
__device__ __forceinline__ void rows_trsm_lt(T *X, int rows, int n, int64_t ldx, const T *L, int64_t ldl, T *lds) {
    T *lt = lds;                   // RB x (RB + 1)
    T *xs = lds + RB * (RB + 1);   // RB x NT
    const int tid = (int) threadIdx.x;
    for (int rbase = 0; rbase < rows; rbase += NT) {
        const int a = rbase + tid;
        const bool act = a < rows;
        for (int j0 = 0; j0 < n; j0 += RB) {
            const int jb = (n - j0) < RB ? (n - j0) : RB;
            T acc[RB];
#pragma unroll
            for (int c = 0; c < RB; ++c) acc[c] = (act && c < jb) ? X[a + (int64_t) (j0 + c) * ldx] : T(0);
            for (int k0 = 0; k0 < j0; k0 += RB) {
                __syncthreads();
                for (int e = tid; e < RB * RB; e += NT) {
                    const int c = e % RB, kk = e / RB;
                    lt[c * (RB + 1) + kk] = c < jb ? L[(j0 + c) + (int64_t) (k0 + kk) * ldl] : T(0);
                }
                __syncthreads();
                if (act) {
                    for (int kk = 0; kk < RB; ++kk) {
                        const T x = X[a + (int64_t) (k0 + kk) * ldx];
#pragma unroll
                        for (int c = 0; c < RB; ++c) acc[c] -= x * lt[c * (RB + 1) + kk];
                    }
                }
            }
            __syncthreads();
            for (int e = tid; e < RB * RB; e += NT) {
                const int c = e % RB, cc = e / RB;
                lt[c * (RB + 1) + cc] = (c < jb && cc < jb) ? L[(j0 + c) + (int64_t) (j0 + cc) * ldl] : T(1);
            }
#pragma unroll
            for (int c = 0; c < RB; ++c) xs[c * NT + tid] = acc[c];
            __syncthreads();
            if (act) {
                for (int c = 0; c < jb; ++c) {
                    T x = xs[c * NT + tid];
                    for (int cc = 0; cc < c; ++cc) x -= xs[cc * NT + tid] * lt[c * (RB + 1) + cc];
                    x /= lt[c * (RB + 1) + c];
                    xs[c * NT + tid] = x;
                    X[a + (int64_t) (j0 + c) * ldx] = x;
                }
            }
        }
    }
}

// Diagonal block of the blocked potrf: unblocked Cholesky of A[j0:j0+jb, j0:j0+jb] in LDS.
// info[0] receives the first non-positive pivot (1-based, whole-tile numbering) unless already set;
// such a pivot is replaced by 1 so that the factorisation runs to the end.
template <typename T>
__global__ void HCAMD_KERNEL k_potrf_diag(T *A, int64_t lda, int j0, int jb, int *info) {
    __shared__ T d[CB * (CB + 1)];
    for (int e = threadIdx.x; e < jb * jb; e += NT) {
        const int i = e % jb, c = e / jb;
        d[i + c * (CB + 1)] = A[(j0 + i) + (int64_t) (j0 + c) * lda];
    }
    __syncthreads();
    for (int j = 0; j < jb; ++j) {
        if (threadIdx.x == 0) {
            const T v = d[j + j * (CB + 1)];
            T p = T(1);
            if (v > T(0)) p = sqrt(v);
            else if (info[0] == 0) info[0] = j0 + j + 1;
            d[j + j * (CB + 1)] = p;
        }
        __syncthreads();
        const T p = d[j + j * (CB + 1)];
        for (int i = j + 1 + threadIdx.x; i < jb; i += NT) d[i + j * (CB + 1)] /= p;
        __syncthreads();
        for (int e = threadIdx.x; e < jb * jb; e += NT) {
            const int i = e % jb, c = e / jb;
            if (c > j && i >= c) d[i + c * (CB + 1)] -= d[i + j * (CB + 1)] * d[c + j * (CB + 1)];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < jb * jb; e += NT) {
        const int i = e % jb, c = e / jb;
        if (i >= c) A[(j0 + i) + (int64_t) (j0 + c) * lda] = d[i + c * (CB + 1)];
    }
}

// Sub-diagonal panel of the blocked potrf: A[j0+jb:n, j0:j0+jb] <- (that) L_jj^{-T}, rows dealt to
// workgroups NT at a time.
template <typename T>
__global__ void HCAMD_KERNEL k_potrf_panel(T *A, int n, int64_t lda, int j0, int jb) {
    __shared__ T lds[ROWS_TRSM_LDS];
    const int r0 = j0 + jb + (int) blockIdx.x * NT;
    const int rows = (n - r0) < NT ? (n - r0) : NT;
    if (rows <= 0) return;
    rows_trsm_lt(A + r0 + (int64_t) j0 * lda, rows, jb, lda, A + j0 + (int64_t) j0 * lda, lda, lds);
}

template <typename T>
__global__ void HCAMD_KERNEL k_zero_upper(T *A, int n, int64_t lda) {
    for (int64_t e = (int64_t) blockIdx.x * NT + threadIdx.x; e < (int64_t) n * n; e += (int64_t) gridDim.x * NT) {
        const int i = (int) (e % n), c = (int) (e / n);
        if (i < c) A[i + (int64_t) c * lda] = T(0);
    }
}

__global__ void k_zero_int(int *p) {
    if (threadIdx.x == 0) *p = 0;
}

// V (r x n, ld r) <- V L^{-T} for lower-triangular L (n x n, ld ldl): one workgroup per tile.
__global__ void HCAMD_KERNEL k_trsm_v(const CholTrsm *jobs) {
    __shared__ double lds[ROWS_TRSM_LDS];
    const CholTrsm jb = jobs[blockIdx.x];
    rows_trsm_lt<double>(jb.v, (int) jb.rank, (int) jb.n, jb.rank, jb.l, jb.ldl, lds);
}

// A (n x n dense) <- A - U (V V^T) U^T, with scratch W (r x r) and T (n x r); the strict upper
// triangle of A is then set to 0, as the reference's Syrk does for a lower update.
__global__ void HCAMD_KERNEL k_syrk(const CholSyrk *jobs) {
    __shared__ double lds[GT_LDS_ELEMS];
    const CholSyrk jb = jobs[blockIdx.x];
    const int n = (int) jb.n, r = (int) jb.rank;
    if (r > 0) {
        wg_gemm<double>(r, r, n, 1.0, jb.v, r, false, jb.v, r, true, 0.0, jb.w, r, lds);
        __syncthreads();
        wg_gemm<double>(n, r, r, 1.0, jb.u, n, false, jb.w, r, false, 0.0, jb.t, n, lds);
        __syncthreads();
        wg_gemm<double>(n, n, r, -1.0, jb.t, n, false, jb.u, n, true, 1.0, jb.a, jb.lda, lds);
        __syncthreads();
    }
    for (int64_t e = threadIdx.x; e < (int64_t) n * n; e += NT) {
        const int i = (int) (e % n), c = (int) (e / n);
        if (i < c) jb.a[i + (int64_t) c * jb.lda] = 0.0;
    }
}

// dst (compressed, [U n x r | V r x m], max_rank cap) = src^T for src [U m x r | V r x n].
__global__ void HCAMD_KERNEL k_transpose_factors(const CholTrans *jobs) {
    const CholTrans jb = jobs[blockIdx.x];
    const int64_t m = jb.m, n = jb.n, r = jb.rank;
    const double *su = jb.src, *sv = jb.src + m * jb.src_max_rank;
    double *du = jb.dst, *dv = jb.dst + n * jb.dst_max_rank;
    for (int64_t e = threadIdx.x; e < n * r; e += NT) {  // du[j, a] = sv[a, j]
        const int64_t j = e % n, a = e / n;
        du[e] = sv[a + j * r];
    }
    for (int64_t e = threadIdx.x; e < r * m; e += NT) {  // dv[a, i] = su[i, a]
        const int64_t a = e % r, i = e / r;
        dv[e] = su[i + a * m];
    }
}

// Blocked right-looking potrf: per block column, the diagonal block (one workgroup), the panel
// below it (row solves over the GPU), and the trailing update A22 -= L21 L21^T as a dense GEMM
// (both triangles; the strict upper part is cleared at the end).
template <typename T>
hipError_t launch_potrf(hipStream_t s, T *A, int n, int64_t lda, int *info) {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, info);
    for (int j0 = 0; j0 < n; j0 += CB) {
        const int jb = (n - j0) < CB ? (n - j0) : CB;
        hipLaunchKernelGGL(k_potrf_diag<T>, dim3(1), dim3(NT), 0, s, A, lda, j0, jb, info);
        const int rest = n - j0 - jb;
        if (rest > 0) {
            hipLaunchKernelGGL(k_potrf_panel<T>, dim3((rest + NT - 1) / NT), dim3(NT), 0, s, A, n, lda, j0, jb);
            T *L21 = A + (j0 + jb) + (int64_t) j0 * lda;
            hipError_t e = launch_gemm<T>(s, rest, rest, jb, T(-1), L21, lda, false, L21, lda, true, T(1),
                                          A + (j0 + jb) + (int64_t) (j0 + jb) * lda, lda);
            if (e != hipSuccess) return e;
        }
    }
    int64_t work = (int64_t) n * n;
    int grid = (int) ((work + NT - 1) / NT);
    if (grid > 1024) grid = 1024;
    hipLaunchKernelGGL(k_zero_upper<T>, dim3(grid > 0 ? grid : 1), dim3(NT), 0, s, A, n, lda);
    return hipGetLastError();
}
template hipError_t launch_potrf<double>(hipStream_t, double *, int, int64_t, int *);
template hipError_t launch_potrf<float>(hipStream_t, float *, int, int64_t, int *);
hipError_t launch_trsm_v(hipStream_t s, const CholTrsm *jobs, int n) {
    if (n > 0) hipLaunchKernelGGL(k_trsm_v, dim3(n), dim3(NT), 0, s, jobs);
    return hipGetLastError();
}
hipError_t launch_syrk(hipStream_t s, const CholSyrk *jobs, int n) {
    if (n > 0) hipLaunchKernelGGL(k_syrk, dim3(n), dim3(NT), 0, s, jobs);
    return hipGetLastError();
}
hipError_t launch_transpose_factors(hipStream_t s, const CholTrans *jobs, int n) {
    if (n > 0) hipLaunchKernelGGL(k_transpose_factors, dim3(n), dim3(NT), 0, s, jobs);
    return hipGetLastError();
}

}  // namespace hcamd
