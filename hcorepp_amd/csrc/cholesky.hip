// Tile kernels of the TLR Cholesky factorisation A = L L^T (lower), in the engine's tile
// convention (compressed tile = U (m x r) V (r x n), stored [U | V], ld m and r):
//   potrf  dense diagonal tile A_kk = L_kk L_kk^T (in place, lower; the strict upper part is zeroed)
//   trsm   L_ik = A_ik L_kk^{-T}: only the V factor of a compressed A_ik changes, V <- V L_kk^{-T}
//   syrk   A_ii <- A_ii - L_ik L_ik^T = A_ii - U (V V^T) U^T for a compressed L_ik (dense A_ii)
//   transpose  compressed B = A^T: U_B = V_A^T, V_B = U_A^T (the B operand of the trailing
//              updates A_ij <- A_ij - L_ik L_jk^T, which then run as ordinary CCC recompressions)
// Each kernel handles a batch of independent tiles, one workgroup per tile.
#include "kernels.hpp"
#include "cholesky.hpp"

namespace hcamd {

// Right-looking unblocked Cholesky of one n x n tile in global memory (the tile of an nb = 512
// TLR matrix is L2-resident).  info[0] receives the first non-positive pivot (1-based), 0 if none.
__global__ void HCAMD_KERNEL k_potrf(double *A, int n, int64_t lda, int *info) {
    __shared__ double s_piv;
    __shared__ int s_bad;
    if (threadIdx.x == 0) s_bad = 0;
    __syncthreads();
    for (int j = 0; j < n; ++j) {
        if (threadIdx.x == 0) {
            const double d = A[j + (int64_t) j * lda];
            if (!(d > 0.0)) {
                if (!s_bad) s_bad = j + 1;
                s_piv = 1.0;
            } else {
                s_piv = sqrt(d);
            }
            A[j + (int64_t) j * lda] = s_piv;
        }
        __syncthreads();
        const double piv = s_piv;
        for (int i = j + 1 + threadIdx.x; i < n; i += NT) A[i + (int64_t) j * lda] /= piv;
        __syncthreads();
        // trailing update of the lower triangle: A[i, c] -= A[i, j] A[c, j] for j < c <= i
        const int m = n - j - 1;
        for (int64_t e = threadIdx.x; e < (int64_t) m * m; e += NT) {
            const int i = j + 1 + (int) (e % m), c = j + 1 + (int) (e / m);
            if (i >= c) A[i + (int64_t) c * lda] -= A[i + (int64_t) j * lda] * A[c + (int64_t) j * lda];
        }
        __syncthreads();
    }
    for (int64_t e = threadIdx.x; e < (int64_t) n * n; e += NT) {
        const int i = (int) (e % n), c = (int) (e / n);
        if (i < c) A[i + (int64_t) c * lda] = 0.0;
    }
    if (threadIdx.x == 0) *info = s_bad;
}

// V (r x n, ld r) <- V L^{-T} for lower-triangular L (n x n, ld ldl): forward substitution
// X L^T = V row by row; lane a owns row a of V.
__global__ void HCAMD_KERNEL k_trsm_v(const CholTrsm *jobs) {
    const CholTrsm jb = jobs[blockIdx.x];
    const int r = (int) jb.rank, n = (int) jb.n;
    double *V = jb.v;
    const double *L = jb.l;
    for (int a = threadIdx.x; a < r; a += NT) {
        for (int j = 0; j < n; ++j) {
            double acc = V[a + (int64_t) j * r];
            for (int k = 0; k < j; ++k) acc -= V[a + (int64_t) k * r] * L[j + (int64_t) k * jb.ldl];
            V[a + (int64_t) j * r] = acc / L[j + (int64_t) j * jb.ldl];
        }
    }
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

hipError_t launch_potrf(hipStream_t s, double *A, int n, int64_t lda, int *info) {
    hipLaunchKernelGGL(k_potrf, dim3(1), dim3(NT), 0, s, A, n, lda, info);
    return hipGetLastError();
}
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
