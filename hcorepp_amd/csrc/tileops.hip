// Tile operations of the reference's Cholesky-flow API with its argument meaning
// (HCore<T>::Potrf / Trsm / Syrk, reference src/api/HCore.cpp:482-647; the BLAS/LAPACK calls
// of src/kernels/omp/kernels.cpp:236-282):
//   potrf  LAPACK potrf of a dense tile for uplo 'L' or 'U'; the other strict triangle is left
//          as it was
//   trsm   BLAS trsm (side, uplo, trans, diag, alpha) applied to the V factor of a compressed
//          tile, V viewed as rows x rank with the tile's leading dimension (HCore.cpp:624-647)
//   syrk   BLAS syrk on a dense tile (n = C.rows, k = C.cols, as HCore.cpp:576-582 passes them);
//          for a compressed A = U V^T (V stored n x rank, the Cholesky convention): C <- beta C -
//          U (V^T V) U^T, then the strict triangle opposite to uplo set to zero (HCore.cpp:484-575)
// plus the plain matrix transpose the Cholesky-convention Gemm uses to move between V stored
// n x r and the engine's r x n.
#include "kernels.hpp"
#include "cholesky.hpp"
#include "tileops.hpp"

namespace hcamd {

namespace {

int grid_for(int64_t work) {
    int64_t g = (work + NT - 1) / NT;
    if (g > 1024) g = 1024;
    return (int) (g > 0 ? g : 1);
}

}  // namespace

// W (n x n, ld n) = the symmetric matrix whose `uplo` triangle A holds, stored in its lower
// triangle (strict upper zero).
template <typename T>
__global__ void HCAMD_KERNEL k_sym_to_lower(const T *A, int n, int64_t lda, int upper, T *W) {
    for (int64_t e = (int64_t) blockIdx.x * NT + threadIdx.x; e < (int64_t) n * n; e += (int64_t) gridDim.x * NT) {
        const int i = (int) (e % n), c = (int) (e / n);
        T v = T(0);
        if (i >= c) v = upper ? A[c + (int64_t) i * lda] : A[i + (int64_t) c * lda];
        W[e] = v;
    }
}

// A's `uplo` triangle from the lower factor L in W: L itself (lower) or U = L^T (upper); the
// other strict triangle of A is not written.
template <typename T>
__global__ void HCAMD_KERNEL k_lower_to_tri(const T *W, int n, T *A, int64_t lda, int upper) {
    for (int64_t e = (int64_t) blockIdx.x * NT + threadIdx.x; e < (int64_t) n * n; e += (int64_t) gridDim.x * NT) {
        const int i = (int) (e % n), c = (int) (e / n);
        if (i < c) continue;
        if (upper) A[c + (int64_t) i * lda] = W[e];
        else A[i + (int64_t) c * lda] = W[e];
    }
}

// Triangular solves with many right-hand sides, one thread per independent vector: for side
// Left the columns of B (m x n) solve op(A) x = alpha b (A m x m); for side Right the rows solve
// x op(A) = alpha b, i.e. op(A)^T x^T = alpha b^T (A n x n).  M = op(A) or op(A)^T is lower or
// upper, forward or back substitution accordingly; unit diagonals are not read.
template <typename T>
__global__ void HCAMD_KERNEL k_trsm_vec(int m, int n, T alpha, const T *A, int64_t lda, T *B, int64_t ldb, int left,
                                        int upper, int trans, int unit) {
    const int len = left ? m : n, count = left ? n : m;
    const int v = (int) (blockIdx.x * NT + threadIdx.x);
    if (v >= count) return;
    const int64_t estride = left ? 1 : ldb;            // stride between elements of one vector
    T *x = B + (left ? (int64_t) v * ldb : (int64_t) v);
    const bool teff = left ? trans != 0 : trans == 0;  // M(i, j) = A(j, i) when set
    const bool lower = (upper == 0) != teff;
    auto M = [&](int i, int j) -> T { return teff ? A[j + (int64_t) i * lda] : A[i + (int64_t) j * lda]; };
    if (lower) {
        for (int i = 0; i < len; ++i) {
            T s = alpha * x[i * estride];
            for (int j = 0; j < i; ++j) s -= M(i, j) * x[j * estride];
            x[i * estride] = unit ? s : s / M(i, i);
        }
    } else {
        for (int i = len - 1; i >= 0; --i) {
            T s = alpha * x[i * estride];
            for (int j = i + 1; j < len; ++j) s -= M(i, j) * x[j * estride];
            x[i * estride] = unit ? s : s / M(i, i);
        }
    }
}

// C's uplo triangle (diagonal included) <- alpha op(A) op(A)^T + beta C for op(A) n x k.
template <typename T>
__global__ void HCAMD_KERNEL k_syrk_dense(int n, int k, T alpha, const T *A, int64_t lda, int trans, T beta, T *C,
                                          int64_t ldc, int upper) {
    for (int64_t e = (int64_t) blockIdx.x * NT + threadIdx.x; e < (int64_t) n * n; e += (int64_t) gridDim.x * NT) {
        const int i = (int) (e % n), j = (int) (e / n);
        if (upper ? i > j : i < j) continue;
        T s = T(0);
        if (trans)
            for (int p = 0; p < k; ++p) s += A[p + (int64_t) i * lda] * A[p + (int64_t) j * lda];
        else
            for (int p = 0; p < k; ++p) s += A[i + (int64_t) p * lda] * A[j + (int64_t) p * lda];
        T *c = C + i + (int64_t) j * ldc;
        *c = (beta == T(0)) ? alpha * s : alpha * s + beta * *c;
    }
}

// Zero the strict upper (upper = 1) or strict lower (0) triangle of the square n x n C.
template <typename T>
__global__ void HCAMD_KERNEL k_zero_tri(T *C, int n, int64_t ldc, int upper) {
    for (int64_t e = (int64_t) blockIdx.x * NT + threadIdx.x; e < (int64_t) n * n; e += (int64_t) gridDim.x * NT) {
        const int i = (int) (e % n), j = (int) (e / n);
        if (upper ? i < j : i > j) C[i + (int64_t) j * ldc] = T(0);
    }
}

// dst (n x m, ld ldd) = src (m x n, ld lds)^T, through LDS tiles of 32 x 32.
template <typename T>
__global__ void HCAMD_KERNEL k_transpose(int m, int n, const T *src, int64_t lds, T *dst, int64_t ldd) {
    __shared__ T tile[32][33];
    const int tm = (m + 31) / 32, tn = (n + 31) / 32;
    for (int t = blockIdx.x; t < tm * tn; t += gridDim.x) {
        const int i0 = (t % tm) * 32, j0 = (t / tm) * 32;
        for (int e = threadIdx.x; e < 32 * 32; e += NT) {
            const int i = e % 32, j = e / 32;
            if (i0 + i < m && j0 + j < n) tile[j][i] = src[(i0 + i) + (int64_t) (j0 + j) * lds];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 32 * 32; e += NT) {
            const int j = e % 32, i = e / 32;
            if (i0 + i < m && j0 + j < n) dst[(j0 + j) + (int64_t) (i0 + i) * ldd] = tile[j][i];
        }
        __syncthreads();
    }
}

template <typename T>
hipError_t launch_tile_potrf(hipStream_t s, int upper, T *A, int n, int64_t lda, T *work, int *info) {
    hipLaunchKernelGGL(k_sym_to_lower<T>, dim3(grid_for((int64_t) n * n)), dim3(NT), 0, s, A, n, lda, upper, work);
    hipError_t e = launch_potrf<T>(s, work, n, n, info);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_lower_to_tri<T>, dim3(grid_for((int64_t) n * n)), dim3(NT), 0, s, work, n, A, lda, upper);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_tile_trsm(hipStream_t s, int left, int upper, int trans, int unit, int m, int n, T alpha,
                            const T *A, int64_t lda, T *B, int64_t ldb) {
    const int count = left ? n : m;
    if (count <= 0 || (left ? m : n) <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_trsm_vec<T>, dim3((count + NT - 1) / NT), dim3(NT), 0, s, m, n, alpha, A, lda, B, ldb, left,
                       upper, trans, unit);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_tile_syrk_dense(hipStream_t s, int upper, int trans, int n, int k, T alpha, const T *A,
                                  int64_t lda, T beta, T *C, int64_t ldc) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_syrk_dense<T>, dim3(grid_for((int64_t) n * n)), dim3(NT), 0, s, n, k, alpha, A, lda, trans,
                       beta, C, ldc, upper);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_zero_tri(hipStream_t s, T *C, int n, int64_t ldc, int upper) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_tri<T>, dim3(grid_for((int64_t) n * n)), dim3(NT), 0, s, C, n, ldc, upper);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_transpose(hipStream_t s, int m, int n, const T *src, int64_t lds, T *dst, int64_t ldd) {
    if (m <= 0 || n <= 0) return hipSuccess;
    const int tiles = ((m + 31) / 32) * ((n + 31) / 32);
    hipLaunchKernelGGL(k_transpose<T>, dim3(tiles < 1024 ? tiles : 1024), dim3(NT), 0, s, m, n, src, lds, dst, ldd);
    return hipGetLastError();
}

#define HCAMD_TILEOPS_INSTANTIATE(T)                                                                             \
    template hipError_t launch_tile_potrf<T>(hipStream_t, int, T *, int, int64_t, T *, int *);                   \
    template hipError_t launch_tile_trsm<T>(hipStream_t, int, int, int, int, int, int, T, const T *, int64_t, T *, \
                                            int64_t);                                                            \
    template hipError_t launch_tile_syrk_dense<T>(hipStream_t, int, int, int, int, T, const T *, int64_t, T, T *, \
                                                  int64_t);                                                      \
    template hipError_t launch_zero_tri<T>(hipStream_t, T *, int, int64_t, int);                                 \
    template hipError_t launch_transpose<T>(hipStream_t, int, int, const T *, int64_t, T *, int64_t);

HCAMD_TILEOPS_INSTANTIATE(double)
HCAMD_TILEOPS_INSTANTIATE(float)

}  // namespace hcamd
