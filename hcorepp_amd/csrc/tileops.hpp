// Launchers of the reference-semantics tile operations (tileops.hip): potrf for either triangle,
// BLAS trsm and syrk on device tiles, and a plain matrix transpose.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hcamd {

// Cholesky factor of the symmetric n x n A held in its upper (upper = 1) or lower triangle,
// written back to that triangle (A = U^T U or L L^T); the other strict triangle is untouched.
// work: n * n elements of device scratch; info: first non-positive pivot (1-based) or 0.
template <typename T>
hipError_t launch_tile_potrf(hipStream_t s, int upper, T *A, int n, int64_t lda, T *work, int *info);
// BLAS trsm on B (m x n, ld ldb): left ? op(A)^{-1} : (.) op(A)^{-1}, times alpha.
template <typename T>
hipError_t launch_tile_trsm(hipStream_t s, int left, int upper, int trans, int unit, int m, int n, T alpha,
                            const T *A, int64_t lda, T *B, int64_t ldb);
// BLAS syrk on the uplo triangle of C (n x n): alpha op(A) op(A)^T + beta C, op(A) n x k.
template <typename T>
hipError_t launch_tile_syrk_dense(hipStream_t s, int upper, int trans, int n, int k, T alpha, const T *A,
                                  int64_t lda, T beta, T *C, int64_t ldc);
// Zero the strict upper (upper = 1) or lower triangle of the square C.
template <typename T>
hipError_t launch_zero_tri(hipStream_t s, T *C, int n, int64_t ldc, int upper);
// dst (n x m, ld ldd) = src (m x n, ld lds)^T.
template <typename T>
hipError_t launch_transpose(hipStream_t s, int m, int n, const T *src, int64_t lds, T *dst, int64_t ldd);

}  // namespace hcamd
