// Workgroup-level tiled GEMM for the factor products and dense tile updates
// (HCoreKernels::Gemm call sites: reference src/api/HCore.cpp:176-264 via
// src/operators/concrete/Dense.cpp:46-108).  One workgroup computes one
// 64x64 tile of C = alpha*op(A)*op(B) + beta*C with a 16-deep K chunk staged
// in LDS; each thread owns a 4x4 register block.
#pragma once

#include "device_common.hpp"

namespace hcamd {

constexpr int GT_M = 64, GT_N = 64, GT_K = 16;
constexpr int GT_LDS_ELEMS = GT_K * (GT_M + 1) + GT_K * (GT_N + 1);

template <typename T>
__device__ void wg_gemm_tile(int M, int N, int Kd, T alpha, const T *A, int64_t lda, bool ta, const T *B,
                             int64_t ldb, bool tb, T beta, T *C, int64_t ldc, int ti, int tj, T *lds,
                             bool upper_b = false) {
    T *As = lds;                      // [GT_K][GT_M + 1]  As[k][i]
    T *Bs = lds + GT_K * (GT_M + 1);  // [GT_K][GT_N + 1]  Bs[k][j]
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int i0 = ti * GT_M, j0 = tj * GT_N;
    T acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = T(0);
    for (int k0 = 0; k0 < Kd; k0 += GT_K) {
        // A tile: i in [i0, i0+64), k in [k0, k0+16)
        for (int e = threadIdx.x; e < GT_M * GT_K; e += NT) {
            int i, k;
            if (!ta) {
                i = e % GT_M;
                k = e / GT_M;
            } else {
                k = e % GT_K;
                i = e / GT_K;
            }
            const int gi = i0 + i, gk = k0 + k;
            T v = T(0);
            if (gi < M && gk < Kd) v = ta ? A[gk + (int64_t) gi * lda] : A[gi + (int64_t) gk * lda];
            As[k * (GT_M + 1) + i] = v;
        }
        for (int e = threadIdx.x; e < GT_N * GT_K; e += NT) {
            int j, k;
            if (!tb) {
                k = e % GT_K;
                j = e / GT_K;
            } else {
                j = e % GT_N;
                k = e / GT_N;
            }
            const int gj = j0 + j, gk = k0 + k;
            T v = T(0);
            if (gj < N && gk < Kd && !(upper_b && gk < gj))
                v = tb ? B[gj + (int64_t) gk * ldb] : B[gk + (int64_t) gj * ldb];
            Bs[k * (GT_N + 1) + j] = v;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < GT_K; ++k) {
            T a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a[q] = As[k * (GT_M + 1) + ty + 16 * q];
                b[q] = Bs[k * (GT_N + 1) + tx + 16 * q];
            }
#pragma unroll
            for (int p = 0; p < 4; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[p][q] += a[p] * b[q];
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int gi = i0 + ty + 16 * p;
        if (gi >= M) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int gj = j0 + tx + 16 * q;
            if (gj >= N) continue;
            T *c = C + gi + (int64_t) gj * ldc;
            *c = (beta == T(0)) ? alpha * acc[p][q] : alpha * acc[p][q] + beta * *c;
        }
    }
}

// Whole-matrix GEMM by one workgroup (loops over all 64x64 tiles).
// `upper_b` treats op(B)(k, j) as zero for k < j (a transposed upper-triangular factor).
template <typename T>
__device__ void wg_gemm(int M, int N, int Kd, T alpha, const T *A, int64_t lda, bool ta, const T *B, int64_t ldb,
                        bool tb, T beta, T *C, int64_t ldc, T *lds) {
    const int tm = (M + GT_M - 1) / GT_M, tn = (N + GT_N - 1) / GT_N;
    for (int t = 0; t < tm * tn; ++t) {
        wg_gemm_tile<T>(M, N, Kd, alpha, A, lda, ta, B, ldb, tb, beta, C, ldc, t % tm, t / tm, lds);
    }
}

}  // namespace hcamd
