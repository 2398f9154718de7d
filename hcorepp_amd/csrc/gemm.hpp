// Workgroup-level tiled GEMM for the factor products and dense tile updates
// (HCoreKernels::Gemm call sites: reference src/api/HCore.cpp:176-264 via
// src/operators/concrete/Dense.cpp:46-108).  One workgroup computes one
// 64x64 tile of C = alpha*op(A)*op(B) + beta*C with a K chunk (32 deep in FP64, 16 in FP32)
// staged in LDS; FP64 on the matrix cores, FP32 with a 4x4 register block per thread.
#pragma once

#include "device_common.hpp"

namespace hcamd {

constexpr int GT_M = 64, GT_N = 64, GT_K = 16;
// K chunk of the FP64 (matrix-core) form: 32 deep, so that each thread has 16 loads in flight
// per chunk and a tile takes half as many memory round trips (the FP32 form and the mixed split
// keep 16: the split's exactness bound counts 16-deep sums).
constexpr int GT_KD = 32;
constexpr int GT_LDS_ELEMS = GT_KD * (GT_M + 1) + GT_KD * (GT_N + 1);
template <typename T>
__host__ __device__ constexpr int gemm_kc() {
    return sizeof(T) == 8 ? GT_KD : GT_K;
}

// The K loop (chunks of gemm_kc<T>() = KC) is software-pipelined: the next chunk's elements are
// loaded into registers (KC / 4 of A and of B per thread) right after the current chunk is in LDS, so their memory latency
// overlaps the chunk's arithmetic.  `fa(i, k)` and `fb(k, j)` return op(A)[i, k] and op(B)[k, j]
// (zero outside the operands); A's chunk is read row-fastest when `a_rows_fast`, B's K-fastest
// unless `b_cols_fast`, so that consecutive threads read consecutive addresses.  `epi(i, j, v)`
// receives each element of the tile inside M x N.
// FP64 runs on the matrix cores (v_mfma_f64_16x16x4f64): wave w computes the 32 x 32 quarter at
// rows 32 (w / 2), columns 32 (w % 2) of the tile as 2 x 2 blocks of 16 x 16; per 4-deep K step
// a lane reads two values of A and two of B from LDS for four MFMAs (the VALU form, a 4 x 4
// block per thread, reads eight for sixteen FMAs).  MFMA lane layout: A[i][k] in lane i + 16 k,
// B[k][j] in lane j + 16 k, D[4 v + l / 16][l % 16] in element v of lane l.
// Only K chunks overlapping [k_lo, k_hi) are visited (operands known to vanish outside, as for
// triangular factors); k_lo is rounded down to a chunk boundary.
template <typename T, typename FA, typename FB, typename EPI>
__device__ __forceinline__ void wg_gemm_tile_core(int M, int N, int Kd, const FA &fa, const FB &fb, bool a_rows_fast,
                                                  bool b_cols_fast, const EPI &epi, int ti, int tj, T *lds,
                                                  int k_lo = 0, int k_hi = -1) {
    constexpr int KC = gemm_kc<T>();
    k_lo = (k_lo / KC) * KC;
    if (k_hi < 0 || k_hi > Kd) k_hi = Kd;
    T *As = lds;                    // [KC][GT_M + 1]  As[k][i]
    T *Bs = lds + KC * (GT_M + 1);  // [KC][GT_N + 1]  Bs[k][j]
    const int i0 = ti * GT_M, j0 = tj * GT_N;
    constexpr int PER = GT_M * KC / NT;  // elements of each operand chunk per thread
    static_assert(GT_M * KC == GT_N * KC && PER * NT == GT_M * KC, "tile shape");
    auto a_pos = [&](int e, int &i, int &k) {
        if (a_rows_fast) {
            i = e % GT_M;
            k = e / GT_M;
        } else {
            k = e % KC;
            i = e / KC;
        }
    };
    auto b_pos = [&](int e, int &j, int &k) {
        if (b_cols_fast) {
            j = e % GT_N;
            k = e / GT_N;
        } else {
            k = e % KC;
            j = e / KC;
        }
    };
    T ra[PER], rb[PER];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = threadIdx.x + q * NT;
            int i, j, k, k2;
            a_pos(e, i, k);
            b_pos(e, j, k2);
            const int gi = i0 + i, gk = k0 + k, gj = j0 + j, gk2 = k0 + k2;
            ra[q] = (gi < M && gk < Kd) ? fa(gi, gk) : T(0);
            rb[q] = (gj < N && gk2 < Kd) ? fb(gk2, gj) : T(0);
        }
    };
    auto stage = [&]() {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = threadIdx.x + q * NT;
            int i, j, k, k2;
            a_pos(e, i, k);
            b_pos(e, j, k2);
            As[k * (GT_M + 1) + i] = ra[q];
            Bs[k2 * (GT_N + 1) + j] = rb[q];
        }
    };
    if constexpr (sizeof(T) == 8) {
        const int l = lane_id(), i16 = l & 15, k4 = l >> 4, w = wave_id();
        const int qi = 32 * (w >> 1), qj = 32 * (w & 1);
        hcamd_doublex4 acc[2][2];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = hcamd_doublex4{0.0, 0.0, 0.0, 0.0};
        if (k_lo < k_hi) load(k_lo);
        for (int k0 = k_lo; k0 < k_hi; k0 += KC) {
            stage();
            __syncthreads();
            if (k0 + KC < k_hi) load(k0 + KC);
#pragma unroll
            for (int s = 0; s < KC / 4; ++s) {
                const int kr = 4 * s + k4;
                double av[2], bv[2];
#pragma unroll
                for (int bi = 0; bi < 2; ++bi) av[bi] = As[kr * (GT_M + 1) + qi + 16 * bi + i16];
#pragma unroll
                for (int bj = 0; bj < 2; ++bj) bv[bj] = Bs[kr * (GT_N + 1) + qj + 16 * bj + i16];
#pragma unroll
                for (int bi = 0; bi < 2; ++bi)
#pragma unroll
                    for (int bj = 0; bj < 2; ++bj)
                        acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[bi], bv[bj], acc[bi][bj], 0, 0, 0);
            }
            __syncthreads();
        }
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int gi = i0 + qi + 16 * bi + 4 * v + k4, gj = j0 + qj + 16 * bj + i16;
                    if (gi < M && gj < N) epi(gi, gj, acc[bi][bj][v]);
                }
    } else {
        const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
        T acc[4][4];
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = 0; q < 4; ++q) acc[p][q] = T(0);
        if (k_lo < k_hi) load(k_lo);
        for (int k0 = k_lo; k0 < k_hi; k0 += KC) {
            stage();
            __syncthreads();
            if (k0 + KC < k_hi) load(k0 + KC);
#pragma unroll
            for (int k = 0; k < KC; ++k) {
                T av[4], bv[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    av[q] = As[k * (GT_M + 1) + ty + 16 * q];
                    bv[q] = Bs[k * (GT_N + 1) + tx + 16 * q];
                }
#pragma unroll
                for (int p = 0; p < 4; ++p)
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[p][q] += av[p] * bv[q];
            }
            __syncthreads();
        }
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int gi = i0 + ty + 16 * p, gj = j0 + tx + 16 * q;
                if (gi < M && gj < N) epi(gi, gj, acc[p][q]);
            }
    }
}

template <typename T>
__device__ void wg_gemm_tile(int M, int N, int Kd, T alpha, const T *A, int64_t lda, bool ta, const T *B,
                             int64_t ldb, bool tb, T beta, T *C, int64_t ldc, int ti, int tj, T *lds,
                             bool upper_b = false) {
    auto fa = [=](int i, int k) -> T { return ta ? A[k + (int64_t) i * lda] : A[i + (int64_t) k * lda]; };
    auto fb = [=](int k, int j) -> T {
        if (upper_b && k < j) return T(0);
        return tb ? B[j + (int64_t) k * ldb] : B[k + (int64_t) j * ldb];
    };
    auto epi = [=](int gi, int gj, T v) {
        T *c = C + gi + (int64_t) gj * ldc;
        *c = (beta == T(0)) ? alpha * v : alpha * v + beta * *c;
    };
    // with upper_b, op(B)(k, j) vanishes for k < j: the tile's columns start at tj * GT_N
    wg_gemm_tile_core<T>(M, N, Kd, fa, fb, !ta, tb, epi, ti, tj, lds, upper_b ? tj * GT_N : 0);
}

// One 64 x 64 tile of C (M x N, ld ldc) = sum_k fa(i, k) fb(k, j) for operands given as element
// functions (each returns zero outside its structure: triangles, gathered rows); K chunks are
// staged in LDS as in wg_gemm_tile.  A's chunk is read row-fastest, B's K-fastest.
template <typename T, typename FA, typename FB>
__device__ void wg_gemm_tile_fn(int M, int N, int Kd, const FA &fa, const FB &fb, T *C, int64_t ldc, int ti, int tj,
                                T *lds, int k_lo = 0, int k_hi = -1) {
    auto epi = [=](int gi, int gj, T v) { C[gi + (int64_t) gj * ldc] = v; };
    wg_gemm_tile_core<T>(M, N, Kd, fa, fb, true, false, epi, ti, tj, lds, k_lo, k_hi);
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

// ---------------------------------------------------------------------------
// Mixed-precision tile GEMM: C = op(A) op(B) (alpha = 1, beta = 0) for FP64 operands with the
// products formed on the FP32 matrix cores (v_mfma_f32_16x16x4f32) with FP32 accumulation, made
// exact by an error-free split of the operands (the Ozaki scheme).  Each row i of op(A) is scaled
// by a power of two 2^ea_i above its largest magnitude and cut into MS_SLICES slices: integers
// x_s of at most MS_BITS bits with op(A)[i, k] = 2^ea_i sum_s x_s 2^(-MS_BITS (s+1)) + (a rest
// below 2^(ea_i - MS_BITS MS_SLICES - 1)); each column of op(B) likewise.  A slice product summed
// over one 16-deep K chunk is an integer of magnitude at most 16 * 2^(2 MS_BITS) = 2^24, exact
// in FP32 whatever order the matrix core sums in; every slice pair (s, t) with s + t < MS_SLICES
// is summed per chunk in FP32 and added to the FP64 result with its power-of-two weight.  The
// error is the dropped slices: about 2^-40 of 2^(ea_i + eb_j) per term (FP64 rounding for the
// rest), against 2^-24 for plain FP32 operands.
// A 64 x 64 output tile per workgroup: wave w computes the 32 x 32 quarter at rows 32 (w / 2),
// columns 32 (w % 2) as 2 x 2 blocks of 16 x 16.
// MFMA lane layout (16x16x4): A[i][k] from lane i + 16 k, B[k][j] from lane j + 16 k,
// D[4 (l / 16) + v][l % 16] in element v of lane l.
// ---------------------------------------------------------------------------
typedef float hcamd_floatx4 __attribute__((ext_vector_type(4)));
constexpr int MS_BITS = 10, MS_SLICES = 4;
constexpr int GTM_SLICE_A = GT_K * (GT_M + 1), GTM_SLICE_B = GT_K * (GT_N + 1);
constexpr int GTM_LDS_FLOATS = MS_SLICES * (GTM_SLICE_A + GTM_SLICE_B) + GT_M + GT_N;
static_assert(2 * MS_BITS + 4 <= 24, "slice products over a 16-deep chunk must be exact in FP32");

// Slices of v in [-1, 1] (a value divided by its row's or column's power-of-two scale).
__device__ __forceinline__ void ms_split(double v, float (&x)[MS_SLICES]) {
    double r = v;
#pragma unroll
    for (int s = 0; s < MS_SLICES; ++s) {
        const double q = rint(ldexp(r, MS_BITS * (s + 1)));
        x[s] = (float) q;
        r -= ldexp(q, -MS_BITS * (s + 1));
    }
}

// Exponent e with |v| < 2^e for the largest magnitude v of a row or column (0 for a zero one).
__device__ __forceinline__ int ms_exponent(double vmax) {
    return vmax > 0.0 ? ilogb(vmax) + 1 : 0;
}

__device__ inline void wg_gemm_tile_mixed(int M, int N, int Kd, const double *A, int64_t lda, bool ta,
                                          const double *B, int64_t ldb, bool tb, double *C, int64_t ldc, int ti,
                                          int tj, float *lds) {
    float *As = lds;                                   // [slice][k][i], GTM_SLICE_A per slice
    float *Bs = lds + MS_SLICES * GTM_SLICE_A;         // [slice][k][j]
    int *ea = reinterpret_cast<int *>(Bs + MS_SLICES * GTM_SLICE_B);  // row exponents of op(A)
    int *eb = ea + GT_M;                                               // column exponents of op(B)
    const int i0 = ti * GT_M, j0 = tj * GT_N;
    const int l = lane_id(), w = wave_id();
    const int wr = 32 * (w / 2), wc = 32 * (w % 2);
    auto opA = [&](int i, int k) -> double { return ta ? A[k + (int64_t) i * lda] : A[i + (int64_t) k * lda]; };
    auto opB = [&](int k, int j) -> double { return tb ? B[j + (int64_t) k * ldb] : B[k + (int64_t) j * ldb]; };
    // power-of-two scales: four threads per row of op(A) (then per column of op(B)) scan K
    {
        const int idx = threadIdx.x / 4, part = threadIdx.x % 4;
        double amax = 0.0, bmax = 0.0;
        if (i0 + idx < M)
            for (int k = part; k < Kd; k += 4) amax = fmax(amax, fabs(opA(i0 + idx, k)));
        if (j0 + idx < N)
            for (int k = part; k < Kd; k += 4) bmax = fmax(bmax, fabs(opB(k, j0 + idx)));
        amax = fmax(amax, __shfl_xor(amax, 1, WAVE));
        amax = fmax(amax, __shfl_xor(amax, 2, WAVE));
        bmax = fmax(bmax, __shfl_xor(bmax, 1, WAVE));
        bmax = fmax(bmax, __shfl_xor(bmax, 2, WAVE));
        if (part == 0) {
            ea[idx] = ms_exponent(amax);
            eb[idx] = ms_exponent(bmax);
        }
    }
    double acc[2][2][4];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[p][q][v] = 0.0;
    __syncthreads();
    for (int k0 = 0; k0 < Kd; k0 += GT_K) {
        for (int e = threadIdx.x; e < GT_M * GT_K; e += NT) {
            int i, k;
            if (ta) {
                k = e % GT_K;
                i = e / GT_K;
            } else {
                i = e % GT_M;
                k = e / GT_M;
            }
            const int gi = i0 + i, gk = k0 + k;
            const double v = (gi < M && gk < Kd) ? ldexp(opA(gi, gk), -ea[i]) : 0.0;
            float x[MS_SLICES];
            ms_split(v, x);
#pragma unroll
            for (int sl = 0; sl < MS_SLICES; ++sl) As[sl * GTM_SLICE_A + k * (GT_M + 1) + i] = x[sl];
        }
        for (int e = threadIdx.x; e < GT_N * GT_K; e += NT) {
            int j, k;
            if (tb) {
                j = e % GT_N;
                k = e / GT_N;
            } else {
                k = e % GT_K;
                j = e / GT_K;
            }
            const int gj = j0 + j, gk = k0 + k;
            const double v = (gj < N && gk < Kd) ? ldexp(opB(gk, gj), -eb[j]) : 0.0;
            float x[MS_SLICES];
            ms_split(v, x);
#pragma unroll
            for (int sl = 0; sl < MS_SLICES; ++sl) Bs[sl * GTM_SLICE_B + k * (GT_N + 1) + j] = x[sl];
        }
        __syncthreads();
#pragma unroll
        for (int sa = 0; sa < MS_SLICES; ++sa) {
#pragma unroll
            for (int sb = 0; sb + sa < MS_SLICES; ++sb) {
                hcamd_floatx4 part[2][2];
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int q = 0; q < 2; ++q) part[p][q] = hcamd_floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kk = 0; kk < GT_K; kk += 4) {
                    const int k = kk + l / 16;
                    float a[2], b[2];
#pragma unroll
                    for (int p = 0; p < 2; ++p) a[p] = As[sa * GTM_SLICE_A + k * (GT_M + 1) + wr + 16 * p + (l % 16)];
#pragma unroll
                    for (int q = 0; q < 2; ++q) b[q] = Bs[sb * GTM_SLICE_B + k * (GT_N + 1) + wc + 16 * q + (l % 16)];
#pragma unroll
                    for (int p = 0; p < 2; ++p)
#pragma unroll
                        for (int q = 0; q < 2; ++q)
                            part[p][q] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[p], b[q], part[p][q], 0, 0, 0);
                }
                const double wgt = ldexp(1.0, -MS_BITS * (sa + sb + 2));
#pragma unroll
                for (int p = 0; p < 2; ++p)
#pragma unroll
                    for (int q = 0; q < 2; ++q)
#pragma unroll
                        for (int v = 0; v < 4; ++v) acc[p][q][v] += (double) part[p][q][v] * wgt;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int ri = wr + 16 * p + 4 * (l / 16) + v, cj = wc + 16 * q + (l % 16);
                const int gi = i0 + ri, gj = j0 + cj;
                if (gi < M && gj < N) C[gi + (int64_t) gj * ldc] = ldexp(acc[p][q][v], ea[ri] + eb[cj]);
            }
    __syncthreads();
}

__device__ inline void wg_gemm_mixed(int M, int N, int Kd, const double *A, int64_t lda, bool ta, const double *B,
                                     int64_t ldb, bool tb, double *C, int64_t ldc, float *lds) {
    const int tm = (M + GT_M - 1) / GT_M, tn = (N + GT_N - 1) / GT_N;
    for (int t = 0; t < tm * tn; ++t)
        wg_gemm_tile_mixed(M, N, Kd, A, lda, ta, B, ldb, tb, C, ldc, t % tm, t / tm, lds);
}

}  // namespace hcamd
