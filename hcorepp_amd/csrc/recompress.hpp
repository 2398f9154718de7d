// Workgroup-level device phases of one compressed-tile update
// C <- alpha * Afin * Bfin + beta * C with C = C_U * C_V (rank rc, capacity max_rank).
//
// The phases compute the truncated SVD that CompressedTile<T>::Gemm computes
// (reference src/operators/concrete/Compressed.cpp:208-694), arranged for CDNA4:
//
//   stack   U_s = [C_U | alpha*Afin] (m x K), V_s = [(beta*C_V)^T | Bfin^T] (n x K)
//           (LaCpy#1/#2, MultiplyByAlpha, ProcessVpointer: :331-349, :375-380)
//   qr_u    U_s = Q_U R_U, blocked Householder, reflectors kept     (Geqrf#1 :351-354)
//   y       Y = V_s R_U^T  (n x mu), so that U_s V_s^T = Q_U Y^T
//   qr_y    Y = Q_Y R_Y                                             (replaces Geqrf#2 :383-385)
//   svd     one-sided block Jacobi on J = R_Y^T (mu x kk): J V1 = Z1, sigma_j = |z_j|,
//           descending sort, rank rule and clamps                   (SVD#3 + rank :454-527)
//   recon   C_U = Q_U [Z1 Sigma^-1; 0]_r ,  C_V = (Q_Y [V1 Sigma; 0]_r)^T
//                                                                   (ungqr + Gemm#3/#4 :549-628)
// with mu = min(m, K), kk = min(n, mu) = min(m, n, K) = size_s.  Since
// U_s V_s^T = Q_U R_Y^T Q_Y^T = Q_U Z1 V1^T Q_Y^T, the left singular vectors are
// Q_U Z1 Sigma^-1 and the right ones Q_Y V1, identical in exact arithmetic to the
// reference's Q_U Unew and Q_V VTnew^T.  R_Y equals the R factor of the QR of
// (R_U R_V^T)^T, a triangular preconditioner that cuts the Jacobi sweep count ~3x.
// "V carries Sigma" (:597-599) is the V1 Sigma factor.
//
// Every phase is executed by a whole workgroup of NT threads; the same code
// runs one phase per launch (plain path) or all phases in one persistent
// launch (batched path).
#pragma once

#include "device_common.hpp"
#include "gemm.hpp"
#include "jacobi.hpp"

// Register-heavy phase variants (Householder panel application, block Jacobi) are
// compiled out of line by default so that each variant gets its own register allocation.
#ifndef HCAMD_PHASE_ATTR
#define HCAMD_PHASE_ATTR __noinline__
#endif

namespace hcamd {

// Sub-phase timers of the timing build: HCAMD_SUB_BEGIN(tv) starts stopwatch tv,
// HCAMD_SUB_END(tv, hdr, slot) adds its elapsed device clock (thread 0) to header word HDR_SUB + slot.
#ifdef HCAMD_PHASE_TIMING
#define HCAMD_SUB_BEGIN(tv) const uint64_t tv = (uint64_t) wall_clock64()
#define HCAMD_SUB_END(tv, hdr, slot)                                                               \
    do {                                                                                          \
        if ((hdr) && threadIdx.x == 0) (hdr)[HDR_SUB + (slot)] += (int64_t) ((uint64_t) wall_clock64() - (tv)); \
    } while (0)
#else
#define HCAMD_SUB_BEGIN(tv) (void) 0
#define HCAMD_SUB_END(tv, hdr, slot) (void) 0
#endif

// Header slots at the start of each update's workspace (int64).
enum : int {
    HDR_K = 0,       // stacked rank K = arank + rc
    HDR_RANK = 1,    // new rank r
    HDR_SIZES = 2,   // size_s = min(m, n, K)
    HDR_SWEEPS = 3,  // Jacobi sweeps used
    HDR_STATUS = 4,  // status bits (STATUS_*)
    HDR_KT = 5,      // columns of the truncated core kept for the SVD (<= size_s)
    HDR_JFRO2 = 6,   // ||J||_F^2 of the truncated core as double bits (multi-workgroup Jacobi)
    HDR_CQR = 7,     // 1: the core was QR-factored (J = Q2 R2) and the Jacobi ran on R2
    HDR_T0 = 8,      // 8 phase timestamps (builds with HCAMD_PHASE_TIMING only)
    HDR_SUB = 16,    // 8 sub-phase cycle accumulators (builds with HCAMD_PHASE_TIMING only)
    HDR_FEAT = 24,   // RANK_FEATURES float32 rank-predictor features (12 words)
    HDR_PRED = 36,   // predicted rank (learned core sizing), 0 when off
    HDR_LEARNED = 37,  // 1: core sized by the prediction, 2: prediction rejected (fallback), 0: off
    HDR_KQ = 38,     // rows of R_Y the QR of Y computed when it stopped early (0: all of them)
    HDR_TRAIL2 = 39,  // squared Frobenius norm of the part it left unfactored, as double bits
    HDR_STOP2 = 40,  // the stop test's threshold (multi-workgroup QR of Y), as double bits
    HDR_WORDS = 48,
};
enum : int { STATUS_NOCONV = 1, STATUS_NAN = 2, STATUS_TOO_BIG = 4 };

constexpr int JACOBI_MAX_SWEEPS = 40;
constexpr int QR_PB_MAX = 16;  // widest Householder panel

// Descriptor of one compressed-C update.  Pointers are device pointers.
template <typename T>
struct RecompDesc {
    const T *afin;  // m x arank, ld ldaf  (the final "A" factor: C.Gemm's aTileA)
    const T *bfin;  // arank x n, ld ldbf  (the final "B" factor: C.Gemm's aTileB)
    T *cdata;       // [U m x max_rank | V max_rank x n]
    int64_t *crank; // in: rc, out: new rank
    T *ws;          // workspace base (layout from ws_layout)
    int64_t ldaf, ldbf;
    int32_t m, n, arank, max_rank;
    int32_t kmax;        // capacity for K used to lay out ws
    int32_t truncated_svd;
    int32_t fixed_rank;
    int32_t rb;          // input rank of B (rank-predictor feature; 0 when B is dense)
    T alpha, beta;
    double accuracy;
    double *sigma_out;   // optional: sorted singular values (size_s entries)
    const uint16_t *rank_model;  // BF16 rank-predictor parameters, or null (exact core sizing)
    double model_tol;            // learned core sizing: accept when the dropped part <= model_tol * accuracy
    int32_t model_margin;        // columns kept beyond the predicted rank
    int32_t pad2;
};

// Workspace layout (element offsets of T) for capacity kmax and tile m x n.
struct WsLayout {
    int64_t hdr, us, vs, y, tau_u, tau_y, t_u, t_y, j, v1, sig, perm, piv, vtmp, j2, tau_j, t_j, total;
};

__host__ __device__ inline int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

template <typename T>
__host__ __device__ inline WsLayout ws_layout(int64_t m, int64_t n, int64_t kmax, int64_t max_rank) {
    WsLayout L;
    int64_t o = 0;
    const int64_t mu = m < kmax ? m : kmax;
    const int64_t kk = n < mu ? n : mu;
    const int64_t hdr_elems = (HDR_WORDS * 8 + sizeof(T) - 1) / sizeof(T);
    L.hdr = o; o += align_up(hdr_elems, 32);
    L.us = o; o += align_up(m * kmax, 32);
    L.vs = o;  // no region: V_s is read in place from C_V and Bfin (StackedV)
    L.y = o; o += align_up(n * mu, 32);
    L.tau_u = o; o += align_up(kmax, 32);
    L.tau_y = o; o += align_up(kmax, 32);
    L.t_u = o; o += align_up(kmax * QR_PB_MAX + QR_PB_MAX * QR_PB_MAX, 32);
    L.t_y = o; o += align_up(kmax * QR_PB_MAX + QR_PB_MAX * QR_PB_MAX, 32);
    L.j = o; o += align_up(mu * kk, 32);
    L.v1 = o; o += align_up(kk * kk, 32);
    L.sig = o; o += align_up(kk, 32);
    L.perm = o; o += align_up((kk * 4 + sizeof(T) - 1) / sizeof(T), 32);
    L.piv = o; o += align_up((kmax * 4 + sizeof(T) - 1) / sizeof(T), 32);
    L.vtmp = o; o += align_up(n * (max_rank > 0 ? max_rank : 1), 32);
    L.j2 = o; o += align_up(kk * kk, 32);
    L.tau_j = o; o += align_up(kk, 32);
    L.t_j = o; o += align_up(kk * QR_PB_MAX + QR_PB_MAX * QR_PB_MAX, 32);
    L.total = o;
    return L;
}

template <typename T>
struct Ws {
    int64_t *hdr;
    T *us, *y, *tau_u, *tau_y, *t_u, *t_y, *j, *v1, *sig, *vtmp, *j2, *tau_j, *t_j;
    int32_t *perm;
    int32_t *piv;  // column permutation of the pivoted QR of Y
};

template <typename T>
__device__ inline Ws<T> ws_of(const RecompDesc<T> &d) {
    const WsLayout L = ws_layout<T>(d.m, d.n, d.kmax, d.max_rank);
    Ws<T> w;
    w.hdr = reinterpret_cast<int64_t *>(d.ws + L.hdr);
    w.us = d.ws + L.us;
    w.y = d.ws + L.y;
    w.tau_u = d.ws + L.tau_u;
    w.tau_y = d.ws + L.tau_y;
    w.t_u = d.ws + L.t_u;
    w.t_y = d.ws + L.t_y;
    w.j = d.ws + L.j;
    w.v1 = d.ws + L.v1;
    w.sig = d.ws + L.sig;
    w.perm = reinterpret_cast<int32_t *>(d.ws + L.perm);
    w.piv = reinterpret_cast<int32_t *>(d.ws + L.piv);
    w.vtmp = d.ws + L.vtmp;
    w.j2 = d.ws + L.j2;
    w.tau_j = d.ws + L.tau_j;
    w.t_j = d.ws + L.t_j;
    return w;
}

struct Dims {
    int K, mu, kk;
    int kt;  // truncated core width (valid after ph_core)
};

template <typename T>
__device__ inline Dims dims_of(const RecompDesc<T> &d, const Ws<T> &w) {
    Dims z;
    z.K = (int) w.hdr[HDR_K];
    z.mu = d.m < z.K ? d.m : z.K;
    z.kk = d.n < z.mu ? d.n : z.mu;
    z.kt = (int) w.hdr[HDR_KT];
    return z;
}

// ---------------------------------------------------------------------------
// Phase 1: stack.  Grid-stride so several workgroups may share one update
// (`wg` in [0, nwg)).
// ---------------------------------------------------------------------------
// COPY false: the header words and U_s's first 16 columns only, the first panel of the QR (the
// batched FP64 form's QR of U_s reads the other columns in place through StackedU in its first
// trailing update, which writes them into the workspace).
template <typename T, bool COPY = true>
__device__ void ph_stack(const RecompDesc<T> &d, int wg, int nwg) {
    Ws<T> w = ws_of(d);
    const int64_t rc = *d.crank;
    const int64_t m = d.m, n = d.n, ra = d.arank;
    const int64_t K = ra + rc;
    if (wg == 0 && threadIdx.x == 0) {
        w.hdr[HDR_K] = K;
        w.hdr[HDR_STATUS] = (K > d.kmax) ? STATUS_TOO_BIG : 0;
        w.hdr[HDR_PRED] = 0;
        w.hdr[HDR_LEARNED] = 0;
        w.hdr[HDR_KQ] = 0;
        w.hdr[HDR_TRAIL2] = 0;
    }
    if (K > d.kmax) return;
    (void) n;
    const T *cu = d.cdata;
    const T *af = d.afin;
    const T alpha = d.alpha;
    const int64_t ldaf = d.ldaf;
    // U_s only: V_s = [beta C_V^T | Bfin] is read in place by its one consumer, Y = V_s R_U^T
    // (ph_y).  m * K < 2^31 (m <= 2048, K <= 768), so the element index is 32-bit; four elements
    // per thread and pass have their loads in flight together.
    const int mm = (int) m, totU = (int) (m * (COPY || K < 16 ? K : 16)), irc = (int) rc;
    constexpr int EPT = 4;
    for (int e0 = wg * NT * EPT + (int) threadIdx.x; e0 < totU; e0 += nwg * NT * EPT) {
        T v[EPT];
#pragma unroll
        for (int q = 0; q < EPT; ++q) {
            const int e = e0 + q * NT;
            if (e < totU) {
                const int i = e % mm, j = e / mm;
                v[q] = (j < irc) ? cu[i + (int64_t) j * mm] : alpha * af[i + (int64_t) (j - irc) * ldaf];
            }
        }
#pragma unroll
        for (int q = 0; q < EPT; ++q)
            if (e0 + q * NT < totU) w.us[e0 + q * NT] = v[q];
    }
}

// V_s[i, k] of the stacked right factor (n x K): beta * C_V[k, i] for k < rc, Bfin[k - rc, i]
// after (ProcessVpointer, reference src/kernels/omp/kernels.cpp:69-76), read in place from C's
// V factor (rc x n, ld rc) and Bfin (ld ldbf).
template <typename T>
struct StackedV {
    const T *cv, *bf;
    T beta;
    int rc;
    int64_t ldbf;
    __device__ T operator()(int i, int k) const {
        return k < rc ? beta * cv[k + (int64_t) i * rc] : bf[(k - rc) + (int64_t) i * ldbf];
    }
};

template <typename T>
__device__ inline StackedV<T> stacked_v(const RecompDesc<T> &d, int K) {
    const int rc = K - d.arank;
    return StackedV<T>{d.cdata + (int64_t) d.m * d.max_rank, d.bfin, d.beta, rc, d.ldbf};
}

// U_s[i, j] of the stacked left factor (m x K): C_U[i, j] for j < rc, alpha * Afin[i, j - rc]
// after, as ph_stack copies it, read in place from C's U factor (m x rc, ld m) and Afin (ld
// ldaf).  The QR of U_s reads it so in its first trailing update, which writes U_s's columns
// beyond the first panel into the workspace as it goes.
template <typename T>
struct StackedU {
    const T *cu, *af;
    T alpha;
    int rc, ldu;
    int64_t ldaf;
    // column j's first element, and the factor its elements take
    __device__ __forceinline__ const T *column(int j, T &scale) const {
        if (j < rc) {
            scale = T(1);
            return cu + (int64_t) j * ldu;
        }
        scale = alpha;
        return af + (int64_t) (j - rc) * ldaf;
    }
};

template <typename T>
__device__ inline StackedU<T> stacked_u(const RecompDesc<T> &d, int K) {
    return StackedU<T>{d.cdata, d.afin, d.alpha, K - d.arank, d.m, d.ldaf};
}

// Leading dimension of a Householder panel of `rows` rows staged in LDS: odd, so that the 16
// lanes reading one row across the panel's columns hit distinct LDS banks.
__host__ __device__ inline int panel_ld(int rows) { return rows | 1; }

// ---------------------------------------------------------------------------
// Compact-WY block reflector application with the panel held in registers.
// Rows are dealt to waves in 64-row chunks (chunk c -> wave c % NWAVES); each
// lane keeps its rows of the panel (RPT chunks x PB columns).  Columns of B
// are processed in batches of 64/PB: per batch, partial V^T b dot products are
// reduced within the wave (one multi-value reduction), across waves through
// LDS, multiplied by T^T (trans) or T, and the update b -= V t is applied.
// `red` is LDS scratch of NWAVES*64 + 2*64 elements.
// ---------------------------------------------------------------------------
template <typename T, int PB, int RPT>
__device__ void apply_wy(const T *P, int pld, const T *Tm, int rows, T *B, int64_t ldb, int c0, int c1, bool trans,
                         T *red) {
    constexpr int NCB = WAVE / PB;  // columns per batch
    const int l = lane_id(), wv = wave_id();
    T pv[RPT][PB];
#pragma unroll
    for (int t = 0; t < RPT; ++t) {
        const int r = l + WAVE * (wv + NWAVES * t);
#pragma unroll
        for (int k = 0; k < PB; ++k) pv[t][k] = (r < rows) ? P[r + k * pld] : T(0);
    }
    T *wfull = red + NWAVES * WAVE;
    T *tvec = wfull + WAVE;
    for (int cb = c0; cb < c1; cb += NCB) {
        T part[NCB * PB];
#pragma unroll
        for (int c = 0; c < NCB; ++c) {
            T bv[RPT];
#pragma unroll
            for (int t = 0; t < RPT; ++t) {
                const int r = l + WAVE * (wv + NWAVES * t);
                bv[t] = (r < rows && cb + c < c1) ? B[r + (int64_t) (cb + c) * ldb] : T(0);
            }
#pragma unroll
            for (int k = 0; k < PB; ++k) {
                T s = T(0);
#pragma unroll
                for (int t = 0; t < RPT; ++t) s += pv[t][k] * bv[t];
                part[c * PB + k] = s;
            }
        }
        const T tot = wave_multi_sum_fast<T, NCB * PB>(part);  // lane l holds index l
        red[wv * WAVE + l] = tot;
        __syncthreads();
        if (threadIdx.x < WAVE) {
            T s = T(0);
#pragma unroll
            for (int w2 = 0; w2 < NWAVES; ++w2) s += red[w2 * WAVE + threadIdx.x];
            wfull[threadIdx.x] = s;
        }
        __syncthreads();
        if (threadIdx.x < WAVE) {
            const int c = threadIdx.x / PB, k = threadIdx.x % PB;
            T s = T(0);
            if (trans) {
#pragma unroll
                for (int j = 0; j < PB; ++j)
                    if (j <= k) s += Tm[j + k * PB] * wfull[c * PB + j];
            } else {
#pragma unroll
                for (int j = 0; j < PB; ++j)
                    if (j >= k) s += Tm[k + j * PB] * wfull[c * PB + j];
            }
            tvec[threadIdx.x] = s;
        }
        __syncthreads();
        for (int c = 0; c < NCB; ++c) {
            if (cb + c >= c1) break;
            T tc[PB];
#pragma unroll
            for (int k = 0; k < PB; ++k) tc[k] = tvec[c * PB + k];
#pragma unroll
            for (int t = 0; t < RPT; ++t) {
                const int r = l + WAVE * (wv + NWAVES * t);
                if (r < rows) {
                    T s = T(0);
#pragma unroll
                    for (int k = 0; k < PB; ++k) s += pv[t][k] * tc[k];
                    B[r + (int64_t) (cb + c) * ldb] -= s;
                }
            }
        }
        __syncthreads();
    }
}

// Row-chunk count dispatch.  The variants are kept out of line so that each
// gets its own register allocation (inlining them all into one kernel makes
// the widest variant's pressure spill the others).
template <typename T, int PB, int RPT>
__device__ HCAMD_PHASE_ATTR void apply_wy_nl(const T *P, int pld, const T *Tm, int rows, T *B, int64_t ldb, int c0,
                                             int c1, bool trans, T *red) {
    apply_wy<T, PB, RPT>(P, pld, Tm, rows, B, ldb, c0, c1, trans, red);
}

// ---------------------------------------------------------------------------
// The same block-reflector application B[:, c0:c1] <- (I - P op(T) P^T) B for 16-column FP64
// panels, with all three products on the matrix cores (v_mfma_f64_16x16x4f64).  Each wave owns
// whole 16-column blocks of B, so the reduction over the rows happens inside the matrix core and
// no barrier or cross-wave exchange is needed:
//   W  = P^T B_blk    (16 x 16, K = rows: rows/4 MFMA steps in four interleaved chains)
//   W' = op(T) W      (4 MFMA steps: W's accumulator is already the B operand layout)
//   B_blk -= P W'     (per 16 rows: 4 MFMA steps of K = 4)
// MFMA lane layout (f64 16x16x4): A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k, and
// D[4 v + l / 16][l % 16] in element v of lane l (row groups of one, unlike the FP32 16x16x4
// instruction's groups of four).  The reduction over rows may visit them in any order, and the
// output rows of the update may be any permutation of the 16 rows of a group, so both products
// are arranged for memory: in W, lane (j, k) supplies rows r0 + 4 k .. r0 + 4 k + 3 of column j
// in steps 0..3; in the update, MFMA output row i is row r0 + 4 (i % 4) + i / 4, which puts
// rows r0 + 4 k .. r0 + 4 k + 3 of column j in lane (j, k).  Each lane thus reads and writes 4
// consecutive elements of its column per group.  P: panel in LDS (rows x 16, ld pld, unit
// lower); Tm: 16 x 16 upper (LDS, zero below the diagonal).  With `cn`, the squared norms of the
// updated columns over rows [nrm_lo, rows) go to cn[column] as the update writes them (the next
// panel's pivoting reads them instead of passing over the trailing matrix once more).  SRC: B's
// columns are read from the stacked left factor `src` (rows from 0: the first trailing update of
// the QR of U_s, B = U_s) and the result is written to B.
// ---------------------------------------------------------------------------
template <bool SRC = false>
__device__ HCAMD_PHASE_ATTR void apply_wy_mfma(const double *P, int pld, const double *Tm, int rows, double *B,
                                               int64_t ldb, int c0, int c1, bool trans, double *cn = nullptr,
                                               int nrm_lo = 0, StackedU<double> src = StackedU<double>{}) {
    const int l = lane_id(), i16 = l & 15, k4 = l >> 4;
    const int nblk = (c1 - c0 + 15) / 16;
    // A operand of W' = op(T) W, step s: op(T)[i16][4 s + k4]
    double ta[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int k = 4 * s + k4;
        ta[s] = trans ? (k <= i16 ? Tm[k + 16 * i16] : 0.0) : (i16 <= k ? Tm[i16 + 16 * k] : 0.0);
    }
    const int r64 = rows & ~63;
    const int prow = 4 * (i16 & 3) + (i16 >> 2);  // row of MFMA output row i16 within its group
    for (int blk = wave_id(); blk < nblk; blk += NWAVES) {
        const int col = c0 + 16 * blk + i16;
        const bool cok = col < c1;
        // lanes past c1 read column c0 (in range) and discard it
        double *bcol = B + (int64_t) (cok ? col : c0) * ldb;
        const double *rcol = bcol;
        double sc = 1.0;
        if constexpr (SRC) rcol = src.column(cok ? col : c0, sc);
        auto bload = [&](int r) -> double {
            if constexpr (SRC) return sc * rcol[r];
            else return bcol[r];
        };
        // W = P^T B_blk over 64-row chunks: the 16 loads of B and of P of a chunk are in flight
        // together, and four accumulators (one per 16-row group) keep four MFMA chains apart
        hcamd_doublex4 acc[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) acc[g] = hcamd_doublex4{0.0, 0.0, 0.0, 0.0};
        int r0 = 0;
        for (; r0 < r64; r0 += 64) {
            double bv[4][4], av[4][4];
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int s = 0; s < 4; ++s) bv[g][s] = bload(r0 + 16 * g + 4 * k4 + s);
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int s = 0; s < 4; ++s) av[g][s] = P[r0 + 16 * g + 4 * k4 + s + i16 * pld];
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int g = 0; g < 4; ++g)
                    acc[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[g][s], cok ? bv[g][s] : 0.0, acc[g], 0, 0, 0);
        }
        for (; r0 < rows; r0 += 16) {
            const int ra = r0 + 4 * k4;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int r = ra + s;
                const double a = r < rows ? P[r + i16 * pld] : 0.0;
                const double b = (r < rows && cok) ? bload(r) : 0.0;
                acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[0], 0, 0, 0);
            }
        }
        // w[v] = W[4 v + k4][i16]: element s is the B operand of step s of W' = op(T) W
        const hcamd_doublex4 w = (acc[0] + acc[1]) + (acc[2] + acc[3]);
        hcamd_doublex4 wp = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < 4; ++s) wp = __builtin_amdgcn_mfma_f64_16x16x4f64(ta[s], w[s], wp, 0, 0, 0);
        // wp[s] = W'[4 s + k4][i16], the B operand of step s of the update B_blk -= P W'; in
        // 64-row chunks the 16 loads of B and of P are in flight together
        double nsq = 0.0;
        int g0 = 0;
        for (; g0 < r64; g0 += 64) {
            double old[4][4], av[4][4];
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int v = 0; v < 4; ++v) old[g][v] = bload(g0 + 16 * g + 4 * k4 + v);
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int s = 0; s < 4; ++s) av[g][s] = P[g0 + 16 * g + prow + (4 * s + k4) * pld];
            hcamd_doublex4 d[4];
#pragma unroll
            for (int g = 0; g < 4; ++g) d[g] = hcamd_doublex4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int g = 0; g < 4; ++g) d[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[g][s], wp[s], d[g], 0, 0, 0);
            if (cok) {
#pragma unroll
                for (int g = 0; g < 4; ++g)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = g0 + 16 * g + 4 * k4 + v;
                        const double nv = old[g][v] - d[g][v];
                        bcol[r] = nv;
                        if (r >= nrm_lo) nsq += nv * nv;
                    }
            }
        }
        for (; g0 < rows; g0 += 16) {
            const int pa = g0 + prow, rq = g0 + 4 * k4;
            double old[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) old[v] = (rq + v < rows) ? bload(rq + v) : 0.0;
            hcamd_doublex4 d = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const double a = pa < rows ? P[pa + (4 * s + k4) * pld] : 0.0;
                d = __builtin_amdgcn_mfma_f64_16x16x4f64(a, wp[s], d, 0, 0, 0);
            }
            if (cok) {
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    if (rq + v < rows) {
                        const double nv = old[v] - d[v];
                        bcol[rq + v] = nv;
                        if (rq + v >= nrm_lo) nsq += nv * nv;
                    }
            }
        }
        if (cn) {
            // a column's rows sit in lanes i16 + 16 k4: sum over k4 (lane bits 4 and 5)
            double a = nsq, b = nsq;
            permlane_swap<32>(a, b);
            nsq = a + b;
            a = nsq;
            b = nsq;
            permlane_swap<16>(a, b);
            nsq = a + b;
            if (k4 == 0 && cok) cn[col] = nsq;
        }
    }
}

// `cn` (FP64 16-column panels only: apply_wy_mfma) receives the updated columns' squared norms
// over rows [nrm_lo, rows).
template <typename T, int PB>
__device__ void apply_wy_dispatch(const T *P, int pld, const T *Tm, int rows, T *B, int64_t ldb, int c0, int c1,
                                  bool trans, T *red, T *cn = nullptr, int nrm_lo = 0) {
    if constexpr (PB == 16 && sizeof(T) == 8) {
        apply_wy_mfma(P, pld, Tm, rows, B, ldb, c0, c1, trans, cn, nrm_lo);
        return;
    }
    (void) cn;
    (void) nrm_lo;
    // rows <= 512 for PB 16, <= 1024 for PB 8, <= 2048 for PB 4 (pick_pb): at most 32 panel values per lane
    const int chunks = (rows + WAVE * NWAVES - 1) / (WAVE * NWAVES);
    if (chunks <= 1) apply_wy_nl<T, PB, 1>(P, pld, Tm, rows, B, ldb, c0, c1, trans, red);
    else if (chunks <= 2) apply_wy_nl<T, PB, 2>(P, pld, Tm, rows, B, ldb, c0, c1, trans, red);
    else if constexpr (PB <= 8) {
        if (chunks <= 4) apply_wy_nl<T, PB, 4>(P, pld, Tm, rows, B, ldb, c0, c1, trans, red);
        else if constexpr (PB <= 4) apply_wy_nl<T, PB, 8>(P, pld, Tm, rows, B, ldb, c0, c1, trans, red);
    }
}

// Build T (PB x PB, upper, column-major in LDS) for the panel P with taus tau[0..jb):
// dlarft forward/columnwise: T[i][i] = tau_i, T[0:i, i] = -tau_i T[0:i,0:i] (V[:,0:i]^T v_i).
// For FP64 panels of 16 columns the Gram matrix P^T P is formed on the matrix cores: wave w sums
// the 16-row groups w, w + 4, ... (A = P^T and B = P take the same lane values), and waves 1..3
// pass their partial sums to wave 0 through the 3 x 256 elements from Tm on (the T factor, Gram
// and reduction areas, which follow one another in the panel workspace and are free here).
template <typename T, int PB>
__device__ void build_t(const T *P, int pld, int rows, const T *tau, int jb, T *Tm, T *gram) {
    const int l = lane_id();
    if constexpr (sizeof(T) == 8 && PB == 16) {
        const int i16 = l & 15, k4 = l >> 4, w = wave_id();
        hcamd_doublex4 acc = {0.0, 0.0, 0.0, 0.0};
        for (int r0 = 16 * w; r0 < rows; r0 += 16 * NWAVES) {
            double a[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int r = r0 + 4 * k4 + s;
                a[s] = r < rows ? P[r + i16 * pld] : 0.0;
            }
#pragma unroll
            for (int s = 0; s < 4; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], a[s], acc, 0, 0, 0);
        }
        // element v of lane l is G[4 v + k4][i16], kept at (4 v + k4) + 16 i16
        T *part = Tm;
        if (w > 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) part[(w - 1) * 256 + (4 * v + k4) + 16 * i16] = acc[v];
        }
        __syncthreads();
        if (w == 0) {
            double g[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int e = (4 * v + k4) + 16 * i16;
                g[v] = acc[v] + part[e] + part[256 + e] + part[512 + e];
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) gram[(4 * v + k4) + 16 * i16] = g[v];
        }
        __syncthreads();
    } else {
        for (int pr = wave_id(); pr < PB * PB; pr += NWAVES) {
            const int a = pr % PB, b = pr / PB;
            if (a >= b || b >= jb) continue;
            T s = T(0);
            for (int r = l; r < rows; r += WAVE) s += P[r + a * pld] * P[r + b * pld];
            s = wave_sum(s);
            if (l == 0) gram[a + b * PB] = s;
        }
        __syncthreads();
    }
    // Lane a < PB builds row a of T: T[a][i] = -tau_i sum_{b=a}^{i-1} T[a][b] gram[b][i] only
    // reads row a, so the rows are independent.
    if (threadIdx.x < PB) {
        const int a = threadIdx.x;
        T trow[PB];
#pragma unroll
        for (int b = 0; b < PB; ++b) trow[b] = T(0);
#pragma unroll
        for (int i = 0; i < PB; ++i) {
            if (i < jb) {
                const T ti = tau[i];
                T acc = T(0);
#pragma unroll
                for (int b = 0; b < i; ++b)
                    if (b >= a) acc += trow[b] * gram[b + i * PB];
                trow[i] = (a < i) ? -ti * acc : (a == i ? ti : T(0));
            }
        }
#pragma unroll
        for (int b = 0; b < PB; ++b) Tm[a + b * PB] = trow[b];
    }
    __syncthreads();
}

// Ordering key of a norm in the rank-by-comparison selections below (pivot choice, singular
// value order): NaN, which NaN or Inf in the input produces, ranks below every number, so that
// the positions counted from the keys always form a permutation of the indices.
template <typename T>
__device__ __forceinline__ T order_key(T v) {
    return v == v ? v : T(-INFINITY);
}

// LDS elements needed by wg_geqrf / wg_apply_q for a panel of PB columns over m rows.
__host__ __device__ inline int64_t qr_lds_elems(int64_t m, int pb) {
    return (m | 1) * pb + 2 * pb * pb + (NWAVES + 2) * WAVE + 64;
}

// Window pivoting of a blocked QR at panel start j0 (column norms from `cn` when given): the jb trailing columns of largest
// norm over rows [j0, m) are moved, in decreasing norm order, to positions [j0, j0 + jb)
// by jb column swaps (whole columns, so the R rows above j0 move with them); piv records
// the permutation.  Uses the panel buffer P as scratch (needs m * PB >= 2 * (ncols - j0) + 2 * PB)
// and, for the move list, the 4 PB + 1 ints after the panel ((m | 1) * PB elements on), where the
// panel's T factor goes later.
template <typename T, int PB>
__device__ T window_pivot(int m, int ncols, T *A, int64_t lda, int j0, int jb, int32_t *piv, T *P,
                          const T *cn = nullptr) {
    const int nt = ncols - j0;
    const int l = lane_id();
    int *mv = reinterpret_cast<int *>(P + (int64_t) (m | 1) * PB);
    T *nrm = P;
    int *at = reinterpret_cast<int *>(nrm + nt);  // at[p]: column (by starting position) now at p
    int *where = at + nt;                         // where[c]: current position of column c
    int *sel = where + nt;                        // sel[t]: t-th largest column; sel[PB + t]: swap source
    if (cn) {
        // squared norms over rows [j0, m) by current column position, left by the previous trailing update
        for (int c = threadIdx.x; c < nt; c += NT) nrm[c] = cn[j0 + c];
    } else {
        // four columns per wave at a time (their loads in flight together), one multi-sum
        constexpr int CW = 4;
        for (int cb = wave_id(); cb < nt; cb += CW * NWAVES) {
            T sum[CW];
#pragma unroll
            for (int u = 0; u < CW; ++u) sum[u] = T(0);
            for (int r = j0 + l; r < m; r += WAVE) {
#pragma unroll
                for (int u = 0; u < CW; ++u) {
                    const int c = cb + u * NWAVES;
                    if (c < nt) {
                        const T v = A[r + (int64_t) (j0 + c) * lda];
                        sum[u] += v * v;
                    }
                }
            }
            const T tot = wave_multi_sum_fast<T, CW>(sum);
            const int u = wave_multi_sum_index<CW>(l), c = cb + u * NWAVES;
            if (l == multi_sum_src_lane<CW>(u) && c < nt) nrm[c] = tot;
        }
    }
    for (int c = threadIdx.x; c < nt; c += NT) {
        at[c] = c;
        where[c] = c;
    }
    __syncthreads();
    // squared Frobenius norm of the trailing matrix A[j0:m, j0:ncols] (the stop test of ph_qr_y)
    T tpart = T(0);
    for (int c = threadIdx.x; c < nt; c += NT) tpart += nrm[c];
    const T trail2 = wg_sum(tpart, reinterpret_cast<T *>(sel + 2 * PB) + 1);
    __syncthreads();
    for (int c = threadIdx.x; c < nt; c += NT) {
        const T v = order_key(nrm[c]);
        int pos = 0;
        for (int k = 0; k < nt; ++k) {
            const T u = order_key(nrm[k]);
            pos += (u > v) || (u == v && k < c);
        }
        if (pos < jb) sel[pos] = c;
    }
    __syncthreads();
    // The jb swaps (t, where[sel[t]]) in turn define the permutation; afterwards at[q] is the
    // position (relative to j0) whose column ends at q.  Only thread 0 runs the swaps, on LDS.
    if (threadIdx.x == 0) {
        for (int t = 0; t < jb; ++t) {
            const int c = sel[t], p = where[c];
            sel[PB + t] = p;
            if (p != t) {
                const int c2 = at[t];
                at[t] = c;
                at[p] = c2;
                where[c] = t;
                where[c2] = p;
            }
        }
    }
    __syncthreads();
    // The positions that changed lie among 0..jb-1 and the swap sources sel[PB..PB+jb); wave 0
    // takes one candidate per lane (2 jb <= 64), keeps each changed position once, and writes the
    // move list mv (pairs destination, source; count in mv[4 PB]), so the columns then move in
    // one pass over the rows.  The pivot vector follows: piv[j0 + q] <- piv[j0 + at[q]] (every
    // lane's load completes before any lane's store, one instruction each).
    if (wave_id() == 0) {
        int q = 0;
        bool keep = false;
        if (l < 2 * jb) {
            q = l < jb ? l : sel[PB + l - jb];
            keep = at[q] != q;
            if (keep && l >= jb) {
                if (q < jb) keep = false;
                for (int k = 0; k < l - jb; ++k) keep = keep && sel[PB + k] != q;
            }
        }
        const unsigned long long bal = __ballot(keep);
        const int src = keep ? at[q] : 0;
        const int32_t pv = keep ? piv[j0 + src] : 0;
        if (keep) {
            const int k = __popcll(bal & ((1ull << l) - 1ull));
            mv[2 * k] = q;
            mv[2 * k + 1] = src;
            piv[j0 + q] = pv;
        }
        if (l == 0) mv[4 * PB] = __popcll(bal);
    }
    __syncthreads();
    const int nmv = mv[4 * PB];
    if (nmv > 0) {
        for (int r = threadIdx.x; r < m; r += NT) {
            T v[2 * PB];
#pragma unroll
            for (int i = 0; i < 2 * PB; ++i)
                if (i < nmv) v[i] = A[r + (int64_t) (j0 + mv[2 * i + 1]) * lda];
#pragma unroll
            for (int i = 0; i < 2 * PB; ++i)
                if (i < nmv) A[r + (int64_t) (j0 + mv[2 * i]) * lda] = v[i];
        }
    }
    __syncthreads();
    return trail2;
}

// Whether window pivoting fits the panel buffer of a QR of m x ncols with panel PB.
__host__ __device__ inline bool window_pivot_fits(int64_t m, int64_t ncols, int pb) {
    return m * pb >= 2 * ncols + 2 * pb + 8;
}

// ---------------------------------------------------------------------------
// Blocked Householder QR (LAPACK dgeqrf convention) of A (m x ncols, ld lda) in
// place; tau[0..min(m,ncols)); the T factor of panel p is saved to tsave[p*PB*PB].
// With piv non-null, each panel starts with window pivoting (window_pivot) and
// piv[0..ncols) receives the column permutation: A P = Q R.
// ---------------------------------------------------------------------------

// One panel of the blocked QR below: window pivoting (when piv is given), Householder
// factorisation of columns [j0, j0 + jb) over rows [j0, m), write-back, and the compact-WY T
// (saved to tsave[j0 / PB] when tsave is given).  Leaves the panel's unit-lower reflectors in
// LDS P (rows x PB, rows = m - j0) and T in LDS Tm; returns jb.
template <typename T, int PB>
__device__ int wg_geqrf_panel(int m, int ncols, T *A, int64_t lda, T *tau, T *tsave, T *lds, int32_t *piv,
                              int64_t *hdr, int j0, const T *cn = nullptr) {
    (void) hdr;
    T *P = lds;
    T *Tm = P + (int64_t) (m | 1) * PB;
    T *gram = Tm + PB * PB;
    T *red = gram + PB * PB;  // (NWAVES + 2) * WAVE
    const int nsteps = m < ncols ? m : ncols;
    const int l = lane_id();
    (void) l;
    const int jb = (nsteps - j0) < PB ? (nsteps - j0) : PB;
    const int rows = m - j0;
    const int pld = panel_ld(rows);
    HCAMD_SUB_BEGIN(t_pan);
    HCAMD_SUB_BEGIN(t_piv);
    if (piv) window_pivot<T, PB>(m, ncols, A, lda, j0, jb, piv, P, cn);
    HCAMD_SUB_END(t_piv, hdr, 5);
    // one row per thread and iteration, the PB loads of a row in flight together
    for (int r = threadIdx.x; r < rows; r += NT) {
        T v[PB];
#pragma unroll
        for (int c = 0; c < PB; ++c) v[c] = (c < jb) ? A[(j0 + r) + (int64_t) (j0 + c) * lda] : T(0);
#pragma unroll
        for (int c = 0; c < PB; ++c) P[r + c * pld] = v[c];
    }
    __syncthreads();
    // Householder steps with one pass over the rows and one barrier per column.  The pass of step
    // jj applies reflector jj to the rows below it and, on the updated values, accumulates column
    // jj+1's products with itself and with the later panel columns: reduced, they give the next
    // column's norm and the dot products of its (still unscaled) reflector with those columns.
    // Every lane then forms the reflector and the update coefficients itself (lane k holds the
    // sums of value k; readlane spreads them).  The reduction slots and the copy of the next
    // pivot row (`xrow`, written by that row's owner) alternate between two buffers by step
    // parity, so one barrier per step separates their writers and readers.
    const int kl = l & (PB - 1);
    auto reduce_next = [&](T (&acc)[PB], int buf) {
        const T s = wave_multi_sum_fast<T, PB>(acc);
        const int idx = wave_multi_sum_index<PB>(l);
        if (l == multi_sum_src_lane<PB>(idx)) red[buf * (NWAVES * PB) + wave_id() * PB + idx] = s;
    };
    T *xrow = red + 2 * NWAVES * PB;  // [2][PB]
    {
        T acc[PB];
#pragma unroll
        for (int k = 0; k < PB; ++k) acc[k] = T(0);
        for (int r = 1 + threadIdx.x; r < rows; r += NT) {
            const T x = P[r];
#pragma unroll
            for (int k = 0; k < PB; ++k)
                if (k < jb) acc[k] += x * P[r + k * pld];
        }
        reduce_next(acc, 0);
        if (threadIdx.x < PB) xrow[threadIdx.x] = P[threadIdx.x * pld];
        __syncthreads();
    }
    for (int jj = 0; jj < jb; ++jj) {
        const int buf = jj & 1;
        T sk = T(0);
#pragma unroll
        for (int wv = 0; wv < NWAVES; ++wv) sk += red[buf * (NWAVES * PB) + wv * PB + kl];
        const T xk = xrow[buf * PB + kl];
        const T xnorm2 = read_lane(sk, jj), alpha = read_lane(xk, jj);
        T tau_j, beta, scal;
        if (xnorm2 == T(0)) {
            tau_j = T(0);
            beta = alpha;
            scal = T(0);
        } else {
            beta = -dsign(sqrt(alpha * alpha + xnorm2), alpha);
            tau_j = (beta - alpha) / beta;
            scal = T(1) / (alpha - beta);
        }
        // w_k = x_k[jj] + sum_{r > jj} v_r x_k[r] with v_r = scal x_jj[r]; update x_k -= tau w_k v
        const T gk = (kl > jj && kl < jb) ? tau_j * (xk + scal * sk) : T(0);
        T g[PB];
#pragma unroll
        for (int k = 0; k < PB; ++k) g[k] = read_lane(gk, k);
        if (threadIdx.x == 0) {
            P[jj + jj * pld] = beta;
            tau[j0 + jj] = tau_j;
#pragma unroll
            for (int k = 0; k < PB; ++k)
                if (k > jj && k < jb) P[jj + k * pld] -= g[k];
        }
        const int nj = jj + 1;
        const bool more = nj < jb;
        T acc[PB];
#pragma unroll
        for (int k = 0; k < PB; ++k) acc[k] = T(0);
        for (int r = nj + threadIdx.x; r < rows; r += NT) {
            T x[PB];
#pragma unroll
            for (int k = 0; k < PB; ++k) x[k] = P[r + k * pld];
            if (tau_j != T(0)) {
                const T v = scal * P[r + jj * pld];
                P[r + jj * pld] = v;
#pragma unroll
                for (int k = 0; k < PB; ++k)
                    if (k > jj && k < jb) {
                        x[k] -= g[k] * v;
                        P[r + k * pld] = x[k];
                    }
            }
            if (more) {
                if (r == nj) {
#pragma unroll
                    for (int k = 0; k < PB; ++k) xrow[(buf ^ 1) * PB + k] = x[k];
                } else {
                    T xn = T(0);  // x[nj] without a register-indexed read
#pragma unroll
                    for (int k = 0; k < PB; ++k)
                        if (k == nj) xn = x[k];
#pragma unroll
                    for (int k = 0; k < PB; ++k)
                        if (k >= nj && k < jb) acc[k] += xn * x[k];
                }
            }
        }
        if (more) reduce_next(acc, buf ^ 1);
        __syncthreads();
    }
    HCAMD_SUB_END(t_pan, hdr, 0);
    // write the factored panel back (R on/above the diagonal, reflectors below), a row per thread
    for (int r = threadIdx.x; r < rows; r += NT) {
        T v[PB];
#pragma unroll
        for (int c = 0; c < PB; ++c) v[c] = P[r + c * pld];
#pragma unroll
        for (int c = 0; c < PB; ++c)
            if (c < jb) A[(j0 + r) + (int64_t) (j0 + c) * lda] = v[c];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < PB; ++c)
        if (c < jb)
            for (int r = threadIdx.x; r <= c && r < rows; r += NT) P[r + c * pld] = (r == c) ? T(1) : T(0);
    __syncthreads();
    HCAMD_SUB_BEGIN(t_bt);
    build_t<T, PB>(P, pld, rows, tau + j0, jb, Tm, gram);
    if (tsave) {
        for (int e = threadIdx.x; e < PB * PB; e += NT) tsave[(int64_t) (j0 / PB) * PB * PB + e] = Tm[e];
    }
    HCAMD_SUB_END(t_bt, hdr, 1);
    return jb;
}

// The whole blocked QR in one workgroup: panels and trailing updates in turn.  from_src (FP64
// 16-column panels only, the QR of U_s in the batched form): A holds only the first 16 columns of
// the stacked left factor `src`; the first trailing update reads the others in place and writes
// them into A.
template <typename T, int PB>
__device__ void wg_geqrf(int m, int ncols, T *A, int64_t lda, T *tau, T *tsave, T *lds, int32_t *piv = nullptr,
                         int64_t *hdr = nullptr, bool from_src = false, StackedU<T> src = StackedU<T>{}) {
    T *P = lds;
    T *Tm = P + (int64_t) (m | 1) * PB;
    T *red = Tm + 2 * PB * PB;
    const int nsteps = m < ncols ? m : ncols;
    if (piv) {
        for (int c = threadIdx.x; c < ncols; c += NT) piv[c] = c;
        __syncthreads();
    }
    for (int j0 = 0; j0 < nsteps; j0 += PB) {
        const bool first_src = from_src && j0 == 0;
        const int jb = wg_geqrf_panel<T, PB>(m, ncols, A, lda, tau, tsave, lds, piv, hdr, j0);
        HCAMD_SUB_BEGIN(t_up);
        if (j0 + jb < ncols) {
            if constexpr (PB == 16 && sizeof(T) == 8) {
                if (first_src)
                    apply_wy_mfma<true>(P, panel_ld(m), Tm, m, A, lda, jb, ncols, true, nullptr, 0, src);
                else
                    apply_wy_dispatch<T, PB>(P, panel_ld(m - j0), Tm, m - j0, A + j0, lda, j0 + jb, ncols, true, red);
            } else {
                apply_wy_dispatch<T, PB>(P, panel_ld(m - j0), Tm, m - j0, A + j0, lda, j0 + jb, ncols, true, red);
            }
        }
        __syncthreads();
        HCAMD_SUB_END(t_up, hdr, 2);
    }
}

// B <- Q B with Q = H_1 ... H_nref the reflectors of a factored matrix A (m rows, ld lda),
// applied to columns [c0, c1) of B (m rows, ld ldb).
template <typename T, int PB>
__device__ void wg_apply_q(int m, int nref, const T *A, int64_t lda, const T *tsave, T *B, int64_t ldb, int c0,
                           int c1, T *lds) {
    T *P = lds;
    T *Tm = P + (int64_t) (m | 1) * PB;
    T *red = Tm + 2 * PB * PB;
    const int npan = (nref + PB - 1) / PB;
    for (int p = npan - 1; p >= 0; --p) {
        const int j0 = p * PB;
        const int jb = (nref - j0) < PB ? (nref - j0) : PB;
        const int rows = m - j0;
        const int pld = panel_ld(rows);
        // unit-lower reflectors: one row per thread and iteration, its PB loads in flight together
        for (int r = threadIdx.x; r < rows; r += NT) {
            T v[PB];
#pragma unroll
            for (int c = 0; c < PB; ++c)
                v[c] = (c < jb && r > c) ? A[(j0 + r) + (int64_t) (j0 + c) * lda] : T(c < jb && r == c ? 1 : 0);
#pragma unroll
            for (int c = 0; c < PB; ++c) P[r + c * pld] = v[c];
        }
        for (int e = threadIdx.x; e < PB * PB; e += NT) Tm[e] = tsave[(int64_t) p * PB * PB + e];
        __syncthreads();
        apply_wy_dispatch<T, PB>(P, pld, Tm, rows, B + j0, ldb, c0, c1, false, red);
        __syncthreads();
    }
}

// Squared threshold below which the trailing part of the pivoted QR of Y may be left unfactored:
// the rows of R_Y it would produce are the ones the core truncation drops anyway -- at the
// rounding level eps ||Y||_F (when the rank threshold is far above it, as in ph_core_size), or at
// the tolerance rule's tol * threshold.  Negative: factor everything (fixed rank, learned sizing
// beyond the eps level, thresholds near rounding).
template <typename T>
__device__ T qr_y_stop2(const RecompDesc<T> &d, T fro2, int kk) {
    if (d.fixed_rank > 0 || kk <= 0) return T(-1);
    const T delta = Eps<T>::value * sqrt(fro2);
    const T thr_lo = d.truncated_svd ? T(d.accuracy) * sqrt(fro2 / T(kk)) : T(d.accuracy);
    T stop2 = (thr_lo > T(64) * delta) ? delta * delta : T(-1);
    if (!d.rank_model && d.model_tol > 0.0) {
        const T lim = T(d.model_tol) * thr_lo;
        if (lim * lim > stop2) stop2 = lim * lim;
    }
    return stop2;
}

// Phase: window-pivoted QR of Y (n x mu); piv is the identity when the panel buffer is too
// small for pivoting (very wide Y).  With pivoting the factorisation stops at the first panel
// whose trailing matrix is below qr_y_stop2 (rows of R_Y computed: HDR_KQ, the unfactored norm:
// HDR_TRAIL2); its Q_Y reflectors beyond that point would only act on rows the core drops.
template <typename T, int PB>
__device__ void ph_qr_y(const RecompDesc<T> &d, T *lds) {
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    const bool pivot = window_pivot_fits(d.n, z.mu, PB);
    if (!pivot) {
        for (int c = threadIdx.x; c < z.mu; c += NT) w.piv[c] = c;
        __syncthreads();
        wg_geqrf<T, PB>(d.n, z.mu, w.y, d.n, w.tau_y, w.t_y, lds, nullptr, w.hdr);
        return;
    }
    const int m = d.n, ncols = z.mu;
    const int nsteps = m < ncols ? m : ncols;
    T *P = lds;
    T *Tm = P + (int64_t) (m | 1) * PB;
    T *red = Tm + 2 * PB * PB;
    for (int c = threadIdx.x; c < ncols; c += NT) w.piv[c] = c;
    __syncthreads();
    // with FP64 16-column panels the trailing update leaves the next panel's column norms in cn
    // (the core buffer, unused until the core is built); otherwise each panel sums them itself
    constexpr bool fused_norms = PB == 16 && sizeof(T) == 8;
    T *cn = w.j;
    T stop2 = T(-1);
    for (int j0 = 0; j0 < nsteps; j0 += PB) {
        const int jb = (nsteps - j0) < PB ? (nsteps - j0) : PB;
        const T trail2 = window_pivot<T, PB>(m, ncols, w.y, m, j0, jb, w.piv, P, (fused_norms && j0 > 0) ? cn : nullptr);
        if (j0 == 0) {
            stop2 = qr_y_stop2<T>(d, trail2, z.kk);
        } else if (trail2 <= stop2) {
            if (threadIdx.x == 0) {
                w.hdr[HDR_KQ] = j0;
                w.hdr[HDR_TRAIL2] = __double_as_longlong((double) trail2);
            }
            __syncthreads();
            return;
        }
        wg_geqrf_panel<T, PB>(m, ncols, w.y, m, w.tau_y, w.t_y, lds, nullptr, w.hdr, j0);
        if (j0 + jb < ncols)
            apply_wy_dispatch<T, PB>(P, panel_ld(m - j0), Tm, m - j0, w.y + j0, m, j0 + jb, ncols, true, red,
                                     fused_norms ? cn : nullptr, jb);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Phase: Y = V_s R_U^T  (n x mu),  R_U = triu(U_s[0:mu, 0:K]).
// Output tiles are dealt to workgroups `wg` of `nwg`.
// ---------------------------------------------------------------------------
template <typename T>
__device__ void ph_y(const RecompDesc<T> &d, int wg, int nwg, T *lds) {
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    const int tm = (d.n + GT_M - 1) / GT_M, tn = (z.mu + GT_N - 1) / GT_N;
    const StackedV<T> fa = stacked_v(d, z.K);  // read K-fastest: consecutive addresses in C_V and Bfin
    const T *us = w.us;
    const int64_t ldu = d.m;
    auto fb = [=](int k, int j) -> T { return k < j ? T(0) : us[j + (int64_t) k * ldu]; };  // R_U^T
    T *y = w.y;
    const int64_t ldy = d.n;
    auto epi = [=](int gi, int gj, T v) { y[gi + (int64_t) gj * ldy] = v; };
    // R_U^T(k, j) vanishes for k < j: the tile's K range starts at its first column
    for (int t = wg; t < tm * tn; t += nwg)
        wg_gemm_tile_core<T>(d.n, z.mu, z.K, fa, fb, /*a_rows_fast=*/false, /*b_cols_fast=*/true, epi, t % tm,
                             t / tm, lds, (t / tm) * GT_N);
}

// Phase (single workgroup): truncate the core and set up the Jacobi SVD.
// R_Y (kk x mu, upper, in the leading rows of Y) gives the core J = R_Y^T (mu x kk).  Rows
// j >= kt of R_Y whose total Frobenius norm is at most eps * ||R_Y||_F are dropped: the
// singular values of J move by at most that much, i.e. by the rounding level of the
// factorisation itself.  No truncation when the requested threshold is within a small factor
// of that level, and never below a requested fixed rank.
// ph_core_size writes kt to the header and leaves in lds: rn[j] = ||R_Y[j:, :]||_F^2 (kk values)
// followed by a lower bound of the rank threshold; ph_core_build then builds
// J = P R_Y[0:kt, :]^T (mu x kt, ld mu; P the pivoting of the QR of Y) and V1 = I (kt x kt).  `lds` holds >= kk + NWAVES values.
template <typename T>
__device__ void ph_core_size(const RecompDesc<T> &d, T *lds) {
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    const int mu = z.mu, kk = z.kk, n = d.n, l = lane_id();
    // rows of R_Y the QR of Y computed (all kk unless it stopped early) and the norm it left
    const int kq = (w.hdr[HDR_KQ] > 0 && w.hdr[HDR_KQ] < kk) ? (int) w.hdr[HDR_KQ] : kk;
    const T trail2 = kq < kk ? (T) __longlong_as_double(w.hdr[HDR_TRAIL2]) : T(0);
    T *rn = lds;  // suffix sums, then the threshold bound at rn[kk]
    for (int j = wave_id(); j < kq; j += NWAVES) {
        T sum = T(0);
        for (int c = j + l; c < mu; c += WAVE) {
            const T v = w.y[j + (int64_t) c * n];
            sum += v * v;
        }
        sum = wave_sum(sum);
        if (l == 0) rn[j] = sum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int j = kq; j < kk; ++j) rn[j] = trail2;  // rows never formed: bounded by what was left
        T acc = trail2;
        for (int j = kq - 1; j >= 0; --j) {
            acc += rn[j];
            rn[j] = acc;  // ||R_Y[j:, :]||_F^2
        }
        const T fro2 = kk > 0 ? rn[0] : T(0);
        const T delta = Eps<T>::value * sqrt(fro2);
        int kt = kk;
        // threshold the rank rule will use, bounded below (sigma_0 >= ||J||_F / sqrt(kk))
        const T thr_lo = d.truncated_svd ? T(d.accuracy) * sqrt(fro2 / T(kk > 0 ? kk : 1)) : T(d.accuracy);
        if (thr_lo > T(64) * delta) {
            for (int j = 1; j < kk; ++j) {
                if (sqrt(rn[j]) <= delta) {
                    kt = j;
                    break;
                }
            }
        }
        if (d.fixed_rank > 0) {
            const int fr = d.fixed_rank < kk ? d.fixed_rank : kk;
            if (kt < fr) kt = fr;
        }
        if (kt > kq) kt = kq;  // only kq rows of R_Y exist (the rest were left below the stop test)
        if (kt < 1) kt = kk < 1 ? 0 : 1;
        w.hdr[HDR_KT] = kt;
        w.hdr[HDR_CQR] = 0;
        rn[kk] = thr_lo;
    }
    __syncthreads();
}

// WITH_V: also set V1 = I for the Jacobi that accumulates the right rotations (the batched
// form's Jacobi keeps no V and skips it).
template <typename T, bool WITH_V = true>
__device__ void ph_core_build(const RecompDesc<T> &d) {
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    const int mu = z.mu, kt = z.kt, n = d.n;
    // J[piv[i], j] = R_Y[j, i]: rows of J in the order of Q_U's basis (undoing the column pivoting
    // of the QR of Y), so the left factor needs no permutation in the reconstruction
    if (mu > 0)
        for (ColWalk p(threadIdx.x, NT, mu); p.c < kt; p.next())
            w.j[w.piv[p.i] + (int64_t) p.c * mu] = (p.i >= p.c) ? w.y[p.c + (int64_t) p.i * n] : T(0);
    if (WITH_V && kt > 0)
        for (ColWalk p(threadIdx.x, NT, kt); p.c < kt; p.next())
            w.v1[p.i + (int64_t) p.c * kt] = (p.i == p.c) ? T(1) : T(0);
    __syncthreads();
}

template <typename T>
__device__ void ph_core(const RecompDesc<T> &d, T *lds) {
    ph_core_size<T>(d, lds);
    ph_core_build<T>(d);
}

// Second QR of a tall core (batched form): J (mu x kt) = Q2 R2, after which the Jacobi runs on the
// kt x kt triangle R2 (copied to j2 with zeros below the diagonal) -- its rows, flops and passes
// over memory shrink by mu / kt, and the extra QR preconditions the Jacobi further.  J keeps Q2's
// reflectors (T factors in t_j) for the reconstruction.  Skipped for cores that are not at least
// a quarter taller than wide.  `lds` holds qr_lds_elems(mu, PB) elements.
constexpr int CORE_QR_MIN_COLS = 16;
template <typename T, int PB>
__device__ void ph_core_qr(const RecompDesc<T> &d, T *lds) {
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    const int mu = z.mu, kt = z.kt;
    if (kt < CORE_QR_MIN_COLS || 4 * mu < 5 * kt) return;
    wg_geqrf<T, PB>(mu, kt, w.j, mu, w.tau_j, w.t_j, lds);
    __syncthreads();
    for (ColWalk p(threadIdx.x, NT, kt); p.c < kt; p.next())
        w.j2[p.i + (int64_t) p.c * kt] = (p.i <= p.c) ? w.j[p.i + (int64_t) p.c * mu] : T(0);
    if (threadIdx.x == 0) w.hdr[HDR_CQR] = 1;
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Phase: Jacobi + sigma + sort + rank (single workgroup).
// ---------------------------------------------------------------------------
template <typename T, int RPL, int RPLV, int TW>
__device__ HCAMD_PHASE_ATTR int block_jacobi_nl(T *J, int rows, int ncols, T *V, T *gbuf) {
    return wg_block_jacobi<T, RPL, RPLV, TW>(J, rows, ncols, rows, V, ncols, JACOBI_MAX_SWEEPS, gbuf);
}

template <typename T, int RPL, int RPLV, int TW>
__device__ HCAMD_PHASE_ATTR int coop_jacobi_nl(T *J, int rows, int ncols, T *V, T *gbuf) {
    return wg_coop_jacobi<T, RPL, RPLV, TW>(J, rows, ncols, rows, V, ncols, JACOBI_MAX_SWEEPS, gbuf);
}

// Tile shape by core rows and (truncated) width.  Per-wave tiles (a wave owns a whole tile of J
// and V columns; four tiles in flight per workgroup) with shapes that fit 256 VGPRs without
// spilling; for 193-256 rows the four waves share 16-column tiles instead (rows split across
// waves, one barrier per rotation round), which measured faster there because it needs half the
// passes over J and V of the widest per-wave tile that fits.  `exact_core` (a core sized at the
// rounding level only) also takes the shared tiles for 129-192 rows: 7 % faster on the nb=512
// benchmark batch, while cores cut by a rank tolerance ran 8 % slower with them.
template <typename T>
__device__ int run_block_jacobi(T *J, int rows, int ncols, T *V, T *gbuf, bool exact_core) {
    if (rows <= 64) return block_jacobi_nl<T, 1, 1, 16>(J, rows, ncols, V, gbuf);
    if (rows <= 128)
        return ncols <= 64 ? block_jacobi_nl<T, 2, 1, 8>(J, rows, ncols, V, gbuf)
                           : block_jacobi_nl<T, 2, 2, 8>(J, rows, ncols, V, gbuf);
    if (rows <= 192) {
        if (exact_core) return coop_jacobi_nl<T, 1, 1, 16>(J, rows, ncols, V, gbuf);
        return ncols <= 128 ? block_jacobi_nl<T, 3, 2, 8>(J, rows, ncols, V, gbuf)
                            : block_jacobi_nl<T, 3, 3, 4>(J, rows, ncols, V, gbuf);
    }
    if (rows <= 256) return coop_jacobi_nl<T, 1, 1, 16>(J, rows, ncols, V, gbuf);
    if (rows <= 320)
        return ncols <= 192 ? block_jacobi_nl<T, 5, 3, 4>(J, rows, ncols, V, gbuf)
                            : block_jacobi_nl<T, 5, 5, 4>(J, rows, ncols, V, gbuf);
    if (rows <= 384) {
        if (ncols <= 192) return block_jacobi_nl<T, 6, 3, 4>(J, rows, ncols, V, gbuf);
        if (ncols <= 256) return block_jacobi_nl<T, 6, 4, 4>(J, rows, ncols, V, gbuf);
        return block_jacobi_nl<T, 6, 6, 4>(J, rows, ncols, V, gbuf);
    }
    if (rows <= 512)
        return ncols <= 256 ? block_jacobi_nl<T, 8, 4, 4>(J, rows, ncols, V, gbuf)
                            : block_jacobi_nl<T, 8, 8, 2>(J, rows, ncols, V, gbuf);
    if (rows <= 640)
        return ncols <= 320 ? block_jacobi_nl<T, 10, 5, 2>(J, rows, ncols, V, gbuf)
                            : block_jacobi_nl<T, 10, 10, 2>(J, rows, ncols, V, gbuf);
    return ncols <= 384 ? block_jacobi_nl<T, 12, 6, 2>(J, rows, ncols, V, gbuf)
                        : block_jacobi_nl<T, 12, 12, 2>(J, rows, ncols, V, gbuf);
}

constexpr int JACOBI_ROWS_MAX = 768;
// LDS of the core-SVD phase: ph_core's row norms (<= JACOBI_ROWS_MAX + NWAVES) and the block Jacobi's
// reduction slots, which reuse the same space.
constexpr int JACOBI_LDS_ELEMS = JACOBI_ROWS_MAX + NWAVES + 8;
static_assert(JACOBI_LDS_ELEMS >= block_jacobi_lds_elems<double, 16>() &&
                      JACOBI_LDS_ELEMS >= coop_jacobi_lds_elems<double, 16>(),
              "core-SVD LDS covers both Jacobi variants");
static_assert(JACOBI_LDS_ELEMS >= JACOBI_ROWS_MAX + NWAVES, "ph_core scratch must fit the core-SVD LDS");

// Tail of the core SVD once the Jacobi has converged (`sweeps` used): singular values, their
// descending order, the rank rule and the header words.
template <typename T>
__device__ void ph_svd_finish(const RecompDesc<T> &d, int sweeps) {
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    const int mu = z.mu, kk = z.kk, kt = z.kt, K = z.K;
    __shared__ int s_nan;
    if (threadIdx.x == 0) s_nan = 0;
    __syncthreads();
    const int l = lane_id();
    // the orthogonalised core: J (mu rows) or, after the core QR, R2 (kt rows)
    const bool cqr = w.hdr[HDR_CQR] != 0;
    const T *zc = cqr ? w.j2 : w.j;
    const int zrows = cqr ? kt : mu;
    for (int j = wave_id(); j < kt; j += NWAVES) {
        const T *zj = zc + (int64_t) j * zrows;
        T s = 0;
        for (int r = l; r < zrows; r += WAVE) s += zj[r] * zj[r];
        s = wave_sum(s);
        if (l == 0) {
            w.sig[j] = sqrt(s);
            if (!(s == s) || s > T(1e300)) atomicOr(&s_nan, 1);
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < kt; j += NT) {
        const T sj = order_key(w.sig[j]);
        int pos = 0;
        for (int k = 0; k < kt; ++k) {
            const T sk = order_key(w.sig[k]);
            pos += (sk > sj) || (sk == sj && k < j);
        }
        w.perm[pos] = j;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int size_s = kk;
        int64_t rk;
        if (d.fixed_rank > 0) {
            rk = d.fixed_rank;
            if (rk > K) rk = K;
            if (rk > size_s) rk = size_s;
        } else {
            // singular values beyond the truncated core width kt are below eps * ||J||_F: zero
            rk = size_s;
            const T s0 = kt > 0 ? w.sig[w.perm[0]] : T(0);
            const T thr = d.truncated_svd ? T(d.accuracy) * s0 : T(d.accuracy);
            for (int i = 1; i < size_s; ++i) {
                const T si = i < kt ? w.sig[w.perm[i]] : T(0);
                if (si < thr) {
                    rk = i;
                    break;
                }
            }
        }
        if (rk < 1) rk = 1;
        if (rk > d.max_rank) rk = d.max_rank;
        if (rk > size_s && size_s > 0) rk = size_s;
        if (rk > kt && kt > 0) rk = kt;
        w.hdr[HDR_RANK] = rk;
        w.hdr[HDR_SIZES] = size_s;
        w.hdr[HDR_SWEEPS] = sweeps;
        int st = 0;
        if (sweeps >= JACOBI_MAX_SWEEPS) st |= STATUS_NOCONV;
        if (s_nan) st |= STATUS_NAN;
        w.hdr[HDR_STATUS] |= st;
        if (d.sigma_out) {
            for (int i = 0; i < size_s; ++i) d.sigma_out[i] = i < kt ? (double) w.sig[w.perm[i]] : 0.0;
        }
    }
    __syncthreads();
}

template <typename T>
__device__ void ph_svd(const RecompDesc<T> &d, T *lds) {
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    int sweeps = 0;
    if (z.mu > JACOBI_ROWS_MAX) {
        if (threadIdx.x == 0) w.hdr[HDR_STATUS] |= STATUS_TOO_BIG;
    } else {
        HCAMD_SUB_BEGIN(t_jac);
        const bool exact_core = !d.rank_model && !(d.model_tol > 0.0);
        if (w.hdr[HDR_CQR]) sweeps = run_block_jacobi<T>(w.j2, z.kt, z.kt, w.v1, lds, exact_core);
        else sweeps = run_block_jacobi<T>(w.j, z.mu, z.kt, w.v1, lds, exact_core);
        HCAMD_SUB_END(t_jac, w.hdr, 3);
#ifdef HCAMD_PHASE_TIMING
        if (threadIdx.x == 0) w.hdr[HDR_SUB + 7] = sweeps + 1000 * z.kt;  // timing build: sweeps and core width
#endif
    }
    ph_svd_finish<T>(d, sweeps);
}

// ---------------------------------------------------------------------------
// Batched form: the core Jacobi without V.  The rotations only orthogonalise the columns of the
// core X (X V1 = Z); the right factor is recovered afterwards from the left one: X = Z V1^T
// gives X^T U = V1 Sigma for U = Z Sigma^-1 (ph_recon_nov).  Dropping V halves the Jacobi's data
// and frees the registers that let the tiles double their width (half the passes over X per
// sweep).  A core of at most BATCH_LDS_BYTES / sizeof(T) - NOV_GBUF elements is moved into the
// kernel's dynamic LDS for the whole iteration, so its sweeps read no global memory.
// ---------------------------------------------------------------------------
constexpr int64_t BATCH_LDS_BYTES = 78 * 1024;  // dynamic LDS per workgroup of the batched kernel
constexpr int NOV_GBUF = 16;                     // reduction slots ahead of the LDS-resident core
static_assert(NOV_GBUF >= block_jacobi_lds_elems<double, 16>(), "the block Jacobi's slots fit before the core");

template <typename T>
__host__ __device__ constexpr int64_t nov_lds_cap() {
    return BATCH_LDS_BYTES / (int64_t) sizeof(T) - NOV_GBUF;
}

// LDS-resident core: J (rows x ncols, ld rows) at element NOV_GBUF of the dynamic LDS, the
// reduction slots before it.  The pointers are formed here from the dynamic LDS symbol so that
// every access of the iteration is an LDS instruction.
template <typename T, int RPL, int TW>
__device__ HCAMD_PHASE_ATTR int jacobi_nov_lds_nl(int rows, int ncols) {
    extern __shared__ unsigned char smem_raw[];
    T *gbuf = reinterpret_cast<T *>(smem_raw);
    return wg_block_jacobi<T, RPL, 0, TW>(gbuf + NOV_GBUF, rows, ncols, rows, nullptr, 0, JACOBI_MAX_SWEEPS, gbuf);
}

// Core in global memory: for up to 128 rows the next tile is loaded while the current one
// rotates (measured on the benchmark batch: 6 % less Jacobi time for 100-128-column cores;
// slower for taller tiles, whose two register tiles no longer fit without spills).
template <typename T, int RPL, int TW>
__device__ HCAMD_PHASE_ATTR int jacobi_nov_nl(T *J, int rows, int ncols, T *gbuf) {
    return wg_block_jacobi<T, RPL, 0, TW, (RPL <= 2)>(J, rows, ncols, rows, nullptr, 0, JACOBI_MAX_SWEEPS, gbuf);
}

// Core in global memory staged through LDS by super-blocks of W columns (wg_staged_jacobi): the
// buffer after the reduction slots holds two super-blocks.
template <typename T, int RPL, int TW>
__device__ HCAMD_PHASE_ATTR int jacobi_nov_staged_nl(T *J, int rows, int ncols, int W) {
    extern __shared__ unsigned char smem_raw[];
    T *gbuf = reinterpret_cast<T *>(smem_raw);
    return wg_staged_jacobi<T, RPL, TW>(J, rows, ncols, gbuf + NOV_GBUF, W, JACOBI_MAX_SWEEPS, gbuf);
}

template <typename T, int RPL, int TW>
__device__ int jacobi_nov_shape(T *J, int rows, int ncols, T *lds, bool resident) {
    if (resident) return jacobi_nov_lds_nl<T, RPL, TW>(rows, ncols);
    const int W = staged_width(rows, TW / 2, nov_lds_cap<T>());
    return W > 0 ? jacobi_nov_staged_nl<T, RPL, TW>(J, rows, ncols, W) : jacobi_nov_nl<T, RPL, TW>(J, rows, ncols, lds);
}

// One-sided Jacobi of J (rows x ncols, ld rows) in place, without V; `lds` is the kernel's
// dynamic LDS (BATCH_LDS_BYTES).  Returns the sweeps used.
template <typename T>
__device__ int run_jacobi_nov(T *J, int rows, int ncols, T *lds) {
    const int64_t ne = (int64_t) rows * ncols;
    const bool resident = ne <= nov_lds_cap<T>();
    T *Jl = lds + NOV_GBUF;
    if (resident) {
        wg_copy(Jl, J, ne);
        __syncthreads();
    }
    int sweeps;
    if (rows <= 64) sweeps = jacobi_nov_shape<T, 1, 16>(J, rows, ncols, lds, resident);
    else if (rows <= 128) sweeps = jacobi_nov_shape<T, 2, 16>(J, rows, ncols, lds, resident);
    else if (rows <= 192) sweeps = jacobi_nov_shape<T, 3, 8>(J, rows, ncols, lds, resident);
    else if (rows <= 256) sweeps = jacobi_nov_shape<T, 4, 8>(J, rows, ncols, lds, resident);
    else if (rows <= 384) sweeps = jacobi_nov_shape<T, 6, 8>(J, rows, ncols, lds, resident);
    else if (rows <= 512) sweeps = jacobi_nov_shape<T, 8, 4>(J, rows, ncols, lds, resident);
    else sweeps = jacobi_nov_shape<T, 12, 4>(J, rows, ncols, lds, resident);
    if (resident) {
        __syncthreads();
        wg_copy(J, Jl, ne);
        __syncthreads();
    }
    return sweeps;
}

// Core SVD of the batched form: the Jacobi without V, then sigma, order and rank (ph_svd_finish).
template <typename T>
__device__ void ph_svd_nov(const RecompDesc<T> &d, T *lds) {
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    int sweeps = 0;
    if (z.mu > JACOBI_ROWS_MAX) {
        if (threadIdx.x == 0) w.hdr[HDR_STATUS] |= STATUS_TOO_BIG;
    } else {
        HCAMD_SUB_BEGIN(t_jac);
        if (w.hdr[HDR_CQR]) sweeps = run_jacobi_nov<T>(w.j2, z.kt, z.kt, lds);
        else sweeps = run_jacobi_nov<T>(w.j, z.mu, z.kt, lds);
        HCAMD_SUB_END(t_jac, w.hdr, 3);
#ifdef HCAMD_PHASE_TIMING
        if (threadIdx.x == 0) w.hdr[HDR_SUB + 7] = sweeps + 1000 * z.kt;  // timing build: sweeps and core width
#endif
    }
    ph_svd_finish<T>(d, sweeps);
}

// Reconstruction after ph_svd_nov, all r columns of the new rank:
//   C_U  = Q_U [Q2] [Z_perm Sigma^-1; 0]       (the left factor, as ph_recon side 0)
//   vtmp = Q_Y [X^T C_U(0:xrows, :); 0]        (X^T U = V1 Sigma: the right factor carrying Sigma)
// X is the core the Jacobi orthogonalised, still intact in the workspace: R2, the upper triangle
// of J after the core QR, or P R_Y[0:kt, :]^T, read from Y through the pivoting.
template <typename T, int PB>
__device__ void ph_recon_nov(const RecompDesc<T> &d, T *lds) {
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    const int r = (int) w.hdr[HDR_RANK];
    if (r < 1) return;
    const int m = d.m, n = d.n, mu = z.mu, kk = z.kk, kt = z.kt;
    const bool cqr = w.hdr[HDR_CQR] != 0;
    const T *zc = cqr ? w.j2 : w.j;
    const int xrows = cqr ? kt : mu;
    T *U = d.cdata;  // ld m
    for (ColWalk p(threadIdx.x, NT, m); p.c < r; p.next()) {
        const int i = p.i, c = p.c;
        T v = T(0);
        if (i < xrows && c < kt) {
            const int j = w.perm[c];
            const T sg = w.sig[j];
            v = (sg > T(0)) ? zc[i + (int64_t) j * xrows] / sg : T(0);
        }
        U[i + (int64_t) c * m] = v;
    }
    if (n > kt)
        for (ColWalk p(threadIdx.x, NT, n - kt); p.c < r; p.next()) w.vtmp[kt + p.i + (int64_t) p.c * n] = T(0);
    __syncthreads();
    HCAMD_SUB_BEGIN(t_rf);
    const int tm = (kt + GT_M - 1) / GT_M, tn = (r + GT_N - 1) / GT_N;
    if (cqr) {
        const T *R2 = w.j;  // R2[i, j] = J[i + j mu] for i <= j
        auto fa = [=](int j, int i) -> T { return i <= j ? R2[i + (int64_t) j * mu] : T(0); };
        auto fb = [=](int i, int c) -> T { return U[i + (int64_t) c * m]; };
        // row j of R2^T vanishes beyond column j: the K range of output rows [64 ti, 64 ti + 64)
        for (int t = 0; t < tm * tn; ++t)
            wg_gemm_tile_fn<T>(kt, r, kt, fa, fb, w.vtmp, n, t % tm, t / tm, lds, 0, (t % tm) * GT_M + GT_M);
    } else {
        const T *Y = w.y;  // X[piv[i], j] = R_Y[j, i] = Y[j + i n] for i >= j
        const int32_t *piv = w.piv;
        auto fa = [=](int j, int i) -> T { return i >= j ? Y[j + (int64_t) i * n] : T(0); };
        auto fb = [=](int i, int c) -> T { return U[piv[i] + (int64_t) c * m]; };
        // row j of R_Y^T P vanishes before column j: the K range starts at the tile's first row
        for (int t = 0; t < tm * tn; ++t)
            wg_gemm_tile_fn<T>(kt, r, mu, fa, fb, w.vtmp, n, t % tm, t / tm, lds, (t % tm) * GT_M);
    }
    __syncthreads();
    HCAMD_SUB_END(t_rf, w.hdr, 6);
    if (cqr) {
        wg_apply_q<T, PB>(mu, kt, w.j, mu, w.t_j, U, m, 0, r, lds);
        __syncthreads();
    }
    HCAMD_SUB_BEGIN(t_aq);
    wg_apply_q<T, PB>(m, mu, w.us, m, w.t_u, U, m, 0, r, lds);
    __syncthreads();
    // [X^T U; 0] vanishes below row kt, where reflectors kt.. of Q_Y act as the identity
    wg_apply_q<T, PB>(n, kt < kk ? kt : kk, w.y, n, w.t_y, w.vtmp, n, 0, r, lds);
    HCAMD_SUB_END(t_aq, w.hdr, 4);
}

// ---------------------------------------------------------------------------
// Phase: reconstruction of columns [c0, c1) of the new rank.
//   side 0: C_U[:, c] = Q_U [z_perm(c) / sigma; 0]  (z: column of J = P R_Y^T after the Jacobi, or
//           of Q2 R2 W after the core QR; P, the pivoting of the QR of Y, is applied when J is built)
//   side 1: vtmp[:, c] = Q_Y [v1_perm(c) * sigma; 0]
// ---------------------------------------------------------------------------
template <typename T, int PB>
__device__ void ph_recon(const RecompDesc<T> &d, int side, int c0, int c1, T *lds) {
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    const int r = (int) w.hdr[HDR_RANK];
    if (c1 > r) c1 = r;
    if (c0 >= c1) return;
    const int m = d.m, n = d.n, mu = z.mu, kk = z.kk, kt = z.kt;
    if (side == 0 && w.hdr[HDR_CQR]) {
        T *B = d.cdata;  // ld m
        // left factor of the core J = Q2 R2: Q2 [Z2 / sigma; 0] in rows [0, mu), zero below
        for (ColWalk p(threadIdx.x, NT, m); p.c < c1 - c0; p.next()) {
            const int i = p.i, c = c0 + p.c;
            T v = T(0);
            if (i < kt && c < kt) {
                const int j = w.perm[c];
                const T sg = w.sig[j];
                v = (sg > T(0)) ? w.j2[i + (int64_t) j * kt] / sg : T(0);
            }
            B[i + (int64_t) c * m] = v;
        }
        __syncthreads();
        wg_apply_q<T, PB>(mu, kt, w.j, mu, w.t_j, B, m, c0, c1, lds);
        __syncthreads();
        HCAMD_SUB_BEGIN(t_aq);
        wg_apply_q<T, PB>(m, mu, w.us, m, w.t_u, B, m, c0, c1, lds);
        HCAMD_SUB_END(t_aq, w.hdr, 4);
    } else if (side == 0) {
        T *B = d.cdata;  // ld m
        // the core's left factor Z / sigma in rows [0, mu) (J's rows are in Q_U's basis order)
        for (ColWalk p(threadIdx.x, NT, m); p.c < c1 - c0; p.next()) {
            const int i = p.i, c = c0 + p.c;
            T v = T(0);
            if (i < mu && c < kt) {
                const int j = w.perm[c];
                const T sg = w.sig[j];
                v = (sg > T(0)) ? w.j[i + (int64_t) j * mu] / sg : T(0);
            }
            B[i + (int64_t) c * m] = v;
        }
        __syncthreads();
        HCAMD_SUB_BEGIN(t_aq);
        wg_apply_q<T, PB>(m, mu, w.us, m, w.t_u, B, m, c0, c1, lds);
        HCAMD_SUB_END(t_aq, w.hdr, 4);
    } else {
        T *B = w.vtmp;  // ld n
        for (ColWalk p(threadIdx.x, NT, n); p.c < c1 - c0; p.next()) {
            const int i = p.i, c = c0 + p.c;
            T v = T(0);
            if (i < kt && c < kt) {
                const int j = w.perm[c];
                v = w.v1[i + (int64_t) j * kt] * w.sig[j];
            }
            B[i + (int64_t) c * n] = v;
        }
        __syncthreads();
        // [v1 sigma; 0] vanishes below row kt, where reflectors kt.. of Q_Y act as the identity
        // (each is zero above its own row), so only the first kt are applied
        HCAMD_SUB_BEGIN(t_aq);
        wg_apply_q<T, PB>(n, kt < kk ? kt : kk, w.y, n, w.t_y, B, n, c0, c1, lds);
        HCAMD_SUB_END(t_aq, w.hdr, 4);
    }
}

// Phase: C_V (r x n, ld r) = vtmp^T.
template <typename T>
__device__ void ph_finish(const RecompDesc<T> &d, int wg, int nwg) {
    Ws<T> w = ws_of(d);
    const int r = (int) w.hdr[HDR_RANK];
    const int n = d.n;
    T *cv = d.cdata + (int64_t) d.m * d.max_rank;
    if (r < 1) return;
    // element (c, i) of C_V is cv[c + i r]: the walk's rows are C_V's rows (the rank)
    for (ColWalk p(wg * NT + threadIdx.x, nwg * NT, r); p.c < n; p.next())
        cv[p.i + (int64_t) p.c * r] = w.vtmp[p.c + (int64_t) p.i * n];
}

}  // namespace hcamd
