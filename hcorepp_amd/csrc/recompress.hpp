// Workgroup-level device phases of one compressed-tile update
// C <- alpha * Afin * Bfin + beta * C with C = C_U * C_V (rank rc, capacity max_rank).
//
// The phases restate CompressedTile<T>::Gemm (reference
// src/operators/concrete/Compressed.cpp:208-694) in a form suited to CDNA4:
//   stack      U_s = [C_U | alpha*Afin], V_s = [(beta*C_V)^T | Bfin^T]      (:331-349, :375-380)
//   qr         Householder QR of U_s and V_s, compact-WY panels in LDS       (:351-354, :383-385)
//   core       X = R_V * R_U^T  (the transpose of R_U * R_V^T, :454-461)
//   svd        one-sided Jacobi on X: X*W = Z, sigma_j = |Z_j|; sort; rank  (:466-527)
//   recon      C_U = Q_U [W_r; 0],  C_V = (Q_V [Z_r; 0])^T                  (:549-628, :656-682)
// Since X = M^T with M = R_U R_V^T, M = W Z^T: the left singular vectors of M
// are the columns of W and the right ones scaled by sigma are the columns of Z,
// so "V carries Sigma" (:597-599) falls out of the rotation without a scaling pass.
//
// Every phase is executed by a whole workgroup of NT threads; the same code
// runs one phase per launch (plain path) or all phases in one persistent
// launch (batched path).
#pragma once

#include "device_common.hpp"

namespace hcamd {

// Header slots at the start of each update's workspace (int64).
enum : int {
    HDR_K = 0,       // stacked rank K = arank + rc
    HDR_RANK = 1,    // new rank r
    HDR_SIZES = 2,   // size_s = min(m, n, K)
    HDR_SWEEPS = 3,  // Jacobi sweeps used
    HDR_STATUS = 4,  // status bits (STATUS_*)
    HDR_WORDS = 8,
};
enum : int { STATUS_NOCONV = 1, STATUS_NAN = 2 };

constexpr int JACOBI_MAX_SWEEPS = 60;

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
    int32_t pad;
    T alpha, beta;
    double accuracy;
    double *sigma_out;   // optional: sorted singular values (size_s entries)
};

// Workspace layout (element offsets of T) for capacity kmax and tile m x n.
struct WsLayout {
    int64_t hdr, us, vs, tau_u, tau_v, t_u, t_v, x, w, sig, perm, vtmp, total;
};

__host__ __device__ inline int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

template <typename T>
__host__ __device__ inline WsLayout ws_layout(int64_t m, int64_t n, int64_t kmax, int64_t max_rank) {
    WsLayout L;
    int64_t o = 0;
    const int64_t hdr_elems = (HDR_WORDS * 8 + sizeof(T) - 1) / sizeof(T);
    L.hdr = o; o += align_up(hdr_elems, 32);
    L.us = o; o += align_up(m * kmax, 32);
    L.vs = o; o += align_up(n * kmax, 32);
    L.tau_u = o; o += align_up(kmax, 32);
    L.tau_v = o; o += align_up(kmax, 32);
    L.t_u = o; o += align_up(kmax * 16 + 256, 32);
    L.t_v = o; o += align_up(kmax * 16 + 256, 32);
    const int64_t mu = m < kmax ? m : kmax, mv = n < kmax ? n : kmax;
    L.x = o; o += align_up(mv * mu, 32);
    L.w = o; o += align_up(mu * mu, 32);
    L.sig = o; o += align_up(mu, 32);
    L.perm = o; o += align_up((mu * 4 + sizeof(T) - 1) / sizeof(T), 32);
    L.vtmp = o; o += align_up(n * (max_rank > 0 ? max_rank : 1), 32);
    L.total = o;
    return L;
}

template <typename T>
struct Ws {
    int64_t *hdr;
    T *us, *vs, *tau_u, *tau_v, *t_u, *t_v, *x, *w, *sig, *vtmp;
    int32_t *perm;
};

template <typename T>
__device__ inline Ws<T> ws_of(const RecompDesc<T> &d) {
    const WsLayout L = ws_layout<T>(d.m, d.n, d.kmax, d.max_rank);
    Ws<T> w;
    w.hdr = reinterpret_cast<int64_t *>(d.ws + L.hdr);
    w.us = d.ws + L.us;
    w.vs = d.ws + L.vs;
    w.tau_u = d.ws + L.tau_u;
    w.tau_v = d.ws + L.tau_v;
    w.t_u = d.ws + L.t_u;
    w.t_v = d.ws + L.t_v;
    w.x = d.ws + L.x;
    w.w = d.ws + L.w;
    w.sig = d.ws + L.sig;
    w.perm = reinterpret_cast<int32_t *>(d.ws + L.perm);
    w.vtmp = d.ws + L.vtmp;
    return w;
}

// ---------------------------------------------------------------------------
// Phase 1: stack.  Grid-stride over (part, block) so that several workgroups
// may share one update: `wg` in [0, nwg).
// ---------------------------------------------------------------------------
template <typename T>
__device__ void ph_stack(const RecompDesc<T> &d, int wg, int nwg) {
    Ws<T> w = ws_of(d);
    const int64_t rc = *d.crank;
    const int64_t m = d.m, n = d.n, ra = d.arank;
    const int64_t K = ra + rc;
    if (wg == 0 && threadIdx.x == 0) {
        w.hdr[HDR_K] = K;
        w.hdr[HDR_STATUS] = 0;
    }
    const T *cu = d.cdata;
    const T *cv = d.cdata + m * d.max_rank;  // ld rc
    const int64_t totU = m * K, totV = n * K;
    const int64_t stride = (int64_t) nwg * NT;
    for (int64_t e = (int64_t) wg * NT + threadIdx.x; e < totU; e += stride) {
        const int64_t i = e % m, j = e / m;
        w.us[e] = (j < rc) ? cu[i + j * m] : d.alpha * d.afin[i + (j - rc) * d.ldaf];
    }
    // V_s[j, i] = beta * C_V[i, j] for i < rc ; V_s[j, rc + i] = Bfin[i, j]  (ProcessVpointer)
    for (int64_t e = (int64_t) wg * NT + threadIdx.x; e < totV; e += stride) {
        const int64_t j = e % n, i = e / n;
        w.vs[e] = (i < rc) ? d.beta * cv[i + j * rc] : d.bfin[(i - rc) + j * d.ldbf];
    }
}

// ---------------------------------------------------------------------------
// Compact-WY helpers.  A panel of PB reflectors for rows [j0, m) lives in LDS
// as P[r + c*rows] (c < PB), unit diagonal and zeros above, columns >= jb zero.
// T (PB x PB upper) lives in LDS too.
// ---------------------------------------------------------------------------

// Apply (I - V T V^T) (trans=false) or (I - V T^T V^T) (trans=true) from the
// left to columns [c0, c1) of B (rows [j0, j0+rows) of a column-major matrix
// with leading dimension ldb).  One wave per column; all waves participate.
template <typename T, int PB>
__device__ void apply_block_reflector(const T *P, const T *Tm, int rows, T *B, int64_t ldb, int c0, int c1,
                                      bool trans) {
    const int l = lane_id();
    const int idx = wave_multi_sum_index<PB>(l);
    for (int c = c0 + wave_id(); c < c1; c += NWAVES) {
        T *col = B + (int64_t) c * ldb;
        T acc[PB];
#pragma unroll
        for (int k = 0; k < PB; ++k) acc[k] = T(0);
        for (int r = l; r < rows; r += WAVE) {
            const T a = col[r];
#pragma unroll
            for (int k = 0; k < PB; ++k) acc[k] += P[r + k * rows] * a;
        }
        const T dot = wave_multi_sum<T, PB>(acc);  // lane l holds sum for index idx
        // t_k = (T^T w)_k = sum_{j<=k} T[j][k] w_j   or   (T w)_k = sum_{j>=k} T[k][j] w_j
        T tk = T(0);
#pragma unroll
        for (int j = 0; j < PB; ++j) {
            // lane holding index j: bits chosen as in wave_multi_sum_index
            int src = 0, cnt = PB, off = 32, jj = j;
#pragma unroll
            for (int round = 0; round < 6; ++round) {
                if (cnt > 1) {
                    const int half = cnt >> 1;
                    if (jj >= half) {
                        src |= off;
                        jj -= half;
                    }
                    cnt = half;
                }
                off >>= 1;
            }
            const T wj = __shfl(dot, src, WAVE);
            const T coef = trans ? ((j <= idx) ? Tm[j + idx * PB] : T(0)) : ((j >= idx) ? Tm[idx + j * PB] : T(0));
            tk += coef * wj;
        }
        // broadcast t_k for all k
        T tv[PB];
#pragma unroll
        for (int k = 0; k < PB; ++k) {
            int src = 0, cnt = PB, off = 32, kk = k;
#pragma unroll
            for (int round = 0; round < 6; ++round) {
                if (cnt > 1) {
                    const int half = cnt >> 1;
                    if (kk >= half) {
                        src |= off;
                        kk -= half;
                    }
                    cnt = half;
                }
                off >>= 1;
            }
            tv[k] = __shfl(tk, src, WAVE);
        }
        for (int r = l; r < rows; r += WAVE) {
            T s = T(0);
#pragma unroll
            for (int k = 0; k < PB; ++k) s += P[r + k * rows] * tv[k];
            col[r] -= s;
        }
    }
}

// Build T (PB x PB, upper, column-major in LDS) for the panel P with taus tau[0..jb):
// dlarft forward/columnwise: T[i][i] = tau_i, T[0:i, i] = -tau_i T[0:i,0:i] (V[:,0:i]^T v_i).
template <typename T, int PB>
__device__ void build_t(const T *P, int rows, const T *tau, int jb, T *Tm, T *gram) {
    // gram[a + b*PB] = v_a^T v_b for a < b < jb
    const int l = lane_id();
    for (int pr = wave_id(); pr < PB * PB; pr += NWAVES) {
        const int a = pr % PB, b = pr / PB;
        if (a >= b || b >= jb) continue;
        T s = T(0);
        for (int r = l; r < rows; r += WAVE) s += P[r + a * rows] * P[r + b * rows];
        s = wave_sum(s);
        if (l == 0) gram[a + b * PB] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 0; i < PB * PB; ++i) Tm[i] = T(0);
        for (int i = 0; i < jb; ++i) {
            const T ti = tau[i];
            Tm[i + i * PB] = ti;
            // tmp = V[:,0:i]^T v_i = gram[0:i, i]; T[0:i, i] = -ti * T[0:i,0:i] * tmp
            for (int a = 0; a < i; ++a) {
                T s = T(0);
                for (int b = a; b < i; ++b) s += Tm[a + b * PB] * gram[b + i * PB];
                Tm[a + i * PB] = -ti * s;
            }
        }
    }
    __syncthreads();
}

// Load the reflector panel [j0, j0+jb) of the factored matrix A (ld lda) into LDS.
template <typename T, int PB>
__device__ void load_reflector_panel(const T *A, int64_t lda, int m, int j0, int jb, T *P) {
    const int rows = m - j0;
    for (int e = threadIdx.x; e < rows * PB; e += NT) {
        const int r = e % rows, c = e / rows;
        T v = T(0);
        if (c < jb) {
            if (r == c) v = T(1);
            else if (r > c) v = A[(j0 + r) + (int64_t) (j0 + c) * lda];
        }
        P[e] = v;
    }
}

// ---------------------------------------------------------------------------
// Phase 2: blocked Householder QR (LAPACK dgeqrf convention) of A (m x ncols,
// ld lda) in place; tau[0..min(m,ncols)); T factors of every panel saved to
// tsave[panel * PB * PB].  `lds` must hold m*PB + 2*PB*PB + 64 elements.
// ---------------------------------------------------------------------------
template <typename T, int PB>
__device__ void wg_geqrf(int m, int ncols, T *A, int64_t lda, T *tau, T *tsave, T *lds) {
    T *P = lds;
    T *Tm = P + (int64_t) m * PB;
    T *gram = Tm + PB * PB;
    T *red = gram + PB * PB;   // NWAVES * PB
    const int nsteps = m < ncols ? m : ncols;
    const int l = lane_id();
    for (int j0 = 0; j0 < nsteps; j0 += PB) {
        const int jb = (nsteps - j0) < PB ? (nsteps - j0) : PB;
        const int rows = m - j0;
        // load panel (all PB columns; columns >= jb are zero, columns beyond ncols too)
        for (int e = threadIdx.x; e < rows * PB; e += NT) {
            const int r = e % rows, c = e / rows;
            P[e] = (c < jb) ? A[(j0 + r) + (int64_t) (j0 + c) * lda] : T(0);
        }
        __syncthreads();
        for (int jj = 0; jj < jb; ++jj) {
            // ||x[jj+1:]||^2
            T part = T(0);
            for (int r = jj + 1 + threadIdx.x; r < rows; r += NT) {
                const T v = P[r + jj * rows];
                part += v * v;
            }
            const T xnorm2 = wg_sum(part, red);
            const T alpha = P[jj + jj * rows];
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
            __syncthreads();
            if (tau_j != T(0)) {
                for (int r = jj + 1 + threadIdx.x; r < rows; r += NT) P[r + jj * rows] *= scal;
            }
            if (threadIdx.x == 0) {
                P[jj + jj * rows] = beta;
                tau[j0 + jj] = tau_j;
            }
            __syncthreads();
            // apply H_jj to panel columns jj+1..jb-1:  w_c = x_c[jj] + sum_{r>jj} v_r x_c[r]
            if (tau_j != T(0) && jj + 1 < jb) {
                T acc[PB];
#pragma unroll
                for (int k = 0; k < PB; ++k) acc[k] = T(0);
                for (int r = jj + 1 + threadIdx.x; r < rows; r += NT) {
                    const T v = P[r + jj * rows];
#pragma unroll
                    for (int k = 0; k < PB; ++k)
                        if (k > jj && k < jb) acc[k] += v * P[r + k * rows];
                }
                const int idx = wave_multi_sum_index<PB>(l);
                const T s = wave_multi_sum<T, PB>(acc);
                // lanes 0..63 hold index idx(l); the first lane with each index writes
                int first = 1;
                {
                    // lane l is the first holder of idx iff its low bits below the index bits are zero
                    int cnt = PB, off = 32, consumed = 0;
                    for (int round = 0; round < 6 && cnt > 1; ++round) {
                        cnt >>= 1;
                        consumed |= off;
                        off >>= 1;
                    }
                    first = ((l & ~consumed) == 0);
                }
                if (first) red[wave_id() * PB + idx] = s;
                __syncthreads();
                if (threadIdx.x < PB) {
                    const int k = threadIdx.x;
                    if (k > jj && k < jb) {
                        T wsum = P[jj + k * rows];
                        for (int wv = 0; wv < NWAVES; ++wv) wsum += red[wv * PB + k];
                        gram[k] = tau_j * wsum;  // reuse gram[0..PB) as the scaled w
                    }
                }
                __syncthreads();
                for (int e = threadIdx.x; e < rows * PB; e += NT) {
                    const int r = e % rows, k = e / rows;
                    if (k > jj && k < jb && r >= jj) {
                        const T v = (r == jj) ? T(1) : P[r + jj * rows];
                        P[e] -= gram[k] * v;
                    }
                }
                __syncthreads();
            }
        }
        // write the factored panel back (R entries on/above diagonal, reflectors below)
        for (int e = threadIdx.x; e < rows * jb; e += NT) {
            const int r = e % rows, c = e / rows;
            A[(j0 + r) + (int64_t) (j0 + c) * lda] = P[e];
        }
        __syncthreads();
        // reflector form of the panel (unit diagonal, zeros above) for T and the update
        for (int e = threadIdx.x; e < rows * PB; e += NT) {
            const int r = e % rows, c = e / rows;
            if (c < jb) {
                if (r == c) P[e] = T(1);
                else if (r < c) P[e] = T(0);
            }
        }
        __syncthreads();
        build_t<T, PB>(P, rows, tau + j0, jb, Tm, gram);
        if (tsave) {
            for (int e = threadIdx.x; e < PB * PB; e += NT) tsave[(int64_t) (j0 / PB) * PB * PB + e] = Tm[e];
        }
        // trailing update: A[j0:, j0+jb:ncols] = (I - V T^T V^T) A[...]   (H^T applied: Q^T A)
        if (j0 + jb < ncols) {
            apply_block_reflector<T, PB>(P, Tm, rows, A + j0, lda, j0 + jb, ncols, true);
        }
        __syncthreads();
    }
}

// Apply Q = H_1 ... H_k (k = nref reflectors of the factored matrix A, ld lda,
// m rows) to B (m x ncols, ld ldb): B <- Q B.  T factors from tsave.
template <typename T, int PB>
__device__ void wg_apply_q(int m, int nref, const T *A, int64_t lda, const T *tsave, T *B, int64_t ldb, int c0,
                           int c1, T *lds) {
    T *P = lds;
    T *Tm = P + (int64_t) m * PB;
    const int npan = (nref + PB - 1) / PB;
    for (int p = npan - 1; p >= 0; --p) {
        const int j0 = p * PB;
        const int jb = (nref - j0) < PB ? (nref - j0) : PB;
        const int rows = m - j0;
        load_reflector_panel<T, PB>(A, lda, m, j0, jb, P);
        for (int e = threadIdx.x; e < PB * PB; e += NT) Tm[e] = tsave[(int64_t) p * PB * PB + e];
        __syncthreads();
        apply_block_reflector<T, PB>(P, Tm, rows, B + j0, ldb, c0, c1, false);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Phase 3: X = R_V R_U^T (minV x minU, ld minV) and W = I (minU x minU).
// Grid-stride over elements so several workgroups may share one update.
// ---------------------------------------------------------------------------
template <typename T>
__device__ void ph_core(const RecompDesc<T> &d, int wg, int nwg) {
    Ws<T> w = ws_of(d);
    const int64_t K = w.hdr[HDR_K];
    const int64_t m = d.m, n = d.n;
    const int64_t mu = m < K ? m : K, mv = n < K ? n : K;
    const int64_t stride = (int64_t) nwg * NT;
    for (int64_t e = (int64_t) wg * NT + threadIdx.x; e < mv * mu; e += stride) {
        const int64_t i = e % mv, j = e / mv;
        const int64_t k0 = i > j ? i : j;
        T s = T(0);
        for (int64_t k = k0; k < K; ++k) s += w.vs[i + k * n] * w.us[j + k * m];
        w.x[e] = s;
    }
    for (int64_t e = (int64_t) wg * NT + threadIdx.x; e < mu * mu; e += stride) {
        const int64_t i = e % mu, j = e / mu;
        w.w[e] = (i == j) ? T(1) : T(0);
    }
}

// ---------------------------------------------------------------------------
// Phase 4: one-sided Jacobi on X (rows mv, cols mu) with W accumulation,
// round-robin (circle) ordering, relative convergence test
// |x_p . x_q| <= tol |x_p||x_q| with tol = sqrt(mv) * eps.  Then
// sigma_j = |x_j|, descending sort (perm), rank rule (omp/kernels.cpp:82-104)
// and clamps (Compressed.cpp:509-527).  Single workgroup.
// ---------------------------------------------------------------------------
template <typename T>
__device__ void ph_svd(const RecompDesc<T> &d, T *lds, int lds_elems) {
    Ws<T> w = ws_of(d);
    const int K = (int) w.hdr[HDR_K];
    const int m = d.m, n = d.n;
    const int mu = m < K ? m : K, mv = n < K ? n : K;
    const int size_s = mu < mv ? mu : mv;
    __shared__ int s_rot;
    __shared__ int s_flag_nan;
    // Stage X and W in LDS when both fit.
    T *X = w.x, *W = w.w;
    const bool in_lds = (int64_t) mv * mu + (int64_t) mu * mu <= lds_elems;
    if (in_lds) {
        X = lds;
        W = lds + (int64_t) mv * mu;
        for (int e = threadIdx.x; e < mv * mu; e += NT) X[e] = w.x[e];
        for (int e = threadIdx.x; e < mu * mu; e += NT) W[e] = w.w[e];
    }
    __syncthreads();
    const T tol = sqrt(T(mv)) * Eps<T>::value;
    const int np = mu + (mu & 1);
    const int l = lane_id();
    // Columns whose squared norm is below (eps * ||X||_F)^2 are numerically zero
    // (they carry singular values at the rounding level of the largest one); pairs
    // involving them are not rotated.  Without this floor a core with more columns
    // than rows (K > n) never meets the relative orthogonality test.
    __shared__ T s_red[NWAVES];
    T fro_part = T(0);
    for (int e = threadIdx.x; e < mv * mu; e += NT) fro_part += X[e] * X[e];
    const T fro2 = wg_sum(fro_part, s_red);
    const T floor2 = Eps<T>::value * Eps<T>::value * fro2;
    int sweep = 0;
    for (; sweep < JACOBI_MAX_SWEEPS && mu > 1; ++sweep) {
        if (threadIdx.x == 0) s_rot = 0;
        __syncthreads();
        int rotated = 0;
        for (int s = 0; s < np - 1; ++s) {
            for (int pr = wave_id(); pr < np / 2; pr += NWAVES) {
                int p, q;
                if (pr == 0) {
                    p = np - 1;
                    q = s;
                } else {
                    p = (s + pr) % (np - 1);
                    q = (s - pr + np - 1) % (np - 1);
                }
                if (p > q) {
                    const int t = p;
                    p = q;
                    q = t;
                }
                if (q >= mu) continue;
                T *xp = X + (int64_t) p * mv, *xq = X + (int64_t) q * mv;
                T app = 0, aqq = 0, apq = 0;
                for (int r = l; r < mv; r += WAVE) {
                    const T a = xp[r], b = xq[r];
                    app += a * a;
                    aqq += b * b;
                    apq += a * b;
                }
                app = wave_sum(app);
                aqq = wave_sum(aqq);
                apq = wave_sum(apq);
                if (!(app > floor2) || !(aqq > floor2)) continue;
                if (fabs(apq) <= tol * sqrt(app) * sqrt(aqq)) continue;
                const T zeta = (aqq - app) / (T(2) * apq);
                const T t = dsign(T(1), zeta) / (fabs(zeta) + sqrt(T(1) + zeta * zeta));
                const T c = T(1) / sqrt(T(1) + t * t);
                const T sn = c * t;
                for (int r = l; r < mv; r += WAVE) {
                    const T a = xp[r], b = xq[r];
                    xp[r] = c * a - sn * b;
                    xq[r] = sn * a + c * b;
                }
                T *wp = W + (int64_t) p * mu, *wq = W + (int64_t) q * mu;
                for (int r = l; r < mu; r += WAVE) {
                    const T a = wp[r], b = wq[r];
                    wp[r] = c * a - sn * b;
                    wq[r] = sn * a + c * b;
                }
                rotated = 1;
            }
            __syncthreads();
        }
        if (rotated && l == 0) atomicOr(&s_rot, 1);
        __syncthreads();
        if (s_rot == 0) break;
        __syncthreads();
    }
    // sigma_j = |x_j|
    if (threadIdx.x == 0) s_flag_nan = 0;
    __syncthreads();
    for (int j = wave_id(); j < mu; j += NWAVES) {
        const T *xj = X + (int64_t) j * mv;
        T s = 0;
        for (int r = l; r < mv; r += WAVE) s += xj[r] * xj[r];
        s = wave_sum(s);
        if (l == 0) {
            w.sig[j] = sqrt(s);
            if (!(s == s)) atomicOr(&s_flag_nan, 1);
        }
    }
    __syncthreads();
    // descending stable sort by rank counting
    for (int j = threadIdx.x; j < mu; j += NT) {
        const T sj = w.sig[j];
        int pos = 0;
        for (int k = 0; k < mu; ++k) {
            const T sk = w.sig[k];
            pos += (sk > sj) || (sk == sj && k < j);
        }
        w.perm[pos] = j;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // rank rule on sorted sigma (first size_s values)
        int64_t rk;
        if (d.fixed_rank > 0) {
            rk = d.fixed_rank;
            if (rk > K) rk = K;
            if (rk > size_s) rk = size_s;
        } else {
            rk = size_s;
            const T s0 = w.sig[w.perm[0]];
            const T thr = d.truncated_svd ? T(d.accuracy) * s0 : T(d.accuracy);
            for (int i = 1; i < size_s; ++i) {
                if (w.sig[w.perm[i]] < thr) {
                    rk = i;
                    break;
                }
            }
        }
        if (rk < 1) rk = 1;
        if (rk > d.max_rank) rk = d.max_rank;
        w.hdr[HDR_RANK] = rk;
        w.hdr[HDR_SIZES] = size_s;
        w.hdr[HDR_SWEEPS] = sweep;
        int st = 0;
        if (sweep >= JACOBI_MAX_SWEEPS) st |= STATUS_NOCONV;
        if (s_flag_nan) st |= STATUS_NAN;
        w.hdr[HDR_STATUS] |= st;
        if (d.sigma_out) {
            for (int i = 0; i < size_s; ++i) d.sigma_out[i] = (double) w.sig[w.perm[i]];
        }
    }
    if (in_lds) {
        __syncthreads();
        for (int e = threadIdx.x; e < mv * mu; e += NT) w.x[e] = X[e];
        for (int e = threadIdx.x; e < mu * mu; e += NT) w.w[e] = W[e];
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Phase 5: reconstruction.  side 0: C_U[:, c0:c1] = Q_U [W(:, perm); 0];
// side 1: vtmp[:, c0:c1] = Q_V [Z(:, perm); 0].  Columns [c0, c1) of the new
// rank are handled by this workgroup.
// ---------------------------------------------------------------------------
template <typename T, int PB>
__device__ void ph_recon(const RecompDesc<T> &d, int side, int c0, int c1, T *lds) {
    Ws<T> w = ws_of(d);
    const int K = (int) w.hdr[HDR_K];
    const int r = (int) w.hdr[HDR_RANK];
    if (c1 > r) c1 = r;
    if (c0 >= c1) return;
    const int m = d.m, n = d.n;
    const int mu = m < K ? m : K, mv = n < K ? n : K;
    if (side == 0) {
        T *B = d.cdata;  // ld m
        for (int64_t e = threadIdx.x; e < (int64_t) m * (c1 - c0); e += NT) {
            const int i = (int) (e % m), c = c0 + (int) (e / m);
            B[i + (int64_t) c * m] = (i < mu) ? w.w[i + (int64_t) w.perm[c] * mu] : T(0);
        }
        __syncthreads();
        wg_apply_q<T, PB>(m, mu, w.us, m, w.t_u, B, m, c0, c1, lds);
    } else {
        T *B = w.vtmp;  // ld n
        for (int64_t e = threadIdx.x; e < (int64_t) n * (c1 - c0); e += NT) {
            const int i = (int) (e % n), c = c0 + (int) (e / n);
            B[i + (int64_t) c * n] = (i < mv) ? w.x[i + (int64_t) w.perm[c] * mv] : T(0);
        }
        __syncthreads();
        wg_apply_q<T, PB>(n, mv, w.vs, n, w.t_v, B, n, c0, c1, lds);
    }
}

// Phase 6: C_V (r x n, ld r) = vtmp^T; publish the new rank.
template <typename T>
__device__ void ph_finish(const RecompDesc<T> &d, int wg, int nwg) {
    Ws<T> w = ws_of(d);
    const int64_t r = w.hdr[HDR_RANK];
    const int64_t n = d.n;
    T *cv = d.cdata + (int64_t) d.m * d.max_rank;
    const int64_t stride = (int64_t) nwg * NT;
    for (int64_t e = (int64_t) wg * NT + threadIdx.x; e < r * n; e += stride) {
        const int64_t c = e % r, i = e / r;
        cv[e] = w.vtmp[i + c * n];
    }
}

}  // namespace hcamd
