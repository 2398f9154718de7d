// Learned rank predictor of the batched recompression: per-update features computed on the
// device right after stacking, and a small MLP (RANK_FEATURES -> RANK_HIDDEN -> RANK_HIDDEN -> 1,
// ReLU) that predicts log2 of the post-truncation rank.  Inference runs inside the batched kernel
// with BF16 weights and activations and FP32 accumulation; training uses the FP32 kernels of
// mlp.hip.  Parameter layout (row-major): W1 [IN][H], b1 [H], W2 [H][H], b2 [H], w3 [H], b3.
#pragma once

#include "recompress.hpp"

namespace hcamd {

constexpr int RANK_FEATURES = 24;
constexpr int RANK_HIDDEN = 64;
constexpr int MLP_W1 = 0;
constexpr int MLP_B1 = MLP_W1 + RANK_FEATURES * RANK_HIDDEN;
constexpr int MLP_W2 = MLP_B1 + RANK_HIDDEN;
constexpr int MLP_B2 = MLP_W2 + RANK_HIDDEN * RANK_HIDDEN;
constexpr int MLP_W3 = MLP_B2 + RANK_HIDDEN;
constexpr int MLP_B3 = MLP_W3 + RANK_HIDDEN;
constexpr int RANK_PARAMS = MLP_B3 + 1;
static_assert(RANK_HIDDEN == WAVE, "one hidden unit per lane");
static_assert((HDR_PRED - HDR_FEAT) * 2 >= RANK_FEATURES, "features fit their header slots");

__host__ __device__ inline float bf16_to_f32(uint16_t h) {
    union {
        uint32_t u;
        float f;
    } c;
    c.u = ((uint32_t) h) << 16;
    return c.f;
}

// Round to nearest even BF16 (finite inputs).
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
    union {
        uint32_t u;
        float f;
    } c;
    c.f = f;
    const uint32_t lsb = (c.u >> 16) & 1u;
    return (uint16_t) ((c.u + 0x7FFFu + lsb) >> 16);
}

__device__ inline float bf16_round(float f) { return bf16_to_f32(f32_to_bf16(f)); }

// Features of one update (after ph_stack), written as float32 to header words HDR_FEAT...:
//  0-2  input ranks ra, rb, rc / 512          3  log2(max(m, n)) / 10     4  -log10(accuracy) / 10
//  5-12  log10 |R_ii| / 8 of the first 8 columns of V_s's C block (beta C_V^T), -5 when absent
//  13-20 the same for V_s's A*B block          21 truncated-SVD flag       22 K / 512     23 0
// |R_ii| comes from the Cholesky factor of the 8 x 8 Gram matrix of those columns.
template <typename T>
__device__ void ph_features(const RecompDesc<T> &d, T *lds) {
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    const int n = d.n, ra = d.arank, rc = z.K - d.arank, l = lane_id();
    const StackedV<T> vs = stacked_v(d, z.K);
    double *G = reinterpret_cast<double *>(lds);  // 2 x 8 x 8
    double *diag = G + 128;                        // 16
    for (int e = wave_id(); e < 72; e += NWAVES) {
        const int blk = e / 36;
        int q = e % 36, a = 0;
        while (q >= 8 - a) {
            q -= 8 - a;
            ++a;
        }
        const int b = a + q;
        const int c0 = blk == 0 ? 0 : rc, nc = blk == 0 ? rc : ra;
        double sum = 0.0;
        if (b < nc) {
            for (int r = l; r < n; r += WAVE) sum += (double) vs(r, c0 + a) * (double) vs(r, c0 + b);
        }
        sum = wave_sum(sum);
        if (l == 0) {
            G[blk * 64 + a * 8 + b] = sum;
            G[blk * 64 + b * 8 + a] = sum;
        }
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const int blk = threadIdx.x;
        double *g = G + blk * 64;
        const int nc = (blk == 0 ? rc : ra) < 8 ? (blk == 0 ? rc : ra) : 8;
        for (int j = 0; j < 8; ++j) {
            if (j >= nc) {
                diag[blk * 8 + j] = 0.0;
                continue;
            }
            double s = g[j * 8 + j];
            for (int k = 0; k < j; ++k) s -= g[j * 8 + k] * g[j * 8 + k];
            const double dj = sqrt(s > 1e-80 ? s : 1e-80);
            g[j * 8 + j] = dj;
            diag[blk * 8 + j] = dj;
            for (int i = j + 1; i < nc; ++i) {
                double t = g[i * 8 + j];
                for (int k = 0; k < j; ++k) t -= g[i * 8 + k] * g[j * 8 + k];
                g[i * 8 + j] = t / dj;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < RANK_FEATURES) {
        const int f = threadIdx.x;
        float v = 0.f;
        if (f == 0) v = ra / 512.f;
        else if (f == 1) v = d.rb / 512.f;
        else if (f == 2) v = rc / 512.f;
        else if (f == 3) v = log2f((float) (d.m > d.n ? d.m : d.n)) / 10.f;
        else if (f == 4) v = d.accuracy > 0.0 ? (float) (-log10(d.accuracy) / 10.0) : 2.f;
        else if (f <= 20) {
            const double dj = diag[f - 5];
            v = dj > 0.0 ? (float) (log10(dj) / 8.0) : -5.f;
        } else if (f == 21) v = d.truncated_svd ? 1.f : 0.f;
        else if (f == 22) v = z.K / 512.f;
        reinterpret_cast<float *>(w.hdr + HDR_FEAT)[f] = v;
    }
    __syncthreads();
}

// The MLP on one update's features, evaluated by one full wave (lane j = hidden unit j).
// BF16 weights and activations, FP32 accumulation.  Returns the predicted log2 rank.
__device__ inline float rank_mlp_wave(const uint16_t *P, const float *x) {
    const int j = lane_id();
    float a = bf16_to_f32(P[MLP_B1 + j]);
    for (int i = 0; i < RANK_FEATURES; ++i) a += bf16_round(x[i]) * bf16_to_f32(P[MLP_W1 + i * RANK_HIDDEN + j]);
    const float h1 = bf16_round(a > 0.f ? a : 0.f);
    float b = bf16_to_f32(P[MLP_B2 + j]);
    for (int k = 0; k < RANK_HIDDEN; ++k) b += read_lane(h1, k) * bf16_to_f32(P[MLP_W2 + k * RANK_HIDDEN + j]);
    const float h2 = bf16_round(b > 0.f ? b : 0.f);
    return wave_sum(h2 * bf16_to_f32(P[MLP_W3 + j])) + bf16_to_f32(P[MLP_B3]);
}

// Learned core sizing (between ph_core_size and ph_core_build, single workgroup): the MLP's
// rank prediction r presizes the core to kc = min(kt, r + margin) columns.  Accepted only when
// the dropped rows of R_Y are small against the rank threshold,
// ||R_Y[kc:, :]||_F <= model_tol * threshold (then every singular value at index >= kc of the
// core is below that bound and the rank rule cannot pick a rank above kc); otherwise the
// exact width kt stays (fallback).  Header: HDR_PRED = prediction, HDR_LEARNED = 1 accepted,
// 2 fallback.  `lds` is the ph_core_size scratch (suffix sums, threshold bound).
//
// Without a model but with model_tol > 0, the same tolerance is applied directly to the pivoted
// R factor (the smallest kc with ||R_Y[kc:, :]||_F <= model_tol * threshold; HDR_LEARNED = 3): the
// rule the prediction is measured against.
template <typename T>
__device__ void ph_learned_size(const RecompDesc<T> &d, T *lds) {
    if (d.fixed_rank > 0 || !(d.model_tol > 0.0)) return;
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    const int kk = z.kk;
    if (!d.rank_model) {
        if (threadIdx.x == 0) {
            const T *rn = lds;
            const T lim = T(d.model_tol) * rn[kk];
            int kc = z.kt;
            for (int j = 1; j < z.kt; ++j) {
                if (sqrt(rn[j]) <= lim) {
                    kc = j;
                    break;
                }
            }
            w.hdr[HDR_KT] = kc;
            w.hdr[HDR_LEARNED] = 3;
        }
        __syncthreads();
        return;
    }
    __shared__ float s_pred;
    if (wave_id() == 0) {
        const float o = rank_mlp_wave(d.rank_model, reinterpret_cast<const float *>(w.hdr + HDR_FEAT));
        if (lane_id() == 0) s_pred = o;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const T *rn = lds;
        const T thr_lo = rn[kk];
        float o = s_pred;
        o = !(o == o) ? 12.f : (o < 0.f ? 0.f : (o > 12.f ? 12.f : o));  // NaN features: no presizing
        int pred = (int) ceilf(exp2f(o));
        if (pred > kk) pred = kk;
        int kc = pred + d.model_margin;
        const int kt = z.kt;
        if (kc > kt) kc = kt;
        if (kc < 1) kc = 1;
        int state = 1;
        if (kc < kt) {
            const T dropped = sqrt(rn[kc]);
            if (dropped <= T(d.model_tol) * thr_lo) w.hdr[HDR_KT] = kc;
            else state = 2;
        }
        w.hdr[HDR_PRED] = pred;
        w.hdr[HDR_LEARNED] = state;
    }
    __syncthreads();
}

}  // namespace hcamd
