// Training kernels of the rank predictor (FP32): forward pass, per-sample deltas, parameter
// gradients, Adam, and BF16 packing of the trained parameters for the in-kernel inference.
// Network: x (RANK_FEATURES) -> ReLU(W1^T x + b1) -> ReLU(W2^T h1 + b2) -> w3 . h2 + b3,
// loss = mean over samples of (out - y)^2.  Sample-major activations in a workspace of
// mlp_workspace_floats(n) floats: out[n] | h1[n][H] | h2[n][H] | d1[n][H] | d2[n][H] | dout[n] | loss partials.
#include "kernels.hpp"
#include "mlp.hpp"

namespace hcamd {

namespace {
constexpr int IN = RANK_FEATURES, H = RANK_HIDDEN;
constexpr int MLP_NT = 256;

struct MlpWs {
    float *out, *h1, *h2, *d1, *d2, *dout, *lossp;
};

__host__ __device__ inline MlpWs mlp_ws(float *ws, int64_t n) {
    MlpWs w;
    w.out = ws;
    w.h1 = w.out + n;
    w.h2 = w.h1 + n * H;
    w.d1 = w.h2 + n * H;
    w.d2 = w.d1 + n * H;
    w.dout = w.d2 + n * H;
    w.lossp = w.dout + n;
    return w;
}
}  // namespace

int64_t mlp_workspace_floats(int64_t n) { return 4 * H * n + 2 * n + (n + MLP_NT - 1) / MLP_NT + 8; }

__global__ void __launch_bounds__(MLP_NT) k_mlp_forward(const float *P, const float *X, int64_t n, float *ws) {
    const int64_t s = (int64_t) blockIdx.x * MLP_NT + threadIdx.x;
    if (s >= n) return;
    MlpWs w = mlp_ws(ws, n);
    const float *x = X + s * IN;
    float xin[IN];
#pragma unroll
    for (int i = 0; i < IN; ++i) xin[i] = x[i];
    float h1[H];
#pragma unroll
    for (int j = 0; j < H; ++j) {
        float a = P[MLP_B1 + j];
#pragma unroll
        for (int i = 0; i < IN; ++i) a += xin[i] * P[MLP_W1 + i * H + j];
        h1[j] = a > 0.f ? a : 0.f;
        w.h1[s * H + j] = h1[j];
    }
    float o = P[MLP_B3];
    for (int j = 0; j < H; ++j) {
        float b = P[MLP_B2 + j];
#pragma unroll
        for (int k = 0; k < H; ++k) b += h1[k] * P[MLP_W2 + k * H + j];
        const float h2 = b > 0.f ? b : 0.f;
        w.h2[s * H + j] = h2;
        o += h2 * P[MLP_W3 + j];
    }
    w.out[s] = o;
}

__global__ void __launch_bounds__(MLP_NT) k_mlp_deltas(const float *P, const float *y, int64_t n, float *ws) {
    __shared__ float red[MLP_NT / WAVE];
    const int64_t s = (int64_t) blockIdx.x * MLP_NT + threadIdx.x;
    MlpWs w = mlp_ws(ws, n);
    float sq = 0.f;
    if (s < n) {
        const float e = w.out[s] - y[s];
        sq = e * e / (float) n;
        const float dout = 2.f * e / (float) n;
        w.dout[s] = dout;
        float d2[H];
#pragma unroll
        for (int j = 0; j < H; ++j) d2[j] = (w.h2[s * H + j] > 0.f) ? dout * P[MLP_W3 + j] : 0.f;
#pragma unroll
        for (int j = 0; j < H; ++j) w.d2[s * H + j] = d2[j];
        for (int k = 0; k < H; ++k) {
            float t = 0.f;
#pragma unroll
            for (int j = 0; j < H; ++j) t += d2[j] * P[MLP_W2 + k * H + j];
            w.d1[s * H + k] = (w.h1[s * H + k] > 0.f) ? t : 0.f;
        }
    }
    // block partial of the loss (whole waves active: the bound check above only guards the body)
    float v = sq;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    if ((threadIdx.x & (WAVE - 1)) == 0) red[threadIdx.x / WAVE] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int i = 0; i < MLP_NT / WAVE; ++i) t += red[i];
        w.lossp[blockIdx.x] = t;
    }
}

// One thread per parameter: sums over all samples (deterministic, no atomics).
__global__ void __launch_bounds__(MLP_NT) k_mlp_param_grads(const float *X, int64_t n, const float *ws_c,
                                                            float *G) {
    const int p = blockIdx.x * MLP_NT + threadIdx.x;
    if (p >= RANK_PARAMS) return;
    MlpWs w = mlp_ws(const_cast<float *>(ws_c), n);
    float g = 0.f;
    if (p < MLP_B1) {
        const int i = p / H, j = p % H;
        for (int64_t s = 0; s < n; ++s) g += X[s * IN + i] * w.d1[s * H + j];
    } else if (p < MLP_W2) {
        const int j = p - MLP_B1;
        for (int64_t s = 0; s < n; ++s) g += w.d1[s * H + j];
    } else if (p < MLP_B2) {
        const int k = (p - MLP_W2) / H, j = (p - MLP_W2) % H;
        for (int64_t s = 0; s < n; ++s) g += w.h1[s * H + k] * w.d2[s * H + j];
    } else if (p < MLP_W3) {
        const int j = p - MLP_B2;
        for (int64_t s = 0; s < n; ++s) g += w.d2[s * H + j];
    } else if (p < MLP_B3) {
        const int j = p - MLP_W3;
        for (int64_t s = 0; s < n; ++s) g += w.h2[s * H + j] * w.dout[s];
    } else {
        for (int64_t s = 0; s < n; ++s) g += w.dout[s];
    }
    G[p] = g;
}

__global__ void k_mlp_adam(float *P, const float *G, float *M, float *V, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= RANK_PARAMS) return;
    const float g = G[p];
    const float m = b1 * M[p] + (1.f - b1) * g;
    const float v = b2 * V[p] + (1.f - b2) * g * g;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

__global__ void k_mlp_pack_bf16(const float *P, uint16_t *out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < RANK_PARAMS) out[p] = f32_to_bf16(P[p]);
}

// The in-kernel inference routine on rows of X (one wave per sample): for testing it against
// the FP32 network and the CPU oracle.
__global__ void __launch_bounds__(WAVE) k_mlp_predict_bf16(const uint16_t *P, const float *X, int64_t n, float *out) {
    const int64_t s = blockIdx.x;
    if (s >= n) return;
    const float o = rank_mlp_wave(P, X + s * IN);
    if (lane_id() == 0) out[s] = o;
}

hipError_t launch_mlp_forward(hipStream_t st, const float *P, const float *X, int64_t n, float *ws) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned) ((n + MLP_NT - 1) / MLP_NT)), dim3(MLP_NT), 0, st, P, X, n,
                       ws);
    return hipGetLastError();
}

hipError_t launch_mlp_backward(hipStream_t st, const float *P, const float *X, const float *y, int64_t n, float *ws,
                               float *G) {
    if (n <= 0) return hipSuccess;
    const unsigned nb = (unsigned) ((n + MLP_NT - 1) / MLP_NT);
    hipLaunchKernelGGL(k_mlp_deltas, dim3(nb), dim3(MLP_NT), 0, st, P, y, n, ws);
    hipLaunchKernelGGL(k_mlp_param_grads, dim3((RANK_PARAMS + MLP_NT - 1) / MLP_NT), dim3(MLP_NT), 0, st, X, n,
                       (const float *) ws, G);
    return hipGetLastError();
}

float *mlp_loss_partials(float *ws, int64_t n) { return mlp_ws(ws, n).lossp; }
int64_t mlp_loss_blocks(int64_t n) { return (n + MLP_NT - 1) / MLP_NT; }

hipError_t launch_mlp_adam(hipStream_t st, float *P, const float *G, float *M, float *V, float lr, float b1, float b2,
                           float eps, int64_t step) {
    const float c1 = 1.f - powf(b1, (float) step), c2 = 1.f - powf(b2, (float) step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((RANK_PARAMS + 255) / 256), dim3(256), 0, st, P, G, M, V, lr, b1, b2, eps,
                       c1, c2);
    return hipGetLastError();
}

hipError_t launch_mlp_pack_bf16(hipStream_t st, const float *P, uint16_t *out) {
    hipLaunchKernelGGL(k_mlp_pack_bf16, dim3((RANK_PARAMS + 255) / 256), dim3(256), 0, st, P, out);
    return hipGetLastError();
}

hipError_t launch_mlp_predict_bf16(hipStream_t st, const uint16_t *P, const float *X, int64_t n, float *out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_predict_bf16, dim3((unsigned) n), dim3(WAVE), 0, st, P, X, n, out);
    return hipGetLastError();
}

}  // namespace hcamd
