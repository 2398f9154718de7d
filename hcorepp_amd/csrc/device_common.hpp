// Device-side building blocks shared by every kernel of the TLR recompression
// engine: workgroup geometry, wave/workgroup reductions, and small helpers.
// All kernels are written for gfx950 (CDNA4): 64-lane wavefronts, 160 KB LDS
// per CU.  Workgroups are NT = 256 threads (4 waves, one per SIMD).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hcamd {

constexpr int NT = 256;        // threads per workgroup for every engine kernel
constexpr int NWAVES = NT / 64;
constexpr int WAVE = 64;

template <typename T>
struct Eps;
template <>
struct Eps<double> {
    static constexpr double value = 2.220446049250313e-16;
};
template <>
struct Eps<float> {
    static constexpr float value = 1.1920929e-07f;
};

__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, WAVE);
        v = v > o ? v : o;
    }
    return v;
}

// Workgroup-wide sum.  `red` is an LDS scratch of at least NWAVES elements.
// All threads of the workgroup must call it; every thread gets the total.
template <typename T>
__device__ __forceinline__ T wg_sum(T v, T *red) {
    v = wave_sum(v);
    __syncthreads();
    if (lane_id() == 0) red[wave_id()] = v;
    __syncthreads();
    T s = 0;
#pragma unroll
    for (int w = 0; w < NWAVES; ++w) s += red[w];
    return s;
}

// Reduce NV independent partial sums held by every lane of a wave so that,
// on return, lane l (l < NV) holds the full wave sum of value index l.
// Uses the halving "transpose" scheme: log2(64) exchange rounds with a total of
// NV-1 + (6 - log2 NV) shuffles instead of NV * 6.
template <typename T, int NV>
__device__ __forceinline__ T wave_multi_sum(T (&v)[NV]) {
    static_assert((NV & (NV - 1)) == 0 && NV <= 64, "NV must be a power of two <= 64");
    const int l = lane_id();
    int cnt = NV;
    int off = 32;
    // Each round halves the number of live values per lane: lanes with bit `off`
    // set keep the upper half, the others the lower half.
#pragma unroll
    for (int round = 0; round < 6; ++round) {
        if (cnt > 1) {
            const int half = cnt >> 1;
            const bool upper = (l & off) != 0;
#pragma unroll
            for (int i = 0; i < NV / 2; ++i) {
                if (i < half) {
                    T send = upper ? v[i] : v[i + half];
                    T keep = upper ? v[i + half] : v[i];
                    T recv = __shfl_xor(send, off, WAVE);
                    v[i] = keep + recv;
                }
            }
            cnt = half;
        } else {
            v[0] += __shfl_xor(v[0], off, WAVE);
        }
        off >>= 1;
    }
    // After the rounds, lane l holds the sum of value index idx(l), where idx is
    // formed from the lane bits consumed while cnt > 1 (bit 5 is the MSB).
    return v[0];
}

// Value index held by lane l after wave_multi_sum<NV>.
template <int NV>
__device__ __forceinline__ int wave_multi_sum_index(int l) {
    int idx = 0;
    int cnt = NV;
    int off = 32;
    int base = 0;
    (void) base;
    for (int round = 0; round < 6 && cnt > 1; ++round) {
        const int half = cnt >> 1;
        if (l & off) idx += half;
        cnt = half;
        off >>= 1;
    }
    return idx;
}

template <typename T>
__device__ __forceinline__ T dsign(T a, T b) {
    return b >= T(0) ? (a >= T(0) ? a : -a) : (a >= T(0) ? -a : a);
}

}  // namespace hcamd
