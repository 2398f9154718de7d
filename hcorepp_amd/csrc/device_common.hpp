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
// on return, every lane holds the full wave sum of one value index:
// index = lane >> (6 - log2 NV) (lane bit 5 is the index MSB).  Each halving
// round exchanges half of the live values with the partner lane, so the total
// cost is NV - 1 + (6 - log2 NV) shuffles instead of 6 * NV.  Written with
// compile-time recursion so that the value array stays in registers.
template <typename T, int NV, int CNT, int OFF>
struct MultiSum {
    static __device__ __forceinline__ void run(T (&v)[NV], int l) {
        if constexpr (CNT > 1) {
            constexpr int HALF = CNT / 2;
            const bool upper = (l & OFF) != 0;
#pragma unroll
            for (int i = 0; i < HALF; ++i) {
                const T send = upper ? v[i] : v[i + HALF];
                const T keep = upper ? v[i + HALF] : v[i];
                v[i] = keep + __shfl_xor(send, OFF, WAVE);
            }
            MultiSum<T, NV, HALF, OFF / 2>::run(v, l);
        } else if constexpr (OFF >= 1) {
            v[0] += __shfl_xor(v[0], OFF, WAVE);
            MultiSum<T, NV, 1, OFF / 2>::run(v, l);
        }
    }
};

template <typename T, int NV>
__device__ __forceinline__ T wave_multi_sum(T (&v)[NV]) {
    static_assert((NV & (NV - 1)) == 0 && NV <= 64, "NV must be a power of two <= 64");
    MultiSum<T, NV, NV, 32>::run(v, lane_id());
    return v[0];
}

template <int NV>
__host__ __device__ __forceinline__ constexpr int log2_const() {
    return NV <= 1 ? 0 : 1 + log2_const<NV / 2>();
}

// Value index held by lane l after wave_multi_sum<NV>.
template <int NV>
__device__ __forceinline__ int wave_multi_sum_index(int l) {
    return l >> (6 - log2_const<NV>());
}

template <typename T>
__device__ __forceinline__ T dsign(T a, T b) {
    return b >= T(0) ? (a >= T(0) ? a : -a) : (a >= T(0) ? -a : a);
}

}  // namespace hcamd
