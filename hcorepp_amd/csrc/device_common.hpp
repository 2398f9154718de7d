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

// A thread's walk over the elements start, start + step, ... of a column-major array of `rows`
// rows (rows > 0), as (row i, column c) pairs: the divisions are done once, at set-up, and each
// step adds the step's row and column parts, carrying one column when the row wraps.  The walk
// covers the first `cols` columns while c < cols.
struct ColWalk {
    int i, c, di, dc, rows;
    __device__ __forceinline__ ColWalk(int start, int step, int rows_)
        : i(start % rows_), c(start / rows_), di(step % rows_), dc(step / rows_), rows(rows_) {}
    __device__ __forceinline__ void next() {
        i += di;
        c += dc;
        if (i >= rows) {
            i -= rows;
            ++c;
        }
    }
};


template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, WAVE);
        v = v > o ? v : o;
    }
    return v;
}

// Wave-wide sum (every lane gets the total); needs the whole wave active.
template <typename T>
__device__ __forceinline__ T wave_sum(T v);

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

// ---------------------------------------------------------------------------
// Cross-lane exchanges on 32- or 64-bit values without the LDS crossbar: DPP row
// permutations (row_mirror, row_half_mirror, quad_perm) and gfx950's permlane swaps.
// ---------------------------------------------------------------------------
template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ double dpp_update(double old, double src) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(src), CTRL, ROW_MASK, BANK_MASK,
                                               false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(src), CTRL, ROW_MASK, BANK_MASK,
                                               false);
    return __hiloint2double(hi, lo);
}
template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ float dpp_update(float old, float src) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL, ROW_MASK,
                                                      BANK_MASK, false));
}

// v_permlane32_swap / v_permlane16_swap on (a, b): with W = 32 (16), the upper half-wave (odd
// rows) of a is exchanged with the lower half-wave (even rows) of b.  Afterwards
// a = [a_lo | b_lo] and b = [a_hi | b_hi] per lane block, so a + b sums each value with its
// partner's copy, the lower block keeping value a and the upper block value b.
template <int W>
__device__ __forceinline__ void permlane_swap(double &a, double &b) {
    if constexpr (W == 32) {
        auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(a), __double2loint(b), false, false);
        auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(a), __double2hiint(b), false, false);
        a = __hiloint2double((int) hi[0], (int) lo[0]);
        b = __hiloint2double((int) hi[1], (int) lo[1]);
    } else {
        auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(a), __double2loint(b), false, false);
        auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(a), __double2hiint(b), false, false);
        a = __hiloint2double((int) hi[0], (int) lo[0]);
        b = __hiloint2double((int) hi[1], (int) lo[1]);
    }
}
template <int W>
__device__ __forceinline__ void permlane_swap(float &a, float &b) {
    if constexpr (W == 32) {
        auto r = __builtin_amdgcn_permlane32_swap(__float_as_int(a), __float_as_int(b), false, false);
        a = __int_as_float((int) r[0]);
        b = __int_as_float((int) r[1]);
    } else {
        auto r = __builtin_amdgcn_permlane16_swap(__float_as_int(a), __float_as_int(b), false, false);
        a = __int_as_float((int) r[0]);
        b = __int_as_float((int) r[1]);
    }
}

constexpr int DPP_QUAD_XOR1 = 0xB1;    // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_XOR2 = 0x4E;    // quad_perm [2,3,0,1]
constexpr int DPP_ROW_MIRROR = 0x140;  // lane i <- lane 15-i within a row of 16
constexpr int DPP_ROW_HALF_MIRROR = 0x141;  // lane i <- lane 7-i within each half row

// One reduce-scatter level of the wave multi-value sum on bit BIT of the lane id: on return
// v[i] (i < CNT/2) holds, for lanes with that bit clear, the pair sum of value i, and for
// lanes with it set, the pair sum of value i + CNT/2.
template <typename T, int NV, int CNT, int BIT>
__device__ __forceinline__ void multi_sum_level(T (&v)[NV], int l) {
    constexpr int H = CNT / 2;
#pragma unroll
    for (int i = 0; i < H; ++i) {
        T a = v[i], b = v[i + H];
        if constexpr (BIT == 5) {
            permlane_swap<32>(a, b);
            v[i] = a + b;
        } else if constexpr (BIT == 4) {
            permlane_swap<16>(a, b);
            v[i] = a + b;
        } else if constexpr (BIT == 3) {
            // lanes 0-7 of a row take the partner's a; lanes 8-15 the partner's b
            const T x = dpp_update<DPP_ROW_MIRROR, 0xF, 0x3>(b, a);
            const T y = dpp_update<DPP_ROW_MIRROR, 0xF, 0xC>(a, b);
            v[i] = x + y;
        } else if constexpr (BIT == 2) {
            const T x = dpp_update<DPP_ROW_HALF_MIRROR, 0xF, 0x5>(b, a);
            const T y = dpp_update<DPP_ROW_HALF_MIRROR, 0xF, 0xA>(a, b);
            v[i] = x + y;
        } else {
            constexpr int CTRL = BIT == 1 ? DPP_QUAD_XOR2 : DPP_QUAD_XOR1;
            const T ta = dpp_update<CTRL, 0xF, 0xF>(a, a);
            const T tb = dpp_update<CTRL, 0xF, 0xF>(b, b);
            v[i] = (l & (1 << BIT)) ? b + tb : a + ta;
        }
    }
}

template <typename T, int NV, int CNT, int BIT>
struct MultiSumFast {
    static __device__ __forceinline__ void run(T (&v)[NV], int l) {
        if constexpr (BIT >= 0) {
            if constexpr (CNT > 1) {
                multi_sum_level<T, NV, CNT, BIT>(v, l);
                MultiSumFast<T, NV, CNT / 2, BIT - 1>::run(v, l);
            } else {
                // one value left: plain butterfly sum on the remaining (low) lane bits
                constexpr int CTRL = BIT == 1 ? DPP_QUAD_XOR2 : DPP_QUAD_XOR1;
                if constexpr (BIT <= 1) {
                    v[0] += dpp_update<CTRL, 0xF, 0xF>(v[0], v[0]);
                } else if constexpr (BIT == 2) {
                    v[0] += dpp_update<DPP_ROW_HALF_MIRROR, 0xF, 0xF>(v[0], v[0]);
                } else if constexpr (BIT == 3) {
                    v[0] += dpp_update<DPP_ROW_MIRROR, 0xF, 0xF>(v[0], v[0]);
                } else {
                    T a = v[0], b = v[0];
                    permlane_swap<BIT == 5 ? 32 : 16>(a, b);
                    v[0] = a + b;
                }
                MultiSumFast<T, NV, 1, BIT - 1>::run(v, l);
            }
        }
    }
};

// Same result convention as wave_multi_sum (lane l holds the total of value index
// l >> (6 - log2 NV)), computed with permlane swaps and DPP instead of ds_bpermute.
template <typename T, int NV>
__device__ __forceinline__ T wave_multi_sum_fast(T (&v)[NV]) {
    static_assert((NV & (NV - 1)) == 0 && NV <= 64, "NV must be a power of two <= 64");
    MultiSumFast<T, NV, NV, 5>::run(v, lane_id());
    return v[0];
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    T a[1] = {v};
    return wave_multi_sum_fast<T, 1>(a);
}

// Value of lane `src` (a compile-time or wave-uniform lane index) in every lane.
__device__ __forceinline__ double read_lane(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float read_lane(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
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

// Accumulator of one FP64 matrix-core operation (v_mfma_f64_16x16x4f64): four doubles per lane.
typedef double hcamd_doublex4 __attribute__((ext_vector_type(4)));

// dst[e] = src[e] for e in [0, n) by the whole workgroup, four loads in flight per thread.
template <typename T>
__device__ __forceinline__ void wg_copy(T *dst, const T *src, int64_t n) {
    int64_t e = threadIdx.x;
    for (; e + 3 * NT < n; e += 4 * NT) {
        const T a = src[e], b = src[e + NT], c = src[e + 2 * NT], d = src[e + 3 * NT];
        dst[e] = a;
        dst[e + NT] = b;
        dst[e + 2 * NT] = c;
        dst[e + 3 * NT] = d;
    }
    for (; e < n; e += NT) dst[e] = src[e];
}

template <typename T>
__device__ __forceinline__ T dsign(T a, T b) {
    return b >= T(0) ? (a >= T(0) ? a : -a) : (a >= T(0) ? -a : a);
}

}  // namespace hcamd
