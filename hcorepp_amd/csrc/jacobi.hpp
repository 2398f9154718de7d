// Register-tiled one-sided block Jacobi SVD for one workgroup (4 waves).
//
// The matrix J (rows x ncols, column-major) is orthogonalised column-wise:
// J V = Z with V orthogonal (accumulated) and the columns of Z mutually
// orthogonal; sigma_j = |z_j|.  Columns are grouped in blocks of TW/2; a wave
// owns one block pair ("tile" of TW columns) at a time, holding the tile in
// registers (RPL rows per lane).  Within a tile, TW/2 disjoint column pairs
// are rotated per round, their 3*TW/2 dot products reduced with one
// multi-value wave reduction.  Block pairs follow the circle (round-robin)
// ordering; step 0 of every sweep rotates all pairs inside the tile (so the
// intra-block pairs of every block are visited once per sweep), later steps
// only the cross pairs.  The V columns of a tile are loaded with its J columns
// and rotated alongside them, so a tile costs one memory round trip.
//
// Convergence test: |z_p . z_q| <= sqrt(rows) * eps * |z_p| |z_q| (relative,
// as LAPACK's one-sided Jacobi), pairs involving a column whose norm is below
// eps * ||J||_F are not rotated (those columns carry rounding-level
// singular values).  A sweep ends the iteration when every pair it rotated had
// a cosine of at most JacobiStopCos before its rotation: Jacobi converges
// quadratically, so the cosines it leaves are of the order of that bound's
// square, and the sweeps that would only shrink them further are skipped.  The
// forms that accumulate V (the plain, one-update-at-a-time form) stop at
// 0.1 sqrt(eps), which leaves U orthonormal to rounding as LAPACK's SVD does; the
// batched form without V stops at 1e-5 in FP64 (cosines left of order 1e-10, U
// orthonormal to about 1e-10, singular values exact to rounding), one sweep fewer.
#pragma once

#include "device_common.hpp"

namespace hcamd {

// Bits of the rotation summary the Jacobi rounds and tiles return: some pair was rotated (the
// tile must be written back), and some rotated pair was far from orthogonal (the sweep has not
// converged).
enum : int { JROT_APPLIED = 1, JROT_LARGE = 2 };

// Convergence: a sweep is the last when every pair it rotated had, before its rotation, a
// cosine |a_p . a_q| / (|a_p| |a_q|) of at most JacobiStopCos.  The one-sided Jacobi converges
// quadratically, so the cosines such a sweep leaves are of the order of its square, and the
// column norms (the singular values) are exact to rounding; the pairs' angles are not used
// because two columns of nearly equal norm take a large rotation even when nearly orthogonal.
template <typename T>
struct JacobiStopCos;
template <>
struct JacobiStopCos<double> {
    static constexpr double tight = 1.4901161193847656e-09;  // 0.1 * sqrt(2^-52)
    static constexpr double loose = 1e-5;
};
template <>
struct JacobiStopCos<float> {
    static constexpr float tight = 3.4526698e-05f;  // 0.1 * sqrt(2^-23)
    static constexpr float loose = 3.4526698e-05f;
};

// Lane that holds value index j after wave_multi_sum<NV>: index bits map to lane bits 5..(6-log2 NV).
template <int NV>
__device__ __forceinline__ constexpr int multi_sum_src_lane(int j) {
    return j << (6 - log2_const<NV>());
}

// Rotation (c, s) that annihilates apq for the Gram entries (app, aqq, apq) of a column pair,
// or (1, 0) when the pair is already orthogonal to tolerance or involves a numerically zero
// column.  t = tan(theta) is the smaller root of t^2 + 2 zeta t - 1 = 0 with
// zeta = (aqq - app) / (2 apq), written as t = sign(d) e / (|d| + hypot(d, e)) for d = aqq - app,
// e = 2 apq (one division and one square root); c = 1/sqrt(1 + t^2), s = c t.
// TIGHT selects the stop cosine (JacobiStopCos: tight for the forms that accumulate V).
template <typename T, bool TIGHT>
__device__ __forceinline__ void jacobi_rotation(T app, T aqq, T apq, T tol, T floor2, T &c, T &s, int &large) {
    c = T(1);
    s = T(0);
    large = 0;
    if (app > floor2 && aqq > floor2 && fabs(apq) > tol * sqrt(app * aqq)) {
        large = fabs(apq) > (TIGHT ? JacobiStopCos<T>::tight : JacobiStopCos<T>::loose) * sqrt(app * aqq);
        const T d = aqq - app, e = T(2) * apq;
        // scale by a power of two so that d*d + e*e neither overflows nor underflows
        const int ex = ilogb(fmax(fabs(d), fabs(e)));
        const T ds = scalbn(d, -ex), es = scalbn(e, -ex);
        const T h = sqrt(ds * ds + es * es);
        const T t = (d >= T(0) ? es : -es) / (fabs(ds) + h);
        c = T(1) / sqrt(T(1) + t * t);
        s = c * t;
    }
}

// Row count of the V register tile: RPLV = 0 runs the Jacobi without accumulating V (the
// caller recovers the right factor from the left one), the array keeps one unused row.
template <int RPLV>
__host__ __device__ constexpr int vrows() {
    return RPLV > 0 ? RPLV : 1;
}

// One round: TW/2 disjoint pairs of the register tile y.  The 3*TW/2 dot products are
// reduce-scattered across the wave with permlane/DPP exchanges; lane k < TW/2 gathers the three
// sums of pair k, computes its rotation, and every lane reads the rotations back with readlane.
// The rotations are applied to the J tile y and to the matching V columns vt (none for RPLV 0).
// The pairs sit in fixed register slots so that the round loop can stay rolled (unrolled rounds
// of a 16-column tile are tens of KB of code, more than the instruction cache holds next to the
// rest of the kernel): pair k is slots (k, TW - 1 - k) in a full tile (all pairs of its columns) and
// (k, TW/2 + k) in a cross tile (pairs between its two blocks); between rounds the caller moves
// the columns to the next slots (jacobi_rotate_slots), the circle method of round-robin pairing.
// The rotation of a pair does not depend on which of its columns is named first.
template <typename T, int TW, bool FULL>
__host__ __device__ __forceinline__ constexpr int jslot_q(int k) {
    return FULL ? TW - 1 - k : TW / 2 + k;
}

// Tile column held by slot `slot` in round rd of the rolled rounds (jacobi_rotate_slots moves the
// content of slot i + 1 to slot i, wrapping within the rotated range).
template <int TW, bool FULL>
__host__ __device__ __forceinline__ constexpr int jslot_col(int rd, int slot) {
    if (FULL) return slot == TW - 1 ? TW - 1 : (slot + rd) % (TW - 1);
    return slot < TW / 2 ? slot : TW / 2 + (slot - TW / 2 + rd) % (TW / 2);
}

// The value a lane's group partner holds in the other half of its group of G = 64 / NP lanes
// (groups of 8: half-row mirror, 16: row mirror, 32: the odd rows by a permlane16 swap, 64: the
// upper half-wave by a permlane32 swap); lanes in the lower half of a group read the upper half.
template <typename T, int NP>
__device__ __forceinline__ T group_partner(T v) {
    constexpr int G = WAVE / NP;
    if constexpr (G == 8) {
        return dpp_update<DPP_ROW_HALF_MIRROR, 0xF, 0xF>(v, v);
    } else if constexpr (G == 16) {
        return dpp_update<DPP_ROW_MIRROR, 0xF, 0xF>(v, v);
    } else {
        static_assert(G == 32 || G == 64, "pair groups of 8 to 64 lanes");
        T a = v, b = v;
        permlane_swap<G / 2>(a, b);
        return b;
    }
}

template <typename T, int RPL, int RPLV, int TW, bool FULL>
__device__ __forceinline__ int jacobi_round_fixed(T (&y)[RPL][TW], T (&vt)[vrows<RPLV>()][TW], T tol, T floor2) {
    constexpr int NP = TW / 2, G = WAVE / NP;
    const int l = lane_id();
    // the squared norms (2 NP values, pair k's at 2k and 2k + 1) and the cross products (NP values)
    // are reduced separately: both counts are powers of two, so no padded slot takes part.  After
    // the reductions the G-lane group of pair k holds its cross product, the group's lower half
    // its first column's squared norm and the upper half its second's: the group's first lane
    // has all three after one exchange within the group, and computes the pair's rotation.
    T nrm[2 * NP], crs[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int q = jslot_q<T, TW, FULL>(k);
        nrm[2 * k] = nrm[2 * k + 1] = crs[k] = T(0);
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const T a = y[t][k], b = y[t][q];
            nrm[2 * k] += a * a;
            nrm[2 * k + 1] += b * b;
            crs[k] += a * b;
        }
    }
    const T app = wave_multi_sum_fast<T, 2 * NP>(nrm);
    const T apq = wave_multi_sum_fast<T, NP>(crs);
    const T aqq = group_partner<T, NP>(app);
    T c, s;
    int lg;
    jacobi_rotation<T, (RPLV > 0)>(app, aqq, apq, tol, floor2, c, s, lg);
    const unsigned long long lgm = __ballot((l % G) == 0 && lg);
    int any = 0;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const T sk = read_lane(s, k * G);
        if (sk != T(0)) {
            const T ck = read_lane(c, k * G);
            const int q = jslot_q<T, TW, FULL>(k);
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                const T a = y[r][k], b = y[r][q];
                y[r][k] = ck * a - sk * b;
                y[r][q] = sk * a + ck * b;
            }
#pragma unroll
            for (int r = 0; r < RPLV; ++r) {
                const T va = vt[r][k], vb = vt[r][q];
                vt[r][k] = ck * va - sk * vb;
                vt[r][q] = sk * va + ck * vb;
            }
            any |= JROT_APPLIED | (((lgm >> (k * G)) & 1ull) ? JROT_LARGE : 0);
        }
    }
    return any;
}

// Columns to their slots for the next round: a full tile keeps slot TW-1 and rotates slots
// 0..TW-2 down by one, so that round rd pairs column TW-1 with column rd and columns rd + k with
// rd - k (mod TW - 1): after TW - 1 rounds every pair has met once and the columns are back in
// place.  A cross tile rotates the second block's slots TW/2..TW-1 down by one (column k of the
// first block meets column (k + rd) mod TW/2 of the second; back in place after TW/2 rounds).
template <typename T, int ROWS, int TW, bool FULL>
__device__ __forceinline__ void jacobi_rotate_slots(T (&y)[ROWS][TW]) {
    constexpr int FIRST = FULL ? 0 : TW / 2, LAST = FULL ? TW - 2 : TW - 1;
#pragma unroll
    for (int t = 0; t < ROWS; ++t) {
        const T first = y[t][FIRST];
#pragma unroll
        for (int c = FIRST; c < LAST; ++c) y[t][c] = y[t][c + 1];
        y[t][LAST] = first;
    }
}

// Process one tile (block pair) with one wave.  Columns [cI, cI+TW/2) and [cJ, cJ+TW/2)
// of J (rows x ncols, ld ldj; RPL rows per lane) and of V (ncols x ncols, ld ldv; RPLV rows
// per lane) are loaded together, so the tile costs one memory round trip; columns >= ncols
// are zero and never stored.
template <typename T, int RPL, int RPLV, int TW>
__device__ __forceinline__ void jtile_load(T (&y)[RPL][TW], T (&vt)[vrows<RPLV>()][TW], const T *J, int rows,
                                           int ncols, int64_t ldj, const T *V, int64_t ldv, int cI, int cJ) {
    constexpr int B = TW / 2;
    const int l = lane_id();
#pragma unroll
    for (int c = 0; c < TW; ++c) {
        const int gc = (c < B) ? cI + c : cJ + (c - B);
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const int r = l + WAVE * t;
            y[t][c] = (r < rows && gc < ncols) ? J[r + (int64_t) gc * ldj] : T(0);
        }
#pragma unroll
        for (int t = 0; t < RPLV; ++t) {
            const int r = l + WAVE * t;
            vt[t][c] = (r < ncols && gc < ncols) ? V[r + (int64_t) gc * ldv] : T(0);
        }
    }
}

template <typename T, int RPL, int RPLV, int TW>
__device__ __forceinline__ void jtile_store(const T (&y)[RPL][TW], const T (&vt)[vrows<RPLV>()][TW], T *J, int rows,
                                            int ncols, int64_t ldj, T *V, int64_t ldv, int cI, int cJ) {
    constexpr int B = TW / 2;
    const int l = lane_id();
#pragma unroll
    for (int c = 0; c < TW; ++c) {
        const int gc = (c < B) ? cI + c : cJ + (c - B);
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const int r = l + WAVE * t;
            if (r < rows && gc < ncols) J[r + (int64_t) gc * ldj] = y[t][c];
        }
#pragma unroll
        for (int t = 0; t < RPLV; ++t) {
            const int r = l + WAVE * t;
            if (r < ncols && gc < ncols) V[r + (int64_t) gc * ldv] = vt[t][c];
        }
    }
}

// The rounds of a full (all pairs of the tile's columns) or cross (pairs between its two
// blocks) tile; returns the JROT flags.
template <typename T, int RPL, int RPLV, int TW>
__device__ __forceinline__ int jtile_rounds(T (&y)[RPL][TW], T (&vt)[vrows<RPLV>()][TW], bool full, T tol, T floor2) {
    int any = 0;
    if (full) {
#pragma unroll 1
        for (int rd = 0; rd < TW - 1; ++rd) {
            any |= jacobi_round_fixed<T, RPL, RPLV, TW, true>(y, vt, tol, floor2);
            jacobi_rotate_slots<T, RPL, TW, true>(y);
            if constexpr (RPLV > 0) jacobi_rotate_slots<T, RPLV, TW, true>(vt);
        }
    } else {
#pragma unroll 1
        for (int rd = 0; rd < TW / 2; ++rd) {
            any |= jacobi_round_fixed<T, RPL, RPLV, TW, false>(y, vt, tol, floor2);
            jacobi_rotate_slots<T, RPL, TW, false>(y);
            if constexpr (RPLV > 0) jacobi_rotate_slots<T, RPLV, TW, false>(vt);
        }
    }
    return any;
}

// Block pair bp of round-robin step `step` over nbp blocks of B columns: its first columns (cI,
// cJ) and whether it exists (the padding block of an odd block count pairs with nothing).
__device__ __forceinline__ bool jacobi_pair_cols(int bp, int step, int nbp, int B, int ncols, int &cI, int &cJ) {
    int bI, bJ;
    if (bp == 0) {
        bI = nbp - 1;
        bJ = step;
    } else {
        bI = (step + bp) % (nbp - 1);
        bJ = (step - bp + nbp - 1) % (nbp - 1);
    }
    if (bI > bJ) {
        const int t = bI;
        bI = bJ;
        bJ = t;
    }
    cI = bI * B;
    cJ = bJ * B;
    return cI < ncols;
}

// LDS elements wg_block_jacobi needs: a reduction slot per wave and a flag.
template <typename T, int TW>
__host__ __device__ constexpr int block_jacobi_lds_elems() {
    return NWAVES + 2;
}

// Whole-workgroup block Jacobi.  J: rows x ncols (ld ldj, rows <= 64 RPL), V: ncols x ncols
// (ld ldv, ncols <= 64 RPLV) must hold the identity on entry.
// `gbuf` is LDS of block_jacobi_lds_elems<T, TW>() elements (all the LDS this routine touches,
// so it may be called out of line).  Returns the sweeps used (max_sweeps when it did not
// converge).
template <typename T, int RPL, int RPLV, int TW, bool PREFETCH = false>
__device__ __forceinline__ int wg_block_jacobi(T *J, int rows, int ncols, int64_t ldj, T *V, int64_t ldv,
                                               int max_sweeps, T *gbuf) {
    constexpr int B = TW / 2;
    T *s_red = gbuf;
    int &s_rot = *reinterpret_cast<int *>(s_red + NWAVES);
    T part = T(0);
    for (int c = 0; c < ncols; ++c)
        for (int r = threadIdx.x; r < rows; r += NT) {
            const T a = J[r + (int64_t) c * ldj];
            part += a * a;
        }
    const T fro2 = wg_sum(part, s_red);
    const T floor2 = Eps<T>::value * Eps<T>::value * fro2;
    const T tol = sqrt(T(rows)) * Eps<T>::value;
    const int nblk = (ncols + B - 1) / B;
    const int nbp = nblk + (nblk & 1);
    int sweep = 0;
    if (ncols <= 1) return 0;
    for (; sweep < max_sweeps; ++sweep) {
        if (threadIdx.x == 0) s_rot = 0;
        __syncthreads();
        int any = 0;
        for (int step = 0; step < nbp - 1; ++step) {
            // the wave's tiles of one step are disjoint: with PREFETCH the next one is loaded
            // while the current one's rounds run (for J in global memory, where a load is a
            // memory round trip)
            int bp = wave_id(), cI = 0, cJ = 0;
            while (bp < nbp / 2 && !jacobi_pair_cols(bp, step, nbp, B, ncols, cI, cJ)) bp += NWAVES;
            if constexpr (PREFETCH) {
                T yn[RPL][TW], vn[vrows<RPLV>()][TW];
                if (bp < nbp / 2) jtile_load<T, RPL, RPLV, TW>(yn, vn, J, rows, ncols, ldj, V, ldv, cI, cJ);
                while (bp < nbp / 2) {
                    T y[RPL][TW], vt[vrows<RPLV>()][TW];
#pragma unroll
                    for (int t = 0; t < RPL; ++t)
#pragma unroll
                        for (int c = 0; c < TW; ++c) y[t][c] = yn[t][c];
#pragma unroll
                    for (int t = 0; t < vrows<RPLV>(); ++t)
#pragma unroll
                        for (int c = 0; c < TW; ++c) vt[t][c] = vn[t][c];
                    const int c0 = cI, c1 = cJ;
                    bp += NWAVES;
                    while (bp < nbp / 2 && !jacobi_pair_cols(bp, step, nbp, B, ncols, cI, cJ)) bp += NWAVES;
                    if (bp < nbp / 2) jtile_load<T, RPL, RPLV, TW>(yn, vn, J, rows, ncols, ldj, V, ldv, cI, cJ);
                    const int a = jtile_rounds<T, RPL, RPLV, TW>(y, vt, step == 0, tol, floor2);
                    if (a & JROT_APPLIED) jtile_store<T, RPL, RPLV, TW>(y, vt, J, rows, ncols, ldj, V, ldv, c0, c1);
                    any |= a;
                }
            } else {
                while (bp < nbp / 2) {
                    T y[RPL][TW], vt[vrows<RPLV>()][TW];
                    jtile_load<T, RPL, RPLV, TW>(y, vt, J, rows, ncols, ldj, V, ldv, cI, cJ);
                    const int a = jtile_rounds<T, RPL, RPLV, TW>(y, vt, step == 0, tol, floor2);
                    if (a & JROT_APPLIED) jtile_store<T, RPL, RPLV, TW>(y, vt, J, rows, ncols, ldj, V, ldv, cI, cJ);
                    any |= a;
                    bp += NWAVES;
                    while (bp < nbp / 2 && !jacobi_pair_cols(bp, step, nbp, B, ncols, cI, cJ)) bp += NWAVES;
                }
            }
            __syncthreads();
        }
        if ((any & JROT_LARGE) && lane_id() == 0) atomicOr(&s_rot, 1);
        __syncthreads();
        const int rot = s_rot;
        __syncthreads();
        if (!rot) break;
    }
    return sweep;
}

// One tile (block pair) of a staged super-block pair, in LDS (ld rows, 2 W columns).
template <typename T, int RPL, int TW>
__device__ __forceinline__ int staged_tile(T *Jl, int rows, int lcols, int lI, int lJ, bool full, T tol, T floor2) {
    T y[RPL][TW], vt[vrows<0>()][TW];
    jtile_load<T, RPL, 0, TW>(y, vt, Jl, rows, lcols, rows, nullptr, 0, lI, lJ);
    const int a = jtile_rounds<T, RPL, 0, TW>(y, vt, full, tol, floor2);
    if (a & JROT_APPLIED) jtile_store<T, RPL, 0, TW>(y, vt, Jl, rows, lcols, rows, nullptr, 0, lI, lJ);
    return a;
}

// Block Jacobi without V for a core too large for LDS, staged through it by super-blocks.  The
// columns form super-blocks of W (a multiple of B = TW / 2) and each sweep runs the round-robin
// ordering over super-blocks.  For a super-block pair (SI, SJ), both are copied into the LDS
// buffer Jl (ld rows; SI in local columns [0, W), SJ in [W, 2 W), columns past ncols zero, so they
// never rotate), the four waves rotate its block pairs there round by round, and the pair is
// copied back.  In the first super-step every super-block appears once, and its pair runs the
// round-robin over all 2 W / B local blocks (in whose first round each block appears once, as a
// full tile: its internal column pairs); later super-steps visit only the block pairs across SI
// and SJ, W / B rounds.  So every column pair is visited once per sweep, as in wg_block_jacobi,
// while the core passes through memory once per super-step instead of once per block step:
// W / B times fewer passes.  J: rows x ncols (ld rows, rows <= 64 RPL); Jl: 2 W rows elements of
// LDS; gbuf: block_jacobi_lds_elems<T, TW>() elements of LDS.  Returns the sweeps used.
template <typename T, int RPL, int TW>
__device__ __forceinline__ int wg_staged_jacobi(T *J, int rows, int ncols, T *Jl, int W, int max_sweeps, T *gbuf) {
    constexpr int B = TW / 2;
    T *s_red = gbuf;
    int &s_rot = *reinterpret_cast<int *>(s_red + NWAVES);
    const int64_t ne = (int64_t) rows * ncols;
    T part = T(0);
    for (int64_t e = threadIdx.x; e < ne; e += NT) {
        const T a = J[e];
        part += a * a;
    }
    const T fro2 = wg_sum(part, s_red);
    const T floor2 = Eps<T>::value * Eps<T>::value * fro2;
    const T tol = sqrt(T(rows)) * Eps<T>::value;
    if (ncols <= 1) return 0;
    const int a = W / B, lcols = 2 * W;
    const int ns = (ncols + W - 1) / W, nsp = ns + (ns & 1);
    const int64_t sw = (int64_t) W * rows;
    int sweep = 0;
    for (; sweep < max_sweeps; ++sweep) {
        if (threadIdx.x == 0) s_rot = 0;
        __syncthreads();
        int any = 0;
        for (int ss = 0; ss < nsp - 1; ++ss) {
            for (int sp = 0; sp < nsp / 2; ++sp) {
                int cI, cJ;
                if (!jacobi_pair_cols(sp, ss, nsp, W, ncols, cI, cJ)) continue;
                const bool solo = cJ >= ncols;  // paired with the padding super-block
                if (ss > 0 && solo) continue;
                const int64_t nI = (int64_t) ((ncols - cI) < W ? (ncols - cI) : W) * rows;
                const int64_t nJ = solo ? 0 : (int64_t) ((ncols - cJ) < W ? (ncols - cJ) : W) * rows;
                wg_copy(Jl, J + (int64_t) cI * rows, nI);
                if (nJ > 0) wg_copy(Jl + sw, J + (int64_t) cJ * rows, nJ);
                for (int64_t e = nI + threadIdx.x; e < sw; e += NT) Jl[e] = T(0);
                for (int64_t e = nJ + threadIdx.x; e < sw; e += NT) Jl[sw + e] = T(0);
                __syncthreads();
                // local block k holds real columns when its first global column is below ncols;
                // a tile is skipped when no pair of it can rotate (a cross tile with an empty
                // block, a full tile with two)
                const int realI = (int) ((nI / rows + B - 1) / B), realJ = (int) ((nJ / rows + B - 1) / B);
                auto real = [&](int lc) { return lc < W ? lc / B < realI : (lc - W) / B < realJ; };
                if (ss == 0) {
                    // the round-robin over the local blocks: all 2 a of them, or the a of SI alone
                    const int nb = solo ? a : 2 * a;
                    for (int t = 0; t < nb - 1; ++t) {
                        for (int p = wave_id(); p < nb / 2; p += NWAVES) {
                            int lI, lJ;
                            jacobi_pair_cols(p, t, nb, B, lcols, lI, lJ);
                            const bool full = t == 0;
                            if (full ? (real(lI) || real(lJ)) : (real(lI) && real(lJ)))
                                any |= staged_tile<T, RPL, TW>(Jl, rows, lcols, lI, lJ, full, tol, floor2);
                        }
                        __syncthreads();
                    }
                } else {
                    for (int t = 0; t < a; ++t) {
                        for (int p = wave_id(); p < a; p += NWAVES) {
                            int q = p + t;
                            if (q >= a) q -= a;
                            if (real(p * B) && real(W + q * B))
                                any |= staged_tile<T, RPL, TW>(Jl, rows, lcols, p * B, W + q * B, false, tol, floor2);
                        }
                        __syncthreads();
                    }
                }
                wg_copy(J + (int64_t) cI * rows, Jl, nI);
                if (nJ > 0) wg_copy(J + (int64_t) cJ * rows, Jl + sw, nJ);
                __syncthreads();
            }
        }
        if ((any & JROT_LARGE) && lane_id() == 0) atomicOr(&s_rot, 1);
        __syncthreads();
        const int rot = s_rot;
        __syncthreads();
        if (!rot) break;
    }
    return sweep;
}

// Super-block width of wg_staged_jacobi for `rows` rows, blocks of B columns and an LDS buffer of
// `cap` elements: the widest with two super-blocks in the buffer whose block count is a multiple
// of the wave count (every inner round then gives each wave the same number of tiles); 0 when
// not even NWAVES blocks fit.
__host__ __device__ inline int staged_width(int rows, int B, int64_t cap) {
    const int64_t a = cap / (2 * (int64_t) (rows > 0 ? rows : 1)) / B / NWAVES * NWAVES;
    return a >= NWAVES ? (int) (a * B) : 0;
}

// ---------------------------------------------------------------------------
// Workgroup-cooperative variant: the four waves work on the same tile of TW columns, each
// owning an interleaved quarter of the rows of J (row lane + 64 (wave + 4 t), t < RPL) and of V
// (t < RPLV).  A round reduce-scatters each wave's partial dot products, exchanges them
// through LDS (double-buffered by round parity, one barrier per round) and every wave derives
// the same rotations from the same sums taken in the same order.  A wave only ever reads the
// rows it wrote, so consecutive tiles need no barrier.  Wider tiles than one wave's registers
// allow mean fewer passes over J and V per sweep (traffic ~ ncols / (TW / 2) passes).
// ---------------------------------------------------------------------------
template <typename T, int RPL, int RPLV, int TW, bool FULL>
__device__ __forceinline__ int coop_round(T (&y)[RPL][TW], T (&vt)[RPLV][TW], T tol, T floor2, T *red, int parity) {
    constexpr int NP = TW / 2;
    constexpr int NV = (3 * NP <= 4) ? 4 : (3 * NP <= 8) ? 8 : (3 * NP <= 16) ? 16 : 32;
    const int l = lane_id(), w = wave_id();
    T part[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) part[k] = T(0);
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int p = k, q = jslot_q<T, TW, FULL>(k);
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const T a = y[t][p], b = y[t][q];
            part[3 * k] += a * a;
            part[3 * k + 1] += b * b;
            part[3 * k + 2] += a * b;
        }
    }
    const T tot = wave_multi_sum_fast<T, NV>(part);
    T *buf = red + parity * (NWAVES * NV);
    const int idx = wave_multi_sum_index<NV>(l);
    if (l == multi_sum_src_lane<NV>(idx)) buf[w * NV + idx] = tot;
    __syncthreads();
    T c = T(1), s = T(0);
    int lg = 0;
    if (l < NP) {
        T app = T(0), aqq = T(0), apq = T(0);
#pragma unroll
        for (int w2 = 0; w2 < NWAVES; ++w2) {
            app += buf[w2 * NV + 3 * l];
            aqq += buf[w2 * NV + 3 * l + 1];
            apq += buf[w2 * NV + 3 * l + 2];
        }
        jacobi_rotation<T, (RPLV > 0)>(app, aqq, apq, tol, floor2, c, s, lg);
    }
    const unsigned long long lgm = __ballot(lg);
    int any = 0;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const T sk = read_lane(s, k);
        if (sk != T(0)) {
            const T ck = read_lane(c, k);
            const int p = k, q = jslot_q<T, TW, FULL>(k);
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                const T a = y[r][p], b = y[r][q];
                y[r][p] = ck * a - sk * b;
                y[r][q] = sk * a + ck * b;
            }
#pragma unroll
            for (int r = 0; r < RPLV; ++r) {
                const T va = vt[r][p], vb = vt[r][q];
                vt[r][p] = ck * va - sk * vb;
                vt[r][q] = sk * va + ck * vb;
            }
            any |= JROT_APPLIED | (((lgm >> k) & 1ull) ? JROT_LARGE : 0);
        }
    }
    return any;
}

template <typename T, int TW>
__host__ __device__ constexpr int coop_jacobi_lds_elems() {
    return 2 * NWAVES * ((3 * (TW / 2) <= 4) ? 4 : (3 * (TW / 2) <= 8) ? 8 : (3 * (TW / 2) <= 16) ? 16 : 32) +
           NWAVES + 2;
}

template <typename T, int RPL, int RPLV, int TW>
__device__ int coop_jacobi_tile(T *J, int rows, int ncols, int64_t ldj, T *V, int64_t ldv, int cI, int cJ, bool full,
                                T tol, T floor2, T *red, int &parity) {
    constexpr int B = TW / 2;
    const int l = lane_id(), w = wave_id();
    T y[RPL][TW], vt[RPLV][TW];
#pragma unroll
    for (int c = 0; c < TW; ++c) {
        const int gc = (c < B) ? cI + c : cJ + (c - B);
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const int r = l + WAVE * (w + NWAVES * t);
            y[t][c] = (r < rows && gc < ncols) ? J[r + (int64_t) gc * ldj] : T(0);
        }
#pragma unroll
        for (int t = 0; t < RPLV; ++t) {
            const int r = l + WAVE * (w + NWAVES * t);
            vt[t][c] = (r < ncols && gc < ncols) ? V[r + (int64_t) gc * ldv] : T(0);
        }
    }
    int any = 0;
    if (full) {
#pragma unroll 1
        for (int rd = 0; rd < TW - 1; ++rd) {
            any |= coop_round<T, RPL, RPLV, TW, true>(y, vt, tol, floor2, red, parity);
            jacobi_rotate_slots<T, RPL, TW, true>(y);
            if constexpr (RPLV > 0) jacobi_rotate_slots<T, RPLV, TW, true>(vt);
            parity ^= 1;
        }
    } else {
#pragma unroll 1
        for (int rd = 0; rd < TW / 2; ++rd) {
            any |= coop_round<T, RPL, RPLV, TW, false>(y, vt, tol, floor2, red, parity);
            jacobi_rotate_slots<T, RPL, TW, false>(y);
            if constexpr (RPLV > 0) jacobi_rotate_slots<T, RPLV, TW, false>(vt);
            parity ^= 1;
        }
    }
    if (any & JROT_APPLIED) {
#pragma unroll
        for (int c = 0; c < TW; ++c) {
            const int gc = (c < B) ? cI + c : cJ + (c - B);
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                const int r = l + WAVE * (w + NWAVES * t);
                if (r < rows && gc < ncols) J[r + (int64_t) gc * ldj] = y[t][c];
            }
#pragma unroll
            for (int t = 0; t < RPLV; ++t) {
                const int r = l + WAVE * (w + NWAVES * t);
                if (r < ncols && gc < ncols) V[r + (int64_t) gc * ldv] = vt[t][c];
            }
        }
    }
    return any;
}

// ---------------------------------------------------------------------------
// Cooperative tile with deferred V: the rotations of the J tile are recorded in LDS and applied
// to the V columns once the tile's rounds are done, streaming V rows through registers (one row
// of the tile per thread at a time).  Only the J tile stays in registers, so 16-column tiles fit
// for up to 768 rows (RPL <= 3) without spilling.  Used by the multi-workgroup Jacobi, where
// each workgroup owns one tile per launch.
// ---------------------------------------------------------------------------
template <typename T, int TW>
__host__ __device__ constexpr int coop_rec_lds_elems() {
    return coop_jacobi_lds_elems<T, TW>() + 2 * (TW - 1) * (TW / 2);
}

template <typename T, int RPL, int TW, bool FULL>
__device__ __forceinline__ int coop_round_rec(T (&y)[RPL][TW], int rd, T tol, T floor2, T *red, int parity,
                                              T *cs) {
    constexpr int NP = TW / 2;
    constexpr int NV = (3 * NP <= 4) ? 4 : (3 * NP <= 8) ? 8 : (3 * NP <= 16) ? 16 : 32;
    const int l = lane_id(), w = wave_id();
    T part[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) part[k] = T(0);
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const int p = k, q = jslot_q<T, TW, FULL>(k);
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const T a = y[t][p], b = y[t][q];
            part[3 * k] += a * a;
            part[3 * k + 1] += b * b;
            part[3 * k + 2] += a * b;
        }
    }
    const T tot = wave_multi_sum_fast<T, NV>(part);
    T *buf = red + parity * (NWAVES * NV);
    const int idx = wave_multi_sum_index<NV>(l);
    if (l == multi_sum_src_lane<NV>(idx)) buf[w * NV + idx] = tot;
    __syncthreads();
    T c = T(1), s = T(0);
    int lg = 0;
    if (l < NP) {
        T app = T(0), aqq = T(0), apq = T(0);
#pragma unroll
        for (int w2 = 0; w2 < NWAVES; ++w2) {
            app += buf[w2 * NV + 3 * l];
            aqq += buf[w2 * NV + 3 * l + 1];
            apq += buf[w2 * NV + 3 * l + 2];
        }
        jacobi_rotation<T, true>(app, aqq, apq, tol, floor2, c, s, lg);
        if (w == 0) {
            cs[2 * (rd * NP + l)] = c;
            cs[2 * (rd * NP + l) + 1] = s;
        }
    }
    const unsigned long long lgm = __ballot(lg);
    int any = 0;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const T sk = read_lane(s, k);
        if (sk != T(0)) {
            const T ck = read_lane(c, k);
            const int p = k, q = jslot_q<T, TW, FULL>(k);
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                const T a = y[r][p], b = y[r][q];
                y[r][p] = ck * a - sk * b;
                y[r][q] = sk * a + ck * b;
            }
            any |= JROT_APPLIED | (((lgm >> k) & 1ull) ? JROT_LARGE : 0);
        }
    }
    return any;
}

// Apply the recorded rotations of one tile (rounds of the FULL or cross ordering) to the V
// columns [cI, cI+TW/2) and [cJ, cJ+TW/2) (V: ncols x ncols, ld ldv), one V row per thread.
template <typename T, int TW, bool FULL>
__device__ __forceinline__ void coop_apply_v(T *V, int ncols, int64_t ldv, int cI, int cJ, const T *cs) {
    constexpr int B = TW / 2, NP = TW / 2, ROUNDS = FULL ? TW - 1 : TW / 2;
#pragma unroll 1
    for (int r = threadIdx.x; r < ncols; r += NT) {
        T v[TW];
#pragma unroll
        for (int c = 0; c < TW; ++c) {
            const int gc = (c < B) ? cI + c : cJ + (c - B);
            v[c] = gc < ncols ? V[r + (int64_t) gc * ldv] : T(0);
        }
#pragma unroll
        for (int rd = 0; rd < ROUNDS; ++rd) {
#pragma unroll
            for (int k = 0; k < NP; ++k) {
                const T sk = cs[2 * (rd * NP + k) + 1];
                const T ck = cs[2 * (rd * NP + k)];
                const int p = jslot_col<TW, FULL>(rd, k), q = jslot_col<TW, FULL>(rd, jslot_q<T, TW, FULL>(k));
                const T a = v[p], b = v[q];
                v[p] = ck * a - sk * b;
                v[q] = sk * a + ck * b;
            }
            // keep the next round's rotation loads from being scheduled early (all would be live)
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int c = 0; c < TW; ++c) {
            const int gc = (c < B) ? cI + c : cJ + (c - B);
            if (gc < ncols) V[r + (int64_t) gc * ldv] = v[c];
        }
    }
}

// One tile of the multi-workgroup Jacobi; `lds` holds coop_rec_lds_elems<T, TW>() elements.
// Returns whether any pair was rotated (uniform over the workgroup).
template <typename T, int RPL, int TW>
__device__ __forceinline__ int coop_jacobi_tile_rec(T *J, int rows, int ncols, int64_t ldj, T *V, int64_t ldv, int cI, int cJ,
                                    bool full, T tol, T floor2, T *lds) {
    constexpr int B = TW / 2;
    T *red = lds;
    T *cs = lds + coop_jacobi_lds_elems<T, TW>();
    const int l = lane_id(), w = wave_id();
    T y[RPL][TW];
#pragma unroll
    for (int c = 0; c < TW; ++c) {
        const int gc = (c < B) ? cI + c : cJ + (c - B);
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const int r = l + WAVE * (w + NWAVES * t);
            y[t][c] = (r < rows && gc < ncols) ? J[r + (int64_t) gc * ldj] : T(0);
        }
    }
    int any = 0, parity = 0;
    if (full) {
#pragma unroll 1
        for (int rd = 0; rd < TW - 1; ++rd) {
            any |= coop_round_rec<T, RPL, TW, true>(y, rd, tol, floor2, red, parity, cs);
            jacobi_rotate_slots<T, RPL, TW, true>(y);
            parity ^= 1;
        }
    } else {
#pragma unroll 1
        for (int rd = 0; rd < TW / 2; ++rd) {
            any |= coop_round_rec<T, RPL, TW, false>(y, rd, tol, floor2, red, parity, cs);
            jacobi_rotate_slots<T, RPL, TW, false>(y);
            parity ^= 1;
        }
    }
    if (any & JROT_APPLIED) {
#pragma unroll
        for (int c = 0; c < TW; ++c) {
            const int gc = (c < B) ? cI + c : cJ + (c - B);
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                const int r = l + WAVE * (w + NWAVES * t);
                if (r < rows && gc < ncols) J[r + (int64_t) gc * ldj] = y[t][c];
            }
        }
        __syncthreads();  // the last round's rotations in cs are visible to every wave
        if (full) coop_apply_v<T, TW, true>(V, ncols, ldv, cI, cJ, cs);
        else coop_apply_v<T, TW, false>(V, ncols, ldv, cI, cJ, cs);
    }
    return any;
}

// Whole-workgroup block Jacobi with cooperative tiles: rows <= 256 RPL, ncols <= 256 RPLV.
// Same contract as wg_block_jacobi; `gbuf` holds coop_jacobi_lds_elems<T, TW>() elements.
template <typename T, int RPL, int RPLV, int TW>
__device__ int wg_coop_jacobi(T *J, int rows, int ncols, int64_t ldj, T *V, int64_t ldv, int max_sweeps, T *gbuf) {
    constexpr int B = TW / 2;
    constexpr int NV = (3 * B <= 4) ? 4 : (3 * B <= 8) ? 8 : (3 * B <= 16) ? 16 : 32;
    T *red = gbuf;                                // 2 x NWAVES x NV
    T *s_red = gbuf + 2 * NWAVES * NV;            // NWAVES
    T part = T(0);
    for (int c = 0; c < ncols; ++c)
        for (int r = threadIdx.x; r < rows; r += NT) {
            const T a = J[r + (int64_t) c * ldj];
            part += a * a;
        }
    const T fro2 = wg_sum(part, s_red);
    const T floor2 = Eps<T>::value * Eps<T>::value * fro2;
    const T tol = sqrt(T(rows)) * Eps<T>::value;
    const int nblk = (ncols + B - 1) / B;
    const int nbp = nblk + (nblk & 1);
    int parity = 0;
    int sweep = 0;
    if (ncols <= 1) return 0;
    for (; sweep < max_sweeps; ++sweep) {
        int any = 0;
        for (int step = 0; step < nbp - 1; ++step) {
            for (int bp = 0; bp < nbp / 2; ++bp) {
                int bI, bJ;
                if (bp == 0) {
                    bI = nbp - 1;
                    bJ = step;
                } else {
                    bI = (step + bp) % (nbp - 1);
                    bJ = (step - bp + nbp - 1) % (nbp - 1);
                }
                if (bI > bJ) {
                    const int t = bI;
                    bI = bJ;
                    bJ = t;
                }
                if (bI * B >= ncols) continue;
                any |= coop_jacobi_tile<T, RPL, RPLV, TW>(J, rows, ncols, ldj, V, ldv, bI * B, bJ * B, step == 0,
                                                          tol, floor2, red, parity);
            }
        }
        // every wave derived the same rotations, so `any` agrees across the workgroup
        if (!(any & JROT_LARGE)) break;
    }
    __syncthreads();
    return sweep;
}

}  // namespace hcamd
