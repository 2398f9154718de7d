// HIP kernels of the TLR recompression engine (gfx950).  Each kernel is a thin
// wrapper that maps workgroups onto the workgroup-level phases of
// recompress.hpp / gemm.hpp; the batched persistent kernel runs every phase of
// one update in one workgroup and pulls updates from a device work queue.
#include "kernels.hpp"

namespace hcamd {


// Phase timestamps of the batched kernel (100 MHz constant clock), recorded only in builds
// with HCAMD_PHASE_TIMING; read back with hcorepp_batch_phase_times.
#ifdef HCAMD_PHASE_TIMING
#define HCAMD_TSTAMP() ((uint64_t) wall_clock64())
#else
#define HCAMD_TSTAMP() ((uint64_t) 0)
#endif

template <typename T>
__global__ void HCAMD_KERNEL k_gemm(int M, int N, int Kd, T alpha, const T *A, int64_t lda, int ta,
                                             const T *B, int64_t ldb, int tb, T beta, T *C, int64_t ldc) {
    __shared__ T lds[GT_LDS_ELEMS];
    wg_gemm_tile<T>(M, N, Kd, alpha, A, lda, ta != 0, B, ldb, tb != 0, beta, C, ldc, blockIdx.x, blockIdx.y, lds);
}

template <typename T>
__global__ void HCAMD_KERNEL k_stack(RecompDesc<T> d) {
    ph_stack<T>(d, blockIdx.x, gridDim.x);
}

template <typename T, int PB>
__global__ void HCAMD_KERNEL k_qr(RecompDesc<T> d, int which) {
    extern __shared__ unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    if (which == 0) wg_geqrf<T, PB>(d.m, z.K, w.us, d.m, w.tau_u, w.t_u, lds);
    else ph_qr_y<T, PB>(d, lds);
}

template <typename T>
__global__ void HCAMD_KERNEL k_y(RecompDesc<T> d) {
    __shared__ T lds[GT_LDS_ELEMS];
    ph_y<T>(d, blockIdx.x, gridDim.x, lds);
}

template <typename T>
__global__ void HCAMD_KERNEL k_svd(RecompDesc<T> d) {
    __shared__ T gbuf[JACOBI_LDS_ELEMS];
    ph_core<T>(d, gbuf);
    ph_svd<T>(d, gbuf);
}

// ---------------------------------------------------------------------------
// Multi-workgroup blocked QR of the plain form: one launch factors a panel (one workgroup,
// window pivoting included for the QR of Y) and writes its unit-lower reflectors to `vpan`;
// the next applies the block reflector to the trailing columns, a few column batches per
// workgroup over the whole GPU.
// ---------------------------------------------------------------------------
template <typename T>
struct QrOperand {
    T *A;
    int m, ncols;
    int64_t lda;
    T *tau, *tsave;
    bool pivot;
};

template <typename T, int PB>
__device__ QrOperand<T> qr_operand(const RecompDesc<T> &d, const Ws<T> &w, const Dims &z, int which) {
    if (which == 0) return QrOperand<T>{w.us, d.m, z.K, d.m, w.tau_u, w.t_u, false};
    return QrOperand<T>{w.y, d.n, z.mu, d.n, w.tau_y, w.t_y, window_pivot_fits(d.n, z.mu, PB)};
}

// Squared norms of the columns of the QR operand `which` (all rows), one wave per column; the
// window pivoting of the first panel reads them.
template <typename T, int PB>
__global__ void HCAMD_KERNEL k_qr_colnorms(RecompDesc<T> d, int which, T *cn) {
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    const QrOperand<T> q = qr_operand<T, PB>(d, w, z, which);
    const int l = lane_id();
    for (int c = blockIdx.x * NWAVES + wave_id(); c < q.ncols; c += gridDim.x * NWAVES) {
        const T *col = q.A + (int64_t) c * q.lda;
        T sum = T(0);
        for (int r = l; r < q.m; r += WAVE) sum += col[r] * col[r];
        sum = wave_sum(sum);
        if (l == 0) cn[c] = sum;
    }
}

template <typename T, int PB>
__global__ void HCAMD_KERNEL k_qr_panel(RecompDesc<T> d, int which, int j0, T *vpan, const T *cn) {
    extern __shared__ unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    const QrOperand<T> q = qr_operand<T, PB>(d, w, z, which);
    if (which == 1 && j0 == 0) {
        for (int c = threadIdx.x; c < q.ncols; c += NT) w.piv[c] = c;
        __syncthreads();
    }
    if (which == 1 && q.pivot) {
        // the QR of Y stops at the first panel whose trailing matrix is below the core-truncation
        // level (as ph_qr_y); the launches after that return at once
        if (w.hdr[HDR_KQ] > 0) return;
        const int nsteps = q.m < q.ncols ? q.m : q.ncols;
        const int jb = (nsteps - j0) < PB ? (nsteps - j0) : PB;
        const T trail2 = window_pivot<T, PB>(q.m, q.ncols, q.A, q.lda, j0, jb, w.piv, lds, cn);
        if (j0 == 0) {
            if (threadIdx.x == 0) w.hdr[HDR_STOP2] = __double_as_longlong((double) qr_y_stop2<T>(d, trail2, z.kk));
        } else if (trail2 <= (T) __longlong_as_double(w.hdr[HDR_STOP2])) {
            if (threadIdx.x == 0) {
                w.hdr[HDR_KQ] = j0;
                w.hdr[HDR_TRAIL2] = __double_as_longlong((double) trail2);
            }
            return;
        }
        __syncthreads();
        wg_geqrf_panel<T, PB>(q.m, q.ncols, q.A, q.lda, q.tau, q.tsave, lds, nullptr, w.hdr, j0);
    } else {
        wg_geqrf_panel<T, PB>(q.m, q.ncols, q.A, q.lda, q.tau, q.tsave, lds, q.pivot ? w.piv : nullptr, w.hdr, j0,
                              q.pivot ? cn : nullptr);
    }
    const int rows = q.m - j0;
    wg_copy(vpan, lds, (int64_t) panel_ld(rows) * PB);
}

template <typename T, int PB, int RPT>
__global__ void HCAMD_KERNEL k_qr_update(RecompDesc<T> d, int which, int j0, int cols_per_wg, const T *vpan,
                                         T *cn) {
    __shared__ T red[(NWAVES + 2) * WAVE];
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    const QrOperand<T> q = qr_operand<T, PB>(d, w, z, which);
    if (which == 1 && w.hdr[HDR_KQ] > 0) return;  // the QR of Y stopped early
    const int nsteps = q.m < q.ncols ? q.m : q.ncols;
    const int jb = (nsteps - j0) < PB ? (nsteps - j0) : PB;
    const int c0 = j0 + jb + (int) blockIdx.x * cols_per_wg;
    const int c1 = (c0 + cols_per_wg) < q.ncols ? (c0 + cols_per_wg) : q.ncols;
    if (c0 >= c1) return;
    apply_wy<T, PB, RPT>(vpan, panel_ld(q.m - j0), q.tsave + (int64_t) (j0 / PB) * PB * PB, q.m - j0, q.A + j0, q.lda,
                         c0, c1, true, red);
    if (q.pivot) {
        // norms over rows [j0 + jb, m) of the updated columns, for the next panel's pivoting
        __syncthreads();
        const int l = lane_id();
        for (int c = c0 + wave_id(); c < c1; c += NWAVES) {
            const T *col = q.A + (int64_t) c * q.lda;
            T sum = T(0);
            for (int r = j0 + jb + l; r < q.m; r += WAVE) sum += col[r] * col[r];
            sum = wave_sum(sum);
            if (l == 0) cn[c] = sum;
        }
    }
}

// Trailing update with the row-chunk count of the panel's rows (rows <= 512 for PB 16, <= 1024
// for PB 8, <= 2048 for PB 4, as pick_pb guarantees).
template <typename T, int PB>
static void launch_qr_update(hipStream_t s, const RecompDesc<T> &d, int which, int j0, int rows, int grid, int cpw,
                             const T *vpan, T *cn) {
    const int chunks = (rows + WAVE * NWAVES - 1) / (WAVE * NWAVES);
    if (chunks <= 1)
        hipLaunchKernelGGL((k_qr_update<T, PB, 1>), dim3(grid), dim3(NT), 0, s, d, which, j0, cpw, vpan, cn);
    else if (chunks <= 2)
        hipLaunchKernelGGL((k_qr_update<T, PB, 2>), dim3(grid), dim3(NT), 0, s, d, which, j0, cpw, vpan, cn);
    else if constexpr (PB <= 8) {
        if (chunks <= 4)
            hipLaunchKernelGGL((k_qr_update<T, PB, 4>), dim3(grid), dim3(NT), 0, s, d, which, j0, cpw, vpan, cn);
        else if constexpr (PB <= 4)
            hipLaunchKernelGGL((k_qr_update<T, PB, 8>), dim3(grid), dim3(NT), 0, s, d, which, j0, cpw, vpan, cn);
    }
}

template <typename T, int PB>
static hipError_t launch_qr_mw_pb(hipStream_t s, const RecompDesc<T> &d, int64_t lds_bytes, int which, int m,
                                  int ncols, T *vpan) {
    constexpr int NCB = WAVE / PB;
    const int nsteps = m < ncols ? m : ncols;
    T *cn = vpan + (int64_t) ((d.m > d.n ? d.m : d.n) + 1) * 16;  // after the largest panel (ld rows | 1)
    if (which == 1)
        hipLaunchKernelGGL((k_qr_colnorms<T, PB>), dim3((ncols + NWAVES - 1) / NWAVES), dim3(NT), 0, s, d, which, cn);
    for (int j0 = 0; j0 < nsteps; j0 += PB) {
        hipLaunchKernelGGL((k_qr_panel<T, PB>), dim3(1), dim3(NT), lds_bytes, s, d, which, j0, vpan, (const T *) cn);
        const int jb = (nsteps - j0) < PB ? (nsteps - j0) : PB;
        const int trail = ncols - (j0 + jb);
        if (trail > 0) {
            int cpw = NCB;
            while ((trail + cpw - 1) / cpw > 512) cpw += NCB;
            launch_qr_update<T, PB>(s, d, which, j0, m - j0, (trail + cpw - 1) / cpw, cpw, vpan, cn);
        }
    }
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_qr_mw(hipStream_t s, const RecompDesc<T> &d, int64_t lds_bytes, int pb, int which, int m,
                               int ncols, T *vpan) {
    switch (pb) {
        case 16: return launch_qr_mw_pb<T, 16>(s, d, lds_bytes, which, m, ncols, vpan);
        case 8: return launch_qr_mw_pb<T, 8>(s, d, lds_bytes, which, m, ncols, vpan);
        default: return launch_qr_mw_pb<T, 4>(s, d, lds_bytes, which, m, ncols, vpan);
    }
}

// ---------------------------------------------------------------------------
// Multi-workgroup core SVD of the plain (one update per launch sequence) form.  A large core
// (hundreds of columns) is the whole cost of an update, so its one-sided block Jacobi is spread
// over the GPU: each launch is one step of the round-robin ordering of column blocks, one
// workgroup per block pair (the four waves of a workgroup share the pair's 16 columns, rows
// split across the waves).  flags[s] records whether sweep s rotated anything; the launches of
// a later sweep return at once when the previous sweep was already rotation-free.
// ---------------------------------------------------------------------------
constexpr int MW_TW = 16;

template <typename T>
__global__ void HCAMD_KERNEL k_core(RecompDesc<T> d, int32_t *flags) {
    __shared__ T gbuf[JACOBI_LDS_ELEMS];
    ph_core<T>(d, gbuf);
    for (int i = threadIdx.x; i < JACOBI_MAX_SWEEPS; i += NT) flags[i] = 0;
    if (threadIdx.x == 0) {
        Ws<T> w = ws_of(d);
        const Dims z = dims_of(d, w);
        // gbuf[j] = ||R_Y[j:, :]||_F^2 after ph_core_size, so ||J||_F^2 = gbuf[0] - gbuf[kt]
        const T fro2 = z.kk > 0 ? gbuf[0] - (z.kt < z.kk ? gbuf[z.kt] : T(0)) : T(0);
        w.hdr[HDR_JFRO2] = __double_as_longlong((double) fro2);
    }
}

template <typename T, int RPL>
__global__ void HCAMD_KERNEL k_jacobi_step(RecompDesc<T> d, int32_t *flags, int sweep, int step) {
    constexpr int B = MW_TW / 2;
    __shared__ T lds[coop_rec_lds_elems<T, MW_TW>()];
    if (sweep > 0 && flags[sweep - 1] == 0) return;  // converged in an earlier sweep
    Ws<T> w = ws_of(d);
    const Dims z = dims_of(d, w);
    const int rows = z.mu, ncols = z.kt;
    if (ncols <= 1) return;
    const int nblk = (ncols + B - 1) / B;
    const int nbp = nblk + (nblk & 1);
    const int bp = (int) blockIdx.x;
    if (step >= nbp - 1 || bp >= nbp / 2) return;
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
    if (bI * B >= ncols) return;
    const T fro2 = (T) __longlong_as_double(w.hdr[HDR_JFRO2]);
    const T floor2 = Eps<T>::value * Eps<T>::value * fro2;
    const T tol = sqrt(T(rows)) * Eps<T>::value;
    const int any = coop_jacobi_tile_rec<T, RPL, MW_TW>(w.j, rows, ncols, rows, w.v1, ncols, bI * B, bJ * B,
                                                         step == 0, tol, floor2, lds);
    if ((any & JROT_LARGE) && threadIdx.x == 0) flags[sweep] = 1;
}

template <typename T>
__global__ void HCAMD_KERNEL k_svd_finish(RecompDesc<T> d, const int32_t *flags, int launched) {
    int sweeps = launched;
    for (int s = 0; s < launched; ++s)
        if (flags[s] == 0) {
            sweeps = s;
            break;
        }
    ph_svd_finish<T>(d, sweeps);
}

template <typename T, int RPL>
static void launch_jacobi_sweep(hipStream_t s, const RecompDesc<T> &d, int32_t *flags, int sweep, int steps,
                                int grid) {
    for (int st = 0; st < steps; ++st)
        hipLaunchKernelGGL((k_jacobi_step<T, RPL>), dim3(grid), dim3(NT), 0, s, d, flags, sweep, st);
}

// Core SVD over many workgroups: core build, Jacobi sweeps launched in chunks (the host reads the
// last sweep's flag between chunks, through the pinned word `host_flag`), then the rank rule.
template <typename T>
static hipError_t launch_core_svd_mw(hipStream_t s, const RecompDesc<T> &d, int mu, int kk, int32_t *flags,
                                     volatile int32_t *host_flag) {
    hipLaunchKernelGGL(k_core<T>, dim3(1), dim3(NT), 0, s, d, flags);
    constexpr int B = MW_TW / 2;
    const int nblk = (kk + B - 1) / B;
    const int nbp = nblk + (nblk & 1);
    const int steps = nbp - 1, grid = nbp / 2;
    const int rpl = (mu + NT - 1) / NT;
    int launched = 0, chunk = 6;
    while (launched < JACOBI_MAX_SWEEPS) {
        const int end = launched + chunk < JACOBI_MAX_SWEEPS ? launched + chunk : JACOBI_MAX_SWEEPS;
        for (int sw = launched; sw < end; ++sw) {
            if (rpl <= 1) launch_jacobi_sweep<T, 1>(s, d, flags, sw, steps, grid);
            else if (rpl == 2) launch_jacobi_sweep<T, 2>(s, d, flags, sw, steps, grid);
            else launch_jacobi_sweep<T, 3>(s, d, flags, sw, steps, grid);
        }
        launched = end;
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        e = hipMemcpyAsync((void *) host_flag, flags + launched - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return e;
        if (*host_flag == 0) break;
        chunk = 2;
    }
    hipLaunchKernelGGL(k_svd_finish<T>, dim3(1), dim3(NT), 0, s, d, flags, launched);
    return hipGetLastError();
}

template <typename T, int PB>
__global__ void HCAMD_KERNEL k_recon(RecompDesc<T> d, int cols_per_wg) {
    extern __shared__ unsigned char smem_raw[];
    const int c0 = blockIdx.x * cols_per_wg;
    ph_recon<T, PB>(d, blockIdx.y, c0, c0 + cols_per_wg, reinterpret_cast<T *>(smem_raw));
}

template <typename T>
__global__ void HCAMD_KERNEL k_finish(RecompDesc<T> d) {
    ph_finish<T>(d, blockIdx.x, gridDim.x);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        Ws<T> w = ws_of(d);
        *d.crank = w.hdr[HDR_RANK];
    }
}

__global__ void k_set_i64(int64_t *p, int64_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// Fill V (n x n, ld n) with the identity (FillIdentityMatrix, reference omp/kernels.cpp:144-151,
// on a zeroed buffer) and copy U = W (m x n): the DDC ReadjustTile (Compressed.cpp:696-734).
template <typename T>
__global__ void HCAMD_KERNEL k_readjust_ddc(const T *W, int64_t m, int64_t n, T *cdata) {
    const int64_t tot = m * n + n * n;
    for (int64_t e = (int64_t) blockIdx.x * NT + threadIdx.x; e < tot; e += (int64_t) gridDim.x * NT) {
        if (e < m * n) {
            cdata[e] = W[e];
        } else {
            const int64_t f = e - m * n;
            cdata[e] = ((f % n) == (f / n)) ? T(1) : T(0);
        }
    }
}

template <typename T>
__global__ void HCAMD_KERNEL k_fill_identity(T *I, int64_t n) {
    for (int64_t e = (int64_t) blockIdx.x * NT + threadIdx.x; e < n * n; e += (int64_t) gridDim.x * NT)
        I[e] = ((e % n) == (e / n)) ? T(1) : T(0);
}

template <typename T>
hipError_t launch_fill_identity(hipStream_t s, T *I, int64_t n) {
    int64_t grid = (n * n + NT - 1) / NT;
    if (grid > 1024) grid = 1024;
    hipLaunchKernelGGL(k_fill_identity<T>, dim3((unsigned) (grid > 0 ? grid : 1)), dim3(NT), 0, s, I, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched persistent kernel.
// ---------------------------------------------------------------------------
static_assert(GTM_LDS_FLOATS * sizeof(float) <= (size_t) BATCH_LDS_BYTES, "mixed tile GEMM LDS");

// Factor products of one update (C.Gemm's temporaries).  mixed: FP32 matrix-core products of
// the operands' error-free slices with FP32 accumulation (FP64 operands and results,
// wg_gemm_mixed), otherwise FP64.
template <typename T>
__device__ void run_temp_gemms(const BatchEntry<T> &e, T *lds, int mixed) {
    for (int g = 0; g < e.ngemm; ++g) {
        const GemmSpec<T> &s = e.gemm[g];
        if constexpr (sizeof(T) == 8) {
            if (mixed) {
                wg_gemm_mixed(s.m, s.n, s.k, s.a, s.lda, s.ta, s.b, s.ldb, s.tb, s.c, s.ldc,
                              reinterpret_cast<float *>(lds));
                __syncthreads();
                continue;
            }
        }
        wg_gemm<T>(s.m, s.n, s.k, T(1), s.a, s.lda, s.ta, s.b, s.ldb, s.tb, T(0), s.c, s.ldc, lds);
        __syncthreads();
    }
}

// The rebased copy of one batch entry: pointers named by `rel` are byte offsets into `slot`.
template <typename P>
__device__ __forceinline__ P rebase(P p, int32_t rel, int32_t bit, char *slot) {
    return (rel & bit) ? reinterpret_cast<P>(slot + reinterpret_cast<uintptr_t>(p)) : p;
}

template <typename T, int PB>
__global__ void HCAMD_KERNEL k_batched(const BatchEntry<T> *entries, const int32_t *order, int64_t n,
                                       BatchCounters *ctr, int mixed, char *slots, int64_t slot_bytes,
                                       int64_t *hdrs) {
    extern __shared__ unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    __shared__ int64_t s_item;
    __shared__ BatchEntry<T> s_ent;  // the current update, its pointers rebased onto this slot
    char *slot = slots + (int64_t) blockIdx.x * slot_bytes;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int64_t) atomicAdd(&ctr->next, 1ULL);
        __syncthreads();
        const int64_t item = s_item;
        if (item >= n) break;
        const int32_t idx = order[item];
        if (threadIdx.x == 0) {
            BatchEntry<T> e = entries[idx];
            const int32_t rel = e.rel;
            e.gemm[0].a = rebase(e.gemm[0].a, rel, REL_G0A, slot);
            e.gemm[0].b = rebase(e.gemm[0].b, rel, REL_G0B, slot);
            e.gemm[0].c = rebase(e.gemm[0].c, rel, REL_G0C, slot);
            e.gemm[1].a = rebase(e.gemm[1].a, rel, REL_G1A, slot);
            e.gemm[1].b = rebase(e.gemm[1].b, rel, REL_G1B, slot);
            e.gemm[1].c = rebase(e.gemm[1].c, rel, REL_G1C, slot);
            e.rd.afin = rebase(e.rd.afin, rel, REL_AFIN, slot);
            e.rd.bfin = rebase(e.rd.bfin, rel, REL_BFIN, slot);
            e.rd.ws = rebase(e.rd.ws, rel, REL_WS, slot);
            s_ent = e;
        }
        __syncthreads();
        const BatchEntry<T> &e = s_ent;
        const RecompDesc<T> &d = e.rd;
        Ws<T> w = ws_of(d);
        // the slot's header still holds the previous update's words: start from zero
        for (int k = threadIdx.x; k < HDR_WORDS; k += NT) w.hdr[k] = 0;
        __syncthreads();
        const uint64_t t_start = HCAMD_TSTAMP();
        run_temp_gemms<T>(e, lds, mixed);
        const uint64_t t_gemm = HCAMD_TSTAMP();
        // FP64 with 16-column panels: the QR of U_s reads U_s beyond its first panel in place
        // (StackedU); only the first panel's 16 columns are stacked
        constexpr bool u_in_place = PB == 16 && sizeof(T) == 8;
        ph_stack<T, !u_in_place>(d, 0, 1);
        __syncthreads();
        const Dims z = dims_of(d, w);
        // features for the rank model, and in the exact form, whose runs are its training data;
        // the tolerance rule without a model reads the core width off R_Y and needs none
        if (z.K <= d.kmax && (d.rank_model || !(d.model_tol > 0.0))) ph_features<T>(d, lds);
        uint64_t t_qru = 0, t_y = 0, t_qry = 0, t_svd = 0, t_ru = 0;
        if (z.K <= d.kmax) {
            wg_geqrf<T, PB>(d.m, z.K, w.us, d.m, w.tau_u, w.t_u, lds, nullptr, w.hdr, u_in_place, stacked_u(d, z.K));
            __syncthreads();
            t_qru = HCAMD_TSTAMP();
            ph_y<T>(d, 0, 1, lds);
            __syncthreads();
            t_y = HCAMD_TSTAMP();
            ph_qr_y<T, PB>(d, lds);
            __syncthreads();
            ph_core_size<T>(d, lds);
            ph_learned_size<T>(d, lds);
            ph_core_build<T, false>(d);
            ph_core_qr<T, PB>(d, lds);
            t_qry = HCAMD_TSTAMP();
            ph_svd_nov<T>(d, lds);
            __syncthreads();
            t_svd = HCAMD_TSTAMP();
        }
        // A core beyond the Jacobi's limit, or NaN/Inf in the factors, leaves C as it was: its
        // factors and rank are kept and the status word reports the update.
        const bool keep_c = (z.K > d.kmax) || (w.hdr[HDR_STATUS] & (STATUS_TOO_BIG | STATUS_NAN)) != 0;
        const int r = keep_c ? (int) *d.crank : (int) w.hdr[HDR_RANK];
        if (!keep_c) {
            ph_recon_nov<T, PB>(d, lds);
            __syncthreads();
            t_ru = HCAMD_TSTAMP();
            ph_finish<T>(d, 0, 1);
            __syncthreads();
        }
#ifdef HCAMD_PHASE_TIMING
        if (threadIdx.x == 0) {
            const uint64_t ts[8] = {t_start, t_gemm, t_qru, t_y, t_qry, t_svd, t_ru, HCAMD_TSTAMP()};
            for (int k = 0; k < 8; ++k) w.hdr[HDR_T0 + k] = (int64_t) ts[k];
        }
        __syncthreads();
#else
        (void) t_start; (void) t_gemm; (void) t_qru; (void) t_y; (void) t_qry; (void) t_svd; (void) t_ru;
#endif
        if (threadIdx.x == 0) {
            *d.crank = r;
            const int64_t st = w.hdr[HDR_STATUS];
            if (st) atomicOr(&ctr->status, (unsigned long long) st);
            if (z.K <= d.kmax && w.hdr[HDR_LEARNED] == 2) atomicAdd(&ctr->fallbacks, 1ULL);
            atomicMax(&ctr->max_sweeps, (unsigned long long) w.hdr[HDR_SWEEPS]);
        }
        // the header outlives the slot: diagnostics read it per update
        for (int k = threadIdx.x; k < HDR_WORDS; k += NT) hdrs[(int64_t) idx * HDR_WORDS + k] = w.hdr[k];
        __syncthreads();
    }
}

__global__ void k_reset_counters(BatchCounters *c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        c->next = 0;
        c->status = 0;
        c->max_sweeps = 0;
        c->fallbacks = 0;
    }
}

// ---------------------------------------------------------------------------
// Host launch wrappers
// ---------------------------------------------------------------------------
template <typename T>
hipError_t launch_gemm(hipStream_t s, int M, int N, int Kd, T alpha, const T *A, int64_t lda, bool ta, const T *B,
                       int64_t ldb, bool tb, T beta, T *C, int64_t ldc) {
    if (M <= 0 || N <= 0) return hipSuccess;
    dim3 grid((M + GT_M - 1) / GT_M, (N + GT_N - 1) / GT_N);
    hipLaunchKernelGGL(k_gemm<T>, grid, dim3(NT), 0, s, M, N, Kd, alpha, A, lda, (int) ta, B, ldb, (int) tb, beta,
                       C, ldc);
    return hipGetLastError();
}

// Householder panel width for a QR of `rows` rows: the widest panel whose register tile in
// apply_wy (rows/256 row chunks x PB per lane) stays within 32 values and whose LDS fits.
int pick_pb(int64_t rows, int64_t lds_elems) {
    for (int p = 16; p >= 4; p /= 2)
        if (pb_fits(rows, p, lds_elems)) return p;
    return 0;
}

// Whether a panel of pb columns is valid for a QR of `rows` rows: an instantiated width, at most
// 32 panel values per lane in apply_wy, and the panel workspace within lds_elems.
bool pb_fits(int64_t rows, int pb, int64_t lds_elems) {
    if (pb != 4 && pb != 8 && pb != 16) return false;
    if (rows > 2048 * 4 / pb) return false;
    return qr_lds_elems(rows, pb) <= lds_elems;
}

template <typename T, int PB>
static hipError_t launch_qr_pb(hipStream_t s, const RecompDesc<T> &d, int64_t lds_bytes, int which) {
    hipLaunchKernelGGL((k_qr<T, PB>), dim3(1), dim3(NT), lds_bytes, s, d, which);
    return hipGetLastError();
}

template <typename T, int PB>
static hipError_t launch_recon_pb(hipStream_t s, const RecompDesc<T> &d, int64_t lds_bytes, int r) {
    const int cols = 64 / PB;  // one column batch of the compact-WY application per workgroup
    const int nch = (r + cols - 1) / cols;
    hipLaunchKernelGGL((k_recon<T, PB>), dim3(nch > 0 ? nch : 1, 2), dim3(NT), lds_bytes, s, d, cols);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_qr(hipStream_t s, const RecompDesc<T> &d, int64_t lds_bytes, int pb, int which) {
    switch (pb) {
        case 16: return launch_qr_pb<T, 16>(s, d, lds_bytes, which);
        case 8: return launch_qr_pb<T, 8>(s, d, lds_bytes, which);
        default: return launch_qr_pb<T, 4>(s, d, lds_bytes, which);
    }
}

static int clamp_grid(int64_t work) {
    int64_t g = (work + NT - 1) / NT;
    if (g > 1024) g = 1024;
    if (g < 1) g = 1;
    return (int) g;
}

template <typename T>
hipError_t launch_recomp_pre(hipStream_t s, const RecompDesc<T> &d, int64_t K, int64_t lds_bytes, int pb,
                             int32_t *jflags, int32_t *host_flag, T *vpan) {
    const int64_t m = d.m, n = d.n;
    const int64_t mu = m < K ? m : K;
    const int64_t kk = n < mu ? n : mu;
    hipLaunchKernelGGL(k_stack<T>, dim3(clamp_grid(m * K)), dim3(NT), 0, s, d);
    hipError_t e = (vpan && K >= MW_MIN_QR_COLS) ? launch_qr_mw<T>(s, d, lds_bytes, pb, 0, (int) m, (int) K, vpan)
                                                 : launch_qr<T>(s, d, lds_bytes, pb, 0);
    if (e != hipSuccess) return e;
    const int ytiles = (int) (((n + GT_M - 1) / GT_M) * ((mu + GT_N - 1) / GT_N));
    hipLaunchKernelGGL(k_y<T>, dim3(ytiles), dim3(NT), 0, s, d);
    e = (vpan && mu >= MW_MIN_QR_COLS) ? launch_qr_mw<T>(s, d, lds_bytes, pb, 1, (int) n, (int) mu, vpan)
                                       : launch_qr<T>(s, d, lds_bytes, pb, 1);
    if (e != hipSuccess) return e;
    if (jflags && host_flag && kk >= MW_MIN_COLS && mu <= JACOBI_ROWS_MAX)
        return launch_core_svd_mw<T>(s, d, (int) mu, (int) kk, jflags, host_flag);
    hipLaunchKernelGGL(k_svd<T>, dim3(1), dim3(NT), 0, s, d);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_recomp_post(hipStream_t s, const RecompDesc<T> &d, int64_t r, int64_t lds_bytes, int pb) {
    hipError_t e;
    switch (pb) {
        case 16: e = launch_recon_pb<T, 16>(s, d, lds_bytes, (int) r); break;
        case 8: e = launch_recon_pb<T, 8>(s, d, lds_bytes, (int) r); break;
        default: e = launch_recon_pb<T, 4>(s, d, lds_bytes, (int) r); break;
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_finish<T>, dim3(clamp_grid(r * d.n)), dim3(NT), 0, s, d);
    return hipGetLastError();
}

hipError_t launch_set_i64(hipStream_t s, int64_t *p, int64_t v) {
    hipLaunchKernelGGL(k_set_i64, dim3(1), dim3(64), 0, s, p, v);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_readjust_ddc(hipStream_t s, const T *W, int64_t m, int64_t n, T *cdata) {
    int nwg = (int) ((m * n + n * n + NT - 1) / NT);
    if (nwg > 1024) nwg = 1024;
    hipLaunchKernelGGL(k_readjust_ddc<T>, dim3(nwg), dim3(NT), 0, s, W, m, n, cdata);
    return hipGetLastError();
}

hipError_t launch_reset_counters(hipStream_t s, BatchCounters *c) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, c);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_batched(hipStream_t s, const BatchEntry<T> *entries, const int32_t *order, int64_t n,
                          BatchCounters *ctr, int grid, int64_t lds_bytes, int pb, int mixed, char *slots,
                          int64_t slot_bytes, int64_t *hdrs) {
    switch (pb) {
        case 16:
            hipLaunchKernelGGL((k_batched<T, 16>), dim3(grid), dim3(NT), lds_bytes, s, entries, order, n, ctr, mixed,
                               slots, slot_bytes, hdrs);
            break;
        case 8:
            hipLaunchKernelGGL((k_batched<T, 8>), dim3(grid), dim3(NT), lds_bytes, s, entries, order, n, ctr, mixed,
                               slots, slot_bytes, hdrs);
            break;
        default:
            hipLaunchKernelGGL((k_batched<T, 4>), dim3(grid), dim3(NT), lds_bytes, s, entries, order, n, ctr, mixed,
                               slots, slot_bytes, hdrs);
            break;
    }
    return hipGetLastError();
}

template <typename T>
hipError_t set_kernel_lds_limits(int64_t lds_bytes) {
    // Allow the dynamic-LDS kernels to use more than the default 64 KB.
    const void *fns[] = {
            (const void *) k_qr<T, 16>, (const void *) k_qr<T, 8>, (const void *) k_qr<T, 4>,
            (const void *) k_qr_panel<T, 16>, (const void *) k_qr_panel<T, 8>, (const void *) k_qr_panel<T, 4>,
            (const void *) k_recon<T, 16>, (const void *) k_recon<T, 8>, (const void *) k_recon<T, 4>,
            (const void *) k_batched<T, 16>, (const void *) k_batched<T, 8>, (const void *) k_batched<T, 4>,
    };
    for (const void *f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds_bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

#define HCAMD_INSTANTIATE(T)                                                                                     \
    template hipError_t launch_gemm<T>(hipStream_t, int, int, int, T, const T *, int64_t, bool, const T *,      \
                                       int64_t, bool, T, T *, int64_t);                                          \
    template hipError_t launch_recomp_pre<T>(hipStream_t, const RecompDesc<T> &, int64_t, int64_t, int, int32_t *, \
                                             int32_t *, T *);                                                    \
    template hipError_t launch_recomp_post<T>(hipStream_t, const RecompDesc<T> &, int64_t, int64_t, int);        \
    template hipError_t launch_readjust_ddc<T>(hipStream_t, const T *, int64_t, int64_t, T *);                   \
    template hipError_t launch_fill_identity<T>(hipStream_t, T *, int64_t);                                      \
    template hipError_t launch_batched<T>(hipStream_t, const BatchEntry<T> *, const int32_t *, int64_t,          \
                                          BatchCounters *, int, int64_t, int, int, char *, int64_t, int64_t *);  \
    template hipError_t set_kernel_lds_limits<T>(int64_t);

HCAMD_INSTANTIATE(double)
HCAMD_INSTANTIATE(float)

}  // namespace hcamd
