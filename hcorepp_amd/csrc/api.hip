// Host side of the C-ABI (include/hcorepp_amd.h): execution contexts with a
// device arena, the plain one-update path (HCore<T>::Gemm restated for the
// device), and the batched variable-rank form.
#include <algorithm>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/hcorepp_amd.h"
#include "kernels.hpp"
#include "mlp.hpp"
#include "cholesky.hpp"
#include "tileops.hpp"
#include "autotune_table.hpp"

using namespace hcamd;

namespace {

thread_local std::string g_last_error;

hcorepp_status fail(hcorepp_status st, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
hcorepp_status fail(hcorepp_status st, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return st;
}

#define HIP_TRY(expr)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return fail(HCOREPP_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// Dynamic LDS per workgroup for the QR / SVD / reconstruction / batched kernels: two
// workgroups of the batched kernel share one CU's 160 KB (the batched kernel's LDS-resident
// Jacobi is sized by the same constant).
constexpr int64_t kLdsBytes = BATCH_LDS_BYTES;

}  // namespace

struct hcorepp_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    char *arena = nullptr;
    size_t arena_cap = 0;
    size_t arena_off = 0;
    int64_t *pinned = nullptr;  // host mirror for rank/status readback (HDR_WORDS int64)
    double *pinned_sigma = nullptr;
    size_t pinned_sigma_cap = 0;
    int num_cus = 0;
    hipEvent_t events[16] = {};
    // last plain recompression (diagnostics: hcorepp_debug_workspace)
    void *last_ws = nullptr;
    int64_t last_info[16] = {};
    // while set (hcorepp_gemm_noalloc_d), the arena functions carve this fixed device region
    // instead of the context's own arena, and never allocate
    char *fixed = nullptr;
    size_t fixed_cap = 0;
};

namespace {

bool g_lds_ready_d = false, g_lds_ready_s = false;
std::mutex g_lds_mutex;

hcorepp_status ensure_lds_limits() {
    std::lock_guard<std::mutex> lk(g_lds_mutex);
    if (!g_lds_ready_d) {
        HIP_TRY(set_kernel_lds_limits<double>(kLdsBytes));
        g_lds_ready_d = true;
    }
    if (!g_lds_ready_s) {
        HIP_TRY(set_kernel_lds_limits<float>(kLdsBytes));
        g_lds_ready_s = true;
    }
    return HCOREPP_OK;
}

// Reserve `bytes` of arena for the next operation: grows (after draining the
// stream) when the request exceeds the current capacity, then resets the bump pointer.
hcorepp_status arena_begin(hcorepp_ctx *ctx, size_t bytes) {
    ctx->arena_off = 0;
    if (ctx->fixed) {
        if (bytes > ctx->fixed_cap)
            return fail(HCOREPP_ERR_CAPACITY, "the update needs %zu bytes of device scratch; the allocation-free "
                        "entry point has %zu", bytes, ctx->fixed_cap);
        return HCOREPP_OK;
    }
    if (bytes <= ctx->arena_cap) return HCOREPP_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->arena) HIP_TRY(hipFree(ctx->arena));
    ctx->arena = nullptr;
    size_t cap = std::max(bytes, ctx->arena_cap * 2);
    if (hipMalloc(&ctx->arena, cap) != hipSuccess) {
        ctx->arena = nullptr;
        ctx->arena_cap = 0;
        return fail(HCOREPP_ERR_NOMEM, "arena allocation of %zu bytes failed", cap);
    }
    ctx->arena_cap = cap;
    return HCOREPP_OK;
}

void *arena_take(hcorepp_ctx *ctx, size_t bytes) {
    size_t off = (ctx->arena_off + 255) & ~size_t(255);
    ctx->arena_off = off + bytes;
    return (ctx->fixed ? ctx->fixed : ctx->arena) + off;
}

// ---------------------------------------------------------------------------
// Device scratch of the allocation-free entry point (hcorepp_gemm_noalloc_d).  Every GPU the
// library runs on holds its own copy of this array until the process ends.  In host code the
// array's name is only a stand-in: its device address is looked up once per device
// (hipGetSymbolAddress, with that device current, when the first context on it is created) and
// only that address is handed to kernels.  All calls on one GPU share its array; calls that
// overlap on one GPU race.
constexpr size_t kScratchBytes = size_t(16) << 20;
constexpr int kMaxDevices = 64;
__device__ __align__(256) static unsigned char scratch[kScratchBytes];
char *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mutex;

// Look up (once per device, on the first allocation-free call there) the address of `scratch` on
// `device`, with that device current.
hcorepp_status scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(HCOREPP_ERR_INVALID, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_scratch_mutex);
    if (g_scratch_ptr[device]) return HCOREPP_OK;
    int cur = 0;
    HIP_TRY(hipGetDevice(&cur));
    if (cur != device) HIP_TRY(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) HIP_TRY(hipSetDevice(cur));
    if (e != hipSuccess) return fail(HCOREPP_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    if (!p) return fail(HCOREPP_ERR_HIP, "no device address for the scratch array on device %d", device);
    g_scratch_ptr[device] = static_cast<char *>(p);
    return HCOREPP_OK;
}

// Points a context's arena functions at its device's scratch array for one call.
struct FixedArena {
    hcorepp_ctx *ctx;
    FixedArena(hcorepp_ctx *c, char *base, size_t cap) : ctx(c) {
        ctx->fixed = base;
        ctx->fixed_cap = cap;
    }
    ~FixedArena() {
        ctx->fixed = nullptr;
        ctx->fixed_cap = 0;
        ctx->arena_off = 0;
    }
};

size_t aligned(size_t b) { return (b + 255) & ~size_t(255); }

// (rows, cols, ld, ptr) view of a tile factor, as DataHolder (reference src/data-units/DataHolder.cpp:12-30).
template <typename T>
struct Holder {
    int64_t rows, cols, ld;
    T *data;
};

template <typename T>
int tile_holders(const hcorepp_tile *t, char op, Holder<T> *h, char *hop) {
    if (t->type == HCOREPP_DENSE) {
        h[0] = {t->rows, t->cols, t->ld, (T *) t->data};
        hop[0] = op;
        return 1;
    }
    Holder<T> u = {t->rows, t->rank, t->rows, (T *) t->data};
    Holder<T> v = {t->rank, t->cols, t->rank, (T *) t->data + t->rows * t->max_rank};
    if (op == 'T') {
        h[0] = v;
        h[1] = u;
    } else {
        h[0] = u;
        h[1] = v;
    }
    hop[0] = hop[1] = op;
    return 2;
}

// DenseTile<T>::Gemm k / ld inference (reference src/operators/concrete/Dense.cpp:55-96, ColMajor).
template <typename T>
hipError_t dense_tile_gemm(hipStream_t s, int64_t m, int64_t n, T *c, int64_t ldc, T alpha, const Holder<T> &a,
                           char opa, const Holder<T> &b, char opb, T beta) {
    int64_t k = std::min(a.cols, b.rows);
    if (opa == 'T') {
        k = std::min(a.rows, b.rows);
        if (opb == 'T') k = std::min(a.rows, b.cols);
    } else if (opb == 'T') {
        k = std::min(a.cols, b.cols);
    }
    const int64_t lda = (opa == 'T') ? std::max(a.cols, k) : std::max(a.rows, m);
    const int64_t ldb = (opb == 'T') ? std::max(b.cols, n) : std::max(b.rows, k);
    return launch_gemm<T>(s, (int) m, (int) n, (int) k, alpha, a.data, lda, opa == 'T', b.data, ldb, opb == 'T',
                          beta, c, ldc);
}

// Flop model of CompressedTile<T>::Gemm, default branch (reference Compressed.cpp:335-625).
uint64_t ref_flops_compressed(uint64_t m, uint64_t n, uint64_t arank, uint64_t crk, uint64_t rk, bool fixed) {
    const uint64_t Um = m, Un = arank + crk, minU = std::min(Um, Un);
    const uint64_t Vm = n, Vn = arank + crk, minV = std::min(Vm, Vn);
    const uint64_t sizeS = std::min(std::min(m, n), arank + crk);
    uint64_t f = 0;
    f += m * crk + m * arank + m * arank;
    f += 2 * Um * Un * Un - (uint64_t) (2.0f * (float) Un * (float) Un * (float) Un / 3.0f) + 2 * Um * Un +
         (uint64_t) (17 * (float) Un / 3.0f);
    f += 2 * (minU * ((Un + 1) / 2));
    f += n * crk + n * arank;
    f += 2 * Vm * Vn * Vn - (uint64_t) (2 * (float) Vn * (float) Vn * (float) Vn / 3.0f) + 2 * Vm * Vn +
         (uint64_t) (17 * (float) Vn / 3.0f);
    f += 2 * (minV * ((Vn + 1) / 2));
    f += minU * minV * (arank + crk);
    f += 22 * minV * minV * minV;
    if (!fixed) f += sizeS;
    f += 2 * (uint64_t) std::pow((double) minU, 2) * (3 * Um - minU);  // (8/3) evaluates to 2
    f += Um * rk * minU;
    f += rk;
    f += 2 * (uint64_t) std::pow((double) minV, 2) * (3 * Vm - minV);
    f += Vm * rk * minV;
    return f;
}

template <typename T>
hcorepp_status validate_tile(const hcorepp_tile *t, const char *name) {
    if (!t) return fail(HCOREPP_ERR_INVALID, "%s is null", name);
    if (t->layout != 0) return fail(HCOREPP_ERR_UNSUPPORTED, "%s: only ColMajor layout is implemented", name);
    if (t->rows < 0 || t->cols < 0) return fail(HCOREPP_ERR_INVALID, "%s: negative size", name);
    if (t->type == HCOREPP_DENSE) {
        if (t->ld < std::max<int64_t>(t->rows, 1)) return fail(HCOREPP_ERR_INVALID, "%s: ld < rows", name);
    } else if (t->type == HCOREPP_COMPRESSED) {
        if (t->rank < 0 || t->max_rank < t->rank)
            return fail(HCOREPP_ERR_INVALID, "%s: rank %lld outside [0, max_rank %lld]", name, (long long) t->rank,
                        (long long) t->max_rank);
    } else {
        return fail(HCOREPP_ERR_INVALID, "%s: unknown tile type %d", name, t->type);
    }
    if (!t->data && t->rows * t->cols > 0) return fail(HCOREPP_ERR_INVALID, "%s: null data", name);
    return HCOREPP_OK;
}

// Arena bytes of one plain recompression (workspace, rank word, sigma, Jacobi flags, QR panel scratch).
template <typename T>
size_t recompress_arena_bytes(int64_t m, int64_t n, int64_t kmax, int64_t max_rank) {
    const WsLayout L = ws_layout<T>(m, n, kmax, max_rank);
    return aligned(sizeof(T) * L.total) + aligned(sizeof(int64_t) * 8) +
           aligned(sizeof(double) * (size_t) std::max<int64_t>(kmax, 1)) +
           aligned(sizeof(int32_t) * JACOBI_MAX_SWEEPS) +
           aligned(sizeof(T) * ((size_t) std::max(m, n) * 16 + 1024));
}

template <typename T>
hcorepp_status recompress_plain(hcorepp_ctx *ctx, T alpha, const Holder<T> &ad, const Holder<T> &bd, T beta,
                                hcorepp_tile *C, const hcorepp_cparams &p, uint64_t *flops, double *sigma_out,
                                int64_t *n_sigma);

// Arena bytes one plain update takes (its temporary factor products, then the recompression
// workspace of a compressed C), the figure the pool-size query returns: a context arena of this
// size never grows for the update.
template <typename T>
size_t plain_arena_need(const hcorepp_tile *A, char opA, const hcorepp_tile *B, char opB, const hcorepp_tile *C) {
    Holder<T> ah[2], bh[2];
    char aop[2], bop[2];
    const int na = tile_holders<T>(A, opA, ah, aop);
    const int nb = tile_holders<T>(B, opB, bh, bop);
    const int total = na + nb;
    const bool ddc = (total == 2 && C->type == HCOREPP_COMPRESSED);
    int nt = (total == 3 || ddc) ? 1 : (total == 4 ? 2 : 0);
    size_t need = 0;
    int ai = na - 1, bi = 0;
    while (nt > 0) {
        const int64_t mm = (aop[ai] == 'T') ? ah[ai].cols : ah[ai].rows;
        const int64_t nn = (bop[bi] == 'T') ? bh[bi].rows : bh[bi].cols;
        need += aligned(sizeof(T) * std::max<int64_t>(mm * nn, 1));
        Holder<T> th = {mm, nn, mm, nullptr};
        if (nt == 2 || (nt == 1 && ai == 0)) {
            ah[ai] = th;
            aop[ai] = 'N';
            bi++;
        } else {
            bh[bi] = th;
            bop[bi] = 'N';
            ai--;
        }
        nt--;
    }
    if (C->type == HCOREPP_COMPRESSED && !ddc) {
        const int64_t arank = ah[ai].cols;
        need += recompress_arena_bytes<T>(C->rows, C->cols, arank + C->rank, C->max_rank);
    }
    return need + 4096;
}

template <typename T>
hcorepp_status gemm_cholesky(hcorepp_ctx *ctx, T alpha, const hcorepp_tile *A, char opA, const hcorepp_tile *B,
                             char opB, T beta, hcorepp_tile *C, const hcorepp_cparams &p, uint64_t *flops);

template <typename T>
hcorepp_status gemm_plain(hcorepp_ctx *ctx, T alpha, const hcorepp_tile *A, char opA, const hcorepp_tile *B,
                          char opB, T beta, hcorepp_tile *C, const hcorepp_cparams *params, int cholesky,
                          uint64_t *flops, double *sigma_out, int64_t *n_sigma) {
    if (!ctx) return fail(HCOREPP_ERR_INVALID, "null context");
    hcorepp_status st;
    if ((st = validate_tile<T>(A, "A")) || (st = validate_tile<T>(B, "B")) || (st = validate_tile<T>(C, "C")))
        return st;
    if ((opA != 'N' && opA != 'T') || (opB != 'N' && opB != 'T'))
        return fail(HCOREPP_ERR_INVALID, "op must be 'N' or 'T'");
    hcorepp_cparams p = params ? *params : hcorepp_cparams{1e-9, 0, 1, 0, HCOREPP_SVD_GESDD, 0};
    if ((st = ensure_lds_limits())) return st;
    if (n_sigma) *n_sigma = 0;
    hipStream_t s = ctx->stream;

    Holder<T> ah[2], bh[2];
    char aop[2], bop[2];
    const int na = tile_holders<T>(A, opA, ah, aop);
    const int nb = tile_holders<T>(B, opB, bh, bop);
    const int nc = (C->type == HCOREPP_DENSE) ? 1 : 2;
    const int total = na + nb;
    const bool ddc = (total == 2 && nc == 2);
    int ntemp = (total == 3 || ddc) ? 1 : (total == 4 ? 2 : 0);

    if (cholesky) return gemm_cholesky<T>(ctx, alpha, A, opA, B, opB, beta, C, p, flops);
    if (C->type == HCOREPP_COMPRESSED && !ddc && (opA != 'N' || opB != 'N'))
        return fail(HCOREPP_ERR_UNSUPPORTED,
                    "compressed C with transposed operands: the reference CompressedTile::Gemm ignores the ops");

    // ---- size the arena for this update ----
    const size_t need = plain_arena_need<T>(A, opA, B, opB, C);
    if ((st = arena_begin(ctx, need))) return st;

    // ---- temporary factor products (HCore.cpp:176-264) ----
    int a_idx = na - 1, b_idx = 0;
    T *temps[2] = {nullptr, nullptr};
    int64_t tm[2] = {0, 0}, tn[2] = {0, 0};
    int ntt = 0;
    const T alpha_local = ddc ? alpha : T(1), beta_local = T(0);
    while (ntemp > 0) {
        Holder<T> ad = ah[a_idx], bd = bh[b_idx];
        const char ao = aop[a_idx], bo = bop[b_idx];
        const int64_t mm = (ao == 'T') ? ad.cols : ad.rows;
        const int64_t nn = (bo == 'T') ? bd.rows : bd.cols;
        T *t = (T *) arena_take(ctx, sizeof(T) * std::max<int64_t>(mm * nn, 1));
        temps[ntt] = t;
        tm[ntt] = mm;
        tn[ntt] = nn;
        ntt++;
        HIP_TRY(dense_tile_gemm<T>(s, mm, nn, t, mm, alpha_local, ad, ao, bd, bo, beta_local));
        if (flops) *flops += 2ull * mm * nn * ad.cols;
        Holder<T> th = {mm, nn, mm, t};
        if (ntemp == 2 || (ntemp == 1 && a_idx == 0)) {
            ah[a_idx] = th;
            aop[a_idx] = 'N';
            b_idx++;
        } else {
            bh[b_idx] = th;
            bop[b_idx] = 'N';
            a_idx--;
        }
        ntemp--;
    }

    if (ddc) {
        // HCore.cpp:272-299: W += beta * C_U * C_V, then ReadjustTile(W, rank = -1).
        T *W = temps[0];
        const int64_t m = tm[0], n = tn[0];
        Holder<T> cu = {C->rows, C->rank, C->rows, (T *) C->data};
        Holder<T> cv = {C->rank, C->cols, C->rank, (T *) C->data + C->rows * C->max_rank};
        HIP_TRY(dense_tile_gemm<T>(s, m, n, W, m, beta, cu, 'N', cv, 'N', T(1)));
        if (flops) *flops += 2ull * m * n * cu.cols;
        if (m < n)
            return fail(HCOREPP_ERR_UNSUPPORTED,
                        "DDC with rows < cols: the reference's identity fill writes past its V buffer");
        if ((C->rows + C->cols) * C->max_rank < (m + n) * n)
            return fail(HCOREPP_ERR_CAPACITY, "DDC result needs max_rank >= cols (%lld)", (long long) n);
        HIP_TRY(launch_readjust_ddc<T>(s, W, m, n, (T *) C->data));
        C->rows = m;
        C->cols = n;
        C->rank = n;
        C->max_rank = n;
        C->ld = m;
        return HCOREPP_OK;
    }

    const Holder<T> ad = ah[a_idx], bd = bh[b_idx];
    if (C->type == HCOREPP_DENSE) {
        HIP_TRY(dense_tile_gemm<T>(s, C->rows, C->cols, (T *) C->data, C->ld, alpha, ad, aop[a_idx], bd,
                                   bop[b_idx], beta));
        if (flops) *flops += 2ull * C->rows * C->cols * ad.cols;
        return HCOREPP_OK;
    }
    // ---- compressed C: recompression (CompressedTile<T>::Gemm) ----
    return recompress_plain<T>(ctx, alpha, ad, bd, beta, C, p, flops, sigma_out, n_sigma);
}

// Recompression of C <- alpha * ad * bd + beta * C (compressed C, ad: m x arank, bd: arank x n) in the
// plain form; the context arena must already hold recompress_arena_bytes for this update.
template <typename T>
hcorepp_status recompress_plain(hcorepp_ctx *ctx, T alpha, const Holder<T> &ad, const Holder<T> &bd, T beta,
                                hcorepp_tile *C, const hcorepp_cparams &p, uint64_t *flops, double *sigma_out,
                                int64_t *n_sigma) {
    hipStream_t s = ctx->stream;
    const int64_t m = C->rows, n = C->cols, arank = ad.cols, rc = C->rank;
    const int64_t K = arank + rc;
    if (ad.rows != m || bd.cols != n || bd.rows != arank)
        return fail(HCOREPP_ERR_INVALID, "operand shapes do not match C (%lldx%lld)", (long long) m,
                    (long long) n);
    const int pb = pick_pb(std::max(m, n), kLdsBytes / (int64_t) sizeof(T));
    if (pb == 0) return fail(HCOREPP_ERR_UNSUPPORTED, "tile too tall for the LDS panel (rows %lld)", (long long) m);
    if (std::min(m, K) > JACOBI_ROWS_MAX)
        return fail(HCOREPP_ERR_UNSUPPORTED, "stacked rank %lld exceeds the core SVD limit %d", (long long) K,
                    JACOBI_ROWS_MAX);
    if (arank < 1) return fail(HCOREPP_ERR_INVALID, "the A factor has no columns");
    const WsLayout L = ws_layout<T>(m, n, K, C->max_rank);
    RecompDesc<T> d{};
    d.afin = ad.data;
    d.ldaf = ad.ld;
    d.bfin = bd.data;
    d.ldbf = arank;  // ProcessVpointer reads B with ld = arank (omp/kernels.cpp:69-76)
    d.cdata = (T *) C->data;
    d.crank = (int64_t *) arena_take(ctx, sizeof(int64_t) * 8);
    d.ws = (T *) arena_take(ctx, sizeof(T) * L.total);
    d.m = (int32_t) m;
    d.n = (int32_t) n;
    d.arank = (int32_t) arank;
    d.max_rank = (int32_t) C->max_rank;
    d.kmax = (int32_t) K;
    d.truncated_svd = p.truncated_svd;
    d.fixed_rank = (int32_t) p.fixed_rank;
    d.alpha = alpha;
    d.beta = beta;
    d.accuracy = p.accuracy;
    d.sigma_out = (double *) arena_take(ctx, sizeof(double) * (size_t) std::max<int64_t>(K, 1));
    HIP_TRY(launch_set_i64(s, d.crank, rc));
    ctx->last_ws = d.ws;
    {
        const int64_t info[16] = {m, n, K, K, C->max_rank, (int64_t) sizeof(T), L.total, L.hdr, L.us, L.vs,
                                  L.y, L.t_u, L.t_y, L.j, L.v1, L.vtmp};
        for (int i = 0; i < 16; ++i) ctx->last_info[i] = info[i];
    }
    int32_t *jflags = (int32_t *) arena_take(ctx, sizeof(int32_t) * JACOBI_MAX_SWEEPS);
    T *vpan = (T *) arena_take(ctx, sizeof(T) * ((size_t) std::max(m, n) * 16 + 1024));
    HIP_TRY(launch_recomp_pre<T>(s, d, K, kLdsBytes, pb, jflags, reinterpret_cast<int32_t *>(ctx->pinned + HDR_WORDS),
                                 vpan));
    // The plain form reads the new rank back with one stream synchronisation,
    // as the reference does (src/kernels/cuda/CudaKernels.cu:656-697).
    const int64_t *hdr = reinterpret_cast<const int64_t *>(d.ws + L.hdr);
    HIP_TRY(hipMemcpyAsync(ctx->pinned, hdr, sizeof(int64_t) * HDR_WORDS, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const int64_t r = ctx->pinned[HDR_RANK];
    const int64_t size_s = ctx->pinned[HDR_SIZES];
    const int64_t status = ctx->pinned[HDR_STATUS];
    if (status & STATUS_NAN) return fail(HCOREPP_ERR_NUMERIC, "NaN/Inf in the core SVD");
    if (status & STATUS_NOCONV) return fail(HCOREPP_ERR_NUMERIC, "core Jacobi SVD did not converge");
    if (status & STATUS_TOO_BIG) return fail(HCOREPP_ERR_CAPACITY, "stacked rank exceeds the workspace");
    if (r < 1 || r > C->max_rank) return fail(HCOREPP_ERR_NUMERIC, "invalid rank %lld", (long long) r);
    HIP_TRY(launch_recomp_post<T>(s, d, r, kLdsBytes, pb));
    if (sigma_out && size_s > 0) {
        HIP_TRY(hipMemcpyAsync(sigma_out, d.sigma_out, sizeof(double) * size_s, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (n_sigma) *n_sigma = size_s;
    }
    if (flops) *flops += ref_flops_compressed(m, n, arank, rc, r, p.fixed_rank != 0);
    C->rank = r;
    C->ld = m;
    return HCOREPP_OK;
}

// Gemm in the Cholesky flow (the reference's aCholesky = true, HCore.cpp:160-174 and the aCholesky
// branches of Compressed.cpp:273-276, 458-476, 529-595, 669-673): compressed tiles are U V^T with
// V stored n x r (ld n), and C <- beta C + alpha A B^T for opA = 'N', opB = 'T'.  With
// A = U_a V_a^T and B = U_b V_b^T, A B^T = U_a [(V_a^T V_b) U_b^T], so the update is the engine's
// recompression with the A factor U_a and the B factor (V_a^T V_b) U_b^T (two small products);
// C's V factor is transposed into the engine's r x n layout before and back after.
template <typename T>
hcorepp_status gemm_cholesky(hcorepp_ctx *ctx, T alpha, const hcorepp_tile *A, char opA, const hcorepp_tile *B,
                             char opB, T beta, hcorepp_tile *C, const hcorepp_cparams &p, uint64_t *flops) {
    if (A->type != HCOREPP_COMPRESSED || B->type != HCOREPP_COMPRESSED || C->type != HCOREPP_COMPRESSED ||
        opA != 'N' || opB != 'T')
        return fail(HCOREPP_ERR_UNSUPPORTED, "the Cholesky-flow Gemm takes compressed A, B, C with opA 'N', opB 'T' "
                    "(the re-plumbed case of reference HCore.cpp:160-174)");
    const int64_t m = C->rows, n = C->cols, ra = A->rank, rb = B->rank, rc = C->rank, kk = A->cols;
    if (A->rows != m || B->rows != n || B->cols != kk)
        return fail(HCOREPP_ERR_INVALID, "shapes: A %lldx%lld, B^T %lldx%lld, C %lldx%lld", (long long) A->rows,
                    (long long) A->cols, (long long) B->cols, (long long) B->rows, (long long) m, (long long) n);
    if (ra < 1 || rb < 1) return fail(HCOREPP_ERR_INVALID, "the A and B factors need rank >= 1");
    hcorepp_status st;
    if ((st = ensure_lds_limits())) return st;
    const size_t need = aligned(sizeof(T) * (size_t) (ra * rb)) + aligned(sizeof(T) * (size_t) (ra * n)) +
                        aligned(sizeof(T) * (size_t) (n * std::max<int64_t>(C->max_rank, 1))) +
                        recompress_arena_bytes<T>(m, n, ra + rc, C->max_rank) + 4096;
    if ((st = arena_begin(ctx, need))) return st;
    hipStream_t s = ctx->stream;
    T *w1 = (T *) arena_take(ctx, sizeof(T) * (size_t) (ra * rb));
    T *w2 = (T *) arena_take(ctx, sizeof(T) * (size_t) (ra * n));
    T *xv = (T *) arena_take(ctx, sizeof(T) * (size_t) (n * std::max<int64_t>(C->max_rank, 1)));
    const T *ua = (const T *) A->data, *va = (const T *) A->data + A->rows * A->max_rank;
    const T *ub = (const T *) B->data, *vb = (const T *) B->data + B->rows * B->max_rank;
    T *cv = (T *) C->data + m * C->max_rank;
    HIP_TRY(launch_gemm<T>(s, (int) ra, (int) rb, (int) kk, T(1), va, kk, true, vb, kk, false, T(0), w1, ra));
    HIP_TRY(launch_gemm<T>(s, (int) ra, (int) n, (int) rb, T(1), w1, ra, false, ub, n, true, T(0), w2, ra));
    if (flops) *flops += 2ull * ra * rb * kk + 2ull * ra * n * rb;
    // C's V: n x rc (ld n) -> rc x n (ld rc)
    HIP_TRY(launch_transpose<T>(s, (int) n, (int) rc, cv, n, xv, std::max<int64_t>(rc, 1)));
    HIP_TRY(hipMemcpyAsync(cv, xv, sizeof(T) * (size_t) (rc * n), hipMemcpyDeviceToDevice, s));
    const Holder<T> ad{m, ra, m, const_cast<T *>(ua)}, bd{ra, n, ra, w2};
    hcorepp_tile Ce = *C;
    st = recompress_plain<T>(ctx, alpha, ad, bd, beta, &Ce, p, flops, nullptr, nullptr);
    const int64_t r = st ? rc : Ce.rank;  // on failure the engine left C's factors in place
    // back to n x r (ld n)
    HIP_TRY(launch_transpose<T>(s, (int) r, (int) n, cv, std::max<int64_t>(r, 1), xv, n));
    HIP_TRY(hipMemcpyAsync(cv, xv, sizeof(T) * (size_t) (r * n), hipMemcpyDeviceToDevice, s));
    if (st) return st;
    C->rank = r;
    C->ld = m;
    return HCOREPP_OK;
}

// HCore<T>::Potrf (reference src/api/HCore.cpp:586-622, lapack::potrf at omp/kernels.cpp:236-243).
template <typename T>
hcorepp_status tile_potrf(hcorepp_ctx *ctx, char uplo, hcorepp_tile *A) {
    hcorepp_status st;
    if (!ctx) return fail(HCOREPP_ERR_INVALID, "null context");
    if ((st = validate_tile<T>(A, "A"))) return st;
    if (A->type != HCOREPP_DENSE) return fail(HCOREPP_ERR_INVALID, "Potrf works only with dense tiles");
    if (A->rows != A->cols) return fail(HCOREPP_ERR_INVALID, "potrf needs a square tile");
    if (uplo != 'L' && uplo != 'U') return fail(HCOREPP_ERR_INVALID, "uplo must be 'L' or 'U'");
    const int64_t n = A->rows;
    if ((st = arena_begin(ctx, aligned(sizeof(T) * (size_t) std::max<int64_t>(n * n, 1)) + 4096))) return st;
    T *work = (T *) arena_take(ctx, sizeof(T) * (size_t) std::max<int64_t>(n * n, 1));
    int *info = (int *) arena_take(ctx, sizeof(int));
    HIP_TRY(launch_tile_potrf<T>(ctx->stream, uplo == 'U', (T *) A->data, (int) n, A->ld, work, info));
    HIP_TRY(hipMemcpyAsync(ctx->pinned, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const int h = *reinterpret_cast<const int *>(ctx->pinned);
    if (h) return fail(HCOREPP_ERR_NUMERIC, "potrf: leading minor %d is not positive", h);
    return HCOREPP_OK;
}

// HCore<T>::Trsm (reference src/api/HCore.cpp:624-647): blas::trsm on the V factor of the
// compressed B, viewed as rows x rank with B's leading dimension.
template <typename T>
hcorepp_status tile_trsm(hcorepp_ctx *ctx, char side, char uplo, char trans, char diag, T alpha,
                         const hcorepp_tile *A, hcorepp_tile *B) {
    hcorepp_status st;
    if (!ctx) return fail(HCOREPP_ERR_INVALID, "null context");
    if ((st = validate_tile<T>(A, "A")) || (st = validate_tile<T>(B, "B"))) return st;
    if (B->type != HCOREPP_COMPRESSED) return fail(HCOREPP_ERR_INVALID, "TRSM: Tile B must be compressed");
    if (A->type != HCOREPP_DENSE || A->rows != A->cols) return fail(HCOREPP_ERR_INVALID, "A must be square dense");
    if ((side != 'L' && side != 'R') || (uplo != 'L' && uplo != 'U') || (trans != 'N' && trans != 'T' &&
        trans != 'C') || (diag != 'N' && diag != 'U'))
        return fail(HCOREPP_ERR_INVALID, "side L/R, uplo L/U, trans N/T/C, diag N/U");
    const int64_t m = B->rows, n = B->rank;
    if (std::max<int64_t>(B->ld, 1) * n > B->cols * B->max_rank || B->ld < m)
        return fail(HCOREPP_ERR_INVALID, "V viewed as %lld x %lld (ld %lld) exceeds B's V slot", (long long) m,
                    (long long) n, (long long) B->ld);
    if (A->rows != (side == 'L' ? m : n))
        return fail(HCOREPP_ERR_INVALID, "A is %lldx%lld, the solve needs order %lld", (long long) A->rows,
                    (long long) A->cols, (long long) (side == 'L' ? m : n));
    T *v = (T *) B->data + B->rows * B->max_rank;
    HIP_TRY(launch_tile_trsm<T>(ctx->stream, side == 'L', uplo == 'U', trans != 'N', diag == 'U', (int) m, (int) n,
                                alpha, (const T *) A->data, A->ld, v, std::max<int64_t>(B->ld, 1)));
    return HCOREPP_OK;
}

// HCore<T>::Syrk (reference src/api/HCore.cpp:484-584): blas::syrk on a dense A; for a
// compressed A (U V^T, V n x r): C <- beta C - U (V^T V) U^T (alpha is not used, as in the
// reference), then the strict triangle opposite to uplo set to zero.
template <typename T>
hcorepp_status tile_syrk(hcorepp_ctx *ctx, T alpha, const hcorepp_tile *A, char opA, char uplo, T beta,
                         hcorepp_tile *C) {
    hcorepp_status st;
    if (!ctx) return fail(HCOREPP_ERR_INVALID, "null context");
    if ((st = validate_tile<T>(A, "A")) || (st = validate_tile<T>(C, "C"))) return st;
    if (C->type != HCOREPP_DENSE) return fail(HCOREPP_ERR_UNSUPPORTED, "syrk writes a dense C");
    if ((opA != 'N' && opA != 'T' && opA != 'C') || (uplo != 'L' && uplo != 'U' && uplo != 'G'))
        return fail(HCOREPP_ERR_INVALID, "opA N/T/C, uplo L/U/G");
    hipStream_t s = ctx->stream;
    if (C->rows != C->cols) return fail(HCOREPP_ERR_INVALID, "syrk needs a square C");
    if (A->type == HCOREPP_DENSE) {
        if (uplo == 'G') return fail(HCOREPP_ERR_INVALID, "syrk of a dense tile needs uplo L or U");
        const int64_t n = C->rows, k = C->cols;
        const int64_t arows = opA == 'N' ? n : k, acols = opA == 'N' ? k : n;
        if (A->rows < arows || A->cols < acols) return fail(HCOREPP_ERR_INVALID, "A is smaller than op(A) n x k");
        HIP_TRY(launch_tile_syrk_dense<T>(s, uplo == 'U', opA != 'N', (int) n, (int) k, alpha, (const T *) A->data,
                                          A->ld, beta, (T *) C->data, C->ld));
        return HCOREPP_OK;
    }
    const int64_t m = A->rows, na = A->cols, r = A->rank;
    if (C->rows != m || C->cols != m) return fail(HCOREPP_ERR_INVALID, "C must be %lldx%lld", (long long) m,
                                                   (long long) m);
    if ((st = arena_begin(ctx, aligned(sizeof(T) * (size_t) std::max<int64_t>(r * r, 1)) +
                                   aligned(sizeof(T) * (size_t) std::max<int64_t>(m * r, 1)) + 4096)))
        return st;
    T *w = (T *) arena_take(ctx, sizeof(T) * (size_t) std::max<int64_t>(r * r, 1));
    T *t = (T *) arena_take(ctx, sizeof(T) * (size_t) std::max<int64_t>(m * r, 1));
    const T *u = (const T *) A->data, *v = (const T *) A->data + m * A->max_rank;
    HIP_TRY(launch_gemm<T>(s, (int) r, (int) r, (int) na, T(1), v, na, true, v, na, false, T(0), w, std::max<int64_t>(r, 1)));
    HIP_TRY(launch_gemm<T>(s, (int) m, (int) r, (int) r, T(1), u, m, false, w, std::max<int64_t>(r, 1), false, T(0), t, m));
    HIP_TRY(launch_gemm<T>(s, (int) m, (int) m, (int) r, T(-1), t, m, false, u, m, true, beta, (T *) C->data, C->ld));
    if (uplo != 'G') HIP_TRY(launch_zero_tri<T>(s, (T *) C->data, (int) m, C->ld, uplo == 'L'));
    return HCOREPP_OK;
}

// Compression constructor on the device: the truncated SVD of the dense tile A is the
// recompression of A * I from an empty C, so it reuses the plain path (QR, window-pivoted QR,
// Jacobi core SVD, the rank rule, reconstruction into the caller's tile buffer).
template <typename T>
hcorepp_status compress_plain(hcorepp_ctx *ctx, int64_t m, int64_t n, const T *A, int64_t lda,
                              const hcorepp_cparams *params, hcorepp_tile *out) {
    hcorepp_status st;
    if (!ctx || !out || (m * n > 0 && !A)) return fail(HCOREPP_ERR_INVALID, "bad arguments");
    if (m < 1 || n < 1 || lda < m) return fail(HCOREPP_ERR_INVALID, "bad tile shape %lldx%lld (ld %lld)",
                                               (long long) m, (long long) n, (long long) lda);
    if ((st = validate_tile<T>(out, "out"))) return st;
    if (out->type != HCOREPP_COMPRESSED || out->rows != m || out->cols != n || out->max_rank < 1 || !out->data)
        return fail(HCOREPP_ERR_INVALID, "out must be a compressed %lldx%lld tile with max_rank >= 1",
                    (long long) m, (long long) n);
    const hcorepp_cparams p = params ? *params : hcorepp_cparams{1e-9, 0, 1, 0, HCOREPP_SVD_GESDD, 0};
    if ((st = ensure_lds_limits())) return st;
    if ((st = arena_begin(ctx, recompress_arena_bytes<T>(m, n, n, out->max_rank) +
                                   aligned(sizeof(T) * (size_t) (n * n)) + 4096)))
        return st;
    T *I = (T *) arena_take(ctx, sizeof(T) * (size_t) (n * n));
    HIP_TRY(launch_fill_identity<T>(ctx->stream, I, n));
    const Holder<T> ad{m, n, lda, const_cast<T *>(A)}, bd{n, n, n, I};
    out->rank = 0;
    out->ld = m;
    return recompress_plain<T>(ctx, T(1), ad, bd, T(0), out, p, nullptr, nullptr, nullptr);
}

}  // namespace

// ===========================================================================
// Batched variable-rank form
// ===========================================================================
struct hcorepp_batch {
    hcorepp_ctx *ctx = nullptr;
    int precision = 0;
    int64_t n = 0;
    int pb = 16;
    int64_t max_rows = 1;  // tallest stacked-factor QR of the batch
    int grid = 1;
    void *entries = nullptr;  // BatchEntry<double>[n]
    int32_t *order = nullptr;
    int64_t *ranks = nullptr;
    BatchCounters *ctr = nullptr;
    // Workspace: one slot per workgroup of the launch (the largest update's temporaries and
    // recompression workspace), reused by every update the workgroup takes.
    char *ws = nullptr;
    size_t slot_bytes = 0;
    int slots = 0;
    int64_t *hdrs = nullptr;  // HDR_WORDS header words per update, copied out of the slot
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<char> ent_host;  // host copy of the entries (re-uploaded when the rank model changes)
};

// Make room for `grid` workspace slots (drains the stream before replacing the allocation).
static hcorepp_status ensure_slots(hcorepp_batch *b, int grid) {
    if (grid <= b->slots) return HCOREPP_OK;
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    if (b->ws) HIP_TRY(hipFree(b->ws));
    b->ws = nullptr;
    b->slots = 0;
    const size_t bytes = b->slot_bytes * (size_t) grid;
    if (hipMalloc(&b->ws, std::max<size_t>(bytes, 256)) != hipSuccess) {
        b->ws = nullptr;
        return fail(HCOREPP_ERR_NOMEM, "batch workspace of %d slots x %zu bytes", grid, b->slot_bytes);
    }
    b->slots = grid;
    return HCOREPP_OK;
}

// Address of header word `word` of update i (device).
static const char *hdr_word(const hcorepp_batch *b, int64_t i, int word) {
    return reinterpret_cast<const char *>(b->hdrs + i * HDR_WORDS + word);
}

// Autotune table row for a batch class (tallest QR, mean stacked rank, updates per CU): the
// first row whose bounds hold.  False when the table has none (or is empty).
static bool tune_lookup(int64_t rows, double mean_k, double per_cu, int *pb, int *wg_per_cu) {
    for (int i = 0; i < kTuneRows; ++i) {
        const TuneRow &r = kTuneTable[i];
        if (rows <= r.max_rows && (r.max_mean_k < 0.0 || mean_k <= r.max_mean_k) &&
            (r.max_per_cu < 0.0 || per_cu <= r.max_per_cu)) {
            *pb = r.panel;
            *wg_per_cu = r.workgroups_per_cu;
            return true;
        }
    }
    return false;
}

// A captured graph holds the launch's grid and panel width: changing either drops it.
static void drop_graph(hcorepp_batch *b) {
    if (b->graph_exec) {
        hipGraphExecDestroy(b->graph_exec);
        b->graph_exec = nullptr;
    }
    if (b->graph) {
        hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
}

extern "C" {

hcorepp_status hcorepp_ctx_create(int device, size_t arena_bytes, hcorepp_ctx **out) {
    if (!out) return fail(HCOREPP_ERR_INVALID, "out is null");
    *out = nullptr;
    HIP_TRY(hipSetDevice(device));
    hcorepp_ctx *c = new (std::nothrow) hcorepp_ctx();
    if (!c) return fail(HCOREPP_ERR_NOMEM, "host allocation failed");
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamDefault);  // ordered with the null stream
    if (e != hipSuccess) {
        delete c;
        return fail(HCOREPP_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    e = hipHostMalloc((void **) &c->pinned, sizeof(int64_t) * 64, hipHostMallocDefault);
    if (e != hipSuccess) {
        hipStreamDestroy(c->stream);
        delete c;
        return fail(HCOREPP_ERR_HIP, "hipHostMalloc: %s", hipGetErrorString(e));
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    if (arena_bytes) {
        hcorepp_status st = arena_begin(c, arena_bytes);
        if (st) {
            hipHostFree(c->pinned);
            hipStreamDestroy(c->stream);
            delete c;
            return st;
        }
    }
    hcorepp_status st = ensure_lds_limits();
    if (st) {
        hipHostFree(c->pinned);
        hipStreamDestroy(c->stream);
        delete c;
        return st;
    }
    *out = c;
    return HCOREPP_OK;
}

hcorepp_status hcorepp_ctx_destroy(hcorepp_ctx *ctx) {
    if (!ctx) return HCOREPP_OK;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    for (auto &ev : ctx->events)
        if (ev) hipEventDestroy(ev);
    if (ctx->arena) hipFree(ctx->arena);
    if (ctx->pinned) hipHostFree(ctx->pinned);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return HCOREPP_OK;
}

hcorepp_status hcorepp_ctx_sync(hcorepp_ctx *ctx) {
    if (!ctx) return fail(HCOREPP_ERR_INVALID, "null context");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_ctx_event_record(hcorepp_ctx *ctx, int slot) {
    if (!ctx || slot < 0 || slot >= 16) return fail(HCOREPP_ERR_INVALID, "bad event slot");
    if (!ctx->events[slot]) HIP_TRY(hipEventCreate(&ctx->events[slot]));
    HIP_TRY(hipEventRecord(ctx->events[slot], ctx->stream));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_ctx_event_elapsed(hcorepp_ctx *ctx, int slot0, int slot1, float *ms) {
    if (!ctx || !ms || slot0 < 0 || slot0 >= 16 || slot1 < 0 || slot1 >= 16 || !ctx->events[slot0] ||
        !ctx->events[slot1])
        return fail(HCOREPP_ERR_INVALID, "bad event slots");
    HIP_TRY(hipEventSynchronize(ctx->events[slot1]));
    HIP_TRY(hipEventElapsedTime(ms, ctx->events[slot0], ctx->events[slot1]));
    return HCOREPP_OK;
}

void *hcorepp_ctx_stream(hcorepp_ctx *ctx) { return ctx ? (void *) ctx->stream : nullptr; }

size_t hcorepp_ctx_arena_bytes(hcorepp_ctx *ctx) { return ctx ? ctx->arena_cap : 0; }

int hcorepp_ctx_compute_units(hcorepp_ctx *ctx) { return ctx ? ctx->num_cus : 0; }

hcorepp_status hcorepp_malloc(size_t bytes, void **out) {
    if (!out) return fail(HCOREPP_ERR_INVALID, "out is null");
    if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) return fail(HCOREPP_ERR_NOMEM, "hipMalloc(%zu)", bytes);
    return HCOREPP_OK;
}

hcorepp_status hcorepp_free(void *ptr) {
    if (ptr) HIP_TRY(hipFree(ptr));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_memcpy(hcorepp_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!bytes) return HCOREPP_OK;
    if (ctx) {
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    } else {
        HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
    }
    return HCOREPP_OK;
}

hcorepp_status hcorepp_memcpy_async(hcorepp_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx) return fail(HCOREPP_ERR_INVALID, "null context");
    if (bytes) HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_kernel_gemm_d(hcorepp_ctx *ctx, char transa, char transb, int64_t m, int64_t n, int64_t k,
                                     double alpha, const double *A, int64_t lda, const double *B, int64_t ldb,
                                     double beta, double *C, int64_t ldc) {
    if (!ctx || m < 0 || n < 0 || k < 0 || (m * n > 0 && !C) || (m * n * k > 0 && (!A || !B)))
        return fail(HCOREPP_ERR_INVALID, "bad arguments");
    if ((transa != 'N' && transa != 'T') || (transb != 'N' && transb != 'T'))
        return fail(HCOREPP_ERR_INVALID, "trans must be 'N' or 'T'");
    if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) return fail(HCOREPP_ERR_UNSUPPORTED, "dimension too large");
    const int64_t ar = transa == 'N' ? m : k, br = transb == 'N' ? k : n;
    if (ldc < std::max<int64_t>(m, 1) || lda < std::max<int64_t>(ar, 1) || ldb < std::max<int64_t>(br, 1))
        return fail(HCOREPP_ERR_INVALID, "leading dimension too small");
    HIP_TRY(launch_gemm<double>(ctx->stream, (int) m, (int) n, (int) k, alpha, A, lda, transa == 'T', B, ldb,
                                transb == 'T', beta, C, ldc));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_memset(hcorepp_ctx *ctx, void *dst, int value, size_t bytes) {
    if (!bytes) return HCOREPP_OK;
    if (ctx) HIP_TRY(hipMemsetAsync(dst, value, bytes, ctx->stream));
    else HIP_TRY(hipMemset(dst, value, bytes));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_gemm_d(hcorepp_ctx *ctx, double alpha, const hcorepp_tile *A, char opA,
                              const hcorepp_tile *B, char opB, double beta, hcorepp_tile *C,
                              const hcorepp_cparams *params, int cholesky, uint64_t *flops, double *sigma_out,
                              int64_t *n_sigma) {
    return gemm_plain<double>(ctx, alpha, A, opA, B, opB, beta, C, params, cholesky, flops, sigma_out, n_sigma);
}

hcorepp_status hcorepp_gemm_s(hcorepp_ctx *ctx, float alpha, const hcorepp_tile *A, char opA,
                              const hcorepp_tile *B, char opB, float beta, hcorepp_tile *C,
                              const hcorepp_cparams *params, int cholesky, uint64_t *flops, double *sigma_out,
                              int64_t *n_sigma) {
    return gemm_plain<float>(ctx, alpha, A, opA, B, opB, beta, C, params, cholesky, flops, sigma_out, n_sigma);
}

hcorepp_status hcorepp_gemm_noalloc_d(hcorepp_ctx *ctx, double alpha, const hcorepp_tile *A, char opA,
                                      const hcorepp_tile *B, char opB, double beta, hcorepp_tile *C,
                                      const hcorepp_cparams *params, uint64_t *flops) {
    if (!ctx) return fail(HCOREPP_ERR_INVALID, "null context");
    hcorepp_status st = scratch_init(ctx->device);
    if (st) return st;
    FixedArena guard(ctx, g_scratch_ptr[ctx->device], kScratchBytes);
    return gemm_plain<double>(ctx, alpha, A, opA, B, opB, beta, C, params, 0, flops, nullptr, nullptr);
}

size_t hcorepp_gemm_noalloc_scratch_bytes(void) { return kScratchBytes; }

hcorepp_status hcorepp_tile_potrf_d(hcorepp_ctx *ctx, char uplo, hcorepp_tile *A) {
    return tile_potrf<double>(ctx, uplo, A);
}
hcorepp_status hcorepp_tile_potrf_s(hcorepp_ctx *ctx, char uplo, hcorepp_tile *A) {
    return tile_potrf<float>(ctx, uplo, A);
}
hcorepp_status hcorepp_tile_trsm_d(hcorepp_ctx *ctx, char side, char uplo, char trans, char diag, double alpha,
                                   const hcorepp_tile *A, hcorepp_tile *B) {
    return tile_trsm<double>(ctx, side, uplo, trans, diag, alpha, A, B);
}
hcorepp_status hcorepp_tile_trsm_s(hcorepp_ctx *ctx, char side, char uplo, char trans, char diag, float alpha,
                                   const hcorepp_tile *A, hcorepp_tile *B) {
    return tile_trsm<float>(ctx, side, uplo, trans, diag, alpha, A, B);
}
hcorepp_status hcorepp_tile_syrk_d(hcorepp_ctx *ctx, double alpha, const hcorepp_tile *A, char opA, char uplo,
                                   double beta, hcorepp_tile *C) {
    return tile_syrk<double>(ctx, alpha, A, opA, uplo, beta, C);
}
hcorepp_status hcorepp_tile_syrk_s(hcorepp_ctx *ctx, float alpha, const hcorepp_tile *A, char opA, char uplo,
                                   float beta, hcorepp_tile *C) {
    return tile_syrk<float>(ctx, alpha, A, opA, uplo, beta, C);
}

hcorepp_status hcorepp_compress_d(hcorepp_ctx *ctx, int64_t rows, int64_t cols, const double *A, int64_t lda,
                                  const hcorepp_cparams *params, hcorepp_tile *out) {
    return compress_plain<double>(ctx, rows, cols, A, lda, params, out);
}

hcorepp_status hcorepp_compress_s(hcorepp_ctx *ctx, int64_t rows, int64_t cols, const float *A, int64_t lda,
                                  const hcorepp_cparams *params, hcorepp_tile *out) {
    return compress_plain<float>(ctx, rows, cols, A, lda, params, out);
}

size_t hcorepp_gemm_pool_size_d(const hcorepp_tile *A, const hcorepp_tile *B, const hcorepp_tile *C,
                                const hcorepp_cparams *params) {
    (void) params;
    if (!A || !B || !C) return 0;
    return (plain_arena_need<double>(A, 'N', B, 'N', C) + sizeof(double) - 1) / sizeof(double);
}

size_t hcorepp_gemm_pool_size_s(const hcorepp_tile *A, const hcorepp_tile *B, const hcorepp_tile *C,
                                const hcorepp_cparams *params) {
    (void) params;
    if (!A || !B || !C) return 0;
    return (plain_arena_need<float>(A, 'N', B, 'N', C) + sizeof(float) - 1) / sizeof(float);
}

// ---------------------------------------------------------------------------
hcorepp_status hcorepp_batch_create(hcorepp_ctx *ctx, int64_t n, const hcorepp_tile *A, const hcorepp_tile *B,
                                    const hcorepp_tile *C, double alpha, double beta,
                                    const hcorepp_cparams *params, int precision, hcorepp_batch **out) {
    if (!ctx || !out || n < 0 || (n > 0 && (!A || !B || !C))) return fail(HCOREPP_ERR_INVALID, "bad arguments");
    *out = nullptr;
    hcorepp_status st;
    if ((st = ensure_lds_limits())) return st;
    if (precision != HCOREPP_PREC_FP64 && precision != HCOREPP_PREC_MIXED)
        return fail(HCOREPP_ERR_INVALID, "unknown precision form %d", precision);
    hcorepp_cparams p = params ? *params : hcorepp_cparams{1e-9, 0, 1, 0, HCOREPP_SVD_GESDD, 0};
    using T = double;
    for (int64_t i = 0; i < n; ++i) {
        if ((st = validate_tile<T>(&A[i], "A[i]")) || (st = validate_tile<T>(&B[i], "B[i]")) ||
            (st = validate_tile<T>(&C[i], "C[i]")))
            return st;
        if (C[i].type != HCOREPP_COMPRESSED)
            return fail(HCOREPP_ERR_UNSUPPORTED, "batched form needs compressed C tiles (update %lld)",
                        (long long) i);
        if (A[i].type == HCOREPP_DENSE && B[i].type == HCOREPP_DENSE)
            return fail(HCOREPP_ERR_UNSUPPORTED, "DDC updates are not batched (update %lld)", (long long) i);
    }
    // Plan one update: fills `e` (when non-null) and returns the bytes of slot workspace the update
    // needs (its temporaries, then the recompression workspace).  Pointers into the slot are byte
    // offsets, marked in e->rel; the kernel adds the slot's base.
    auto plan = [&](int64_t i, BatchEntry<T> *e, double *cost, int64_t *rows, int64_t *kst) -> size_t {
        const hcorepp_tile *a = &A[i], *b = &B[i], *c = &C[i];
        Holder<T> ah[2], bh[2];
        char aop[2], bop[2];
        bool atmp[2] = {false, false}, btmp[2] = {false, false};  // holder lives in the slot
        const int na = tile_holders<T>(a, 'N', ah, aop);
        const int nb = tile_holders<T>(b, 'N', bh, bop);
        int ntemp = (na + nb == 3) ? 1 : 2;
        int a_idx = na - 1, b_idx = 0, g = 0;
        int32_t rel = 0;
        size_t off = 0;
        while (ntemp > 0) {
            const Holder<T> ad = ah[a_idx], bd = bh[b_idx];
            const int64_t mm = ad.rows, nn = bd.cols, kk = std::min(ad.cols, bd.rows);
            T *dst = reinterpret_cast<T *>((uintptr_t) off);
            if (e) {
                GemmSpec<T> &gs = e->gemm[g];
                gs.a = ad.data;
                gs.b = bd.data;
                gs.lda = std::max(ad.rows, mm);
                gs.ldb = std::max(bd.rows, kk);
                gs.ldc = mm;
                gs.m = (int32_t) mm;
                gs.n = (int32_t) nn;
                gs.k = (int32_t) kk;
                gs.ta = gs.tb = 0;
                gs.c = dst;
                rel |= (atmp[a_idx] ? (g ? REL_G1A : REL_G0A) : 0) | (btmp[b_idx] ? (g ? REL_G1B : REL_G0B) : 0) |
                       (g ? REL_G1C : REL_G0C);
            }
            g++;
            off += aligned(sizeof(T) * std::max<int64_t>(mm * nn, 1));
            const Holder<T> th = {mm, nn, mm, dst};
            if (ntemp == 2 || (ntemp == 1 && a_idx == 0)) {
                ah[a_idx] = th;
                atmp[a_idx] = true;
                b_idx++;
            } else {
                bh[b_idx] = th;
                btmp[b_idx] = true;
                a_idx--;
            }
            ntemp--;
        }
        const Holder<T> ad = ah[a_idx], bd = bh[b_idx];
        const int64_t m = c->rows, nn = c->cols, arank = ad.cols;
        const int64_t kmax = arank + c->max_rank;
        const WsLayout L = ws_layout<T>(m, nn, kmax, c->max_rank);
        if (e) {
            e->ngemm = g;
            RecompDesc<T> &d = e->rd;
            d.afin = ad.data;
            d.ldaf = ad.ld;
            d.bfin = bd.data;
            d.ldbf = arank;
            d.cdata = (T *) c->data;
            d.ws = reinterpret_cast<T *>((uintptr_t) off);
            rel |= (atmp[a_idx] ? REL_AFIN : 0) | (btmp[b_idx] ? REL_BFIN : 0) | REL_WS;
            e->rel = rel;
            d.m = (int32_t) m;
            d.n = (int32_t) nn;
            d.arank = (int32_t) arank;
            d.max_rank = (int32_t) c->max_rank;
            d.kmax = (int32_t) kmax;
            d.truncated_svd = p.truncated_svd;
            d.fixed_rank = (int32_t) p.fixed_rank;
            d.rb = (b->type == HCOREPP_COMPRESSED) ? (int32_t) b->rank : 0;
            d.rank_model = nullptr;
            d.model_tol = 0.0;
            d.model_margin = 0;
            d.alpha = alpha;
            d.beta = beta;
            d.accuracy = p.accuracy;
            d.sigma_out = nullptr;
        }
        if (cost) {
            const double kk = (double) (arank + c->rank);
            *cost = kk * kk * (double) std::max(m, nn) + 20.0 * kk * kk * kk;
        }
        if (rows) *rows = std::max(m, nn);
        if (kst) *kst = arank + c->rank;
        if (e && (ad.rows != m || bd.cols != nn || bd.rows != arank)) return 0;
        return off + aligned(sizeof(T) * L.total);
    };
    size_t slot_bytes = 256;
    std::vector<double> cost((size_t) n);
    int64_t max_rows = 1;
    double sum_k = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        int64_t rows = 0, kst = 0;
        slot_bytes = std::max(slot_bytes, aligned(plan(i, nullptr, &cost[(size_t) i], &rows, &kst)));
        max_rows = std::max(max_rows, rows);
        sum_k += (double) kst;
        const int64_t mu = std::min<int64_t>(std::max<int64_t>(C[i].rows, 0), kst);
        if (mu > JACOBI_ROWS_MAX)
            return fail(HCOREPP_ERR_UNSUPPORTED, "update %lld: stacked rank %lld exceeds the core SVD limit %d",
                        (long long) i, (long long) kst, JACOBI_ROWS_MAX);
    }
    const double mean_k = n > 0 ? sum_k / (double) n : 0.0;
    const int64_t lds_elems = kLdsBytes / (int64_t) sizeof(T);
    // Panel width and residency from the autotune table for this (rows, K, size) class; the
    // built-in rule when the table has no entry or its panel does not fit these tiles.
    int pb = 0, wg_per_cu = 2;  // two 78 KB-LDS, 256-VGPR workgroups fill a CU
    int tuned_pb = 0, tuned_wg = 0;
    const double per_cu = (double) n / (double) std::max(ctx->num_cus, 1);
    if (tune_lookup(max_rows, mean_k, per_cu, &tuned_pb, &tuned_wg) && pb_fits(max_rows, tuned_pb, lds_elems)) {
        pb = tuned_pb;
        wg_per_cu = tuned_wg;
    } else {
        pb = pick_pb(max_rows, lds_elems);
    }
    if (pb == 0) return fail(HCOREPP_ERR_UNSUPPORTED, "tile too tall for the LDS panel");

    hcorepp_batch *bt = new (std::nothrow) hcorepp_batch();
    if (!bt) return fail(HCOREPP_ERR_NOMEM, "host allocation failed");
    bt->ctx = ctx;
    bt->precision = precision;
    bt->n = n;
    bt->pb = pb;
    bt->max_rows = max_rows;
    bt->slot_bytes = slot_bytes;
    bt->grid = (int) std::max<int64_t>(1, std::min<int64_t>(n, (int64_t) std::max(ctx->num_cus, 1) * wg_per_cu));
    auto cleanup = [&](hcorepp_status s2) {
        hcorepp_batch_destroy(bt);
        return s2;
    };
    if ((st = ensure_slots(bt, bt->grid))) return cleanup(st);
    if (hipMalloc(&bt->hdrs, sizeof(int64_t) * HDR_WORDS * (size_t) std::max<int64_t>(n, 1)) != hipSuccess)
        return cleanup(fail(HCOREPP_ERR_NOMEM, "batch headers"));
    if (hipMalloc(&bt->entries, sizeof(BatchEntry<T>) * (size_t) std::max<int64_t>(n, 1)) != hipSuccess ||
        hipMalloc(&bt->order, sizeof(int32_t) * (size_t) std::max<int64_t>(n, 1)) != hipSuccess ||
        hipMalloc(&bt->ranks, sizeof(int64_t) * (size_t) std::max<int64_t>(n, 1)) != hipSuccess ||
        hipMalloc(&bt->ctr, sizeof(BatchCounters)) != hipSuccess)
        return cleanup(fail(HCOREPP_ERR_NOMEM, "batch descriptors"));
    std::vector<BatchEntry<T>> ent((size_t) n);
    std::vector<int64_t> ranks((size_t) n);
    for (int64_t i = 0; i < n; ++i) {
        BatchEntry<T> &e = ent[(size_t) i];
        memset(&e, 0, sizeof(e));
        if (plan(i, &e, nullptr, nullptr, nullptr) == 0)
            return cleanup(fail(HCOREPP_ERR_INVALID, "operand shapes do not match C (update %lld)", (long long) i));
        e.rd.crank = bt->ranks + i;
        ranks[(size_t) i] = C[i].rank;
    }
    bt->ent_host.assign((const char *) ent.data(), (const char *) ent.data() + sizeof(BatchEntry<T>) * (size_t) n);
    std::vector<int32_t> order((size_t) n);
    for (int64_t i = 0; i < n; ++i) order[(size_t) i] = (int32_t) i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return cost[x] > cost[y]; });
    hipStream_t s = ctx->stream;
    // counters start at zero, so that status reads before the first pass (or of an empty batch) are clean
    if (launch_reset_counters(s, bt->ctr) != hipSuccess ||
        hipMemsetAsync(bt->hdrs, 0, sizeof(int64_t) * HDR_WORDS * (size_t) std::max<int64_t>(n, 1), s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return cleanup(fail(HCOREPP_ERR_HIP, "batch counter reset"));
    if (n > 0) {
        if (hipMemcpyAsync(bt->entries, ent.data(), sizeof(BatchEntry<T>) * n, hipMemcpyHostToDevice, s) !=
                    hipSuccess ||
            hipMemcpyAsync(bt->order, order.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(bt->ranks, ranks.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return cleanup(fail(HCOREPP_ERR_HIP, "batch descriptor upload"));
    }
    *out = bt;
    return HCOREPP_OK;
}

hcorepp_status hcorepp_batch_run(hcorepp_batch *b) {
    if (!b) return fail(HCOREPP_ERR_INVALID, "null batch");
    if (b->n == 0) return HCOREPP_OK;
    hipStream_t s = b->ctx->stream;
    HIP_TRY(launch_reset_counters(s, b->ctr));
    if (b->grid > b->slots) return fail(HCOREPP_ERR_INVALID, "grid %d exceeds the %d workspace slots", b->grid, b->slots);
    HIP_TRY(launch_batched<double>(s, (const BatchEntry<double> *) b->entries, b->order, b->n, b->ctr, b->grid,
                                   kLdsBytes, b->pb, b->precision == HCOREPP_PREC_MIXED ? 1 : 0, b->ws,
                                   (int64_t) b->slot_bytes, b->hdrs));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_batch_set_grid(hcorepp_batch *b, int workgroups) {
    if (!b || workgroups < 1) return fail(HCOREPP_ERR_INVALID, "bad arguments");
    const int grid = (int) std::min<int64_t>(workgroups, std::max<int64_t>(b->n, 1));
    drop_graph(b);
    hcorepp_status st = ensure_slots(b, grid);
    if (st) return st;
    b->grid = grid;
    return HCOREPP_OK;
}

int hcorepp_batch_get_grid(hcorepp_batch *b) { return b ? b->grid : 0; }

hcorepp_status hcorepp_batch_set_panel(hcorepp_batch *b, int pb) {
    if (!b) return fail(HCOREPP_ERR_INVALID, "null batch");
    if (!pb_fits(b->max_rows, pb, kLdsBytes / (int64_t) sizeof(double)))
        return fail(HCOREPP_ERR_INVALID, "panel width %d does not fit a QR of %lld rows", pb, (long long) b->max_rows);
    b->pb = pb;
    drop_graph(b);
    return HCOREPP_OK;
}

int hcorepp_batch_get_panel(hcorepp_batch *b) { return b ? b->pb : 0; }

int hcorepp_autotune_lookup(int64_t rows, double mean_k, double updates_per_cu, int *pb, int *workgroups_per_cu) {
    int p = 0, w = 0;
    const int hit = tune_lookup(rows, mean_k, updates_per_cu, &p, &w) ? 1 : 0;
    if (pb) *pb = hit ? p : pick_pb(rows, kLdsBytes / (int64_t) sizeof(double));
    if (workgroups_per_cu) *workgroups_per_cu = hit ? w : 2;
    return hit;
}

hcorepp_status hcorepp_batch_phase_times(hcorepp_batch *b, int64_t *times) {
    if (!b || (!times && b->n > 0)) return fail(HCOREPP_ERR_INVALID, "bad arguments");
#ifndef HCAMD_PHASE_TIMING
    return fail(HCOREPP_ERR_UNSUPPORTED, "library built without HCAMD_PHASE_TIMING");
#else
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    for (int64_t i = 0; i < b->n; ++i)
        HIP_TRY(hipMemcpy(times + 16 * i, hdr_word(b, i, HDR_T0), 16 * sizeof(int64_t), hipMemcpyDeviceToHost));
    return HCOREPP_OK;
#endif
}

hcorepp_status hcorepp_batch_capture_graph(hcorepp_batch *b) {
    if (!b) return fail(HCOREPP_ERR_INVALID, "null batch");
    hipStream_t s = b->ctx->stream;
    drop_graph(b);
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hcorepp_status st = hcorepp_batch_run(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (st) {
        if (g) hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return fail(HCOREPP_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_batch_launch_graph(hcorepp_batch *b) {
    if (!b || !b->graph_exec) return fail(HCOREPP_ERR_INVALID, "no captured graph");
    HIP_TRY(hipGraphLaunch(b->graph_exec, b->ctx->stream));
    return HCOREPP_OK;
}

int64_t *hcorepp_batch_rank_dev(hcorepp_batch *b) { return b ? b->ranks : nullptr; }

hcorepp_status hcorepp_batch_ranks(hcorepp_batch *b, int64_t *ranks_out) {
    if (!b || !ranks_out) return fail(HCOREPP_ERR_INVALID, "bad arguments");
    if (b->n == 0) return HCOREPP_OK;
    HIP_TRY(hipMemcpyAsync(ranks_out, b->ranks, sizeof(int64_t) * b->n, hipMemcpyDeviceToHost, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_batch_set_ranks(hcorepp_batch *b, const int64_t *ranks) {
    if (!b || !ranks) return fail(HCOREPP_ERR_INVALID, "bad arguments");
    if (b->n == 0) return HCOREPP_OK;
    HIP_TRY(hipMemcpyAsync(b->ranks, ranks, sizeof(int64_t) * b->n, hipMemcpyHostToDevice, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_batch_status(hcorepp_batch *b, int64_t *max_sweeps) {
    if (!b) return fail(HCOREPP_ERR_INVALID, "null batch");
    HIP_TRY(hipMemcpyAsync(b->ctx->pinned, b->ctr, sizeof(BatchCounters), hipMemcpyDeviceToHost, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    const BatchCounters *c = reinterpret_cast<const BatchCounters *>(b->ctx->pinned);
    if (max_sweeps) *max_sweeps = (int64_t) c->max_sweeps;
    if (c->status) return fail(HCOREPP_ERR_NUMERIC, "batch status bits 0x%llx", (unsigned long long) c->status);
    return HCOREPP_OK;
}

size_t hcorepp_batch_workspace_bytes(hcorepp_batch *b) {
    return b ? b->slot_bytes * (size_t) b->slots + sizeof(int64_t) * HDR_WORDS * (size_t) b->n : 0;
}

hcorepp_status hcorepp_batch_set_rank_model(hcorepp_batch *b, const uint16_t *model_bf16, double tol, int margin) {
    if (!b || tol < 0.0 || (model_bf16 && (!(tol > 0.0) || margin < 0))) return fail(HCOREPP_ERR_INVALID, "bad arguments");
    BatchEntry<double> *ent = reinterpret_cast<BatchEntry<double> *>(b->ent_host.data());
    for (int64_t i = 0; i < b->n; ++i) {
        ent[i].rd.rank_model = model_bf16;
        ent[i].rd.model_tol = tol;
        ent[i].rd.model_margin = margin;
    }
    if (b->n > 0) {
        HIP_TRY(hipMemcpyAsync(b->entries, ent, sizeof(BatchEntry<double>) * (size_t) b->n, hipMemcpyHostToDevice,
                               b->ctx->stream));
        HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    }
    return HCOREPP_OK;
}

hcorepp_status hcorepp_batch_features(hcorepp_batch *b, float *features) {
    if (!b || (!features && b->n > 0)) return fail(HCOREPP_ERR_INVALID, "bad arguments");
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    for (int64_t i = 0; i < b->n; ++i)
        HIP_TRY(hipMemcpy(features + (size_t) RANK_FEATURES * i, hdr_word(b, i, HDR_FEAT),
                          sizeof(float) * RANK_FEATURES, hipMemcpyDeviceToHost));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_batch_learned_stats(hcorepp_batch *b, int64_t *fallbacks, int64_t *predicted) {
    if (!b) return fail(HCOREPP_ERR_INVALID, "null batch");
    HIP_TRY(hipMemcpyAsync(b->ctx->pinned, b->ctr, sizeof(BatchCounters), hipMemcpyDeviceToHost, b->ctx->stream));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    const BatchCounters *c = reinterpret_cast<const BatchCounters *>(b->ctx->pinned);
    if (fallbacks) *fallbacks = (int64_t) c->fallbacks;
    if (predicted) {
        for (int64_t i = 0; i < b->n; ++i)
            HIP_TRY(hipMemcpy(predicted + i, hdr_word(b, i, HDR_PRED), sizeof(int64_t), hipMemcpyDeviceToHost));
    }
    return HCOREPP_OK;
}

hcorepp_status hcorepp_batch_core_widths(hcorepp_batch *b, int64_t *widths) {
    if (!b || (b->n > 0 && !widths)) return fail(HCOREPP_ERR_INVALID, "bad arguments");
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    for (int64_t i = 0; i < b->n; ++i)
        HIP_TRY(hipMemcpy(widths + i, hdr_word(b, i, HDR_KT), sizeof(int64_t), hipMemcpyDeviceToHost));
    return HCOREPP_OK;
}

// ---------------------------------------------------------------------------
// TLR Cholesky tile operations (FP64).
extern "C++" {
namespace {
template <typename J>
hcorepp_status upload_jobs(hcorepp_ctx *ctx, const std::vector<J> &jobs, J **dev) {
    *dev = (J *) arena_take(ctx, sizeof(J) * std::max<size_t>(jobs.size(), 1));
    if (!jobs.empty()) HIP_TRY(hipMemcpy(*dev, jobs.data(), sizeof(J) * jobs.size(), hipMemcpyHostToDevice));
    return HCOREPP_OK;
}
}  // namespace
}

hcorepp_status hcorepp_potrf_d(hcorepp_ctx *ctx, hcorepp_tile *A) {
    hcorepp_status st;
    if (!ctx) return fail(HCOREPP_ERR_INVALID, "null context");
    if ((st = validate_tile<double>(A, "A"))) return st;
    if (A->type != HCOREPP_DENSE || A->rows != A->cols)
        return fail(HCOREPP_ERR_INVALID, "potrf needs a square dense tile");
    if ((st = arena_begin(ctx, 4096))) return st;
    int *info = (int *) arena_take(ctx, sizeof(int));
    HIP_TRY(launch_potrf(ctx->stream, (double *) A->data, (int) A->rows, A->ld, info));
    int h = 0;
    HIP_TRY(hipMemcpyAsync(&h, info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (h) return fail(HCOREPP_ERR_NUMERIC, "potrf: tile not positive definite (pivot %d)", h);
    return HCOREPP_OK;
}

hcorepp_status hcorepp_trsm_batch_d(hcorepp_ctx *ctx, int64_t n, const hcorepp_tile *L, hcorepp_tile *B) {
    hcorepp_status st;
    if (!ctx || n < 0 || (n > 0 && !B)) return fail(HCOREPP_ERR_INVALID, "bad arguments");
    if ((st = validate_tile<double>(L, "L"))) return st;
    if (L->type != HCOREPP_DENSE || L->rows != L->cols) return fail(HCOREPP_ERR_INVALID, "L must be square dense");
    std::vector<CholTrsm> jobs;
    for (int64_t i = 0; i < n; ++i) {
        if ((st = validate_tile<double>(&B[i], "B[i]"))) return st;
        if (B[i].type != HCOREPP_COMPRESSED || B[i].cols != L->rows)
            return fail(HCOREPP_ERR_INVALID, "trsm: B[%lld] must be compressed with cols == L rows", (long long) i);
        if (B[i].rank == 0) continue;
        jobs.push_back(CholTrsm{(double *) B[i].data + B[i].rows * B[i].max_rank, (const double *) L->data,
                                B[i].rank, L->rows, L->ld});
    }
    if ((st = arena_begin(ctx, sizeof(CholTrsm) * (jobs.size() + 1) + 4096))) return st;
    CholTrsm *dj;
    if ((st = upload_jobs(ctx, jobs, &dj))) return st;
    HIP_TRY(launch_trsm_v(ctx->stream, dj, (int) jobs.size()));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_syrk_batch_d(hcorepp_ctx *ctx, int64_t n, const hcorepp_tile *A, hcorepp_tile *C) {
    hcorepp_status st;
    if (!ctx || n < 0 || (n > 0 && (!A || !C))) return fail(HCOREPP_ERR_INVALID, "bad arguments");
    std::vector<CholSyrk> jobs;
    size_t scratch = 0;
    for (int64_t i = 0; i < n; ++i) {
        if ((st = validate_tile<double>(&A[i], "A[i]")) || (st = validate_tile<double>(&C[i], "C[i]"))) return st;
        if (A[i].type != HCOREPP_COMPRESSED || C[i].type != HCOREPP_DENSE || C[i].rows != C[i].cols ||
            A[i].rows != C[i].rows)
            return fail(HCOREPP_ERR_INVALID, "syrk: A[%lld] compressed, C[%lld] square dense of A's rows",
                        (long long) i, (long long) i);
        scratch += aligned(sizeof(double) * (size_t) (A[i].rank * A[i].rank)) +
                   aligned(sizeof(double) * (size_t) (A[i].rows * A[i].rank));
    }
    if ((st = arena_begin(ctx, scratch + sizeof(CholSyrk) * (size_t) (n + 1) + 4096))) return st;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t r = A[i].rank, m = A[i].rows;
        double *w = (double *) arena_take(ctx, sizeof(double) * (size_t) (r * r));
        double *t = (double *) arena_take(ctx, sizeof(double) * (size_t) (m * r));
        const double *u = (const double *) A[i].data;
        jobs.push_back(CholSyrk{(double *) C[i].data, u, u + m * A[i].max_rank, w, t, m, r, C[i].ld});
    }
    CholSyrk *dj;
    if ((st = upload_jobs(ctx, jobs, &dj))) return st;
    HIP_TRY(launch_syrk(ctx->stream, dj, (int) jobs.size()));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_transpose_batch_d(hcorepp_ctx *ctx, int64_t n, const hcorepp_tile *src, hcorepp_tile *dst) {
    hcorepp_status st;
    if (!ctx || n < 0 || (n > 0 && (!src || !dst))) return fail(HCOREPP_ERR_INVALID, "bad arguments");
    std::vector<CholTrans> jobs;
    for (int64_t i = 0; i < n; ++i) {
        if ((st = validate_tile<double>(&src[i], "src[i]")) || (st = validate_tile<double>(&dst[i], "dst[i]")))
            return st;
        if (src[i].type != HCOREPP_COMPRESSED || dst[i].type != HCOREPP_COMPRESSED || dst[i].rows != src[i].cols ||
            dst[i].cols != src[i].rows || dst[i].max_rank < src[i].rank)
            return fail(HCOREPP_ERR_INVALID, "transpose: dst[%lld] must be a compressed %lldx%lld tile of capacity "
                        ">= %lld", (long long) i, (long long) src[i].cols, (long long) src[i].rows,
                        (long long) src[i].rank);
        dst[i].rank = src[i].rank;
        jobs.push_back(CholTrans{(const double *) src[i].data, (double *) dst[i].data, src[i].rows, src[i].cols,
                                 src[i].rank, src[i].max_rank, dst[i].max_rank});
    }
    if ((st = arena_begin(ctx, sizeof(CholTrans) * (jobs.size() + 1) + 4096))) return st;
    CholTrans *dj;
    if ((st = upload_jobs(ctx, jobs, &dj))) return st;
    HIP_TRY(launch_transpose_factors(ctx->stream, dj, (int) jobs.size()));
    return HCOREPP_OK;
}

// ---------------------------------------------------------------------------
// Rank predictor training (FP32 on the device).
int hcorepp_rank_model_features(void) { return RANK_FEATURES; }
size_t hcorepp_rank_model_params(void) { return (size_t) RANK_PARAMS; }
size_t hcorepp_mlp_workspace_floats(int64_t n) { return (size_t) mlp_workspace_floats(n); }

hcorepp_status hcorepp_mlp_forward_f32(hcorepp_ctx *ctx, const float *params, const float *X, int64_t n, float *ws) {
    if (!ctx || !params || !ws || (n > 0 && !X)) return fail(HCOREPP_ERR_INVALID, "bad arguments");
    HIP_TRY(launch_mlp_forward(ctx->stream, params, X, n, ws));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_mlp_grad_f32(hcorepp_ctx *ctx, const float *params, const float *X, const float *y, int64_t n,
                                    float *ws, float *grads, float *loss) {
    if (!ctx || !params || !ws || !grads || n <= 0 || !X || !y) return fail(HCOREPP_ERR_INVALID, "bad arguments");
    HIP_TRY(launch_mlp_forward(ctx->stream, params, X, n, ws));
    HIP_TRY(launch_mlp_backward(ctx->stream, params, X, y, n, ws, grads));
    if (loss) {
        const int64_t nb = mlp_loss_blocks(n);
        std::vector<float> part((size_t) nb);
        HIP_TRY(hipMemcpyAsync(part.data(), mlp_loss_partials(ws, n), sizeof(float) * (size_t) nb,
                               hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        double t = 0.0;
        for (float v : part) t += v;
        *loss = (float) t;
    }
    return HCOREPP_OK;
}

hcorepp_status hcorepp_mlp_adam_f32(hcorepp_ctx *ctx, float *params, const float *grads, float *m, float *v, float lr,
                                    float beta1, float beta2, float eps, int64_t step) {
    if (!ctx || !params || !grads || !m || !v || step < 1) return fail(HCOREPP_ERR_INVALID, "bad arguments");
    HIP_TRY(launch_mlp_adam(ctx->stream, params, grads, m, v, lr, beta1, beta2, eps, step));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_mlp_pack_bf16(hcorepp_ctx *ctx, const float *params, uint16_t *out) {
    if (!ctx || !params || !out) return fail(HCOREPP_ERR_INVALID, "bad arguments");
    HIP_TRY(launch_mlp_pack_bf16(ctx->stream, params, out));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_mlp_predict_bf16(hcorepp_ctx *ctx, const uint16_t *model_bf16, const float *X, int64_t n,
                                        float *out) {
    if (!ctx || !model_bf16 || (n > 0 && (!X || !out))) return fail(HCOREPP_ERR_INVALID, "bad arguments");
    HIP_TRY(launch_mlp_predict_bf16(ctx->stream, model_bf16, X, n, out));
    return HCOREPP_OK;
}

hcorepp_status hcorepp_batch_destroy(hcorepp_batch *b) {
    if (!b) return HCOREPP_OK;
    if (b->ctx) hipStreamSynchronize(b->ctx->stream);
    if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
    if (b->graph) hipGraphDestroy(b->graph);
    if (b->ws) hipFree(b->ws);
    if (b->hdrs) hipFree(b->hdrs);
    if (b->entries) hipFree(b->entries);
    if (b->order) hipFree(b->order);
    if (b->ranks) hipFree(b->ranks);
    if (b->ctr) hipFree(b->ctr);
    delete b;
    return HCOREPP_OK;
}

// Diagnostics: copy the workspace of the last plain recompression on this context to `host`
// (at most `cap` bytes) and its layout to info[16] = {m, n, K, kmax, max_rank, sizeof(T), total,
// hdr, us, vs, y, t_u, t_y, j, v1, vtmp} (offsets in elements).  Returns bytes copied or < 0.
int64_t hcorepp_debug_workspace(hcorepp_ctx *ctx, void *host, size_t cap, int64_t *info) {
    if (!ctx || !ctx->last_ws) return -1;
    for (int i = 0; i < 16; ++i) info[i] = ctx->last_info[i];
    size_t bytes = (size_t) (ctx->last_info[6] * ctx->last_info[5]);
    if (bytes > cap) bytes = cap;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return -2;
    if (hipMemcpy(host, ctx->last_ws, bytes, hipMemcpyDeviceToHost) != hipSuccess) return -3;
    return (int64_t) bytes;
}

const char *hcorepp_last_error(void) { return g_last_error.c_str(); }

const char *hcorepp_version(void) { return "hcorepp-amd 0.1 (gfx950)"; }

}  // extern "C"
