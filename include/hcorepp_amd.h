/*
 * hcorepp_amd.h -- C-ABI of the MI355X-native TLR tile GEMM + recompression engine.
 *
 * This is the drop-in boundary: plain pointers and sizes, no C++ or torch types.
 * The C++ surface in include/hcorepp/ (HCore<T>::Gemm, DenseTile<T>,
 * CompressedTile<T>, CompressionParameters, RunContext, MemoryUnit) forwards to
 * these entry points.  Each declaration names the reference interface it
 * replaces (paths relative to the ecrc/hcorepp source tree).
 *
 * Conventions
 *  - All matrices are column-major (blas::Layout::ColMajor).
 *  - Tile data pointers are device pointers (HBM).  A compressed tile's buffer
 *    holds [U rows x max_rank (ld rows) | V max_rank x cols (ld rank)]
 *    (reference src/operators/concrete/Compressed.cpp:175-185).
 *  - Functions return HCOREPP_OK (0) or a negative hcorepp_status; nothing
 *    throws across the boundary.  hcorepp_last_error() describes the last
 *    failure on the calling thread.
 *  - One hcorepp_ctx per host thread.  Calls on different contexts are
 *    thread-safe; calls sharing a context must be serialised by the caller.
 */
#ifndef HCOREPP_AMD_H
#define HCOREPP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    HCOREPP_OK = 0,
    HCOREPP_ERR_INVALID = -1,      /* bad argument (null pointer, inconsistent sizes) */
    HCOREPP_ERR_UNSUPPORTED = -2,  /* combination the reference does not define (see hcorepp_gemm_d) */
    HCOREPP_ERR_CAPACITY = -3,     /* output does not fit the tile's max_rank capacity */
    HCOREPP_ERR_HIP = -4,          /* HIP runtime error */
    HCOREPP_ERR_NUMERIC = -5,      /* Jacobi SVD did not converge or NaN/Inf encountered */
    HCOREPP_ERR_NOMEM = -6         /* device allocation failed */
} hcorepp_status;

/* TileType (reference include/hcorepp/operators/interface/Tile.hpp:21-24). */
typedef enum { HCOREPP_DENSE = 0, HCOREPP_COMPRESSED = 1 } hcorepp_tile_type;

/* SVD implementation selector (reference include/hcorepp/common/Definitions.hpp:101-104).
 * The device path uses one-sided Jacobi for both; the value is accepted for parity. */
typedef enum { HCOREPP_SVD_GESVD = 0, HCOREPP_SVD_GESDD = 1 } hcorepp_svd_type;

/* Precision form of a batch (north-star forms). */
typedef enum {
    HCOREPP_PREC_FP64 = 0,   /* everything in FP64 */
    HCOREPP_PREC_MIXED = 1   /* factor products (the temporary tiles of HCore.cpp:176-264) on the FP32
                                matrix cores with FP32 accumulation; stacking, both QRs, the core SVD, the
                                truncation decision and the reconstruction in FP64 */
} hcorepp_precision;

/* TileMetadata + data pointer (reference include/hcorepp/operators/interface/Tile.hpp:30-52).
 * `rank` is written back by the callee for compressed outputs. */
typedef struct {
    int64_t rows;      /* mNumOfRows */
    int64_t cols;      /* mNumOfCols */
    int64_t rank;      /* mMatrixRank (compressed tiles) */
    int64_t max_rank;  /* mMaxRank: capacity of the U/V slots */
    int64_t ld;        /* mLeadingDimension (dense tiles) */
    int32_t layout;    /* 0 = ColMajor (the only layout the reference's kernels implement) */
    int32_t type;      /* hcorepp_tile_type */
    void *data;        /* device pointer */
} hcorepp_tile;

/* CompressionParameters (reference include/hcorepp/operators/helpers/CompressionParameters.hpp:44-46).
 * trmm / ungqr / svd_type select LAPACK call sequences in the reference; the device path computes
 * the same truncated SVD regardless and accepts them for parity. */
typedef struct {
    double accuracy;        /* absolute truncation threshold (default 1e-4) */
    int32_t trmm;           /* default 0 */
    int32_t ungqr;          /* default 1 */
    int32_t truncated_svd;  /* default 0: relative threshold accuracy*sigma_0 when 1 */
    int32_t svd_type;       /* hcorepp_svd_type, default GESDD */
    int64_t fixed_rank;     /* default 0: truncate to this rank when > 0 */
} hcorepp_cparams;

/* Opaque execution context: one HIP stream, one growable device arena, pinned host mirror.
 * Replaces RunContext (reference src/kernels/cuda/RunContext.cpp:11-78) and the pool
 * MemoryUnit (reference include/hcorepp/data-units/memory-handlers/pool/MemoryHandler.hpp:22-157). */
typedef struct hcorepp_ctx hcorepp_ctx;

/* Replaces RunContext() + MemoryHandler::Initialize (reference src/kernels/cuda/RunContext.cpp:11-20,
 * src/data-units/memory-handlers/pool/MemoryHandler.cpp:23-35).  arena_bytes may be 0 (grown on demand). */
hcorepp_status hcorepp_ctx_create(int device, size_t arena_bytes, hcorepp_ctx **out);
hcorepp_status hcorepp_ctx_destroy(hcorepp_ctx *ctx);
/* RunContext::Sync (reference src/kernels/cuda/RunContext.cpp:48-50). */
hcorepp_status hcorepp_ctx_sync(hcorepp_ctx *ctx);
/* Returns the context's hipStream_t (as void*) so callers can order their own work on it. */
void *hcorepp_ctx_stream(hcorepp_ctx *ctx);
/* Bytes currently reserved by the context arena. */
size_t hcorepp_ctx_arena_bytes(hcorepp_ctx *ctx);
/* Compute units of the context's device (no counterpart in the reference). */
int hcorepp_ctx_compute_units(hcorepp_ctx *ctx);

/* Device memory helpers (reference include/hcorepp/kernels/memory.hpp:46-121:
 * AllocateArray, DestroyArray, Memcpy (cudaMemcpyDefault semantics), Memset). */
hcorepp_status hcorepp_malloc(size_t bytes, void **out);
hcorepp_status hcorepp_free(void *ptr);
hcorepp_status hcorepp_memcpy(hcorepp_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Stream-ordered copy on the context stream, no synchronisation. */
hcorepp_status hcorepp_memcpy_async(hcorepp_ctx *ctx, void *dst, const void *src, size_t bytes);

/* Dense column-major C = alpha op(A) op(B) + beta C on device pointers, ordered on the context
 * stream (no synchronisation); trans 'N' or 'T'.  Replaces HCoreKernels<T>::Gemm
 * (reference include/hcorepp/kernels/kernels.hpp:27-29), ColMajor only. */
hcorepp_status hcorepp_kernel_gemm_d(hcorepp_ctx *ctx, char transa, char transb, int64_t m, int64_t n, int64_t k,
                                     double alpha, const double *A, int64_t lda, const double *B, int64_t ldb,
                                     double beta, double *C, int64_t ldc);
hcorepp_status hcorepp_memset(hcorepp_ctx *ctx, void *dst, int value, size_t bytes);

/* Stream-ordered timing on the context stream: record event `slot` (0..15) and read the
 * elapsed milliseconds between two recorded slots (synchronises on the second). */
hcorepp_status hcorepp_ctx_event_record(hcorepp_ctx *ctx, int slot);
hcorepp_status hcorepp_ctx_event_elapsed(hcorepp_ctx *ctx, int slot0, int slot1, float *ms);

/* ---------------------------------------------------------------------------
 * One tile update C = alpha * op(A) * op(B) + beta * C.
 * Replaces hcorepp::api::HCore<T>::Gemm (reference include/hcorepp/api/HCore.hpp:41-46,
 * src/api/HCore.cpp:22-344) including DenseTile::Gemm (src/operators/concrete/Dense.cpp:46-108)
 * and CompressedTile::Gemm (src/operators/concrete/Compressed.cpp:208-694).
 *  - All 8 dense/compressed combinations.  Compressed C requires opA = opB = 'N' (the
 *    reference's CompressedTile::Gemm ignores the ops, Compressed.cpp:331-380); other ops
 *    return HCOREPP_ERR_UNSUPPORTED.  Dense C supports 'N' and 'T'.
 *  - cholesky != 0: the Cholesky-flow Gemm (the reference's aCholesky, HCore.cpp:160-174):
 *    compressed A, B, C in that flow's convention (V stored cols x rank), opA 'N', opB 'T',
 *    C <- beta C + alpha A B^T; other combinations return HCOREPP_ERR_UNSUPPORTED.  (The
 *    batched TLR Cholesky driver instead passes the transposed panel tile explicitly,
 *    hcorepp_transpose_batch_d below, and stays in the [U | V r x n] convention.)
 *  - Plain form: the recompression runs as a chain of launches on the context stream and the
 *    new rank is read back with one stream synchronisation, as the reference does
 *    (src/kernels/cuda/CudaKernels.cu:656-697).  C->rank is updated on return.
 *  - flops (may be NULL) is incremented with the reference's own flop model.
 *  - sigma_out (may be NULL, host pointer) receives the sorted core singular values
 *    (size_s = min(rows, cols, K) entries) for compressed C; *n_sigma their count. */
hcorepp_status hcorepp_gemm_d(hcorepp_ctx *ctx, double alpha, const hcorepp_tile *A, char opA,
                              const hcorepp_tile *B, char opB, double beta, hcorepp_tile *C,
                              const hcorepp_cparams *params, int cholesky, uint64_t *flops,
                              double *sigma_out, int64_t *n_sigma);
/* FP32 instantiation (reference HCOREPP_INSTANTIATE_CLASS, include/hcorepp/common/Definitions.hpp:10-13). */
hcorepp_status hcorepp_gemm_s(hcorepp_ctx *ctx, float alpha, const hcorepp_tile *A, char opA,
                              const hcorepp_tile *B, char opB, float beta, hcorepp_tile *C,
                              const hcorepp_cparams *params, int cholesky, uint64_t *flops,
                              double *sigma_out, int64_t *n_sigma);

/* Allocation-free variant of hcorepp_gemm_d (FP64, no cholesky, no sigma output).  No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Instead of
 * the context arena, every kernel's scratch (temporary factor products, the recompression
 * workspace, panel and flag buffers) lives in a fixed device array of
 * hcorepp_gemm_noalloc_scratch_bytes() bytes that each GPU holds for the life of the process;
 * the call never allocates or frees device memory.  An update that needs more scratch returns
 * HCOREPP_ERR_CAPACITY (16 MiB covers nb = 512 tiles with stacked ranks up to about 400).
 * All calls on one GPU share that GPU's array, which keeps the last call's contents: calls that
 * overlap on one GPU (several contexts or threads) race and must be serialised by the caller.
 * Same results as hcorepp_gemm_d; it synchronises the context stream once, as that does. */
hcorepp_status hcorepp_gemm_noalloc_d(hcorepp_ctx *ctx, double alpha, const hcorepp_tile *A, char opA,
                                      const hcorepp_tile *B, char opB, double beta, hcorepp_tile *C,
                                      const hcorepp_cparams *params, uint64_t *flops);
/* Size in bytes of the per-GPU scratch array of hcorepp_gemm_noalloc_d. */
size_t hcorepp_gemm_noalloc_scratch_bytes(void);

/* Compression constructor on the device: replaces CompressedTile<T>(rows, cols, apData, ld,
 * CompressionParameters, layout, context) (reference src/operators/concrete/Compressed.cpp:75-146).
 * A: device pointer, rows x cols column-major (ld lda).  out: compressed tile descriptor with
 * rows/cols equal to A's, a device buffer of capacity [U rows x max_rank | V max_rank x cols]
 * and max_rank >= 1 (the reference uses max(min(rows, cols) / 3, 1)).  Rank rule as the
 * reference: fixed_rank if set (clamped to min(rows, cols)), else the first singular value
 * below accuracy (relative to sigma_0 with truncated_svd), at least 1, at most max_rank;
 * U orthonormal, V = Sigma V^T.  out->rank and out->ld are set; synchronises the stream once.
 * min(rows, cols) <= 768. */
hcorepp_status hcorepp_compress_d(hcorepp_ctx *ctx, int64_t rows, int64_t cols, const double *A, int64_t lda,
                                  const hcorepp_cparams *params, hcorepp_tile *out);
hcorepp_status hcorepp_compress_s(hcorepp_ctx *ctx, int64_t rows, int64_t cols, const float *A, int64_t lda,
                                  const hcorepp_cparams *params, hcorepp_tile *out);

/* Workspace (elements of T) one update C = A B + C (opA = opB = 'N') takes from the context arena:
 * reference HCore<T>::CalculateMemoryPoolSize (src/api/HCore.cpp:417-480), which the reference's
 * stream drivers use to presize one pool per stream (examples/matrix_multiplication/
 * par_fixed_rank_streams_main.cpp:86-122).  The engine's layout differs, so the number is this
 * engine's; a context created with arena_bytes = this * sizeof(T) never grows its arena for the
 * update. */
size_t hcorepp_gemm_pool_size_d(const hcorepp_tile *A, const hcorepp_tile *B, const hcorepp_tile *C,
                                const hcorepp_cparams *params);
size_t hcorepp_gemm_pool_size_s(const hcorepp_tile *A, const hcorepp_tile *B, const hcorepp_tile *C,
                                const hcorepp_cparams *params);

/* ---------------------------------------------------------------------------
 * Batched variable-rank form (no counterpart call in the reference: it replaces the
 * reference drivers' loop of HCore::Gemm calls over independent C tiles,
 * examples/matrix_multiplication/omp_main.cpp:112-126 and
 * par_fixed_rank_streams_main.cpp:104-137).  Each update is C_i = alpha*A_i*B_i + beta*C_i
 * with opA = opB = 'N'; every C_i must be distinct.  Ranks of C tiles live on the device
 * (rank_dev) and are updated in place; host descriptors are not touched by hcorepp_batch_run.
 * ------------------------------------------------------------------------- */
typedef struct hcorepp_batch hcorepp_batch;

hcorepp_status hcorepp_batch_create(hcorepp_ctx *ctx, int64_t n, const hcorepp_tile *A, const hcorepp_tile *B,
                                    const hcorepp_tile *C, double alpha, double beta,
                                    const hcorepp_cparams *params, int precision, hcorepp_batch **out);
/* Enqueue one pass over the batch on the context stream (capturable: no host sync, no allocation). */
hcorepp_status hcorepp_batch_run(hcorepp_batch *b);
/* Capture hcorepp_batch_run into a hipGraph; later calls of hcorepp_batch_launch_graph replay it. */
hcorepp_status hcorepp_batch_capture_graph(hcorepp_batch *b);
hcorepp_status hcorepp_batch_launch_graph(hcorepp_batch *b);
/* Device array (int64, n entries) holding the current rank of every C tile. */
int64_t *hcorepp_batch_rank_dev(hcorepp_batch *b);
/* Copy ranks to host (synchronises the context stream). */
hcorepp_status hcorepp_batch_ranks(hcorepp_batch *b, int64_t *ranks_out);
/* Reset the device ranks of the C tiles to the values given (host array). */
hcorepp_status hcorepp_batch_set_ranks(hcorepp_batch *b, const int64_t *ranks);
/* Worst status over the last pass (synchronises): 0, or HCOREPP_ERR_NUMERIC. Also max Jacobi sweeps. */
hcorepp_status hcorepp_batch_status(hcorepp_batch *b, int64_t *max_sweeps);
/* Launch shape of the persistent kernel (no counterpart in the reference): number of
 * workgroups pulling updates from the queue (default: the autotune table's workgroups per CU
 * for the batch's class times the CU count, at most n).  Setting it drops a captured graph;
 * capture again afterwards. */
hcorepp_status hcorepp_batch_set_grid(hcorepp_batch *b, int workgroups);
int hcorepp_batch_get_grid(hcorepp_batch *b);
/* Householder panel width of the batch's QRs (no counterpart in the reference; the
 * reference's QR is LAPACK's, whose block size is LAPACK's own choice).  Default: the
 * autotune table's width for the batch's class (tallest QR, mean stacked rank K = ra + rc,
 * updates per CU).
 * 4, 8 or 16, within the register/LDS limits for the tallest QR (<= 512 rows for 16, <= 1024
 * for 8, <= 2048 for 4); anything else is HCOREPP_ERR_INVALID.  Drops a captured graph. */
hcorepp_status hcorepp_batch_set_panel(hcorepp_batch *b, int panel);
int hcorepp_batch_get_panel(hcorepp_batch *b);
/* The compiled-in autotune table's choice for a batch whose tallest QR has `rows` rows, whose
 * mean stacked rank is mean_k and which holds updates_per_cu updates per compute unit (no
 * counterpart in the reference).  Returns 1 and the table's panel width and workgroups per CU
 * when the table covers the class, else 0 and the built-in rule's (widest fitting panel, 2
 * workgroups per CU). */
int hcorepp_autotune_lookup(int64_t rows, double mean_k, double updates_per_cu, int *panel,
                            int *workgroups_per_cu);
/* Diagnostics (no counterpart in the reference): per-update phase timestamps of the last pass,
 * 16 int64 per update in ticks of the 100 MHz device wall clock: 8 timestamps (start, after the
 * factor products, after QR(U_s), after Y, after QR(Y) and the core set-up, after the core SVD,
 * after U, end) and 8 accumulated sub-phase durations (Householder panels, T factors, trailing
 * updates, core SVD sweeps, Q application, 3 spare).  Returns
 * HCOREPP_ERR_UNSUPPORTED unless the library was built with -DHCAMD_PHASE_TIMING. */
hcorepp_status hcorepp_batch_phase_times(hcorepp_batch *b, int64_t *times);
/* ---------------------------------------------------------------------------
 * TLR Cholesky tile operations (FP64, lower, A = L L^T), batched over independent tiles where
 * the factorisation allows it.  The trailing updates A_ij -= L_ik L_jk^T are hcorepp_gemm_d /
 * batch updates with alpha = -1 whose B operand is the transposed tile from
 * hcorepp_transpose_batch_d.
 * Replace HCore<T>::Potrf (reference src/api/HCore.cpp:586-622, dense tile, Uplo::Lower): */
hcorepp_status hcorepp_potrf_d(hcorepp_ctx *ctx, hcorepp_tile *A);
/* Replace HCore<T>::Trsm (HCore.cpp:624-647; Side::Right, Uplo::Lower, Op::Trans, NonUnit, alpha 1) on
 * n compressed tiles B[i] = U V: V <- V L^{-T}. */
hcorepp_status hcorepp_trsm_batch_d(hcorepp_ctx *ctx, int64_t n, const hcorepp_tile *L, hcorepp_tile *B);
/* Replace HCore<T>::Syrk (HCore.cpp:484-584; compressed A, alpha -1, beta 1) on n dense tiles:
 * C[i] <- C[i] - A[i] A[i]^T (full tile written). */
hcorepp_status hcorepp_syrk_batch_d(hcorepp_ctx *ctx, int64_t n, const hcorepp_tile *A, hcorepp_tile *C);
/* Compressed dst[i] = src[i]^T (U_dst = V_src^T, V_dst = U_src^T; dst[i].rank is set).  No
 * counterpart in the reference, whose aCholesky Gemm branch (HCore.cpp:160-174) re-plumbs the
 * operands instead. */
hcorepp_status hcorepp_transpose_batch_d(hcorepp_ctx *ctx, int64_t n, const hcorepp_tile *src, hcorepp_tile *dst);

/* ---------------------------------------------------------------------------
 * Cholesky-flow tile operations with the reference's argument meaning and its tile convention
 * for that flow: a compressed tile is U V^T with V stored cols x rank (ld cols) in the V slot.
 * The Gemm of that flow is hcorepp_gemm_d / _s with cholesky = 1 (compressed A, B, C; opA 'N',
 * opB 'T': C <- beta C + alpha A B^T), replacing HCore<T>::Gemm(..., aCholesky = true)
 * (reference src/api/HCore.cpp:160-174, src/operators/concrete/Compressed.cpp:273-276, 458-476,
 * 529-595, 669-673).
 * Replaces HCore<T>::Potrf (reference include/hcorepp/api/HCore.hpp:78-90, src/api/HCore.cpp:586-622):
 * LAPACK potrf of a dense square tile, uplo 'L' (A = L L^T) or 'U' (A = U^T U); the other strict
 * triangle is left untouched.  A non-positive leading minor returns HCOREPP_ERR_NUMERIC (the
 * reference ignores LAPACK's info).  Synchronises the context stream. */
hcorepp_status hcorepp_tile_potrf_d(hcorepp_ctx *ctx, char uplo, hcorepp_tile *A);
hcorepp_status hcorepp_tile_potrf_s(hcorepp_ctx *ctx, char uplo, hcorepp_tile *A);
/* Replaces HCore<T>::Trsm (reference include/hcorepp/api/HCore.hpp:92-106, src/api/HCore.cpp:624-647):
 * BLAS trsm(side 'L'/'R', uplo 'L'/'U', trans 'N'/'T'/'C', diag 'N'/'U', alpha) with the dense
 * square A on the V factor of the compressed B, viewed as B.rows x B.rank with B's leading
 * dimension: V <- alpha op(A)^{-1} V (left) or alpha V op(A)^{-1} (right). */
hcorepp_status hcorepp_tile_trsm_d(hcorepp_ctx *ctx, char side, char uplo, char trans, char diag, double alpha,
                                   const hcorepp_tile *A, hcorepp_tile *B);
hcorepp_status hcorepp_tile_trsm_s(hcorepp_ctx *ctx, char side, char uplo, char trans, char diag, float alpha,
                                   const hcorepp_tile *A, hcorepp_tile *B);
/* Replaces HCore<T>::Syrk (reference include/hcorepp/api/HCore.hpp:66-77, src/api/HCore.cpp:484-584)
 * into a dense C.  Dense A: BLAS syrk on C's uplo ('L'/'U') triangle, C <- alpha op(A) op(A)^T +
 * beta C with n = C.rows and k = C.cols, as the reference passes them.  Compressed A = U V^T:
 * C <- beta C - U (V^T V) U^T (alpha is not used, as in the reference), then the strict upper
 * (uplo 'L') or lower ('U') triangle set to zero; uplo 'G' keeps both. */
hcorepp_status hcorepp_tile_syrk_d(hcorepp_ctx *ctx, double alpha, const hcorepp_tile *A, char opA, char uplo,
                                   double beta, hcorepp_tile *C);
hcorepp_status hcorepp_tile_syrk_s(hcorepp_ctx *ctx, float alpha, const hcorepp_tile *A, char opA, char uplo,
                                   float beta, hcorepp_tile *C);

/* ---------------------------------------------------------------------------
 * Learned rank predictor (no counterpart in the reference).  Per-update features are computed on
 * the device right after stacking (hcorepp_batch_features copies them, RANK_FEATURES floats per
 * update, after a pass with a model set or in the exact form -- the tolerance rule without a
 * model skips them and leaves zeros); an MLP of hcorepp_rank_model_params() parameters predicts log2 of the
 * post-truncation rank.  With a model set (BF16 parameters in device memory, from
 * hcorepp_mlp_pack_bf16), each update's core is presized to the prediction + margin columns and
 * accepted only if the dropped rows of the pivoted R factor have Frobenius norm <= tol * threshold;
 * otherwise the exact core is used (counted as a fallback).  model_bf16 = NULL with tol > 0 applies
 * the same tolerance directly to the pivoted R factor (no prediction); NULL with tol = 0 switches
 * both off (exact eps-level core). */
hcorepp_status hcorepp_batch_set_rank_model(hcorepp_batch *b, const uint16_t *model_bf16, double tol, int margin);
hcorepp_status hcorepp_batch_features(hcorepp_batch *b, float *features);
/* Fallbacks of the last pass (synchronises); predicted (may be NULL, host, n entries) gets each
 * update's predicted rank (0 when no model was set). */
hcorepp_status hcorepp_batch_learned_stats(hcorepp_batch *b, int64_t *fallbacks, int64_t *predicted);
/* Width of the core each update of the last pass handed to its SVD (synchronises; host, n entries):
 * the Jacobi's work grows with its square, so this is what learned or rule sizing saves. */
hcorepp_status hcorepp_batch_core_widths(hcorepp_batch *b, int64_t *widths);
int hcorepp_rank_model_features(void);
size_t hcorepp_rank_model_params(void);
/* FP32 training kernels on device buffers: X (n x features, row-major), y (n), workspace of
 * hcorepp_mlp_workspace_floats(n) floats.  grad computes the mean-squared-error loss (to host,
 * synchronises when loss is non-NULL) and its gradient; adam applies one Adam step (step >= 1);
 * pack_bf16 rounds the parameters to BF16 for inference; predict_bf16 runs the in-kernel BF16
 * inference on rows of X. */
size_t hcorepp_mlp_workspace_floats(int64_t n);
hcorepp_status hcorepp_mlp_forward_f32(hcorepp_ctx *ctx, const float *params, const float *X, int64_t n, float *ws);
hcorepp_status hcorepp_mlp_grad_f32(hcorepp_ctx *ctx, const float *params, const float *X, const float *y, int64_t n,
                                    float *ws, float *grads, float *loss);
hcorepp_status hcorepp_mlp_adam_f32(hcorepp_ctx *ctx, float *params, const float *grads, float *m, float *v, float lr,
                                    float beta1, float beta2, float eps, int64_t step);
hcorepp_status hcorepp_mlp_pack_bf16(hcorepp_ctx *ctx, const float *params, uint16_t *out);
hcorepp_status hcorepp_mlp_predict_bf16(hcorepp_ctx *ctx, const uint16_t *model_bf16, const float *X, int64_t n,
                                        float *out);
/* Bytes of device workspace the batch holds. */
size_t hcorepp_batch_workspace_bytes(hcorepp_batch *b);
/* Frees the batch; it synchronises the batch's context stream, so destroy every batch before
 * the context it was created on. */
hcorepp_status hcorepp_batch_destroy(hcorepp_batch *b);

/* Diagnostics (no counterpart in the reference): copy the device workspace of the last plain
 * recompression on `ctx` to host memory and describe its layout in info[16]. */
int64_t hcorepp_debug_workspace(hcorepp_ctx *ctx, void *host, size_t cap, int64_t *info);

/* Human-readable description of the last error on this thread. */
const char *hcorepp_last_error(void);
/* Library build/version string. */
const char *hcorepp_version(void);

#ifdef __cplusplus
}
#endif
#endif /* HCOREPP_AMD_H */
