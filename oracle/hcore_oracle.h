/*
 * hcore_oracle.h -- CPU restatement of HCore++'s TLR GEMM + recompression path.
 * Test infrastructure only (the checker); never linked into the product library.
 */
#ifndef HCORE_ORACLE_H
#define HCORE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_DENSE = 0, ORC_COMPRESSED = 1 };
enum { ORC_SVD_GESVD = 0, ORC_SVD_GESDD = 1 };
enum { ORC_ERR_UNSUPPORTED = -100, ORC_ERR_CAPACITY = -101 };

/* Mirrors TileMetadata + data pointer (include/hcorepp/operators/interface/Tile.hpp:30-52).
 * Compressed data layout: [U rows x max_rank (ld rows) | V max_rank x cols (ld rank)]. */
typedef struct {
    int64_t rows, cols, rank, max_rank, ld;
    int32_t type;
    int32_t pad;
    double *data;
} orc_tile;

/* Mirrors CompressionParameters (include/hcorepp/operators/helpers/CompressionParameters.hpp:44-46). */
typedef struct {
    double accuracy;
    int32_t trmm, ungqr, truncated_svd;
    int32_t svd_type;
    int64_t fixed_rank;
} orc_params;

int orc_init(const char *lapack_path);
const char *orc_lapack_path(void);
int orc_has_latms(void);
void orc_set_lapack_threads(int n);

int64_t orc_calculate_new_rank(int truncated_svd, const double *sigma, int64_t size_s, double accuracy);
void orc_process_vpointer(int64_t n, int64_t crank, int64_t vm, double beta, const double *cv, int64_t ldcv,
                          double *v, int64_t arank, const double *b);
void orc_calculate_uvptr(int64_t rank, int64_t vm, double *uvptr, const double *vnew);
void orc_calculate_vtnew(int64_t rk, int64_t min_vm_vn, const double *sigma, double *vtnew, int64_t size_s);
void orc_lacpy(char uplo, int64_t m, int64_t n, const double *a, int64_t lda, double *b, int64_t ldb);
void orc_laset(char uplo, int64_t m, int64_t n, double off, double diag, double *a, int64_t lda);
int orc_geqrf(int64_t m, int64_t n, double *a, int64_t lda, double *tau);
int orc_svd(int svd_type, int64_t m, int64_t n, double *a, int64_t lda, double *s, double *u, int64_t ldu,
            double *vt, int64_t ldvt);
void orc_dgemm(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
               const double *B, int64_t ldb, double beta, double *C, int64_t ldc);

int orc_hcore_gemm(double alpha, const orc_tile *A, char opA, const orc_tile *B, char opB, double beta, orc_tile *C,
                   const orc_params *p, double *sigma_out, int64_t *size_s_out);
int orc_compress(int64_t m, int64_t n, const double *a, int64_t lda, const orc_params *p, orc_tile *out,
                 double *sigma_out);
int64_t orc_compress_dense_matrix(int64_t m, int64_t n, const double *a, int64_t lda, double accuracy, double *uv);
int orc_latms_dense(int64_t m, int64_t n, double *a, int64_t lda, int32_t *iseed, double cond);
int orc_dpotrf(char uplo, int64_t n, double *a, int64_t lda);

/* Batch driver for the CPU baseline: runs `n` independent updates with OpenMP over
 * updates (MKL sequential per thread).  Returns wall seconds. */
double orc_batch_gemm(int64_t n, double alpha, const orc_tile *A, const orc_tile *B, double beta, orc_tile *C,
                      const orc_params *p, int threads, int64_t *ranks_out);
double orc_tlr_gemm_omp(int64_t ti, int64_t tj, int64_t tk, double alpha, const orc_tile *A, const orc_tile *B,
                        double beta, orc_tile *C, const orc_params *p, int omp_threads, int lapack_threads);

#ifdef __cplusplus
}
#endif
#endif
