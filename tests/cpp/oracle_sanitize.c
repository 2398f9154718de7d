/* Drives the CPU oracle (oracle/hcore_oracle.c) through every dense/compressed combination of
 * HCore::Gemm, the compression constructor and the tile-GEMM driver, so that a build with
 * -fsanitize=address,undefined checks its memory accesses and arithmetic.  Exit 0 when every call
 * returned 0 and the results are finite. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/hcore_oracle.h"

static double urand(unsigned long long *s) {
    *s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double) (*s >> 11) / 9007199254740992.0 - 0.5;
}

static orc_tile make_tile(int compressed, int64_t m, int64_t n, int64_t r, int64_t mr, unsigned long long *seed) {
    orc_tile t;
    memset(&t, 0, sizeof(t));
    t.rows = m;
    t.cols = n;
    t.type = compressed ? ORC_COMPRESSED : ORC_DENSE;
    if (compressed) {
        t.rank = r;
        t.max_rank = mr;
        t.ld = m;
        t.data = calloc((size_t) ((m + n) * mr), sizeof(double));
        for (int64_t e = 0; e < m * r; ++e) t.data[e] = urand(seed);
        for (int64_t e = 0; e < r * n; ++e) t.data[m * mr + e] = urand(seed) * pow(0.8, (double) (e % r));
    } else {
        t.rank = t.max_rank = m < n ? m : n;
        t.ld = m;
        t.data = calloc((size_t) (m * n), sizeof(double));
        for (int64_t e = 0; e < m * n; ++e) t.data[e] = urand(seed);
    }
    return t;
}

static int all_finite(const orc_tile *t) {
    const int64_t cnt = t->type == ORC_DENSE ? t->ld * t->cols : (t->rows + t->cols) * t->max_rank;
    for (int64_t e = 0; e < cnt; ++e)
        if (!isfinite(t->data[e])) return 0;
    return 1;
}

int main(void) {
    if (orc_init(NULL) != 0) {
        printf("oracle_sanitize: no LAPACK runtime\n");
        return 2;
    }
    unsigned long long seed = 12345;
    orc_params p = {1e-8, 0, 1, 0, ORC_SVD_GESDD, 0};
    const char *kinds[] = {"DDD", "DDC", "DCD", "CDD", "CCD", "CDC", "DCC", "CCC"};
    int bad = 0;
    for (int k = 0; k < 8; ++k) {
        const int64_t m = 48;
        orc_tile A = make_tile(kinds[k][0] == 'C', m, m, 6, 6, &seed);
        orc_tile B = make_tile(kinds[k][1] == 'C', m, m, 5, 5, &seed);
        const int cc = kinds[k][2] == 'C';
        orc_tile C = make_tile(cc, m, m, 7, kinds[k][0] == 'D' && kinds[k][1] == 'D' ? m : 30, &seed);
        double sigma[256];
        int64_t ns = 0;
        const int info = orc_hcore_gemm(1.5, &A, 'N', &B, 'N', 0.5, &C, &p, sigma, &ns);
        if (info != 0 || !all_finite(&C)) {
            printf("oracle_sanitize: %s info %d\n", kinds[k], info);
            bad = 1;
        }
        free(A.data);
        free(B.data);
        free(C.data);
    }
    /* compression constructor and the tile-GEMM driver */
    orc_tile D = make_tile(0, 40, 40, 0, 0, &seed);
    orc_tile Z;
    memset(&Z, 0, sizeof(Z));
    Z.rows = Z.cols = 40;
    Z.max_rank = 13;
    Z.ld = 40;
    Z.type = ORC_COMPRESSED;
    Z.data = calloc(80 * 13, sizeof(double));
    double s2[40];
    if (orc_compress(40, 40, D.data, 40, &p, &Z, s2) != 0 || !all_finite(&Z)) bad = 1;
    orc_tile tiles[3][4];
    for (int g = 0; g < 3; ++g)
        for (int i = 0; i < 4; ++i) tiles[g][i] = make_tile(1, 32, 32, 4, g == 2 ? 16 : 4, &seed);
    const double sec = orc_tlr_gemm_omp(2, 2, 2, 1.0, tiles[0], tiles[1], 1.0, tiles[2], &p, 2, 1);
    for (int i = 0; i < 4; ++i)
        if (!all_finite(&tiles[2][i])) bad = 1;
    for (int g = 0; g < 3; ++g)
        for (int i = 0; i < 4; ++i) free(tiles[g][i].data);
    free(D.data);
    free(Z.data);
    printf("oracle_sanitize: %s (%.3f s tile GEMM)\n", bad ? "FAIL" : "ok", sec);
    return bad;
}
