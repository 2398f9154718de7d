// Compiles a caller written against the reference's C++ API (HCore<T>::Gemm with
// DenseTile / CompressedTile / CompressionParameters / RunContext / MemoryUnit, as in
// the reference's tests/api/TestHCore.cpp) against the drop-in headers.  Built and
// linked on the CPU; it runs GPU work only when invoked with the argument "run".
#include <cstdio>
#include <cstring>
#include <vector>

#include <hcorepp/api/HCore.hpp>
#include <hcorepp/operators/interface/TilePacker.hpp>

using namespace hcorepp::api;
using namespace hcorepp::operators;
using namespace hcorepp::kernels;
using namespace hcorepp::dataunits;

// "chol <in> <out>": the right-looking TLR Cholesky of a T x T tile matrix through the reference's
// Cholesky-flow API (HCore<double>::Potrf / Trsm / Syrk and Gemm with aCholesky = true; compressed
// tiles U V^T with V stored nb x rank).  File format (int64 / double, column-major): T, nb, acc;
// the T dense diagonal tiles; then for i > k in row order: rank, max_rank, U (nb x rank), V (nb x rank).
// The output file holds the factor in the same format (final ranks).
static int run_cholesky(const char *in_path, const char *out_path) {
    FILE *f = std::fopen(in_path, "rb");
    if (!f) return 2;
    int64_t T = 0, nb = 0;
    double acc = 0;
    bool ok = std::fread(&T, 8, 1, f) == 1 && std::fread(&nb, 8, 1, f) == 1 && std::fread(&acc, 8, 1, f) == 1;
    RunContext ctx;
    std::vector<DenseTile<double> *> D((size_t) T, nullptr);
    std::vector<CompressedTile<double> *> L((size_t) (T * T), nullptr);
    std::vector<double> buf((size_t) (2 * nb * nb));
    for (int64_t k = 0; ok && k < T; ++k) {
        ok = std::fread(buf.data(), 8, (size_t) (nb * nb), f) == (size_t) (nb * nb);
        D[(size_t) k] = new DenseTile<double>((size_t) nb, (size_t) nb, buf.data(), (size_t) nb, ctx);
    }
    for (int64_t i = 0; ok && i < T; ++i)
        for (int64_t k = 0; ok && k < i; ++k) {
            int64_t r = 0, mr = 0;
            ok = std::fread(&r, 8, 1, f) == 1 && std::fread(&mr, 8, 1, f) == 1 &&
                 std::fread(buf.data(), 8, (size_t) (2 * nb * r), f) == (size_t) (2 * nb * r);
            L[(size_t) (i * T + k)] = new CompressedTile<double>((size_t) nb, (size_t) nb, buf.data(), (size_t) nb,
                                                                 (size_t) r, (size_t) mr, ctx);
        }
    std::fclose(f);
    if (!ok) return 3;
    size_t flops = 0;
    MemoryUnit<double> &unit = MemoryHandler<double>::GetInstance().GetMemoryUnit();
    const CompressionParameters params(acc);
    for (int64_t k = 0; k < T; ++k) {
        HCore<double>::Potrf(*D[(size_t) k], blas::Uplo::Lower, ctx, flops, unit);
        for (int64_t i = k + 1; i < T; ++i)
            HCore<double>::Trsm(blas::Side::Left, blas::Uplo::Lower, blas::Op::NoTrans, blas::Diag::NonUnit, 1.0,
                                *D[(size_t) k], *L[(size_t) (i * T + k)], ctx, flops, unit);
        for (int64_t i = k + 1; i < T; ++i)
            HCore<double>::Syrk(-1.0, *L[(size_t) (i * T + k)], blas::Op::NoTrans, blas::Uplo::Lower, 1.0,
                                *D[(size_t) i], ctx, flops, unit);
        for (int64_t j = k + 1; j < T; ++j)
            for (int64_t i = j + 1; i < T; ++i)
                HCore<double>::Gemm(-1.0, *L[(size_t) (i * T + k)], blas::Op::NoTrans, *L[(size_t) (j * T + k)],
                                    blas::Op::Trans, 1.0, *L[(size_t) (i * T + j)], ctx, flops, unit, params, true);
    }
    ctx.Sync();
    FILE *o = std::fopen(out_path, "wb");
    if (!o) return 4;
    std::fwrite(&T, 8, 1, o);
    std::fwrite(&nb, 8, 1, o);
    std::fwrite(&acc, 8, 1, o);
    for (int64_t k = 0; k < T; ++k) {
        ThrowIfError(hcorepp_memcpy(ctx.Handle(), buf.data(), D[(size_t) k]->GetData(), nb * nb * sizeof(double)));
        std::fwrite(buf.data(), 8, (size_t) (nb * nb), o);
    }
    for (int64_t i = 0; i < T; ++i)
        for (int64_t k = 0; k < i; ++k) {
            CompressedTile<double> *t = L[(size_t) (i * T + k)];
            const int64_t r = (int64_t) t->GetTileRank(), mr = (int64_t) t->GetMaxRank();
            ThrowIfError(hcorepp_memcpy(ctx.Handle(), buf.data(), t->GetUMatrix(), nb * r * sizeof(double)));
            ThrowIfError(hcorepp_memcpy(ctx.Handle(), buf.data() + nb * r, t->GetVMatrix(), nb * r * sizeof(double)));
            std::fwrite(&r, 8, 1, o);
            std::fwrite(&mr, 8, 1, o);
            std::fwrite(buf.data(), 8, (size_t) (2 * nb * r), o);
        }
    std::fclose(o);
    for (auto *d : D) delete d;
    for (auto *t : L) delete t;
    std::printf("dropin_compile: cholesky %lld x %lld tiles nb=%lld, flops %zu\n", (long long) T, (long long) T,
                (long long) nb, flops);
    return 0;
}

int main(int argc, char **argv) {
    if (argc == 4 && std::strcmp(argv[1], "chol") == 0) return run_cholesky(argv[2], argv[3]);
    if (argc < 2 || std::strcmp(argv[1], "run") != 0) {
        std::printf("dropin_compile: built (%s)\n", hcorepp_version());
        return 0;
    }
    // Compressed A (2x1 rank 1), B (1x2 as U 1x1, V 1x2), C dense 2x2: C = A*B + C.
    const size_t m = 4, n = 4, r = 1;
    std::vector<double> a(m * r + r * n), b(m * r + r * n), c(m * n, 0.0);
    for (size_t i = 0; i < a.size(); ++i) a[i] = 1.0 + i;
    for (size_t i = 0; i < b.size(); ++i) b[i] = 0.5 * (i + 1);
    RunContext ctx;
    CompressedTile<double> A(m, n, a.data(), m, r, ctx);
    CompressedTile<double> B(m, n, b.data(), m, r, ctx);
    DenseTile<double> C(m, n, c.data(), m, ctx);
    size_t flops = 0;
    MemoryUnit<double> &unit = MemoryHandler<double>::GetInstance().GetMemoryUnit();
    HCore<double>::Gemm(1.0, A, blas::Op::NoTrans, B, blas::Op::NoTrans, 0.0, C, ctx, flops, unit,
                        CompressionParameters(1e-9));
    ctx.Sync();
    std::vector<double> out(m * n);
    ThrowIfError(hcorepp_memcpy(ctx.Handle(), out.data(), C.GetTileSubMatrix(0), out.size() * sizeof(double)));
    // expected: (Au Av)(Bu Bv)
    double err = 0;
    for (size_t j = 0; j < n; ++j)
        for (size_t i = 0; i < m; ++i) {
            double s = 0;
            for (size_t k = 0; k < m; ++k) {
                const double av = a[i] * a[m + k], bv = b[k] * b[m + j];
                s += av * bv;
            }
            err = std::max(err, std::abs(out[i + m * j] - s));
        }
    // TilePacker round trip (TestTilePacker.cpp): metadata as in the reference, zero-copy repack
    auto md = TilePacker<double>::UnPackTile(A, ctx);
    bool ok = md.first->mNumOfRows == m && md.first->mNumOfCols == n && md.first->mMatrixRank == r &&
              md.first->mType == COMPRESSED && md.second == A.GetUMatrix();
    auto mdd = TilePacker<double>::UnPackTile(C, ctx);
    ok = ok && mdd.first->mMatrixRank == 0 && mdd.first->mType == DENSE && mdd.first->mLeadingDimension == m;
    Tile<double> *packed = TilePacker<double>::PackTile(*md.first, md.second, ctx);
    ok = ok && packed->isCompressed() && packed->GetData() == A.GetUMatrix();
    delete packed;
    delete md.first;
    delete mdd.first;
    // Compression constructor (Compressed.cpp:75-146) on a rank-2 dense 12x9 tile from host memory
    const size_t cm = 12, cn = 9;
    std::vector<double> dense(cm * cn);
    for (size_t j = 0; j < cn; ++j)
        for (size_t i = 0; i < cm; ++i) dense[i + cm * j] = (1.0 + i) * (2.0 - 0.1 * j) + 0.5 * i * i * (j % 3);
    CompressedTile<double> Z(cm, cn, dense.data(), cm, CompressionParameters(1e-10), blas::Layout::ColMajor, ctx);
    std::vector<double> zu(cm * Z.GetTileRank()), zv(Z.GetTileRank() * cn);
    ThrowIfError(hcorepp_memcpy(ctx.Handle(), zu.data(), Z.GetUMatrix(), zu.size() * sizeof(double)));
    ThrowIfError(hcorepp_memcpy(ctx.Handle(), zv.data(), Z.GetVMatrix(), zv.size() * sizeof(double)));
    double cerr = 0;
    for (size_t j = 0; j < cn; ++j)
        for (size_t i = 0; i < cm; ++i) {
            double s = 0;
            for (size_t k = 0; k < Z.GetTileRank(); ++k) s += zu[i + cm * k] * zv[k + Z.GetTileRank() * j];
            cerr = std::max(cerr, std::abs(s - dense[i + cm * j]));
        }
    const bool cok = Z.GetTileRank() == 2 && Z.GetMaxRank() == 3 && cerr < 1e-9;
    std::printf("dropin_compile: max abs err %.3e flops %zu packer %s compress rank %zu err %.3e\n", err, flops,
                ok ? "ok" : "FAIL", Z.GetTileRank(), cerr);
    return (err < 1e-10 && ok && cok) ? 0 : 1;
}
