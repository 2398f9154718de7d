"""GPU tests of the Cholesky-flow tile operations with the reference's argument meaning
(HCore<T>::Potrf / Trsm / Syrk and Gemm(..., aCholesky = true), reference src/api/HCore.cpp:160-174,
482-647): against the reference's own known-answer vectors (tests/kernels/TestKernels.cpp:1070-1248),
LAPACK/BLAS semantics restated with numpy/scipy (potrf of either triangle with the other left
untouched, all 16 trsm variants, syrk on either triangle), and the Cholesky-flow Gemm against the
oracle's recompression of the same product (ranks and U V^T) and the dense FP64 product.
Tolerances: 1e-12 relative for FP64 factorisations and solves, 1e-4 for FP32, the oracle's
recompression bar (ranks equal unless a tie, products 1e-10) for the Gemm."""
import json
import os

import numpy as np
import pytest
from scipy.linalg import solve_triangular

from helpers import rank_tie, rel_fro

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))["kernels"]


def _H():
    import hcorepp_amd as H
    return H


def _spd(rng, n):
    g = rng.standard_normal((n, n))
    return g @ g.T + n * np.eye(n)


def test_reference_vectors_through_tile_api(ctx):
    H = _H()
    v = GOLD["potrf_lower"]
    t = H.DenseTile(np.array(v["a"], dtype=float))
    H.HCore.potrf(t, "L", ctx)
    got = t.to_numpy()
    want = np.array(v["expected"], dtype=float)
    assert np.allclose(np.tril(got), np.tril(want), rtol=1e-12, atol=1e-14)
    v = GOLD["trsm_left_upper"]
    A = H.DenseTile(np.array(v["a"], dtype=float))
    b = np.array(v["b"], dtype=float)[:, 0]
    B = H.CompressedTile.from_cholesky_factors(np.ones((len(b), 1)), b[:, None])
    H.HCore.trsm("L", "U", "N", "N", 1.0, A, B, ctx)
    _, x = B.cholesky_factors()
    assert np.allclose(x[:, 0], np.array(v["expected"], dtype=float)[:, 0], rtol=0, atol=1e-13)
    v = GOLD["syrk_lower"]
    Ad = np.array(v["a"], dtype=float)
    Cd = H.DenseTile(np.zeros((4, 4)))
    H.HCore.syrk(1.0, H.DenseTile(Ad), "N", "L", 0.0, Cd, ctx)
    assert np.allclose(np.tril(Cd.to_numpy()), np.tril(np.array(v["expected"], dtype=float)), rtol=1e-11)


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_potrf_either_triangle(ctx, uplo, dtype):
    H = _H()
    rng = np.random.default_rng(3)
    n = 200
    S = _spd(rng, n)
    junk = rng.standard_normal((n, n))
    a = np.where(np.tril(np.ones((n, n))) if uplo == "L" else np.triu(np.ones((n, n))), S, junk)
    t = H.DenseTile(a, dtype=dtype)
    H.HCore.potrf(t, uplo, ctx)
    got = t.to_numpy().astype(np.float64)
    L = np.linalg.cholesky(S)
    tol = 1e-12 if dtype == np.float64 else 1e-4
    if uplo == "L":
        assert rel_fro(np.tril(got), L) < tol
        assert np.array_equal(np.triu(got, 1), np.triu(a.astype(dtype), 1).astype(np.float64))
    else:
        assert rel_fro(np.triu(got), L.T) < tol
        assert np.array_equal(np.tril(got, -1), np.tril(a.astype(dtype), -1).astype(np.float64))
    bad = H.DenseTile(-np.eye(8), dtype=dtype)
    with pytest.raises(H.HCoreError, match="NUMERIC"):
        H.HCore.potrf(bad, uplo, ctx)


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_trsm_variants(ctx, side, uplo, trans, diag):
    H = _H()
    rng = np.random.default_rng(hash((side, uplo, trans, diag)) % 1000)
    rows, r = 150, 37
    order = rows if side == "L" else r
    a = rng.standard_normal((order, order)) / order + 3.0 * np.eye(order)
    a = np.tril(a) if uplo == "L" else np.triu(a)
    v = rng.standard_normal((rows, r))
    u = rng.standard_normal((rows, r))
    B = H.CompressedTile.from_cholesky_factors(u, v, max_rank=r + 5)
    H.HCore.trsm(side, uplo, trans, diag, 1.5, H.DenseTile(a), B, ctx)
    ug, vg = B.cholesky_factors()
    assert np.array_equal(ug, u)
    opa = a.T if trans == "T" else a
    if diag == "U":
        opa = opa - np.diag(np.diag(opa)) + np.eye(order)
    want = np.linalg.solve(opa, 1.5 * v) if side == "L" else np.linalg.solve(opa.T, 1.5 * v.T).T
    assert rel_fro(vg, want) < 1e-12


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("op", ["N", "T"])
def test_syrk_dense(ctx, uplo, op):
    H = _H()
    rng = np.random.default_rng(17)
    n = 130
    a = rng.standard_normal((n, n))
    c = rng.standard_normal((n, n))
    C = H.DenseTile(c)
    H.HCore.syrk(0.7, H.DenseTile(a), op, uplo, -0.4, C, ctx)
    opa = a if op == "N" else a.T
    want = 0.7 * opa @ opa.T - 0.4 * c
    got = C.to_numpy()
    tri = np.tril if uplo == "L" else np.triu
    other = (lambda x: np.triu(x, 1)) if uplo == "L" else (lambda x: np.tril(x, -1))
    assert rel_fro(tri(got), tri(want)) < 1e-12
    assert np.array_equal(other(got), other(c))


def test_syrk_compressed(ctx):
    """Compressed A = U V^T (V n x r): C <- beta C - A A^T, strict upper zeroed for uplo L."""
    H = _H()
    rng = np.random.default_rng(19)
    n, r = 160, 23
    u, v = rng.standard_normal((n, r)), rng.standard_normal((n, r))
    c = _spd(rng, n)
    C = H.DenseTile(c)
    H.HCore.syrk(5.0, H.CompressedTile.from_cholesky_factors(u, v), "N", "L", 1.0, C, ctx)
    a = u @ v.T
    want = np.tril(c - a @ a.T)
    assert rel_fro(C.to_numpy(), want) < 1e-12


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cholesky_flow_gemm_matches_oracle(ctx, dtype):
    """Gemm(aCholesky = true): C <- beta C + alpha A B^T on Cholesky-flow tiles; the oracle runs
    the same product as an ordinary recompression (A = (U_a, V_a^T), B^T = (V_b, U_b^T))."""
    H = _H()
    from oracle import oracle as O
    from hcorepp_amd.workloads import lowrank_factors
    rng = np.random.default_rng(23)
    m, ra, rb, rc, mr, acc = 256, 30, 25, 40, 120, 1e-8
    ua, va_t = lowrank_factors(rng, m, m, ra, 0.9)
    ub, vb_t = lowrank_factors(rng, m, m, rb, 0.9)
    uc, vc_t = lowrank_factors(rng, m, m, rc, 0.9)
    A = H.CompressedTile.from_cholesky_factors(ua, va_t.T, dtype=dtype)
    B = H.CompressedTile.from_cholesky_factors(ub, vb_t.T, dtype=dtype)
    Ct = H.CompressedTile.from_cholesky_factors(uc, vc_t.T, max_rank=mr, dtype=dtype)
    H.HCore.gemm(-1.0, A, B, 1.0, Ct, ctx, H.CompressionParameters(acc), opA="N", opB="T", cholesky=True)
    oC = O.Tile.compressed(uc, vc_t, max_rank=mr)
    _, so = O.hcore_gemm(-1.0, O.Tile.compressed(ua, va_t), O.Tile.compressed(vb_t.T.copy(), ub.T.copy()), 1.0, oC,
                         O.Params(acc), want_sigma=True)
    ug, vg = Ct.cholesky_factors()
    got = ug.astype(np.float64) @ vg.astype(np.float64).T
    dense = uc @ vc_t - (ua @ va_t) @ (ub @ vb_t).T
    if dtype == np.float64:
        assert rank_tie(so, oC.rank, Ct.rank, acc), (oC.rank, Ct.rank)
        if Ct.rank == oC.rank:
            assert rel_fro(got, oC.to_dense()) < 1e-10
        tail = np.sqrt(np.sum(so[Ct.rank:] ** 2))
        assert np.linalg.norm(got - dense) <= 1.01 * tail + 1e-12 * np.linalg.norm(dense) * m
    else:
        assert rel_fro(got, dense) < 1e-4
    # the other operand layouts are not part of that flow
    with pytest.raises(H.HCoreError, match="UNSUPPORTED"):
        H.HCore.gemm(1.0, A, B, 1.0, Ct, ctx, H.CompressionParameters(acc), opA="N", opB="N", cholesky=True)
