"""GPU parity tests: the HIP path (through the C-ABI) against the reference's own
known-answer vectors and against the CPU oracle on seeded inputs.

Tolerances
  * known-answer products: 1 % relative per element, the reference's own bar
    (tests/helpers/src/testHelpers.cpp:83-86); exact-rank cases 1e-12 norm-wise.
  * oracle parity (FP64): ranks equal unless a singular value of the oracle lies
    within 1e3*eps*sigma_0 of the threshold (a rounding tie); reassembled U*V equal to
    1e-11 * ||C_oracle||_F (plus the size of sigma_{r} when a tie changed the rank).
"""
import json
import os

import numpy as np
import pytest

from helpers import EPS, rank_tie, rel_fro

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))


def _H():
    import hcorepp_amd as H
    return H


def _orc():
    from oracle import oracle as O
    return O


def _mk_tiles(case, dev_dtype=np.float64):
    """Build (A, B, C) device tiles and oracle tiles for a TestGemm case."""
    H, O = _H(), _orc()
    k = case["kind"]
    arr = {key: np.array(case[key], dtype=np.float64) for key in ("A", "B", "Au", "Av", "Bu", "Bv") if key in case}
    if k[0] == "D":
        A, oA = H.DenseTile(arr["A"], dtype=dev_dtype), O.Tile.dense(arr["A"])
    else:
        A, oA = H.CompressedTile(arr["Au"], arr["Av"], dtype=dev_dtype), O.Tile.compressed(arr["Au"], arr["Av"])
    if k[1] == "D":
        B, oB = H.DenseTile(arr["B"], dtype=dev_dtype), O.Tile.dense(arr["B"])
    else:
        B, oB = H.CompressedTile(arr["Bu"], arr["Bv"], dtype=dev_dtype), O.Tile.compressed(arr["Bu"], arr["Bv"])
    out = np.array(case["out"], dtype=np.float64)
    m, n = out.shape
    if k[2] == "D":
        C, oC = H.DenseTile(np.zeros((m, n)), dtype=dev_dtype), O.Tile.dense(np.zeros((m, n)))
    else:
        cr = case["c_rank"]
        C = H.CompressedTile(np.zeros((m, cr)), np.zeros((cr, n)), dtype=dev_dtype)
        oC = O.Tile.compressed(np.zeros((m, cr)), np.zeros((cr, n)))
    return A, B, C, oA, oB, oC, out


@pytest.mark.parametrize("case", GOLD["gemm"], ids=[c["name"] for c in GOLD["gemm"]])
def test_reference_known_answers(ctx, case):
    H, O = _H(), _orc()
    print(f"[case] {case['name']} ({case['src']})", flush=True)
    A, B, C, oA, oB, oC, out = _mk_tiles(case)
    acc = EPS if case.get("acc") == "eps" else 1e-9
    H.HCore.gemm(1.0, A, B, 1.0, C, ctx, H.CompressionParameters(acc))
    got = C.to_dense()
    # the reference's bar: 1 % relative per element
    np.testing.assert_allclose(got, out, rtol=1e-2)
    # and the oracle restatement gives the same
    O.hcore_gemm(1.0, oA, oB, 1.0, oC, O.Params(acc))
    assert C.rank == oC.rank
    assert rel_fro(got, oC.to_dense()) < 1e-10


def test_compressed_tile_direct_gemm(ctx):
    """CompressedTile::Gemm called directly (TestCompressedTile.cpp:124-260), expressed as a CDC
    update with A_U = A, A_V = I so that the final factors are exactly (A, B)."""
    H = _H()
    g = GOLD["compressed_direct"]
    A = np.array(g["A"], float)
    Bm = np.array(g["B"], float)
    At = H.CompressedTile(A, np.eye(3))
    Bt = H.DenseTile(Bm)
    C = H.CompressedTile(np.zeros((3, 2)), np.zeros((2, 2)))
    H.HCore.gemm(1.0, At, Bt, 1.0, C, ctx, H.CompressionParameters(EPS))
    np.testing.assert_allclose(C.to_dense(), np.array(g["out"], float), rtol=1e-2)
    assert C.rank == 2


def _random_update(rng, m, n, ra, rb, rc, kind="CCC", decay=0.7, scale_c=1.0):
    from hcorepp_amd.workloads import lowrank_factors
    au, av = lowrank_factors(rng, m, n, ra, decay)
    bu, bv = lowrank_factors(rng, n, n, rb, decay)
    cu, cv = lowrank_factors(rng, m, n, rc, decay, scale=scale_c)
    return au, av, bu, bv, cu, cv


def _run_pair(ctx, kind, au, av, bu, bv, cu, cv, alpha, beta, params_kw, max_rank=None, dtype=np.float64):
    H, O = _H(), _orc()
    m, n = cu.shape[0], cv.shape[1]
    if kind[0] == "C":
        A, oA = H.CompressedTile(au, av, dtype=dtype), O.Tile.compressed(au, av)
    else:
        Ad = au @ av
        A, oA = H.DenseTile(Ad, dtype=dtype), O.Tile.dense(Ad)
    if kind[1] == "C":
        B, oB = H.CompressedTile(bu, bv, dtype=dtype), O.Tile.compressed(bu, bv)
    else:
        Bd = bu @ bv
        B, oB = H.DenseTile(Bd, dtype=dtype), O.Tile.dense(Bd)
    mr = max_rank if max_rank is not None else cu.shape[1]
    C = H.CompressedTile(cu, cv, max_rank=mr, dtype=dtype)
    oC = O.Tile.compressed(cu, cv, max_rank=mr)
    dense_ref = beta * (cu @ cv) + alpha * (oA.to_dense() @ oB.to_dense())
    _, sig_gpu = H.HCore.gemm(alpha, A, B, beta, C, ctx, H.CompressionParameters(**params_kw), want_sigma=True)
    info, sig_orc = O.hcore_gemm(alpha, oA, oB, beta, oC, O.Params(**params_kw), want_sigma=True)
    assert info == 0
    return C, oC, sig_gpu, sig_orc, dense_ref


CASES = [
    # kind, m, ra, rb, rc, max_rank, acc
    ("CCC", 64, 5, 7, 9, 40, 1e-8),
    ("CCC", 100, 12, 9, 20, 60, 1e-6),
    ("CDC", 96, 10, 0, 14, 50, 1e-8),
    ("DCC", 96, 0, 11, 13, 50, 1e-8),
    ("CCC", 200, 30, 25, 40, 100, 1e-8),
    ("CCC", 256, 40, 35, 60, 120, 1e-10),
    ("CCC", 512, 64, 48, 80, 170, 1e-8),
    # cores of >= 96 columns take the multi-workgroup Jacobi (rows <= 256, <= 512, <= 768 per tile)
    ("CCC", 512, 170, 150, 170, 170, 1e-8),
    ("CCC", 1024, 341, 300, 341, 341, 1e-10),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-m{c[1]}-ra{c[2]}-rc{c[4]}" for c in CASES])
def test_oracle_parity_fp64(ctx, case):
    kind, m, ra, rb, rc, mr, acc = case
    print(f"[case] {case}", flush=True)
    rng = np.random.default_rng(1000 + m + ra)
    ra_, rb_ = max(ra, 1), max(rb, 1)
    au, av, bu, bv, cu, cv = _random_update(rng, m, m, ra_, rb_, rc)
    C, oC, sg, so, dense_ref = _run_pair(ctx, kind, au, av, bu, bv, cu, cv, 1.0, 1.0, dict(accuracy=acc),
                                         max_rank=mr)
    assert len(sg) == len(so)
    # singular values agree to rounding relative to sigma_0
    assert np.max(np.abs(sg - so)) <= 1e3 * EPS * so[0]
    assert rank_tie(so, oC.rank, C.rank, acc), (oC.rank, C.rank)
    got, ref = C.to_dense(), oC.to_dense()
    if C.rank == oC.rank:
        assert rel_fro(got, ref) < 1e-10
    # Eckart-Young: truncation error of the update
    tail = np.sqrt(np.sum(so[C.rank:] ** 2)) if C.rank < len(so) else 0.0
    assert np.linalg.norm(got - dense_ref) <= tail * 1.01 + 1e-12 * np.linalg.norm(dense_ref) * m
    # U has orthonormal columns to rounding (the plain form's Jacobi stops at cosines 0.1 sqrt(eps))
    u, _ = C.factors()
    assert np.linalg.norm(u.T @ u - np.eye(C.rank)) < 1e-13 * np.sqrt(m)


@pytest.mark.parametrize("m,r", [(512, 170), (1024, 300)])
def test_wide_core_multi_workgroup_jacobi(ctx, m, r):
    """Slowly decaying spectra keep the whole core (kt = K = 2r columns): the plain path's
    multi-workgroup Jacobi must give the oracle's singular values and rank."""
    print(f"[case] wide core m={m} r={r}", flush=True)
    rng = np.random.default_rng(77 + m)
    au, av, bu, bv, cu, cv = _random_update(rng, m, m, r, r, r, decay=0.985)
    C, oC, sg, so, dense_ref = _run_pair(ctx, "CCC", au, av, bu, bv, cu, cv, 1.0, 1.0, dict(accuracy=1e-8),
                                         max_rank=2 * r)
    assert len(sg) == len(so)
    assert np.max(np.abs(sg - so)) <= 1e3 * EPS * so[0]
    assert rank_tie(so, oC.rank, C.rank, 1e-8), (oC.rank, C.rank)
    if C.rank == oC.rank:
        assert rel_fro(C.to_dense(), oC.to_dense()) < 1e-10
    u, _ = C.factors()
    assert np.linalg.norm(u.T @ u - np.eye(C.rank)) < 1e-13 * np.sqrt(m)


@pytest.mark.parametrize("m,gap", [(256, 0.0), (512, 1e-12)])
def test_clustered_spectrum(ctx, m, gap):
    """Clusters of equal (gap 0) or nearly equal singular values: the singular vectors inside a
    cluster are not unique, so the product, the singular values and the rank are compared with the
    oracle, and U must still be orthonormal to rounding (equal-norm, nearly orthogonal column
    pairs are the Jacobi's hardest convergence case)."""
    print(f"[case] clustered spectrum m={m} gap={gap}", flush=True)
    rng = np.random.default_rng(91 + m)
    sig = np.concatenate([1.0 + gap * np.arange(24), 0.25 + gap * np.arange(24), 1e-3 * np.ones(16)])
    r = len(sig)
    q1 = np.linalg.qr(rng.standard_normal((m, r)))[0]
    q2 = np.linalg.qr(rng.standard_normal((m, r)))[0]
    cu, cv = np.asfortranarray(q1), np.asfortranarray(sig[:, None] * q2.T)
    # a small update (singular values <= 1e-11) that leaves the clusters in place
    au, av, bu, bv, _, _ = _random_update(rng, m, m, 6, 6, 1)
    av = av * 1e-11
    C, oC, sg, so, dense_ref = _run_pair(ctx, "CCC", au, av, bu, bv, cu, cv, 1.0, 1.0, dict(accuracy=1e-8),
                                         max_rank=r + 12)
    assert np.max(np.abs(sg - so)) <= 1e3 * EPS * so[0]
    assert rank_tie(so, oC.rank, C.rank, 1e-8), (oC.rank, C.rank)
    assert rel_fro(C.to_dense(), dense_ref) < 1e-10
    if C.rank == oC.rank:
        assert rel_fro(C.to_dense(), oC.to_dense()) < 1e-10
    u, _ = C.factors()
    assert np.linalg.norm(u.T @ u - np.eye(C.rank)) < 1e-13 * np.sqrt(m)


def test_alpha_beta_placement(ctx):
    """alpha scales only the A block of U_s, beta only the C block of V_s (Compressed.cpp:344-380)."""
    rng = np.random.default_rng(7)
    au, av, bu, bv, cu, cv = _random_update(rng, 80, 80, 6, 5, 8)
    for alpha, beta in [(3.5, 2.5), (0.0, 1.0), (1.0, 0.0), (-2.0, 0.5)]:
        C, oC, sg, so, dense_ref = _run_pair(ctx, "CCC", au, av, bu, bv, cu, cv, alpha, beta,
                                             dict(accuracy=1e-12), max_rank=30)
        assert C.rank == oC.rank
        assert rel_fro(C.to_dense(), dense_ref) < 1e-10


def test_rank_rules(ctx):
    """truncated-SVD (relative) rule, fixed rank, rank >= 1 and max_rank clamps."""
    rng = np.random.default_rng(11)
    au, av, bu, bv, cu, cv = _random_update(rng, 72, 72, 10, 9, 12, decay=0.6)
    for kw, mr in [(dict(accuracy=1e-6, truncated_svd=True), 40), (dict(accuracy=1e-8, fixed_rank=7), 40),
                   (dict(accuracy=1e3), 40), (dict(accuracy=1e-14), 12)]:
        C, oC, sg, so, _ = _run_pair(ctx, "CCC", au, av, bu, bv, cu, cv, 1.0, 1.0, kw, max_rank=mr)
        if mr == 12:
            assert C.rank == 12  # clamped to max_rank
        assert rank_tie(so, oC.rank, C.rank, kw["accuracy"], kw.get("truncated_svd", False)), (kw, C.rank, oC.rank)
        if C.rank == oC.rank:
            assert rel_fro(C.to_dense(), oC.to_dense()) < 1e-9


def test_dense_c_and_transposes(ctx):
    H, O = _H(), _orc()
    rng = np.random.default_rng(3)
    for opA, opB in [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")]:
        a = rng.standard_normal((50, 50))
        b = rng.standard_normal((50, 50))
        c = rng.standard_normal((50, 50))
        A, B, Cd = H.DenseTile(a), H.DenseTile(b), H.DenseTile(c)
        H.HCore.gemm(1.5, A, B, 0.5, Cd, ctx, opA=opA, opB=opB)
        oa = a.T if opA == "T" else a
        ob = b.T if opB == "T" else b
        np.testing.assert_allclose(Cd.to_numpy(), 1.5 * oa @ ob + 0.5 * c, rtol=1e-12, atol=1e-12)


def test_float32_path(ctx):
    """FP32 instantiation (HCOREPP_INSTANTIATE_CLASS) against the FP64 oracle at FP32 tolerance."""
    rng = np.random.default_rng(5)
    au, av, bu, bv, cu, cv = _random_update(rng, 64, 64, 6, 6, 8)
    C, oC, sg, so, dense_ref = _run_pair(ctx, "CCC", au, av, bu, bv, cu, cv, 1.0, 1.0, dict(accuracy=1e-4),
                                         max_rank=30, dtype=np.float32)
    assert abs(C.rank - oC.rank) <= 1
    assert rel_fro(C.to_dense(), oC.to_dense()) < 1e-3


def test_batched_matches_plain(ctx):
    """Batched persistent form == plain form (same kernels' math, different orchestration)."""
    H = _H()
    from hcorepp_amd.workloads import batch_ccc
    ups = batch_ccc(12, 128, seed=99, lo=4, hi=40)
    acc = 1e-8
    Cs_b, Cs_p, As, Bs = [], [], [], []
    for u in ups:
        As.append(H.CompressedTile(u["au"], u["av"]))
        Bs.append(H.CompressedTile(u["bu"], u["bv"]))
        Cs_b.append(H.CompressedTile(u["cu"], u["cv"], max_rank=90))
        Cs_p.append(H.CompressedTile(u["cu"], u["cv"], max_rank=90))
    b = H.Batch(ctx, As, Bs, Cs_b, 1.0, 1.0, H.CompressionParameters(acc))
    b.run()
    rk = b.sync_ranks_to_tiles()
    st, sweeps = b.status()
    assert st == 0, st
    for i in range(len(ups)):
        H.HCore.gemm(1.0, As[i], Bs[i], 1.0, Cs_p[i], ctx, H.CompressionParameters(acc))
        assert rk[i] == Cs_p[i].rank
        assert rel_fro(Cs_b[i].to_dense(), Cs_p[i].to_dense()) < 1e-12


def test_batched_wide_cores_match_oracle(ctx):
    """Cores too large for the workgroup's LDS (more than 9968 elements: kt >= 100 here), whose
    Jacobi is staged through LDS by super-block pairs, in the exact and the rule form against the
    CPU oracle: ranks equal (up to rounding ties at the threshold), exact form within 1e-10 and
    rule form within the accuracy (1e-8) of the oracle's result."""
    H, O = _H(), _orc()
    from hcorepp_amd.workloads import batch_ccc
    ups = batch_ccc(12, 512, seed=61, lo=120, hi=192)
    acc = 1e-8
    want = []
    for u in ups:
        oC = O.Tile.compressed(u["cu"], u["cv"], max_rank=384)
        _, sig = O.hcore_gemm(1.0, O.Tile.compressed(u["au"], u["av"]), O.Tile.compressed(u["bu"], u["bv"]), 1.0,
                              oC, O.Params(acc), want_sigma=True)
        want.append((oC, sig))
    for tol, bound in ((0.0, 1e-10), (0.1, acc)):
        As = [H.CompressedTile(u["au"], u["av"]) for u in ups]
        Bs = [H.CompressedTile(u["bu"], u["bv"]) for u in ups]
        Cs = [H.CompressedTile(u["cu"], u["cv"], max_rank=384) for u in ups]
        b = H.Batch(ctx, As, Bs, Cs, 1.0, 1.0, H.CompressionParameters(acc))
        b.set_rank_model(None, tol=tol, margin=0)
        b.run()
        rk = b.sync_ranks_to_tiles()
        assert b.status()[0] == 0
        kw = b.core_widths()
        print(f"tol {tol}: core widths {int(kw.min())}-{int(kw.max())}, ranks {int(rk.min())}-{int(rk.max())}")
        assert int(kw.max()) >= 100  # the staged path ran
        for i, (oC, sig) in enumerate(want):
            if rk[i] != oC.rank:
                assert rank_tie(sig, oC.rank, rk[i], acc), (tol, i, rk[i], oC.rank)
                continue
            assert rel_fro(Cs[i].to_dense(), oC.to_dense()) < bound, (tol, i)
        b.close()


@pytest.mark.parametrize("ra,rc", [(120, 120), (180, 180), (240, 240)])
def test_batched_tall_cores_match_oracle(ctx, ra, rc):
    """Slowly decaying spectra keep the whole stacked rank K = ra + rc in the core (no second QR:
    the core is K x K), so the batched Jacobi runs on 240, 360 and 480 rows: 4, 6 and 8 rows per
    lane, staged through LDS by 16- and 8-column super-blocks (240 and 480 rows) or streamed from
    memory (360 rows).  Exact form against the CPU oracle: ranks equal up to rounding ties, results
    within 1e-10."""
    H, O = _H(), _orc()
    from hcorepp_amd.workloads import lowrank_factors
    print(f"[case] ra={ra} rc={rc}", flush=True)
    rng = np.random.default_rng(500 + ra)
    ups = []
    for _ in range(2):
        au, av = lowrank_factors(rng, 512, 512, ra, 0.995)
        bu, bv = lowrank_factors(rng, 512, 512, ra, 0.995)
        cu, cv = lowrank_factors(rng, 512, 512, rc, 0.995)
        ups.append(dict(au=au, av=av, bu=bu, bv=bv, cu=cu, cv=cv))
    acc = 1e-8
    As = [H.CompressedTile(u["au"], u["av"]) for u in ups]
    Bs = [H.CompressedTile(u["bu"], u["bv"]) for u in ups]
    Cs = [H.CompressedTile(u["cu"], u["cv"], max_rank=384) for u in ups]
    b = H.Batch(ctx, As, Bs, Cs, 1.0, 1.0, H.CompressionParameters(acc))
    b.run()
    rk = b.sync_ranks_to_tiles()
    assert b.status()[0] == 0
    kw = b.core_widths()
    print(f"core widths {kw.tolist()}, ranks {rk.tolist()}")
    assert int(kw.min()) > 0.9 * (ra + rc)  # the whole stacked rank stays in the core
    for i, u in enumerate(ups):
        oC = O.Tile.compressed(u["cu"], u["cv"], max_rank=384)
        _, sig = O.hcore_gemm(1.0, O.Tile.compressed(u["au"], u["av"]), O.Tile.compressed(u["bu"], u["bv"]), 1.0,
                              oC, O.Params(acc), want_sigma=True)
        if rk[i] != oC.rank:
            assert rank_tie(sig, oC.rank, rk[i], acc), (i, rk[i], oC.rank)
            continue
        assert rel_fro(Cs[i].to_dense(), oC.to_dense()) < 1e-10, i
    b.close()


def test_batched_mixed_precision_meets_1e8(ctx):
    """Mixed form: the factor products on FP32 matrix cores with FP32 accumulation, over the
    operands' error-free slices (10-bit integer slices per power-of-two row/column scale, exact
    FP32 partial sums, FP64 sums of the slice pairs).  Its result stays within the form's 1e-8
    bar of the FP64 oracle (the dropped slices are ~2^-40 of each term), and the FP64 batched form
    of the same inputs at FP64 rounding."""
    H, O = _H(), _orc()
    from hcorepp_amd.workloads import batch_ccc
    ups = batch_ccc(16, 256, seed=7, lo=8, hi=96)
    acc = 1e-8
    res = {}
    for prec in (H.PREC_FP64, H.PREC_MIXED):
        As = [H.CompressedTile(u["au"], u["av"]) for u in ups]
        Bs = [H.CompressedTile(u["bu"], u["bv"]) for u in ups]
        Cs = [H.CompressedTile(u["cu"], u["cv"], max_rank=200) for u in ups]
        b = H.Batch(ctx, As, Bs, Cs, 1.0, 1.0, H.CompressionParameters(acc), prec)
        b.run()
        b.sync_ranks_to_tiles()
        assert b.status()[0] == 0
        errs = []
        for i, u in enumerate(ups):
            oC = O.Tile.compressed(u["cu"], u["cv"], max_rank=200)
            O.hcore_gemm(1.0, O.Tile.compressed(u["au"], u["av"]), O.Tile.compressed(u["bu"], u["bv"]), 1.0, oC,
                         O.Params(acc))
            errs.append(rel_fro(Cs[i].to_dense(), oC.to_dense()))
        res[prec] = max(errs)
        b.close()
    assert res[H.PREC_FP64] < 1e-12
    # FP32 matrix-core products of the error-free operand slices: within the form's 1e-8 bar
    assert res[H.PREC_MIXED] < 1e-8, res[H.PREC_MIXED]


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_kernel_gemm_matches_numpy(ctx, ta, tb):
    """hcorepp_kernel_gemm_d (HCoreKernels<T>::Gemm) on ragged sizes, all transposes."""
    H = _H()
    rng = np.random.default_rng(5)
    m, n, k = 131, 77, 45
    A = rng.standard_normal((m, k) if ta == "N" else (k, m))
    B = rng.standard_normal((k, n) if tb == "N" else (n, k))
    Cm = rng.standard_normal((m, n))
    dA = H.DeviceBuffer.from_numpy(np.asfortranarray(A).reshape(-1, order="F"))
    dB = H.DeviceBuffer.from_numpy(np.asfortranarray(B).reshape(-1, order="F"))
    dC = H.DeviceBuffer.from_numpy(np.asfortranarray(Cm).reshape(-1, order="F"))
    H._check(H.lib().hcorepp_kernel_gemm_d(ctx.h, ta.encode(), tb.encode(), m, n, k, -0.5, dA.ptr, A.shape[0],
                                           dB.ptr, B.shape[0], 2.0, dC.ptr, m))
    ctx.sync()
    got = dC.download(np.float64).reshape(m, n, order="F")
    ref = -0.5 * (A if ta == "N" else A.T) @ (B if tb == "N" else B.T) + 2.0 * Cm
    assert rel_fro(got, ref) < 1e-14
