"""GPU tests of the TLR Cholesky tile kernels and driver (hcorepp_amd.cholesky) against the
reference's known-answer vectors (tests/kernels/TestKernels.cpp:1070-1248) and the CPU
restatement of the tile sequence (oracle/cholesky.py)."""
import json
import os

import numpy as np
import pytest

from helpers import rel_fro

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.json")))["kernels"]


def _mods():
    import hcorepp_amd as H
    from hcorepp_amd import cholesky as HC
    from oracle import cholesky as OC
    return H, HC, OC


def test_potrf_known_answer(ctx):
    H, HC, OC = _mods()
    import ctypes as C
    v = GOLD["potrf_lower"]
    t = H.DenseTile(np.array(v["a"], dtype=float))
    d = t.c()
    H._check(H.lib().hcorepp_potrf_d(ctx.h, C.byref(d)))
    ctx.sync()
    assert np.array_equal(t.to_numpy(), np.array(v["expected"], dtype=float))


def test_potrf_rejects_indefinite(ctx):
    H, HC, OC = _mods()
    import ctypes as C
    t = H.DenseTile(np.array([[1.0, 2.0], [2.0, 1.0]]))
    d = t.c()
    with pytest.raises(H.HCoreError):
        H._check(H.lib().hcorepp_potrf_d(ctx.h, C.byref(d)))


def test_trsm_known_answer(ctx):
    """Reference vector A x = b (A upper) restated as X L^T = V with L = J A J lower."""
    H, HC, OC = _mods()
    import ctypes as C
    v = GOLD["trsm_left_upper"]
    A = np.array(v["a"], dtype=float)
    b = np.array(v["b"], dtype=float)[:, 0]
    L = H.DenseTile(A[::-1, ::-1].copy())
    B = H.CompressedTile(np.ones((5, 1)), b[::-1][None, :].copy(), max_rank=2)
    arr = (H._Tile * 1)(B.c())
    H._check(H.lib().hcorepp_trsm_batch_d(ctx.h, 1, C.byref(L.c()), arr))
    ctx.sync()
    u, vv = B.factors()
    assert np.array_equal(u, np.ones((5, 1)))
    assert np.allclose(vv[0][::-1], np.array(v["expected"], dtype=float)[:, 0], rtol=0, atol=1e-14)


def test_syrk_known_answer(ctx):
    H, HC, OC = _mods()
    import ctypes as C
    v = GOLD["syrk_lower"]
    A = np.array(v["a"], dtype=float)
    Ct = H.DenseTile(np.full((4, 4), 7.0))
    Ut = H.CompressedTile(A, np.eye(4))
    H._check(H.lib().hcorepp_syrk_batch_d(ctx.h, 1, (H._Tile * 1)(Ut.c()), (H._Tile * 1)(Ct.c())))
    ctx.sync()
    out = Ct.to_numpy()
    exp = np.tril(np.array(v["expected"], dtype=float))
    assert np.allclose(np.tril(7.0 - out), exp, rtol=1e-11, atol=0)
    assert np.array_equal(np.triu(out, 1), np.zeros((4, 4)))


@pytest.mark.parametrize("nb", [128, 200, 512])
def test_tile_ops_random_match_oracle(ctx, nb):
    """potrf / trsm / syrk / transpose on random tiles of several ranks, batched (tile sizes on and
    off the 32-column block of the blocked potrf and the 16-column block of the row solves)."""
    H, HC, OC = _mods()
    import ctypes as C
    print(f"[case] tile ops nb={nb}", flush=True)
    rng = np.random.default_rng(11 + nb)
    G = rng.standard_normal((nb, nb))
    S = G @ G.T + nb * np.eye(nb)
    D = H.DenseTile(S)
    H._check(H.lib().hcorepp_potrf_d(ctx.h, C.byref(D.c())))
    ctx.sync()
    Lg = D.to_numpy()
    Lo = OC.potrf(S)
    assert rel_fro(Lg, Lo) < 1e-13
    ranks = [0, 1, 7, 33, 64, 300] if nb == 512 else [0, 1, 7, 33, 64]
    us = [rng.standard_normal((nb, r)) for r in ranks]
    vs = [rng.standard_normal((r, nb)) for r in ranks]
    Bs = [H.CompressedTile(u, v, max_rank=max(ranks) + 16) for u, v in zip(us, vs)]
    H._check(H.lib().hcorepp_trsm_batch_d(ctx.h, len(Bs), C.byref(D.c()), (H._Tile * len(Bs))(*[b.c() for b in Bs])))
    Cs = [H.DenseTile(S) for _ in ranks]
    H._check(H.lib().hcorepp_syrk_batch_d(ctx.h, len(Bs), (H._Tile * len(Bs))(*[b.c() for b in Bs]),
                                          (H._Tile * len(Cs))(*[c.c() for c in Cs])))
    Ts = [H.CompressedTile(np.zeros((nb, 0)), np.zeros((0, nb)), max_rank=max(ranks) + 16) for _ in ranks]
    src = (H._Tile * len(Bs))(*[b.c() for b in Bs])
    dst = (H._Tile * len(Ts))(*[t.c() for t in Ts])
    H._check(H.lib().hcorepp_transpose_batch_d(ctx.h, len(Bs), src, dst))
    for t, d in zip(Ts, dst):
        t.update_from(d)
    ctx.sync()
    for r, u, v, b, c, t in zip(ranks, us, vs, Bs, Cs, Ts):
        vo = OC.trsm_v(Lo, v)
        ug, vg = b.factors()
        assert np.array_equal(ug, u)
        assert rel_fro(vg, vo) < 1e-12 if r else vg.size == 0
        co = OC.syrk_compressed(S, u, vo)
        assert rel_fro(c.to_numpy(), co) < 1e-13
        assert t.rank == r
        tu, tv = t.factors()
        assert np.array_equal(tu, vg.T) and np.array_equal(tv, ug.T)


def _small(acc, n=512):
    from hcorepp_amd.workloads import TLRMatrix
    M = TLRMatrix(n, 128, 2, 0.1, acc, 20261205, nugget=1e-2)
    diag = [M.tile(k, k)[1] for k in range(M.T)]
    off = {(i, k): M.tile(i, k)[1:] for i in range(M.T) for k in range(i)}
    return M, OC_factor(diag, off, acc)


def OC_factor(diag, off, acc):
    from oracle import cholesky as OC
    return OC.TLRCholesky(diag, off, acc).factor()


@pytest.mark.parametrize("form", ["plain", "batched", "hybrid"])
@pytest.mark.parametrize("acc", [1e-5, 1e-8])
def test_tlr_cholesky_matches_oracle(ctx, form, acc):
    """The device factor equals the oracle's tile sequence to rounding (ranks may differ by a
    tie at the threshold; the factors then differ by O(acc)); backward error O(acc).  'batched'
    runs every step's trailing updates as one batch; 'hybrid' (8 x 8 tiles) the default split,
    batches for the first steps and the plain form for the last ones."""
    H, HC, OC = _mods()
    print(f"[case] cholesky {form} acc {acc}", flush=True)
    M, F_o = _small(acc, n=1024 if form == "hybrid" else 512)
    F = HC.TLRCholesky(ctx, M)
    p = H.CompressionParameters(acc)
    if form == "plain":
        F.factor_plain(p)
    elif form == "batched":
        F.PLAIN_STEP_MAX_UPDATES = 0
        F.factor_batched(p)
    else:
        F.factor_batched(p)
    Ld, Lo = F.dense_factor(), F_o.dense_factor()
    rg, ro = F.ranks(), F_o.ranks()
    diff = sum(abs(rg[k] - ro[k]) for k in ro)
    print("ranks gpu", rg, "oracle", ro)
    assert diff <= 2 + len(ro) // 10
    assert rel_fro(Ld, Lo) < (1e-12 if diff == 0 else 10 * acc)
    A = M.dense_block_range(0, M.n)
    be = np.linalg.norm(A - Ld @ Ld.T) / np.linalg.norm(A)
    be_o = np.linalg.norm(A - Lo @ Lo.T) / np.linalg.norm(A)
    assert be <= 1.5 * be_o + 1e-14, (be, be_o)


def test_tlr_cholesky_batched_mixed_backward_error(ctx):
    """Mixed form (FP32 matrix-core factor products): backward error at FP32 level."""
    H, HC, OC = _mods()
    M, F_o = _small(1e-5)
    F = HC.TLRCholesky(ctx, M)
    F.factor_batched(H.CompressionParameters(1e-5), precision=H.PREC_MIXED)
    Ld = F.dense_factor()
    A = M.dense_block_range(0, M.n)
    be = np.linalg.norm(A - Ld @ Ld.T) / np.linalg.norm(A)
    assert be < 5e-6, be
