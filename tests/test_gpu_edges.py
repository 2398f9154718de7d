"""GPU edge cases of the C-ABI: empty and ragged batches, empty C tiles, rank capped by the
tile capacity, and the documented error paths (shape mismatch, sizes beyond the limits,
non-square potrf)."""
import ctypes as C

import numpy as np
import pytest

from helpers import rel_fro

pytestmark = pytest.mark.gpu


def _H():
    import hcorepp_amd as H
    return H


def _lowrank(rng, m, n, r, decay=0.7):
    from hcorepp_amd.workloads import lowrank_factors
    return lowrank_factors(rng, m, n, r, decay)


def test_empty_batch(ctx):
    H = _H()
    b = H.Batch(ctx, [], [], [], 1.0, 1.0, H.CompressionParameters(1e-8))
    b.run()
    assert b.status()[0] == 0
    assert len(b.ranks()) == 0
    b.close()


def test_ragged_batch_matches_plain(ctx):
    """Tiles of different shapes and ranks in one batch, including an empty C (rank 0)."""
    H = _H()
    rng = np.random.default_rng(21)
    shapes = [(64, 96, 5, 7, 0), (200, 128, 20, 15, 30), (128, 200, 12, 9, 11), (33, 47, 3, 4, 6),
              (512, 300, 40, 30, 64), (96, 96, 1, 1, 1)]
    As, Bs, Cb, Cp = [], [], [], []
    for (m, n, ra, rb, rc) in shapes:
        au, av = _lowrank(rng, m, n, ra)
        bu, bv = _lowrank(rng, n, n, rb)
        cu, cv = _lowrank(rng, m, n, rc) if rc else (np.zeros((m, 0)), np.zeros((0, n)))
        As.append(H.CompressedTile(au, av))
        Bs.append(H.CompressedTile(bu, bv))
        Cb.append(H.CompressedTile(cu, cv, max_rank=min(m, n) // 2))
        Cp.append(H.CompressedTile(cu, cv, max_rank=min(m, n) // 2))
    b = H.Batch(ctx, As, Bs, Cb, 1.0, 1.0, H.CompressionParameters(1e-8))
    b.run()
    rk = b.sync_ranks_to_tiles()
    assert b.status()[0] == 0
    for i, s in enumerate(shapes):
        print(f"[case] ragged {s}", flush=True)
        H.HCore.gemm(1.0, As[i], Bs[i], 1.0, Cp[i], ctx, H.CompressionParameters(1e-8))
        assert rk[i] == Cp[i].rank
        assert rel_fro(Cb[i].to_dense(), Cp[i].to_dense()) < 1e-11
    b.close()


def test_rank_capped_by_capacity(ctx):
    """A result rank above the C tile's max_rank is cut to max_rank (Compressed.cpp rank clamp):
    the truncation error is then the dropped tail of the singular values."""
    H = _H()
    rng = np.random.default_rng(5)
    m = 128
    au, av = _lowrank(rng, m, m, 40, decay=0.95)
    bu, bv = _lowrank(rng, m, m, 40, decay=0.95)
    cu, cv = _lowrank(rng, m, m, 10, decay=0.95)
    Ct = H.CompressedTile(cu, cv, max_rank=16)
    A, B = H.CompressedTile(au, av), H.CompressedTile(bu, bv)
    _, sig = H.HCore.gemm(1.0, A, B, 1.0, Ct, ctx, H.CompressionParameters(1e-12), want_sigma=True)
    assert Ct.rank == 16
    dense = cu @ cv + au @ av @ bu @ bv
    assert np.linalg.norm(Ct.to_dense() - dense) <= 1.01 * np.sqrt(np.sum(sig[16:] ** 2)) + 1e-12


def test_error_paths(ctx):
    H = _H()
    rng = np.random.default_rng(8)
    au, av = _lowrank(rng, 64, 64, 4)
    A = H.CompressedTile(au, av)
    Cbad = H.CompressedTile(np.zeros((32, 2)), np.zeros((2, 64)), max_rank=8)  # rows do not match A
    with pytest.raises(H.HCoreError, match="INVALID"):
        H.HCore.gemm(1.0, A, A, 1.0, Cbad, ctx, H.CompressionParameters(1e-8))
    # stacked rank beyond the core-SVD limit (768)
    m = 1024
    bu, bv = _lowrank(rng, m, m, 500, decay=0.999)
    cu, cv = _lowrank(rng, m, m, 300, decay=0.999)
    big = H.CompressedTile(bu, bv)
    Cbig = H.CompressedTile(cu, cv, max_rank=341)
    with pytest.raises(H.HCoreError, match="UNSUPPORTED"):
        H.HCore.gemm(1.0, big, big, 1.0, Cbig, ctx, H.CompressionParameters(1e-8))
    # potrf needs a square dense tile
    d = H.DenseTile(np.eye(4, 6)).c()
    with pytest.raises(H.HCoreError, match="INVALID"):
        H._check(H.lib().hcorepp_potrf_d(ctx.h, C.byref(d)))
    # compression beyond the core-SVD limit
    with pytest.raises(H.HCoreError, match="UNSUPPORTED"):
        H.CompressedTile.compress(ctx, rng.standard_normal((800, 800)), H.CompressionParameters(1e-8))


@pytest.mark.parametrize("kw", [dict(fixed_rank=9), dict(truncated_svd=True, accuracy=1e-6)])
def test_batched_rank_rules_match_plain(ctx, kw):
    """Fixed rank and the relative (truncated_svd) rule in the batched form: same ranks and
    results as the plain form."""
    H = _H()
    rng = np.random.default_rng(31)
    acc = kw.pop("accuracy", 1e-8)
    p = H.CompressionParameters(acc, **kw)
    As, Bs, Cb, Cp = [], [], [], []
    for m, ra, rb, rc in [(96, 12, 10, 14), (128, 20, 18, 25), (256, 30, 20, 40)]:
        au, av = _lowrank(rng, m, m, ra, decay=0.9)
        bu, bv = _lowrank(rng, m, m, rb, decay=0.9)
        cu, cv = _lowrank(rng, m, m, rc, decay=0.9)
        As.append(H.CompressedTile(au, av))
        Bs.append(H.CompressedTile(bu, bv))
        Cb.append(H.CompressedTile(cu, cv, max_rank=m // 2))
        Cp.append(H.CompressedTile(cu, cv, max_rank=m // 2))
    b = H.Batch(ctx, As, Bs, Cb, 1.0, 1.0, p)
    b.run()
    rk = b.sync_ranks_to_tiles()
    assert b.status()[0] == 0
    for i in range(len(As)):
        H.HCore.gemm(1.0, As[i], Bs[i], 1.0, Cp[i], ctx, p)
        assert rk[i] == Cp[i].rank
        if "fixed_rank" in kw:
            assert rk[i] == kw["fixed_rank"]
        assert rel_fro(Cb[i].to_dense(), Cp[i].to_dense()) < 1e-11
    b.close()
