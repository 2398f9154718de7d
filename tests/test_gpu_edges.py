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


def test_nan_factors_report_numeric_and_keep_c(ctx):
    """NaN in A's factors: the plain form returns HCOREPP_ERR_NUMERIC (the device status word's
    NaN bit), and in a batch that update reports NUMERIC through the batch status while its C
    keeps its factors and rank; the other updates complete."""
    H = _H()
    rng = np.random.default_rng(41)
    m = 96
    au, av = _lowrank(rng, m, m, 8)
    bu, bv = _lowrank(rng, m, m, 6)
    cu, cv = _lowrank(rng, m, m, 10)
    bad = au.copy()
    bad[3, 2] = np.nan
    A_bad, A_ok, B = H.CompressedTile(bad, av), H.CompressedTile(au, av), H.CompressedTile(bu, bv)
    Cp = H.CompressedTile(cu, cv, max_rank=40)
    with pytest.raises(H.HCoreError, match="NUMERIC"):
        H.HCore.gemm(1.0, A_bad, B, 1.0, Cp, ctx, H.CompressionParameters(1e-8))
    C0 = H.CompressedTile(cu, cv, max_rank=40)
    C1 = H.CompressedTile(cu, cv, max_rank=40)
    b = H.Batch(ctx, [A_bad, A_ok], [B, B], [C0, C1], 1.0, 1.0, H.CompressionParameters(1e-8))
    b.run()
    st, _ = b.status()
    assert H.STATUS[st] == "NUMERIC", st
    rk = b.sync_ranks_to_tiles()
    assert rk[0] == 10
    u0, v0 = C0.factors()
    assert np.array_equal(u0, cu) and np.array_equal(v0, cv)
    Cref = H.CompressedTile(cu, cv, max_rank=40)
    H.HCore.gemm(1.0, A_ok, B, 1.0, Cref, ctx, H.CompressionParameters(1e-8))
    assert rk[1] == Cref.rank
    assert rel_fro(C1.to_dense(), Cref.to_dense()) < 1e-11
    b.close()


def test_batch_core_beyond_limit_keeps_c(ctx):
    """A batched update whose stacked rank exceeds the core-SVD limit (768 rows) is refused at
    creation; one that only exceeds it after its C rank was raised (set_ranks) reports the
    status bit at run time and leaves C unchanged."""
    H = _H()
    rng = np.random.default_rng(43)
    m = 1024
    au, av = _lowrank(rng, m, m, 400, decay=0.999)
    bu, bv = _lowrank(rng, m, m, 20)
    cu, cv = _lowrank(rng, m, m, 400, decay=0.999)
    A, B = H.CompressedTile(au, av), H.CompressedTile(bu, bv)
    Cbig = H.CompressedTile(cu, cv, max_rank=500)
    with pytest.raises(H.HCoreError, match="UNSUPPORTED"):
        H.Batch(ctx, [A], [B], [Cbig], 1.0, 1.0, H.CompressionParameters(1e-8))
    # created at rank 300 (K = 700), run at rank 400 (K = 800 > 768)
    Cs = H.CompressedTile(cu[:, :300], cv[:300], max_rank=500)
    b = H.Batch(ctx, [A], [B], [Cs], 1.0, 1.0, H.CompressionParameters(1e-8))
    full = H.CompressedTile(cu, cv, max_rank=500)  # the rank-400 factors in C's buffer
    H.lib().hcorepp_memcpy(ctx.h, Cs.buf.ptr, full.buf.ptr, full.buf.nbytes)
    b.set_ranks(np.array([400], dtype=np.int64))
    b.run()
    st, _ = b.status()
    assert H.STATUS[st] == "NUMERIC", st
    assert b.ranks()[0] == 400
    Cs.rank = 400
    u, v = Cs.factors()
    assert np.array_equal(u, cu) and np.array_equal(v, cv)
    b.close()


@pytest.mark.parametrize("kind", ["CCC", "DCC", "CDC", "CCD", "DDD", "DDC"])
def test_pool_size_presizes_the_arena(kind):
    """A context whose arena is presized with the pool-size query (HCore<T>::CalculateMemoryPoolSize,
    reference src/api/HCore.cpp:417-480, as par_fixed_rank_streams_main.cpp:86-122 uses it) never
    grows it for that update."""
    H = _H()
    rng = np.random.default_rng(51)
    m = 256
    au, av = _lowrank(rng, m, m, 30)
    bu, bv = _lowrank(rng, m, m, 25)
    cu, cv = _lowrank(rng, m, m, 40)
    A = H.CompressedTile(au, av) if kind[0] == "C" else H.DenseTile(rng.standard_normal((m, m)))
    B = H.CompressedTile(bu, bv) if kind[1] == "C" else H.DenseTile(rng.standard_normal((m, m)))
    Ct = H.CompressedTile(cu, cv, max_rank=m if kind == "DDC" else 120) if kind[2] == "C" else \
        H.DenseTile(rng.standard_normal((m, m)))
    p = H.CompressionParameters(1e-8)
    elems = H.lib().hcorepp_gemm_pool_size_d(C.byref(A.c()), C.byref(B.c()), C.byref(Ct.c()), C.byref(p.c()))
    ctx = H.RunContext(0, arena_bytes=elems * 8)
    assert ctx.arena_bytes() == elems * 8
    H.HCore.gemm(1.0, A, B, 1.0, Ct, ctx, p)
    ctx.sync()
    assert ctx.arena_bytes() == elems * 8, (kind, ctx.arena_bytes(), elems * 8)
    ctx.close()
