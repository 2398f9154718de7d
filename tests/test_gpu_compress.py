"""GPU tests of the compression constructor on the device (hcorepp_compress_d/_s, reference
src/operators/concrete/Compressed.cpp:75-146) against the oracle's restatement of the
constructor (full SVD with MKL, the same rank rule), on covariance tiles and on random tiles
of known spectrum."""
import numpy as np
import pytest

from helpers import EPS, rank_tie, rel_fro

pytestmark = pytest.mark.gpu


def _mods():
    import hcorepp_amd as H
    from oracle import oracle as O
    return H, O


def _spectrum_tile(rng, m, n, decay):
    q1, _ = np.linalg.qr(rng.standard_normal((m, min(m, n))))
    q2, _ = np.linalg.qr(rng.standard_normal((n, min(m, n))))
    s = decay ** np.arange(min(m, n))
    return (q1 * s) @ q2.T, s


@pytest.mark.parametrize("m,n,decay,acc", [(256, 256, 0.9, 1e-8), (512, 512, 0.97, 1e-8), (300, 200, 0.8, 1e-6),
                                           (200, 300, 0.8, 1e-6), (512, 512, 0.5, 1e-12)])
def test_compress_matches_oracle(ctx, m, n, decay, acc):
    H, O = _mods()
    print(f"[case] compress {m}x{n} decay {decay} acc {acc}", flush=True)
    rng = np.random.default_rng(m * 7 + n)
    a, s = _spectrum_tile(rng, m, n, decay)
    t = H.CompressedTile.compress(ctx, a, H.CompressionParameters(acc))
    ot, osig = O.compress(a, O.Params(acc))
    assert t.max_rank == ot.max_rank == max(min(m, n) // 3, 1)
    assert rank_tie(osig, ot.rank, t.rank, acc), (ot.rank, t.rank)
    u, v = t.factors()
    assert np.linalg.norm(u.T @ u - np.eye(t.rank)) < 1e-13 * np.sqrt(m)
    if t.rank == ot.rank:
        assert rel_fro(u @ v, ot.to_dense()) < 1e-10
    # Eckart-Young: the truncation error is the dropped singular values
    tail = np.sqrt(np.sum(osig[t.rank:] ** 2))
    assert np.linalg.norm(u @ v - a) <= 1.01 * tail + 1e-12 * np.linalg.norm(a)


def test_compress_covariance_tiles(ctx):
    """Off-diagonal tiles of the exp(-r/0.1) covariance (the Cholesky benchmark's matrix) at
    acc 1e-8 and 1e-4: ranks and products as the oracle's constructor."""
    H, O = _mods()
    from hcorepp_amd.workloads import TLRMatrix
    M = TLRMatrix(2048, 256, 3, 0.1, 1e-8, 20261205, nugget=1e-2)
    for (i, j) in [(1, 0), (4, 0), (7, 0), (5, 3)]:
        blk = M.dense_block(i, j)
        for acc in (1e-8, 1e-4):
            t = H.CompressedTile.compress(ctx, blk, H.CompressionParameters(acc))
            ot, osig = O.compress(blk, O.Params(acc))
            assert rank_tie(osig, ot.rank, t.rank, acc), ((i, j), acc, ot.rank, t.rank)
            if t.rank == ot.rank:
                assert rel_fro(t.to_dense(), ot.to_dense()) < 1e-10


def test_compress_fixed_and_relative_rank(ctx):
    H, O = _mods()
    rng = np.random.default_rng(3)
    a, s = _spectrum_tile(rng, 256, 256, 0.9)  # max_rank 85
    t = H.CompressedTile.compress(ctx, a, H.CompressionParameters(1e-8, fixed_rank=17))
    assert t.rank == 17
    assert abs(np.linalg.norm(t.to_dense() - a) - np.sqrt(np.sum(s[17:] ** 2))) < 1e-10
    t = H.CompressedTile.compress(ctx, a, H.CompressionParameters(1e-3, truncated_svd=True))
    # relative rule: first sigma_i < 1e-3 * sigma_0 with sigma_i = 0.9^i, i.e. i = 66
    assert t.rank == 66
    t = H.CompressedTile.compress(ctx, np.zeros((64, 64)), H.CompressionParameters(1e-8))
    assert t.rank == 1 and np.all(t.to_dense() == 0.0)


def test_compress_float32(ctx):
    H, O = _mods()
    rng = np.random.default_rng(9)
    a, s = _spectrum_tile(rng, 128, 128, 0.8)
    t = H.CompressedTile.compress(ctx, a.astype(np.float32), H.CompressionParameters(1e-4), dtype=np.float32)
    ot, osig = O.compress(a, O.Params(1e-4))
    assert abs(t.rank - ot.rank) <= 1
    assert np.linalg.norm(t.to_dense() - a) / np.linalg.norm(a) < 1e-4
