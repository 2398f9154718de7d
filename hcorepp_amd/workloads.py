"""Synthetic workloads for the TLR tile GEMM + recompression path.

* ``covariance_points`` / ``covariance_matrix``: points uniform in the unit
  square/cube from a seed, sorted along a Morton (Z-order) curve, and the
  exponential covariance exp(-|x_i - x_j| / ell) (+ nugget * I).
* ``lowrank_factors``: U = orth(G_{m x r}), V = diag(decay^i) orth(G_{n x r})^T,
  i.e. U has orthonormal columns and V carries the singular values, the
  reference's compressed-tile convention (src/operators/concrete/Compressed.cpp:121-135).
* ``batch_ccc``: independent CCC updates with ranks drawn log-uniform.
* ``TLRMatrix``: a tile-low-rank covariance matrix built tile by tile on demand: diagonal
  tiles dense, off-diagonal tiles compressed by the compression-constructor rule.

All generation is host-side numpy with ``numpy.random.default_rng(seed)``.
"""
from __future__ import annotations

import numpy as np

BASE_SEED = 20261202


def _spread_bits(x: np.ndarray, dims: int) -> np.ndarray:
    x = x.astype(np.uint64)
    out = np.zeros_like(x)
    for b in range(21 if dims == 3 else 32):
        out |= ((x >> np.uint64(b)) & np.uint64(1)) << np.uint64(dims * b)
    return out


def morton_order(points: np.ndarray) -> np.ndarray:
    """Permutation sorting ``points`` (n x d in [0,1)^d, d in {2,3}) along a Morton curve."""
    n, d = points.shape
    bits = 21 if d == 3 else 31
    q = np.clip(np.floor(points * (1 << bits)), 0, (1 << bits) - 1).astype(np.uint64)
    code = np.zeros(n, dtype=np.uint64)
    for k in range(d):
        code |= _spread_bits(q[:, k], d) << np.uint64(k)
    return np.argsort(code, kind="stable")


def covariance_points(n: int, dims: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = rng.random((n, dims))
    return p[morton_order(p)]


def covariance_matrix(points: np.ndarray, ell: float, nugget: float = 0.0, rows=None, cols=None) -> np.ndarray:
    """exp(-|x_i - x_j| / ell) (+ nugget on the diagonal) for row/col index ranges."""
    rows = slice(None) if rows is None else rows
    cols = slice(None) if cols is None else cols
    X, Y = points[rows], points[cols]
    d2 = np.zeros((X.shape[0], Y.shape[0]))
    for k in range(points.shape[1]):
        d2 += (X[:, k:k + 1] - Y[None, :, k]) ** 2
    A = np.exp(-np.sqrt(d2) / ell)
    if nugget:
        ri = np.arange(points.shape[0])[rows]
        ci = np.arange(points.shape[0])[cols]
        A[ri[:, None] == ci[None, :]] += nugget
    return np.asfortranarray(A)


def _orth(rng, m, r):
    q, _ = np.linalg.qr(rng.standard_normal((m, r)))
    return q


def lowrank_factors(rng, m, n, r, decay=0.85, scale=1.0):
    """U (m x r, orthonormal columns), V (r x n) = diag(scale*decay^i) Q^T."""
    u = _orth(rng, m, r)
    q = _orth(rng, n, r)
    s = scale * decay ** np.arange(r)
    v = (s[:, None] * q.T)
    return np.asfortranarray(u), np.asfortranarray(v)


def loguniform_ranks(rng, n, lo=8, hi=192):
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=n))).astype(np.int64)


def _ccc_update(seed: int, i: int, nb: int, ra: int, rb: int, rc: int, decay: float):
    rng = np.random.default_rng([seed, i])
    au, av = lowrank_factors(rng, nb, nb, ra, decay)
    bu, bv = lowrank_factors(rng, nb, nb, rb, decay)
    cu, cv = lowrank_factors(rng, nb, nb, rc, decay)
    return dict(au=au, av=av, bu=bu, bv=bv, cu=cu, cv=cv)


def batch_ccc(n: int, nb: int, seed: int, lo=8, hi=192, decay=0.85, threads: int = 8):
    """Independent CCC updates: list of dicts with (au, av, bu, bv, cu, cv).  Each factor pair
    is U = orth(G_{nb x r}), V = diag(decay^i) orth(G_{nb x r})^T with r log-uniform in [lo, hi].
    Update i draws from its own generator seeded with (seed, i), so the result does not depend
    on the thread count."""
    import concurrent.futures as cf
    rng = np.random.default_rng(seed)
    ra = loguniform_ranks(rng, n, lo, hi)
    rb = loguniform_ranks(rng, n, lo, hi)
    rc = loguniform_ranks(rng, n, lo, hi)
    args = [(seed, i, nb, int(ra[i]), int(rb[i]), int(rc[i]), decay) for i in range(n)]
    if threads <= 1 or n < 16:
        return [_ccc_update(*a) for a in args]
    try:
        from threadpoolctl import threadpool_limits
        limit = threadpool_limits(limits=1)
    except Exception:  # pragma: no cover - optional dependency
        limit = None
    try:
        with cf.ThreadPoolExecutor(max_workers=threads) as ex:
            return list(ex.map(lambda a: _ccc_update(*a), args))
    finally:
        if limit is not None:
            limit.restore_original_limits()


def compress_tile(a: np.ndarray, acc: float, max_rank: int | None = None):
    """Dense tile -> (U, V) by the rule of the compressed-tile constructor
    (reference src/operators/concrete/Compressed.cpp:75-146): full SVD, rank = first i >= 1 with
    sigma_i < acc (absolute), at least 1, at most maxRank = max(min(m, n) / 3, 1); U = U_r,
    V = Sigma_r V_r^T."""
    m, n = a.shape
    mr = max(min(m, n) // 3, 1) if max_rank is None else max_rank
    u, s, vt = np.linalg.svd(a, full_matrices=False)
    below = np.nonzero(s[1:] < acc)[0]
    r = int(below[0]) + 1 if len(below) else len(s)
    r = min(max(r, 1), mr)
    return np.asfortranarray(u[:, :r]), np.asfortranarray(s[:r, None] * vt[:r]), mr


class TLRMatrix:
    """T x T tiles of nb rows of the covariance exp(-|x_i - x_j| / ell) over `n` points drawn
    uniform in the unit square (dims=2) or cube (dims=3) from `seed` and sorted along a Morton
    curve.  tile(i, j) returns ('D', dense) on the diagonal and ('C', U, V, max_rank) off it,
    compressed at absolute accuracy `acc`; tiles are built on first use and cached."""

    def __init__(self, n: int, nb: int, dims: int, ell: float, acc: float, seed: int, nugget: float = 0.0):
        self.n, self.nb, self.acc, self.ell, self.nugget = n, nb, acc, ell, nugget
        self.T = n // nb
        self.points = covariance_points(n, dims, seed)
        self._cache = {}

    def dense_block(self, i: int, j: int) -> np.ndarray:
        nb = self.nb
        return covariance_matrix(self.points, self.ell, self.nugget, slice(i * nb, (i + 1) * nb),
                                 slice(j * nb, (j + 1) * nb))

    def dense_block_range(self, r0: int, r1: int) -> np.ndarray:
        """Dense rows/cols [r0, r1) of the matrix (with the nugget on the diagonal)."""
        return covariance_matrix(self.points, self.ell, self.nugget, slice(r0, r1), slice(r0, r1))

    def tile(self, i: int, j: int):
        key = (i, j)
        if key not in self._cache:
            blk = self.dense_block(i, j)
            if i == j:
                self._cache[key] = ("D", blk)
            else:
                u, v, mr = compress_tile(blk, self.acc)
                self._cache[key] = ("C", u, v, mr)
        return self._cache[key]


def batch_ccc_pool(n: int, nb: int, seed: int, lo=8, hi=192, decay=0.85, pool: int = 96):
    """Same distribution as batch_ccc (ranks log-uniform, U orthonormal, V = diag(decay^i) Q^T)
    but fast: every orthonormal factor is a random column subset of one of `pool` Haar-random
    nb x nb orthogonal matrices (a column subset of a Haar matrix is a uniformly random
    orthonormal basis); the six factors of an update come from six different pool matrices.
    Used to generate rank-predictor training data in bulk."""
    rng = np.random.default_rng(seed)
    Q = []
    for _ in range(pool):
        q, r = np.linalg.qr(rng.standard_normal((nb, nb)))
        Q.append(q * np.sign(np.diag(r)))
    ra = loguniform_ranks(rng, n, lo, hi)
    rb = loguniform_ranks(rng, n, lo, hi)
    rc = loguniform_ranks(rng, n, lo, hi)
    out = []
    for i in range(n):
        pk = rng.choice(pool, 6, replace=False)

        def fac(k0, k1, r):
            u = Q[pk[k0]][:, rng.choice(nb, r, replace=False)]
            v = (decay ** np.arange(r))[:, None] * Q[pk[k1]][:, rng.choice(nb, r, replace=False)].T
            return np.asfortranarray(u), np.asfortranarray(v)

        au, av = fac(0, 1, int(ra[i]))
        bu, bv = fac(2, 3, int(rb[i]))
        cu, cv = fac(4, 5, int(rc[i]))
        out.append(dict(au=au, av=av, bu=bu, bv=bv, cu=cu, cv=cv))
    return out


def prefetch_tiles(M: "TLRMatrix", keys, threads: int = 8):
    """Build (and cache) the listed tiles of a TLRMatrix in parallel (one BLAS thread each)."""
    import concurrent.futures as cf
    keys = [k for k in keys if k not in M._cache]
    try:
        from threadpoolctl import threadpool_limits
        limit = threadpool_limits(limits=1)
    except Exception:  # pragma: no cover
        limit = None
    try:
        with cf.ThreadPoolExecutor(max_workers=threads) as ex:
            list(ex.map(lambda k: M.tile(*k), keys))
    finally:
        if limit is not None:
            limit.restore_original_limits()
