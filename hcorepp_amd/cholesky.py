"""TLR Cholesky A = L L^T of a symmetric positive definite tile-low-rank matrix (lower tiles).

Right-looking tile sequence, step k:
    L_kk = potrf(A_kk)                                    (dense)
    L_ik = A_ik L_kk^{-T}              i > k              (compressed: V <- V L_kk^{-T})
    A_ii = A_ii - L_ik L_ik^T          i > k              (dense syrk)
    A_ij = A_ij - L_ik L_jk^T          i > j > k          (compressed GEMM + recompression,
                                                           B operand = L_jk^T as a tile)
The plain form issues one call per tile operation (the trailing updates through HCore.gemm);
the batched form issues, per step, one launch for all trsm, one for all syrk, one for the
transposes and one persistent recompression batch for all trailing updates (FP64 or mixed).
Reference tile operations: src/api/HCore.cpp:484-647 (Syrk, Potrf, Trsm) and the aCholesky
Gemm branch :160-174.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import (PREC_FP64, Batch, CompressedTile, CompressionParameters, DenseTile, HCore, _check, _Tile, lib)


def _arr(tiles):
    a = (_Tile * max(len(tiles), 1))()
    for i, t in enumerate(tiles):
        a[i] = t.c()
    return a


def _sync_back(tiles, arr):
    for i, t in enumerate(tiles):
        t.update_from(arr[i])


class TLRCholesky:
    """Device copy of the lower tiles of a TLRMatrix and its in-place factorisation."""

    def __init__(self, ctx, M):
        self.ctx, self.M, self.T, self.nb = ctx, M, M.T, M.nb
        self.D = [DenseTile(M.tile(k, k)[1]) for k in range(self.T)]
        self.L = {}
        for i in range(self.T):
            for k in range(i):
                t = M.tile(i, k)
                self.L[(i, k)] = CompressedTile(t[1], t[2], max_rank=t[3])
        mr = max((t.max_rank for t in self.L.values()), default=1)
        self.LT = {key: CompressedTile(np.zeros((self.nb, 0)), np.zeros((0, self.nb)), max_rank=mr)
                   for key in self.L}

    def _potrf(self, k):
        d = self.D[k].c()
        _check(lib().hcorepp_potrf_d(self.ctx.h, C.byref(d)))

    def _trsm(self, k, rows):
        tiles = [self.L[(i, k)] for i in rows]
        L = self.D[k].c()
        arr = _arr(tiles)
        _check(lib().hcorepp_trsm_batch_d(self.ctx.h, len(tiles), C.byref(L), arr))

    def _syrk(self, k, rows):
        A = [self.L[(i, k)] for i in rows]
        Cd = [self.D[i] for i in rows]
        _check(lib().hcorepp_syrk_batch_d(self.ctx.h, len(A), _arr(A), _arr(Cd)))

    def _transpose(self, k, rows):
        src = [self.L[(j, k)] for j in rows]
        dst = [self.LT[(j, k)] for j in rows]
        sa, da = _arr(src), _arr(dst)
        _check(lib().hcorepp_transpose_batch_d(self.ctx.h, len(src), sa, da))
        _sync_back(dst, da)

    def factor_plain(self, params: CompressionParameters, progress=None):
        """One call per tile operation; trailing updates through the plain HCore.gemm.
        `progress(k)` is called after step k."""
        for k in range(self.T):
            self._potrf(k)
            for i in range(k + 1, self.T):
                self._trsm(k, [i])
            for i in range(k + 1, self.T):
                self._syrk(k, [i])
            for j in range(k + 1, self.T):
                self._transpose(k, [j])
            for j in range(k + 1, self.T):
                for i in range(j + 1, self.T):
                    HCore.gemm(-1.0, self.L[(i, k)], self.LT[(j, k)], 1.0, self.L[(i, j)], self.ctx, params)
            if progress is not None:
                progress(k)
        self.ctx.sync()

    # Steps with at most this many trailing updates run them in the plain form (whose QR and
    # Jacobi phases spread over the whole GPU), PLAIN_STREAMS at a time on contexts of their own: a
    # persistent batch lasts as long as its largest update on one workgroup (~80 ms at nb = 512
    # when the core keeps all K = 340 columns), the plain form takes a fraction of that per update.
    PLAIN_STEP_MAX_UPDATES = 16
    PLAIN_STREAMS = 4

    def _plain_updates(self, k, pairs, params):
        """The trailing updates of step k in the plain form, spread over PLAIN_STREAMS contexts
        (distinct C tiles, so independent); the main stream's work is finished first."""
        import concurrent.futures as cf
        from . import RunContext
        self.ctx.sync()
        if not hasattr(self, "_aux"):
            self._aux = [RunContext(self.ctx.device) for _ in range(self.PLAIN_STREAMS)]

        def share(t):
            for i, j in pairs[t::self.PLAIN_STREAMS]:
                HCore.gemm(-1.0, self.L[(i, k)], self.LT[(j, k)], 1.0, self.L[(i, j)], self._aux[t], params)

        n = min(self.PLAIN_STREAMS, len(pairs))
        with cf.ThreadPoolExecutor(max_workers=n) as ex:
            list(ex.map(share, range(n)))

    def factor_batched(self, params: CompressionParameters, precision=PREC_FP64, tol=0.0):
        """Per step one launch per tile operation kind; the trailing updates as one persistent
        recompression batch (optionally with the pivoted-R tolerance rule); the last steps, with
        a handful of updates, in the plain form (FP64, exact core) on several streams."""
        for k in range(self.T):
            rows = list(range(k + 1, self.T))
            self._potrf(k)
            if not rows:
                continue
            self._trsm(k, rows)
            self._syrk(k, rows)
            self._transpose(k, rows)
            pairs = [(i, j) for j in rows for i in range(j + 1, self.T)]
            if 0 < len(pairs) <= self.PLAIN_STEP_MAX_UPDATES:
                self._plain_updates(k, pairs, params)
            elif pairs:
                b = Batch(self.ctx, [self.L[(i, k)] for i, j in pairs], [self.LT[(j, k)] for i, j in pairs],
                          [self.L[(i, j)] for i, j in pairs], -1.0, 1.0, params, precision)
                if tol > 0:
                    b.set_rank_model(None, tol=tol, margin=0)
                b.run()
                st, _ = b.status()
                if st != 0:
                    raise RuntimeError(f"trailing update batch of step {k}: status {st}")
                b.sync_ranks_to_tiles()
                b.close()
        self.ctx.sync()

    def dense_factor(self) -> np.ndarray:
        """L assembled densely (n x n, lower)."""
        T, nb = self.T, self.nb
        Ld = np.zeros((T * nb, T * nb))
        for k in range(T):
            Ld[k * nb:(k + 1) * nb, k * nb:(k + 1) * nb] = np.tril(self.D[k].to_numpy())
            for i in range(k + 1, T):
                Ld[i * nb:(i + 1) * nb, k * nb:(k + 1) * nb] = self.L[(i, k)].to_dense()
        return Ld

    def ranks(self):
        return {key: t.rank for key, t in self.L.items()}


def backward_error(M, Ld: np.ndarray) -> float:
    """||A - L L^T||_F / ||A||_F with A the dense matrix of M."""
    n = Ld.shape[0]
    A = M.dense_block_range(0, n)
    R = A - Ld @ Ld.T
    return float(np.linalg.norm(R) / np.linalg.norm(A))
