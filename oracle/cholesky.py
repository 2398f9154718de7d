"""CPU restatement of the TLR Cholesky tile sequence (right-looking, lower).

TEST INFRASTRUCTURE ONLY: this module is the checker of hcorepp_amd.cholesky.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may import it.

Tile operations, in the engine's compressed-tile convention A = U V (U m x r, V r x n):
  potrf(A_kk)           dense Cholesky, strict upper part zeroed       src/api/HCore.cpp:586-622
  trsm(L_kk, A_ik)      V <- V L_kk^{-T}, U untouched                   src/api/HCore.cpp:624-647
  syrk(A_ik, A_ii)      A_ii <- A_ii - U (V V^T) U^T, strict upper := 0 src/api/HCore.cpp:484-584
  gemm(A_ik, A_jk^T, A_ij)  A_ij <- A_ij - L_ik L_jk^T as an ordinary recompression (oracle.hcore_gemm,
                        src/api/HCore.cpp:22-344) with B = (V_jk^T, U_jk^T)
The reference's Trsm applies blas::trsm to the V factor of the compressed tile (HCore.cpp:636-641);
its Syrk forms V V^T, U (V V^T), then C -= (U V V^T) U^T and zeroes the other triangle
(HCore.cpp:522-566).  Its Gemm with aCholesky=true re-plumbs the operands for a V stored n x r
(HCore.cpp:160-174); here the transposed factor is an explicit tile instead, so the update is a
plain CCC recompression.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import solve_triangular

from . import oracle as O


def potrf(a: np.ndarray) -> np.ndarray:
    """Lower Cholesky factor with zeroed strict upper triangle; raises on a non-SPD tile."""
    return np.linalg.cholesky(np.asarray(a, dtype=np.float64))


def trsm_v(L: np.ndarray, v: np.ndarray) -> np.ndarray:
    """V L^{-T} for lower-triangular L: solves X L^T = V."""
    return solve_triangular(L, v.T, lower=True).T


def syrk_compressed(c: np.ndarray, u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """C - U (V V^T) U^T with the strict upper triangle set to 0."""
    out = c - (u @ (v @ v.T)) @ u.T
    return np.tril(out)


class TLRCholesky:
    """Tiles of a lower TLR matrix as oracle tiles; factor() runs the tile sequence in place."""

    def __init__(self, diag, off, acc: float):
        """diag: list of T dense nb x nb tiles; off: {(i, k): (U, V, max_rank)} for i > k."""
        self.T = len(diag)
        self.acc = acc
        self.D = [np.array(d, dtype=np.float64, order="F") for d in diag]
        self.L = {key: O.Tile.compressed(u, v, mr) for key, (u, v, mr) in off.items()}

    def factor(self):
        T = self.T
        p = O.Params(self.acc)
        for k in range(T):
            self.D[k] = potrf(self.D[k])
            for i in range(k + 1, T):
                t = self.L[(i, k)]
                v = trsm_v(self.D[k], t.V)
                self.L[(i, k)] = O.Tile.compressed(t.U, v, t.max_rank)
            for i in range(k + 1, T):
                t = self.L[(i, k)]
                self.D[i] = syrk_compressed(self.D[i], t.U, t.V)
            for j in range(k + 1, T):
                tj = self.L[(j, k)]
                bt = O.Tile.compressed(np.asfortranarray(tj.V.T), np.asfortranarray(tj.U.T))
                for i in range(j + 1, T):
                    info = O.hcore_gemm(-1.0, self.L[(i, k)], bt, 1.0, self.L[(i, j)], p)
                    if info != 0:
                        raise RuntimeError(f"oracle gemm ({i},{j},{k}) info={info}")
        return self

    def dense_factor(self) -> np.ndarray:
        T, nb = self.T, self.D[0].shape[0]
        Ld = np.zeros((T * nb, T * nb))
        for k in range(T):
            Ld[k * nb:(k + 1) * nb, k * nb:(k + 1) * nb] = self.D[k]
            for i in range(k + 1, T):
                Ld[i * nb:(i + 1) * nb, k * nb:(k + 1) * nb] = self.L[(i, k)].to_dense()
        return Ld

    def ranks(self):
        return {key: t.rank for key, t in self.L.items()}
