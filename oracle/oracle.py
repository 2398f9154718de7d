"""ctypes front end of the CPU oracle (oracle/hcore_oracle.c).

TEST INFRASTRUCTURE ONLY: this module is the checker.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may import it.

Tiles are numpy-backed mirrors of the reference's TileMetadata + data pointer
(include/hcorepp/operators/interface/Tile.hpp:30-52): a compressed tile's buffer
is ``[U (rows x max_rank, ld rows) | V (max_rank x cols, ld rank)]``
(src/operators/concrete/Compressed.cpp:175-185).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libhcore_oracle.so")

DENSE, COMPRESSED = 0, 1
GESVD, GESDD = 0, 1
ERR_UNSUPPORTED, ERR_CAPACITY = -100, -101


class _Tile(C.Structure):
    _fields_ = [("rows", C.c_int64), ("cols", C.c_int64), ("rank", C.c_int64), ("max_rank", C.c_int64),
                ("ld", C.c_int64), ("type", C.c_int32), ("pad", C.c_int32), ("data", C.c_void_p)]


class _Params(C.Structure):
    _fields_ = [("accuracy", C.c_double), ("trmm", C.c_int32), ("ungqr", C.c_int32),
                ("truncated_svd", C.c_int32), ("svd_type", C.c_int32), ("fixed_rank", C.c_int64)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc only)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        L.orc_init.argtypes = [C.c_char_p]
        if L.orc_init(None) != 0:
            raise RuntimeError("hcore_oracle: no LAPACK runtime (libmkl_rt / libtorch_cpu) found")
        L.orc_lapack_path.restype = C.c_char_p
        L.orc_calculate_new_rank.restype = C.c_int64
        L.orc_calculate_new_rank.argtypes = [C.c_int, C.c_void_p, C.c_int64, C.c_double]
        L.orc_process_vpointer.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_void_p, C.c_int64,
                                           C.c_void_p, C.c_int64, C.c_void_p]
        L.orc_calculate_uvptr.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p]
        L.orc_calculate_vtnew.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64]
        L.orc_lacpy.argtypes = [C.c_char, C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64]
        L.orc_laset.argtypes = [C.c_char, C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_void_p, C.c_int64]
        L.orc_geqrf.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p]
        L.orc_svd.argtypes = [C.c_int, C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                              C.c_int64, C.c_void_p, C.c_int64]
        L.orc_dgemm.argtypes = [C.c_char, C.c_char, C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_void_p,
                                C.c_int64, C.c_void_p, C.c_int64, C.c_double, C.c_void_p, C.c_int64]
        L.orc_hcore_gemm.argtypes = [C.c_double, C.POINTER(_Tile), C.c_char, C.POINTER(_Tile), C.c_char,
                                     C.c_double, C.POINTER(_Tile), C.POINTER(_Params), C.c_void_p,
                                     C.POINTER(C.c_int64)]
        L.orc_compress.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.POINTER(_Params),
                                   C.POINTER(_Tile), C.c_void_p]
        L.orc_compress_dense_matrix.restype = C.c_int64
        L.orc_compress_dense_matrix.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_double,
                                                C.c_void_p]
        L.orc_latms_dense.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_double]
        L.orc_dpotrf.argtypes = [C.c_char, C.c_int64, C.c_void_p, C.c_int64]
        L.orc_batch_gemm.restype = C.c_double
        L.orc_batch_gemm.argtypes = [C.c_int64, C.c_double, C.POINTER(_Tile), C.POINTER(_Tile), C.c_double,
                                     C.POINTER(_Tile), C.POINTER(_Params), C.c_int, C.c_void_p]
        L.orc_tlr_gemm_omp.restype = C.c_double
        L.orc_tlr_gemm_omp.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.c_double, C.POINTER(_Tile),
                                       C.POINTER(_Tile), C.c_double, C.POINTER(_Tile), C.POINTER(_Params), C.c_int,
                                       C.c_int]
        _lib = L
    return _lib


def lapack_path() -> str:
    return lib().orc_lapack_path().decode()


def _ptr(a: np.ndarray):
    assert a.flags["F_CONTIGUOUS"] or a.ndim == 1 or a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


class Params:
    """CompressionParameters mirror; reference defaults (CompressionParameters.hpp:44-46)."""

    def __init__(self, accuracy=1e-4, trmm=False, ungqr=True, truncated_svd=False, fixed_rank=0, svd_type=GESDD):
        self.accuracy, self.trmm, self.ungqr = float(accuracy), bool(trmm), bool(ungqr)
        self.truncated_svd, self.fixed_rank, self.svd_type = bool(truncated_svd), int(fixed_rank), int(svd_type)

    def c(self) -> _Params:
        return _Params(self.accuracy, int(self.trmm), int(self.ungqr), int(self.truncated_svd), self.svd_type,
                       self.fixed_rank)


class Tile:
    """Host tile (numpy float64 buffer) in the reference layout."""

    def __init__(self, rows, cols, rank, max_rank, ld, kind, data):
        self.rows, self.cols, self.rank, self.max_rank, self.ld = int(rows), int(cols), int(rank), int(
            max_rank), int(ld)
        self.type = int(kind)
        self.data = np.ascontiguousarray(data, dtype=np.float64).reshape(-1)

    # -- constructors -------------------------------------------------------
    @staticmethod
    def dense(a: np.ndarray) -> "Tile":
        a = np.asfortranarray(a, dtype=np.float64)
        m, n = a.shape
        return Tile(m, n, min(m, n), min(m, n), m, DENSE, a.reshape(-1, order="F"))

    @staticmethod
    def compressed(u: np.ndarray, v: np.ndarray, max_rank: int | None = None) -> "Tile":
        """UV tile: ``u`` rows x r, ``v`` r x cols; max_rank defaults to r
        (src/operators/concrete/Compressed.cpp:19-47)."""
        u = np.asarray(u, dtype=np.float64)
        v = np.asarray(v, dtype=np.float64)
        m, r = u.shape
        r2, n = v.shape
        assert r == r2
        mr = r if max_rank is None else int(max_rank)
        assert mr >= r
        buf = np.zeros((m + n) * mr)
        buf[: m * r] = u.reshape(-1, order="F")
        buf[m * mr: m * mr + r * n] = v.reshape(-1, order="F")
        return Tile(m, n, r, mr, m, COMPRESSED, buf)

    def copy(self) -> "Tile":
        return Tile(self.rows, self.cols, self.rank, self.max_rank, self.ld, self.type, self.data.copy())

    # -- views --------------------------------------------------------------
    @property
    def U(self) -> np.ndarray:
        return self.data[: self.rows * self.rank].reshape(self.rows, self.rank, order="F")

    @property
    def V(self) -> np.ndarray:
        off = self.rows * self.max_rank
        return self.data[off: off + self.rank * self.cols].reshape(self.rank, self.cols, order="F")

    def dense_view(self) -> np.ndarray:
        return self.data[: self.ld * self.cols].reshape(self.ld, self.cols, order="F")[: self.rows]

    def to_dense(self) -> np.ndarray:
        if self.type == DENSE:
            return np.array(self.dense_view())
        return self.U @ self.V

    def c(self) -> _Tile:
        return _Tile(self.rows, self.cols, self.rank, self.max_rank, self.ld, self.type, 0,
                     self.data.ctypes.data_as(C.c_void_p))

    def update_from(self, t: _Tile):
        self.rows, self.cols, self.rank, self.max_rank, self.ld = t.rows, t.cols, t.rank, t.max_rank, t.ld


def hcore_gemm(alpha, A: Tile, B: Tile, beta, Cm: Tile, params: Params, opA="N", opB="N", want_sigma=False):
    """HCore<double>::Gemm restated (src/api/HCore.cpp:22-344).  Updates ``Cm`` in place.
    Returns (info, sigma) where sigma are the core singular values (compressed C only)."""
    L = lib()
    ca, cb, cc = A.c(), B.c(), Cm.c()
    sig = np.zeros(max(Cm.rows, Cm.cols, 1) + A.rows + A.cols + B.cols + 8)
    ns = C.c_int64(0)
    info = L.orc_hcore_gemm(float(alpha), C.byref(ca), opA.encode(), C.byref(cb), opB.encode(), float(beta),
                            C.byref(cc), C.byref(params.c()), _ptr(sig), C.byref(ns))
    Cm.update_from(cc)
    if want_sigma:
        return info, sig[: ns.value].copy()
    return info


def compress(a: np.ndarray, params: Params):
    """CompressedTile compression constructor (src/operators/concrete/Compressed.cpp:75-146)."""
    a = np.asfortranarray(a, dtype=np.float64)
    m, n = a.shape
    mr = max(min(m, n) // 3, 1)
    t = Tile(m, n, 0, mr, m, COMPRESSED, np.zeros((m + n) * mr))
    ct = t.c()
    sig = np.zeros(min(m, n))
    info = lib().orc_compress(m, n, _ptr(a), m, C.byref(params.c()), C.byref(ct), _ptr(sig))
    if info != 0:
        raise RuntimeError(f"orc_compress info={info}")
    t.update_from(ct)
    return t, sig


def compress_dense_matrix(a: np.ndarray, accuracy: float):
    """compress_dense_matrix (src/helpers/MatrixHelpers.cpp:49-100) -> UV tile with max_rank = rank."""
    a = np.asfortranarray(a, dtype=np.float64)
    m, n = a.shape
    uv = np.zeros((m + n) * min(m, n))
    rk = lib().orc_compress_dense_matrix(m, n, _ptr(a), m, float(accuracy), _ptr(uv))
    u = uv[: m * rk].reshape(m, rk, order="F")
    v = uv[m * rk: m * rk + rk * n].reshape(rk, n, order="F")
    return Tile.compressed(u, v)


def latms_dense(m: int, n: int, iseed, cond: float = 1.0) -> np.ndarray:
    """generate_dense_matrix (src/helpers/MatrixHelpers.cpp:22-36): dlatms with D_i = 10^-i."""
    a = np.zeros((m, n), order="F")
    seed = np.asarray(iseed, dtype=np.int32)
    info = lib().orc_latms_dense(m, n, _ptr(a), m, _ptr(seed), float(cond))
    if info != 0:
        raise RuntimeError(f"dlatms info={info}")
    iseed[:] = seed.tolist()
    return a


def calculate_new_rank(sigma, accuracy, truncated_svd=False) -> int:
    s = np.ascontiguousarray(sigma, dtype=np.float64)
    return int(lib().orc_calculate_new_rank(int(truncated_svd), _ptr(s), len(s), float(accuracy)))


def batch_gemm(alpha, As, Bs, beta, Cs, params: Params, threads: int):
    """Throughput CPU baseline: independent updates over OpenMP threads.  Returns seconds."""
    n = len(As)
    TA = (_Tile * n)(*[a.c() for a in As])
    TB = (_Tile * n)(*[b.c() for b in Bs])
    TC = (_Tile * n)(*[c.c() for c in Cs])
    ranks = np.zeros(n, dtype=np.int64)
    sec = lib().orc_batch_gemm(n, float(alpha), TA, TB, float(beta), TC, C.byref(params.c()), int(threads),
                               _ptr(ranks))
    for i in range(n):
        Cs[i].update_from(TC[i])
    return sec, ranks


def tlr_gemm_omp(alpha, A, B, beta, Cg, params: Params, omp_threads: int, lapack_threads: int):
    """Tile GEMM C(i, j) += sum_k A(i, k) B(k, j) with OpenMP over the C tiles and `lapack_threads`
    MKL threads per tile (the reference's OpenMP example, omp_main.cpp:85-126).  A, B, Cg are lists
    of lists of tiles (ti x tk, tk x tj, ti x tj); the C tiles are updated.  Returns seconds."""
    ti, tk, tj = len(A), len(A[0]), len(B[0])
    TA = (_Tile * (ti * tk))(*[A[i][k].c() for i in range(ti) for k in range(tk)])
    TB = (_Tile * (tk * tj))(*[B[k][j].c() for k in range(tk) for j in range(tj)])
    TC = (_Tile * (ti * tj))(*[Cg[i][j].c() for i in range(ti) for j in range(tj)])
    sec = lib().orc_tlr_gemm_omp(ti, tj, tk, float(alpha), TA, TB, float(beta), TC, C.byref(params.c()),
                                 int(omp_threads), int(lapack_threads))
    for i in range(ti):
        for j in range(tj):
            Cg[i][j].update_from(TC[i * tj + j])
    return sec
