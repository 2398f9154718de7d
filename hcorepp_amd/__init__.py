"""hcorepp_amd -- MI355X-native TLR tile GEMM + recompression (HCore++ hot path).

Python host-side mirror of the reference's operator interface, over the C-ABI
of ``include/hcorepp_amd.h`` (library ``hcorepp_amd/libhcorepp_amd.so``):

    reference (C++)                               here
    hcorepp::api::HCore<T>::Gemm                  HCore.gemm / gemm()
    hcorepp::operators::DenseTile<T>              DenseTile
    hcorepp::operators::CompressedTile<T>         CompressedTile
    hcorepp::operators::CompressionParameters     CompressionParameters
    hcorepp::kernels::RunContext (+ MemoryUnit)   RunContext (stream + device arena)

Device buffers are allocated through the library itself (hcorepp_malloc /
hcorepp_memcpy), so one HIP runtime serves the whole process.  The library must
be built (``hcorepp_amd.build.build()``); every compute entry point raises if it
is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libhcorepp_amd.so")

DENSE, COMPRESSED = 0, 1
GESVD, GESDD = 0, 1
PREC_FP64, PREC_MIXED = 0, 1

STATUS = {0: "OK", -1: "INVALID", -2: "UNSUPPORTED", -3: "CAPACITY", -4: "HIP", -5: "NUMERIC", -6: "NOMEM"}


class HCoreError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class _Tile(C.Structure):
    _fields_ = [("rows", C.c_int64), ("cols", C.c_int64), ("rank", C.c_int64), ("max_rank", C.c_int64),
                ("ld", C.c_int64), ("layout", C.c_int32), ("type", C.c_int32), ("data", C.c_void_p)]


class _CParams(C.Structure):
    _fields_ = [("accuracy", C.c_double), ("trmm", C.c_int32), ("ungqr", C.c_int32),
                ("truncated_svd", C.c_int32), ("svd_type", C.c_int32), ("fixed_rank", C.c_int64)]


# Every symbol include/hcorepp_amd.h declares.
EXPORTED = [
    "hcorepp_ctx_create", "hcorepp_ctx_destroy", "hcorepp_ctx_sync", "hcorepp_ctx_stream",
    "hcorepp_ctx_arena_bytes", "hcorepp_ctx_compute_units", "hcorepp_ctx_event_record", "hcorepp_ctx_event_elapsed", "hcorepp_malloc", "hcorepp_free", "hcorepp_memcpy",
    "hcorepp_memcpy_async", "hcorepp_memset", "hcorepp_kernel_gemm_d", "hcorepp_compress_d", "hcorepp_compress_s",
    "hcorepp_gemm_d", "hcorepp_gemm_s", "hcorepp_gemm_pool_size_d", "hcorepp_gemm_pool_size_s",
    "hcorepp_gemm_noalloc_d",
    "hcorepp_gemm_noalloc_scratch_bytes", "hcorepp_tile_potrf_d", "hcorepp_tile_potrf_s", "hcorepp_tile_trsm_d",
    "hcorepp_tile_trsm_s", "hcorepp_tile_syrk_d", "hcorepp_tile_syrk_s",
    "hcorepp_batch_create", "hcorepp_batch_run", "hcorepp_batch_capture_graph", "hcorepp_batch_launch_graph",
    "hcorepp_batch_rank_dev", "hcorepp_batch_ranks", "hcorepp_batch_set_ranks", "hcorepp_batch_status",
    "hcorepp_batch_workspace_bytes", "hcorepp_batch_phase_times", "hcorepp_batch_set_grid", "hcorepp_batch_get_grid",
    "hcorepp_batch_set_panel", "hcorepp_batch_get_panel", "hcorepp_autotune_lookup",
    "hcorepp_batch_set_rank_model", "hcorepp_batch_features", "hcorepp_batch_core_widths", "hcorepp_batch_learned_stats",
    "hcorepp_rank_model_features", "hcorepp_rank_model_params", "hcorepp_mlp_workspace_floats",
    "hcorepp_mlp_forward_f32", "hcorepp_mlp_grad_f32", "hcorepp_mlp_adam_f32", "hcorepp_mlp_pack_bf16",
    "hcorepp_mlp_predict_bf16", "hcorepp_potrf_d", "hcorepp_trsm_batch_d", "hcorepp_syrk_batch_d",
    "hcorepp_transpose_batch_d", "hcorepp_batch_destroy", "hcorepp_debug_workspace", "hcorepp_last_error",
    "hcorepp_version",
]

_lib = None
_live = weakref.WeakSet()   # device-owning objects, released before the HIP runtime shuts down
_shutting_down = False


@atexit.register
def _release_all():
    """Free device objects while the HIP runtime is still alive; later finalizers become no-ops."""
    global _shutting_down
    order = {"Batch": 0, "DeviceBuffer": 1, "RunContext": 2}
    for obj in sorted(list(_live), key=lambda o: order.get(type(o).__name__, 1)):
        try:
            obj._release()
        except Exception:
            pass
    _shutting_down = True


def lib():
    """Load the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with hcorepp_amd.build.build()")
    L = C.CDLL(LIB_PATH)
    P, I64, I32, D = C.c_void_p, C.c_int64, C.c_int32, C.c_double
    L.hcorepp_ctx_create.argtypes = [C.c_int, C.c_size_t, C.POINTER(P)]
    L.hcorepp_ctx_destroy.argtypes = [P]
    L.hcorepp_ctx_sync.argtypes = [P]
    L.hcorepp_ctx_stream.argtypes = [P]
    L.hcorepp_ctx_stream.restype = P
    L.hcorepp_ctx_arena_bytes.argtypes = [P]
    L.hcorepp_ctx_arena_bytes.restype = C.c_size_t
    L.hcorepp_ctx_compute_units.argtypes = [P]
    L.hcorepp_ctx_compute_units.restype = C.c_int
    L.hcorepp_memcpy.argtypes = [P, P, P, C.c_size_t]
    L.hcorepp_memset.argtypes = [P, P, C.c_int, C.c_size_t]
    L.hcorepp_memcpy_async.argtypes = [P, P, P, C.c_size_t]
    L.hcorepp_kernel_gemm_d.argtypes = [P, C.c_char, C.c_char, I64, I64, I64, D, P, I64, P, I64, D, P, I64]
    L.hcorepp_compress_d.argtypes = [P, I64, I64, P, I64, C.POINTER(_CParams), C.POINTER(_Tile)]
    L.hcorepp_compress_s.argtypes = [P, I64, I64, P, I64, C.POINTER(_CParams), C.POINTER(_Tile)]
    L.hcorepp_malloc.argtypes = [C.c_size_t, C.POINTER(P)]
    L.hcorepp_free.argtypes = [P]
    L.hcorepp_ctx_event_record.argtypes = [P, C.c_int]
    L.hcorepp_ctx_event_elapsed.argtypes = [P, C.c_int, C.c_int, C.POINTER(C.c_float)]
    L.hcorepp_gemm_d.argtypes = [P, D, C.POINTER(_Tile), C.c_char, C.POINTER(_Tile), C.c_char, D,
                                 C.POINTER(_Tile), C.POINTER(_CParams), C.c_int, C.POINTER(C.c_uint64), P,
                                 C.POINTER(I64)]
    L.hcorepp_gemm_s.argtypes = [P, C.c_float, C.POINTER(_Tile), C.c_char, C.POINTER(_Tile), C.c_char,
                                 C.c_float, C.POINTER(_Tile), C.POINTER(_CParams), C.c_int,
                                 C.POINTER(C.c_uint64), P, C.POINTER(I64)]
    L.hcorepp_gemm_noalloc_d.argtypes = [P, D, C.POINTER(_Tile), C.c_char, C.POINTER(_Tile), C.c_char, D,
                                         C.POINTER(_Tile), C.POINTER(_CParams), C.POINTER(C.c_uint64)]
    L.hcorepp_gemm_noalloc_scratch_bytes.restype = C.c_size_t
    for t, F in (("d", D), ("s", C.c_float)):
        getattr(L, f"hcorepp_tile_potrf_{t}").argtypes = [P, C.c_char, C.POINTER(_Tile)]
        getattr(L, f"hcorepp_tile_trsm_{t}").argtypes = [P, C.c_char, C.c_char, C.c_char, C.c_char, F,
                                                         C.POINTER(_Tile), C.POINTER(_Tile)]
        getattr(L, f"hcorepp_tile_syrk_{t}").argtypes = [P, F, C.POINTER(_Tile), C.c_char, C.c_char, F,
                                                         C.POINTER(_Tile)]
    for t in ("d", "s"):
        getattr(L, f"hcorepp_gemm_pool_size_{t}").argtypes = [C.POINTER(_Tile)] * 3 + [C.POINTER(_CParams)]
        getattr(L, f"hcorepp_gemm_pool_size_{t}").restype = C.c_size_t
    L.hcorepp_batch_create.argtypes = [P, I64, C.POINTER(_Tile), C.POINTER(_Tile), C.POINTER(_Tile), D, D,
                                       C.POINTER(_CParams), C.c_int, C.POINTER(P)]
    for name in ("hcorepp_batch_run", "hcorepp_batch_capture_graph", "hcorepp_batch_launch_graph",
                 "hcorepp_batch_destroy"):
        getattr(L, name).argtypes = [P]
    L.hcorepp_batch_rank_dev.argtypes = [P]
    L.hcorepp_batch_rank_dev.restype = P
    L.hcorepp_batch_ranks.argtypes = [P, P]
    L.hcorepp_batch_set_ranks.argtypes = [P, P]
    L.hcorepp_batch_status.argtypes = [P, C.POINTER(I64)]
    L.hcorepp_batch_phase_times.argtypes = [P, P]
    L.hcorepp_batch_set_grid.argtypes = [P, C.c_int]
    L.hcorepp_batch_set_rank_model.argtypes = [P, P, D, C.c_int]
    L.hcorepp_batch_features.argtypes = [P, P]
    L.hcorepp_batch_learned_stats.argtypes = [P, C.POINTER(I64), P]
    L.hcorepp_batch_core_widths.argtypes = [P, P]
    L.hcorepp_rank_model_params.restype = C.c_size_t
    L.hcorepp_mlp_workspace_floats.argtypes = [I64]
    L.hcorepp_mlp_workspace_floats.restype = C.c_size_t
    L.hcorepp_mlp_forward_f32.argtypes = [P, P, P, I64, P]
    L.hcorepp_mlp_grad_f32.argtypes = [P, P, P, P, I64, P, P, C.POINTER(C.c_float)]
    L.hcorepp_mlp_adam_f32.argtypes = [P, P, P, P, P, C.c_float, C.c_float, C.c_float, C.c_float, I64]
    L.hcorepp_mlp_pack_bf16.argtypes = [P, P, P]
    L.hcorepp_mlp_predict_bf16.argtypes = [P, P, P, I64, P]
    L.hcorepp_potrf_d.argtypes = [P, C.POINTER(_Tile)]
    L.hcorepp_trsm_batch_d.argtypes = [P, I64, C.POINTER(_Tile), C.POINTER(_Tile)]
    L.hcorepp_syrk_batch_d.argtypes = [P, I64, C.POINTER(_Tile), C.POINTER(_Tile)]
    L.hcorepp_transpose_batch_d.argtypes = [P, I64, C.POINTER(_Tile), C.POINTER(_Tile)]
    L.hcorepp_batch_get_grid.argtypes = [P]
    L.hcorepp_batch_get_grid.restype = C.c_int
    L.hcorepp_batch_set_panel.argtypes = [P, C.c_int]
    L.hcorepp_batch_get_panel.argtypes = [P]
    L.hcorepp_batch_get_panel.restype = C.c_int
    L.hcorepp_autotune_lookup.argtypes = [I64, D, D, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.hcorepp_autotune_lookup.restype = C.c_int
    L.hcorepp_batch_workspace_bytes.argtypes = [P]
    L.hcorepp_batch_workspace_bytes.restype = C.c_size_t
    L.hcorepp_debug_workspace.argtypes = [P, P, C.c_size_t, P]
    L.hcorepp_debug_workspace.restype = I64
    L.hcorepp_last_error.restype = C.c_char_p
    L.hcorepp_version.restype = C.c_char_p
    _lib = L
    return L


def _check(st: int):
    if st != 0:
        raise HCoreError(st, lib().hcorepp_last_error().decode())


def autotune_lookup(rows: int, mean_k: float, updates_per_cu: float):
    """The compiled-in autotune table's choice for a batch whose tallest QR has `rows` rows,
    whose mean stacked rank (ra + rc) is mean_k and which holds updates_per_cu updates per
    compute unit: (from_table, panel width, workgroups per CU).  Without a table entry for the
    class, the built-in rule's choice and from_table False."""
    pb, wg = C.c_int(0), C.c_int(0)
    hit = lib().hcorepp_autotune_lookup(int(rows), float(mean_k), float(updates_per_cu), C.byref(pb),
                                        C.byref(wg))
    return bool(hit), pb.value, wg.value


class DeviceBuffer:
    """Device allocation owned by Python (hcorepp_malloc / hcorepp_free)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        h = C.c_void_p()
        _check(lib().hcorepp_malloc(max(self.nbytes, 1), C.byref(h)))
        self.ptr = h.value
        _live.add(self)

    @staticmethod
    def from_numpy(a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = DeviceBuffer(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, offset_bytes: int = 0):
        a = np.ascontiguousarray(a)
        assert offset_bytes + a.nbytes <= self.nbytes
        _check(lib().hcorepp_memcpy(None, self.ptr + offset_bytes, a.ctypes.data_as(C.c_void_p), a.nbytes))

    def download(self, dtype, count: int | None = None, offset_bytes: int = 0) -> np.ndarray:
        itemsize = np.dtype(dtype).itemsize
        if count is None:
            count = (self.nbytes - offset_bytes) // itemsize
        out = np.empty(count, dtype=dtype)
        if count:
            _check(lib().hcorepp_memcpy(None, out.ctypes.data_as(C.c_void_p), self.ptr + offset_bytes,
                                        out.nbytes))
        return out

    def copy_from(self, other: "DeviceBuffer", nbytes: int | None = None):
        n = self.nbytes if nbytes is None else nbytes
        _check(lib().hcorepp_memcpy(None, self.ptr, other.ptr, n))

    def data_ptr(self) -> int:
        return self.ptr

    def _release(self):
        if getattr(self, "ptr", None) and not _shutting_down:
            lib().hcorepp_free(self.ptr)
        self.ptr = None

    free = _release

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass


class BufferView:
    """A byte range [offset, offset + nbytes) of a DeviceBuffer (shares its lifetime)."""

    def __init__(self, parent: DeviceBuffer, offset: int, nbytes: int):
        assert offset + nbytes <= parent.nbytes
        self.parent, self.offset, self.nbytes = parent, int(offset), int(nbytes)

    @property
    def ptr(self) -> int:
        return self.parent.ptr + self.offset

    def data_ptr(self) -> int:
        return self.ptr

    def upload(self, a: np.ndarray, offset_bytes: int = 0):
        self.parent.upload(a, self.offset + offset_bytes)

    def download(self, dtype, count: int | None = None, offset_bytes: int = 0) -> np.ndarray:
        itemsize = np.dtype(dtype).itemsize
        if count is None:
            count = (self.nbytes - offset_bytes) // itemsize
        return self.parent.download(dtype, count, self.offset + offset_bytes)


class TileArena:
    """One device allocation holding many tiles back to back (256-byte aligned), so that a
    whole set of tiles can be saved/restored with a single device-to-device copy."""

    def __init__(self, nbytes: int):
        self.buf = DeviceBuffer(nbytes)
        self.off = 0

    def take(self, nbytes: int) -> BufferView:
        off = (self.off + 255) // 256 * 256
        v = BufferView(self.buf, off, nbytes)
        self.off = off + nbytes
        return v

    @staticmethod
    def compressed_bytes(rows, cols, max_rank, dtype=np.float64):
        return (rows + cols) * max(int(max_rank), 1) * np.dtype(dtype).itemsize + 256


class CompressionParameters:
    """Mirror of hcorepp::operators::CompressionParameters
    (include/hcorepp/operators/helpers/CompressionParameters.hpp:44-46 defaults)."""

    def __init__(self, accuracy=1e-4, trmm=False, ungqr=True, truncated_svd=False, fixed_rank=0,
                 svd_type=GESDD):
        self.accuracy = float(accuracy)
        self.trmm, self.ungqr, self.truncated_svd = bool(trmm), bool(ungqr), bool(truncated_svd)
        self.fixed_rank, self.svd_type = int(fixed_rank), int(svd_type)

    def c(self) -> _CParams:
        return _CParams(self.accuracy, int(self.trmm), int(self.ungqr), int(self.truncated_svd),
                        self.svd_type, self.fixed_rank)


class RunContext:
    """One HIP stream + device arena (RunContext + pool MemoryUnit of the reference)."""

    def __init__(self, device: int = 0, arena_bytes: int = 0):
        self.device = device
        h = C.c_void_p()
        _check(lib().hcorepp_ctx_create(device, arena_bytes, C.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()  # batches on this context, released before it
        _live.add(self)

    def sync(self):
        _check(lib().hcorepp_ctx_sync(self.h))

    @property
    def stream_ptr(self) -> int:
        return lib().hcorepp_ctx_stream(self.h) or 0

    def arena_bytes(self) -> int:
        return lib().hcorepp_ctx_arena_bytes(self.h)

    def compute_units(self) -> int:
        return lib().hcorepp_ctx_compute_units(self.h)

    def record(self, slot: int):
        _check(lib().hcorepp_ctx_event_record(self.h, slot))

    def elapsed_ms(self, slot0: int, slot1: int) -> float:
        ms = C.c_float(0)
        _check(lib().hcorepp_ctx_event_elapsed(self.h, slot0, slot1, C.byref(ms)))
        return float(ms.value)

    def _release(self):
        for child in list(getattr(self, "_children", ())):
            child._release()
        if getattr(self, "h", None) and not _shutting_down:
            lib().hcorepp_ctx_destroy(self.h)
        self.h = None

    close = _release

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass


_DTYPES = {"d": np.float64, "s": np.float32}


class _TileBase:
    dtype = np.float64

    def c(self) -> _Tile:
        return _Tile(self.rows, self.cols, self.rank, self.max_rank, self.ld, 0, self.type, self.buf.data_ptr())


class DenseTile(_TileBase):
    """Dense tile in HBM (DenseTile<T>, reference src/operators/concrete/Dense.cpp:19-29)."""
    type = DENSE

    def __init__(self, a, device=0, dtype=np.float64):
        a = np.asfortranarray(np.asarray(a, dtype=dtype))
        self.rows, self.cols = a.shape
        self.ld = max(self.rows, 1)
        self.rank = self.max_rank = min(self.rows, self.cols)
        self.dtype = dtype
        self.buf = DeviceBuffer.from_numpy(a.reshape(-1, order="F"))

    def to_numpy(self) -> np.ndarray:
        h = self.buf.download(self.dtype)
        return h[: self.ld * self.cols].reshape(self.ld, self.cols, order="F")[: self.rows].copy()

    to_dense = to_numpy


class CompressedTile(_TileBase):
    """Compressed tile U*V in HBM with the reference layout
    [U rows x max_rank | V max_rank x cols] (src/operators/concrete/Compressed.cpp:19-47,175-185)."""
    type = COMPRESSED

    def __init__(self, u, v, max_rank=None, device=0, dtype=np.float64, arena: "TileArena | None" = None):
        u = np.asarray(u, dtype=dtype)
        v = np.asarray(v, dtype=dtype)
        self.rows, r = u.shape
        r2, self.cols = v.shape
        assert r == r2, "U and V ranks differ"
        self.rank = r
        self.max_rank = r if max_rank is None else int(max_rank)
        assert self.max_rank >= r
        self.ld = self.rows
        self.dtype = dtype
        host = np.zeros((self.rows + self.cols) * max(self.max_rank, 1), dtype=dtype)
        host[: self.rows * r] = u.reshape(-1, order="F")
        off = self.rows * self.max_rank
        host[off: off + r * self.cols] = v.reshape(-1, order="F")
        if arena is None:
            self.buf = DeviceBuffer.from_numpy(host)
        else:
            self.buf = arena.take(host.nbytes)
            self.buf.upload(host)

    @classmethod
    def compress(cls, ctx: "RunContext", a, params: "CompressionParameters | None" = None, max_rank=None,
                 dtype=np.float64) -> "CompressedTile":
        """Compression constructor on the device (CompressedTile(rows, cols, data, ld, params, ...),
        reference src/operators/concrete/Compressed.cpp:75-146): max_rank defaults to the
        reference's max(min(rows, cols) / 3, 1)."""
        a = np.asfortranarray(np.asarray(a, dtype=dtype))
        m, n = a.shape
        mr = max(min(m, n) // 3, 1) if max_rank is None else int(max_rank)
        t = cls(np.zeros((m, 0), dtype=dtype), np.zeros((0, n), dtype=dtype), max_rank=mr, dtype=dtype)
        src = DeviceBuffer.from_numpy(a.reshape(-1, order="F"))
        d = t.c()
        p = (params or CompressionParameters(1e-9)).c()
        f = lib().hcorepp_compress_s if dtype == np.float32 else lib().hcorepp_compress_d
        _check(f(ctx.h, m, n, src.ptr, m, C.byref(p), C.byref(d)))
        t.update_from(d)
        src.free()
        return t

    @classmethod
    def from_cholesky_factors(cls, u, v_nr, max_rank=None, dtype=np.float64) -> "CompressedTile":
        """A tile in the Cholesky-flow convention (reference src/operators/concrete/Compressed.cpp
        aCholesky branches): A = U V^T with V stored cols x rank (ld cols) in the V slot."""
        u = np.asarray(u, dtype=dtype)
        v_nr = np.asarray(v_nr, dtype=dtype)
        m, r = u.shape
        n, r2 = v_nr.shape
        assert r == r2
        t = cls(np.zeros((m, 0), dtype=dtype), np.zeros((0, n), dtype=dtype),
                max_rank=r if max_rank is None else max_rank, dtype=dtype)
        mr = t.max_rank
        host = np.zeros((m + n) * max(mr, 1), dtype=dtype)
        host[: m * r] = u.reshape(-1, order="F")
        host[m * mr: m * mr + n * r] = v_nr.reshape(-1, order="F")
        t.buf.upload(host)
        t.rank = r
        return t

    def cholesky_factors(self):
        """(U, V) of a Cholesky-flow tile: U rows x rank, V cols x rank (A = U V^T)."""
        h = self.buf.download(self.dtype)
        u = h[: self.rows * self.rank].reshape(self.rows, self.rank, order="F").copy()
        off = self.rows * self.max_rank
        v = h[off: off + self.cols * self.rank].reshape(self.cols, self.rank, order="F").copy()
        return u, v

    def factors(self):
        h = self.buf.download(self.dtype)
        u = h[: self.rows * self.rank].reshape(self.rows, self.rank, order="F").copy()
        off = self.rows * self.max_rank
        v = h[off: off + self.rank * self.cols].reshape(self.rank, self.cols, order="F").copy()
        return u, v

    def to_dense(self) -> np.ndarray:
        u, v = self.factors()
        return u.astype(np.float64) @ v.astype(np.float64)

    def update_from(self, t: _Tile):
        self.rows, self.cols, self.rank, self.max_rank, self.ld = t.rows, t.cols, t.rank, t.max_rank, t.ld


class HCore:
    """hcorepp::api::HCore<T> (reference include/hcorepp/api/HCore.hpp:23-128)."""

    @staticmethod
    def gemm(alpha, A, B, beta, Ct, ctx: RunContext, params: CompressionParameters | None = None,
             opA="N", opB="N", cholesky=False, want_sigma=False):
        """C = alpha*op(A)*op(B) + beta*C (reference src/api/HCore.cpp:22-344).  Returns the
        reference flop count, and the sorted core singular values if ``want_sigma``."""
        L = lib()
        p = (params or CompressionParameters(1e-9)).c()
        ta, tb, tc = A.c(), B.c(), Ct.c()
        flops = C.c_uint64(0)
        nsig = C.c_int64(0)
        sig = np.zeros(max(Ct.rows, Ct.cols, 1) + 1024) if want_sigma else None
        sp = sig.ctypes.data_as(C.c_void_p) if want_sigma else None
        f = L.hcorepp_gemm_s if Ct.dtype == np.float32 else L.hcorepp_gemm_d
        st = f(ctx.h, alpha, C.byref(ta), opA.encode(), C.byref(tb), opB.encode(), beta, C.byref(tc),
               C.byref(p), int(bool(cholesky)), C.byref(flops), sp, C.byref(nsig))
        _check(st)
        if isinstance(Ct, CompressedTile):
            Ct.update_from(tc)
        if want_sigma:
            return flops.value, sig[: nsig.value].copy()
        return flops.value


    @staticmethod
    def potrf(A: "DenseTile", uplo: str, ctx: RunContext):
        """HCore<T>::Potrf (reference src/api/HCore.cpp:586-622): LAPACK potrf of the dense tile A
        in place, uplo 'L' or 'U' (the other strict triangle is left as it was)."""
        t = A.c()
        f = lib().hcorepp_tile_potrf_s if A.dtype == np.float32 else lib().hcorepp_tile_potrf_d
        _check(f(ctx.h, uplo.encode(), C.byref(t)))

    @staticmethod
    def trsm(side: str, uplo: str, trans: str, diag: str, alpha, A: "DenseTile", B: "CompressedTile",
             ctx: RunContext):
        """HCore<T>::Trsm (reference src/api/HCore.cpp:624-647): BLAS trsm on the V factor of the
        compressed B (Cholesky-flow convention: V stored cols x rank)."""
        ta, tb = A.c(), B.c()
        f = lib().hcorepp_tile_trsm_s if B.dtype == np.float32 else lib().hcorepp_tile_trsm_d
        _check(f(ctx.h, side.encode(), uplo.encode(), trans.encode(), diag.encode(), alpha, C.byref(ta),
                 C.byref(tb)))

    @staticmethod
    def syrk(alpha, A, opA: str, uplo: str, beta, Ct: "DenseTile", ctx: RunContext):
        """HCore<T>::Syrk (reference src/api/HCore.cpp:484-584) into the dense tile Ct."""
        ta, tc = A.c(), Ct.c()
        f = lib().hcorepp_tile_syrk_s if Ct.dtype == np.float32 else lib().hcorepp_tile_syrk_d
        _check(f(ctx.h, alpha, C.byref(ta), opA.encode(), uplo.encode(), beta, C.byref(tc)))


gemm = HCore.gemm


class Batch:
    """Batched variable-rank form: n independent updates C_i = alpha*A_i*B_i + beta*C_i
    run by one persistent launch (optionally replayed from a captured hipGraph)."""

    def __init__(self, ctx: RunContext, As, Bs, Cs, alpha=1.0, beta=1.0,
                 params: CompressionParameters | None = None, precision=PREC_FP64):
        L = lib()
        n = len(Cs)
        self.ctx, self.n = ctx, n
        self._tiles = (As, Bs, Cs)  # keep device buffers alive
        TA = (_Tile * n)(*[a.c() for a in As])
        TB = (_Tile * n)(*[b.c() for b in Bs])
        TC = (_Tile * n)(*[c.c() for c in Cs])
        p = (params or CompressionParameters(1e-9)).c()
        h = C.c_void_p()
        _check(L.hcorepp_batch_create(ctx.h, n, TA, TB, TC, float(alpha), float(beta), C.byref(p),
                                      int(precision), C.byref(h)))
        self.h = h
        _live.add(self)
        ctx._children.add(self)

    def run(self):
        _check(lib().hcorepp_batch_run(self.h))

    def capture(self):
        _check(lib().hcorepp_batch_capture_graph(self.h))

    def launch_graph(self):
        _check(lib().hcorepp_batch_launch_graph(self.h))

    def ranks(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.int64)
        _check(lib().hcorepp_batch_ranks(self.h, out.ctypes.data_as(C.c_void_p)))
        return out

    def set_ranks(self, ranks):
        r = np.ascontiguousarray(ranks, dtype=np.int64)
        _check(lib().hcorepp_batch_set_ranks(self.h, r.ctypes.data_as(C.c_void_p)))

    def status(self):
        sw = C.c_int64(0)
        st = lib().hcorepp_batch_status(self.h, C.byref(sw))
        return st, sw.value

    def workspace_bytes(self) -> int:
        return lib().hcorepp_batch_workspace_bytes(self.h)

    def set_grid(self, workgroups: int):
        """Number of workgroups of the persistent kernel (drops a captured graph)."""
        _check(lib().hcorepp_batch_set_grid(self.h, int(workgroups)))

    def grid(self) -> int:
        return lib().hcorepp_batch_get_grid(self.h)

    def set_panel(self, panel: int):
        """Householder panel width of the batch's QRs: 4, 8 or 16 within the limits for its
        tallest QR (drops a captured graph)."""
        _check(lib().hcorepp_batch_set_panel(self.h, int(panel)))

    def panel(self) -> int:
        return lib().hcorepp_batch_get_panel(self.h)

    def set_rank_model(self, model_bf16, tol: float = 0.1, margin: int = 40):
        """Learned core sizing with a BF16 parameter buffer (DeviceBuffer).  model_bf16=None with
        tol > 0 applies the tolerance directly to the pivoted R factor; None with tol=0 switches
        both off.  Rewrites the device descriptors; a captured graph stays valid."""
        ptr = model_bf16.ptr if model_bf16 is not None else None
        self._rank_model = model_bf16
        _check(lib().hcorepp_batch_set_rank_model(self.h, ptr, float(tol), int(margin)))

    def features(self) -> np.ndarray:
        out = np.zeros((self.n, lib().hcorepp_rank_model_features()), dtype=np.float32)
        _check(lib().hcorepp_batch_features(self.h, out.ctypes.data_as(C.c_void_p)))
        return out

    def learned_stats(self):
        """(fallbacks of the last pass, predicted rank per update)."""
        fb = C.c_int64(0)
        pred = np.zeros(self.n, dtype=np.int64)
        _check(lib().hcorepp_batch_learned_stats(self.h, C.byref(fb), pred.ctypes.data_as(C.c_void_p)))
        return int(fb.value), pred

    def core_widths(self) -> np.ndarray:
        """Core width each update of the last pass handed to its SVD."""
        out = np.zeros(self.n, dtype=np.int64)
        _check(lib().hcorepp_batch_core_widths(self.h, out.ctypes.data_as(C.c_void_p)))
        return out

    def phase_times(self) -> np.ndarray:
        """Per-update phase timestamps and sub-phase durations of the last pass (n x 16,
        100 MHz ticks); needs a library built with HCAMD_PHASE_TIMING."""
        out = np.zeros((self.n, 16), dtype=np.int64)
        _check(lib().hcorepp_batch_phase_times(self.h, out.ctypes.data_as(C.c_void_p)))
        return out

    def sync_ranks_to_tiles(self):
        rk = self.ranks()
        for c, r in zip(self._tiles[2], rk):
            c.rank = int(r)
        return rk

    def _release(self):
        if getattr(self, "h", None) and not _shutting_down:
            lib().hcorepp_batch_destroy(self.h)
        self.h = None

    close = _release

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass
