"""CPU checks of the drop-in boundary: the HIP library loads without a GPU and
exports every function include/hcorepp_amd.h declares (no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hcorepp_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hcorepp_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("hcorepp_gemm_d", "hcorepp_gemm_s", "hcorepp_ctx_create", "hcorepp_batch_create",
                     "hcorepp_batch_run", "hcorepp_gemm_pool_size_d"):
        assert required in names


def test_library_exports_every_declared_symbol():
    import hcorepp_amd as H
    lib = H.lib()  # raises if the library was not built
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(H.EXPORTED) == set(declared_functions())


def test_version_and_error_strings_without_gpu():
    import hcorepp_amd as H
    lib = H.lib()
    assert b"gfx950" in lib.hcorepp_version()
    assert isinstance(lib.hcorepp_last_error(), bytes)


def test_pool_size_query_is_host_only():
    import hcorepp_amd as H
    A = H._Tile(512, 512, 40, 40, 512, 0, H.COMPRESSED, None)
    B = H._Tile(512, 512, 30, 30, 512, 0, H.COMPRESSED, None)
    C = H._Tile(512, 512, 50, 170, 512, 0, H.COMPRESSED, None)
    p = H.CompressionParameters(1e-8).c()
    n = H.lib().hcorepp_gemm_pool_size_d(ctypes.byref(A), ctypes.byref(B), ctypes.byref(C), ctypes.byref(p))
    assert n > 512 * 90 * 2


def test_ops_fail_loudly_without_library(tmp_path, monkeypatch):
    import hcorepp_amd as H
    monkeypatch.setattr(H, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(H, "_lib", None)
    with pytest.raises(RuntimeError):
        H.lib()


DROPIN_SRC = os.path.join(ROOT, "tests", "cpp", "dropin_compile.cpp")


def build_dropin(out_dir):
    """Compile a caller written against the reference's C++ API (HCore<T>::Gemm,
    DenseTile, CompressedTile, CompressionParameters, RunContext, MemoryUnit) against the
    drop-in headers in include/hcorepp and link it to the HIP library."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = os.path.join(str(out_dir), "dropin_compile")
    libdir = os.path.join(ROOT, "hcorepp_amd")
    subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), DROPIN_SRC,
                    "-L", libdir, "-lhcorepp_amd", "-Wl,-rpath," + libdir, "-o", exe], check=True)
    return exe


def test_cpp_dropin_headers_compile_and_link(tmp_path):
    import subprocess
    import hcorepp_amd as H
    H.lib()
    exe = build_dropin(tmp_path)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    assert "built" in out


@pytest.mark.gpu
def test_cpp_dropin_gemm_runs(tmp_path):
    import subprocess
    exe = build_dropin(tmp_path)
    res = subprocess.run([exe, "run"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "max abs err" in res.stdout


def test_argument_checks_fail_before_any_device_work():
    """Entry points reject bad arguments with HCOREPP_ERR_INVALID and a message, before they
    touch the GPU (so these run on the CPU)."""
    import hcorepp_amd as H
    L = H.lib()
    t = H._Tile(4, 4, 1, 1, 4, 0, H.COMPRESSED, None)
    p = H.CompressionParameters(1e-8).c()
    flops = ctypes.c_uint64(0)
    st = L.hcorepp_gemm_d(None, 1.0, ctypes.byref(t), b"N", ctypes.byref(t), b"N", 1.0, ctypes.byref(t),
                          ctypes.byref(p), 0, ctypes.byref(flops), None, None)
    assert st == -1 and b"context" in L.hcorepp_last_error()
    assert L.hcorepp_potrf_d(None, ctypes.byref(t)) == -1
    assert L.hcorepp_trsm_batch_d(None, 1, ctypes.byref(t), ctypes.byref(t)) == -1
    assert L.hcorepp_syrk_batch_d(None, 1, ctypes.byref(t), ctypes.byref(t)) == -1
    assert L.hcorepp_transpose_batch_d(None, 1, ctypes.byref(t), ctypes.byref(t)) == -1
    assert L.hcorepp_compress_d(None, 4, 4, None, 4, ctypes.byref(p), ctypes.byref(t)) == -1
    assert L.hcorepp_kernel_gemm_d(None, b"N", b"N", 2, 2, 2, 1.0, None, 2, None, 2, 0.0, None, 2) == -1
    out = ctypes.c_void_p()
    assert L.hcorepp_batch_create(None, 1, ctypes.byref(t), ctypes.byref(t), ctypes.byref(t), 1.0, 1.0,
                                  ctypes.byref(p), 0, ctypes.byref(out)) == -1


@pytest.mark.gpu
def test_cpp_dropin_tlr_cholesky_matches_oracle(tmp_path):
    """A 3 x 3-tile TLR Cholesky (nb = 128, exp(-r/0.1) + 1e-2 I over points in the unit cube, acc
    1e-8) run by a C++ caller through the reference's Cholesky-flow API (HCore<double>::Potrf,
    Trsm, Syrk and Gemm with aCholesky = true, tiles U V^T with V stored nb x rank), against the
    oracle's tile sequence (oracle/cholesky.py)."""
    import subprocess
    import numpy as np
    from hcorepp_amd.workloads import TLRMatrix
    from oracle import cholesky as OC
    exe = build_dropin(tmp_path)
    nb, T, acc = 128, 3, 1e-8
    M = TLRMatrix(T * nb, nb, 3, 0.1, acc, 20261207, nugget=1e-2)
    inp, out = str(tmp_path / "chol_in.bin"), str(tmp_path / "chol_out.bin")
    with open(inp, "wb") as f:
        np.array([T, nb], dtype=np.int64).tofile(f)
        np.array([acc], dtype=np.float64).tofile(f)
        for k in range(T):
            np.asfortranarray(M.tile(k, k)[1]).reshape(-1, order="F").tofile(f)
        for i in range(T):
            for k in range(i):
                _, u, v, mr = M.tile(i, k)
                np.array([u.shape[1], mr], dtype=np.int64).tofile(f)
                u.reshape(-1, order="F").tofile(f)
                np.asfortranarray(v.T).reshape(-1, order="F").tofile(f)  # V stored nb x rank
    res = subprocess.run([exe, "chol", inp, out], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    raw = open(out, "rb").read()
    pos = 24
    Ld = np.zeros((T * nb, T * nb))
    ranks = {}
    for k in range(T):
        d = np.frombuffer(raw, dtype=np.float64, count=nb * nb, offset=pos).reshape(nb, nb, order="F")
        Ld[k * nb:(k + 1) * nb, k * nb:(k + 1) * nb] = np.tril(d)
        pos += 8 * nb * nb
    for i in range(T):
        for k in range(i):
            r, _ = np.frombuffer(raw, dtype=np.int64, count=2, offset=pos)
            pos += 16
            u = np.frombuffer(raw, dtype=np.float64, count=nb * r, offset=pos).reshape(nb, r, order="F")
            v = np.frombuffer(raw, dtype=np.float64, count=nb * r, offset=pos + 8 * nb * r).reshape(nb, r, order="F")
            pos += 16 * nb * r
            Ld[i * nb:(i + 1) * nb, k * nb:(k + 1) * nb] = u @ v.T
            ranks[(i, k)] = int(r)
    diag = [M.tile(k, k)[1].copy(order="F") for k in range(T)]
    off = {(i, k): M.tile(i, k)[1:] for i in range(T) for k in range(i)}
    Fo = OC.TLRCholesky(diag, off, acc).factor()
    Lo = Fo.dense_factor()
    ro = Fo.ranks()
    diff = sum(abs(ranks[k] - ro[k]) for k in ro)
    assert diff <= 1, (ranks, ro)
    assert np.linalg.norm(Ld - Lo) / np.linalg.norm(Lo) < (1e-11 if diff == 0 else 10 * acc)
    A = M.dense_block_range(0, M.n)
    be = np.linalg.norm(A - Ld @ Ld.T) / np.linalg.norm(A)
    be_o = np.linalg.norm(A - Lo @ Lo.T) / np.linalg.norm(A)
    assert be <= 1.5 * be_o + 1e-14, (be, be_o)
