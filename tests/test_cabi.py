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
