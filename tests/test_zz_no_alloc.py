"""GPU test of the allocation-free entry point hcorepp_gemm_noalloc_d, called through the C-ABI:
its results equal hcorepp_gemm_d's on the same inputs (same kernels, so the same ranks and
factors to rounding), on two calls in a row whose second is smaller (the scratch array keeps the
first call's contents), with the context's own arena never grown; an update beyond the scratch
size is refused with HCOREPP_ERR_CAPACITY and leaves C as it was."""
import ctypes as C

import numpy as np
import pytest

from helpers import rel_fro

pytestmark = pytest.mark.gpu


def _call(H, ctx, fn, A, B, Ct, acc):
    p = H.CompressionParameters(acc).c()
    ta, tb, tc = A.c(), B.c(), Ct.c()
    flops = C.c_uint64(0)
    if fn == "noalloc":
        st = H.lib().hcorepp_gemm_noalloc_d(ctx.h, 1.0, C.byref(ta), b"N", C.byref(tb), b"N", 1.0, C.byref(tc),
                                            C.byref(p), C.byref(flops))
    else:
        st = H.lib().hcorepp_gemm_d(ctx.h, 1.0, C.byref(ta), b"N", C.byref(tb), b"N", 1.0, C.byref(tc),
                                    C.byref(p), 0, C.byref(flops), None, None)
    if isinstance(Ct, H.CompressedTile):
        Ct.update_from(tc)
    return st, flops.value


def test_noalloc_gemm_matches_gemm():
    import hcorepp_amd as H
    from hcorepp_amd.workloads import lowrank_factors
    ctx = H.RunContext(0)   # allocation-free calls only: its arena must stay empty
    ref = H.RunContext(0)   # the ordinary entry point on the same inputs
    assert ctx.arena_bytes() == 0
    assert H.lib().hcorepp_gemm_noalloc_scratch_bytes() == 16 << 20
    rng = np.random.default_rng(2027)
    cases = [("CCC", 512, 90, 70, 120, 300), ("CCC", 128, 12, 9, 20, 60), ("DCC", 256, 0, 30, 40, 100),
             ("CCD", 200, 25, 20, 0, 0)]
    for kind, m, ra, rb, rc, mr in cases:
        print(f"[case] noalloc {kind} m={m}", flush=True)
        if kind[0] == "C":
            au, av = lowrank_factors(rng, m, m, ra, 0.9)
            A = H.CompressedTile(au, av)
        else:
            A = H.DenseTile(rng.standard_normal((m, m)))
        bu, bv = lowrank_factors(rng, m, m, rb, 0.9)
        B = H.CompressedTile(bu, bv)
        if kind[2] == "C":
            cu, cv = lowrank_factors(rng, m, m, rc, 0.9)
            C1, C2 = H.CompressedTile(cu, cv, max_rank=mr), H.CompressedTile(cu, cv, max_rank=mr)
        else:
            c = rng.standard_normal((m, m))
            C1, C2 = H.DenseTile(c), H.DenseTile(c)
        st2, f2 = _call(H, ctx, "noalloc", A, B, C2, 1e-8)
        assert st2 == 0, H.lib().hcorepp_last_error()
        assert ctx.arena_bytes() == 0
        st1, f1 = _call(H, ref, "gemm", A, B, C1, 1e-8)
        assert st1 == 0
        assert f1 == f2
        if kind[2] == "C":
            assert C1.rank == C2.rank
        assert rel_fro(C2.to_dense(), C1.to_dense()) < 1e-13
    # beyond the scratch array: refused, C untouched
    ctx2 = H.RunContext(0)
    au, av = lowrank_factors(rng, 1024, 1024, 300, 0.99)
    cu, cv = lowrank_factors(rng, 1024, 1024, 300, 0.99)
    A, Cb = H.CompressedTile(au, av), H.CompressedTile(cu, cv, max_rank=341)
    st, _ = _call(H, ctx2, "noalloc", A, A, Cb, 1e-8)
    assert H.STATUS[st] == "CAPACITY", st
    assert ctx2.arena_bytes() == 0
    assert Cb.rank == 300
    u, v = Cb.factors()
    assert np.array_equal(u, cu) and np.array_equal(v, cv)
    ctx2.close()
    ref.close()
    ctx.close()
