"""The batched form at the benchmark's full size: 4096 CCC updates, nb = 512, ranks log-uniform in
[8, 192], decay 0.85, max_rank 384, accuracy 1e-8 (absolute), one graph-captured persistent launch
per form.  Checked through properties that do not need the oracle at this size: a clean status,
the rule form's ranks equal to the exact form's on every update, and on a sample of updates the
rank equal to the number of singular values of the dense FP64 result above the accuracy (away from
the threshold), the truncation error within its bound, orthonormal U columns, and the rule form's
result equal to the exact form's within the accuracy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NB, N, LO, HI, DECAY, MAXR, ACC = 512, 4096, 8, 192, 0.85, 384, 1e-8


def test_full_size_batch_properties(ctx):
    import hcorepp_amd as H
    from hcorepp_amd.workloads import batch_ccc_pool
    print("[case] full-size batch: generating 4096 updates", flush=True)
    ups = batch_ccc_pool(N, NB, seed=777, lo=LO, hi=HI, decay=DECAY, pool=48)
    As = [H.CompressedTile(u["au"], u["av"]) for u in ups]
    Bs = [H.CompressedTile(u["bu"], u["bv"]) for u in ups]
    arena = H.TileArena(N * H.TileArena.compressed_bytes(NB, NB, MAXR))
    Cs = [H.CompressedTile(u["cu"], u["cv"], max_rank=MAXR, arena=arena) for u in ups]
    c0 = H.DeviceBuffer(arena.off)
    c0.copy_from(arena.buf, arena.off)
    r0 = np.array([c.rank for c in Cs], dtype=np.int64)
    b = H.Batch(ctx, As, Bs, Cs, 1.0, 1.0, H.CompressionParameters(ACC), H.PREC_FP64)
    b.capture()
    sample = list(range(0, N, N // 48))
    out = {}
    for form, tol in (("exact", 0.0), ("rule", 0.1)):
        print(f"[case] full-size batch: {form} form", flush=True)
        H.lib().hcorepp_memcpy_async(ctx.h, arena.buf.ptr, c0.ptr, arena.off)
        b.set_ranks(r0)
        b.set_rank_model(None, tol=tol, margin=0)
        b.launch_graph()
        st, sweeps = b.status()
        assert st == 0, (form, st)
        rk = b.ranks().copy()
        facs = []
        for i in sample:
            Cs[i].rank = int(rk[i])
            facs.append(Cs[i].factors())
        out[form] = (rk, facs)
    rk_e, fac_e = out["exact"]
    rk_r, fac_r = out["rule"]
    assert np.array_equal(rk_e, rk_r)
    assert rk_e.min() >= 1 and rk_e.max() <= MAXR
    for j, i in enumerate(sample):
        u = ups[i]
        dense = u["cu"] @ u["cv"] + u["au"] @ (u["av"] @ u["bu"]) @ u["bv"]
        s = np.linalg.svd(dense, compute_uv=False)
        ue, ve = fac_e[j]
        ur, vr = fac_r[j]
        ce, cr = ue @ ve, ur @ vr
        # rank: singular values of the dense result above the absolute accuracy (unless one lies
        # within 1e-6 relative of it, where FP64 roundoff may decide either way)
        near = np.any(np.abs(s - ACC) <= 1e-6 * ACC)
        if not near:
            assert rk_e[i] == int(np.sum(s >= ACC)), (i, rk_e[i], int(np.sum(s >= ACC)))
        # truncation error: the dropped singular values, plus rounding at the result's scale
        bound = np.sqrt(np.sum(s[rk_e[i]:] ** 2)) + 1e-12 * s[0]
        assert np.linalg.norm(ce - dense) <= 1.01 * bound + 1e-14, i
        assert np.linalg.norm(ue.T @ ue - np.eye(ue.shape[1])) <= 1e-12 * NB, i
        assert np.linalg.norm(cr - ce) <= ACC * max(1.0, np.sqrt(rk_e[i])), i
    b.close()
