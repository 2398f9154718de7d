"""The batched kernel's launch shape (Householder panel width, workgroups per CU): the default
comes from the compiled-in autotune table for the batch's class, and every valid panel width
gives the same ranks and the same factors as the others (to the batched Jacobi's stopping level)."""
import numpy as np
import pytest

from helpers import rel_fro

pytestmark = pytest.mark.gpu


def _batch(H, ctx, n, nb, lo, hi, seed, max_rank):
    from hcorepp_amd.workloads import batch_ccc_pool
    ups = batch_ccc_pool(n, nb, seed, lo, hi, 0.85, pool=12)
    As = [H.CompressedTile(u["au"], u["av"]) for u in ups]
    Bs = [H.CompressedTile(u["bu"], u["bv"]) for u in ups]
    Cs = [H.CompressedTile(u["cu"], u["cv"], max_rank=max_rank) for u in ups]
    mean_k = float(np.mean([u["au"].shape[1] + u["cu"].shape[1] for u in ups]))
    return ups, As, Bs, Cs, mean_k


def test_default_shape_from_table(ctx):
    """A new batch takes the table's panel width and residency for its class."""
    import hcorepp_amd as H
    for nb, lo, hi in [(128, 4, 24), (512, 8, 192), (1024, 8, 96)]:
        print(f"[case] default shape nb={nb} ranks {lo}..{hi}", flush=True)
        ups, As, Bs, Cs, mean_k = _batch(H, ctx, 8, nb, lo, hi, 3, min(nb, 2 * hi))
        b = H.Batch(ctx, As, Bs, Cs, 1.0, 1.0, H.CompressionParameters(1e-8))
        hit, pb, wg = H.autotune_lookup(nb, mean_k, len(ups) / ctx.compute_units())
        assert b.panel() == pb
        assert b.grid() == min(len(ups), wg * ctx.compute_units())
        b.close()


@pytest.mark.parametrize("nb,lo,hi", [(256, 8, 96), (512, 8, 192)])
def test_panel_widths_agree(ctx, nb, lo, hi):
    """Panel widths 4, 8 and 16 (and both residencies) give the same ranks, and factors equal to
    the Jacobi's stopping level, on one batch; a captured graph is dropped by the change and
    captured again."""
    import hcorepp_amd as H
    n = 48
    ups, As, Bs, Cs, _ = _batch(H, ctx, n, nb, lo, hi, 11, min(nb, 2 * hi))
    r0 = np.array([c.rank for c in Cs], dtype=np.int64)
    img0 = [c.buf.download(np.float64) for c in Cs]  # the C tiles' input images
    b = H.Batch(ctx, As, Bs, Cs, 1.0, 1.0, H.CompressionParameters(1e-8))
    results = {}
    for pb in (16, 8, 4):
        for wg in (1, 2):
            print(f"[case] nb={nb} panel {pb} x {wg} WG/CU", flush=True)
            for c, img, r in zip(Cs, img0, r0):
                c.buf.upload(img)
                c.rank = int(r)
            b.set_ranks(r0)
            b.set_panel(pb)
            b.set_grid(wg * ctx.compute_units())
            assert b.panel() == pb
            with pytest.raises(H.HCoreError):
                b.launch_graph()  # the change dropped the graph
            b.capture()
            b.launch_graph()
            rk = b.sync_ranks_to_tiles()
            assert b.status()[0] == 0
            results[(pb, wg)] = (rk.copy(), [c.to_dense() for c in Cs])
    rk_ref, d_ref = results[(16, 2)]
    for key, (rk, d) in results.items():
        assert np.array_equal(rk, rk_ref), key
        # the batched Jacobi stops once a sweep's cosines are <= 1e-5, leaving U orthonormal to
        # about 1e-10; QRs that differ in rounding can stop a sweep apart, so the results agree
        # to that level, not to the QR's roundoff
        diff = max(rel_fro(a, r) for a, r in zip(d, d_ref))
        print(f"panel {key}: max relative difference {diff:.2e}", flush=True)
        assert diff < 1e-10, key
    b.close()


def test_panel_width_limits(ctx):
    """Widths outside {4, 8, 16}, or too wide for the tallest QR, are rejected."""
    import hcorepp_amd as H
    ups, As, Bs, Cs, _ = _batch(H, ctx, 4, 1024, 8, 32, 5, 64)
    b = H.Batch(ctx, As, Bs, Cs, 1.0, 1.0, H.CompressionParameters(1e-8))
    for bad in (0, 2, 12, 16, 32):
        with pytest.raises(H.HCoreError):
            b.set_panel(bad)
    for ok in (8, 4):
        b.set_panel(ok)
        assert b.panel() == ok
    b.close()
