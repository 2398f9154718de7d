"""Offline autotuner of the batched persistent kernel's launch shape.

A batch's class is its tallest stacked-factor QR (rows = tile size nb), the mean stacked rank
K = ra + rc of its updates, and its size in updates per CU.  For representative batches of each
class the tuner times one persistent, graph-captured launch under each candidate shape:

  * panel width of the blocked Householder QRs: 4, 8 or 16 columns (the widths the kernel is
    instantiated for, within the register/LDS limits for the class's rows),
  * residency: 1 or 2 workgroups per CU (the grid of the persistent kernel; two 256-thread
    workgroups with 256 VGPRs and 78 KB of LDS each fill a CU).  With few updates per
    workgroup the launch ends with its largest updates, which run faster with a CU to
    themselves; with many, two workgroups per CU hide each other's barriers and latencies.

The workgroup size (256 threads, 4 waves) is fixed: the QR register tiles, the wave-cooperative
Jacobi tiles and the reductions are written for four waves.

The result goes to ``hcorepp_amd/autotune.json``: every timed point (``sweep``) and the chosen
shape per class (``table``), which the build compiles into the library as the default of
``hcorepp_batch_create``.  A candidate replaces the built-in rule's shape (widest panel that
fits, two workgroups per CU) only when it is faster by more than ``--margin`` (default 2 %), so
timing noise does not flip the table.  Every candidate's ranks and factors are compared with the
first candidate's on the same batch and the largest differences are recorded.

Run on the GPU (writes the JSON; rebuild afterwards to compile it in)::

    python -m hcorepp_amd.autotune --out hcorepp_amd/autotune.json
"""
from __future__ import annotations

import argparse
import datetime
import json
import math
import os
import sys
import time

import numpy as np

# K-classes per tile size: (nb, lo, hi), ranks of A, B, C drawn log-uniform in [lo, hi].  The
# nb = 512 middle class is the distribution of the 4096-update nb = 512 benchmark (ranks 8..192).
K_CLASSES = [
    (128, 4, 16), (128, 8, 48), (128, 24, 64),
    (256, 4, 32), (256, 8, 96), (256, 48, 128),
    (512, 4, 32), (512, 8, 192), (512, 64, 256),
    (1024, 4, 32), (1024, 8, 192), (1024, 64, 384),
]
# Batch sizes per K-class, in updates per CU: one per CU (the trailing-update batches of a TLR
# Cholesky step are of this order), a small batch (4 per CU: two per workgroup at two workgroups
# per CU) and a large one (16 per CU; 8 for the largest class, whose workspace would otherwise
# pass 100 GB).
SIZES_PER_CU = (1, 4, 16)
LARGEST_CLASS_PER_CU = 8
PANELS = (16, 8, 4)
RESIDENCY = (1, 2)
ACCURACY, DECAY = 1e-8, 0.85


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _time_batch(H, ctx, batch, arena, c0, r0, steps):
    """Median event time (ms) of `steps` graph launches of batch, C restored before each."""
    L = H.lib()

    def restore():
        L.hcorepp_memcpy_async(ctx.h, arena.buf.ptr, c0.ptr, arena.off)
        batch.set_ranks(r0)

    batch.capture()
    restore()
    batch.launch_graph()
    ctx.sync()
    st, _ = batch.status()
    if st != 0:
        raise RuntimeError(f"numerical status {st}")
    ms = []
    for _ in range(steps):
        restore()
        ctx.record(0)
        batch.launch_graph()
        ctx.record(1)
        ms.append(ctx.elapsed_ms(0, 1))
    ctx.sync()
    return float(np.median(ms))


def _results(H, batch, cs, nchk):
    rk = np.asarray(batch.ranks())
    fac = []
    for i in range(nchk):
        cs[i].rank = int(rk[i])
        u, v = cs[i].factors()
        fac.append(u @ v)
    return rk, fac


def sweep_class(H, ctx, nb, n, lo, hi, seed, steps, num_cus):
    from hcorepp_amd.workloads import batch_ccc_pool
    t0 = time.time()
    ups = batch_ccc_pool(n, nb, seed, lo, hi, DECAY, pool=48)
    max_rank = min(nb, 2 * hi)
    params = H.CompressionParameters(ACCURACY)
    As = [H.CompressedTile(u["au"], u["av"]) for u in ups]
    Bs = [H.CompressedTile(u["bu"], u["bv"]) for u in ups]
    arena = H.TileArena(n * H.TileArena.compressed_bytes(nb, nb, max_rank))
    Cs = [H.CompressedTile(u["cu"], u["cv"], max_rank=max_rank, arena=arena) for u in ups]
    c0 = H.DeviceBuffer(arena.off)
    c0.copy_from(arena.buf, arena.off)
    r0 = np.array([c.rank for c in Cs], dtype=np.int64)
    mean_k = float(np.mean([u["au"].shape[1] + u["cu"].shape[1] for u in ups]))
    batch = H.Batch(ctx, As, Bs, Cs, 1.0, 1.0, params, H.PREC_FP64)
    log(f"[autotune] nb={nb} n={n} ({n / num_cus:.0f} per CU) ranks {lo}..{hi}: mean K {mean_k:.1f}, built in {time.time() - t0:.1f}s")
    nchk = min(64, n)
    points = []
    ref = None
    builtin = None  # the built-in rule: the widest panel that fits, 2 workgroups per CU
    for pb in PANELS:
        try:
            batch.set_panel(pb)
        except H.HCoreError:
            continue
        builtin = builtin or [pb, 2]
        for wg in RESIDENCY:
            batch.set_grid(wg * num_cus)
            ms = _time_batch(H, ctx, batch, arena, c0, r0, steps)
            rk, fac = _results(H, batch, Cs, nchk)
            if ref is None:
                ref = (rk, fac)
            rank_diff = int(np.sum(rk != ref[0]))
            rel = max(float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)) for a, b in zip(fac, ref[1]))
            points.append(dict(panel=pb, workgroups_per_cu=wg, ms=round(ms, 3), rank_diffs=rank_diff,
                               max_rel_diff=rel))
            log(f"[autotune]   panel {pb:2d}, {wg} WG/CU: {ms:8.2f} ms  (rank diffs {rank_diff}, "
                f"max rel diff {rel:.1e})")
    batch.close()
    del c0, arena, Cs, As, Bs
    return dict(nb=nb, updates=n, per_cu=round(n / num_cus, 3), ranks=[lo, hi], mean_k=round(mean_k, 2),
                builtin=builtin, points=points)


def choose(cls, margin):
    """The class's shape: the fastest point, unless the built-in shape is within `margin` of it."""
    pts = cls["points"]
    best = min(pts, key=lambda p: p["ms"])
    pb0, wg0 = cls["builtin"]
    base = [p for p in pts if p["panel"] == pb0 and p["workgroups_per_cu"] == wg0]
    if base and base[0]["ms"] <= best["ms"] * (1.0 + margin):
        best = base[0]
    return best


def _split(values):
    """Upper bounds between sorted class values: geometric means of neighbours, None for the last."""
    return [round(math.sqrt(values[i] * values[i + 1]), 2) if i + 1 < len(values) else None
            for i in range(len(values))]


def build_table(classes, margin):
    """Table rows sorted by (rows, K, updates per CU).  Per tile size, the K bounds lie at the
    geometric means of adjacent K-classes' mean K; per K-class, the size bounds at the geometric
    means of adjacent batch sizes (updates per CU); the last of each is unbounded."""
    rows = []
    for nb in sorted({c["nb"] for c in classes}):
        kcls = sorted({tuple(c["ranks"]) for c in classes if c["nb"] == nb},
                      key=lambda r: min(c["mean_k"] for c in classes if c["nb"] == nb and tuple(c["ranks"]) == r))
        kmean = [min(c["mean_k"] for c in classes if c["nb"] == nb and tuple(c["ranks"]) == r) for r in kcls]
        for rk, kb in zip(kcls, _split(kmean)):
            cl = sorted([c for c in classes if c["nb"] == nb and tuple(c["ranks"]) == rk], key=lambda c: c["per_cu"])
            for c, sb in zip(cl, _split([c["per_cu"] for c in cl])):
                b = choose(c, margin)
                base = [p["ms"] for p in c["points"]
                        if [p["panel"], p["workgroups_per_cu"]] == list(c["builtin"])]
                rows.append(dict(max_rows=nb, max_mean_k=kb, max_per_cu=sb, panel=b["panel"],
                                 workgroups_per_cu=b["workgroups_per_cu"], ms=b["ms"],
                                 builtin_ms=base[0] if base else None))
    return rows


def _args_without_out(argv):
    """The sweep's arguments minus --out (where the file was written says nothing about it)."""
    out, skip = [], False
    for a in argv:
        if skip:
            skip = False
        elif a == "--out":
            skip = True
        elif not a.startswith("--out="):
            out.append(a)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune.json"))
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--margin", type=float, default=0.02)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--nbs", default="", help="comma-separated tile sizes to sweep (default: all classes)")
    args = ap.parse_args(argv)
    import hcorepp_amd as H

    ctx = H.RunContext(0)
    num_cus = ctx.compute_units()
    want = {int(x) for x in args.nbs.split(",") if x}
    classes = []
    t0 = time.time()
    for i, (nb, lo, hi) in enumerate(K_CLASSES):
        if want and nb not in want:
            continue
        for per_cu in SIZES_PER_CU:
            if i == len(K_CLASSES) - 1:
                per_cu = min(per_cu, LARGEST_CLASS_PER_CU)
            classes.append(sweep_class(H, ctx, nb, per_cu * num_cus, lo, hi, args.seed + i, args.steps, num_cus))
    table = build_table(classes, args.margin)
    doc = dict(
        description="Launch shape of the batched persistent kernel per (tallest QR rows, mean stacked rank K, "
                    "updates per CU) class: Householder panel width and workgroups per CU, timed on one "
                    "graph-captured launch of a representative batch per class (FP64, accuracy 1e-8).",
        command=" ".join(["python -m hcorepp_amd.autotune"] + _args_without_out(argv if argv is not None else sys.argv[1:])),
        generated=datetime.date.today().isoformat(),
        arch="gfx950", compute_units=num_cus, steps=args.steps, margin=args.margin,
        table=table, sweep=classes)
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    log(f"[autotune] {len(classes)} classes in {time.time() - t0:.1f}s -> {args.out}")
    for r in table:
        log(f"[autotune] rows <= {r['max_rows']}, mean K <= {r['max_mean_k']}, updates/CU <= {r['max_per_cu']}: "
            f"panel {r['panel']}, "
            f"{r['workgroups_per_cu']} WG/CU ({r['ms']} ms; built-in {r['builtin_ms']} ms)")


if __name__ == "__main__":
    main()
