#!/usr/bin/env python3
"""Benchmark: TLR tile GEMM + recompression updates/s on one MI355X.

Workload: 4096 independent CCC tile updates C_i = A_i B_i + C_i with nb = 512,
input ranks of A_i, B_i, C_i drawn log-uniform in [8, 192] from a seed, factors
U = orth(Gaussian), V = diag(0.85^i) orth(Gaussian)^T (U orthonormal, V carries
Sigma), C capacity max_rank = 384, absolute accuracy 1e-8.

One step = restore the C tiles' input factors (device-to-device copy) and run
the whole batch through the batched persistent kernel (replayed from a captured
hipGraph).  Inputs are resident in HBM before the timed region.  Also reported:
the plain one-update-per-launch FP64 form on a sample of the same updates, the
CPU baseline (the oracle, OpenMP over updates, MKL sequential per thread) on a
bounded sample, the relative Frobenius error against the FP64 oracle on that
sample, and the roofline of the dominant kernel measured with HIP events.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NB, N_UPDATES, LO, HI, DECAY, MAX_RANK, ACC, SEED = 512, 4096, 8, 192, 0.85, 384, 1e-8, 20261204
PEAK_FP64_TFLOPS = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_flops(nb, ra, rb, rc, r):
    """Algorithmic flops of one CCC update: factor products, two stacked QRs, core product,
    core SVD (22 K^3), Q formation and reconstruction."""
    K = ra + rc
    return (4.0 * ra * rb * nb + 4.0 * (2.0 * nb * K * K - 2.0 / 3.0 * K ** 3) + 2.0 * K ** 3 + 22.0 * K ** 3
            + 4.0 * nb * K * r)


def alg_bytes(nb, ra, rb, rc, r):
    return 8.0 * nb * (2 * ra + 2 * rb + 2 * rc + 2 * r)


def alg_flops_kind(kind, nb, ra, rb, rc, r):
    """Algorithmic flops of one update into a compressed C of rank rc: the factor products of
    the combination (CCC: A_V B_U and the product with B_V; DCC: A B_U; CDC: A_V B), then the
    recompression of the stacked rank K."""
    if kind == "CCC":
        prod, ka = 4.0 * ra * rb * nb, ra
    elif kind == "DCC":
        prod, ka = 2.0 * nb * nb * rb, rb
    else:  # CDC
        prod, ka = 2.0 * ra * nb * nb, ra
    K = ka + rc
    return prod + 4.0 * (2.0 * nb * K * K - 2.0 / 3.0 * K ** 3) + 2.0 * K ** 3 + 22.0 * K ** 3 + 4.0 * nb * K * r


def alg_bytes_kind(kind, nb, ra, rb, rc, r):
    a = nb * nb if kind == "DCC" else 2 * nb * ra
    b = nb * nb if kind == "CDC" else 2 * nb * rb
    return 8.0 * (a + b + 2 * nb * rc + 2 * nb * r)


TLR_N, TLR_NB, TLR_ELL, TLR_ACC, TLR_SEED = 16384, 1024, 0.2, 1e-10, 20261203


def plain_tlr_leg(H, ctx, n_ctiles):
    """Plain FP64 form, one update per launch sequence, on a 16 x 16-tile TLR covariance GEMM
    C += A B with nb = 1024 (A = B = C = exp(-r/0.2) over 16384 points in the unit cube, diagonal
    tiles dense, off-diagonal compressed at 1e-10, maxRank 341).  A sample of C tiles receives its
    16 updates in k order (14 CCC, 1 DCC, 1 CDC each)."""
    from hcorepp_amd.workloads import TLRMatrix
    t0 = time.time()
    M = TLRMatrix(TLR_N, TLR_NB, 3, TLR_ELL, TLR_ACC, TLR_SEED)
    picks = [(0, 1), (9, 4), (14, 15), (3, 12)][:n_ctiles]
    params = H.CompressionParameters(TLR_ACC)
    dev = {}

    def devtile(i, j):
        if (i, j) not in dev:
            t = M.tile(i, j)
            dev[(i, j)] = H.DenseTile(t[1]) if t[0] == "D" else H.CompressedTile(t[1], t[2], max_rank=t[3])
        return dev[(i, j)]

    jobs = []
    for (i, j) in picks:
        tc = M.tile(i, j)
        C = H.CompressedTile(tc[1], tc[2], max_rank=tc[3])
        ups = []
        for k in range(M.T):
            A, B = devtile(i, k), devtile(k, j)
            kind = ("D" if i == k else "C") + ("D" if k == j else "C") + "C"
            ups.append((kind, A, B))
        jobs.append(((i, j), C, ups))
    log(f"[bench] plain TLR nb={TLR_NB}: {len(jobs)} C tiles, {sum(len(j[2]) for j in jobs)} updates, tiles built in "
        f"{time.time() - t0:.1f}s")
    # warm-up on a throw-away copy of the first C tile
    tc = M.tile(*picks[0])
    Cw = H.CompressedTile(tc[1], tc[2], max_rank=tc[3])
    H.HCore.gemm(1.0, jobs[0][2][2][1], jobs[0][2][2][2], 1.0, Cw, ctx, params)
    ctx.sync()
    flops = byts = 0.0
    n_up = 0
    ranks_seen = []
    t1 = time.perf_counter()
    for _, C, ups in jobs:
        for kind, A, B in ups:
            rc = C.rank
            ra = A.rank if kind[0] == "C" else 0
            rb = B.rank if kind[1] == "C" else 0
            H.HCore.gemm(1.0, A, B, 1.0, C, ctx, params)
            flops += alg_flops_kind(kind, TLR_NB, ra, rb, rc, C.rank)
            byts += alg_bytes_kind(kind, TLR_NB, ra, rb, rc, C.rank)
            ranks_seen.append(C.rank)
            n_up += 1
    ctx.sync()
    dt = time.perf_counter() - t1
    # accuracy: each C tile against the dense FP64 result C + sum_k A(i,k) B(k,j)
    errs = []
    for (i, j), C, ups in jobs:
        ref = M.dense_block(i, j).copy()
        for k in range(M.T):
            ta, tb = M.tile(i, k), M.tile(k, j)
            a = ta[1] if ta[0] == "D" else ta[1] @ ta[2]
            b = tb[1] if tb[0] == "D" else tb[1] @ tb[2]
            ref += a @ b
        errs.append(float(np.linalg.norm(C.to_dense() - ref) / np.linalg.norm(ref)))
    tf = flops / dt / 1e12
    out = {"updates_per_s": round(n_up / dt, 3), "updates": n_up, "seconds": round(dt, 3),
           "fp64_tflops": round(tf, 4), "fp64_frac_of_peak": round(tf / PEAK_FP64_TFLOPS, 5),
           "alg_hbm_gb_per_s": round(byts / dt / 1e9, 2), "ranks_min_max": [int(min(ranks_seen)), int(max(ranks_seen))],
           "rel_err_vs_dense_fp64_max": max(errs),
           "workload": f"{TLR_N // TLR_NB}x{TLR_N // TLR_NB}-tile TLR GEMM, nb={TLR_NB}, exp(-r/{TLR_ELL}) over "
                       f"{TLR_N} points in the unit cube (seed {TLR_SEED}), acc {TLR_ACC}, maxRank {TLR_NB // 3}; "
                       f"C tiles {picks}, all 16 updates each"}
    log(f"[bench] plain TLR nb={TLR_NB}: {n_up} updates in {dt:.2f}s -> {n_up / dt:.2f} updates/s, "
        f"{tf:.3f} TF/s FP64 ({100 * tf / PEAK_FP64_TFLOPS:.2f}% of peak), {byts / dt / 1e9:.1f} GB/s algorithmic; "
        f"ranks {min(ranks_seen)}-{max(ranks_seen)}; rel err vs dense max {max(errs):.2e}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--updates", type=int, default=N_UPDATES)
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--cpu-sample", type=int, default=N_UPDATES)
    ap.add_argument("--check-sample", type=int, default=256)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--plain-tlr-tiles", type=int, default=2,
                    help="C tiles of the nb=1024 TLR covariance GEMM timed in the plain form (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import hcorepp_amd as H
    from hcorepp_amd.workloads import batch_ccc

    t0 = time.time()
    ups = batch_ccc(args.updates, NB, SEED, LO, HI, DECAY, threads=min(16, os.cpu_count() or 8))
    log(f"[bench] generated {len(ups)} updates in {time.time() - t0:.1f}s")
    ctx = H.RunContext(0)
    params = H.CompressionParameters(ACC)
    As = [H.CompressedTile(u["au"], u["av"]) for u in ups]
    Bs = [H.CompressedTile(u["bu"], u["bv"]) for u in ups]
    # all C tiles live in one arena so that one device-to-device copy restores them
    arena = H.TileArena(args.updates * H.TileArena.compressed_bytes(NB, NB, MAX_RANK))
    Cs = [H.CompressedTile(u["cu"], u["cv"], max_rank=MAX_RANK, arena=arena) for u in ups]
    C0 = H.DeviceBuffer(arena.off)
    C0.copy_from(arena.buf, arena.off)
    r0 = np.array([c.rank for c in Cs], dtype=np.int64)
    ra = np.array([u["au"].shape[1] for u in ups])
    rb = np.array([u["bu"].shape[1] for u in ups])
    rc = r0.copy()
    log(f"[bench] uploaded; mean K = {np.mean(ra + rc):.1f}, max K = {np.max(ra + rc)}")

    batch = H.Batch(ctx, As, Bs, Cs, 1.0, 1.0, params, H.PREC_FP64)
    log(f"[bench] batch workspace {batch.workspace_bytes() / 2**30:.2f} GiB")
    L = H.lib()

    def restore():
        L.hcorepp_memcpy_async(ctx.h, arena.buf.ptr, C0.ptr, arena.off)
        batch.set_ranks(r0)

    use_graph = not args.no_graph
    if use_graph:
        batch.capture()

    def run_batch():
        if use_graph:
            batch.launch_graph()
        else:
            batch.run()

    for _ in range(args.warmup):
        restore()
        run_batch()
        ctx.sync()
    st, sweeps = batch.status()
    if st != 0:
        raise SystemExit(f"batched run reported numerical status {st}")

    kernel_ms = []
    ctx.sync()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        restore()
        ctx.record(0)
        run_batch()
        ctx.record(1)
        kernel_ms.append(ctx.elapsed_ms(0, 1))
    ctx.sync()
    t_total = time.perf_counter() - t_start
    ranks = batch.ranks()
    st, sweeps = batch.status()
    ms_per_step = 1e3 * t_total / args.steps
    value = args.updates / (t_total / args.steps)
    kavg = float(np.mean(kernel_ms))
    log(f"[bench] batched FP64: {ms_per_step:.1f} ms/step, kernel {kavg:.1f} ms, {value:.0f} updates/s, "
        f"max sweeps {sweeps}")

    # ---- mixed form on the same updates: factor products on FP32 matrix cores (FP32 accumulate)
    mixed = None
    if not args.no_mixed:
        ranks_fp64 = ranks.copy()
        # timing: rerun the whole batch in mixed form on the arena tiles
        bm = H.Batch(ctx, As, Bs, Cs, 1.0, 1.0, params, H.PREC_MIXED)
        if use_graph:
            bm.capture()
        restore()
        bm.set_ranks(r0)
        (bm.launch_graph() if use_graph else bm.run())
        ctx.sync()
        mk = []
        tm0 = time.perf_counter()
        for _ in range(args.steps):
            restore()
            bm.set_ranks(r0)
            ctx.record(2)
            (bm.launch_graph() if use_graph else bm.run())
            ctx.record(3)
            mk.append(ctx.elapsed_ms(2, 3))
        ctx.sync()
        tm = (time.perf_counter() - tm0) / args.steps
        ranks_mixed = bm.ranks()
        stm, swm = bm.status()
        # the arena now holds the mixed results: keep the first check_sample for the error check
        mixed_fac = []
        for i in range(min(args.check_sample, args.updates)):
            Cs[i].rank = int(ranks_mixed[i])
            mixed_fac.append(Cs[i].factors())
        mixed = dict(ms=tm * 1e3, kernel_ms=float(np.mean(mk)), ups=args.updates / tm, status=stm,
                     ranks=ranks_mixed, fac=mixed_fac)
        log(f"[bench] batched mixed: {tm * 1e3:.1f} ms/step, kernel {np.mean(mk):.1f} ms, "
            f"{args.updates / tm:.0f} updates/s, status {stm}")
        bm.close()
        # restore the FP64 results for the checks below
        restore()
        batch.set_ranks(r0)
        run_batch()
        ctx.sync()
        ranks = batch.ranks()
        assert np.array_equal(ranks, ranks_fp64)

    flops = float(np.sum([alg_flops(NB, a, b, c, r) for a, b, c, r in zip(ra, rb, rc, ranks)]))
    traffic_alg = float(np.sum([alg_bytes(NB, a, b, c, r) for a, b, c, r in zip(ra, rb, rc, ranks)]))
    achieved_tflops = flops / (kavg * 1e-3) / 1e12

    # ---- plain one-update-per-launch FP64 form on a sample of the same workload
    ns = min(args.plain_sample, args.updates)
    Cp = [H.CompressedTile(u["cu"], u["cv"], max_rank=MAX_RANK) for u in ups[:ns]]
    H.HCore.gemm(1.0, As[0], Bs[0], 1.0, H.CompressedTile(ups[0]["cu"], ups[0]["cv"], max_rank=MAX_RANK), ctx,
                 params)
    ctx.sync()
    tp = time.perf_counter()
    for i in range(ns):
        H.HCore.gemm(1.0, As[i], Bs[i], 1.0, Cp[i], ctx, params)
    ctx.sync()
    tp = time.perf_counter() - tp
    plain_ups = ns / tp
    # the sample's share of the batch work, to compare like for like
    frac = float(np.sum([alg_flops(NB, a, b, c, 0) for a, b, c in zip(ra[:ns], rb[:ns], rc[:ns])])) / float(
        np.sum([alg_flops(NB, a, b, c, 0) for a, b, c in zip(ra, rb, rc)]))
    plain_full_est = args.updates / (tp / frac)
    log(f"[bench] plain FP64 (sample {ns}): {plain_ups:.1f} updates/s; work-weighted estimate over the "
        f"batch {plain_full_est:.1f} updates/s")

    plain_tlr = plain_tlr_leg(H, ctx, args.plain_tlr_tiles) if args.plain_tlr_tiles > 0 else None

    # ---- CPU baseline (the oracle) on a bounded sample; its outputs also check the GPU results
    from oracle import oracle as O
    ncpu = min(args.cpu_sample, args.updates)
    cores = min(len(os.sched_getaffinity(0)), 16)
    oA = [O.Tile.compressed(u["au"], u["av"]) for u in ups[:ncpu]]
    oB = [O.Tile.compressed(u["bu"], u["bv"]) for u in ups[:ncpu]]
    oC = [O.Tile.compressed(u["cu"], u["cv"], max_rank=MAX_RANK) for u in ups[:ncpu]]
    nwarm = min(32, ncpu)  # warm the thread pool and MKL on copies of the first updates
    O.batch_gemm(1.0, oA[:nwarm], oB[:nwarm], 1.0,
                 [O.Tile.compressed(u["cu"], u["cv"], max_rank=MAX_RANK) for u in ups[:nwarm]], O.Params(ACC), cores)
    sec, orc_ranks = O.batch_gemm(1.0, oA, oB, 1.0, oC, O.Params(ACC), cores)
    frac_cpu = float(np.sum([alg_flops(NB, a, b, c, 0) for a, b, c in zip(ra[:ncpu], rb[:ncpu], rc[:ncpu])])) / float(
        np.sum([alg_flops(NB, a, b, c, 0) for a, b, c in zip(ra, rb, rc)]))
    cpu_ups = ncpu / sec
    cpu_full = args.updates / (sec / frac_cpu)
    log(f"[bench] CPU oracle ({cores} threads, sample {ncpu}): {sec:.1f}s -> {cpu_ups:.1f} updates/s "
        f"(work-weighted over the batch {cpu_full:.1f})")
    rank_mis = int(np.sum(np.asarray(ranks[:ncpu]) != np.asarray(orc_ranks)))
    errs, errs_dense = [], []
    nchk = min(args.check_sample, ncpu)
    for i in range(nchk):
        ref = oC[i].to_dense()
        Cs[i].rank = int(ranks[i])
        u_, v_ = Cs[i].factors()
        g = u_ @ v_
        errs.append(np.linalg.norm(g - ref) / np.linalg.norm(ref))
        dense = ups[i]["cu"] @ ups[i]["cv"] + ups[i]["au"] @ (ups[i]["av"] @ ups[i]["bu"]) @ ups[i]["bv"]
        errs_dense.append(np.linalg.norm(g - dense) / np.linalg.norm(dense))
    log(f"[bench] rel err vs FP64 oracle (recompressed, first {nchk}): max {max(errs):.2e}; vs dense FP64: max "
        f"{max(errs_dense):.2e}; rank mismatches {rank_mis}/{ncpu}")
    mixed_out = None
    if mixed is not None:
        merrs = []
        for i in range(min(nchk, len(mixed["fac"]))):
            u_, v_ = mixed["fac"][i]
            merrs.append(np.linalg.norm(u_ @ v_ - oC[i].to_dense()) / np.linalg.norm(oC[i].to_dense()))
        mis = int(np.sum(np.asarray(mixed["ranks"][:ncpu]) != np.asarray(orc_ranks)))
        log(f"[bench] mixed: rel err vs FP64 oracle max {max(merrs):.2e} (first {len(merrs)}); rank mismatches "
            f"{mis}/{ncpu}")
        mixed_out = {"updates_per_s": round(mixed["ups"], 2), "ms_per_step": round(mixed["ms"], 3),
                     "kernel_ms": round(mixed["kernel_ms"], 3), "rel_err_vs_fp64_oracle_max": float(max(merrs)),
                     "rank_mismatches_vs_oracle": f"{mis}/{ncpu}",
                     "meets_1e-8": bool(max(merrs) <= 1e-8),
                     "form": "batched, factor products on FP32 MFMA with FP32 accumulation; QR/SVD/truncation FP64"}

    out = {
        "metric": "TLR tile GEMM+recompression updates/s at rel err <=1e-8 (batched vs plain FP64)",
        "value": round(value, 2),
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"batched CCC tile updates, {args.updates} per step, nb={NB}, ranks log-uniform "
                        f"[{LO},{HI}], decay {DECAY}, max_rank {MAX_RANK}, acc {ACC}",
            "form": "batched FP64, persistent kernel" + (" in hipGraph" if use_graph else ""),
            "global_batch": args.updates,
            "parallelism": "single GPU",
        },
        "plain_fp64_updates_per_s": round(plain_full_est, 2),
        "speedup_vs_plain_fp64": round(value / plain_full_est, 2),
        "rel_err_vs_fp64_oracle_max": float(max(errs)),
        "rel_err_vs_dense_fp64_max": float(max(errs_dense)),
        "rank_mismatches_vs_oracle": f"{rank_mis}/{ncpu}",
        "rel_err_checked_updates": nchk,
        "max_jacobi_sweeps": int(sweeps),
        "mixed_form": mixed_out,
        "plain_tlr_nb1024": plain_tlr,
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved_tflops, 4),
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved_tflops / PEAK_FP64_TFLOPS, 5),
            "traffic": None,
            "kernel": "k_batched (persistent batched recompression)",
            "kernel_ms": round(kavg, 3),
            "alg_flops_per_launch": flops,
            "alg_bytes_per_launch": traffic_alg,
        },
        "cpu_baseline": {
            "value": round(cpu_full, 3),
            "unit": "updates/s",
            "cores": cores,
            "kind": "port",
            "sample": (f"all {ncpu} updates of the batch, one pass after a {nwarm}-update warm-up; OpenMP over "
                       f"updates ({cores} threads), MKL sequential per thread ({sec:.2f} s)"
                       if ncpu == args.updates else
                       f"first {ncpu} updates of the batch, OpenMP over updates, MKL sequential per thread; "
                       f"rate scaled to the whole batch by algorithmic-flop share ({ncpu / sec:.2f} updates/s "
                       f"on the sample itself)"),
        },
    }
    print(json.dumps(out), flush=True)
    batch.close()
    ctx.close()


if __name__ == "__main__":
    main()
