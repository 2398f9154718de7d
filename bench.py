#!/usr/bin/env python3
"""Benchmark: TLR tile GEMM + recompression updates/s on one MI355X.

Workload: 4096 independent CCC tile updates C_i = A_i B_i + C_i with nb = 512,
input ranks of A_i, B_i, C_i drawn log-uniform in [8, 192] from a seed, factors
U = orth(Gaussian), V = diag(0.85^i) orth(Gaussian)^T (U orthonormal, V carries
Sigma), C capacity max_rank = 384, absolute accuracy 1e-8.

One step = restore the C tiles' input factors (device-to-device copy) and run
the whole batch through the batched persistent kernel (replayed from a captured
hipGraph).  Inputs are resident in HBM before the timed region.  Forms measured
on the same updates, each checked against the FP64 oracle (relative Frobenius
error on a sample, ranks on all updates):
  fp64_exact    exact core (truncated only at the rounding level of its QR)
  fp64_learned  core presized by the trained BF16 rank MLP, exact acceptance test
  fp64_rule     the same tolerance applied directly to the pivoted R factor
  mixed_rule    factor products on FP32 matrix cores (FP32 accumulation over error-free slices
                of the FP64 operands), core truncated by the same tolerance rule as fp64_rule
  mixed_learned the same products with the learned core sizing of fp64_learned
The headline `value` is the fastest form whose error stays <= 1e-8.  Also
reported: the plain one-update-per-launch FP64 form on a sample of the same
updates and on the nb=1024 TLR covariance GEMM, the CPU baseline (the oracle,
OpenMP over updates, MKL sequential per thread) on the whole batch, and the
roofline of the dominant kernel measured with HIP events.

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

# Hardware queues of this process: HIP maps streams onto GPU_MAX_HW_QUEUES queues (its default is
# 4), so the concurrent plain leg's PLAIN_STREAMS contexts only run side by side with at least as
# many queues.  Read once when the HIP runtime starts, hence set before anything initialises it;
# a larger value from the caller is kept.
HW_QUEUES = 16
if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < HW_QUEUES:
    os.environ["GPU_MAX_HW_QUEUES"] = str(HW_QUEUES)

NB, N_UPDATES, LO, HI, DECAY, MAX_RANK, ACC, SEED = 512, 4096, 8, 192, 0.85, 384, 1e-8, 20261204
LEARNED_TOL, LEARNED_MARGIN = 0.1, 40  # accept a presized core when the dropped part <= 0.1 * accuracy
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
# C tiles of the nb=1024 TLR GEMM sample: a 4 x 4 block of off-diagonal tiles, 16 updates each
TLR_PICKS = [(i, j) for i in range(4) for j in range(8, 12)]
PLAIN_STREAMS = 8  # host threads / contexts (streams) of the concurrent plain leg


def host_info():
    """The CPU the baselines run on: logical CPUs, the affinity mask, model, thread settings."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)), "cpu_model": model,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"), "MKL_NUM_THREADS": os.environ.get("MKL_NUM_THREADS")}


def cpu_threads():
    """Threads the CPU baselines use: the affinity mask, capped at OMP_NUM_THREADS when it is set
    (the GPU host's CPU share for one GPU)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def tlr_tiles_dev(H, M, keys):
    out = {}
    for key in keys:
        t = M.tile(*key)
        out[key] = H.DenseTile(t[1]) if t[0] == "D" else H.CompressedTile(t[1], t[2], max_rank=t[3])
    return out


def plain_tlr_leg(H, ctx, M, picks):
    """Plain FP64 form, one update per launch sequence, on the 16 x 16-tile TLR covariance GEMM
    C += A B with nb = 1024 (A = B = C = exp(-r/0.2) over 16384 points in the unit cube, diagonal
    tiles dense, off-diagonal compressed at 1e-10, maxRank 341).  A sample of C tiles receives its
    16 updates in k order (14 CCC, 1 DCC, 1 CDC each)."""
    t0 = time.time()
    params = H.CompressionParameters(TLR_ACC)
    keys = sorted({(i, k) for i, _ in picks for k in range(M.T)} | {(k, j) for _, j in picks for k in range(M.T)})
    dev = tlr_tiles_dev(H, M, keys)
    jobs = []
    for (i, j) in picks:
        tc = M.tile(i, j)
        C = H.CompressedTile(tc[1], tc[2], max_rank=tc[3])
        ups = [(("D" if i == k else "C") + ("D" if k == j else "C") + "C", dev[(i, k)], dev[(k, j)])
               for k in range(M.T)]
        jobs.append(((i, j), C, ups))
    log(f"[bench] plain TLR nb={TLR_NB}: {len(jobs)} C tiles, {sum(len(j[2]) for j in jobs)} updates, uploaded in "
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
    # the same C tiles again with PLAIN_STREAMS host threads, each on its own context (stream and
    # arena) and its own share of the C tiles, as the reference's stream drivers run independent
    # tiles concurrently (par_fixed_rank_streams_main.cpp:104-137)
    import concurrent.futures as cf
    ctxs = [H.RunContext(0) for _ in range(PLAIN_STREAMS)]
    jobs2 = []
    for (i, j), _, ups in jobs:
        tc = M.tile(i, j)
        jobs2.append((H.CompressedTile(tc[1], tc[2], max_rank=tc[3]), ups))

    def run_share(t):
        for C, ups in jobs2[t::PLAIN_STREAMS]:
            for _, A, B in ups:
                H.HCore.gemm(1.0, A, B, 1.0, C, ctxs[t], params)
        ctxs[t].sync()

    for c in ctxs:
        c.sync()
    t2 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=PLAIN_STREAMS) as ex:
        list(ex.map(run_share, range(PLAIN_STREAMS)))
    dt_streams = time.perf_counter() - t2
    same = all(C2.rank == C.rank and np.allclose(C2.to_dense(), C.to_dense(), rtol=0, atol=1e-12 * max(1.0, float(
        np.abs(C.to_dense()).max()))) for (_, C, _), (C2, _) in zip(jobs[:4], jobs2[:4]))
    for c in ctxs:
        c.close()
    # accuracy: each C tile against the dense FP64 result C + sum_k A(i,k) B(k,j)
    errs = []
    for (i, j), C, ups in jobs[:4]:
        ref = M.dense_block(i, j).copy()
        for k in range(M.T):
            ta, tb = M.tile(i, k), M.tile(k, j)
            a = ta[1] if ta[0] == "D" else ta[1] @ ta[2]
            b = tb[1] if tb[0] == "D" else tb[1] @ tb[2]
            ref += a @ b
        errs.append(float(np.linalg.norm(C.to_dense() - ref) / np.linalg.norm(ref)))
    tf = flops / dt / 1e12
    out = {"updates_per_s": round(n_up / dt, 3), "updates": n_up, "seconds": round(dt, 3),
           f"updates_per_s_{PLAIN_STREAMS}_streams": round(n_up / dt_streams, 3),
           "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
           "streams_results_equal_one_stream": bool(same),
           "fp64_tflops": round(tf, 4), "fp64_frac_of_peak": round(tf / PEAK_FP64_TFLOPS, 5),
           "alg_hbm_gb_per_s": round(byts / dt / 1e9, 2), "ranks_min_max": [int(min(ranks_seen)), int(max(ranks_seen))],
           "rel_err_vs_dense_fp64_max": max(errs), "rel_err_checked_c_tiles": len(errs),
           "workload": f"{TLR_N // TLR_NB}x{TLR_N // TLR_NB}-tile TLR GEMM, nb={TLR_NB}, exp(-r/{TLR_ELL}) over "
                       f"{TLR_N} points in the unit cube (seed {TLR_SEED}), acc {TLR_ACC}, maxRank {TLR_NB // 3}; "
                       f"{len(picks)} C tiles (rows 0-3 x cols 8-11), all 16 updates each"}
    for b in dev.values():
        b.buf.free()
    log(f"[bench] plain TLR nb={TLR_NB}: {n_up} updates in {dt:.2f}s -> {n_up / dt:.2f} updates/s, "
        f"{tf:.3f} TF/s FP64 ({100 * tf / PEAK_FP64_TFLOPS:.2f}% of peak), {byts / dt / 1e9:.1f} GB/s algorithmic; "
        f"ranks {min(ranks_seen)}-{max(ranks_seen)}; rel err vs dense max {max(errs):.2e}; with {PLAIN_STREAMS} "
        f"streams {n_up / dt_streams:.2f} updates/s")
    return out


CPU_N, CPU_NB, CPU_ELL, CPU_ACC, CPU_SEED = 1024, 256, 0.1, 1e-8, 20261202


def _orc_tile(O, t, copy=False):
    if t[0] == "D":
        return O.Tile.dense(t[1].copy(order="F") if copy else t[1])
    return O.Tile.compressed(t[1], t[2], max_rank=t[3])


def gpu_tlr4_leg(H, ctx):
    """The 4 x 4-tile nb=256 covariance GEMM of the CPU reference-like leg through the plain FP64
    form on the GPU (one update per launch sequence, the reference's loop order; fresh C tiles
    uploaded before each run; one warm-up, median of 5)."""
    from hcorepp_amd.workloads import TLRMatrix
    M = TLRMatrix(CPU_N, CPU_NB, 2, CPU_ELL, CPU_ACC, CPU_SEED)
    T = M.T
    params = H.CompressionParameters(CPU_ACC)
    dev = tlr_tiles_dev(H, M, [(i, k) for i in range(T) for k in range(T)])
    secs = []
    for rep in range(6):
        Cg = tlr_tiles_dev(H, M, [(i, j) for i in range(T) for j in range(T)])
        ctx.sync()
        t0 = time.perf_counter()
        for i in range(T):
            for j in range(T):
                for k in range(T):
                    H.HCore.gemm(1.0, dev[(i, k)], dev[(k, j)], 1.0, Cg[(i, j)], ctx, params)
        ctx.sync()
        if rep:
            secs.append(time.perf_counter() - t0)
    sec = float(np.median(secs))
    log(f"[bench] GPU plain 4x4 nb=256: {sec * 1e3:.1f} ms median -> {T ** 3 / sec:.1f} updates/s")
    return {"updates_per_s": round(T ** 3 / sec, 2), "updates": T ** 3, "seconds_median": round(sec, 5),
            "runs": "1 warm-up + median of 5", "form": "plain FP64, one update per launch sequence"}


def cpu_tlr_legs(O, M1024, picks, threads):
    """CPU runs of the oracle (the reference's CPU algorithm over MKL) on the tile GEMMs:
    reference-like mode on the 4 x 4-tile nb=256 covariance GEMM (OpenMP over the 16 C tiles, MKL
    threads = ceil(threads / 16), as omp_main.cpp:85-126; one warm-up, median of 5), and throughput
    mode on the nb=1024 sample of the 16 x 16-tile GEMM (OpenMP over its C tiles, MKL sequential;
    one warm-up, median of 3)."""
    from hcorepp_amd.workloads import TLRMatrix, prefetch_tiles
    out = {}
    M = TLRMatrix(CPU_N, CPU_NB, 2, CPU_ELL, CPU_ACC, CPU_SEED)
    T = M.T
    A = [[_orc_tile(O, M.tile(i, k), copy=True) for k in range(T)] for i in range(T)]
    B = [[_orc_tile(O, M.tile(k, j), copy=True) for j in range(T)] for k in range(T)]
    p = O.Params(CPU_ACC)
    mkl = max(1, -(-threads // (T * T)))
    secs = []
    for rep in range(6):
        Cg = [[_orc_tile(O, M.tile(i, j), copy=True) for j in range(T)] for i in range(T)]
        sec = O.tlr_gemm_omp(1.0, A, B, 1.0, Cg, p, threads, mkl)
        if rep:
            secs.append(sec)
    sec = float(np.median(secs))
    ranks = [Cg[i][j].rank for i in range(T) for j in range(T) if i != j]
    out["reference_like_4x4_nb256"] = {
        "updates_per_s": round(T ** 3 / sec, 2), "updates": T ** 3, "seconds_median": round(sec, 5),
        "omp_threads": threads, "mkl_threads_per_tile": mkl, "runs": "1 warm-up + median of 5",
        "final_ranks_min_max": [int(min(ranks)), int(max(ranks))],
        "workload": f"{T}x{T}-tile TLR GEMM C += A B, nb={CPU_NB}, exp(-r/{CPU_ELL}) over {CPU_N} points in the "
                    f"unit square (seed {CPU_SEED}), acc {CPU_ACC}, maxRank {CPU_NB // 3}, A = B = C"}
    log(f"[bench] CPU reference-like 4x4 nb=256: {sec * 1e3:.1f} ms median -> {T ** 3 / sec:.1f} updates/s "
        f"({threads} OpenMP threads x {mkl} MKL)")
    # throughput mode on the nb=1024 sample
    ii = sorted({i for i, _ in picks})
    jj = sorted({j for _, j in picks})
    A2 = [[_orc_tile(O, M1024.tile(i, k), copy=True) for k in range(M1024.T)] for i in ii]
    B2 = [[_orc_tile(O, M1024.tile(k, j), copy=True) for j in jj] for k in range(M1024.T)]
    p2 = O.Params(TLR_ACC)
    secs = []
    for rep in range(4):
        Cg = [[_orc_tile(O, M1024.tile(i, j), copy=True) for j in jj] for i in ii]
        sec = O.tlr_gemm_omp(1.0, A2, B2, 1.0, Cg, p2, threads, 1)
        if rep:
            secs.append(sec)
    sec = float(np.median(secs))
    n_up = len(ii) * len(jj) * M1024.T
    out["throughput_16x16_nb1024_sample"] = {
        "updates_per_s": round(n_up / sec, 3), "updates": n_up, "seconds_median": round(sec, 4),
        "omp_threads": min(threads, len(ii) * len(jj)), "mkl_threads_per_tile": 1, "runs": "1 warm-up + median of 3",
        "workload": f"the nb={TLR_NB} GEMM's C tiles rows {ii[0]}-{ii[-1]} x cols {jj[0]}-{jj[-1]}, 16 updates each "
                    f"(OpenMP over the {len(ii) * len(jj)} C tiles, each tile's updates in k order)"}
    log(f"[bench] CPU throughput nb=1024 sample: {n_up} updates in {sec:.2f}s -> {n_up / sec:.2f} updates/s")
    return out


CHOL_N, CHOL_NB, CHOL_ELL, CHOL_ACC, CHOL_SEED, CHOL_NUGGET = 16384, 512, 0.1, 1e-8, 20261205, 1e-2


def cholesky_leg(H, ctx, forms):
    """TLR Cholesky A = L L^T of a 32 x 32-tile SPD matrix (nb = 512): exp(-r/0.1) + 1e-2 I over
    16384 points in the unit cube, off-diagonal tiles compressed at 1e-8 (maxRank 170).  Each
    form factors a fresh device copy; the step loop (potrf, trsm, syrk, transposes, trailing
    updates) is timed end to end.  Backward error ||A - L L^T||_F / ||A||_F against the dense
    matrix and against the compressed input, with L reassembled densely."""
    import ctypes as C
    from hcorepp_amd.cholesky import TLRCholesky
    from hcorepp_amd.workloads import TLRMatrix, prefetch_tiles
    t0 = time.time()
    M = TLRMatrix(CHOL_N, CHOL_NB, 3, CHOL_ELL, CHOL_ACC, CHOL_SEED, nugget=CHOL_NUGGET)
    T = M.T
    prefetch_tiles(M, [(i, j) for i in range(T) for j in range(i + 1)], min(16, os.cpu_count() or 8))
    ranks0 = [M.tile(i, j)[1].shape[1] for i in range(T) for j in range(i)]
    log(f"[bench] TLR Cholesky {T}x{T} tiles nb={CHOL_NB}: tiles built in {time.time() - t0:.1f}s, input ranks "
        f"{min(ranks0)}-{max(ranks0)} (mean {np.mean(ranks0):.1f})")
    n, nb = CHOL_N, CHOL_NB
    A = M.dense_block_range(0, n)  # symmetric: its C-order buffer is also its column-major buffer
    nA = float(np.linalg.norm(A))
    At = np.zeros_like(A)  # the compressed input, assembled
    for i in range(T):
        for j in range(i + 1):
            t = M.tile(i, j)
            blk = t[1] if t[0] == "D" else t[1] @ t[2]
            At[i * nb:(i + 1) * nb, j * nb:(j + 1) * nb] = blk
            if i != j:
                At[j * nb:(j + 1) * nb, i * nb:(i + 1) * nb] = blk.T
    nAt = float(np.linalg.norm(At))
    input_err = float(np.linalg.norm(A - At)) / nA
    dA, dAt = H.DeviceBuffer.from_numpy(A.reshape(-1)), H.DeviceBuffer.from_numpy(At.reshape(-1))
    dR, dL = H.DeviceBuffer(A.nbytes), H.DeviceBuffer(A.nbytes)
    del At

    def residual(src, dLbuf, norm):
        """||S - L L^T||_F / norm on the device; dLbuf holds L^T column-major (L in C order)."""
        dR.copy_from(src, A.nbytes)
        H._check(H.lib().hcorepp_kernel_gemm_d(ctx.h, b"T", b"N", n, n, n, -1.0, dLbuf.ptr, n, dLbuf.ptr, n, 1.0,
                                               dR.ptr, n))
        ctx.sync()
        return float(np.linalg.norm(dR.download(np.float64))) / norm
    params = H.CompressionParameters(CHOL_ACC)
    n_upd = sum((T - 1 - k) * (T - 2 - k) // 2 for k in range(T))
    out = {"workload": f"TLR Cholesky, {T}x{T} tiles, nb={CHOL_NB}: exp(-r/{CHOL_ELL}) + {CHOL_NUGGET} I over "
                       f"{CHOL_N} points in the unit cube (seed {CHOL_SEED}), acc {CHOL_ACC}, maxRank {CHOL_NB // 3}; "
                       f"{T} potrf, {T * (T - 1) // 2} trsm, {T * (T - 1) // 2} syrk, {n_upd} trailing updates",
           "input_compression_rel_err": input_err, "forms": {}}
    for name in forms:
        F = TLRCholesky(ctx, M)
        ctx.sync()
        tt = time.perf_counter()
        if name == "plain_fp64":
            F.factor_plain(params, progress=lambda k: log(f"[bench] TLR Cholesky plain_fp64: step {k}/{T} done "
                                                          f"after {time.perf_counter() - tt:.1f}s") if k % 8 == 0
                           else None)
        elif name == "batched_fp64":
            F.factor_batched(params)
        elif name == "batched_rule":
            F.factor_batched(params, tol=LEARNED_TOL)
        elif name == "batched_mixed":
            F.factor_batched(params, precision=H.PREC_MIXED)
        elif name == "batched_mixed_rule":
            F.factor_batched(params, precision=H.PREC_MIXED, tol=LEARNED_TOL)
        ctx.sync()
        dt = time.perf_counter() - tt
        dL.upload(np.ascontiguousarray(F.dense_factor()).reshape(-1))
        be = residual(dA, dL, nA)
        be_t = residual(dAt, dL, nAt)
        rk = list(F.ranks().values())
        out["forms"][name] = {"seconds": round(dt, 3), "trailing_updates_per_s": round(n_upd / dt, 1),
                              "backward_error_vs_A": be, "backward_error_vs_compressed_input": be_t,
                              "final_ranks_min_max_mean": [int(min(rk)), int(max(rk)), round(float(np.mean(rk)), 1)]}
        log(f"[bench] TLR Cholesky {name}: {dt:.2f}s ({n_upd / dt:.0f} trailing updates/s); backward error "
            f"{be:.2e} vs A, {be_t:.2e} vs the compressed input; ranks {min(rk)}-{max(rk)}")
        del F
    if "plain_fp64" in out["forms"]:
        base = out["forms"]["plain_fp64"]["seconds"]
        for name, f in out["forms"].items():
            f["speedup_vs_plain_fp64"] = round(base / f["seconds"], 2)
    for b in (dA, dAt, dR, dL):
        b.free()
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
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--forms", default="fp64_exact,fp64_learned,fp64_rule,mixed_rule,mixed_learned",
                    help="comma-separated forms to measure (profiling runs pick one)")
    ap.add_argument("--plain-tlr-tiles", type=int, default=len(TLR_PICKS),
                    help="C tiles of the nb=1024 TLR covariance GEMM sample timed in the plain form and on the "
                         "CPU (0 = skip)")
    ap.add_argument("--no-cpu-tlr", action="store_true", help="skip the CPU runs of the tile GEMMs")
    ap.add_argument("--learned-margin", type=int, default=LEARNED_MARGIN,
                    help="columns added to the rank model's prediction when presizing the core")
    ap.add_argument("--cholesky-forms", default="batched_fp64,batched_rule,batched_mixed_rule,plain_fp64",
                    help="TLR Cholesky forms to time (empty = skip the Cholesky leg)")
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
    nchk = min(args.check_sample, args.updates)
    log(f"[bench] uploaded; mean K = {np.mean(ra + rc):.1f}, max K = {np.max(ra + rc)}")
    L = H.lib()
    use_graph = not args.no_graph

    batch = H.Batch(ctx, As, Bs, Cs, 1.0, 1.0, params, H.PREC_FP64)
    log(f"[bench] batch workspace {batch.workspace_bytes() / 2**30:.2f} GiB")
    # launch shape: the autotune table's entry for this class (tallest QR NB rows, mean K, updates per CU)
    tuned, _, _ = H.autotune_lookup(NB, float(np.mean(ra + rc)), args.updates / ctx.compute_units())
    shape = {"panel": batch.panel(), "workgroups": batch.grid(),
             "source": "autotune table (hcorepp_amd/autotune.json)" if tuned else "built-in rule"}
    log(f"[bench] launch shape: panel {shape['panel']}, {shape['workgroups']} workgroups ({shape['source']})")
    if use_graph:
        batch.capture()
    forms = {}

    def measure(name, b):
        """Time `steps` passes of batch b (after `warmup`), keep its ranks and the first nchk results."""
        def restore():
            L.hcorepp_memcpy_async(ctx.h, arena.buf.ptr, C0.ptr, arena.off)
            b.set_ranks(r0)

        def launch():
            b.launch_graph() if use_graph else b.run()

        for _ in range(max(args.warmup, 1)):
            restore()
            launch()
            ctx.sync()
        st, _ = b.status()
        if st != 0:
            raise SystemExit(f"form {name}: numerical status {st}")
        kms = []
        ctx.sync()
        tt = time.perf_counter()
        for _ in range(args.steps):
            restore()
            ctx.record(0)
            launch()
            ctx.record(1)
            kms.append(ctx.elapsed_ms(0, 1))
        ctx.sync()
        dt = (time.perf_counter() - tt) / args.steps
        st, sweeps = b.status()
        rk = b.ranks()
        fac = []
        for i in range(nchk):
            Cs[i].rank = int(rk[i])
            fac.append(Cs[i].factors())
        f = dict(ups=args.updates / dt, ms=dt * 1e3, kernel_ms=float(np.mean(kms)), ranks=rk, fac=fac,
                 sweeps=int(sweeps))
        forms[name] = f
        log(f"[bench] {name}: {dt * 1e3:.1f} ms/step, kernel {f['kernel_ms']:.1f} ms, {f['ups']:.0f} updates/s, "
            f"max sweeps {sweeps}")
        return f

    want = set(args.forms.split(","))
    if args.no_learned:
        want -= {"fp64_learned"}
    if args.no_mixed:
        want -= {"mixed_rule", "mixed_learned"}
    if "fp64_exact" in want:
        measure("fp64_exact", batch)
    model_path = os.path.join(ROOT, "hcorepp_amd", "rank_mlp.npz")
    model = None
    if os.path.exists(model_path) and ({"fp64_learned", "mixed_learned"} & want):
        from hcorepp_amd.learned import RankPredictor
        model = RankPredictor.load(ctx, model_path)
    if "fp64_learned" in want and model is not None:
        batch.set_rank_model(model.bf16(), tol=LEARNED_TOL, margin=args.learned_margin)
        f = measure("fp64_learned", batch)
        fb, pred = batch.learned_stats()
        f["fallbacks"] = fb
        f["mae"] = float(np.mean(np.abs(pred - f["ranks"])))
        log(f"[bench] fp64_learned: fallbacks {fb}/{args.updates}, mean |predicted - rank| {f['mae']:.2f}")
    if "fp64_rule" in want:
        batch.set_rank_model(None, tol=LEARNED_TOL, margin=0)
        measure("fp64_rule", batch)
    batch.set_rank_model(None, tol=0.0, margin=0)
    if {"mixed_rule", "mixed_learned"} & want:
        bm = H.Batch(ctx, As, Bs, Cs, 1.0, 1.0, params, H.PREC_MIXED)
        if use_graph:
            bm.capture()
        if "mixed_rule" in want:
            bm.set_rank_model(None, tol=LEARNED_TOL, margin=0)
            measure("mixed_rule", bm)
        if "mixed_learned" in want and model is not None:
            bm.set_rank_model(model.bf16(), tol=LEARNED_TOL, margin=args.learned_margin)
            f = measure("mixed_learned", bm)
            fb, pred = bm.learned_stats()
            f["fallbacks"] = fb
            f["mae"] = float(np.mean(np.abs(pred - f["ranks"])))
        bm.close()

    # ---- plain one-update-per-launch FP64 form on a sample of the same workload
    ns = max(min(args.plain_sample, args.updates), 1)
    Cp = [H.CompressedTile(u["cu"], u["cv"], max_rank=MAX_RANK) for u in ups[:ns]]
    H.HCore.gemm(1.0, As[0], Bs[0], 1.0, H.CompressedTile(ups[0]["cu"], ups[0]["cv"], max_rank=MAX_RANK), ctx,
                 params)
    ctx.sync()
    tp = time.perf_counter()
    for i in range(ns):
        H.HCore.gemm(1.0, As[i], Bs[i], 1.0, Cp[i], ctx, params)
    ctx.sync()
    tp = time.perf_counter() - tp
    # the sample's share of the batch work, to compare like for like
    frac = float(np.sum([alg_flops(NB, a, b, c, 0) for a, b, c in zip(ra[:ns], rb[:ns], rc[:ns])])) / float(
        np.sum([alg_flops(NB, a, b, c, 0) for a, b, c in zip(ra, rb, rc)]))
    plain_full_est = args.updates / (tp / frac)
    log(f"[bench] plain FP64 (sample {ns}): {ns / tp:.1f} updates/s; work-weighted estimate over the batch "
        f"{plain_full_est:.1f} updates/s")
    picks = TLR_PICKS[:args.plain_tlr_tiles]
    M1024 = None
    if picks:
        from hcorepp_amd.workloads import TLRMatrix, prefetch_tiles
        t0 = time.time()
        M1024 = TLRMatrix(TLR_N, TLR_NB, 3, TLR_ELL, TLR_ACC, TLR_SEED)
        keys = sorted({(i, k) for i, _ in picks for k in range(M1024.T)} | {(k, j) for _, j in picks
                                                                             for k in range(M1024.T)} | set(picks))
        prefetch_tiles(M1024, keys, cpu_threads())
        log(f"[bench] nb={TLR_NB} TLR tiles: {len(keys)} built in {time.time() - t0:.1f}s")
    plain_tlr = plain_tlr_leg(H, ctx, M1024, picks) if picks else None
    plain_tlr4 = gpu_tlr4_leg(H, ctx) if not args.no_cpu_tlr else None
    chol_forms = [f for f in args.cholesky_forms.split(",") if f]
    if chol_forms:
        batch.close()  # free the batch workspace before the Cholesky leg
        chol = cholesky_leg(H, ctx, chol_forms)
    else:
        chol = None

    # ---- CPU baseline (the oracle) on the whole batch; its outputs also check every form
    from oracle import oracle as O
    ncpu = min(args.cpu_sample, args.updates)
    cores = cpu_threads()
    oA = [O.Tile.compressed(u["au"], u["av"]) for u in ups[:ncpu]]
    oB = [O.Tile.compressed(u["bu"], u["bv"]) for u in ups[:ncpu]]
    oC = [O.Tile.compressed(u["cu"], u["cv"], max_rank=MAX_RANK) for u in ups[:ncpu]]
    nwarm = min(32, ncpu)  # warm the thread pool and MKL on copies of the first updates
    O.batch_gemm(1.0, oA[:nwarm], oB[:nwarm], 1.0,
                 [O.Tile.compressed(u["cu"], u["cv"], max_rank=MAX_RANK) for u in ups[:nwarm]], O.Params(ACC), cores)
    sec, orc_ranks = O.batch_gemm(1.0, oA, oB, 1.0, oC, O.Params(ACC), cores)
    frac_cpu = float(np.sum([alg_flops(NB, a, b, c, 0) for a, b, c in zip(ra[:ncpu], rb[:ncpu], rc[:ncpu])])) / float(
        np.sum([alg_flops(NB, a, b, c, 0) for a, b, c in zip(ra, rb, rc)]))
    cpu_full = args.updates / (sec / frac_cpu)
    log(f"[bench] CPU oracle ({cores} threads, {ncpu} updates): {sec:.2f}s -> {ncpu / sec:.1f} updates/s")
    cpu_legs = None
    if not args.no_cpu_tlr and picks:
        cpu_legs = cpu_tlr_legs(O, M1024, picks, cores)
    orc_dense = [oC[i].to_dense() for i in range(min(nchk, ncpu))]
    for name, f in forms.items():
        errs = [np.linalg.norm(u_ @ v_ - orc_dense[i]) / np.linalg.norm(orc_dense[i])
                for i, (u_, v_) in enumerate(f["fac"][:len(orc_dense)])]
        f["err"] = float(max(errs))
        f["rank_mis"] = int(np.sum(np.asarray(f["ranks"][:ncpu]) != np.asarray(orc_ranks)))
        log(f"[bench] {name}: rel err vs FP64 oracle max {f['err']:.2e} (first {len(errs)}); rank mismatches "
            f"{f['rank_mis']}/{ncpu}")
    errs_dense = []
    ref_form = "fp64_exact" if "fp64_exact" in forms else next(iter(forms))
    for i, (u_, v_) in enumerate(forms[ref_form]["fac"]):
        dense = ups[i]["cu"] @ ups[i]["cv"] + ups[i]["au"] @ (ups[i]["av"] @ ups[i]["bu"]) @ ups[i]["bv"]
        errs_dense.append(np.linalg.norm(u_ @ v_ - dense) / np.linalg.norm(dense))

    # headline: the fastest form within the metric's error bound whose ranks all equal the oracle's
    valid = {k: f for k, f in forms.items() if f["err"] <= 1e-8 and f["rank_mis"] == 0}
    best = max(valid, key=lambda k: valid[k]["ups"])
    hb = forms[best]
    ranks = hb["ranks"]
    flops = float(np.sum([alg_flops(NB, a, b, c, r) for a, b, c, r in zip(ra, rb, rc, ranks)]))
    traffic_alg = float(np.sum([alg_bytes(NB, a, b, c, r) for a, b, c, r in zip(ra, rb, rc, ranks)]))
    achieved_tflops = flops / (hb["kernel_ms"] * 1e-3) / 1e12
    # HBM bytes per launch: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this kernel and form, calibrated
    # against a copy of known size (the committed profile gives the commands); bench.py cannot
    # collect counters itself, so the figure is quoted from there for the same form
    traffic, traffic_src = None, None
    prof = os.path.join(ROOT, "profiles", f"k_batched_ccc{args.updates}_nb{NB}_{best}.md")
    if os.path.exists(prof):
        for line in open(prof):
            if line.startswith("traffic_bytes_per_launch:"):
                traffic = float(line.split(":")[1])
                traffic_src = f"rocprofv3 PMC, {os.path.relpath(prof, ROOT)}"
    # executed FP64 VALU flops and matrix-core busy fraction of this form (rocprofv3 SQ counters;
    # quoted from the committed summary, which gives the command)
    executed, mfma_busy, counters_src = None, None, None
    exec_valu, exec_mfma = None, None
    cprof = os.path.join(ROOT, "profiles", f"k_batched_ccc{args.updates}_nb{NB}_valu_mfma_counters.md")
    if os.path.exists(cprof):
        for line in open(cprof):
            if line.startswith(f"{best}.valu_fp64_flop_per_launch:"):
                exec_valu = float(line.split(":")[1])
                counters_src = f"rocprofv3 PMC, {os.path.relpath(cprof, ROOT)}"
            elif line.startswith(f"{best}.mfma_f64_flop_per_launch:"):
                exec_mfma = float(line.split(":")[1])
            elif line.startswith(f"{best}.mfma_busy_frac:"):
                mfma_busy = float(line.split(":")[1])
    if exec_valu is not None:
        # FP64 work the kernel executes: VALU FP64 instructions plus FP64 matrix-core operations
        executed = exec_valu + (exec_mfma or 0.0)
    describe = {
        "fp64_exact": "batched FP64, exact core, persistent kernel",
        "fp64_learned": f"batched FP64, core presized by the BF16 rank MLP (margin {args.learned_margin}, accepted when "
                        f"the dropped part <= {LEARNED_TOL} * accuracy, else exact core), persistent kernel",
        "fp64_rule": f"batched FP64, core truncated where the pivoted R factor's dropped part <= {LEARNED_TOL} * "
                     f"accuracy, persistent kernel",
        "mixed_rule": "batched, factor products on FP32 matrix cores with FP32 accumulation over error-free slices "
                      f"of the FP64 operands, QR/SVD/truncation FP64, core truncated as fp64_rule (tol {LEARNED_TOL}), "
                      "persistent kernel",
        "mixed_learned": "batched, mixed-precision factor products as mixed_rule, core presized by the BF16 rank MLP "
                         f"(margin {args.learned_margin}, tol {LEARNED_TOL}, else exact core), persistent kernel",
    }

    def form_json(k, f):
        d = {"updates_per_s": round(f["ups"], 2), "ms_per_step": round(f["ms"], 3),
             "kernel_ms": round(f["kernel_ms"], 3), "rel_err_vs_fp64_oracle_max": f["err"],
             "rank_mismatches_vs_oracle": f"{f['rank_mis']}/{ncpu}", "meets_1e-8": f["err"] <= 1e-8,
             "max_jacobi_sweeps": f["sweeps"], "form": describe[k]}
        if "fallbacks" in f:
            d["fallbacks"] = f"{f['fallbacks']}/{args.updates}"
            d["mean_abs_rank_prediction_error"] = round(f["mae"], 3)
        return d

    out = {
        "metric": "TLR tile GEMM+recompression updates/s at rel err <=1e-8 (batched vs plain FP64)",
        "value": round(hb["ups"], 2),
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(hb["ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"batched CCC tile updates, {args.updates} per step, nb={NB}, ranks log-uniform "
                        f"[{LO},{HI}], decay {DECAY}, max_rank {MAX_RANK}, acc {ACC}",
            "form": describe[best] + (" in hipGraph" if use_graph else ""),
            "form_name": best,
            "global_batch": args.updates,
            "parallelism": "single GPU",
            "launch_shape": shape,
        },
        "plain_fp64_updates_per_s": round(plain_full_est, 2),
        "speedup_vs_plain_fp64": round(hb["ups"] / plain_full_est, 2),
        "rel_err_vs_fp64_oracle_max": hb["err"],
        f"rel_err_vs_dense_fp64_max_{ref_form}": float(max(errs_dense)),
        "rank_mismatches_vs_oracle": f"{hb['rank_mis']}/{ncpu}",
        "rel_err_checked_updates": len(orc_dense),
        "forms": {k: form_json(k, f) for k, f in forms.items()},
        "plain_tlr_nb1024": plain_tlr,
        "plain_tlr_4x4_nb256": plain_tlr4,
        "tlr_cholesky_nb512": chol,
        "roofline": {
            "bound": "mfma",
            "bound_note": "FP64 compute roof (dense FP64 matrix-core peak); the kernel's HBM traffic is "
                          "traffic_gb_per_s against 8000 GB/s, so neither roof is reached: the limit is the "
                          "latency of the dependent Householder/Jacobi steps per update",
            "achieved": round(achieved_tflops, 4),
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved_tflops / PEAK_FP64_TFLOPS, 5),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "traffic_gb_per_s": round(traffic / (hb["kernel_ms"] * 1e-3) / 1e9, 1) if traffic else None,
            "executed_fp64_flop_per_launch": executed,
            "executed_fp64_valu_flop_per_launch": exec_valu,
            "executed_fp64_mfma_flop_per_launch": exec_mfma,
            "executed_fp64_tflops": round(executed / (hb["kernel_ms"] * 1e-3) / 1e12, 4) if executed else None,
            "mfma_busy_frac": mfma_busy,
            "counters_source": counters_src,
            "kernel": "k_batched (persistent batched recompression)",
            "kernel_ms": round(hb["kernel_ms"], 3),
            "alg_flops_per_launch": flops,
            "alg_bytes_per_launch": traffic_alg,
        },
        "cpu_baseline": {
            "value": round(cpu_full, 3),
            "unit": "updates/s",
            "cores": cores,
            "kind": "port",
            "host": dict(host_info(), lapack=O.lapack_path(),
                         threads_rule="affinity mask capped at OMP_NUM_THREADS (the GPU host's CPU share)"),
            "tile_gemm_legs": cpu_legs,
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
