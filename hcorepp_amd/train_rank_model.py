"""Train the rank predictor of the batched recompression on recompressions the method performs.

Generates batches of independent CCC updates (nb = 512, ranks log-uniform in [8, 192], decay 0.85,
accuracy 1e-8: the batched benchmark's distribution, with other seeds), runs each batch through
the exact FP64 batched form on the GPU, collects the device-computed features and the exact
post-truncation ranks, trains the MLP with the library's FP32 kernels, and writes the parameters
to hcorepp_amd/rank_mlp.npz (with validation statistics).

    python -m hcorepp_amd.train_rank_model [--batches 8] [--per-batch 4096] [--epochs 120]

Collection can be split over several processes: `--collect-only FILE --first-batch K` writes the
features and ranks of batches K .. K+batches-1 to FILE (no training), and `--from FILE [FILE ...]`
trains on collected files (the last batch of the last file is the validation set).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

MODEL_PATH = os.path.join(HERE, "rank_mlp.npz")
NB, LO, HI, DECAY, MAX_RANK, ACC = 512, 8, 192, 0.85, 384, 1e-8
TRAIN_SEED0 = 7100


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def collect(ctx, n, seed):
    """Features and exact ranks of one batch of n updates."""
    import hcorepp_amd as H
    from hcorepp_amd.workloads import batch_ccc
    ups = batch_ccc(n, NB, seed, LO, HI, DECAY, threads=min(16, os.cpu_count() or 8))
    As = [H.CompressedTile(u["au"], u["av"]) for u in ups]
    Bs = [H.CompressedTile(u["bu"], u["bv"]) for u in ups]
    arena = H.TileArena(n * H.TileArena.compressed_bytes(NB, NB, MAX_RANK))
    Cs = [H.CompressedTile(u["cu"], u["cv"], max_rank=MAX_RANK, arena=arena) for u in ups]
    b = H.Batch(ctx, As, Bs, Cs, 1.0, 1.0, H.CompressionParameters(ACC), H.PREC_FP64)
    b.run()
    st, _ = b.status()
    if st != 0:
        raise SystemExit(f"batch status {st}")
    X, r = b.features(), b.ranks()
    b.close()
    return X, r.astype(np.float64)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--batches", type=int, default=8)
    ap.add_argument("--per-batch", type=int, default=4096)
    ap.add_argument("--epochs", type=int, default=120)
    ap.add_argument("--out", default=MODEL_PATH)
    ap.add_argument("--first-batch", type=int, default=0, help="index of the first batch (seed offset)")
    ap.add_argument("--collect-only", default=None, help="write features and ranks here, do not train")
    ap.add_argument("--from", dest="sources", nargs="+", default=None, help="train on collected files")
    args = ap.parse_args()
    import hcorepp_amd as H
    from hcorepp_amd.learned import RankPredictor
    ctx = H.RunContext(0)
    Xs, ys = [], []
    t0 = time.time()
    if args.sources:
        for f in args.sources:
            with np.load(f) as z:
                for k in range(int(z["batches"])):
                    Xs.append(z[f"x{k}"])
                    ys.append(z[f"y{k}"])
        log(f"[train] {len(Xs)} collected batches from {len(args.sources)} file(s)")
    else:
        # the last batch is held out for validation (collection: simply the last one written)
        nb_ = args.batches if args.collect_only else args.batches + 1
        for k in range(nb_):
            X, r = collect(ctx, args.per_batch, TRAIN_SEED0 + args.first_batch + k)
            Xs.append(X)
            ys.append(np.log2(r))
            log(f"[train] batch {args.first_batch + k}: {args.per_batch} recompressions, ranks "
                f"{int(r.min())}-{int(r.max())} ({time.time() - t0:.1f}s)")
        if args.collect_only:
            np.savez(args.collect_only, batches=len(Xs), first_batch=args.first_batch,
                     **{f"x{k}": x for k, x in enumerate(Xs)}, **{f"y{k}": y for k, y in enumerate(ys)})
            log(f"[train] wrote {args.collect_only}")
            ctx.close()
            return
    Xtr, ytr = np.concatenate(Xs[:-1]), np.concatenate(ys[:-1])
    Xva, yva = Xs[-1], ys[-1]
    model = RankPredictor(ctx, seed=1)
    t1 = time.time()
    hist = model.train(Xtr, ytr, epochs=args.epochs, batch=4096, lr=2e-3, log=log)
    pred = model.predict_log2(Xva)  # the in-kernel BF16 inference
    err = np.abs(np.exp2(pred) - np.exp2(yva))
    under = np.mean(np.ceil(np.exp2(pred)) + 8 < np.exp2(yva))
    log(f"[train] {len(ytr)} samples, {args.epochs} epochs in {time.time() - t1:.1f}s; validation: rank MAE "
        f"{err.mean():.2f}, median {np.median(err):.1f}, under-predicted by > 8: {100 * under:.1f}%")
    model.save(args.out, train_samples=len(ytr), epochs=args.epochs, final_train_mse=hist[-1],
               val_rank_mae=float(err.mean()), val_under_margin=float(under), seed0=TRAIN_SEED0)
    log(f"[train] wrote {args.out}")
    ctx.close()


if __name__ == "__main__":
    main()
