"""Train the learned troubled-stencil switch on the GPU.

1. Harvest: seeded four-quadrant problems (rho ~ U[0.1, 2], p ~ U[0.05, 2],
   u, v ~ U[-1.5, 1.5] per quadrant) at N^2, WENO-Z RK3 CFL 0.6, run to
   t_end; at each snapshot time every face of both directions gives 4 samples
   (one per characteristic field: 10 normalised stencil values, label
   max |omega - d| > 0.2), kept with probability keep_pos / keep_neg.
   Training seeds 0 .. seeds-val-1, validation seeds after them.
2. Train: the library's FP32 forward/backward/Adam kernels, weighted BCE.
3. Threshold: the logit at which the deployed bf16 MFMA network keeps the
   requested recall on the validation positives.

Writes <out>.npz (params float32[1441], threshold) and <out>.json (settings,
loss curve, validation counts).

    python switch_train.py                       # 256 seeds at 512^2 (the default)
    python switch_train.py --seeds 16 --n 128    # a quick run
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import wenofd as W  # noqa: E402

DEFAULT_OUT = os.path.join(HERE, "switch", "switch_mlp")


def log(msg):
    print(f"[switch_train] {msg}", flush=True)


def harvest_run(m, seed, n, t_end, snapshots, keep_pos, keep_neg):
    cfg = W.config_text(TESTCASE=5, SCHEME=1, RKMETH=2, CFL=0.6, ELEMNUMX=n, ELEMNUMY=n, QUAD_SEED=seed,
                        TEND="%.17g" % t_end)
    added = 0
    with W.Solver(cfg) as s:
        for k, ts in enumerate(snapshots):
            try:
                while s.info().now < ts - 1e-12:
                    if s.advance(10) == 0:
                        break
            except W.WenofdError as e:  # the plain solver itself failed on this problem: keep what we have
                if e.code != W.E_NAN_DT:
                    raise
                log(f"seed {seed}: NaN time step before t = {ts}; the run stops here")
                break
            added += m.harvest(s, keep_pos, keep_neg, seed=seed * 16 + k)
    return added


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=256)
    ap.add_argument("--val", type=int, default=16, help="the last VAL seeds are held out for validation")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--t-end", type=float, default=0.25)
    ap.add_argument("--snapshots", default="0.05,0.15,0.25")
    ap.add_argument("--keep-pos", type=float, default=0.25)
    ap.add_argument("--keep-neg", type=float, default=0.02)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--recall", type=float, default=0.99)
    ap.add_argument("--init-seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args(argv)
    snaps = [float(v) for v in args.snapshots.split(",")]
    t0 = time.perf_counter()
    train = W.Mlp(seed=args.init_seed)
    val = W.Mlp(seed=args.init_seed)
    ntrain = args.seeds - args.val
    for seed in range(args.seeds):
        target = train if seed < ntrain else val
        harvest_run(target, seed, args.n, args.t_end, snaps, args.keep_pos, args.keep_neg)
        if seed % 16 == 15 or seed == args.seeds - 1:
            log(f"harvested seeds 0..{seed}: train {train.size()}, validation {val.size()} "
                f"(samples, positives), {time.perf_counter() - t0:.1f} s")
    n, pos = train.size()
    pw = float((n - pos) / max(pos, 1))
    log(f"training on {n} samples ({pos} positive), pos_weight {pw:.3f}")
    losses = list(map(float, train.train(args.epochs, batch=args.batch, lr=args.lr, pos_weight=pw, seed=1)))
    losses += list(map(float, train.train(max(1, args.epochs // 3), batch=args.batch, lr=args.lr / 3,
                                          pos_weight=pw, seed=2)))
    log("epoch losses " + " ".join(f"{v:.4f}" for v in losses))
    P = train.params
    val.params = P
    Xv, Yv = val.samples()
    zv = val.infer(Xv)
    zpos = np.sort(zv[Yv == 1])
    thr = float(zpos[int(np.floor((1.0 - args.recall) * len(zpos)))] - 1e-6) if len(zpos) else 0.0
    counts = val.evaluate(threshold=thr)
    counts0 = val.evaluate(threshold=0.0)
    log(f"validation at threshold {thr:.4f}: {counts}; at 0: {counts0}")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.savez(args.out + ".npz", params=P, threshold=np.float32(thr))
    meta = {"seeds": args.seeds, "validation_seeds": args.val, "n": args.n, "t_end": args.t_end, "snapshots": snaps,
            "keep_pos": args.keep_pos, "keep_neg": args.keep_neg, "epochs": len(losses), "batch": args.batch,
            "lr": args.lr, "pos_weight": pw, "train_samples": int(n), "train_positives": int(pos),
            "validation_samples": int(len(Yv)), "validation_positives": int(Yv.sum()), "recall_target": args.recall,
            "threshold": thr, "validation_counts": counts, "validation_counts_at_0": counts0, "epoch_losses": losses,
            "seconds": time.perf_counter() - t0}
    with open(args.out + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    log(f"wrote {args.out}.npz / .json in {meta['seconds']:.1f} s")
    train.close()
    val.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
