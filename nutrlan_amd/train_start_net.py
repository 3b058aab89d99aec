"""Trains the learned start-vector net and writes the shipped weights
(nutrlan_amd/data/trlnet_anderson2d.npy, with a log beside it).

Training data is made by the solver itself: `count` periodic 2D Anderson
lattices (L = 64, hopping -1, on-site U[-W/2, W/2], W ~ U[1, 12]) are solved
for their 8 lowest eigenpairs by the batched mixed-precision solver
(trlan_batch), and the net learns sqrt(sum_k psi_k^2) of those states from
the potential (nutrlan_amd/learned.py).  All of it runs on one GPU:

    python -m nutrlan_amd.train_start_net --count 20000 --epochs 12
"""
import argparse
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from nutrlan_amd import learned as LN  # noqa: E402
else:
    from . import learned as LN


def make_set(count, L, seed, chunk=1024, log=print, kind="density"):
    rng = np.random.default_rng(seed)
    X = np.zeros((count, L, L), np.float32)
    Y = np.zeros((count, L, L), np.float32)
    for s in range(0, count, chunk):
        k = min(chunk, count - s)
        models = [LN.lattice(L, float(rng.uniform(1.0, 12.0)), seed * 1000003 + s + j) for j in range(k)]
        t0 = time.perf_counter()
        tg, res = LN.solver_targets(models, seed=seed + s, kind=kind)
        log(f"solved {s + k}/{count} lattices ({time.perf_counter() - t0:.1f} s, "
            f"{sum(r['matvec'] for r in res)} operator applications)")
        for j, ((csr, d), t) in enumerate(zip(models, tg)):
            X[s + j] = d * LN.SCALE
            Y[s + j] = t
    return X, Y


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=20000)
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=LN.WEIGHTS)
    ap.add_argument("--target", choices=["density", "signed"], default="density",
                    help="density: sqrt(sum_k psi_k^2); signed: sum_k psi_k, each psi_k's "
                         "largest entry positive")
    args = ap.parse_args()
    lines = []

    def log(msg):
        print(msg, flush=True)
        lines.append(msg)

    t0 = time.perf_counter()
    X, Y = make_set(args.count, args.L, args.seed, log=log, kind=args.target)
    log(f"training set: {args.count} lattices {args.L}x{args.L} in {time.perf_counter() - t0:.1f} s")
    net = LN.Net(seed=args.seed)
    t1 = time.perf_counter()
    hist = []
    for ep in range(args.epochs):
        lr = args.lr * (0.5 if ep >= args.epochs * 2 // 3 else 1.0)
        hist += LN.train(net, X, Y, epochs=1, batch=args.batch, lr=lr, seed=args.seed + ep)
        log(f"epoch {ep} loss {hist[-1]:.4f} ({time.perf_counter() - t1:.1f} s)")
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    np.save(args.out, net.params())
    with open(os.path.splitext(args.out)[0] + ".txt", "w") as f:
        f.write("command: python -m nutrlan_amd.train_start_net " + " ".join(sys.argv[1:]) + "\n")
        f.write("\n".join(lines) + "\n")
    log(f"wrote {args.out}")
    net.close()


if __name__ == "__main__":
    main()
