"""Trains the learned start-vector net and writes the shipped weights
(nutrlan_amd/data/trlnet_anderson2d.npy, with a log beside it).

Training data is made by the solver itself: periodic 2D Anderson lattices
(hopping -1, on-site U[-W/2, W/2], W ~ U[1, 12]) are solved for their 8
lowest eigenpairs by the batched mixed-precision solver (trlan_batch), and
the net learns sqrt(sum_k psi_k^2) of those states from the potential
(nutrlan_amd/learned.py).  The net is convolutional, so one set of weights
serves every lattice size; --sizes trains on several (count at the first
size, and the same number of lattice sites at each further size), each
epoch visiting every size.  All of it runs on one GPU:

    python -m nutrlan_amd.train_start_net --count 20000 --epochs 12
    python -m nutrlan_amd.train_start_net --sizes 64,96,128,192,256 --count 12000
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
    ap.add_argument("--sizes", default=None,
                    help="comma-separated lattice sizes (overrides --L): --count lattices at the "
                         "first, as many sites at each further size")
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
    sizes = [int(x) for x in args.sizes.split(",")] if args.sizes else [args.L]
    sets = []
    for i, L in enumerate(sizes):
        cnt = max(1, args.count * sizes[0] ** 2 // (L * L))
        X, Y = make_set(cnt, L, args.seed + 7919 * i, log=log, kind=args.target)
        sets.append((L, X, Y))
        log(f"training set: {cnt} lattices {L}x{L} ({time.perf_counter() - t0:.1f} s so far)")
    net = LN.Net(seed=args.seed)
    t1 = time.perf_counter()
    hist = []
    for ep in range(args.epochs):
        lr = args.lr * (0.5 if ep >= args.epochs * 2 // 3 else 1.0)
        for i, (L, X, Y) in enumerate(sets):
            # the same number of sites per step at every size
            b = max(1, args.batch * sizes[0] ** 2 // (L * L))
            hist += LN.train(net, X, Y, epochs=1, batch=b, lr=lr, seed=args.seed + 31 * ep + i)
            log(f"epoch {ep} L {L} loss {hist[-1]:.4f} ({time.perf_counter() - t1:.1f} s)")
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    np.save(args.out, net.params())
    with open(os.path.splitext(args.out)[0] + ".txt", "w") as f:
        f.write("command: python -m nutrlan_amd.train_start_net " + " ".join(sys.argv[1:]) + "\n")
        f.write("\n".join(lines) + "\n")
    log(f"wrote {args.out}")
    net.close()


if __name__ == "__main__":
    main()
