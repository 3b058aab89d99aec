"""Throughput of the device-resident thick-restart Lanczos solver on one GPU.

Workload (one step = one complete trlan() solve, CSR resident in HBM):
  periodic 3D 7-point Laplacian on 128^3 (n = 2,097,152, nnz = 14,680,064),
  off-diagonal -1, diagonal 6 + U[0,1) (seed 0), FP64 CSR on the device;
  nev = 16 largest, maxlan = 64, tol = 1e-8, restart scheme 1,
  start vector U[0,1) (seed 1).
value = eigenpairs with ||A v - lambda v|| <= 1e-8 ||A|| (checked on the
host after the timed solves) per second, for the path given by --path
(default: mixed precision); the other path is timed the same way and
reported beside it, with the ratio and the largest eigenvalue difference.
Each solve is synchronous (trlan() returns after its last device
synchronisation), so the wall clock around the K solves brackets all GPU
work.

Prints one JSON line, including
  roofline: the dominant kernel's algorithmic bytes / HIP-event time vs the
            8 TB/s HBM peak, plus the SpMV + recurrence + Gram-Schmidt
            aggregate, and the rocprofv3-measured HBM traffic per launch;
  cpu_baseline: the compiled reference trlan (oracle/_ref, OpenBLAS +
            OpenMP CSR operator) on a bounded sample of the same solve.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
L = 128
NEV, MAXLAN, TOL, RESTART, MAXMV = 16, 64, 1e-8, 1, 20000
STATS_EVERY = 7
# the headline metric; this line measures its plain-FP64 side on the 128^3
# Laplacian (config "path" says which solver path ran)
METRIC = "eigenpairs/sec at resid<=1e-8||A||: mixed+learned-start vs plain FP64, 1 GPU"


# device kernel behind each timed kernel class, per solver path
KERNEL_NAMES = {
    "plain": {"proj": "k_proj<double, double, 1, 8, 4>", "update": "k_update<double, double, 1, 4>",
              "spmv": "k_spmv<double, double, 1, true>"},
    "mixed": {"proj": "k_proj<float, double, 1, 8, 4>", "update": "k_update<float, double, 1, 4>",
              "spmv": "k_spmv<double, double, 1, true>"},
}


def calibration_copy(T, n):
    """One plain copy of n doubles on the device (k_scale_copy<double>: 8n
    bytes read, 8n written), the first kernel dispatch of the process, so a
    rocprofv3 --pmc run of this script holds a kernel of known traffic to
    scale FETCH_SIZE / WRITE_SIZE by."""
    import numpy as np

    c = T.Ctx(n, 2)
    c.set(0, np.ones(n))
    c.scale_copy(1, 0, 1.0)
    assert c.get(1)[-1] == 1.0
    c.close()


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_laplace3d_128_fp64.txt")
    if kernel is None or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch ") and line.split(" ", 1)[1].rsplit(" ", 1)[0] == kernel:
            return int(line.split()[-1])
    return None


def build_problem():
    from nutrlan_amd import workloads as W

    rp, ci, v = W.laplace3d_csr(L, seed=0)
    x0 = W.start_vector(L ** 3, seed=1)
    return rp, ci, v, x0


def residuals(rp, ci, v, evec, ev):
    import scipy.sparse as sp

    n = len(rp) - 1
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    R = A @ evec - evec * ev[None, :]
    return (R * R).sum(axis=0) ** 0.5


def cpu_baseline(sample_matvec, full_matvec, nthreads):
    """Reference trlan on the same problem capped at `sample_matvec`
    operator applications, in a child process (OpenBLAS thread count is
    fixed when the library loads).  Returns (eigenpairs/s extrapolated to a
    full solve, seconds, sample description) or None."""
    from oracle import ref

    if not ref.available():
        return None
    code = (
        "import sys, json; sys.path.insert(0, %r)\n"
        "import bench\n"
        "from oracle import ref\n"
        "rp, ci, v, x0 = bench.build_problem()\n"
        "r = ref.solve_csr(rp, ci, v, 1, bench.NEV, bench.MAXLAN, bench.TOL, bench.RESTART, %d,"
        " bench.NEV, x0, nthreads=%d)\n"
        "print(json.dumps(dict(wall=r['wall_s'], matvec=r['matvec'], north=r['north'],"
        " stat=r['stat'])))\n" % (ROOT, sample_matvec, nthreads))
    env = dict(os.environ, OPENBLAS_NUM_THREADS=str(nthreads), OMP_NUM_THREADS=str(nthreads))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         timeout=600)
    if out.returncode != 0:
        print(out.stderr[-2000:], file=sys.stderr)
        return None
    r = json.loads(out.stdout.strip().splitlines()[-1])
    per_mv = r["wall"] / max(r["matvec"], 1)
    full_s = per_mv * full_matvec
    return NEV / full_s, r["wall"], (
        f"reference trlan, first {r['matvec']} operator applications of the same solve "
        f"({r['wall']:.1f} s, {r['north']} Gram-Schmidt passes); eigenpairs/s extrapolated to "
        f"the full solve's {full_matvec} applications")


def leg(T, A, x0, steps, rp, ci, v, stats_every=0):
    """Time `steps` complete solves on operator A (inputs resident); returns
    (seconds, last result, eigenpairs at resid <= 1e-8 ||A||, stats)."""
    def solve():
        return T.solve_csr(A, 1, NEV, MAXLAN, TOL, RESTART, MAXMV, NEV, x0)

    if stats_every:
        T.stats_enable(True, every=stats_every)
    t0 = time.perf_counter()
    outs = [solve() for _ in range(steps)]
    t1 = time.perf_counter()
    stats = T.stats_get() if stats_every else None
    if stats_every:
        T.stats_enable(False)
    g = outs[-1]
    assert g["stat"] == 0, g["stat"]
    nec = g["nec"]
    res = residuals(rp, ci, v, g["evec"][:, :nec], g["eval"][:nec])
    good = int((res <= 1e-8 * g["anrm"]).sum())
    return t1 - t0, g, good, stats


def batched_anderson2d(args):
    """The batched workload: 1024 independent periodic 2D Anderson lattices,
    L ~ U{64..256}, hopping -1, on-site U[-W/2, W/2] with W ~ U[1, 12]
    (seeded), 8 lowest eigenpairs each, maxlan 32, tol 1e-8.  value = the
    mixed-precision batched solver started from the learned net (inference
    inside the timed region); beside it the plain FP64 solver one lattice at
    a time (timed on a sample of the same lattices), the plain FP64 batched
    solver and the mixed batched solver from random starts."""
    import numpy as np
    import scipy.sparse as sp

    from nutrlan_amd import learned as LN
    from nutrlan_amd import trlan as T
    from nutrlan_amd import workloads as W

    rng = np.random.default_rng(5)
    count = args.models
    Ls = rng.integers(64, 257, count)
    Ws = rng.uniform(1.0, 12.0, count)
    models = [LN.lattice(int(L), float(w), 7000 + i) for i, (L, w) in enumerate(zip(Ls, Ws))]
    nev, maxlan, tol = LN.NEV, 32, 1e-8
    starts = [W.start_vector(len(c[0]) - 1, seed=9000 + i) for i, (c, d) in enumerate(models)]
    plain = [T.Csr(*c) for c, d in models]
    mixed = [T.Csr(*c, mixed=True) for c, d in models]
    net = LN.load_net()
    assert net is not None, "learned-start weights missing (run nutrlan_amd/train_start_net.py)"

    def good_pairs(res, mods=models):
        g = 0
        for (c, d), r in zip(mods, res):
            rp, ci, v = c
            n = len(rp) - 1
            A = sp.csr_matrix((v, ci, rp), shape=(n, n))
            X = r["evec"][:, :r["nec"]]
            R = A @ X - X * r["eval"][:r["nec"]]
            g += int(((R * R).sum(axis=0) ** 0.5 <= 1e-8 * r["anrm"]).sum())
        return g

    def batch(ops, x0s):
        probs = [(A, -1, nev, maxlan, tol, 1, 100000, nev, x) for A, x in zip(ops, x0s)]
        t0 = time.perf_counter()
        rc, res, _ = T.solve_batch(probs)
        el = time.perf_counter() - t0
        assert rc == 0, rc
        return el, res

    batch(mixed[:64], starts[:64])  # warm-up
    # headline: learned start + mixed precision, batched
    t0 = time.perf_counter()
    x_net = LN.start_vectors(net, [d for c, d in models])
    t_inf = time.perf_counter() - t0
    el, res = batch(mixed, x_net)
    el += t_inf
    good = good_pairs(res)
    mv_net = sum(r["matvec"] for r in res)
    el_r, res_r = batch(mixed, starts)
    good_r = good_pairs(res_r)
    mv_rand = sum(r["matvec"] for r in res_r)
    el_p, res_p = batch(plain, starts)
    good_p = good_pairs(res_p)
    # plain FP64, one lattice at a time, on a sample
    ns = min(args.plain_sample, count)
    t0 = time.perf_counter()
    res_s = [T.solve_csr(plain[i], -1, nev, maxlan, tol, 1, 100000, nev, starts[i]) for i in range(ns)]
    el_s = time.perf_counter() - t0
    g_s = good_pairs(res_s, models[:ns])
    plain_single = g_s / el_s
    value = good / el
    for A in plain + mixed:
        A.close()
    net.close()
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "eigenpairs/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": round(1e3 * el, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "anderson2d_batch_%d" % count, "L": "U{64..256}",
                   "W": "U[1,12]", "sites_total": int(sum(len(c[0]) - 1 for c, d in models)),
                   "nev": nev, "lohi": -1, "maxlan": maxlan, "tol": tol,
                   "path": "batched, mixed precision, learned start",
                   "converged_at_1e-8": good, "matvec_total": mv_net,
                   "learned_start_inference_s": round(t_inf, 3)},
        "mixed_batched_random_start": {"value": round(good_r / el_r, 1), "s": round(el_r, 3),
                                       "matvec_total": mv_rand, "converged_at_1e-8": good_r},
        "plain_batched_random_start": {"value": round(good_p / el_p, 1), "s": round(el_p, 3),
                                       "converged_at_1e-8": good_p},
        "plain_fp64_one_at_a_time": {"value": round(plain_single, 2),
                                     "sample": f"first {ns} lattices, {el_s:.2f} s"},
        "vs_plain_fp64": round(value / plain_single, 2),
        "learned_vs_random_matvec": round(mv_net / mv_rand, 4),
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--path", choices=["mixed", "plain"], default="mixed",
                    help="solver path measured for `value` (the other is reported beside it)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=96)
    ap.add_argument("--workload", choices=["laplace3d", "anderson2d-batch"], default="laplace3d",
                    help="laplace3d: the 128^3 Laplacian (default line); anderson2d-batch: "
                         "1024 uneven 2D Anderson lattices through the batched solver")
    ap.add_argument("--models", type=int, default=1024)
    ap.add_argument("--plain-sample", type=int, default=64)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}): the multi-GPU "
              "row-block sharding is a design, not built here.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from nutrlan_amd import trlan as T

    assert T.device_count() > 0, "no GPU"
    if args.workload == "anderson2d-batch":
        batched_anderson2d(args)
        return
    rp, ci, v, x0 = build_problem()
    n = len(rp) - 1
    calibration_copy(T, n)
    ops = {"plain": T.Csr(rp, ci, v), "mixed": T.Csr(rp, ci, v, mixed=True)}
    other = "plain" if args.path == "mixed" else "mixed"
    for _ in range(args.warmup):
        for k in ops:
            T.solve_csr(ops[k], 1, NEV, MAXLAN, TOL, RESTART, MAXMV, NEV, x0)
    # the path beside the headline one, timed on its own (no events)
    o_el, o_g, o_good, _ = leg(T, ops[other], x0, args.steps, rp, ci, v)
    # headline path: one launch in STATS_EVERY of each kernel kind (a hashed,
    # non-periodic pick) carries HIP start/stop events stamped from its own
    # dispatch on the solver's stream; timing every launch would cost ~10%.
    elapsed, g, good, stats = leg(T, ops[args.path], x0, args.steps, rp, ci, v, STATS_EVERY)
    value = good * args.steps / elapsed
    other_value = o_good * args.steps / o_el
    nec = g["nec"]
    dl = float(np.abs(np.sort(g["eval"][:nec]) - np.sort(o_g["eval"][:o_g["nec"]])).max()
               / g["anrm"]) if o_g["nec"] == nec else None

    # roofline: dominant kernel by time, plus the path aggregate
    dom = max(stats, key=lambda k: stats[k][1])
    launches, ms, byts = stats[dom]
    achieved = byts / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    path = ["spmv", "recur", "proj", "update"]
    pb = sum(stats[k][2] for k in path)
    pt = sum(stats[k][1] for k in path)
    path_gbs = pb / (pt * 1e-3) / 1e9 if pt > 0 else 0.0
    roofline = {
        "bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(KERNEL_NAMES[args.path].get(dom)),
        "avg_launch_us": round(ms * 1e3 / max(launches, 1), 2),
        "bytes_per_launch": round(byts / max(launches, 1)),
        "path_spmv_recur_cgs": {"achieved": round(path_gbs, 1),
                                "frac": round(path_gbs / HBM_PEAK_GBS, 4)},
        "kernels": {k: {"launches": s[0], "ms": round(s[1], 3),
                        "GBps": round(s[2] / (s[1] * 1e-3) / 1e9, 1) if s[1] > 0 else 0.0}
                    for k, s in stats.items()},
    }
    cpu = None
    if not args.no_cpu_baseline:
        nthreads = min(os.cpu_count() or 1, 16)
        cb = cpu_baseline(args.cpu_sample, o_g["matvec"] if args.path == "mixed" else g["matvec"],
                          nthreads)
        if cb is not None:
            cpu = {"value": round(cb[0], 4), "unit": "eigenpairs/s", "cores": nthreads,
                   "kind": "reference", "sample": cb[2]}
    desc = {"mixed": "mixed: FP32 shadow basis for Gram-Schmidt, FP64 SpMV/recurrence/basis/"
                     "restart GEMM, FP64 Rayleigh-quotient refinement",
            "plain": "plain FP64 device basis"}
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "laplace3d_128_periodic_fp64", "n": n, "nnz": int(len(ci)),
                   "nev": NEV, "lohi": 1, "maxlan": MAXLAN, "tol": TOL, "restart": RESTART,
                   "path": desc[args.path], "learned_start": "none (random start; the learned "
                   "start applies to the batched 2D Anderson models)",
                   "matvec": g["matvec"], "nloop": g["nloop"], "north": g["north"],
                   "converged": nec, "converged_at_1e-8": good},
        other + "_path": {"value": round(other_value, 3),
                          "ms_per_step": round(1e3 * o_el / args.steps, 2),
                          "matvec": o_g["matvec"], "converged_at_1e-8": o_good,
                          "path": desc[other]},
        "mixed_vs_plain": round((value / other_value) if args.path == "mixed"
                                else (other_value / value), 3),
        "max_eigenvalue_diff_vs_" + other: dl,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    for A in ops.values():
        A.close()


if __name__ == "__main__":
    main()
