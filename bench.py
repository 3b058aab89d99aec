"""Throughput of the device-resident thick-restart Lanczos solver on one GPU.

Default workload (--workload anderson2d-batch): 1024 independent periodic
2D Anderson lattices, L ~ U{64..256} (n = 4096..65536), hopping -1, on-site
U[-W/2, W/2] with W ~ U[1, 12], 8 lowest eigenpairs each, maxlan 32,
tol 1e-8, CSR operators resident in HBM.  One step = learned-start
inference for every lattice + one batched mixed-precision solve of all of
them (trlan_batch).  value = eigenpairs with ||A v - lambda v|| <=
1e-8 ||A|| (checked on the host after the timed steps) per second.  Beside
it: the same batch from random starts, the plain FP64 batched solver, and
the plain FP64 solver one lattice at a time (vs_plain_fp64 = value / that).

--workload laplace3d: periodic 3D 7-point Laplacian on 128^3 (n =
2,097,152, nnz = 14,680,064), diagonal 6 + U[0,1), nev = 16 largest,
maxlan = 64, tol = 1e-8, restart 1; one step = one complete solve, the
mixed-precision path (--path) with plain FP64 timed beside it.

Prints one JSON line, including
  roofline: the dominant kernel's algorithmic bytes / HIP-event time (a
            hashed sample of launches, events stamped from each launch's
            own dispatch) vs the 8 TB/s HBM peak, the SpMV + recurrence +
            Gram-Schmidt aggregate, and the rocprofv3-measured HBM traffic
            per launch from the committed PMC summary;
  cpu_baseline: the compiled reference trlan (oracle/_ref) on a bounded
            sample of the same workload on the host cores.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6  # dense FP64 matrix-core peak of one MI355X
L = 128
NEV, MAXLAN, TOL, RESTART, MAXMV = 16, 64, 1e-8, 1, 20000
STATS_EVERY = 7
# the headline metric; this line measures its plain-FP64 side on the 128^3
# Laplacian (config "path" says which solver path ran)
METRIC = "eigenpairs/sec at resid<=1e-8||A||: mixed+learned-start vs plain FP64, 1 GPU"


def kernel_names(T, path, A=None):
    """Device kernel(s) behind each timed kernel class of a solver path, as
    rocprofv3 names the dispatches (the projection's columns per block come
    from the library's current tile setting; the SpMV from operator A's
    format: column panels run two kernels per product, joined by '+')."""
    cb64, cb32 = T.proj_tiles_get()
    if path == "complex":
        return {"proj": "k_proj<double, double, 2, 4, 4>", "update": "k_update<double, double, 2, 4>",
                "spmv": "k_spmv<double, double, 2, true>"}
    vs, cb = ("float", cb32) if path == "mixed" else ("double", cb64)
    # real operators up to 2^20 rows: the recurrence runs fused with the
    # first projection (k_recproj, over the FP32 shadow or the FP64 basis),
    # timed as the projection
    fused = A is not None and A.n <= 1 << 20
    spmv = "k_spmv<double, double, 1, true>"
    if A is not None and A.paneled:
        spmv = "k_spmv_panel<true>+k_spmv_panel_rows<true>"
    elif A is not None and A.stencil[0]:
        # up to 2048 row blocks of at most 256 rows: four row blocks per workgroup
        D, L = A.stencil
        spmv = f"k_spmv_st_rb<{D}, 4>" if L ** D <= 2048 * 256 else f"k_spmv_st<{D}, true>"
    elif A is not None and A.packed:
        spmv = "k_spmv_pk<true>"
    return {"proj": f"k_recproj<{cb}, {vs}>" if fused else f"k_proj<{vs}, double, 1, {cb}, 4>",
            "update": f"k_update<{vs}, double, 1, 4>", "spmv": spmv}


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


LAPLACE_TRAFFIC = "hbm_traffic_laplace3d_128_fp64.txt"
BATCH_TRAFFIC = "hbm_traffic_anderson2d_batch_1024.txt"


def measured_traffic(kernel, summary=LAPLACE_TRAFFIC):
    """HBM bytes per launch of `kernel` ('a+b': the sum over kernels a and
    b, dispatched once each) from a committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), or None (with a
    warning on stderr when the summary exists but does not list a kernel: a
    renamed kernel must not pass silently)."""
    path = os.path.join(ROOT, "profiles", summary)
    if kernel is None or not os.path.exists(path):
        return None
    found = {}
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch "):
            name, val = line.split(" ", 1)[1].rsplit(" ", 1)
            found[name] = int(val)
    total = 0
    for k in kernel.split("+"):
        if k not in found:
            print(f"bench.py: profiles/{summary} has no traffic line for kernel {k!r}",
                  file=sys.stderr, flush=True)
            return None
        total += found[k]
    return total


# Single-solve workloads (one step = one complete trlan()/ztrlan() solve):
# builder kwargs, start-vector seed, lohi, nev, maxlan, tol, restart, maxmv,
# complex, and the committed PMC summary for the roofline's traffic field.
SINGLE = {
    "laplace3d": dict(name="laplace3d_128_periodic_fp64", kind="laplace3d", x0_seed=1, lohi=1,
                      nev=16, maxlan=64, tol=1e-8, restart=1, maxmv=20000, cplx=False,
                      traffic=None),
    "diag1000": dict(name="diag_1000", kind="diag", x0_seed=1, lohi=1, nev=5, maxlan=40,
                     tol=1.4e-8, restart=1, maxmv=5000, cplx=False, traffic=None),
    "anderson3d-c": dict(name="anderson3d_96_peierls_c128", kind="anderson3d", x0_seed=1,
                         lohi=-1, nev=24, maxlan=80, tol=1e-8, restart=1, maxmv=40000,
                         cplx=True, traffic=None),
    "rand8m": dict(name="random_sym_8388608_deg16", kind="rand", x0_seed=5, lohi=1, nev=32,
                   maxlan=128, tol=1e-8, restart=1, maxmv=40000, cplx=False, traffic=None),
}
SINGLE["laplace3d"]["traffic"] = "hbm_traffic_laplace3d_128_fp64.txt"
SINGLE["anderson3d-c"]["traffic"] = "hbm_traffic_anderson3d_96_c128.txt"
SINGLE["rand8m"]["traffic"] = "hbm_traffic_rand8m.txt"


def single_problem(workload):
    """(row_ptr, col_idx, values, start vector) of a single-solve workload:
    laplace3d = periodic 7-point Laplacian on 128^3, diagonal 6 + U[0,1)
    (seed 0); diag1000 = diag(1..1000); anderson3d-c = complex Hermitian 96^3
    Anderson model, Peierls phase exp(2 pi i x / 16) on y-bonds, on-site
    U[-2, 2] (seed 0); rand8m = symmetric random sparse n = 2^23, row degree
    ~Poisson(16), values N(0,1)/4, diagonal U[-1,1] (seed 4)."""
    from nutrlan_amd import workloads as W

    sp = SINGLE[workload]
    if sp["kind"] == "laplace3d":
        rp, ci, v = W.laplace3d_csr(L, seed=0)
    elif sp["kind"] == "diag":
        rp, ci, v = W.diag_csr(1000)
    elif sp["kind"] == "anderson3d":
        rp, ci, v = W.anderson3d_csr(96, seed=0)
    else:
        rp, ci, v = W.random_sym_csr(1 << 23, seed=4)
    x0 = W.start_vector(len(rp) - 1, seed=sp["x0_seed"], complex_=sp["cplx"])
    return rp, ci, v, x0


def build_problem():
    """The 128^3 Laplacian workload (kept for callers of the older name)."""
    return single_problem("laplace3d")


def residuals(rp, ci, v, evec, ev):
    import scipy.sparse as sp

    n = len(rp) - 1
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    R = A @ evec - evec * ev[None, :]
    return np.sqrt((np.abs(R) ** 2).sum(axis=0))


def cpu_baseline(workload, sample_matvec, full_matvec, nthreads):
    """Reference trlan/ztrlan on the same problem capped at `sample_matvec`
    operator applications, in a child process (OpenBLAS thread count is
    fixed when the library loads).  Returns (eigenpairs/s extrapolated to a
    full solve, seconds, sample description) or None."""
    from oracle import ref

    if not ref.available():
        return None
    sp = SINGLE[workload]
    code = (
        "import sys, json; sys.path.insert(0, %r)\n"
        "import bench\n"
        "from oracle import ref\n"
        "sp = bench.SINGLE[%r]\n"
        "rp, ci, v, x0 = bench.single_problem(%r)\n"
        "r = ref.solve_csr(rp, ci, v, sp['lohi'], sp['nev'], sp['maxlan'], sp['tol'], sp['restart'],"
        " %d, sp['nev'], x0, nthreads=%d, complex_=sp['cplx'])\n"
        "print(json.dumps(dict(wall=r['wall_s'], matvec=r['matvec'], north=r['north'],"
        " stat=r['stat'])))\n" % (ROOT, workload, workload, sample_matvec, nthreads))
    env = dict(os.environ, OPENBLAS_NUM_THREADS=str(nthreads), OMP_NUM_THREADS=str(nthreads))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         timeout=900)
    if out.returncode != 0:
        print(out.stderr[-2000:], file=sys.stderr)
        return None
    r = json.loads(out.stdout.strip().splitlines()[-1])
    per_mv = r["wall"] / max(r["matvec"], 1)
    full_s = per_mv * full_matvec
    return sp["nev"] / full_s, r["wall"], (
        f"reference {'ztrlan' if sp['cplx'] else 'trlan'}, first {r['matvec']} operator "
        f"applications of the same solve ({r['wall']:.1f} s, {r['north']} Gram-Schmidt passes); "
        f"eigenpairs/s extrapolated to the full solve's {full_matvec} applications")


def leg(T, A, x0, steps, rp, ci, v, stats_every=0, sp=None):
    """Time `steps` complete solves on operator A (inputs resident); returns
    (seconds, last result, eigenpairs at resid <= 1e-8 ||A||, stats)."""
    sp = sp or SINGLE["laplace3d"]

    def solve():
        return T.solve_csr(A, sp["lohi"], sp["nev"], sp["maxlan"], sp["tol"], sp["restart"],
                           sp["maxmv"], sp["nev"], x0)

    if stats_every:
        T.stats_enable(True, every=stats_every)
    t0 = time.perf_counter()
    outs = [solve() for _ in range(steps)]
    t1 = time.perf_counter()
    stats = T.stats_get() if stats_every else None
    if stats_every:
        stats["_flops"] = T.stats_flops()
        T.stats_enable(False)
    g = outs[-1]
    assert g["stat"] == 0, g["stat"]
    nec = g["nec"]
    res = residuals(rp, ci, v, g["evec"][:, :nec], g["eval"][:nec])
    good = int((res <= 1e-8 * g["anrm"]).sum())
    return t1 - t0, g, good, stats


def mfma_ceiling():
    """The FP64 matrix-core rate this GPU sustains (TFLOP/s), from
    nutrlan_amd/mfma_f64_probe (register-only MFMA streams filling every CU),
    or None when the probe is not built."""
    exe = os.path.join(ROOT, "nutrlan_amd", "mfma_f64_probe")
    if not os.path.exists(exe):
        return None
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    for line in out.stdout.splitlines():
        if line.startswith("mfma_f64_tflops "):
            return float(line.split()[1])
    return None


def restart_gemm(stats, flops, key, kernel, ceiling=None):
    """FP64-MFMA use of the restart GEMM (V(:, lck:lck+k) <- V Y): its
    2 n jl k flops per timed launch over their event time, against the
    dense FP64 matrix-core peak and the rate the probe measured on this GPU,
    with its HBM rate beside it."""
    launches, ms, byts = stats[key]
    if launches <= 0 or ms <= 0:
        return None
    tf = flops[key] / (ms * 1e-3) / 1e12
    return {"kernel": kernel, "launches_timed": launches, "avg_launch_us": round(ms * 1e3 / launches, 2),
            "achieved_tflops": round(tf, 3), "peak_tflops": FP64_MFMA_PEAK_TFLOPS,
            "mfma_frac": round(tf / FP64_MFMA_PEAK_TFLOPS, 4),
            "measured_ceiling_tflops": ceiling,
            "ceiling_frac": round(tf / ceiling, 4) if ceiling else None,
            "GBps": round(byts / (ms * 1e-3) / 1e9, 1),
            "hbm_frac": round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}


def anderson_params(count):
    """(L, W, potential seed) of the batched workload's lattices: L ~
    U{64..256}, W ~ U[1, 12], seeded."""
    import numpy as np

    m = max(count, 1024)  # lattice i is the same for every count
    rng = np.random.default_rng(5)
    Ls = rng.integers(64, 257, m)
    Ws = rng.uniform(1.0, 12.0, m)
    return [(int(L), float(w), 7000 + i) for i, (L, w) in enumerate(zip(Ls[:count], Ws[:count]))]


def good_pairs(csr, r, tol=1e-8):
    """Eigenpairs of result r with ||A v - lambda v|| <= tol ||A||."""
    import scipy.sparse as sp

    rp, ci, v = csr
    n = len(rp) - 1
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    X = r["evec"][:, :r["nec"]]
    R = A @ X - X * r["eval"][:r["nec"]]
    return int(((R * R).sum(axis=0) ** 0.5 <= tol * r["anrm"]).sum())


def cpu_solve_one(i):
    """Reference trlan (oracle/_ref, 1 thread) on lattice i of the batched
    workload; returns (eigenpairs at 1e-8, operator applications)."""
    from nutrlan_amd import workloads as W
    from oracle import ref

    L, w, seed = anderson_params(i + 1)[i]  # same lattice as in the batch
    csr, _ = W.anderson2d_csr(L, w, seed=seed)
    x0 = W.start_vector(L * L, seed=9000 + i)
    r = ref.solve_csr(*csr, -1, 8, 32, 1e-8, 1, 100000, 8, x0, nthreads=1)
    return good_pairs(csr, r), r["matvec"]


def cpu_baseline_batch(ns, nproc):
    """Reference trlan on the first ns lattices, nproc single-threaded
    solves at a time (a process pool in a child interpreter that never
    touches the GPU).  Returns (eigenpairs/s, description) or None."""
    from oracle import ref

    if not ref.available():
        return None
    code = (
        "import sys, json, time; sys.path.insert(0, %r)\n"
        "from multiprocessing import get_context\n"
        "import bench\n"
        "t0 = time.perf_counter()\n"
        "with get_context('fork').Pool(%d) as pool:\n"
        "    out = pool.map(bench.cpu_solve_one, range(%d), chunksize=1)\n"
        "wall = time.perf_counter() - t0\n"
        "print(json.dumps(dict(wall=wall, good=sum(o[0] for o in out), "
        "matvec=sum(o[1] for o in out))))\n" % (ROOT, nproc, ns))
    env = dict(os.environ, OPENBLAS_NUM_THREADS="1", OMP_NUM_THREADS="1")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         timeout=900)
    if out.returncode != 0:
        print(out.stderr[-2000:], file=sys.stderr)
        return None
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return r["good"] / r["wall"], (
        f"reference trlan (1 thread each) on the first {ns} of the same lattices, {nproc} solves "
        f"at a time: {r['good']} eigenpairs at 1e-8, {r['matvec']} operator applications, "
        f"{r['wall']:.1f} s")


def batched_anderson2d(args):
    """The batched workload: 1024 independent periodic 2D Anderson lattices,
    L ~ U{64..256} (n 4096..65536, 28.7M sites in all), hopping -1, on-site
    U[-W/2, W/2] with W ~ U[1, 12] (seeded), 8 lowest eigenpairs each,
    maxlan 32, tol 1e-8, CSR resident in HBM.  One step = learned-start
    inference for every lattice + the mixed-precision batched solve
    (trlan_batch).  Beside it: the same batch from random starts, the plain
    FP64 batched solver, and the plain FP64 solver one lattice at a time
    (the drop-in trlan() path, timed on a sample of the same lattices)."""
    import numpy as np

    from nutrlan_amd import learned as LN
    from nutrlan_amd import trlan as T
    from nutrlan_amd import workloads as W

    ceiling = mfma_ceiling()  # its own process, before this one's first dispatch
    calibration_copy(T, 1 << 21)
    params = anderson_params(args.models)
    models = [LN.lattice(L, w, seed) for L, w, seed in params]
    count = len(models)
    nev, maxlan, tol = LN.NEV, 32, 1e-8
    starts = [W.start_vector(len(c[0]) - 1, seed=9000 + i) for i, (c, d) in enumerate(models)]
    plain = [T.Csr(*c) for c, d in models]
    mixed = [T.Csr(*c, mixed=True) for c, d in models]
    net = LN.load_net(args.start_weights)
    assert net is not None, "learned-start weights missing (run nutrlan_amd/train_start_net.py)"

    def total_good(res, n=None):
        return sum(good_pairs(c, r) for (c, d), r in zip(models[:n or count], res))

    # output arrays allocated (and paged in) once, reused by every leg, as a
    # serving caller would: trlan() writes into the caller's eval/evec/wrk
    bufs = [T.output_buffers(len(c[0]) - 1, nev) for c, d in models]

    def batch(ops, x0s, dstarts=None):
        probs = [(A, -1, nev, maxlan, tol, 1, 100000, nev, None if x0s is None else x0s[i])
                 for i, A in enumerate(ops)]
        rc, res, _ = T.solve_batch(probs, out=bufs, dstarts=dstarts)
        assert rc == 0, rc
        return res

    t_inf = [0.0]

    def step():
        # the net's outputs stay on the device and go into the batch as the
        # jobs' start vectors (trl_batch_job.dstart)
        t0 = time.perf_counter()
        x_dev = LN.start_vectors_device(net, [d for c, d in models])
        t_inf[0] += time.perf_counter() - t0
        return batch(mixed, None, x_dev)

    def progress(msg):
        print(f"bench.py: {msg}", file=sys.stderr, flush=True)

    progress(f"{count} lattices built; warm-up")
    for _ in range(args.warmup):
        step()
    progress("timed steps")
    t_inf[0] = 0.0
    # headline: learned start + mixed precision, batched; one launch in
    # STATS_EVERY of each kernel kind carries HIP start/stop events
    T.stats_enable(True, every=args.stats_every)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = step()
    elapsed = time.perf_counter() - t0
    stats = T.stats_get()
    flops = T.stats_flops()
    T.stats_enable(False)
    good = total_good(res)
    mv_net = sum(r["matvec"] for r in res)
    res = [dict(r, eval=r["eval"].copy(), evec=None) for r in res]  # the buffers are reused
    value = good * args.steps / elapsed
    band_ratio = {}
    if args.headline_only:  # a profiling run: no legs beside the headline
        el_r = el_p = el_s = float("nan")
        good_r = mv_rand = good_p = mv_p = 0
        plain_single = dl = float("nan")
        dstat = None
        ns = 0
    else:  # beside it, untimed by events
        t0 = time.perf_counter()
        progress("mixed batch from random starts")
        res_r = batch(mixed, starts)
        el_r = time.perf_counter() - t0
        good_r, mv_rand = total_good(res_r), sum(r["matvec"] for r in res_r)
        # operator applications, learned over random start, by lattice side
        for lo, hi in ((64, 112), (112, 160), (160, 208), (208, 257)):
            sel = [i for i, (L, _, _) in enumerate(params) if lo <= L < hi]
            a = sum(res[i]["matvec"] for i in sel)
            b = sum(res_r[i]["matvec"] for i in sel)
            band_ratio[f"L{lo}-{hi - 1}"] = {"lattices": len(sel),
                                             "ratio": round(a / b, 4) if b else None}
        t0 = time.perf_counter()
        progress("plain FP64 batch")
        res_p = batch(plain, starts)
        el_p = time.perf_counter() - t0
        good_p, mv_p = total_good(res_p), sum(r["matvec"] for r in res_p)
        res_p = [dict(r, eval=r["eval"].copy(), evec=None) for r in res_p]
        ns = min(args.plain_sample, count)
        t0 = time.perf_counter()
        progress(f"plain FP64 one lattice at a time ({ns})")
        res_s = [T.solve_csr(plain[i], -1, nev, maxlan, tol, 1, 100000, nev, starts[i],
                             out=bufs[i]) for i in range(ns)]
        el_s = time.perf_counter() - t0
        plain_single = total_good(res_s, ns) / el_s
        # eigenvalues of the headline path against plain FP64 (same lattices),
        # every eigenpair: |dlambda| / ||A||
        dall = np.concatenate([np.abs(np.sort(a["eval"][:nev]) - np.sort(b["eval"][:nev])) / a["anrm"]
                               for a, b in zip(res, res_p)])
        dl = float(dall.max())
        dstat = {"pairs": int(dall.size), "max": dl, "median": float(np.median(dall)),
                 "pairs_over_1e-10": int((dall > 1e-10).sum())}
    packed = all(A.packed for A in mixed)
    stencil = all(A.stencil[0] == 2 for A in mixed)
    for A in plain + mixed:
        A.close()
    net.close()

    # the batched kernels (b_*: one launch steps every solve of a lane)
    path = ["b_spmv", "b_recur", "b_proj", "b_update"]
    dom = max(path, key=lambda k: stats[k][1])
    launches, ms, byts = stats[dom]
    achieved = byts / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    pb = sum(stats[k][2] for k in path)
    pt = sum(stats[k][1] for k in path)
    path_gbs = pb / (pt * 1e-3) / 1e9 if pt > 0 else 0.0
    kname = {"b_spmv": "kb_spmv_st<2>" if stencil else "kb_spmv_pk" if packed else "kb_spmv_stream",
             "b_recur": "kb_recur",
             "b_proj": f"kb_recproj<{T.proj_tiles_get()[1]}, float>",
             "b_update": "kb_update<float>"}[dom]
    roofline = {
        "bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(kname, BATCH_TRAFFIC),
        "avg_launch_us": round(ms * 1e3 / max(launches, 1), 2),
        "bytes_per_launch": round(byts / max(launches, 1)),
        "path_spmv_recur_cgs": {"achieved": round(path_gbs, 1),
                                "frac": round(path_gbs / HBM_PEAK_GBS, 4)},
        "kernels": {k: {"launches": s[0], "ms": round(s[1], 3), "bytes": round(s[2]),
                        "GBps": round(s[2] / (s[1] * 1e-3) / 1e9, 1) if s[1] > 0 else 0.0}
                    for k, s in stats.items() if s[0] > 0},
        "restart_gemm": restart_gemm(stats, flops, "b_ritz", "kb_ritz", ceiling),
    }
    cpu = cpu_seq = None
    if not args.no_cpu_baseline and not args.headline_only:
        nproc = min(os.cpu_count() or 1, 16)
        cb = cpu_baseline_batch(min(args.cpu_sample, count), nproc)
        if cb is not None:
            cpu = {"value": round(cb[0], 2), "unit": "eigenpairs/s", "cores": nproc,
                   "kind": "reference", "sample": cb[1]}
        cb = cpu_baseline_batch(min(args.cpu_seq_sample, count), 1)
        if cb is not None:
            cpu_seq = {"value": round(cb[0], 2), "unit": "eigenpairs/s", "cores": 1,
                       "kind": "reference", "sample": cb[1]}
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "eigenpairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "anderson2d_batch_%d" % count, "L": "U{64..256}",
                   "W": "U[1,12]", "sites_total": int(sum(len(c[0]) - 1 for c, d in models)),
                   "nev": nev, "lohi": -1, "maxlan": maxlan, "tol": tol,
                   "path": "batched (trlan_batch), mixed precision (FP32 shadow basis for "
                           "Gram-Schmidt, FP64 recurrence/restart/refinement), learned start "
                           "(BF16-MFMA conv net, inference inside each step)",
                   "converged_at_1e-8": good, "of": nev * count, "matvec_total": mv_net,
                   "learned_start_inference_ms_per_step": round(1e3 * t_inf[0] / args.steps, 1)},
        "vs_plain_fp64": round(value / plain_single, 2) if plain_single == plain_single else None,
        "plain_fp64_one_at_a_time": {"value": round(plain_single, 2),
                                     "sample": f"first {ns} lattices, {el_s:.2f} s"},
        "plain_fp64_batched": {"value": round(good_p / el_p, 1), "s": round(el_p, 3),
                               "matvec_total": mv_p, "converged_at_1e-8": good_p},
        "mixed_batched_random_start": {"value": round(good_r / el_r, 1), "s": round(el_r, 3),
                                       "matvec_total": mv_rand, "converged_at_1e-8": good_r},
        "learned_vs_random_matvec": round(mv_net / mv_rand, 4) if mv_rand else None,
        "learned_vs_random_matvec_by_L": band_ratio or None,
        "max_eigenvalue_diff_vs_plain_fp64": dl,
        "eigenvalue_diff_vs_plain_fp64_per_pair": dstat,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_baseline_sequential": cpu_seq,
    }
    if args.headline_only:
        for k in ("plain_fp64_one_at_a_time", "plain_fp64_batched", "mixed_batched_random_start",
                  "max_eigenvalue_diff_vs_plain_fp64", "eigenvalue_diff_vs_plain_fp64_per_pair"):
            line[k] = None
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--path", choices=["mixed", "plain"], default="mixed",
                    help="solver path measured for `value` (the other is reported beside it)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=None,
                    help="CPU baseline sample: operator applications (laplace3d, default 96) "
                         "or lattices (anderson2d-batch, default 96)")
    ap.add_argument("--workload", choices=["anderson2d-batch"] + list(SINGLE),
                    default="anderson2d-batch",
                    help="anderson2d-batch (default): 1024 uneven 2D Anderson lattices through "
                         "the batched mixed-precision learned-start solver; laplace3d (128^3 "
                         "Laplacian), diag1000, anderson3d-c (complex 96^3), rand8m (n = 2^23 "
                         "random sparse): one solve per step, mixed vs plain FP64 (complex: "
                         "plain c128)")
    ap.add_argument("--cpu-seq-sample", type=int, default=24,
                    help="anderson2d-batch: lattices of the sequential (one process, one "
                         "thread) reference baseline")
    ap.add_argument("--models", type=int, default=1024)
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--stats-every", type=int, default=STATS_EVERY,
                    help="time one launch in N of each kernel kind with HIP events (1: all)")
    ap.add_argument("--start-weights", default=None,
                    help="anderson2d-batch: learned-start weights (.npy; default the shipped ones)")
    ap.add_argument("--headline-only", action="store_true",
                    help="anderson2d-batch: time the headline path only (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU (asked for {args.gpus}): the multi-GPU "
              "row-block sharding is a design, not built here.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from nutrlan_amd import trlan as T

    assert T.device_count() > 0, "no GPU"
    if args.workload == "anderson2d-batch":
        if args.cpu_sample is None:
            args.cpu_sample = 96
        batched_anderson2d(args)
        return
    if args.cpu_sample is None:
        args.cpu_sample = 96
    single_solve(T, args)


def single_solve(T, args):
    """One of the single-solve workloads (SINGLE): the --path solver timed
    with sampled HIP events, the other real path (plain or mixed) timed
    beside it; complex operators run the plain c128 path only."""
    sp = SINGLE[args.workload]
    print(f"building {sp['name']}", file=sys.stderr, flush=True)
    rp, ci, v, x0 = single_problem(args.workload)
    n = len(rp) - 1
    ceiling = mfma_ceiling()
    calibration_copy(T, min(n, 1 << 21) if n >= 1 << 16 else 1 << 16)
    cplx = sp["cplx"]
    path = "plain" if cplx else args.path
    ops = {"plain": T.Csr(rp, ci, v, complex_=cplx)}
    if not cplx:
        ops["mixed"] = T.Csr(rp, ci, v, mixed=True)
    other = None if cplx else ("plain" if path == "mixed" else "mixed")
    print("bench.py: warm-up", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        for k in ops:
            leg(T, ops[k], x0, 1, rp, ci, v, 0, sp)
    o = None
    if other:  # the path beside the headline one, timed on its own (no events)
        print(f"bench.py: {other} path", file=sys.stderr, flush=True)
        o = leg(T, ops[other], x0, args.steps, rp, ci, v, 0, sp)
    print(f"bench.py: {path} path (headline)", file=sys.stderr, flush=True)
    # headline path: one launch in stats_every of each kernel kind (a hashed,
    # non-periodic pick) carries HIP start/stop events stamped from its own
    # dispatch on the solver's stream; timing every launch would cost ~10%.
    elapsed, g, good, stats = leg(T, ops[path], x0, args.steps, rp, ci, v, args.stats_every, sp)
    flops = stats.pop("_flops")
    value = good * args.steps / elapsed
    nec = g["nec"]

    # roofline: dominant kernel by time, plus the path aggregate
    dom = max(("spmv", "recur", "proj", "update", "ritz"), key=lambda k: stats[k][1])
    launches, ms, byts = stats[dom]
    achieved = byts / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    pk = ["spmv", "recur", "proj", "update"]
    pb = sum(stats[k][2] for k in pk)
    pt = sum(stats[k][1] for k in pk)
    path_gbs = pb / (pt * 1e-3) / 1e9 if pt > 0 else 0.0
    kname = kernel_names(T, "complex" if cplx else path, ops[path]).get(dom)
    roofline = {
        "bound": "hbm", "kernel": kname or dom, "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(kname, sp["traffic"]) if sp["traffic"] else None,
        "avg_launch_us": round(ms * 1e3 / max(launches, 1), 2),
        "bytes_per_launch": round(byts / max(launches, 1)),
        "path_spmv_recur_cgs": {"achieved": round(path_gbs, 1),
                                "frac": round(path_gbs / HBM_PEAK_GBS, 4)},
        "kernels": {k: {"launches": s_[0], "ms": round(s_[1], 3), "bytes": round(s_[2]),
                        "GBps": round(s_[2] / (s_[1] * 1e-3) / 1e9, 1) if s_[1] > 0 else 0.0}
                    for k, s_ in stats.items() if s_[0] > 0},
        "restart_gemm": restart_gemm(stats, flops, "ritz", "k_ritz_mfma", ceiling),
    }
    cpu = None
    if not args.no_cpu_baseline:
        nthreads = min(os.cpu_count() or 1, 16)
        full_mv = o[1]["matvec"] if (o and path == "mixed") else g["matvec"]
        cb = cpu_baseline(args.workload, min(args.cpu_sample, full_mv), full_mv, nthreads)
        if cb is not None:
            cpu = {"value": round(cb[0], 4), "unit": "eigenpairs/s", "cores": nthreads,
                   "kind": "reference", "sample": cb[2]}
    desc = {"mixed": "mixed: FP32 shadow basis for Gram-Schmidt, FP64 SpMV/recurrence/basis/"
                     "restart GEMM, FP64 Rayleigh-quotient refinement",
            "plain": "plain FP64 device basis" + (" (complex c128)" if cplx else "")}
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
        "dtype": "c128" if cplx else "f64",
        "data": "synthetic",
        "config": {"workload": sp["name"], "n": n, "nnz": int(len(ci)), "nev": sp["nev"],
                   "lohi": sp["lohi"], "maxlan": sp["maxlan"], "tol": sp["tol"],
                   "restart": sp["restart"], "path": desc[path],
                   "learned_start": "none (random start; the learned start applies to the "
                                    "batched 2D Anderson models)",
                   "matvec": g["matvec"], "nloop": g["nloop"], "north": g["north"],
                   "converged": nec, "converged_at_1e-8": good},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if o:
        o_el, o_g, o_good, _ = o
        other_value = o_good * args.steps / o_el
        line[other + "_path"] = {"value": round(other_value, 3),
                                 "ms_per_step": round(1e3 * o_el / args.steps, 2),
                                 "matvec": o_g["matvec"], "converged_at_1e-8": o_good,
                                 "path": desc[other]}
        line["mixed_vs_plain"] = round((value / other_value) if path == "mixed"
                                       else (other_value / value), 3)
        line["max_eigenvalue_diff_vs_" + other] = (
            float(np.abs(np.sort(g["eval"][:nec]) - np.sort(o_g["eval"][:o_g["nec"]])).max()
                  / g["anrm"]) if o_g["nec"] == nec else None)
    print(json.dumps(line), flush=True)
    for A in ops.values():
        A.close()


if __name__ == "__main__":
    main()
