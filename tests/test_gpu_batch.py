"""GPU checks of trlan_batch: many independent solves advancing together
(one batched launch per step phase, one synchronisation per round).  Each
job must make exactly the decisions of a lone trlan() call, so its
eigenvalues and counters equal the single solve's bit for bit, and its
eigenvalues match the oracle's to 1e-10 ||A||."""
import numpy as np
import pytest

from nutrlan_amd import workloads as W
from oracle import trlan_oracle as O

pytestmark = pytest.mark.gpu


def _jobs(T, count, mixed_every=0):
    rng = np.random.default_rng(11)
    out = []
    for i in range(count):
        L = int(rng.integers(24, 72))
        Wd = float(rng.uniform(1.0, 12.0))
        rp, ci, v = W.anderson2d_csr(L, Wd, seed=100 + i)[0]
        n = len(rp) - 1
        mixed = bool(mixed_every) and i % mixed_every == 0
        A = T.Csr(rp, ci, v, mixed=mixed)
        x0 = W.start_vector(n, seed=200 + i)
        out.append(((rp, ci, v), A, (A, -1, 8, 32, 1e-8, 1, 40000, 8, x0), x0))
    return out


def test_batch_matches_single_solves(gpu_lib):
    T = gpu_lib
    jobs = _jobs(T, 24, mixed_every=3)
    rc, res, wall = T.solve_batch([j[2] for j in jobs])
    assert rc == 0
    for (csr, A, args, x0), g in zip(jobs, res):
        s = T.solve_csr(*args)
        assert g["stat"] == 0 and s["stat"] == 0
        keys = ["nec", "matvec", "nloop", "north", "nrand", "locked"]
        assert [g[k] for k in keys] == [s[k] for k in keys]
        assert np.array_equal(g["eval"], s["eval"])
        assert np.array_equal(g["evec"], s["evec"])
    for csr, A, args, x0 in jobs:
        A.close()


def test_batch_vs_oracle(gpu_lib):
    T = gpu_lib
    jobs = _jobs(T, 6)
    rc, res, wall = T.solve_batch([j[2] for j in jobs])
    assert rc == 0
    for (csr, A, args, x0), g in zip(jobs, res):
        rp, ci, v = csr
        n = len(rp) - 1
        info = O.init_info(n, 32, -1, 8, 1e-8, 1, 40000)
        O.set_iguess(info, 0, 1)
        o = O.trlan(W.csr_matvec(rp, ci, v), info, n, 8, x0)
        assert g["nec"] == info.nec == 8
        assert (g["matvec"], g["nloop"], g["north"]) == (info.matvec, info.nloop, info.north)
        assert np.abs(g["eval"][:8] - o["eval"][:8]).max() <= 1e-10 * info.anrm
        A.close()


def test_batch_mixed_operators_and_early_exit(gpu_lib):
    """Jobs with different lanes-per-row groups (2D and 3D stencils), a job
    with nothing to do (ned <= nec) and a job with an argument error run in
    one batch; the good ones still equal their single solves."""
    T = gpu_lib
    a = W.laplace3d_csr(14, seed=1)
    b = W.anderson2d_csr(40, 5.0, seed=2)[0]
    Aa, Ab = T.Csr(*a), T.Csr(*b)
    xa = W.start_vector(len(a[0]) - 1, seed=3)
    xb = W.start_vector(len(b[0]) - 1, seed=4)
    probs = [(Aa, 1, 6, 30, 1e-8, 1, 20000, 6, xa), (Ab, -1, 5, 24, 1e-8, 1, 20000, 5, xb),
             (Ab, -1, 0, 24, 1e-8, 1, 20000, 5, xb)]
    rc, res, wall = T.solve_batch(probs)
    assert rc == 0
    for p, g in zip(probs[:2], res[:2]):
        s = T.solve_csr(*p)
        assert g["matvec"] == s["matvec"] and np.array_equal(g["eval"], s["eval"])
    assert res[2]["matvec"] == 0
    Aa.close()
    Ab.close()


def test_batch_device_start_vectors(gpu_lib):
    """Start vectors handed over on the device (trl_batch_job.dstart, FP32;
    the library converts the FP64 potentials into the net's input itself,
    trlnet_forward_many_device_f64) give exactly the solves that the host
    path's FP32 values give from evec(:, 0) (converted to FP64)."""
    from nutrlan_amd import learned as LN

    T = gpu_lib
    net = LN.load_net()
    assert net is not None
    rng = np.random.default_rng(31)
    models = [LN.lattice(int(L), float(rng.uniform(1, 12)), 800 + i)
              for i, L in enumerate((64, 80, 64, 96, 72))]
    pots = [d for c, d in models]
    host = LN.start_vectors(net, pots)
    dptrs = LN.start_vectors_device(net, pots)
    for mixed in (False, True):
        ops = [T.Csr(*c, mixed=mixed) for c, d in models]
        probs = [(A, -1, 8, 32, 1e-8, 1, 100000, 8, x) for A, x in zip(ops, host)]
        rc_h, res_h, _ = T.solve_batch(probs)
        rc_d, res_d, _ = T.solve_batch([p[:-1] + (None,) for p in probs], dstarts=dptrs)
        assert rc_h == 0 and rc_d == 0
        for a, b in zip(res_h, res_d):
            assert a["matvec"] == b["matvec"] and a["nec"] == b["nec"] == 8
            assert np.array_equal(a["eval"], b["eval"])
            assert np.array_equal(a["evec"], b["evec"])
        for A in ops:
            A.close()
    net.close()


def test_batch_heterogeneous_groups(gpu_lib, monkeypatch):
    """One batch whose steps fall into four launch groups (packed and
    unpacked CSR-stream SpMV, FP64 and mixed-precision bases) staged in one
    transfer: every job still equals its single solve bit for bit."""
    T = gpu_lib
    a = W.anderson2d_csr(40, 6.0, seed=5)[0]
    b = W.random_sym_csr(3000, seed=6)
    c = W.laplace3d_csr(12, seed=7)
    d = W.anderson2d_csr(48, 9.0, seed=8)[0]
    e = W.random_sym_csr(2500, seed=9)
    ops = [T.Csr(*a), T.Csr(*b), T.Csr(*c, mixed=True)]
    monkeypatch.setenv("TRLHIP_CSR_PACK", "0")
    ops += [T.Csr(*d, mixed=True), T.Csr(*e, mixed=True)]
    monkeypatch.delenv("TRLHIP_CSR_PACK")
    assert [A.packed for A in ops] == [True, False, True, False, False]
    probs = [(A, -1, 5, 30, 1e-8, 1, 20000, 5, W.start_vector(A.n, seed=40 + i))
             for i, A in enumerate(ops)]
    print("heterogeneous batch: n =", [A.n for A in ops], flush=True)
    rc, res, wall = T.solve_batch(probs)
    assert rc == 0
    for p, g in zip(probs, res):
        s = T.solve_csr(*p)
        assert g["stat"] == 0 and s["stat"] == 0 and g["nec"] == s["nec"] == 5
        assert (g["matvec"], g["nloop"], g["north"]) == (s["matvec"], s["nloop"], s["north"])
        assert np.array_equal(g["eval"], s["eval"])
        assert np.array_equal(g["evec"], s["evec"])
    for A in ops:
        A.close()


def test_batch_chained_steps(gpu_lib, monkeypatch):
    """Chained inner steps (several Lanczos steps of a solve in one round,
    each run only when the device finds the solver would accept the step
    before it) change nothing: every job equals its single solve bit for
    bit with chains of 8 (the default), of 2 and with none, including jobs
    whose chains stop early -- a basis as large as the problem (the last
    step takes the local re-orthogonalisation, n = 30 = maxlan), frequent
    convergence tests (ned close to nec), and mixed-precision bases."""
    T = gpu_lib
    specs = []
    for i, (L, Wd, nev, maxlan, mixed) in enumerate([(24, 3.0, 8, 32, False), (40, 9.0, 8, 32, True),
                                                     (56, 1.5, 4, 20, False), (32, 6.0, 12, 40, True),
                                                     (64, 11.0, 8, 32, False)]):
        rp, ci, v = W.anderson2d_csr(L, Wd, seed=300 + i)[0]
        specs.append(((rp, ci, v), mixed, nev, maxlan))
    rp, ci, v = W.random_sym_csr(30, seed=310)
    specs.append(((rp, ci, v), False, 6, 30))
    ops, probs = [], []
    for i, (csr, mixed, nev, maxlan) in enumerate(specs):
        A = T.Csr(*csr, mixed=mixed)
        x0 = W.start_vector(len(csr[0]) - 1, seed=320 + i)
        ops.append(A)
        probs.append((A, -1, nev, maxlan, 1e-8, 1, 40000, nev, x0))
    singles = [T.solve_csr(*p) for p in probs]
    keys = ["stat", "nec", "matvec", "nloop", "north", "nrand", "locked"]
    for chain in (None, "2", "0"):
        if chain is None:
            monkeypatch.delenv("TRLHIP_BATCH_CHAIN", raising=False)
        else:
            monkeypatch.setenv("TRLHIP_BATCH_CHAIN", chain)
        rc, res, wall = T.solve_batch(probs)
        assert rc == 0
        print(f"chain {chain or 'default'}: matvec {[g['matvec'] for g in res]}", flush=True)
        for g, s in zip(res, singles):
            assert [g[k] for k in keys] == [s[k] for k in keys]
            assert np.array_equal(g["eval"], s["eval"])
            assert np.array_equal(g["evec"], s["evec"])
    for A in ops:
        A.close()


def test_batch_polish_matches_single_solves(gpu_lib):
    """Mixed-precision jobs whose refined residual misses tol take the
    Rayleigh-Ritz polish over [X, residuals] (trlhip_residuals and
    trlhip_gram, a device round trip each) inside the batch; with a few
    jobs that do not, every job equals its lone trlan() call bit for bit,
    and each returned pair's true residual meets tol (numpy)."""
    import scipy.sparse as sp

    T = gpu_lib
    specs = [(80, 11.35921550186609, 7718, 9718),  # ground state needs the polish
             (75, 9.673365812110031, 7380, 9380),  # polish with a singular [X, R]
             (64, 6.0, 31, 41), (48, 3.0, 32, 42)]
    jobs = []
    for L, Wd, seed, sseed in specs:
        rp, ci, v = W.anderson2d_csr(L, Wd, seed=seed)[0]
        n = len(rp) - 1
        A = T.Csr(rp, ci, v, mixed=True)
        x0 = W.start_vector(n, seed=sseed)
        jobs.append(((rp, ci, v), A, (A, -1, 8, 32, 1e-8, 1, 40000, 8, x0)))
    print(f"batch polish jobs {[s[:2] for s in specs]}", flush=True)
    rc, res, wall = T.solve_batch([j[2] for j in jobs])
    assert rc == 0
    for ((rp, ci, v), A, args), g in zip(jobs, res):
        s = T.solve_csr(*args)
        assert g["stat"] == 0 and s["stat"] == 0
        keys = ["nec", "matvec", "nloop", "north", "nrand", "locked"]
        assert [g[k] for k in keys] == [s[k] for k in keys]
        assert np.array_equal(g["eval"], s["eval"])
        assert np.array_equal(g["evec"], s["evec"])
        n = len(rp) - 1
        M = sp.csr_matrix((v, ci, rp), shape=(n, n))
        anrm = abs(g["anrm"])
        for i in range(g["nec"]):
            x = g["evec"][:, i]
            r = np.linalg.norm(M @ x - g["eval"][i] * x) / np.linalg.norm(x)
            assert r <= 1e-8 * anrm, (i, r / anrm)
    for _, A, _ in jobs:
        A.close()
