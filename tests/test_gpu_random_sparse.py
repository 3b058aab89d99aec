"""The random-sparse workload's parameters on the GPU: symmetric random
pattern (row degree ~Poisson(16), values N(0,1)/4 mirrored, diagonal
U[-1,1]), 32 largest eigenpairs, maxlan 128, tol 1e-8.

1. At an oracle-sized n (50,000): the plain FP64 path against the oracle
   (the reference restated; integer trajectory and eigenvalues within
   1e-10 ||A||), and the mixed-precision path's eigenvalues against the same
   FP64 oracle, every error printed, with true residuals recomputed here.
2. At the full size, n = 2^23 (142.6 M nonzeros), where the oracle is too
   slow: size-independent properties of the mixed-precision solve from two
   start vectors: every pair's true residual ||A v - lambda v|| <= tol
   ||A|| (computed with torch's CPU sparse product, not this library),
   orthonormal Ritz vectors, eigenvalues inside the Gershgorin bound, and
   the two solves' eigenvalues equal within 2 tol ||A||."""
import numpy as np
import pytest

from nutrlan_amd import workloads as W
from oracle import trlan_oracle as O

pytestmark = pytest.mark.gpu

NEV, MAXLAN, TOL = 32, 128, 1e-8


def _true_residuals(rp, ci, v, X, lam):
    import torch

    n = len(rp) - 1
    A = torch.sparse_csr_tensor(torch.from_numpy(rp.astype(np.int64)),
                                torch.from_numpy(ci.astype(np.int64)),
                                torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)),
                                size=(n, n))
    AX = (A @ torch.from_numpy(np.ascontiguousarray(X))).numpy()
    return np.linalg.norm(AX - X * lam[None, :], axis=0) / np.linalg.norm(X, axis=0)


def test_random_sparse_oracle_size(gpu_lib):
    T = gpu_lib
    n = 50000
    rp, ci, v = W.random_sym_csr(n, seed=11)
    x0 = W.start_vector(n, seed=12)
    print(f"random sparse n={n} nnz={len(v)} nev={NEV} maxlan={MAXLAN}", flush=True)
    P = T.Csr(rp, ci, v)
    p = T.solve_csr(P, 1, NEV, MAXLAN, TOL, 1, 40000, NEV, x0)
    P.close()
    M = T.Csr(rp, ci, v, mixed=True)
    g = T.solve_csr(M, 1, NEV, MAXLAN, TOL, 1, 40000, NEV, x0)
    M.close()
    info = O.init_info(n, MAXLAN, 1, NEV, TOL, 1, 40000)
    O.set_iguess(info, 0, 1)
    o = O.trlan(W.csr_matvec(rp, ci, v), info, n, NEV, x0)
    assert info.stat == 0 and info.nec == NEV
    anrm = info.anrm
    # plain FP64: the oracle's trajectory and eigenvalues
    assert p["stat"] == 0 and p["nec"] == NEV
    assert (p["matvec"], p["nloop"], p["north"], p["locked"]) == (
        info.matvec, info.nloop, info.north, info.locked)
    assert np.abs(p["eval"][:NEV] - o["eval"][:NEV]).max() <= 1e-10 * anrm
    # mixed precision: eigenvalues against the FP64 oracle, true residuals
    assert g["stat"] == 0 and g["nec"] == NEV
    err = []
    for i in range(NEV):
        j = np.abs(o["eval"][:NEV] - g["eval"][i]).argmin()
        err.append(abs(o["eval"][j] - g["eval"][i]) / anrm)
    print("mixed |dlambda|/||A|| per pair: " + " ".join(f"{e:.1e}" for e in err), flush=True)
    assert max(err) <= 1e-10
    for r in (p, g):
        res = _true_residuals(rp, ci, v, r["evec"][:, :NEV], r["eval"][:NEV])
        assert res.max() <= TOL * anrm, res.max() / anrm
    print(f"matvec plain {p['matvec']} mixed {g['matvec']} oracle {info.matvec}", flush=True)


def test_random_sparse_full_size_invariants(gpu_lib):
    T = gpu_lib
    n = 1 << 23
    rp, ci, v = W.random_sym_csr(n, seed=4)
    print(f"random sparse full size n={n} nnz={len(v)}", flush=True)
    A = T.Csr(rp, ci, v, mixed=True)
    runs = []
    for s in (5, 6):
        x0 = W.start_vector(n, seed=s)
        g = T.solve_csr(A, 1, NEV, MAXLAN, TOL, 1, 40000, NEV, x0)
        print(f"start seed {s}: stat {g['stat']} nec {g['nec']} matvec {g['matvec']} "
              f"{g['wall_s']:.2f} s", flush=True)
        runs.append(g)
    A.close()
    bound = float(np.add.reduceat(np.abs(v), rp[:-1]).max())
    for g in runs:
        assert g["stat"] == 0 and g["nec"] == NEV
        X = g["evec"][:, :NEV]
        lam = g["eval"][:NEV]
        res = _true_residuals(rp, ci, v, X, lam)
        assert res.max() <= TOL * g["anrm"], res.max() / g["anrm"]
        assert np.abs(X.T @ X - np.eye(NEV)).max() <= 1e-6
        assert np.abs(lam).max() <= bound * (1 + 1e-12)
    e1, e2 = (np.sort(g["eval"][:NEV]) for g in runs)
    assert np.abs(e1 - e2).max() <= 2 * TOL * runs[0]["anrm"]
