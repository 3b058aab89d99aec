"""GPU parity of the complex Hermitian path: ztrlan() on a device c128 CSR
operator against the oracle (reference ztrlan restated) and LAPACK.

The device projection computes the exact V^H r, so the oracle runs with
gdot_quirk=False (the reference's two-column ztrl_g_dot_ mis-computation,
SRC/ztrl_comm_none.c:100,117-119, is not reproduced).  Bars: eigenvalues
within 1e-10 ||A|| of the oracle's and of LAPACK's, residuals
||A v - lambda v|| <= tol ||A||, ||V^H V - I|| <= 1e-12 on the returned
Ritz vectors.  Locking decisions of the complex path compare residuals
against eps*|lambda| and may tip at roundoff (see tests/test_oracle.py), so
matvec counts are held to 10% of the oracle's instead of exactly.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from nutrlan_amd import workloads as W
from oracle import trlan_oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    # name, L, seed, lohi, ned, maxlan, tol, restart, maxmv, mev, start seed
    ("zand8_sm6", 8, 3, -1, 6, 40, 1e-8, 1, 20000, 6, 4),
    ("zand10_lg6", 10, 5, 1, 6, 40, 1e-8, 1, 20000, 6, 6),
    ("zand12_sm24_m80", 12, 7, -1, 24, 80, 1e-8, 1, 40000, 24, 8),   # c3 shape, small lattice
    ("zand9_both8", 9, 2, 0, 8, 48, 1e-8, 1, 40000, 8, 3),
]


def mv(rp, ci, v):
    return W.csr_matvec(rp, ci, v)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_ztrlan_device_vs_oracle_and_lapack(gpu_lib, case):
    T = gpu_lib
    name, L, seed, lohi, ned, maxlan, tol, restart, maxmv, mev, sseed = case
    rp, ci, v = W.anderson3d_csr(L, seed=seed)
    n = len(rp) - 1
    x0 = W.start_vector(n, seed=sseed, complex_=True)
    print(f"case {name} n={n}", flush=True)
    A = T.Csr(rp, ci, v, complex_=True)
    g = T.solve_csr(A, lohi, ned, maxlan, tol, restart, maxmv, mev, x0)
    A.close()
    assert g["stat"] == 0, g["stat"]

    info = O.init_info(n, maxlan, lohi, ned, tol, restart, maxmv)
    O.set_iguess(info, 0, 1)
    o = O.trlan(mv(rp, ci, v), info, n, mev, x0, complex_=True, gdot_quirk=False)
    assert info.stat == 0
    nec = g["nec"]
    assert nec == info.nec == ned
    anrm = max(abs(info.anrm), 1e-300)
    assert np.abs(g["eval"][:nec] - o["eval"][:nec]).max() <= 1e-10 * anrm
    assert abs(g["matvec"] - info.matvec) <= 0.1 * info.matvec

    # LAPACK on the dense Hermitian matrix
    H = W.csr_dense(rp, ci, v)
    w = sla.eigh(H, eigvals_only=True)
    if lohi < 0:
        exact = w[:ned]
    elif lohi > 0:
        exact = w[-ned:]
    else:
        exact = None
    got = np.sort(g["eval"][:nec])
    if exact is not None:
        assert np.abs(got - exact).max() <= 1e-10 * abs(w).max()
    else:
        # both ends: every returned value is an eigenvalue of H
        assert np.abs(got[:, None] - w[None, :]).min(axis=1).max() <= 1e-10 * abs(w).max()

    X = g["evec"][:, :nec]
    G = X.conj().T @ X
    assert np.abs(G - np.eye(nec)).max() <= 1e-12
    f = mv(rp, ci, v)
    for i in range(nec):
        r = f(X[:, i]) - g["eval"][i] * X[:, i]
        assert np.linalg.norm(r) <= 1.01 * tol * anrm + 1e-12 * anrm


def test_ztrlan_real_operator_matches_trlan(gpu_lib):
    """A real symmetric matrix stored as complex gives the trlan()
    eigenvalues (same spectrum, independent code path through ztrlan)."""
    T = gpu_lib
    rp, ci, v = W.laplace3d_csr(9, seed=12)
    n = len(rp) - 1
    x0 = W.start_vector(n, seed=13)
    A = T.Csr(rp, ci, v)
    g = T.solve_csr(A, 1, 6, 40, 1e-9, 1, 20000, 6, x0)
    A.close()
    Z = T.Csr(rp, ci, v.astype(np.complex128), complex_=True)
    z = T.solve_csr(Z, 1, 6, 40, 1e-9, 1, 20000, 6, x0.astype(np.complex128))
    Z.close()
    assert g["stat"] == 0 and z["stat"] == 0 and g["nec"] == z["nec"] == 6
    assert np.abs(np.sort(g["eval"][:6]) - np.sort(z["eval"][:6])).max() <= 1e-10 * g["anrm"]
