"""GPU checks of the post-processing entry points that evaluate Ritz pairs
with the operator (reference SRC/trlan.c:805-1108, 1111-1215, 1438-1678 and
SRC/ztrlan.c:225-500), against numpy on the same vectors."""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sla

from nutrlan_amd import workloads as W

pytestmark = pytest.mark.gpu
P = ctypes.c_void_p


def _p(a):
    return a.ctypes.data_as(P)


def _problem(T, n_side=10, seed=1, complex_=False):
    if complex_:
        rp, ci, v = W.anderson3d_csr(n_side, seed=seed)
    else:
        rp, ci, v = W.laplace3d_csr(n_side, seed=seed)
    n = len(rp) - 1
    A = T.Csr(rp, ci, v, complex_=complex_)
    H = W.csr_dense(rp, ci, v)
    return A, H, n


def test_rayleigh_quotients(gpu_lib):
    T = gpu_lib
    A, H, n = _problem(T)
    w, Z = sla.eigh(H)
    rng = np.random.default_rng(3)
    k = 5
    X = (Z[:, -k:] + 1e-3 * rng.standard_normal((n, k))) * rng.uniform(0.5, 2.0, k)
    X[:, 2] = 0.0  # a zero vector gets -DBL_MAX
    evec = np.asfortranarray(X.copy())
    eres = np.zeros(2 * k)
    info = T.init_info(n, 30, 1, k, 1e-8, 1, 1000)
    info.mvparam = A.handle.value
    T.lib().trl_rayleigh_quotients(A.op, ctypes.byref(info), k, _p(evec), n, _p(eres), None)
    assert info.stat == 0 and info.matvec == k
    for i in range(k):
        x = X[:, i]
        vv = x @ x
        if vv == 0.0:
            assert eres[i] == -np.finfo(float).max and eres[k + i] == -np.finfo(float).max
            continue
        rq = x @ H @ x / vv
        assert abs(eres[i] - rq) <= 1e-13 * abs(w).max()
        res = np.linalg.norm(H @ x - rq * x) / np.sqrt(vv)
        assert abs(eres[k + i] - res) <= 1e-10 * max(res, 1e-6)
        assert np.abs(evec[:, i] - x / np.sqrt(vv)).max() <= 1e-14
    A.close()


def test_check_ritz_converged_pairs(gpu_lib, capfd):
    """Pairs from a converged solve pass the residual and Rayleigh-quotient
    checks (check == 0)."""
    T = gpu_lib
    A, H, n = _problem(T, 9, seed=4)
    x0 = W.start_vector(n, seed=5)
    g = T.solve_csr(A, 1, 6, 40, 1e-10, 1, 20000, 6, x0)
    assert g["stat"] == 0 and g["nec"] == 6
    nec = 6
    evec = np.asfortranarray(g["evec"][:, :nec])
    alpha = g["eval"][:nec].copy()
    beta = g["resid"][:nec].copy()
    info = T.init_info(n, 40, 1, 6, 1e-10, 1, 20000)
    info.mvparam = A.handle.value
    info.nec = nec
    check = ctypes.c_int(99)
    wrk = np.zeros(4 * n)
    T.lib().trl_check_ritz(A.op, ctypes.byref(info), n, nec, _p(evec), n, _p(alpha),
                           ctypes.byref(check), _p(beta), None, _p(wrk), len(wrk))
    capfd.readouterr()
    assert check.value == 0
    A.close()


def test_check_ritz_flags_bad_pair(gpu_lib, capfd):
    """A perturbed vector and a wrong Ritz value are counted as failures."""
    T = gpu_lib
    A, H, n = _problem(T, 8, seed=6)
    w, Z = sla.eigh(H)
    k = 3
    X = np.asfortranarray(Z[:, -k:].copy())
    alpha = w[-k:].copy()
    beta = np.array([np.linalg.norm(H @ X[:, i] - alpha[i] * X[:, i]) for i in range(k)])
    alpha_bad = alpha.copy()
    alpha_bad[1] += 1e-2
    info = T.init_info(n, 30, 1, k, 1e-10, 1, 1000)
    info.mvparam = A.handle.value
    info.nec = k
    check = ctypes.c_int(0)
    wrk = np.zeros(4 * n)
    T.lib().trl_check_ritz(A.op, ctypes.byref(info), n, k, _p(X), n, _p(alpha_bad),
                           ctypes.byref(check), _p(beta), _p(alpha), _p(wrk), len(wrk))
    capfd.readouterr()
    assert check.value < 0
    A.close()


def test_ritz_projection(gpu_lib):
    """Rayleigh-Ritz on a perturbed invariant subspace: values are the
    eigenvalues of the projected pencil, vectors A-orthogonal Ritz vectors,
    residuals ||A v - d v||."""
    T = gpu_lib
    A, H, n = _problem(T, 10, seed=7)
    w, Z = sla.eigh(H)
    m = 4
    rng = np.random.default_rng(8)
    X = Z[:, -m:] @ rng.standard_normal((m, m)) + 1e-4 * rng.standard_normal((n, m))
    evec = np.asfortranarray(X.copy())
    eres = np.zeros(2 * m)
    wrk = np.zeros(4 * m * m + 16)
    info = T.init_info(n, 30, 1, m, 1e-8, 1, 1000)
    info.mvparam = A.handle.value
    T.lib().trl_ritz_projection(A.op, ctypes.byref(info), m, _p(evec), n, _p(eres), _p(wrk),
                                len(wrk), None)
    assert info.stat == 0
    G = X.T @ X
    Hp = X.T @ H @ X
    d_ref = sla.eigh(Hp, G, eigvals_only=True)
    assert np.abs(np.sort(eres[:m]) - d_ref).max() <= 1e-10 * abs(w).max()
    V = np.asarray(evec)
    assert np.abs(V.T @ V - np.eye(m)).max() <= 1e-10
    for i in range(m):
        r = np.linalg.norm(H @ V[:, i] - eres[i] * V[:, i])
        assert abs(eres[m + i] - r) <= 1e-9 * max(r, 1e-8)
    A.close()


def test_ztrl_check_ritz(gpu_lib, capfd):
    T = gpu_lib
    A, H, n = _problem(T, 7, seed=9, complex_=True)
    w, Z = sla.eigh(H)
    k = 3
    X = np.asfortranarray(Z[:, :k].copy())
    alpha = w[:k].copy()
    beta = np.array([np.linalg.norm(H @ X[:, i] - alpha[i] * X[:, i]) for i in range(k)])
    info = T.init_info(n, 30, -1, k, 1e-10, 1, 1000)
    info.mvparam = A.handle.value
    info.nec = k
    check = ctypes.c_int(99)
    wrk = np.zeros(4 * n, dtype=np.complex128)
    T.lib().ztrl_check_ritz(A.op, ctypes.byref(info), n, k, _p(X), n, _p(alpha),
                            ctypes.byref(check), _p(beta), _p(alpha), _p(wrk), len(wrk))
    capfd.readouterr()
    assert check.value == 0
    A.close()
