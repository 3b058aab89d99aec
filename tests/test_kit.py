"""Host FP64 kit of the driver (trlkit_* in libnutrlan_amd.so) against the
oracle and LAPACK, on the CPU.

The kit's dense eigen-solvers are this project's own (implicit QL,
bottom-up Householder); they are checked against scipy's LAPACK.  Its
decision logic (paired sorts, convergence test, restart schemes, locking)
restates the reference and is checked against the oracle's restatement on
randomized inputs, including ties and every restart scheme.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sla

from nutrlan_amd import trlan as T
from oracle import restart_oracle, sorts
from oracle import trlan_oracle as O

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double


@pytest.fixture(scope="module")
def kit():
    L = T.lib()
    L.trlkit_sort2.argtypes = [I, I, D, P, P, P]
    L.trlkit_tridiag_ql.argtypes = [I, P, P, P, I]
    L.trlkit_tridiag_ql.restype = I
    L.trlkit_householder.argtypes = [I, P, P, P]
    L.trlkit_syev.argtypes = [I, P, P]
    L.trlkit_syev.restype = I
    L.trlkit_convergence_test.argtypes = [I, P, P, P, ctypes.POINTER(T.TrlInfo)]
    L.trlkit_shuffle_eig.argtypes = [I, I, P, P, P, ctypes.POINTER(T.TrlInfo), I]
    L.trlkit_shuffle_eig.restype = I
    L.trlkit_set_locking.argtypes = [I, I, P, P, P, D]
    L.trlkit_set_locking.restype = I
    L.trlkit_sort_eig.argtypes = [I, I, I, D, P, P, P]
    L.trlkit_smooth_rr.argtypes = [I, P]
    return L


def p(a):
    return a.ctypes.data_as(P)


@pytest.mark.parametrize("n", [1, 2, 3, 7, 40, 129])
def test_tridiag_ql_vs_lapack(kit, n):
    rng = np.random.default_rng(n)
    d = rng.standard_normal(n)
    e = rng.standard_normal(max(n - 1, 0))
    w_ref, Z_ref = sla.eigh_tridiagonal(d, e)
    dd, ee, z = d.copy(), np.concatenate([e, [0.0]]), np.zeros(n)
    assert kit.trlkit_tridiag_ql(n, p(dd), p(ee), p(z), 1) == 0
    scale = max(abs(w_ref).max(), 1.0)
    assert np.abs(dd - w_ref).max() <= 1e-13 * scale
    assert np.abs(np.abs(z) - np.abs(Z_ref[-1])).max() <= 1e-12
    # full vectors from identity: A Z = Z diag(w)
    dd, ee = d.copy(), np.concatenate([e, [0.0]])
    Z = np.eye(n).copy(order="F")
    assert kit.trlkit_tridiag_ql(n, p(dd), p(ee), p(Z), 2) == 0
    A = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    Zc = Z  # Fortran order: column j = eigenvector j
    assert np.abs(A @ Zc - Zc * dd).max() <= 1e-12 * scale
    assert np.abs(Zc.T @ Zc - np.eye(n)).max() <= 1e-13 * max(n, 1)


@pytest.mark.parametrize("n", [2, 5, 33, 64])
def test_householder_arrowhead(kit, n):
    """Arrowhead -> tridiagonal with Q fixing the last coordinate (the
    dsytrd('U') + dorgtr convention): A = Q T Q^T."""
    rng = np.random.default_rng(100 + n)
    alpha = rng.standard_normal(n)
    beta = rng.standard_normal(n)
    A = np.diag(alpha)
    A[:-1, -1] = beta[:-1]
    A[-1, :-1] = beta[:-1]
    a = np.asfortranarray(A.copy())
    d, e = np.zeros(n), np.zeros(n)
    kit.trlkit_householder(n, p(a), p(d), p(e))
    Q = a  # column-major n x n in Fortran order
    Tm = np.diag(d) + np.diag(e[:-1], 1) + np.diag(e[:-1], -1)
    assert np.abs(Q @ Tm @ Q.T - A).max() <= 1e-13 * n * max(abs(A).max(), 1)
    assert np.abs(Q.T @ Q - np.eye(n)).max() <= 1e-13 * n
    assert np.allclose(Q[:, -1], np.eye(n)[:, -1]) and np.allclose(Q[-1], np.eye(n)[-1])
    assert abs(d[-1] - alpha[-1]) == 0.0


@pytest.mark.parametrize("n", [1, 4, 31, 100])
def test_syev_vs_lapack(kit, n):
    rng = np.random.default_rng(200 + n)
    B = rng.standard_normal((n, n))
    A = np.asfortranarray(B + B.T)
    a = A.copy(order="F")
    w = np.zeros(n)
    assert kit.trlkit_syev(n, p(a), p(w)) == 0
    w_ref = np.linalg.eigvalsh(A)
    assert np.abs(w - w_ref).max() <= 1e-12 * abs(w_ref).max()
    assert np.abs(A @ a - a * w).max() <= 1e-11 * abs(w_ref).max()


@pytest.mark.parametrize("key", [0, 1, 2, 3, 4])
def test_sort2_matches_reference_shell_sort(kit, key):
    rng = np.random.default_rng(key)
    for n in [0, 1, 2, 9, 50]:
        a1 = np.round(rng.standard_normal(n) * 3) / 2  # ties
        a2 = np.arange(n, dtype=np.float64)
        s = 0.7
        x1, x2 = a1.copy(), a2.copy()
        tag = np.arange(n, dtype=np.int32)
        kit.trlkit_sort2(n, key, s, p(x1), p(x2), p(tag))
        y1, y2 = list(a1), list(a2)
        fn = [lambda: sorts.dsort2(n, y1, y2), lambda: sorts.dsort2a(n, y1, y2),
              lambda: sorts.dsort2s(n, s, y1, y2), lambda: sorts.dsort2su(n, s, y1, y2),
              lambda: sorts.dsort2sd(n, s, y1, y2)][key]
        fn()
        assert list(x1) == y1 and list(x2) == y2
        assert list(tag) == [int(v) for v in y2]


def mk_info(o):
    """TrlInfo with the fields of an oracle Info."""
    i = T.TrlInfo()
    for f, _ in T.TrlInfo._fields_:
        if hasattr(o, f) and f not in ("mpicomp", "log_fp", "cpt_fp", "mvparam"):
            v = getattr(o, f)
            if isinstance(v, (int, float, np.floating, np.integer)):
                setattr(i, f, type(getattr(i, f))(v) if not isinstance(getattr(i, f), bytes) else v)
    return i


def rand_state(seed, nd, lohi, restart, ned):
    rng = np.random.default_rng(seed)
    o = O.Info()
    o.lohi, o.ned, o.restart = lohi, ned, restart
    o.tol, o.anrm = 1e-8, 10.0
    o.maxlan = nd + 4
    o.klan = nd
    o.maxmv = 5000
    o.matvec = int(rng.integers(nd, 2000))
    o.nloop = int(rng.integers(0, 5))
    o.nec = int(rng.integers(0, ned))
    o.locked = int(rng.integers(0, 3))
    o.crat = float(rng.uniform(0.5, 0.99))
    o.cfac = float(rng.uniform(0.01, 0.9))
    o.ptres = float(rng.uniform(1e-6, 1e-2))
    o.avgm = float(rng.uniform(5, 20))
    o.k = int(rng.integers(0, nd // 2))
    o.rfact = 1.5
    lam = np.sort(rng.standard_normal(nd) * 3)
    res = 10.0 ** rng.uniform(-14, -1, nd)
    o.trgt = float(lam[int(rng.integers(0, nd))])
    return o, lam, res


@pytest.mark.parametrize("lohi", [-4, -3, -2, -1, 0, 1])
def test_convergence_test_vs_oracle(kit, lohi):
    for seed in range(20):
        nd = 5 + seed
        o, lam, res = rand_state(77 + seed, nd, lohi, 1, 4)
        o.tmv = 3 if seed % 2 else -1
        o.tres = 1e-3
        o.ref = 0.3
        res[: seed % 4] = 1e-13
        info = mk_info(o)
        l1, r1 = lam.copy(), res.copy()
        tag = np.arange(nd, dtype=np.int32)
        kit.trlkit_convergence_test(nd, p(l1), p(r1), p(tag), ctypes.byref(info))
        l2, r2 = lam.copy(), res.copy()
        O.convergence_test(nd, l2, r2, o)
        assert np.array_equal(l1, l2) and np.array_equal(r1, r2)
        for f in ["nec", "trgt", "tres", "anrm", "crat", "cfac", "tmv", "ptres"]:
            assert getattr(info, f) == getattr(o, f), f


def test_set_locking_and_sort_eig_vs_oracle(kit):
    rng = np.random.default_rng(5)
    for trial in range(30):
        nlam, jnd = 6 + trial % 5, 20
        lam = np.sort(rng.standard_normal(nlam))
        res = 10.0 ** rng.uniform(-18, -8, nlam)
        yy = np.asfortranarray(rng.standard_normal((jnd, nlam)))
        a1, b1, y1 = lam.copy(), res.copy(), yy.copy(order="F")
        k1 = kit.trlkit_set_locking(jnd, nlam, p(a1), p(b1), p(y1), 3.0)
        a2, b2, y2 = lam.copy(), res.copy(), yy.copy(order="F")
        k2 = O.set_locking(jnd, nlam, a2, b2, y2, 3.0)
        assert k1 == k2 and np.array_equal(a1, a2) and np.array_equal(b1, b2)
        assert np.array_equal(y1, y2)
    for lohi in [-4, -3, -2, -1, 0, 1]:
        lam = rng.standard_normal(12)
        res = rng.uniform(0, 1, 12)
        a1, b1, tag = lam.copy(), res.copy(), np.arange(12, dtype=np.int32)
        kit.trlkit_sort_eig(12, lohi, 5, 0.2, p(a1), p(b1), p(tag))
        a2, b2 = lam.copy(), res.copy()
        O.sort_eig(12, lohi, 5, 0.2, a2, b2)
        assert np.array_equal(a1[:5], a2[:5]) and np.array_equal(b1[:5], b2[:5])
        assert all(a1[k] == lam[tag[k]] for k in range(5))


# ------------------------------------------------------------ checkpoints

COLCB = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double))


def _ck_bind(kit):
    kit.trlkit_write_checkpoint.argtypes = [ctypes.c_char_p, I, I, P, P, I, COLCB, P, P]
    kit.trlkit_write_checkpoint.restype = I
    kit.trlkit_read_checkpoint.argtypes = [ctypes.c_char_p, I, I, I, ctypes.POINTER(I), P, P,
                                          COLCB, P, P]
    kit.trlkit_read_checkpoint.restype = I


def _ck_write(kit, path, V, alpha, beta):
    nrow, ncol = V.shape
    buf = np.zeros(nrow)

    def get(_u, i, b):
        np.ctypeslib.as_array(b, shape=(nrow,))[:] = V[:, i]
        return 0

    cb = COLCB(get)
    return kit.trlkit_write_checkpoint(path.encode(), nrow, len(alpha), p(alpha), p(beta), ncol,
                                       cb, None, p(buf))


def _ck_read(kit, path, nrow, mev, maxjnd):
    cols = {}
    buf = np.zeros(nrow)

    def put(_u, i, b):
        cols[i] = np.ctypeslib.as_array(b, shape=(nrow,)).copy()
        return 0

    cb = COLCB(put)
    jnd = I(0)
    alpha = np.zeros(maxjnd + 1)
    beta = np.zeros(maxjnd + 1)
    rc = kit.trlkit_read_checkpoint(path.encode(), nrow, mev, maxjnd, ctypes.byref(jnd), p(alpha),
                                    p(beta), cb, None, p(buf))
    return rc, jnd.value, alpha[:jnd.value], beta[:jnd.value], cols


def test_checkpoint_round_trip_and_layout(kit, tmp_path):
    """Byte layout of the reference's trl_write_checkpoint (SRC/trlaux.c:
    780-880): int nrow, int jnd, alpha, beta, then the columns."""
    _ck_bind(kit)
    rng = np.random.default_rng(3)
    nrow, jnd = 37, 5
    V = rng.standard_normal((nrow, jnd + 1))
    alpha, beta = rng.standard_normal(jnd), rng.standard_normal(jnd)
    path = str(tmp_path / "cp0")
    assert _ck_write(kit, path, V, alpha, beta) == 0
    raw = open(path, "rb").read()
    assert np.frombuffer(raw[:8], dtype=np.int32).tolist() == [nrow, jnd]
    body = np.frombuffer(raw[8:], dtype=np.float64)
    assert np.array_equal(body[:jnd], alpha) and np.array_equal(body[jnd:2 * jnd], beta)
    assert np.array_equal(body[2 * jnd:].reshape(jnd + 1, nrow).T, V)
    # read back: mev > jnd -> jnd + 1 columns
    rc, j, a2, b2, cols = _ck_read(kit, path, nrow, 8, 10)
    assert rc == 0 and j == jnd and np.array_equal(a2, alpha) and np.array_equal(b2, beta)
    assert sorted(cols) == list(range(jnd + 1))
    assert all(np.array_equal(cols[i], V[:, i]) for i in cols)
    # mev <= jnd -> jnd columns (the reference reads j1 + j2 = jnd)
    rc, j, _, _, cols = _ck_read(kit, path, nrow, 3, 10)
    assert rc == 0 and sorted(cols) == list(range(jnd))
    # error codes of trl_read_checkpoint
    assert _ck_read(kit, path, nrow + 1, 8, 10)[0] == -213
    assert _ck_read(kit, path, nrow, 8, jnd - 1)[0] == -214
    assert _ck_read(kit, str(tmp_path / "missing0"), nrow, 8, 10)[0] == -212
    open(str(tmp_path / "short0"), "wb").write(raw[: len(raw) // 2])
    assert _ck_read(kit, str(tmp_path / "short0"), nrow, 8, 10)[0] in (-215, -216)


@pytest.mark.skipif(not __import__("oracle.ref", fromlist=["available"]).available(),
                    reason="compiled reference not built here")
def test_checkpoint_files_interchange_with_reference(kit, tmp_path):
    """Files written here are read by the reference's trl_read_checkpoint
    and files the reference writes are read here, with identical data."""
    from oracle import ref

    _ck_bind(kit)
    R = ref.lib()
    R.trl_write_checkpoint.argtypes = [ctypes.c_char_p, I, P, P, P, I, I, P, I, I]
    R.trl_write_checkpoint.restype = I
    R.trl_read_checkpoint.argtypes = [ctypes.c_char_p, I, P, I, I, ctypes.POINTER(I), P, I, I,
                                      ctypes.POINTER(I), I, P, I, P]
    R.trl_read_checkpoint.restype = I
    rng = np.random.default_rng(5)
    nrow, me, nb = 23, 3, 2          # evec part and base part
    jnd = me + nb - 1
    E = np.asfortranarray(rng.standard_normal((nrow, me)))
    B = np.asfortranarray(rng.standard_normal((nrow, nb)))
    alpha, beta = rng.standard_normal(jnd), rng.standard_normal(jnd)
    # reference -> here
    path = str(tmp_path / "ref0")
    assert R.trl_write_checkpoint(path.encode(), nrow, p(alpha), p(beta), p(E), nrow, me, p(B),
                                  nrow, nb) == 0
    rc, j, a2, b2, cols = _ck_read(kit, path, nrow, 10, 10)
    assert rc == 0 and j == jnd and np.array_equal(a2, alpha) and np.array_equal(b2, beta)
    V = np.hstack([E, B])
    assert all(np.array_equal(cols[i], V[:, i]) for i in range(jnd + 1))
    # here -> reference
    path2 = str(tmp_path / "ours0")
    assert _ck_write(kit, path2, V, alpha, beta) == 0
    mev, nbas = 10, 4
    ev = np.zeros(nrow * mev)
    bs = np.zeros(nrow * nbas)
    j1, j2 = I(0), I(0)
    al = np.zeros(mev + nbas)
    be = np.zeros(mev + nbas)
    rc = R.trl_read_checkpoint(path2.encode(), nrow, p(ev), nrow, mev, ctypes.byref(j1), p(bs),
                               nrow, nbas, ctypes.byref(j2), mev + nbas - 1, p(al),
                               mev + nbas - 1, p(be))
    assert rc == 0 and j1.value == jnd and j2.value == 0
    assert np.array_equal(al[:jnd], alpha) and np.array_equal(be[:jnd], beta)
    assert np.array_equal(ev[: nrow * (jnd + 1)].reshape(jnd + 1, nrow).T, V)


# ------------------------------------------- compiled-reference fixtures
# tests/golden/kit_cases.json holds inputs and outputs of the reference's
# own routines, called directly (tests/golden/make_kit_golden.py).  Both the
# library's kit and the numpy oracle must reproduce them exactly.

import json as _json
import os as _os

KIT_GOLD = _json.load(open(_os.path.join(_os.path.dirname(_os.path.abspath(__file__)),
                                         "golden", "kit_cases.json")))
FIELDS = KIT_GOLD["info_fields"]


def _o_info(d):
    o = O.Info()
    for f in FIELDS:
        if hasattr(o, f):
            setattr(o, f, type(getattr(o, f))(0) if isinstance(getattr(o, f), (int, float)) else 0)
    for f, v in d.items():
        setattr(o, f, v)
    return o


def _c_info(d):
    i = T.TrlInfo()
    for f, v in d.items():
        setattr(i, f, v)
    return i


def _expect_info(case):
    e = dict(case["info"])
    e.update(case["info_out"])
    return e


def _check_info(get, case, what):
    e = _expect_info(case)
    for f in FIELDS:
        got = get(f)
        want = e.get(f, 0)
        assert got == want or (got != got and want != want), (what, f, got, want)


SHUF = KIT_GOLD["shuffle"]


def _workspace(c):
    """The reference's workspace at a restart, as the fixture was made
    (tests/golden/make_kit_golden.py: [pre, lambda, res, |res|, lambda])."""
    lam, res = np.array(c["lam"]), np.array(c["res"])
    return np.concatenate([[c["pre"]], lam, res, np.abs(res), lam])


@pytest.mark.parametrize("i", range(len(SHUF)))
def test_shuffle_eig_vs_reference(kit, i):
    c = SHUF[i]
    nd, mnd, kept_in, locked = c["nd"], c["mnd"], c["kept_in"], c["locked"]
    jnd = nd + locked
    lo, ro = 1 + locked, 1 + jnd + locked
    w = _workspace(c)
    k = c["kept"]
    # the library's kit, on windows of one array like the driver's
    info = _c_info(c["info"])
    w1 = w.copy()
    tag = np.arange(nd + 2, dtype=np.int32)
    base = w1.ctypes.data
    kc = kit.trlkit_shuffle_eig(nd, mnd, P(base + 8 * lo), P(base + 8 * ro), p(tag),
                                ctypes.byref(info), kept_in)
    assert kc == k
    assert w1[lo:lo + k].tolist() == c["lam_out"][:k] and w1[ro:ro + k].tolist() == c["res_out"][:k]
    assert all(w1[lo + j] == w[lo + tag[j]] for j in range(k))
    _check_info(lambda f: getattr(info, f), c, "kit")
    # the oracle, on C-indexed windows of one array
    o = _o_info(c["info"])
    w2 = w.copy()
    with np.errstate(all="ignore"):
        ko = restart_oracle.shuffle_eig(nd, mnd, restart_oracle.Strip(w2, lo),
                                        restart_oracle.Strip(w2, ro), o, kept_in, locked)
    assert ko == k
    assert w2[lo:lo + k].tolist() == c["lam_out"][:k] and w2[ro:ro + k].tolist() == c["res_out"][:k]
    _check_info(lambda f: getattr(o, f), c, "oracle")


@pytest.mark.parametrize("i", range(len(KIT_GOLD["convergence"])))
def test_convergence_test_vs_reference(kit, i):
    c = KIT_GOLD["convergence"][i]
    nd = c["nd"]
    lam, res = np.array(c["lam"]), np.array(c["res"])
    info = _c_info(c["info"])
    l1, r1 = lam.copy(), res.copy()
    tag = np.arange(nd, dtype=np.int32)
    kit.trlkit_convergence_test(nd, p(l1), p(r1), p(tag), ctypes.byref(info))
    assert l1.tolist() == c["lam_out"] and r1.tolist() == c["res_out"]
    _check_info(lambda f: getattr(info, f), c, "kit")
    o = _o_info(c["info"])
    l2, r2 = lam.copy(), res.copy()
    O.convergence_test(nd, l2, r2, o)
    assert l2.tolist() == c["lam_out"] and r2.tolist() == c["res_out"]
    _check_info(lambda f: getattr(o, f), c, "oracle")


def test_set_locking_vs_reference(kit):
    for c in KIT_GOLD["locking"]:
        jnd, nlam = c["jnd"], c["nlam"]
        lam, res = np.array(c["lam"]), np.array(c["res"])
        yy = np.array(c["yy"])
        a1, b1, y1 = lam.copy(), res.copy(), yy.copy()
        assert kit.trlkit_set_locking(jnd, nlam, p(a1), p(b1), p(y1), c["anrm"]) == c["locked"]
        assert a1.tolist() == c["lam_out"] and b1.tolist() == c["res_out"]
        assert y1.tolist() == c["yy_out"]
        a2, b2 = lam.copy(), res.copy()
        y2 = yy.reshape(nlam, jnd).T.copy(order="F")
        assert O.set_locking(jnd, nlam, a2, b2, y2, c["anrm"]) == c["locked"]
        assert a2.tolist() == c["lam_out"] and b2.tolist() == c["res_out"]
        assert y2.T.ravel().tolist() == c["yy_out"]


def test_sort_eig_vs_reference(kit):
    for c in KIT_GOLD["sort_eig"]:
        nd, nec = c["nd"], c["nec"]
        lam, res = np.array(c["lam"]), np.array(c["res"])
        a1, b1, tag = lam.copy(), res.copy(), np.arange(nd, dtype=np.int32)
        kit.trlkit_sort_eig(nd, c["lohi"], nec, c["ref"], p(a1), p(b1), p(tag))
        a2, b2 = lam.copy(), res.copy()
        O.sort_eig(nd, c["lohi"], nec, c["ref"], a2, b2)
        # the reference's contract is the first nec pairs (the rest is work
        # space order)
        m = min(nec, nd)
        assert a1[:m].tolist() == c["lam_out"][:m] and b1[:m].tolist() == c["res_out"][:m]
        assert a2[:m].tolist() == c["lam_out"][:m] and b2[:m].tolist() == c["res_out"][:m]


def test_smooth_vs_reference(kit):
    kit.trlkit_smooth_zz.argtypes = [I, P]
    for c in KIT_GOLD["smooth"]:
        x = np.array(c["x"])
        y = x.copy()
        kit.trlkit_smooth_rr(len(y), p(y))
        assert y.tolist() == c["x_out"]
        z = list(x)
        O.smooth_rr(z := np.array(z))
        assert z.tolist() == c["x_out"]
        zc = np.array(c["z"]).view(np.complex128)
        w = zc.copy()
        kit.trlkit_smooth_zz(len(w), p(w))
        assert w.view(np.float64).tolist() == c["z_out"]
        w2 = zc.copy()
        O.smooth_zz(w2)
        assert w2.view(np.float64).tolist() == c["z_out"]


def test_sort2_vs_reference(kit):
    fn = [lambda n, s, y1, y2: sorts.dsort2(n, y1, y2),
          lambda n, s, y1, y2: sorts.dsort2a(n, y1, y2),
          lambda n, s, y1, y2: sorts.dsort2s(n, s, y1, y2),
          lambda n, s, y1, y2: sorts.dsort2su(n, s, y1, y2),
          lambda n, s, y1, y2: sorts.dsort2sd(n, s, y1, y2)]
    for c in KIT_GOLD["sort2"]:
        a1, a2 = np.array(c["a1"]), np.array(c["a2"])
        n = len(a1)
        x1, x2 = a1.copy(), a2.copy()
        tag = np.arange(n, dtype=np.int32)
        kit.trlkit_sort2(n, c["key"], c["s"], p(x1), p(x2), p(tag))
        assert x1.tolist() == c["a1_out"] and x2.tolist() == c["a2_out"]
        y1, y2 = list(a1), list(a2)
        fn[c["key"]](n, c["s"], y1, y2)
        assert y1 == c["a1_out"] and y2 == c["a2_out"]


@pytest.mark.skipif(not __import__("oracle.ref", fromlist=["available"]).available(),
                    reason="compiled reference not built here")
def test_complex_checkpoint_files_interchange_with_reference(kit, tmp_path):
    """Complex checkpoint files (interleaved (re, im) columns,
    SRC/ztrlaux.c:69-307) written by the kit are read by the reference's
    ztrl_read_checkpoint and the reference's ztrl_write_checkpoint files are
    read by the kit, with identical data."""
    from oracle import ref

    R = ref.lib()
    R.ztrl_write_checkpoint.argtypes = [ctypes.c_char_p, I, P, P, P, I, I, P, I, I]
    R.ztrl_write_checkpoint.restype = I
    R.ztrl_read_checkpoint.argtypes = [ctypes.c_char_p, I, P, I, I, ctypes.POINTER(I), P, I, I,
                                       ctypes.POINTER(I), I, P, I, P]
    R.ztrl_read_checkpoint.restype = I
    COL = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                           ctypes.POINTER(ctypes.c_double))
    kit.trlkit_write_checkpoint_e.argtypes = [ctypes.c_char_p, I, I, I, P, P, I, COL, P, P]
    kit.trlkit_write_checkpoint_e.restype = I
    kit.trlkit_read_checkpoint_e.argtypes = [ctypes.c_char_p, I, I, I, I, ctypes.POINTER(I), P, P,
                                             COL, P, P]
    kit.trlkit_read_checkpoint_e.restype = I
    rng = np.random.default_rng(6)
    nrow, me, nb = 19, 3, 2
    jnd = me + nb - 1
    Z = rng.standard_normal((nrow, me + nb)) + 1j * rng.standard_normal((nrow, me + nb))
    E, B = np.asfortranarray(Z[:, :me]), np.asfortranarray(Z[:, me:])
    alpha, beta = rng.standard_normal(jnd), rng.standard_normal(jnd)
    # reference -> kit
    path = str(tmp_path / "zref0")
    assert R.ztrl_write_checkpoint(path.encode(), nrow, p(alpha), p(beta), p(E), nrow, me, p(B),
                                   nrow, nb) == 0
    cols = {}
    buf = np.zeros(2 * nrow)

    def put(_u, i, b):
        cols[i] = np.ctypeslib.as_array(b, shape=(2 * nrow,)).copy().view(np.complex128)
        return 0

    cb = COL(put)
    j = I(0)
    a2, b2 = np.zeros(16), np.zeros(16)
    assert kit.trlkit_read_checkpoint_e(path.encode(), nrow, 2, 10, 10, ctypes.byref(j), p(a2),
                                        p(b2), cb, None, p(buf)) == 0
    assert j.value == jnd and np.array_equal(a2[:jnd], alpha) and np.array_equal(b2[:jnd], beta)
    assert all(np.array_equal(cols[i], Z[:, i]) for i in range(jnd + 1))
    # kit -> reference
    path2 = str(tmp_path / "zours0")

    def get(_u, i, b):
        np.ctypeslib.as_array(b, shape=(2 * nrow,))[:] = np.ascontiguousarray(Z[:, i]).view(np.float64)
        return 0

    cb2 = COL(get)
    assert kit.trlkit_write_checkpoint_e(path2.encode(), nrow, 2, jnd, p(alpha), p(beta), jnd + 1,
                                         cb2, None, p(buf)) == 0
    assert open(path2, "rb").read() == open(path, "rb").read()
    mev, nbas = 10, 4
    ev = np.zeros(nrow * mev, dtype=np.complex128)
    bs = np.zeros(nrow * nbas, dtype=np.complex128)
    j1, j2 = I(0), I(0)
    al, be = np.zeros(mev + nbas), np.zeros(mev + nbas)
    assert R.ztrl_read_checkpoint(path2.encode(), nrow, p(ev), nrow, mev, ctypes.byref(j1), p(bs),
                                  nrow, nbas, ctypes.byref(j2), mev + nbas - 1, p(al),
                                  mev + nbas - 1, p(be)) == 0
    assert j1.value == jnd and np.array_equal(al[:jnd], alpha)
    assert np.array_equal(ev[: nrow * (jnd + 1)].reshape(jnd + 1, nrow).T, Z)


def test_debug_fixture_structure():
    """The reference's debug logs (tests/golden/debug_cases.json) have the
    shape its call sites give (SRC/trlcore.c:1025-1029, 1199-1210,
    1497-1505): per restart cycle one orthogonality check after the inner
    loop (maxlan + 1 vectors) and one after the restart (kept + 1), plus
    one after the cycle's first step at debug level 9; recurrence checks at
    level 8 and up, each over one vector fewer than its orthogonality
    check; per-vector levels at level 8 and up."""
    import json
    import os

    path = os.path.join(os.path.dirname(__file__), "golden", "debug_cases.json")
    for c in json.load(open(path))["cases"]:
        log, nloop = c["log"], c["counters"]["nloop"]
        per = 3 if c["verbose"] > 8 else 2
        assert len(log["orth"]) == per * nloop, c["name"]
        after_inner = log["orth"][per - 2::per]
        assert all(j == c["maxlan"] + 1 for _, j in after_inner[:-1]), c["name"]
        assert log["orth"][-1][0] == c["counters"]["matvec"]
        if c["verbose"] > 7:
            assert log["vlevel"] == sum(j for _, j in log["orth"])
            assert len(log["rec"]) + log["rec_too_small"] == 2 * nloop
            if log["rec"]:
                assert log["rec"] == [j - 1 for k, (_, j) in enumerate(log["orth"])
                                      if k % per != (0 if per == 3 else -1)][:len(log["rec"])]
        else:
            assert log["vlevel"] == 0 and log["rec"] == [] and log["rec_too_small"] == 0
        assert 0 < log["max_fro"] < 1e-13
