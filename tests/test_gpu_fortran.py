"""The pointer-argument callback convention (the reference's
TRL_FORTRAN_COMPATIBLE build, INC/trlan.h:9-36, INC/ztrlan.h:30-31): each
*_fc entry point, given a pointer-argument operator (and user_ortho),
returns what the ordinary entry point returns for the same operator written
as a C callback, bit for bit, and leaves the caller's info->mvparam as it
was.  Host operators (a dense numpy matrix), so every operator call crosses
the adapter."""
import ctypes

import numpy as np
import pytest

from nutrlan_amd import trlan as T
from nutrlan_amd import workloads as W

pytestmark = pytest.mark.gpu
P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
PI, PD = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)

C_OP = ctypes.CFUNCTYPE(None, I, I, PD, I, PD, I, P)
F_OP = ctypes.CFUNCTYPE(None, PI, PI, PD, PI, PD, PI)
C_UO = ctypes.CFUNCTYPE(None, I, PD, PD)
F_UO = ctypes.CFUNCTYPE(None, PI, PD, PD)

SENTINEL = 0x5EED


def _p(a):
    return a.ctypes.data_as(P)


def _apply(H, x, ldx, y, ldy, n, ncol, cplx):
    s = 2 if cplx else 1
    xs = np.ctypeslib.as_array(x, shape=(s * ldx * ncol,))
    ys = np.ctypeslib.as_array(y, shape=(s * ldy * ncol,))
    for j in range(ncol):
        xv = xs[s * j * ldx:s * (j * ldx + n)]
        if cplx:
            xv = xv.view(np.complex128)
            ys[s * j * ldy:s * (j * ldy + n)] = (H @ xv).view(np.float64)
        else:
            ys[j * ldy:j * ldy + n] = H @ xv


def _ops(H, bad, cplx=False):
    """The same operator as a C callback (which receives the caller's
    mvparam) and as a pointer-argument one; anything unexpected goes to
    `bad` (an exception raised in a ctypes callback would only be printed)."""
    n = H.shape[0]

    def c_op(nrow, ncol, x, ldx, y, ldy, mvparam):
        if mvparam != SENTINEL or nrow != n:
            bad.append(("c_op", nrow, mvparam))
        _apply(H, x, ldx, y, ldy, nrow, ncol, cplx)

    def f_op(pn, pnc, x, pldx, y, pldy):
        if pn[0] != n:
            bad.append(("f_op", pn[0]))
        _apply(H, x, pldx[0], y, pldy[0], pn[0], pnc[0], cplx)

    return C_OP(c_op), F_OP(f_op)


def _uos(u):
    """user_ortho removing the component along unit vector u, both ways."""
    def body(n, x, y):
        xv = np.ctypeslib.as_array(x, shape=(n,))
        c = float(u @ xv)
        xv -= c * u
        if y:
            y[0] = abs(c)

    return C_UO(lambda n, x, y: body(n, x, y)), F_UO(lambda pn, x, y: body(pn[0], x, y))


def _info(n, maxlan, ned, lohi=-1):
    info = T.init_info(n, maxlan, lohi, ned, 1e-10, 1, 4000)
    T.lib().trl_set_iguess(ctypes.byref(info), 0, 1, 0, None)
    info.mvparam = SENTINEL
    return info


def _bind(L):
    L.trlan_fc.argtypes = [P, P, ctypes.POINTER(T.TrlInfo), I, I, P, P, I, I, P]
    L.trlan_fc.restype = None
    L.trl_check_ritz_fc.argtypes = [P, ctypes.POINTER(T.TrlInfo), I, I, P, I, P, PI, P, P, P, I]
    L.trl_check_ritz_fc.restype = None
    L.trl_rayleigh_quotients_fc.argtypes = [P, ctypes.POINTER(T.TrlInfo), I, P, I, P, P]
    L.trl_rayleigh_quotients_fc.restype = None
    L.trl_ritz_projection_fc.argtypes = [P, ctypes.POINTER(T.TrlInfo), I, P, I, P, P, I, P]
    L.trl_ritz_projection_fc.restype = None
    L.ztrlan_fc.argtypes = [P, ctypes.POINTER(T.TrlInfo), I, I, P, P, I, P, I, P, I]
    L.ztrlan_fc.restype = None
    L.ztrl_check_ritz_fc.argtypes = [P, ctypes.POINTER(T.TrlInfo), I, I, P, I, P, PI, P, P, P, I]
    L.ztrl_check_ritz_fc.restype = None
    return L


COUNTERS = ("stat", "nec", "matvec", "nloop", "north", "nrand", "locked")


def _printed(capfd):
    """What the library printed to stdout since the last call."""
    ctypes.CDLL(None).fflush(None)
    return capfd.readouterr().out


@pytest.mark.parametrize("with_uo", [False, True])
def test_trlan_fc_equals_trlan(gpu_lib, with_uo, capfd):
    L = _bind(T.lib())
    rp, ci, v = W.laplace3d_csr(8, seed=2)
    H = W.csr_dense(rp, ci, v)
    n, mev, ned, maxlan = H.shape[0], 8, 6, 30
    x0 = W.start_vector(n, seed=3)
    bad = []
    c_op, f_op = _ops(H, bad)
    u = np.zeros(n)
    u[0] = 1.0
    c_uo, f_uo = _uos(u) if with_uo else (None, None)
    outs = []
    for fn, op, uo in ((L.trlan, c_op, c_uo), (L.trlan_fc, f_op, f_uo)):
        info = _info(n, maxlan, ned)
        ev = np.zeros(mev)
        evec = np.zeros((mev, n))
        evec[0] = x0
        wrk = np.zeros(4 * maxlan * maxlan)
        fn(ctypes.cast(op, P), ctypes.cast(uo, P) if uo else None, ctypes.byref(info), n, mev,
           _p(ev), _p(evec), n, len(wrk), _p(wrk))
        assert info.mvparam == SENTINEL
        outs.append((info, ev, evec, wrk))
    (a, eva, va, wa), (b, evb, vb, wb) = outs
    print(f"trlan vs trlan_fc (user_ortho {with_uo}): nec {a.nec} matvec {a.matvec} "
          f"nloop {a.nloop}", flush=True)
    assert a.stat == 0 and a.nec >= ned
    for f in COUNTERS:
        assert getattr(a, f) == getattr(b, f), f
    assert np.array_equal(eva, evb) and np.array_equal(va, vb) and np.array_equal(wa, wb)

    # the post-processing entry points on the returned pairs
    nec = a.nec
    chk_a, chk_b = ctypes.c_int(0), ctypes.c_int(0)
    _printed(capfd)
    res = []
    for fn, op, chk in ((L.trl_check_ritz, c_op, chk_a), (L.trl_check_ritz_fc, f_op, chk_b)):
        info = _info(n, maxlan, ned)
        info.nec = nec
        rv = np.ascontiguousarray(va[:nec].copy())
        beta = np.zeros(nec)
        w = np.zeros(4 * n)
        fn(ctypes.cast(op, P), ctypes.byref(info), n, nec, _p(rv), n, _p(eva.copy()),
           ctypes.byref(chk), _p(beta), _p(eva.copy()), _p(w), len(w))
        assert info.mvparam == SENTINEL
        res.append(_printed(capfd))
    assert "TRL_CHECK_RITZ" in res[0]
    assert chk_a.value == chk_b.value and res[0] == res[1]
    res = []
    for fn, op in ((L.trl_rayleigh_quotients, c_op), (L.trl_rayleigh_quotients_fc, f_op)):
        info = _info(n, maxlan, ned)
        rv = np.ascontiguousarray(va[:nec].copy())
        er = np.zeros(2 * nec)
        fn(ctypes.cast(op, P), ctypes.byref(info), nec, _p(rv), n, _p(er), None)
        assert info.mvparam == SENTINEL
        res.append((rv, er))
    assert all(np.array_equal(x, y) for x, y in zip(*res))
    res = []
    for fn, op in ((L.trl_ritz_projection, c_op), (L.trl_ritz_projection_fc, f_op)):
        info = _info(n, maxlan, ned)
        rv = np.ascontiguousarray(va[:nec].copy())
        er = np.zeros(2 * nec)
        w = np.zeros(2 * nec * nec + 8 * nec)
        fn(ctypes.cast(op, P), ctypes.byref(info), nec, _p(rv), n, _p(er), _p(w), len(w), None)
        assert info.mvparam == SENTINEL
        res.append((rv, er))
    assert all(np.array_equal(x, y) for x, y in zip(*res))
    assert not bad, bad[:5]


def test_ztrlan_fc_equals_ztrlan(gpu_lib, capfd):
    L = _bind(T.lib())
    rp, ci, v = W.anderson3d_csr(6, seed=4)
    H = W.csr_dense(rp, ci, v)
    n, mev, ned, maxlan = H.shape[0], 8, 5, 30
    x0 = W.start_vector(n, seed=5, complex_=True)
    bad = []
    c_op, f_op = _ops(H, bad, cplx=True)
    outs = []
    for fn, op in ((L.ztrlan, c_op), (L.ztrlan_fc, f_op)):
        info = _info(n, maxlan, ned)
        ev = np.zeros(mev)
        evec = np.zeros((mev, n), np.complex128)
        evec[0] = x0
        misc = np.zeros(mev * 4, np.complex128)
        dw = np.zeros(4 * maxlan * maxlan)
        fn(ctypes.cast(op, P), ctypes.byref(info), n, mev, _p(ev), _p(evec), n, _p(misc),
           len(misc), _p(dw), len(dw))
        assert info.mvparam == SENTINEL
        outs.append((info, ev, evec, misc))
    (a, eva, va, ma), (b, evb, vb, mb) = outs
    print(f"ztrlan vs ztrlan_fc: nec {a.nec} matvec {a.matvec} nloop {a.nloop}", flush=True)
    assert a.stat == 0 and a.nec >= ned
    for f in COUNTERS:
        assert getattr(a, f) == getattr(b, f), f
    assert np.array_equal(eva, evb) and np.array_equal(va, vb) and np.array_equal(ma, mb)
    nec = a.nec
    _printed(capfd)
    res = []
    for fn, op in ((L.ztrl_check_ritz, c_op), (L.ztrl_check_ritz_fc, f_op)):
        info = _info(n, maxlan, ned)
        info.nec = nec
        chk = ctypes.c_int(0)
        rv = np.ascontiguousarray(va[:nec].copy())
        beta = np.zeros(nec)
        w = np.zeros(4 * n, np.complex128)
        fn(ctypes.cast(op, P), ctypes.byref(info), n, nec, _p(rv), n, _p(eva.copy()),
           ctypes.byref(chk), _p(beta), _p(eva.copy()), _p(w), len(w))
        assert info.mvparam == SENTINEL
        res.append((chk.value, _printed(capfd)))
    assert "CHECK_RITZ" in res[0][1].upper()
    assert res[0] == res[1]
    assert not bad, bad[:5]
