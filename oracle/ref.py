"""ctypes wrapper of the compiled reference library oracle/_ref/libtrlan_ref.so.

Test infrastructure only (pins the restatement; CPU baseline for the bench).
Build it with `make -C oracle` (needs the reference sources; the built .so
itself only needs scipy's bundled OpenBLAS at run time).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "_ref", "libtrlan_ref.so")
_lib = None


def available():
    return os.path.exists(LIBPATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError(LIBPATH + " (run `make -C oracle`)")
        _lib = ctypes.CDLL(LIBPATH)
        P = ctypes.c_void_p
        I = ctypes.c_int
        D = ctypes.c_double
        for name in ("ref_trlan_csr", "ref_ztrlan_csr"):
            f = getattr(_lib, name)
            f.restype = I
            f.argtypes = [I, P, P, P, I, I, I, D, I, I, I, P, I, P, P, P, P, P]
        _lib.ref_info_layout.restype = I
        _lib.ref_info_layout.argtypes = [P, I]
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def info_layout():
    out = np.zeros(64, dtype=np.int64)
    k = lib().ref_info_layout(_ptr(out), 64)
    return out[:k].tolist()


def solve_csr(rp, ci, val, lohi, ned, maxlan, tol, restart, maxmv, mev, start,
              nthreads=1, complex_=False, guess=1, checkpoint=None, resume=None, debug=None):
    """Run the reference trlan()/ztrlan() on a CSR operator (info->guess =
    `guess`; 1: `start` is the start vector).  checkpoint=(cpflag, prefix)
    writes checkpoint files (trl_set_checkpoint); resume=prefix restarts
    from the file under that prefix (guess = 2, info->oldcpf);
    debug=(verbose, prefix): trl_set_debug, the log in prefix + "0".

    Returns dict(eval, evec (n x mev), resid, stat, nec, matvec, nloop, north,
    nrand, locked, klan, maxlan, ned, anrm, wall_s, crat, tol).
    """
    n = len(rp) - 1
    rp = np.ascontiguousarray(rp, dtype=np.int32)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    if complex_:
        v = np.ascontiguousarray(np.asarray(val, dtype=np.complex128)).view(np.float64)
        st = np.ascontiguousarray(np.asarray(start, dtype=np.complex128)).view(np.float64)
        evec = np.zeros(2 * n * mev)
    else:
        v = np.ascontiguousarray(val, dtype=np.float64)
        st = np.ascontiguousarray(start, dtype=np.float64)
        evec = np.zeros(n * mev)
    ev = np.zeros(mev)
    res = np.zeros(mev)
    iout = np.zeros(10, dtype=np.int32)
    dout = np.zeros(4)
    f = lib().ref_ztrlan_csr if complex_ else lib().ref_trlan_csr
    env = {"TRL_REF_GUESS": str(2 if resume is not None else int(guess))}
    if checkpoint is not None:
        env["TRL_REF_CPFLAG"] = str(int(checkpoint[0]))
        env["TRL_REF_CPFILE"] = str(checkpoint[1])
    if resume is not None:
        env["TRL_REF_OLDCPF"] = str(resume)
    if debug is not None:
        env["TRL_REF_VERBOSE"] = str(int(debug[0]))
        env["TRL_REF_LOGFILE"] = str(debug[1])
    for k in ("TRL_REF_CPFLAG", "TRL_REF_CPFILE", "TRL_REF_OLDCPF", "TRL_REF_VERBOSE",
              "TRL_REF_LOGFILE"):
        os.environ.pop(k, None)
    os.environ.update(env)
    f(n, _ptr(rp), _ptr(ci), _ptr(v), lohi, ned, maxlan, tol, restart, maxmv, mev,
      _ptr(st), nthreads, _ptr(ev), _ptr(evec), _ptr(res), _ptr(iout), _ptr(dout))
    if complex_:
        E = evec.view(np.complex128).reshape(mev, n).T
    else:
        E = evec.reshape(mev, n).T
    keys = ["stat", "nec", "matvec", "nloop", "north", "nrand", "locked", "klan",
            "maxlan", "ned"]
    out = {k: int(iout[i]) for i, k in enumerate(keys)}
    out.update(eval=ev, evec=E, resid=res, anrm=float(dout[0]), wall_s=float(dout[1]),
               crat=float(dout[2]), tol=float(dout[3]))
    return out


# ---- single routines of the compiled reference (pin the restatements) ----

def _kit_bind():
    L = lib()
    if getattr(L, "_kit_bound", False):
        return L
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.trl_shuffle_eig.argtypes = [I, I, P, P, P, P, I]
    L.trl_shuffle_eig.restype = None
    L.trl_convergence_test.argtypes = [I, P, P, P, P]
    L.trl_convergence_test.restype = None
    L.trl_sort_eig.argtypes = [I, I, I, D, P, P]
    L.trl_sort_eig.restype = None
    L.trl_set_locking.argtypes = [I, I, P, P, P, D, P]
    L.trl_set_locking.restype = None
    L.trl_smooth_rr.argtypes = [I, P]
    L.trl_smooth_rr.restype = None
    L.trl_smooth_zz.argtypes = [I, P]
    L.trl_smooth_zz.restype = None
    for name in ("dsort2", "dsort2a"):
        getattr(L, name).argtypes = [I, P, P]
        getattr(L, name).restype = None
    for name in ("dsort2s", "dsort2su_", "dsort2sd"):
        getattr(L, name).argtypes = [I, D, P, P]
        getattr(L, name).restype = None
    L._kit_bound = True
    return L


def shuffle_eig(nd, mnd, w, lam_off, res_off, info, kept, locked):
    """SRC/restart.c:1410 trl_shuffle_eig with lambda = w + lam_off and
    res = w + res_off, windows of one float64 array (changed in place), and
    a ctypes trl_info (800-byte layout); returns the new `kept`."""
    L = _kit_bind()
    k = ctypes.c_int(kept)
    base = w.ctypes.data
    L.trl_shuffle_eig(nd, mnd, ctypes.c_void_p(base + 8 * lam_off),
                      ctypes.c_void_p(base + 8 * res_off), ctypes.byref(info), ctypes.byref(k),
                      locked)
    return k.value


def convergence_test(nd, lam, res, info):
    """SRC/trlcore.c:2149 trl_convergence_test (lam/res in place)."""
    L = _kit_bind()
    wrk = np.zeros(2 * nd + 2)
    L.trl_convergence_test(nd, _ptr(lam), _ptr(res), ctypes.byref(info), _ptr(wrk))


def sort_eig(nd, lohi, nec, refv, lam, res):
    """SRC/trlcore.c:2308 trl_sort_eig (in place)."""
    _kit_bind().trl_sort_eig(nd, lohi, nec, refv, _ptr(lam), _ptr(res))


def set_locking(jnd, nlam, lam, res, yy, anrm):
    """SRC/trlcore.c:2676 trl_set_locking (in place; yy column-major
    jnd x nlam); returns the number locked."""
    k = ctypes.c_int(0)
    _kit_bind().trl_set_locking(jnd, nlam, _ptr(lam), _ptr(res), _ptr(yy), anrm, ctypes.byref(k))
    return k.value


def smooth_rr(x):
    """SRC/trlcore.c:1534 trl_smooth_rr on a float64 array (in place)."""
    _kit_bind().trl_smooth_rr(len(x), _ptr(x))


def smooth_zz(z):
    """SRC/ztrlcore.c:1289 trl_smooth_zz on a complex128 array (in place)."""
    _kit_bind().trl_smooth_zz(len(z), _ptr(z))


def sort2(key, s, a1, a2):
    """SRC/dsort2.c: dsort2 / dsort2a / dsort2s / dsort2su_ / dsort2sd by
    key 0..4 (in place)."""
    L = _kit_bind()
    n = len(a1)
    if key == 0:
        L.dsort2(n, _ptr(a1), _ptr(a2))
    elif key == 1:
        L.dsort2a(n, _ptr(a1), _ptr(a2))
    else:
        getattr(L, ["dsort2s", "dsort2su_", "dsort2sd"][key - 2])(n, s, _ptr(a1), _ptr(a2))
