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
              nthreads=1, complex_=False):
    """Run the reference trlan()/ztrlan() on a CSR operator.

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
