"""ctypes bindings to the LAPACK routines the reference calls (test oracle only).

The reference thick-restart Lanczos calls LAPACK directly for its small dense
FP64 kit (trlcore.c:2059 dsytrd_, 2066 dorgtr_, 2454 dstein_, 2577 dsyev_; its
own dstqrb.c is a one-row variant of dsteqr).  The image has no system LAPACK,
so these bind the LP64 OpenBLAS 0.3.28 that scipy bundles (symbols prefixed
``scipy_``), the same library the compiled reference in oracle/_ref links.
"""
import ctypes
import glob
import os

import numpy as np
import scipy

_libdir = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
_cands = sorted(glob.glob(os.path.join(_libdir, "libscipy_openblas*.so")))
if not _cands:
    raise ImportError("scipy's bundled OpenBLAS not found under " + _libdir)
_lib = ctypes.CDLL(_cands[0])

_i = ctypes.c_int
_pi = ctypes.POINTER(_i)
_pd = ctypes.POINTER(ctypes.c_double)
_pc = ctypes.c_char_p
_sz = ctypes.c_size_t


def _fn(name, nstr):
    f = getattr(_lib, "scipy_" + name + "_")
    f.restype = None
    return f


_dsytrd = _fn("dsytrd", 1)
_dorgtr = _fn("dorgtr", 1)
_dstein = _fn("dstein", 0)
_dsyev = _fn("dsyev", 2)
_dsteqr = _fn("dsteqr", 1)


def _p(a):
    return a.ctypes.data_as(_pd)


def _ip(a):
    return a.ctypes.data_as(_pi)


def _ref_int(v):
    return ctypes.byref(_i(v))


def dsytrd_upper(a):
    """Householder tridiagonalization (uplo='U') of a column-major nd x nd array.

    Returns (a_out, d, e, tau, info) with a_out holding the reflectors.
    """
    nd = a.shape[0]
    A = np.asfortranarray(a.copy())
    d = np.zeros(nd)
    e = np.zeros(max(nd - 1, 1))
    tau = np.zeros(max(nd - 1, 1))
    lwork = max(1, 64 * nd)
    work = np.zeros(lwork)
    info = _i(0)
    _dsytrd(_pc(b"U"), _ref_int(nd), _p(A), _ref_int(nd), _p(d), _p(e), _p(tau),
            _p(work), _ref_int(lwork), ctypes.byref(info), _sz(1))
    return A, d, e, tau, info.value


def dorgtr_upper(a, tau):
    nd = a.shape[0]
    A = np.asfortranarray(a.copy())
    lwork = max(1, 64 * nd)
    work = np.zeros(lwork)
    info = _i(0)
    _dorgtr(_pc(b"U"), _ref_int(nd), _p(A), _ref_int(nd), _p(tau), _p(work),
            _ref_int(lwork), ctypes.byref(info), _sz(1))
    return A, info.value


def dsteqr_vectors(d, e):
    """Implicit QL/QR on a symmetric tridiagonal; returns (w, Z, info)."""
    nd = len(d)
    D = np.array(d, dtype=np.float64)
    E = np.zeros(max(nd - 1, 1))
    E[: nd - 1] = e[: nd - 1]
    Z = np.zeros((nd, nd), order="F")
    work = np.zeros(max(1, 2 * nd - 2))
    info = _i(0)
    _dsteqr(_pc(b"I"), _ref_int(nd), _p(D), _p(E), _p(Z), _ref_int(nd), _p(work),
            ctypes.byref(info), _sz(1))
    return D, Z, info.value


def dstein(d, e, w, iblock, isplit):
    """Inverse iteration (dstein) for eigenvectors of a symmetric tridiagonal.

    iblock/isplit are int32 arrays (caller-owned, like the reference's iwrk);
    returns (Z (nd x m), iwork, ifail, info).
    """
    nd = len(d)
    m = len(w)
    D = np.ascontiguousarray(d, dtype=np.float64)
    E = np.zeros(max(nd, 1))
    E[: nd - 1] = e[: nd - 1]
    W = np.ascontiguousarray(w, dtype=np.float64)
    Z = np.zeros((nd, max(m, 1)), order="F")
    work = np.zeros(5 * nd + 1)
    iwork = np.zeros(nd + 1, dtype=np.int32)
    ifail = np.zeros(max(m, 1), dtype=np.int32)
    info = _i(0)
    _dstein(_ref_int(nd), _p(D), _p(E), _ref_int(m), _p(W), _ip(iblock), _ip(isplit),
            _p(Z), _ref_int(nd), _p(work), _ip(iwork), _ip(ifail), ctypes.byref(info))
    return Z[:, :m], iwork, ifail, info.value


def dsyev_vectors(a):
    """Symmetric eigensolve (jobz='V', uplo='U'); returns (w, Z, info)."""
    nd = a.shape[0]
    A = np.asfortranarray(a.copy())
    w = np.zeros(nd)
    lwork = max(1, 64 * nd)
    work = np.zeros(lwork)
    info = _i(0)
    _dsyev(_pc(b"V"), _pc(b"U"), _ref_int(nd), _p(A), _ref_int(nd), _p(w), _p(work),
           _ref_int(lwork), ctypes.byref(info), _sz(1), _sz(1))
    return w, A, info.value
