"""trlhip_cgs_pass_noalloc (no allocation; partials in a fixed device array)
against trlhip_cgs_pass on a basis context holding the same data: r, h,
||h||^2 and ||r||^2 must agree bit for bit, on two calls in a row, the
second with a smaller input, for FP64, complex FP64 and the FP32 shadow of
a mixed-precision basis.  Device buffers for the allocation-free call come
from the HIP runtime directly (hipMalloc through ctypes), not from the
library."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DevBuf:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc_fn(L):
    f = L.trlhip_cgs_pass_noalloc
    P = ctypes.c_void_p
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P, ctypes.c_long, P, P, P, P, P]
    f.restype = ctypes.c_int
    return f


def _case(T, hip, f, n, m, kind, ld_pad, seed):
    rng = np.random.default_rng(seed)
    complex_ = kind == "c128"
    E = 2 if complex_ else 1
    V = rng.standard_normal((n, m + 1))
    if complex_:
        V = V + 1j * rng.standard_normal((n, m + 1))
    r = V[:, m].copy()
    ld = (n + 3) // 4 * 4 + ld_pad if not complex_ else n + ld_pad
    print(f"noalloc n={n} m={m} kind={kind} ld={ld}", flush=True)
    # the existing entry point on a context
    c = T.Ctx(n, m + 2, complex_, mixed=(kind == "r32"))
    for j in range(m + 1):
        c.set(j, V[:, j])
    hh_ref, rr2_ref, h_ref = c.cgs_pass(0, m, m)
    r_ref = c.get(m)
    c.close()
    # the allocation-free entry point on hipMalloc'ed buffers
    vt = {"r64": T.TRLHIP_R64, "c128": T.TRLHIP_C128, "r32": T.TRLHIP_R32}[kind]
    if kind == "r32":
        Vh = np.zeros((m, ld), np.float32)
        Vh[:, :n] = V[:, :m].T.astype(np.float32)
    elif complex_:
        Vh = np.zeros((m, ld), np.complex128)
        Vh[:, :n] = V[:, :m].T
    else:
        Vh = np.zeros((m, ld), np.float64)
        Vh[:, :n] = V[:, :m].T
    dV = DevBuf(hip, Vh.nbytes)
    dr = DevBuf(hip, r.nbytes)
    dV.put(Vh)
    dr.put(r)
    hh, rr2 = ctypes.c_double(), ctypes.c_double()
    h = np.zeros(E * m)
    rc = f(n, m, vt, dV.p, ld, dr.p, ctypes.byref(hh), ctypes.byref(rr2), h.ctypes.data, None)
    assert rc == 0, rc
    r_got = dr.get(np.zeros_like(r))
    dV.free()
    dr.free()
    h_got = h.view(np.complex128) if complex_ else h
    assert np.array_equal(r_got, r_ref)
    assert np.array_equal(h_got, h_ref)
    assert hh.value == hh_ref and rr2.value == rr2_ref
    # and against numpy FP64 (FP32-rounded basis for the shadow kind)
    Vm = V[:, :m].astype(np.float32).astype(np.float64) if kind == "r32" else V[:, :m]
    href = Vm.conj().T @ r
    assert np.abs(h_got - href).max() <= 1e-12 * np.abs(href).max() * np.sqrt(n)
    rn = r - Vm @ h_got
    assert np.abs(r_got - rn).max() <= 1e-11 * np.abs(r).max() * m


@pytest.mark.parametrize("kind", ["r64", "c128", "r32"])
def test_noalloc_matches_cgs_pass(gpu_lib, kind):
    T = gpu_lib
    hip = _hip()
    f = _noalloc_fn(T.lib())
    # two calls in a row: the second smaller (fewer rows and columns), so
    # stale partials of the first would show up if read
    _case(T, hip, f, 300007, 97, kind, 64 if kind != "c128" else 3, seed=1)
    _case(T, hip, f, 20001, 9, kind, 4 if kind != "c128" else 0, seed=2)


def test_noalloc_rejects_sizes(gpu_lib):
    T = gpu_lib
    hip = _hip()
    f = _noalloc_fn(T.lib())
    d = DevBuf(hip, 8 * 1024)
    hh, rr2 = ctypes.c_double(), ctypes.c_double()
    # more columns than the fixed scratch holds: a clear range error
    assert f(8, 257, T.TRLHIP_R64, d.p, 8, d.p, ctypes.byref(hh), ctypes.byref(rr2), None, None) == -5
    assert f(8, 0, T.TRLHIP_R64, d.p, 8, d.p, None, None, None, None) == -5
    # ld below n, misaligned real ld
    assert f(8, 2, T.TRLHIP_R64, d.p, 4, d.p, None, None, None, None) == -1
    assert f(8, 2, T.TRLHIP_R64, d.p, 10, d.p, None, None, None, None) == -1
    d.free()
