"""Python mirror of the library's C API (include/trlan.h, ztrlan.h, trlhip.h).

Loads the in-tree libnutrlan_amd.so (built by `make -C nutrlan_amd/csrc`)
and exposes:
  TrlInfo                 ctypes mirror of trl_info (same 800-byte layout)
  Csr                     a device CSR operator (trlhip_csr_create)
  solve_csr(...)          trlan()/ztrlan() with the device operator
  solve_hostop(...)       trlan() with a Python host operator callback
  stats_enable/stats_get  per-kernel HIP-event timing of the device layer

There is no fallback: if the library is missing or a GPU call fails, these
raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libnutrlan_amd.so")

TRLHIP_R64, TRLHIP_C128, TRLHIP_R32 = 0, 1, 2
TRLHIP_STAT_DEVICE = -1201
KERNELS = ["spmv", "recur", "proj", "update", "ritz", "misc",
           "b_spmv", "b_recur", "b_proj", "b_update", "b_ritz"]  # b_: batched kernels

_lib = None


class TrlInfo(ctypes.Structure):
    """Field-for-field mirror of trl_info (include/trlan.h)."""

    _fields_ = [
        ("stat", ctypes.c_int), ("lohi", ctypes.c_int), ("ned", ctypes.c_int),
        ("nec", ctypes.c_int), ("tol", ctypes.c_double), ("mpicomp", ctypes.c_void_p),
        ("maxlan", ctypes.c_int), ("klan", ctypes.c_int), ("maxmv", ctypes.c_int),
        ("restart", ctypes.c_int), ("locked", ctypes.c_int), ("guess", ctypes.c_int),
        ("matvec", ctypes.c_int), ("nloop", ctypes.c_int), ("north", ctypes.c_int),
        ("nrand", ctypes.c_int), ("flop", ctypes.c_int), ("flop_h", ctypes.c_int),
        ("flop_r", ctypes.c_int), ("rflp", ctypes.c_double), ("rflp_h", ctypes.c_double),
        ("rflp_r", ctypes.c_double), ("clk_rate", ctypes.c_long), ("clk_max", ctypes.c_long),
        ("clk_tot", ctypes.c_long), ("clk_op", ctypes.c_long), ("clk_orth", ctypes.c_long),
        ("clk_res", ctypes.c_long), ("tick_t", ctypes.c_double), ("tick_o", ctypes.c_double),
        ("tick_h", ctypes.c_double), ("tick_r", ctypes.c_double), ("clk_in", ctypes.c_int),
        ("wrds_in", ctypes.c_int), ("clk_out", ctypes.c_int), ("wrds_out", ctypes.c_int),
        ("anrm", ctypes.c_double), ("my_pe", ctypes.c_int), ("npes", ctypes.c_int),
        ("nloc", ctypes.c_int), ("ntot", ctypes.c_int), ("verbose", ctypes.c_int),
        ("crat", ctypes.c_double), ("trgt", ctypes.c_double), ("tres", ctypes.c_double),
        ("tmv", ctypes.c_int), ("avgm", ctypes.c_double), ("cfac", ctypes.c_double),
        ("ptres", ctypes.c_double), ("predicted_crate", ctypes.c_double),
        ("mgap", ctypes.c_double), ("mgamma", ctypes.c_double), ("gamma0", ctypes.c_double),
        ("old_target", ctypes.c_double), ("target_id", ctypes.c_int),
        ("old_locked", ctypes.c_int), ("k1", ctypes.c_int), ("k2", ctypes.c_int),
        ("k", ctypes.c_int), ("rfact", ctypes.c_double), ("ref", ctypes.c_double),
        ("log_io", ctypes.c_int), ("log_fp", ctypes.c_void_p), ("log_file", ctypes.c_char * 128),
        ("cpflag", ctypes.c_int), ("cpio", ctypes.c_int), ("cpt_fp", ctypes.c_void_p),
        ("cpfile", ctypes.c_char * 128), ("oldcpf", ctypes.c_char * 128),
        ("mvparam", ctypes.c_void_p),
    ]


MATVEC = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                          ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                          ctypes.c_void_p)


def available():
    return os.path.exists(LIBPATH)


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not available():
            raise RuntimeError(
                "libnutrlan_amd.so not built: run `make -C nutrlan_amd/csrc` "
                "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIBPATH)
        P, I, D, Lg = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
        L.trl_init_info.argtypes = [ctypes.POINTER(TrlInfo), I, I, I, I, D, I, I, P]
        L.trl_init_info.restype = None
        L.trl_set_iguess.argtypes = [ctypes.POINTER(TrlInfo), I, I, I, ctypes.c_char_p]
        L.trl_set_checkpoint.argtypes = [ctypes.POINTER(TrlInfo), I, ctypes.c_char_p]
        L.trl_set_checkpoint.restype = None
        L.trl_set_iguess.restype = None
        L.trl_set_restart.argtypes = [ctypes.POINTER(TrlInfo), D]
        L.trl_set_restart.restype = None
        L.trlan.argtypes = [P, P, ctypes.POINTER(TrlInfo), I, I, P, P, I, I, P]
        L.trlan.restype = None
        L.ztrlan.argtypes = [P, ctypes.POINTER(TrlInfo), I, I, P, P, I, P, I, P, I]
        L.ztrlan.restype = None
        L.trlhip_csr_create.argtypes = [ctypes.POINTER(P), I, Lg, P, P, P, I]
        L.trlhip_csr_create.restype = I
        L.trlhip_csr_destroy.argtypes = [P]
        L.trlhip_csr_destroy.restype = I
        L.trlhip_stats_enable.argtypes = [I]
        L.trlhip_stats_enable.restype = None
        L.trlhip_stats_get.argtypes = [P, I]
        L.trlhip_stats_get.restype = I
        L.trlhip_device_count.argtypes = []
        L.trlhip_device_count.restype = I
        L.trlhip_csr_set_spmv.argtypes = [P, I]
        L.trlhip_csr_set_spmv.restype = I
        L.trlhip_proj_tiles.argtypes = [I, I]
        L.trlhip_proj_tiles.restype = I
        L.trlhip_proj_tiles_get.argtypes = [P, P]
        L.trlhip_proj_tiles_get.restype = I
        L.trlhip_proj_autotune.argtypes = [I, I, I, P, P]
        L.trlhip_proj_autotune.restype = I
        L.trl_check_ritz.argtypes = [P, ctypes.POINTER(TrlInfo), I, I, P, I, P,
                                     ctypes.POINTER(I), P, P, P, I]
        L.trl_check_ritz.restype = None
        L.trl_rayleigh_quotients.argtypes = [P, ctypes.POINTER(TrlInfo), I, P, I, P, P]
        L.trl_rayleigh_quotients.restype = None
        L.trl_ritz_projection.argtypes = [P, ctypes.POINTER(TrlInfo), I, P, I, P, P, I, P]
        L.trl_ritz_projection.restype = None
        L.ztrl_check_ritz.argtypes = [P, ctypes.POINTER(TrlInfo), I, I, P, I, P,
                                      ctypes.POINTER(I), P, P, P, I]
        L.ztrl_check_ritz.restype = None
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def fn_addr(name):
    return ctypes.cast(getattr(lib(), name), ctypes.c_void_p).value


class Csr:
    """Device CSR operator (trlhip_csr); usable as mvparam with the op
    returned by .op (trlhip_csr_matvec / trlhip_csr_zmatvec)."""

    def __init__(self, rp, ci, val, complex_=False, mixed=False):
        """mixed: mark the operator for the mixed-precision solver
        (TRLHIP_R32: FP32 shadow basis for Gram-Schmidt, FP64 elsewhere)."""
        n = len(rp) - 1
        self.n = n
        self.nnz = len(ci)
        self.complex = complex_
        self.mixed = bool(mixed) and not complex_
        rp = np.ascontiguousarray(rp, dtype=np.int32)
        ci = np.ascontiguousarray(ci, dtype=np.int32)
        if complex_:
            v = np.ascontiguousarray(np.asarray(val, dtype=np.complex128)).view(np.float64)
            vt = TRLHIP_C128
        else:
            v = np.ascontiguousarray(val, dtype=np.float64)
            vt = TRLHIP_R32 if self.mixed else TRLHIP_R64
        h = ctypes.c_void_p()
        rc = lib().trlhip_csr_create(ctypes.byref(h), n, len(ci), _ptr(rp), _ptr(ci), _ptr(v), vt)
        if rc != 0:
            raise RuntimeError(f"trlhip_csr_create failed ({rc})")
        self.handle = h
        self.op = fn_addr("trlhip_csr_zmatvec" if complex_ else "trlhip_csr_matvec")

    @property
    def packed(self):
        """True when the CSR-stream SpMV reads this matrix's packed nonzeros
        (trlhip_csr_is_packed)."""
        L = lib()
        L.trlhip_csr_is_packed.argtypes = [ctypes.c_void_p]
        L.trlhip_csr_is_packed.restype = ctypes.c_int
        return bool(L.trlhip_csr_is_packed(self.handle))

    @property
    def paneled(self):
        """True when the SpMV in use is the column-panel kernel
        (trlhip_csr_is_paneled)."""
        L = lib()
        L.trlhip_csr_is_paneled.argtypes = [ctypes.c_void_p]
        L.trlhip_csr_is_paneled.restype = ctypes.c_int
        return bool(L.trlhip_csr_is_paneled(self.handle))

    @property
    def stencil(self):
        """(D, L) when the operator is a periodic nearest-neighbour stencil on
        an L^D lattice whose SpMV forms its columns from the lattice
        (trlhip_csr_stencil), else (0, 0)."""
        L = lib()
        L.trlhip_csr_stencil.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.trlhip_csr_stencil.restype = ctypes.c_int
        side = ctypes.c_int(0)
        d = L.trlhip_csr_stencil(self.handle, ctypes.byref(side))
        return (d, side.value) if d else (0, 0)

    def set_spmv(self, lpr):
        """Use the SpMV kernel with `lpr` lanes per row (0: CSR-stream), as
        a cached autotuner choice (trlhip_csr_set_spmv)."""
        rc = lib().trlhip_csr_set_spmv(self.handle, int(lpr))
        if rc != 0:
            raise RuntimeError(f"trlhip_csr_set_spmv failed ({rc})")

    def autotune(self, reps=10):
        """Time the SpMV kernels on this matrix and keep the fastest
        (trlhip_csr_autotune); returns (lanes per row, 0 = CSR-stream;
        ms per SpMV)."""
        L = lib()
        L.trlhip_csr_autotune.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)]
        L.trlhip_csr_autotune.restype = ctypes.c_int
        lpr, ms = ctypes.c_int(), ctypes.c_float()
        rc = L.trlhip_csr_autotune(self.handle, reps, ctypes.byref(lpr), ctypes.byref(ms))
        if rc != 0:
            raise RuntimeError(f"trlhip_csr_autotune failed ({rc})")
        return lpr.value, ms.value

    def close(self):
        if self.handle:
            lib().trlhip_csr_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def init_info(n, maxlan, lohi, ned, tol, restart, maxmv):
    info = TrlInfo()
    lib().trl_init_info(ctypes.byref(info), n, maxlan, lohi, ned, tol, restart, maxmv, None)
    return info


def _result(info, ev, evec, wrk, wall):
    keys = ["stat", "nec", "matvec", "nloop", "north", "nrand", "locked", "klan", "maxlan", "ned"]
    out = {k: int(getattr(info, k)) for k in keys}
    out.update(eval=ev, evec=evec, resid=wrk, anrm=float(info.anrm), wall_s=wall,
               crat=float(info.crat), tol=float(info.tol))
    return out


def proj_tiles(cb_fp64=0, cb_fp32=0):
    """Columns per Gram-Schmidt projection block for FP64 and FP32-shadow
    bases (4, 8 or 16; 0 = the default 8), process-wide
    (trlhip_proj_tiles)."""
    rc = lib().trlhip_proj_tiles(int(cb_fp64), int(cb_fp32))
    if rc != 0:
        raise RuntimeError(f"trlhip_proj_tiles failed ({rc})")


def proj_tiles_get():
    a, b = ctypes.c_int(), ctypes.c_int()
    lib().trlhip_proj_tiles_get(ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def proj_autotune(n, m, reps=10):
    """Time the projection tiles on an n x m basis and keep the fastest
    (trlhip_proj_autotune); returns ((cb_fp64, cb_fp32), (ms_fp64,
    ms_fp32))."""
    cb = (ctypes.c_int * 2)()
    ms = (ctypes.c_float * 2)()
    rc = lib().trlhip_proj_autotune(int(n), int(m), int(reps), cb, ms)
    if rc != 0:
        raise RuntimeError(f"trlhip_proj_autotune failed ({rc})")
    return (cb[0], cb[1]), (ms[0], ms[1])


def autotune_cached(path, n, m, csrs=(), reps=10):
    """The autotuner's choices through a JSON cache at `path`: projection
    tiles for an n x m basis and the SpMV kernel of each operator in csrs
    (keyed by rows and nonzeros); missing entries are measured and added.
    Returns the cache dict."""
    import json
    import os

    cache = {"proj": {}, "spmv": {}}
    if os.path.exists(path):
        with open(path) as f:
            cache.update(json.load(f))
    key = f"{int(n)}x{int(m)}"
    if key in cache["proj"]:
        proj_tiles(*cache["proj"][key])
    else:
        cache["proj"][key] = list(proj_autotune(n, m, reps)[0])
    for A in csrs:
        k = f"{A.n}:{A.nnz}"
        if k in cache["spmv"]:
            A.set_spmv(cache["spmv"][k])
        else:
            cache["spmv"][k] = A.autotune(reps)[0]
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    return cache


def output_buffers(n, mev):
    """(eval, evec, wrk) arrays for one real solve, pages touched, for
    reuse across calls (out= of solve_csr / solve_batch)."""
    return np.zeros(mev), np.ones((mev, n)), np.zeros(mev)


def solve_csr(A, lohi, ned, maxlan, tol, restart, maxmv, mev, start, rfact=0.0,
              checkpoint=None, resume=None, out=None, guess=1):
    """trlan()/ztrlan() on the device CSR operator A with start vector
    `start` (guess = 1; guess <= 0: the reference's own start, `start` is
    not read).  checkpoint=(cpflag, prefix) writes checkpoint
    files (trl_set_checkpoint); resume=prefix restarts from the file written
    under that prefix (guess = 2, oldcpf).  out = output_buffers(n, mev)
    reuses the caller's arrays (real operators; the returned dict refers to
    them).  Returns the same dict keys as oracle/ref.py."""
    import time

    n = A.n
    info = init_info(n, maxlan, lohi, ned, tol, restart, maxmv)
    if checkpoint is not None:
        info.verbose = -1
        lib().trl_set_checkpoint(ctypes.byref(info), int(checkpoint[0]),
                                 str(checkpoint[1]).encode())
    if resume is not None:
        info.oldcpf = str(resume).encode()
        lib().trl_set_iguess(ctypes.byref(info), 0, 2, 0, None)
    else:
        lib().trl_set_iguess(ctypes.byref(info), 0, int(guess), 0, None)
    if rfact:
        lib().trl_set_restart(ctypes.byref(info), rfact)
    info.mvparam = A.handle.value
    ev = np.zeros(mev)
    if A.complex:
        evec = np.zeros((mev, n), dtype=np.complex128)
        evec[0] = start
        misc = np.zeros(max(mev, 16), dtype=np.complex128)
        t0 = time.perf_counter()
        lib().ztrlan(A.op, ctypes.byref(info), n, mev, _ptr(ev), _ptr(evec), n, _ptr(misc),
                     len(misc), None, 0)
        wall = time.perf_counter() - t0
        wrk = misc.real[:mev].copy()
    else:
        if out is not None:
            ev, evec, wrk = out
        else:
            evec = np.zeros((mev, n))
            wrk = np.zeros(mev)
        evec[0] = start
        t0 = time.perf_counter()
        lib().trlan(A.op, None, ctypes.byref(info), n, mev, _ptr(ev), _ptr(evec), n, mev,
                    _ptr(wrk))
        wall = time.perf_counter() - t0
    if info.stat == TRLHIP_STAT_DEVICE:
        raise RuntimeError("device failure inside trlan()")
    return _result(info, ev, evec.T, wrk, wall)


class TrlBatchJob(ctypes.Structure):
    """Mirror of trl_batch_job (include/trlan_batch.h)."""
    _fields_ = [("A", ctypes.c_void_p), ("info", ctypes.POINTER(TrlInfo)),
                ("nrow", ctypes.c_int), ("mev", ctypes.c_int),
                ("eval", ctypes.c_void_p), ("evec", ctypes.c_void_p),
                ("lde", ctypes.c_int), ("lwrk", ctypes.c_int), ("wrk", ctypes.c_void_p),
                ("dstart", ctypes.c_void_p)]


def solve_batch(problems, out=None, dstarts=None):
    """trlan_batch over problems = [(A, lohi, ned, maxlan, tol, restart,
    maxmv, mev, start), ...] with real device operators A (Csr, mixed or
    not).  out = [output_buffers(n, mev) per problem] reuses the caller's
    arrays (the returned dicts refer to them).  dstarts = [device address of
    an FP32 start vector per problem] (trl_batch_job.dstart) replaces the
    host start vectors (start may then be None).  Returns (rc, [result dict
    per problem], wall seconds)."""
    import time

    L = lib()
    L.trlan_batch.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.trlan_batch.restype = ctypes.c_int
    jobs = (TrlBatchJob * len(problems))()
    keep = []
    for j, (A, lohi, ned, maxlan, tol, restart, maxmv, mev, start) in enumerate(problems):
        n = A.n
        info = init_info(n, maxlan, lohi, ned, tol, restart, maxmv)
        L.trl_set_iguess(ctypes.byref(info), 0, 1, 0, None)
        if out is not None:
            ev, evec, wrk = out[j]
        else:
            ev, evec, wrk = np.zeros(mev), np.zeros((mev, n)), np.zeros(mev)
        if dstarts is not None:
            jobs[j].dstart = dstarts[j]
        else:
            evec[0] = start
        keep.append((info, ev, evec, wrk))
        jobs[j].A = A.handle.value
        jobs[j].info = ctypes.pointer(info)
        jobs[j].nrow = n
        jobs[j].mev = mev
        jobs[j].eval = ev.ctypes.data
        jobs[j].evec = evec.ctypes.data
        jobs[j].lde = n
        jobs[j].lwrk = mev
        jobs[j].wrk = wrk.ctypes.data
    t0 = time.perf_counter()
    rc = L.trlan_batch(jobs, len(problems))
    wall = time.perf_counter() - t0
    return rc, [_result(info, ev, evec.T, wrk, wall) for info, ev, evec, wrk in keep], wall


def solve_hostop(matvec, n, lohi, ned, maxlan, tol, restart, maxmv, mev, start):
    """trlan() with a Python host operator y = matvec(x) (staged through
    host memory every step; the basis still lives on the device)."""
    import time

    def cb(nrow, ncol, x, ldx, y, ldy, _p):
        for j in range(ncol):
            xv = np.ctypeslib.as_array(x, shape=(ldx * (j + 1),))[j * ldx:j * ldx + nrow]
            yv = np.ctypeslib.as_array(y, shape=(ldy * (j + 1),))
            yv[j * ldy:j * ldy + nrow] = matvec(xv.copy())

    cfn = MATVEC(cb)
    info = init_info(n, maxlan, lohi, ned, tol, restart, maxmv)
    lib().trl_set_iguess(ctypes.byref(info), 0, 1, 0, None)
    ev = np.zeros(mev)
    evec = np.zeros((mev, n))
    evec[0] = start
    wrk = np.zeros(mev)
    t0 = time.perf_counter()
    lib().trlan(ctypes.cast(cfn, ctypes.c_void_p), None, ctypes.byref(info), n, mev, _ptr(ev),
                _ptr(evec), n, mev, _ptr(wrk))
    wall = time.perf_counter() - t0
    if info.stat == TRLHIP_STAT_DEVICE:
        raise RuntimeError("device failure inside trlan()")
    return _result(info, ev, evec.T, wrk, wall)


def stats_enable(on=True, every=1):
    """Per-kernel event timing of every `every`-th launch of each kind."""
    lib().trlhip_stats_enable(int(every) if on else 0)


def stats_get():
    """{kernel: (timed launches, their total ms, their algorithmic bytes)}
    since stats_enable."""
    out = np.zeros(3 * len(KERNELS))
    lib().trlhip_stats_get(_ptr(out), len(out))
    return {k: (int(out[3 * i]), float(out[3 * i + 1]), float(out[3 * i + 2]))
            for i, k in enumerate(KERNELS)}


def stats_flops():
    """{kernel: floating-point operations of the timed launches} since
    stats_enable (the restart GEMMs' 2 n jl k per launch; 0 for kernels
    whose bound is bandwidth)."""
    L = lib()
    L.trlhip_stats_flops.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.trlhip_stats_flops.restype = ctypes.c_int
    out = np.zeros(len(KERNELS))
    L.trlhip_stats_flops(_ptr(out), len(out))
    return {k: float(out[i]) for i, k in enumerate(KERNELS)}


def device_count():
    return int(lib().trlhip_device_count())


class Ctx:
    """A device basis context (trlhip_ctx) for kernel-level checks."""

    def __init__(self, n, ncols, complex_=False, mixed=False):
        L = lib()
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        for name, args in [
            ("trlhip_ctx_create", [ctypes.POINTER(P), I, I, I]),
            ("trlhip_ctx_destroy", [P]),
            ("trlhip_col_set", [P, I, P]),
            ("trlhip_col_get", [P, I, P]),
            ("trlhip_col_copy", [P, I, I]),
            ("trlhip_col_scale", [P, I, D]),
            ("trlhip_spmv", [P, P, I, I, P]),
            ("trlhip_lanczos_step", [P, P, I, I, D, I, P, P]),
            ("trlhip_restart_step", [P, P, I, P, I, I, P, P]),
            ("trlhip_cgs_pass", [P, I, I, I, P, P, P]),
            ("trlhip_proj", [P, I, I, I, P]),
            ("trlhip_update", [P, I, I, P, I, P]),
            ("trlhip_ritz", [P, I, I, I, P, I]),
            ("trlhip_col_scale_copy", [P, I, I, D]),
            ("trlhip_step_cgs", [P, P, I, I, I, D, P, I, I, P]),
            ("trlhip_refine", [P, P, I, I, P, P]),
            ("trlhip_residuals", [P, P, I, P, P]),
            ("trlhip_cols_get", [P, I, I, P, ctypes.c_long]),
            ("trlhip_gram", [P, P, I, I, P, P]),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = I
        self.n, self.ncols, self.complex = n, ncols, complex_
        self.E = 2 if complex_ else 1
        h = ctypes.c_void_p()
        vt = TRLHIP_C128 if complex_ else (TRLHIP_R32 if mixed else TRLHIP_R64)
        rc = L.trlhip_ctx_create(ctypes.byref(h), n, ncols, vt)
        if rc != 0:
            raise RuntimeError(f"trlhip_ctx_create failed ({rc})")
        self.h = h

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed ({rc})")

    def set(self, col, x):
        x = np.ascontiguousarray(x, dtype=np.complex128 if self.complex else np.float64)
        self._chk(lib().trlhip_col_set(self.h, col, _ptr(x)), "col_set")

    def get(self, col):
        x = np.zeros(self.n, dtype=np.complex128 if self.complex else np.float64)
        self._chk(lib().trlhip_col_get(self.h, col, _ptr(x)), "col_get")
        return x

    def proj(self, c0, m, rcol):
        h = np.zeros(m * self.E)
        self._chk(lib().trlhip_proj(self.h, c0, m, rcol, _ptr(h)), "proj")
        return h.view(np.complex128) if self.complex else h

    def update(self, c0, m, h, rcol):
        hv = np.ascontiguousarray(h, dtype=np.complex128 if self.complex else np.float64)
        hv = hv if m > 0 else np.zeros(2)
        rr2 = ctypes.c_double()
        self._chk(lib().trlhip_update(self.h, c0, m, _ptr(hv), rcol, ctypes.byref(rr2)), "update")
        return rr2.value

    def cgs_pass(self, c0, m, rcol):
        hh, rr2 = ctypes.c_double(), ctypes.c_double()
        h = np.zeros(m * self.E)
        self._chk(lib().trlhip_cgs_pass(self.h, c0, m, rcol, ctypes.byref(hh), ctypes.byref(rr2),
                                        _ptr(h)), "cgs_pass")
        return hh.value, rr2.value, (h.view(np.complex128) if self.complex else h)

    def spmv(self, A, xcol, ycol):
        a = ctypes.c_double()
        self._chk(lib().trlhip_spmv(self.h, A.handle, xcol, ycol, ctypes.byref(a)), "spmv")
        return a.value

    def scale_copy(self, dst, src, scale):
        """dst = scale * src; src may be the scratch column (index ncols)."""
        self._chk(lib().trlhip_col_scale_copy(self.h, dst, src, scale), "scale_copy")

    def step_cgs(self, A, first, qa, qb, beta_prev, beta, jnd, rcol):
        """One fused Lanczos step (trlhip_step_cgs); returns (alpha, ||r||^2,
        ||h||^2, ||r - V h||^2) with r left in rcol and r - V h in the
        scratch column."""
        out = np.zeros(4)
        b = np.ascontiguousarray(beta if beta is not None else np.zeros(1), dtype=np.float64)
        self._chk(lib().trlhip_step_cgs(self.h, A.handle if A is not None else None, first, qa,
                                        qb, beta_prev, _ptr(b), jnd, rcol, _ptr(out)), "step_cgs")
        return out

    def ritz(self, lck, jl, k, Y):
        Y = np.asfortranarray(Y, dtype=np.float64)
        self._chk(lib().trlhip_ritz(self.h, lck, jl, k, _ptr(Y), Y.shape[0]), "ritz")

    def refine(self, A, k, avoid):
        """FP64 Rayleigh quotients and residual norms of columns 0..k-1
        (trlhip_refine); returns (rho, res)."""
        rho = np.zeros(k)
        res = np.zeros(k)
        self._chk(lib().trlhip_refine(self.h, A.handle, k, avoid, _ptr(rho), _ptr(res)), "refine")
        return rho, res

    def cols_get(self, c0, k):
        """Columns c0 .. c0+k-1 in one transfer (trlhip_cols_get), as an
        (n, k) array."""
        dt = np.complex128 if self.complex else np.float64
        x = np.zeros((k, self.n), dtype=dt)
        self._chk(lib().trlhip_cols_get(self.h, c0, k, _ptr(x), self.n), "cols_get")
        return x.T

    def residuals(self, A, rho):
        """Columns k+i = A v_i - rho[i] v_i for i < k = len(rho) and their
        squared norms (trlhip_residuals)."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        rr2 = np.zeros(len(rho))
        self._chk(lib().trlhip_residuals(self.h, A.handle, len(rho), _ptr(rho), _ptr(rr2)), "residuals")
        return rr2

    def gram(self, A, m, tmp):
        """(Q'Q, Q'AQ) over columns [0, m), A Q_j formed in column tmp
        (trlhip_gram)."""
        G = np.zeros(m * m)
        H = np.zeros(m * m)
        self._chk(lib().trlhip_gram(self.h, A.handle, m, tmp, _ptr(G), _ptr(H)), "gram")
        return G.reshape(m, m, order="F"), H.reshape(m, m, order="F")

    def close(self):
        if self.h:
            lib().trlhip_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
