"""The C-ABI boundary, checked without a GPU.

* libnutrlan_amd.so loads and exports every function include/*.h declares.
* trl_info has the reference's layout (size and field offsets, as the
  compiled reference reported them) in both the C header and the Python
  mirror.
* trl_init_info applies the reference's defaults/clamps (oracle restatement).
* Device entry points fail cleanly (no crash) when no GPU is visible.
"""
import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

from nutrlan_amd import trlan as T
from oracle import trlan_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "ref_cases.json")))


def declared_functions():
    names = set()
    for h in ("trlan.h", "ztrlan.h", "trlhip.h", "trlan_batch.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(",
                             txt, flags=re.M):
            if m.group(0).lstrip().startswith("typedef"):
                continue
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    T.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", T.LIBPATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    decl = declared_functions()
    assert len(decl) >= 30
    missing = [d for d in decl if d not in exported]
    assert not missing, missing


def test_info_layout_matches_reference(tmp_path):
    lay = GOLD["info_layout"]  # as the compiled reference reported it
    fields = ["stat", "lohi", "ned", "nec", "tol", "mpicomp", "maxlan", "klan", "maxmv",
              "restart", "locked", "guess", "matvec", "nloop", "north", "nrand", "anrm",
              "nloc", "ntot", "verbose", "trgt", "tres", "ref", "log_fp", "log_file", "cpfile",
              "oldcpf", "mvparam"]
    assert ctypes.sizeof(T.TrlInfo) == lay[0]
    for f, off in zip(fields, lay[1:]):
        assert getattr(T.TrlInfo, f).offset == off, f
    # the C header itself, compiled here
    src = str(tmp_path / "layout_probe.c")
    with open(src, "w") as fh:
        fh.write('#include <stddef.h>\n#include <stdio.h>\n#include "trlan.h"\n'
                 'int main(void){printf("%zu %zu %zu %zu\\n", sizeof(trl_info),'
                 ' offsetof(trl_info, guess), offsetof(trl_info, anrm),'
                 ' offsetof(trl_info, mvparam)); return 0;}\n')
    exe = src[:-2]
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
    got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True,
                                          check=True).stdout.split()]
    assert got == [lay[0], lay[12], lay[17], lay[-1]]


def test_fortran_compatible_build_maps_to_exported_entry_points(tmp_path):
    """Sources written for the reference's TRL_FORTRAN_COMPATIBLE build
    (pointer-argument callbacks, INC/trlan.h:9-36, INC/ztrlan.h:30-31)
    compile unchanged against include/, with -Werror, and every function
    they call resolves to a symbol the library exports."""
    src = str(tmp_path / "fortran_probe.c")
    with open(src, "w") as fh:
        fh.write(
            '#include "trlan.h"\n#include "ztrlan.h"\n'
            "static void op(int *n, int *nc, double *x, int *lx, double *y, int *ly)\n"
            "{ for (int j = 0; j < *nc; ++j) for (int i = 0; i < *n; ++i)"
            " y[j * *ly + i] = 2.0 * x[j * *lx + i]; }\n"
            "static void uo(int *n, double *x, double *y) { (void)n; (void)x;"
            " if (y) *y = 0.0; }\n"
            "static void zop(int *n, int *nc, trl_dcomplex *x, int *lx, trl_dcomplex *y,"
            " int *ly) { (void)nc; (void)lx; (void)ly; for (int i = 0; i < *n; ++i) y[i] = x[i]; }\n"
            "void probe(trl_info *info, double *ev, double *evec, double *wrk, int *chk,"
            " trl_dcomplex *zv, trl_dcomplex *zw)\n"
            "{\n"
            "  trlan(op, uo, info, 10, 4, ev, evec, 10, 4, wrk);\n"
            "  trl_check_ritz(op, info, 10, 2, evec, 10, ev, chk, wrk, ev, wrk, 40);\n"
            "  trl_rayleigh_quotients(op, info, 2, evec, 10, wrk, wrk);\n"
            "  trl_ritz_projection(op, info, 4, evec, 10, wrk, wrk, 40, wrk);\n"
            "  ztrlan(zop, info, 10, 4, ev, zv, 10, zw, 4, wrk, 40);\n"
            "  ztrl_check_ritz(zop, info, 10, 2, zv, 10, ev, chk, wrk, ev, zw, 40);\n"
            "}\n")
    obj = src[:-2] + ".o"
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-DTRL_FORTRAN_COMPATIBLE", "-I",
                    os.path.join(ROOT, "include"), "-c", src, "-o", obj], check=True)
    undef = {line.split()[-1] for line in subprocess.run(
        ["nm", "-u", obj], capture_output=True, text=True, check=True).stdout.splitlines()
        if line.strip()}
    want = {"trlan_fc", "trl_check_ritz_fc", "trl_rayleigh_quotients_fc",
            "trl_ritz_projection_fc", "ztrlan_fc", "ztrl_check_ritz_fc"}
    assert want <= undef, undef
    assert not ({"trlan", "ztrlan", "trl_check_ritz"} & undef)
    T.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", T.LIBPATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert want <= exported


@pytest.mark.parametrize("args", [
    (1000, 40, 1, 5, 1.4e-8, 1, 5000),
    (500, 5, -1, 8, 0.0, 0, 0),         # maxlan <= ned, default tol and maxmv
    (2000, 100, 0, 10, 2.0, 9, 100),    # tol > 1 -> min(0.1, 1/tol)
    (300, 20, -1, 4, 1e-320, -3, 7),    # tol <= DBL_MIN -> eps, restart <= 0 -> 0
])
def test_init_info_defaults(args):
    n, maxlan, lohi, ned, tol, restart, maxmv = args
    i = T.init_info(n, maxlan, lohi, ned, tol, restart, maxmv)
    o = O.init_info(n, maxlan, lohi, ned, tol, restart, maxmv)
    for f in ["tol", "restart", "maxmv", "maxlan", "lohi", "ned", "nloc", "ntot", "tmv",
              "trgt", "tres", "crat"]:
        assert getattr(i, f) == getattr(o, f), f
    assert i.guess == 0 and i.nec == 0 and i.mvparam is None


def test_trlan_validation_errors_without_gpu():
    """Argument errors are reported in info->stat before any device use
    (reference trl_clear_counter codes -1..-4)."""
    L = T.lib()
    n = 100
    ev = np.zeros(5)
    evec = np.zeros(5 * n)
    wrk = np.zeros(5)
    P = ctypes.c_void_p

    def run(info, nrow, mev, lde):
        L.trlan(P(T.fn_addr("trlhip_csr_matvec")), None, ctypes.byref(info), nrow, mev,
                ev.ctypes.data_as(P), evec.ctypes.data_as(P), lde, 5, wrk.ctypes.data_as(P))
        return info.stat

    assert run(T.init_info(n, 20, 1, 3, 1e-8, 1, 100), n + 1, 5, n + 1) == -1   # nrow != nloc
    assert run(T.init_info(n, 20, 1, 3, 1e-8, 1, 100), n, 5, n - 1) == -2       # lde < nrow
    assert run(T.init_info(8, 20, 1, 3, 1e-8, 1, 100), 8, 5, 8) == -4           # ntot < 10
    if T.device_count() == 0:
        # valid arguments but no device: a clean device-failure status
        assert run(T.init_info(n, 20, 1, 3, 1e-8, 1, 100), n, 5, n) == T.TRLHIP_STAT_DEVICE


def test_device_entry_points_fail_cleanly_without_gpu():
    if T.device_count() > 0:
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    rp = np.arange(11, dtype=np.int32)
    ci = np.arange(10, dtype=np.int32)
    v = np.ones(10)
    L = T.lib()
    rc = L.trlhip_csr_create(ctypes.byref(h), 10, 10, rp.ctypes.data_as(ctypes.c_void_p),
                             ci.ctypes.data_as(ctypes.c_void_p),
                             v.ctypes.data_as(ctypes.c_void_p), T.TRLHIP_R64)
    assert rc == -4  # TRLHIP_ENODEV


def test_shipped_start_net_weights():
    """The trained learned-start weights load without executing anything
    (numpy, allow_pickle=False) and have the net's parameter count."""
    import numpy as np

    from nutrlan_amd import learned as LN

    p = np.load(LN.WEIGHTS, allow_pickle=False)
    assert p.dtype == np.float32 and p.shape == (19105,)
    assert np.all(np.isfinite(p)) and np.abs(p).max() < 100


def test_projection_tile_setting():
    """trlhip_proj_tiles takes 4, 8 or 16 columns per block (0 = default 8)
    for each basis precision and rejects anything else (host-only)."""
    assert T.proj_tiles_get() == (8, 8)
    try:
        T.proj_tiles(4, 16)
        assert T.proj_tiles_get() == (4, 16)
        for bad in ((5, 8), (8, 32), (-4, 8)):
            with pytest.raises(RuntimeError):
                T.proj_tiles(*bad)
        assert T.proj_tiles_get() == (4, 16)
    finally:
        T.proj_tiles()
    assert T.proj_tiles_get() == (8, 8)
