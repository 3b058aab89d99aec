"""The basis debug checks (trl_check_orth / trl_check_recurrence and their
complex forms, SRC/trlaux.c:331-778, SRC/ztrlaux.c:308-830) on the device
basis, against the compiled reference's own logs of the same solves
(tests/golden/debug_cases.json, made by tests/golden/make_debug_golden.py):

* real cases: the same solve trajectory, so the same checks at the same
  points: every TRL_CHECK_ORTH block's (matvec, vectors) pair, the number
  of per-vector blocks, the entries of every recurrence check (including
  the points where the reference's workspace is too small for it);
* the levels our checks report are roundoff-sized (the reference's are
  too), and the recurrence differences are;
* turning the checks on changes nothing in the solve (bit for bit);
* complex: the checks' structure follows our own trajectory (two checks
  per restart cycle), with roundoff-sized levels.
"""
import ctypes
import json
import os
import sys

import numpy as np
import pytest

from nutrlan_amd import trlan as T
from nutrlan_amd import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_debug_golden import build, parse_log  # noqa: E402

CASES = json.load(open(os.path.join(HERE, "golden", "debug_cases.json")))["cases"]
P = ctypes.c_void_p


def _p(a):
    return a.ctypes.data_as(P)


def _solve(case, tmp_path, verbose):
    """Our trlan()/ztrlan() on the case with the reference driver's
    workspace sizes; returns (info, eval, evec, log text)."""
    L = T.lib()
    L.trl_set_debug.argtypes = [ctypes.POINTER(T.TrlInfo), ctypes.c_int, ctypes.c_char_p]
    L.trl_set_debug.restype = None
    rp, ci, v = build(case["gen"], case["args"])
    n = len(rp) - 1
    cplx = case["complex"]
    x0 = W.start_vector(n, seed=case["seed"], complex_=cplx)
    mev, maxlan = case["mev"], case["maxlan"]
    info = T.init_info(n, maxlan, case["lohi"], case["ned"], case["tol"], case["restart"],
                       case["maxmv"])
    L.trl_set_iguess(ctypes.byref(info), 0, 1, 0, None)
    prefix = str(tmp_path / f"{case['name']}_v{verbose}_")
    if verbose:
        L.trl_set_debug(ctypes.byref(info), verbose, prefix.encode())
    A = T.Csr(rp, ci, v, complex_=cplx)
    ev = np.zeros(mev)
    if cplx:
        evec = np.zeros((mev, n), np.complex128)
        evec[0] = x0
        nmis = max(10 * n, maxlan * (maxlan + 10))
        misc = np.zeros(nmis, np.complex128)
        dw = np.zeros(maxlan * (maxlan + 10))
        info.mvparam = A.handle
        L.ztrlan(P(T.fn_addr("trlhip_csr_zmatvec")), ctypes.byref(info), n, mev, _p(ev),
                 _p(evec), n, _p(misc), nmis, _p(dw), len(dw))
    else:
        evec = np.zeros((mev, n))
        evec[0] = x0
        ldb = (n + 3) // 4 * 4 + 4
        lwrk = max(maxlan - mev + 1, 1) * ldb + maxlan * (maxlan + 10)
        wrk = np.zeros(lwrk)
        info.mvparam = A.handle
        L.trlan(P(T.fn_addr("trlhip_csr_matvec")), None, ctypes.byref(info), n, mev, _p(ev),
                _p(evec), n, lwrk, _p(wrk))
    A.close()
    text = open(prefix + "0").read() if verbose and os.path.exists(prefix + "0") else ""
    return info, ev, evec, text


COUNTERS = ("stat", "nec", "matvec", "nloop", "north", "locked")


@pytest.mark.parametrize("case", [c for c in CASES if not c["complex"]], ids=lambda c: c["name"])
def test_debug_checks_match_reference_log(gpu_lib, case, tmp_path):
    info, ev, evec, text = _solve(case, tmp_path, case["verbose"])
    got = parse_log(text)
    ref = case["log"]
    print(f"{case['name']}: matvec {info.matvec} (reference {case['counters']['matvec']}), "
          f"{len(got['orth'])} orthogonality checks (reference {len(ref['orth'])}), "
          f"recurrence entries {got['rec'][:6]}, largest level {got['max_fro']:.2e} "
          f"(reference {ref['max_fro']:.2e})", flush=True)
    for k in COUNTERS:
        assert getattr(info, k) == case["counters"][k], k
    assert got["orth"] == ref["orth"]
    assert got["vlevel"] == ref["vlevel"]
    assert got["rec"] == ref["rec"]
    assert got["orth_too_small"] == ref["orth_too_small"]
    assert got["rec_too_small"] == ref["rec_too_small"]
    assert got["max_fro"] <= 1e-12 and got["max_level"] <= 1e-12
    anrm = info.anrm
    assert got["max_dalpha"] <= 1e-10 * anrm and got["max_dbeta"] <= 1e-10 * anrm
    assert got["max_rec_err"] <= 1e-10 * anrm
    # the checks change nothing in the solve
    info0, ev0, evec0, _ = _solve(case, tmp_path, 0)
    for k in COUNTERS:
        assert getattr(info0, k) == getattr(info, k), k
    assert np.array_equal(ev0, ev) and np.array_equal(evec0, evec)


@pytest.mark.parametrize("case", [c for c in CASES if c["complex"]], ids=lambda c: c["name"])
def test_complex_debug_checks(gpu_lib, case, tmp_path):
    info, ev, evec, text = _solve(case, tmp_path, case["verbose"])
    got = parse_log(text)
    print(f"{case['name']}: matvec {info.matvec} nloop {info.nloop} (reference "
          f"{case['counters']['matvec']} / {case['counters']['nloop']}), "
          f"{len(got['orth'])} checks, largest level {got['max_fro']:.2e}", flush=True)
    assert info.stat == 0 and info.nec >= case["ned"]
    # two checks per cycle: after its inner loop and after its restart
    assert len(got["orth"]) == 2 * info.nloop == len(got["rec"])
    assert got["orth"][-1][0] == info.matvec
    assert all(m <= info.matvec for m, _ in got["orth"])
    assert got["vlevel"] == sum(j for _, j in got["orth"])
    assert got["max_fro"] <= 1e-12
    anrm = info.anrm
    assert got["max_dalpha"] <= 1e-10 * anrm and got["max_rec_err"] <= 1e-10 * anrm
    if all(info.__getattribute__(k) == case["counters"][k] for k in COUNTERS):
        assert got["orth"] == case["log"]["orth"] and got["rec"] == case["log"]["rec"]
