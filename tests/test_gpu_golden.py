"""GPU solves against the compiled reference's own recorded outputs
(tests/golden/ref_cases.json, made by tests/golden/make_golden.py from
oracle/_ref): the device trlan()/ztrlan() on the same operator and start
as the reference, with no oracle in between.

Bars:
- integer trajectory (matvec, nloop, north, locked, nrand) equal to the
  reference's, eigenvalues within 1e-10 ||A|| of its eigenvalues, for the
  real cases whose decisions have margin (schemes 0-2, 4-6, 7/8 on
  diagonal spectra, guess = 0 starts);
- eigenvalues within 1e-10 ||A|| for the adaptive schemes 7/8 away from
  diagonal spectra (their arg-min follows the last bits of the Ritz
  values, so their trajectories do too), and operator applications within
  10 % beside that for scheme 3 (it compares operator with
  orthogonalization wall time) and the complex path (locking compares
  residuals with eps |lambda|);
- every returned pair has ||A v - lambda v|| <= tol ||A|| and the Ritz
  vectors are orthonormal to 1e-12.
Checkpoint cases: a solve cut at the reference's maxmv writes a checkpoint
with the reference's basis size, and the guess = 2 solve resuming from it
makes the reference's resumed trajectory."""
import json
import os
import sys

import numpy as np
import pytest

from nutrlan_amd import workloads as W

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden as MG  # noqa: E402

GOLD = json.load(open(os.path.join(HERE, "golden", "ref_cases.json")))
# adaptive schemes 7/8 off diagonal spectra: trajectories (and operator
# counts: 1094 here against the reference's 1442 on the degenerate 10^3
# Laplacian) follow roundoff, so only eigenvalues and residuals are held
ADAPTIVE = {"lap10_lg8_r7", "and2d48_sm8_r7", "and2d48_sm8_r8", "rand3000_lg10_r7"}
# scheme 3 reads wall-clock ratios; counts within 10 %
CLOCKED = {"lap10_lg8_r3", "diag2000_sm10_r3"}
CASES = GOLD["cases"]
KEYS = ["matvec", "nloop", "north", "locked", "nrand"]


def _check_pairs(g, c, rp, ci, v, nec):
    f = W.csr_matvec(rp, ci, v)
    X = g["evec"][:, :nec]
    assert np.abs(X.conj().T @ X - np.eye(nec)).max() <= 1e-12
    for i in range(nec):
        r = f(X[:, i]) - g["eval"][i] * X[:, i]
        assert np.linalg.norm(r) <= 1.01 * c["tol"] * g["anrm"] + 1e-12 * g["anrm"]


@pytest.mark.parametrize("c", CASES, ids=[c["name"] for c in CASES])
def test_device_solve_vs_reference(gpu_lib, c):
    T = gpu_lib
    rp, ci, v = MG.make_operator(c["generator"], c["args"])
    n = len(rp) - 1
    x0 = MG.make_start(c["start"], n)
    guess = MG.start_guess(c["start"])
    print(f"golden {c['name']} n={n} guess={guess}", flush=True)
    A = T.Csr(rp, ci, v, complex_=c["complex"])
    g = T.solve_csr(A, c["lohi"], c["ned"], c["maxlan"], c["tol"], c["restart"], c["maxmv"],
                    c["mev"], x0.astype(np.complex128) if c["complex"] else x0, guess=guess)
    A.close()
    assert g["stat"] == 0 and g["nec"] == c["nec"]
    nec = c["nec"]
    got = g["eval"][:nec]
    want = np.array(c["eval"])
    assert np.abs(got - want).max() <= 1e-10 * c["anrm"]
    print(f"matvec {g['matvec']} (reference {c['matvec']}), nloop {g['nloop']} "
          f"({c['nloop']})", flush=True)
    if c["name"] in ADAPTIVE:
        assert g["matvec"] <= c["maxmv"]
    elif c["complex"] or c["name"] in CLOCKED:
        assert abs(g["matvec"] - c["matvec"]) <= 0.1 * c["matvec"]
    else:
        assert [g[k] for k in KEYS] == [c[k] for k in KEYS]
    _check_pairs(g, c, rp, ci, v, nec)


CK = GOLD["checkpoint_cases"]


@pytest.mark.parametrize("c", CK, ids=[c["name"] for c in CK])
def test_checkpoint_resume_vs_reference(gpu_lib, c, tmp_path):
    T = gpu_lib
    rp, ci, v = MG.make_operator(c["generator"], c["args"])
    n = len(rp) - 1
    x0 = MG.make_start(c["start"], n)
    cx = c["complex"]
    args = (c["lohi"], c["ned"], c["maxlan"], c["tol"], c["restart"])
    print(f"checkpoint {c['name']} n={n}", flush=True)
    A = T.Csr(rp, ci, v, complex_=cx)
    prefix = str(tmp_path / "cp")
    part = T.solve_csr(A, *args, c["maxmv_cut"], c["mev"], x0, checkpoint=(c["cpflag"], prefix))
    keys = ["stat", "nec", "matvec", "nloop", "north", "locked"]
    if cx:  # complex locking tips at roundoff (see above): counts to 10 %
        assert part["stat"] == 0 and abs(part["matvec"] - c["part"]["matvec"]) <= 0.1 * c["part"]["matvec"]
    else:
        assert [part[k] for k in keys] == [c["part"][k] for k in keys]
    raw = open(prefix + "0", "rb").read()
    nrow, jnd = np.frombuffer(raw[:8], dtype=np.int32)
    assert nrow == n and (cx or jnd == c["file_jnd"])
    es = 16 if cx else 8  # bytes per element: ztrl_write_checkpoint writes (re, im) pairs
    assert len(raw) == 8 + 16 * jnd + es * n * (jnd + 1)
    res = T.solve_csr(A, *args, c["maxmv_resume"], c["mev"], np.zeros(n, dtype=x0.dtype),
                      resume=prefix)
    A.close()
    want = c["resume"]
    assert res["stat"] == 0 and res["nec"] == want["nec"]
    # a residual that is not read back is replaced by a random vector in
    # both (the reference's drand48 and this library's own generator)
    assert res["nrand"] == want["nrand"]
    if want["nrand"] == 0 and not cx:
        assert [res[k] for k in keys] == [want[k] for k in keys]
    else:
        assert abs(res["matvec"] - want["matvec"]) <= max(2, 0.1 * want["matvec"])
    nec = want["nec"]
    assert np.abs(np.sort(res["eval"][:nec]) - np.sort(want["eval"])).max() <= 1e-10 * want["anrm"]
    _check_pairs(res, c, rp, ci, v, nec)
