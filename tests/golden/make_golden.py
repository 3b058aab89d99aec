"""Writes tests/golden/ref_cases.json: outputs of the compiled reference
trlan()/ztrlan() (oracle/_ref, built by `make -C oracle` from the reference
sources) on small seeded problems, so the oracle and the GPU path can be
checked against the reference where the reference is not available.

Each case stores its inputs (generator name + parameters, start-vector
seed or info->guess) and the reference's outputs: eigenvalues, residual estimates, the
integer trajectory (matvec, nloop, north, nrand, locked) and ||A||, plus a
checksum of the Ritz vectors (sum of |v_i| per vector, sign-free).
Checkpoint cases also record a solve cut short at maxmv that writes the
reference's checkpoint file and the guess = 2 solve that resumes from it.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from nutrlan_amd import workloads as W  # noqa: E402
from oracle import ref  # noqa: E402

CASES = [
    # name, generator, args, lohi, ned, maxlan, tol, restart, maxmv, mev, start
    ("diag1000_lg5_r1", "diag", [1000], 1, 5, 40, 1.4e-8, 1, 5000, 5, ["uniform", 1]),
    ("diag2000_sm10_r1", "diag", [2000], -1, 10, 100, 1.4901e-8, 1, 2000, 10, ["ones"]),
    ("diag2000_sm10_r2", "diag", [2000], -1, 10, 100, 1.4901e-8, 2, 2000, 10, ["ones"]),
    ("diag2000_sm10_r4", "diag", [2000], -1, 10, 100, 1.4901e-8, 4, 2000, 10, ["ones"]),
    ("diag2000_sm10_r5", "diag", [2000], -1, 10, 100, 1.4901e-8, 5, 2000, 10, ["ones"]),
    ("diag2000_sm10_r6", "diag", [2000], -1, 10, 100, 1.4901e-8, 6, 2000, 10, ["ones"]),
    ("diag2000_sm10_r7", "diag", [2000], -1, 10, 100, 1.4901e-8, 7, 2000, 10, ["ones"]),
    ("diag2000_sm10_r8", "diag", [2000], -1, 10, 100, 1.4901e-8, 8, 2000, 10, ["ones"]),
    ("simple_c_diagsq1897", "diagsq", [1897], -1, 5, 200, 1.4901e-8, 1, 2000000, 100, ["ones"]),
    ("lap10_lg8_r1", "laplace3d", [10, 3], 1, 8, 40, 1e-8, 1, 20000, 8, ["uniform", 2]),
    ("lap10_sm8_r1", "laplace3d", [10, 3], -1, 8, 40, 1e-8, 1, 20000, 8, ["uniform", 2]),
    ("lap10_both8_r1", "laplace3d", [10, 3], 0, 8, 40, 1e-8, 1, 20000, 8, ["uniform", 2]),
    ("lap12_sm6_r1", "laplace3d", [12, 4], -1, 6, 32, 1e-8, 1, 20000, 6, ["uniform", 5]),
    ("rand3000_lg10_r1", "randsym", [3000, 4], 1, 10, 48, 1e-8, 1, 20000, 10, ["uniform", 5]),
    ("and2d48_sm8_r1", "anderson2d", [48, 6.0, 6], -1, 8, 32, 1e-8, 1, 20000, 8, ["uniform", 7]),
    # adaptive schemes away from diagonal spectra (their trajectories are
    # sensitive to roundoff in the Ritz values; tests hold eigenvalues)
    ("lap10_lg8_r7", "laplace3d", [10, 3], 1, 8, 40, 1e-8, 7, 20000, 8, ["uniform", 2]),
    ("and2d48_sm8_r7", "anderson2d", [48, 6.0, 6], -1, 8, 32, 1e-8, 7, 20000, 8, ["uniform", 7]),
    ("and2d48_sm8_r8", "anderson2d", [48, 6.0, 6], -1, 8, 32, 1e-8, 8, 20000, 8, ["uniform", 7]),
    ("rand3000_lg10_r7", "randsym", [3000, 4], 1, 10, 48, 1e-8, 7, 20000, 10, ["uniform", 5]),
    # scheme 3 (operator vs orthogonalization time; the operator here is
    # cheap, so its time test stays on one side)
    ("lap10_lg8_r3", "laplace3d", [10, 3], 1, 8, 40, 1e-8, 3, 20000, 8, ["uniform", 2]),
    ("diag2000_sm10_r3", "diag", [2000], -1, 10, 100, 1.4901e-8, 3, 2000, 10, ["ones"]),
    # guess = 0: the reference's own all-ones start (no random draws)
    ("diag2000_sm10_r1_g0", "diag", [2000], -1, 10, 100, 1.4901e-8, 1, 2000, 10, ["guess", 0]),
    ("lap10_lg8_r1_g0", "laplace3d", [10, 3], 1, 8, 40, 1e-8, 1, 20000, 8, ["guess", 0]),
]

ZCASES = [
    ("zdiag1000_sm20_r7", "zdiag", [1000], -1, 20, 60, 1.4901e-8, 7, 500000, 20, ["zones"]),
    ("zand8_sm6_r1", "anderson3d", [8, 3], -1, 6, 40, 1e-8, 1, 20000, 6, ["zuniform", 4]),
    ("zand10_lg6_r1", "anderson3d", [10, 5], 1, 6, 40, 1e-8, 1, 20000, 6, ["zuniform", 6]),
    # the complex workload's shape (24 lowest, maxlan 80) on a 12^3 lattice
    ("zand12_sm24_r1", "anderson3d", [12, 9], -1, 24, 80, 1e-8, 1, 40000, 24, ["zuniform", 10]),
    ("zand8_sm6_r1_g0", "anderson3d", [8, 3], -1, 6, 40, 1e-8, 1, 20000, 6, ["guess", 0]),
]

# checkpoint / resume: a solve cut short at maxmv writes the reference's
# checkpoint file (cpflag 1), then a guess = 2 solve resumes from it
CKPT_CASES = [
    # name, generator, args, lohi, ned, maxlan, tol, restart, cut fraction,
    # resume maxmv, mev, start
    ("lap14_ckpt_cut", "laplace3d", [14, 21], 1, 8, 40, 1e-9, 1, 0.6, 40000, 8, ["uniform", 22]),
    # ned = mev and converged: the file's basis fills evec, so its residual
    # is not read back (SRC/trlaux.c:982-1007) and the resumed solve
    # starts its next vector from a zero column
    ("lap12_ckpt_full_ned_eq_mev", "laplace3d", [12, 4], -1, 6, 32, 1e-8, 1, 1.0, 20000, 6,
     ["uniform", 5]),
    # complex (ztrl_write_checkpoint / ztrl_read_checkpoint)
    ("zand8_ckpt_cut", "anderson3d", [8, 3], -1, 6, 40, 1e-8, 1, 0.6, 20000, 6,
     ["zuniform", 4]),
]


def make_operator(gen, args):
    if gen == "diag":
        return W.diag_csr(args[0])
    if gen == "diagsq":
        rp, ci, v = W.diag_csr(args[0])
        return rp, ci, v * v
    if gen == "zdiag":
        rp, ci, v = W.diag_csr(args[0])
        return rp, ci, (v * v).astype(np.complex128)
    if gen == "laplace3d":
        return W.laplace3d_csr(args[0], seed=args[1])
    if gen == "randsym":
        return W.random_sym_csr(args[0], seed=args[1])
    if gen == "anderson2d":
        return W.anderson2d_csr(args[0], args[1], seed=args[2])[0]
    if gen == "anderson3d":
        return W.anderson3d_csr(args[0], seed=args[1])
    raise ValueError(gen)


def make_start(spec, n):
    if spec[0] == "ones":
        return np.ones(n)
    if spec[0] == "zones":
        return np.ones(n, dtype=np.complex128)
    if spec[0] == "uniform":
        return W.start_vector(n, seed=spec[1])
    if spec[0] == "zuniform":
        return W.start_vector(n, seed=spec[1], complex_=True)
    if spec[0] == "guess":  # info->guess = spec[1]: no start vector is read
        return np.zeros(n)
    raise ValueError(spec)


def start_guess(spec):
    """info->guess a case runs with."""
    return spec[1] if spec[0] == "guess" else 1


def run_case(case, complex_):
    name, gen, args, lohi, ned, maxlan, tol, restart, maxmv, mev, start = case
    rp, ci, v = make_operator(gen, args)
    n = len(rp) - 1
    x0 = make_start(start, n)
    if complex_ and start[0] == "guess":
        x0 = x0.astype(np.complex128)
    r = ref.solve_csr(rp, ci, v, lohi, ned, maxlan, tol, restart, maxmv, mev, x0,
                      complex_=complex_, guess=start_guess(start))
    nec = r["nec"]
    return dict(
        name=name, complex=complex_, generator=gen, args=args, lohi=lohi, ned=ned,
        maxlan=maxlan, tol=tol, restart=restart, maxmv=maxmv, mev=mev, start=start,
        stat=r["stat"], nec=nec, matvec=r["matvec"], nloop=r["nloop"], north=r["north"],
        nrand=r["nrand"], locked=r["locked"], anrm=r["anrm"],
        eval=r["eval"][:nec].tolist(), resid=r["resid"][:nec].tolist(),
        vec_abs_sum=np.abs(r["evec"][:, :nec]).sum(axis=0).tolist())


def run_ckpt_case(case):
    import tempfile

    name, gen, args, lohi, ned, maxlan, tol, restart, frac, maxmv2, mev, start = case
    rp, ci, v = make_operator(gen, args)
    n = len(rp) - 1
    x0 = make_start(start, n)
    cx = np.iscomplexobj(x0)
    full = ref.solve_csr(rp, ci, v, lohi, ned, maxlan, tol, restart, 40000, mev, x0, complex_=cx)
    cut = int(full["matvec"] * frac) if frac < 1.0 else 40000
    with tempfile.TemporaryDirectory() as d:
        prefix = os.path.join(d, "cp")
        part = ref.solve_csr(rp, ci, v, lohi, ned, maxlan, tol, restart, cut, mev, x0,
                             checkpoint=(1, prefix), complex_=cx)
        raw = open(prefix + "0", "rb").read()
        jnd = int(np.frombuffer(raw[4:8], dtype=np.int32)[0])
        res = ref.solve_csr(rp, ci, v, lohi, ned, maxlan, tol, restart, maxmv2, mev,
                            np.zeros(n, dtype=x0.dtype), resume=prefix, complex_=cx)
    keys = ["stat", "nec", "matvec", "nloop", "north", "nrand", "locked"]
    return dict(
        name=name, complex=cx, generator=gen, args=args, lohi=lohi, ned=ned, maxlan=maxlan, tol=tol,
        restart=restart, maxmv_cut=cut, maxmv_resume=maxmv2, mev=mev, start=start, cpflag=1,
        full={k: full[k] for k in keys}, part={k: part[k] for k in keys}, file_jnd=jnd,
        resume=dict({k: res[k] for k in keys}, anrm=res["anrm"],
                    eval=res["eval"][:res["nec"]].tolist(),
                    resid=res["resid"][:res["nec"]].tolist()))


def main():
    out = dict(
        source="compiled reference nu-TRLan (oracle/_ref) with scipy's OpenBLAS 0.3.28",
        info_layout=ref.info_layout(),
        cases=[run_case(c, False) for c in CASES] + [run_case(c, True) for c in ZCASES],
        checkpoint_cases=[run_ckpt_case(c) for c in CKPT_CASES])
    path = os.path.join(HERE, "ref_cases.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path, len(out["cases"]), "cases")


if __name__ == "__main__":
    main()
