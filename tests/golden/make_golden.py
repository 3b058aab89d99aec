"""Writes tests/golden/ref_cases.json: outputs of the compiled reference
trlan()/ztrlan() (oracle/_ref, built by `make -C oracle` from the reference
sources) on small seeded problems, so the oracle and the GPU path can be
checked against the reference where the reference is not available.

Each case stores its inputs (generator name + parameters, start-vector
seed) and the reference's outputs: eigenvalues, residual estimates, the
integer trajectory (matvec, nloop, north, nrand, locked) and ||A||, plus a
checksum of the Ritz vectors (sum of |v_i| per vector, sign-free).

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
]

ZCASES = [
    ("zdiag1000_sm20_r7", "zdiag", [1000], -1, 20, 60, 1.4901e-8, 7, 500000, 20, ["zones"]),
    ("zand8_sm6_r1", "anderson3d", [8, 3], -1, 6, 40, 1e-8, 1, 20000, 6, ["zuniform", 4]),
    ("zand10_lg6_r1", "anderson3d", [10, 5], 1, 6, 40, 1e-8, 1, 20000, 6, ["zuniform", 6]),
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
    raise ValueError(spec)


def run_case(case, complex_):
    name, gen, args, lohi, ned, maxlan, tol, restart, maxmv, mev, start = case
    rp, ci, v = make_operator(gen, args)
    n = len(rp) - 1
    x0 = make_start(start, n)
    r = ref.solve_csr(rp, ci, v, lohi, ned, maxlan, tol, restart, maxmv, mev, x0,
                      complex_=complex_)
    nec = r["nec"]
    return dict(
        name=name, complex=complex_, generator=gen, args=args, lohi=lohi, ned=ned,
        maxlan=maxlan, tol=tol, restart=restart, maxmv=maxmv, mev=mev, start=start,
        stat=r["stat"], nec=nec, matvec=r["matvec"], nloop=r["nloop"], north=r["north"],
        nrand=r["nrand"], locked=r["locked"], anrm=r["anrm"],
        eval=r["eval"][:nec].tolist(), resid=r["resid"][:nec].tolist(),
        vec_abs_sum=np.abs(r["evec"][:, :nec]).sum(axis=0).tolist())


def main():
    out = dict(
        source="compiled reference nu-TRLan (oracle/_ref) with scipy's OpenBLAS 0.3.28",
        info_layout=ref.info_layout(),
        cases=[run_case(c, False) for c in CASES] + [run_case(c, True) for c in ZCASES])
    path = os.path.join(HERE, "ref_cases.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path, len(out["cases"]), "cases")


if __name__ == "__main__":
    main()
