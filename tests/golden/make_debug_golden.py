"""Writes tests/golden/debug_cases.json: what the compiled reference
(oracle/_ref, built by `make -C oracle` from the reference sources) logs
from its basis debug checks, trl_check_orth and trl_check_recurrence
(SRC/trlaux.c:331-778; ztrl_* in SRC/ztrlaux.c:308-830), at debug levels
7, 8 and 9 on small seeded problems.

For each case the fixture holds the inputs (generator + parameters, start
seed, debug level), the solve's counters, and from the log (parse_log):
the (matvec, vectors) pair of every TRL_CHECK_ORTH block in order, the
number of per-vector orthogonality blocks, the number of entries of every
recurrence check, and the largest orthogonality level and alpha/beta
difference the reference reported (roundoff-sized; the GPU's own must be
as small, not equal).

    python tests/golden/make_debug_golden.py
"""
import json
import os
import re
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from nutrlan_amd import workloads as W  # noqa: E402

# name, generator, args, lohi, ned, maxlan, tol, restart, maxmv, mev, start seed, verbose, complex
CASES = [
    ("lap8_sm6_v7", "laplace3d", [8, 2], -1, 6, 30, 1e-10, 1, 4000, 8, 3, 7, False),
    ("lap8_sm6_v8", "laplace3d", [8, 2], -1, 6, 30, 1e-10, 1, 4000, 8, 3, 8, False),
    ("lap8_lg4_v9", "laplace3d", [8, 2], 1, 4, 20, 1e-10, 1, 4000, 6, 5, 9, False),
    ("zand6_sm5_v8", "anderson3d", [6, 4], -1, 5, 30, 1e-10, 1, 4000, 8, 5, 8, True),
]

_NUM = r"[-+]?\d+\.\d+e[-+]\d+"


def build(gen, args):
    if gen == "laplace3d":
        return W.laplace3d_csr(args[0], seed=args[1])
    return W.anderson3d_csr(args[0], seed=args[1])


def parse_log(text):
    """The debug checks' blocks of a log: dict(orth=[[matvec, vectors]...],
    vlevel=<per-vector blocks>, rec=[entries...], max_fro, max_level,
    max_dalpha, max_dbeta, max_rec_err, and the counts of the checks'
    "workspace too small" / "not enough workspace" lines)."""
    lines = text.splitlines()
    out = dict(orth=[], vlevel=0, rec=[], max_fro=0.0, max_level=0.0, max_dalpha=0.0,
               max_dbeta=0.0, max_rec_err=0.0,
               orth_too_small=text.count("TRL_CHECK_ORTH: workspace too small."),
               rec_too_small=text.count("TRL_CHECK_RECURRENCE: not enough workspace."))
    i = 0

    def numbers_after(i):
        vals = []
        j = i + 1
        while j < len(lines) and not lines[j].startswith(("PE ", "TRL_", "Orthogonality",
                                                          "Frobenius", "Maximum")):
            vals += [float(x) for x in re.findall(_NUM, lines[j])]
            j += 1
        return vals

    while i < len(lines):
        ln = lines[i]
        m = re.match(r"Frobenius norm of orthogonality level\s+(\d+)\s+(\d+)\s+(" + _NUM + ")", ln)
        if m:
            out["orth"].append([int(m.group(1)), int(m.group(2))])
            out["max_fro"] = max(out["max_fro"], float(m.group(3)))
        m = re.match(r"Maximum abs\. value of orthogonality level is\s+(" + _NUM + ")", ln)
        if m:
            out["max_level"] = max(out["max_level"], abs(float(m.group(1))))
        if ln.startswith("Orthogonality level of v("):
            out["vlevel"] += 1
        if ln.startswith("PE ") and ln.endswith("Alpha computed by TRLAN .."):
            out["rec"].append(len(numbers_after(i)))
        if ln.startswith("PE ") and ln.endswith("Differences in alpha .."):
            v = numbers_after(i)
            out["max_dalpha"] = max([out["max_dalpha"]] + [abs(x) for x in v])
        if ln.startswith("PE ") and ln.endswith("Differences in beta .."):
            v = numbers_after(i)
            out["max_dbeta"] = max([out["max_dbeta"]] + [abs(x) for x in v])
        m = re.match(r"PE \d+ : Error in Lanczos recurrence \(overall\) =\s*(" + _NUM + ")", ln)
        if m:
            out["max_rec_err"] = max(out["max_rec_err"], abs(float(m.group(1))))
        i += 1
    return out


def main():
    from oracle import ref

    assert ref.available(), "build the reference first: make -C oracle"
    cases = []
    with tempfile.TemporaryDirectory() as d:
        for (name, gen, args, lohi, ned, maxlan, tol, restart, maxmv, mev, seed, verbose,
             cplx) in CASES:
            rp, ci, v = build(gen, args)
            n = len(rp) - 1
            x0 = W.start_vector(n, seed=seed, complex_=cplx)
            prefix = os.path.join(d, name + "_")
            r = ref.solve_csr(rp, ci, v, lohi, ned, maxlan, tol, restart, maxmv, mev, x0,
                              complex_=cplx, debug=(verbose, prefix))
            log = parse_log(open(prefix + "0").read())
            print(name, r["matvec"], r["nloop"], len(log["orth"]), log["rec"][:4],
                  log["orth_too_small"], log["rec_too_small"],
                  f"fro {log['max_fro']:.2e} dalpha {log['max_dalpha']:.2e}", flush=True)
            cases.append(dict(name=name, gen=gen, args=args, lohi=lohi, ned=ned, maxlan=maxlan,
                              tol=tol, restart=restart, maxmv=maxmv, mev=mev, seed=seed,
                              verbose=verbose, complex=cplx,
                              counters={k: r[k] for k in ("stat", "nec", "matvec", "nloop",
                                                          "north", "locked")},
                              log=log))
    with open(os.path.join(HERE, "debug_cases.json"), "w") as f:
        json.dump(dict(source="compiled reference trlan/ztrlan with trl_set_debug",
                       cases=cases), f, indent=1)


if __name__ == "__main__":
    main()
