"""Writes tests/golden/kit_cases.json: inputs and outputs of single routines
of the compiled reference (oracle/_ref, built by `make -C oracle` from the
reference sources), called directly on seeded random states:

  trl_shuffle_eig   restart schemes 0-9 (9 = out of range, which the
                    reference maps to 7), lohi -2..1, tiny bases (nd <= 5),
                    first restarts (nloop 0), and scheme 3 on both sides of
                    its operator-time branch (SRC/restart.c:1052-1054)
  trl_convergence_test, trl_sort_eig, trl_set_locking
  trl_smooth_rr, trl_smooth_zz (their end-of-vector quirks)
  dsort2, dsort2a, dsort2s, dsort2su_, dsort2sd (ties included)

The host kit of the library (trlkit_*) and the numpy oracle are both
checked against these records (tests/test_kit.py), so neither is pinned
only by the other.

    python tests/golden/make_kit_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from nutrlan_amd.trlan import TrlInfo  # noqa: E402  (ctypes layout only)
from oracle import ref  # noqa: E402

# trl_info fields the restart, convergence and locking routines read or write
INFO_FIELDS = ["lohi", "ned", "nec", "tol", "maxlan", "klan", "maxmv", "restart", "locked",
               "matvec", "nloop", "clk_op", "clk_orth", "clk_res", "tick_o", "tick_h", "tick_r",
               "anrm", "ntot", "crat", "trgt", "tres", "tmv", "avgm", "cfac", "ptres",
               "predicted_crate", "mgap", "mgamma", "gamma0", "old_target", "target_id",
               "old_locked", "k1", "k2", "k", "rfact", "ref"]


def info_from(d):
    i = TrlInfo()
    for f, v in d.items():
        setattr(i, f, v)
    return i


def info_dict(i, before):
    """The fields of i that differ from the input dict `before` (fields not
    in `before` start at trl_info's zero)."""
    out = {}
    for f in INFO_FIELDS:
        v = getattr(i, f)
        v = float(v) if isinstance(v, float) else int(v)
        if v != before.get(f, 0):
            out[f] = v
    return out


def shuffle_state(rng, nd, lohi, restart, ned, clk):
    """A restart's state: nd = jml unlocked Ritz pairs after `locked`
    locked ones, and the workspace array the reference keeps them in."""
    locked = int(rng.integers(0, 3))
    jnd = nd + locked
    lam = np.sort(rng.standard_normal(jnd) * 3)
    res = 10.0 ** rng.uniform(-14, -1, jnd)
    res[:locked] = 0.0
    # a few converged pairs at either end now and then
    if rng.uniform() < 0.5:
        res[locked: locked + int(rng.integers(0, 3))] = 1e-15
    if rng.uniform() < 0.5:
        res[jnd - int(rng.integers(0, 3)):] = 1e-15
    d = dict(lohi=lohi, ned=ned, restart=restart, tol=1e-8, anrm=10.0, maxlan=jnd + 4, klan=jnd,
             maxmv=5000, matvec=int(rng.integers(nd, 2000)), nloop=int(rng.integers(0, 5)),
             nec=int(rng.integers(0, ned)), locked=locked,
             crat=float(rng.uniform(0.5, 0.99)), cfac=float(rng.uniform(0.01, 0.9)),
             ptres=float(rng.uniform(1e-6, 1e-2)), avgm=float(rng.uniform(5, 20)),
             k=int(rng.integers(0, max(1, nd // 2))), rfact=1.5,
             trgt=float(lam[locked + int(rng.integers(0, nd))]), ntot=100000)
    if clk is not None:
        # operator time above (True) or below (False) ten times the
        # orthogonalization + restart time
        big, small = 5_000_000, 40_000
        d.update(clk_op=big if clk else small, tick_o=0.25,
                 clk_orth=small if clk else big, tick_h=0.0, clk_res=small, tick_r=0.0,
                 matvec=d["maxlan"] + 50 + int(rng.integers(0, 500)))
    # the reference's workspace at a restart (SRC/trlcore.c:868, 1200-1201,
    # 1256, 1283): betrot's last element, lambda[0:jnd), res[0:jnd), then
    # the convergence test's |res| and lambda (lambda is already sorted)
    w = shuffle_workspace(float(rng.uniform(0, 2)), lam, res)
    return d, w, locked, jnd


def shuffle_workspace(pre, lam, res):
    """[pre, lambda, res, |res|, lambda]: the reference's workspace at a
    restart for sorted lambda (tests rebuild it from pre, lam and res)."""
    return np.concatenate([[pre], lam, res, np.abs(res), lam])


def shuffle_cases():
    cases = []
    for restart in range(10):
        for lohi in ([-2, -1, 0, 1] if restart in (1, 7) else [-1, 0, 1]):
            for seed in range(8):
                rng = np.random.default_rng(10_000 * restart + 100 * (lohi + 2) + seed)
                nd = 8 + 2 * seed
                clk = (seed % 2 == 0) if restart == 3 else None
                d, w, locked, jnd = shuffle_state(rng, nd, lohi, restart, 6, clk)
                if seed == 7:
                    d["nloop"] = 0  # first restart: schemes 3-8 fall back to scheme 1
                cases.append(_shuffle_one(d, nd, w, locked, jnd, kept=3))
    # tiny bases take the nd <= 5 branch
    for nd in range(1, 6):
        rng = np.random.default_rng(900 + nd)
        d, w, locked, jnd = shuffle_state(rng, nd, -1, 1, 2, None)
        cases.append(_shuffle_one(d, nd, w, locked, jnd, kept=1))
    return cases


def _shuffle_one(d, nd, w, locked, jnd, kept):
    info = info_from(d)
    w2 = w.copy()
    lo, ro = 1 + locked, 1 + jnd + locked
    k = ref.shuffle_eig(nd, nd, w2, lo, ro, info, kept, locked)
    return dict(nd=nd, mnd=nd, kept_in=kept, locked=locked, info=d, pre=w[0],
                lam=w[1:1 + jnd].tolist(), res=w[1 + jnd:1 + 2 * jnd].tolist(), kept=k,
                lam_out=w2[lo:lo + nd].tolist(), res_out=w2[ro:ro + nd].tolist(),
                info_out=info_dict(info, d))


def convergence_cases():
    cases = []
    for lohi in [-4, -3, -2, -1, 0, 1]:
        for seed in range(10):
            rng = np.random.default_rng(500 + 10 * (lohi + 4) + seed)
            nd = 5 + 2 * seed
            lam = np.sort(rng.standard_normal(nd) * 3)
            res = 10.0 ** rng.uniform(-14, -1, nd)
            res[: seed % 4] = 1e-13
            d = dict(lohi=lohi, ned=4, restart=1, tol=1e-8, anrm=float(rng.uniform(0, 8)),
                     maxlan=nd + 4, klan=nd, maxmv=5000, matvec=int(rng.integers(nd, 2000)),
                     nloop=int(rng.integers(0, 5)), nec=int(rng.integers(0, 4)),
                     crat=float(rng.uniform(0.5, 0.99)), cfac=float(rng.uniform(0.01, 0.9)),
                     ptres=float(rng.uniform(1e-6, 1e-2)), tmv=3 if seed % 2 else -1,
                     tres=1e-3, ref=0.3, trgt=0.0, ntot=100000)
            info = info_from(d)
            l2, r2 = lam.copy(), res.copy()
            ref.convergence_test(nd, l2, r2, info)
            cases.append(dict(nd=nd, info=d, lam=lam.tolist(), res=res.tolist(),
                              lam_out=l2.tolist(), res_out=r2.tolist(),
                              info_out=info_dict(info, d)))
    return cases


def locking_cases():
    cases = []
    rng = np.random.default_rng(5)
    for trial in range(30):
        nlam, jnd = 6 + trial % 5, 8
        lam = np.sort(rng.standard_normal(nlam))
        res = 10.0 ** rng.uniform(-18, -8, nlam)
        yy = rng.standard_normal(jnd * nlam)
        l2, r2, y2 = lam.copy(), res.copy(), yy.copy()
        k = ref.set_locking(jnd, nlam, l2, r2, y2, 3.0)
        cases.append(dict(jnd=jnd, nlam=nlam, anrm=3.0, lam=lam.tolist(), res=res.tolist(),
                          yy=yy.tolist(), locked=k, lam_out=l2.tolist(), res_out=r2.tolist(),
                          yy_out=y2.tolist()))
    return cases


def sort_eig_cases():
    cases = []
    rng = np.random.default_rng(6)
    for lohi in [-4, -3, -2, -1, 0, 1]:
        for nec in [0, 1, 5, 12]:
            lam = np.round(rng.standard_normal(12) * 4) / 4  # ties
            res = rng.uniform(0, 1, 12)
            l2, r2 = lam.copy(), res.copy()
            ref.sort_eig(12, lohi, nec, 0.2, l2, r2)
            cases.append(dict(nd=12, lohi=lohi, nec=nec, ref=0.2, lam=lam.tolist(),
                              res=res.tolist(), lam_out=l2.tolist(), res_out=r2.tolist()))
    return cases


def smooth_cases():
    cases = []
    rng = np.random.default_rng(7)
    for n in [3, 4, 5, 7, 16, 100]:
        x = rng.standard_normal(n)
        y = x.copy()
        ref.smooth_rr(y)
        z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        w = z.copy()
        ref.smooth_zz(w)
        cases.append(dict(n=n, x=x.tolist(), x_out=y.tolist(), z=z.view(np.float64).tolist(),
                          z_out=w.view(np.float64).tolist()))
    return cases


def sort2_cases():
    cases = []
    rng = np.random.default_rng(8)
    for key in range(5):
        for n in [0, 1, 2, 9, 50]:
            a1 = np.round(rng.standard_normal(n) * 3) / 2  # ties
            a2 = np.arange(n, dtype=np.float64)
            b1, b2 = a1.copy(), a2.copy()
            ref.sort2(key, 0.7, b1, b2)
            cases.append(dict(key=key, s=0.7, a1=a1.tolist(), a2=a2.tolist(), a1_out=b1.tolist(),
                              a2_out=b2.tolist()))
    return cases


def main():
    out = dict(source="compiled reference nu-TRLan (oracle/_ref), routines called directly",
               info_fields=INFO_FIELDS, shuffle=shuffle_cases(),
               convergence=convergence_cases(), locking=locking_cases(),
               sort_eig=sort_eig_cases(), smooth=smooth_cases(), sort2=sort2_cases())
    path = os.path.join(HERE, "kit_cases.json")
    with open(path, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", path, {k: len(v) for k, v in out.items() if isinstance(v, list)})


if __name__ == "__main__":
    main()
