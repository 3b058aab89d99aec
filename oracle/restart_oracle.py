"""Restart policy of the reference (SRC/restart.c), restated.  Test oracle only.

trl_shuffle_eig (restart.c:1410-1651) decides which Ritz pairs survive a
thick restart: [0..kl] from the left and [kr..nd) from the right, then
compacts them to the front.  The scheme-specific choice of (kl, kr) follows
the reference's schemes 0-8.  Integer arithmetic follows C: `nint(x)` is
(int)(x+0.5) truncating toward zero, `/` on ints is truncating division.
Scheme 3 compares operator time with orthogonalization + restart time
(restart.c:1052-1054); given the same trl_info clock fields it decides as the
reference does (pinned on both sides of that branch by the compiled
reference's trl_shuffle_eig, tests/golden/kit_cases.json), but in a solve
those fields are wall-clock readings, so a scheme-3 trajectory is not
reproducible run to run, in the reference or here.
"""
import math

import numpy as np

from .sorts import dsort2

DBL_EPSILON = 2.220446049250313e-16
DBL_MAX = 1.7976931348623157e308




# C math library semantics for the scalar formulas of the schemes: the
# platform libm (Python's math module calls it, as the reference's C does),
# with IEEE results (inf, nan) where Python would raise.  numpy's own
# vectorized exp/log/atan may differ from libm in the last bit, which can
# flip a scheme's arg-min over many restarts.
def _log(x):
    x = float(x)
    if x > 0.0:
        return np.float64(math.log(x))
    return np.float64(-np.inf if x == 0.0 else np.nan)


def _sqrt(x):
    x = float(x)
    return np.float64(math.sqrt(x)) if x >= 0.0 else np.float64(np.nan)


def _exp(x):
    try:
        return np.float64(math.exp(float(x)))
    except OverflowError:
        return np.float64(np.inf)


def _atan(x):
    return np.float64(math.atan(float(x)))


def _pow(x, y):
    try:
        return np.float64(math.pow(float(x), float(y)))
    except (OverflowError, ValueError):
        with np.errstate(all="ignore"):
            return np.power(np.float64(x), np.float64(y))

class Strip:
    """A window of a flat float64 array with C pointer indexing: index i is
    element off + i of the array, for i < 0 and past the window's end too.
    The reference's restart schemes index their lambda/res arguments one
    element beyond either end (restart.c:633, 641, 681; 781, 788, 980),
    reading the neighbouring workspace; windows of one array reproduce
    that."""

    def __init__(self, base, off):
        self.base = base
        self.off = off

    def _at(self, i):
        j = self.off + i
        if j < 0 or j >= len(self.base):
            raise IndexError(f"window index {i} outside the laid-out workspace")
        return j

    def __getitem__(self, i):
        return self.base[self._at(i)]

    def __setitem__(self, i, v):
        self.base[self._at(i)] = v


def _nint(x):
    return int(x + 0.5)


def _idiv(a, b):
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def _search_range(nd, lam, res, info, ncl, ncr, tind):
    """trl_restart_search_range_ (restart.c:26-91); returns (lohi, klm, krm)."""
    klm = max(ncl, 0)
    krm = min(ncr, nd - 1)
    bnd = info.tol * info.anrm
    lohi = info.lohi
    if info.lohi > 0:
        krm = min(max(info.maxlan - info.ned, _idiv(info.maxlan + info.nec, 2)) - 1,
                  min(krm, tind - 1))
        while krm + krm > ncl + ncr and res[krm] < bnd:
            krm -= 1
    elif info.lohi < 0:
        klm = max(min(info.ned, _idiv(info.maxlan + info.nec, 2)) - 1, max(tind + 1, klm))
        while klm + klm < ncl + ncr and res[klm] < bnd:
            klm += 1
    else:
        if tind - klm < krm - tind:
            lohi = -1
            klm = tind + 1
        else:
            lohi = 1
            krm = tind - 1
        j = info.locked + klm + nd - krm + 1
        if j > 0:
            j = j // 2
            klm = max(0, klm - j)
            krm = min(krm + j, nd - 1)
    return lohi, klm, krm


def _min_gap_ratio(info, nd, tind, res):
    """trl_min_gap_ratio_ (restart.c:94-117)."""
    gamma = info.maxmv * (info.nec + 1.0) / info.ned - info.matvec
    if gamma < info.klan:
        gamma = max((info.maxmv - info.matvec) / float(info.ned - info.nec), 2.0)
    return min(_log(res[tind] / (info.tol * info.anrm)) / gamma, 0.5)


def _gap_ratio(i, j, tind, lam):
    return (lam[i] - lam[tind]) / (lam[j] - lam[tind])


def _fixed(nd, mnd, tind, lam, res, info, kl, kr):
    """trl_restart_fixed (restart.c:560-689)."""
    kli, kri = kl, kr
    kl0, kr0 = kli, kri
    extra = _nint((mnd - info.nec) * (0.4 + 0.1 * info.ned / float(mnd)))
    if extra > info.ned + info.ned and extra > 5:
        gmin = np.float64(mnd) / np.float64(info.ned)
        extra = _nint((extra + (_log(gmin) * info.ned) * gmin) / (1.0 + gmin))
    minsep = max(3, max(nd // 5, nd - 4 * info.ned))
    gmin = _min_gap_ratio(info, nd, tind, res)
    if info.lohi > 0:
        kri = min(tind - 1, kri) - extra
        kli = -1
    elif info.lohi < 0:
        kli = max(tind + 1, kli) + extra
        kri = nd
    else:
        extra += 1
        kli = kli + extra // 2
        kri = kri - extra // 2
        i = 0
        while kli > kl0 and kri < kr0 and i == 0:
            if 10.0 * res[kli] < res[kri]:
                if res[kli + 1] < res[kri + 1]:
                    kli += 1
                    kri += 1
                else:
                    i = -1
            elif 10.0 * res[kri] < res[kli]:
                if res[kri - 1] < res[kli - 1]:
                    kri -= 1
                    kli -= 1
                else:
                    i = -1
            else:
                i = -1
    while kli + minsep < kri and _gap_ratio(max(0, kli), min(kri, nd - 1), tind, lam) < gmin:
        if info.lohi < 0:
            kli += 1
        elif info.lohi > 0:
            kri -= 1
        elif res[kli] < res[kri]:
            kli += 1
        else:
            kri += 1
    if info.lohi > 0:
        i = kri - 1
        while (i > kli + minsep and lam[kri] - res[kri] < lam[i]
               and lam[kri] + res[kri] < lam[i] + res[i]):
            i -= 1
        kri = i + 1
    else:
        i = kli + 1
        while (i < kri - minsep and lam[kli] + res[kli] > lam[i]
               and lam[kli] - res[kli] > lam[i] - res[i]):
            i += 1
        kli = i - 1
    return kli, kri


def _scan(nd, res, info, kept, kl, kr):
    """trl_restart_scan (restart.c:692-804)."""
    kli, kri = kl, kr
    if info.lohi < 0:
        kri = nd
        kli = kept + min(max(info.nec, 1), (nd - kept) // 2) - 1
        if kli <= 0:
            if nd > 6:
                kli = nd // 2 - 1
            elif nd > 2:
                kli = 1
        elif kli + 3 > nd:
            kli = info.nec + min(info.ned, min(10, (nd - info.ned) // 2)) - 1
    elif info.lohi > 0:
        kli = -1
        kri = kept + min(max(info.nec, 1), (nd - kept) // 2) - 1
        if kri <= 0:
            if nd > 6:
                kri = nd // 2 - 1
            elif nd > 2:
                kri = 1
        elif kri + 4 > nd:
            kri = info.nec + min(info.ned, min(10, (nd - info.ned) // 2))
        kri = nd - kri - 1
    else:
        kl0, kr0 = kli, kri
        extra = kept + min(info.nec, (nd - kept) // 2) + 1
        if extra <= 0:
            if nd > 6:
                extra = nd // 2
            elif nd > 2:
                extra = 2
        elif extra + 3 > nd:
            extra = info.nec + min(info.ned, min(10, (nd - info.ned) // 2))
        kli = max(kli, (extra // 2) - 1)
        kri = min(kri, (nd - extra // 2) - 1)
        i = 0
        while kli > kl0 and kri < kr0 and i == 0:
            if 10.0 * res[kli] < res[kri]:
                if res[kli + 1] < res[kri + 1]:
                    kli += 1
                    kri += 1
                else:
                    i = -1
            elif 10.0 * res[kri] < res[kli]:
                if res[kri - 1] < res[kli - 1]:
                    kri -= 1
                    kli -= 1
                else:
                    i = -1
            else:
                i = -1
    return kli, kri


def _small_res(nd, mnd, tind, lam, res, info, kl, kr):
    """trl_restart_small_res (restart.c:840-988)."""
    minsep = max(3, max(nd // 5, nd - 4 * info.ned))
    extra = _nint((mnd - info.nec) * (0.4 + 0.1 * info.ned / float(mnd)))
    if extra > info.ned + info.ned and extra > 5:
        gmin = np.float64(mnd) / np.float64(info.ned)
        extra = _nint((extra + (_log(gmin) * info.ned) * gmin) / (1.0 + gmin))
    kli, kri = kl, kr
    kl0, kr0 = kli, kri
    resmax = max(res[i] for i in range(nd)) if nd > 0 else 0.0
    acpt = resmax / res[tind]
    if info.lohi > 0:
        if 0.0 <= acpt < 0.999:
            ii = tind - 1
            acpt = max(_sqrt(acpt) * res[tind], res[ii] + res[ii])
            acpt = min(acpt, resmax)
            kri = ii - 1
            while res[kri] < acpt and kri > kli + 3:
                kri -= 1
        else:
            kri = kr0 - extra
        kri = max(2, kri)
        kli = min(kli, kri - 2)
    elif info.lohi < 0:
        if 0.0 <= acpt < 0.999:
            ii = tind + 1
            acpt = max(_sqrt(acpt) * res[tind], res[ii] + res[ii])
            acpt = min(acpt, resmax)
            kli = ii + 1
            while res[kli] < acpt and kli < kri - 3:
                kli += 1
        else:
            kli = kl0 + extra
        kli = min(nd - 4, kli)
        kri = max(kri, kli + 2)
    else:
        i = kli + 1
        j = kri - 1
        for _ in range(1, extra + 1):
            if res[i] < res[j]:
                kli = i
                i += 1
            elif res[i] > res[j]:
                kri = j
                j -= 1
            elif i < nd - j:
                kli = i
                i += 1
            else:
                kri = j
                j -= 1
    kl0, kr0 = kli, kri
    gmin = _min_gap_ratio(info, nd, tind, res)
    done = 0
    while (kli + minsep < kri and _gap_ratio(max(0, kli), min(nd - 1, kri), tind, lam) < gmin
           and done == 0):
        if info.lohi < 0:
            kli += 1
        elif info.lohi > 0:
            kri -= 1
        elif res[kli] < res[kri]:
            kli += 1
        elif kri < nd - 1:
            kri += 1
        else:
            done = 1
    if info.lohi > 0:
        i = kr0 - 1
        while (i > kli + minsep and lam[kri] - res[kri] < lam[i]
               and lam[kri] + res[kri] < lam[i] + res[i]):
            i -= 1
        kri = min(kri, i + 1)
    else:
        i = kl0 + 1
        while (i < kri - minsep and lam[kli] + res[kli] > lam[i]
               and lam[kli] - res[kli] > lam[i] - res[i]):
            i += 1
        kli = max(kli, i - 1)
    return kli, kri


def _max_gap_ratio(nd, tind, kept, lam, res, info, kl, kr):
    """trl_restart_max_gap_ratio (restart.c:991-1094)."""
    def gratio(i, j):
        return DBL_MAX if lam[j] == info.trgt else (lam[i] - info.trgt) / (lam[j] - info.trgt)

    lohi, klm, krm = _search_range(nd, lam, res, info, kl, kr, tind)
    kli, kri = klm, krm
    igap = max(min(nd - info.ned, _nint((krm - klm) * 0.4)), 2)
    if igap > 2 and info.matvec > info.maxlan:
        if info.clk_op + info.tick_o > 10.0 * (info.clk_orth + info.tick_h + info.clk_res
                                              + info.tick_r):
            igap = max(2, nd - kept - 1)
        else:
            bnd = _min_gap_ratio(info, nd, tind, res)
            if info.crat < bnd:
                igap = max(2, nd - kept - 1)
    if kli + igap > kri:
        return kli, kri
    if lohi > 0:
        bnd = gratio(kri, kli)
        for i in range(klm, krm - igap + 1):
            j = i + igap
            t = gratio(j, i)
            if t > bnd:
                kli, kri, bnd = i, j, t
    else:
        bnd = gratio(kli, kri)
        for i in range(klm, krm - igap + 1):
            j = i + igap
            t = gratio(i, j)
            if t > bnd:
                kli, kri, bnd = i, j, t
    return kli, kri


def _max_progress(nd, tind, kept, lam, res, info, kl, kr):
    """trl_restart_max_progress (restart.c:1097-1185)."""
    def merit(i, j):
        return (lam[i] - info.trgt) * abs(j - i) / (lam[j] - info.trgt)

    lohi, klm, krm = _search_range(nd, lam, res, info, kl, kr, tind)
    kli, kri = klm, krm
    igap = max(min(nd - info.ned, _nint((kri - kli) * 0.4)), 2)
    if igap > 2 and igap + kept > nd and info.crat > 0.0:
        ss = _min_gap_ratio(info, nd, tind, res)
        if ss > info.crat:
            igap = max(2, nd - kept - 1)
    if lohi > 0:
        ss = merit(kri, kli)
        for i in range(klm, krm - igap + 1):
            for j in range(i + igap, krm + 1):
                t = merit(j, i)
                if t > ss:
                    ss, kli, kri = t, i, j
    else:
        ss = merit(kli, kri)
        for i in range(klm, krm - igap + 1):
            for j in range(i + igap, krm + 1):
                t = merit(i, j)
                if t > ss:
                    ss, kli, kri = t, i, j
    return kli, kri


def _dyn_gap(info, span):
    """Dynamic minimum-gap heuristic shared by schemes 5 and 7
    (restart.c:1318-1343 and 199-226); updates info.avgm."""
    def1, def2, up, dw = 1.0, 0.7, 1.0, 0.7
    info.avgm = ((info.avgm * (float(info.nloop) - 1) + float(info.klan - info.k + 1))
                 / float(info.nloop))
    z1 = info.ptres / (info.tol * info.anrm)
    z2 = np.float64(1.0) / np.float64(info.cfac)
    if z2 < 1.0:
        return int(span * def2)
    if z1 < 1.0:
        return int(span * def1)
    tmp3 = _log(z1 + _sqrt(z1 - 1.0) * _sqrt(z1 + 1.0)) / (2.0 * info.avgm)
    tmp2 = _log(z2 + _sqrt(z2 - 1.0) * _sqrt(z2 + 1.0)) / (info.klan - info.k + 1)
    tmp = tmp2 / tmp3
    tmp = _pow(tmp, 2.0)
    tmp = _atan(tmp) * (2.0 / 3.141592654)
    tmp = dw + (up - dw) * tmp
    return int(span * tmp)


def _max_reduction(nd, tind, kept, lam, res, info, kl, kr):
    """trl_restart_max_reduction (restart.c:1280-1408)."""
    def merit(i, j):
        return _sqrt((lam[i] - info.trgt) / (lam[j] - lam[i])) * abs(j - i)

    lohi, klm, krm = _search_range(nd, lam, res, info, kl, kr, tind)
    kli, kri = klm, krm
    igap = _dyn_gap(info, abs(krm - klm))
    if igap > 2 and igap + kept > nd and info.crat > 0.0:
        ss = _min_gap_ratio(info, nd, tind, res)
        if ss > info.crat:
            igap = max(2, nd - kept - 1)
    if lohi > 0:
        ss = merit(kri, kli)
        for i in range(klm, krm - igap + 1):
            for j in range(i + igap, krm + 1):
                t = merit(j, i)
                if t > ss:
                    ss, kli, kri = t, i, j
    else:
        ss = merit(kli, kri)
        for i in range(klm, krm - igap + 1):
            for j in range(i + igap, krm + 1):
                t = merit(i, j)
                if t > ss:
                    ss, kli, kri = t, i, j
    return kli, kri


def _max_gap_cost_ratio(n, tind, info, lam, res, kl, kr):
    """trl_restart_max_gap_cost_ratio (restart.c:121-299); sets info.klan."""
    nd = info.ned
    mn = info.maxlan
    lohi, k1, k2 = _search_range(n, lam, res, info, kl, kr, tind)
    info.lohi = lohi
    t = _dyn_gap(info, abs(k2 - k1))
    min_val = 0.0
    min_k1, min_k2, min_m = k1, k2, info.klan
    min_gamma = min_gamma0 = 0.0
    min_l = 0
    for i in range(k1, k2 - t + 1):
        for j in range(i + t, k2 + 1):
            gamma0 = _sqrt((lam[i + 1] - lam[tind]) / (lam[j - 1] - lam[i + 1]))
            l1 = (i + 1) + (n - j)
            l2 = l1 + info.locked
            mn2 = l2 + 1
            if mn2 > mn:
                mn2 = mn
            if mn2 < nd:
                mn2 = nd
            for k in range(mn2, mn + 1):
                gamma = (k - l2) * gamma0
                val = ((k - l2) * (k + l2 - 1) + k * l1) / gamma
                if min_val == 0.0 or val < min_val:
                    min_k1, min_k2, min_m = i, j, k
                    min_val = val
                    min_gamma, min_gamma0, min_l = gamma, gamma0, l2
    info.klan = min_m
    info.k1, info.k2, info.k = min_k1, min_k2, min_l
    info.predicted_crate = _exp(-min_gamma)
    info.mgamma = min_gamma
    info.gamma0 = min_gamma0
    info.old_target = lam[tind]
    info.target_id = tind
    info.old_locked = info.locked
    return min_k1, min_k2


def _max_gap_cost_ratio_static(n, tind, info, lam, res, kl, kr):
    """trl_restart_max_gap_cost_ratio_static (restart.c:302-432)."""
    nd = info.ned
    mn = info.maxlan
    lohi, k1, k2 = _search_range(n, lam, res, info, kl, kr, tind)
    info.lohi = lohi
    if info.crat > 0.0:
        tmp2 = -_log(info.crat)
        tmp3 = -(tmp2 * info.maxlan) / _log(info.tol * info.anrm)
        if tmp3 >= 0.5 or info.klan < info.maxlan:
            t = int(abs(k2 - k1) * 0.8)
        else:
            tmp = _min_gap_ratio(info, nd, tind, res)
            tmp = tmp2 / tmp
            tmp = _pow(tmp, 0.25)
            tmp = _atan(tmp) * (4.0 / 3.141592654)
            t = int(abs(k2 - k1) * tmp)
    else:
        t = int(abs(k2 - k1) * 0.8)
    mn1 = _nint(2.0 * info.klan / 5.0)
    min_val = 0.0
    min_k1, min_k2, min_m = k1, k2, info.klan
    for i in range(k1, k2 - t + 1):
        for j in range(i + t, k2 + 1):
            gamma0 = _sqrt((lam[i + 1] - lam[tind]) / (lam[j - 1] - lam[i + 1]))
            l1 = (i + 1) + (n - j)
            l2 = l1 + info.locked
            mn2 = mn1
            if mn1 < info.rfact * l2:
                mn2 = int(info.rfact * l2)
            if mn2 > mn:
                mn2 = mn
            if mn2 < nd:
                mn2 = nd
            gamma = (mn2 - l2 + 1) * gamma0
            val = (mn2 * mn2 - l2 * l2 + 2 * mn2 * l1) / gamma
            if min_val == 0.0 or val < min_val:
                min_k1, min_k2, min_m = i, j, mn2
                min_val = val
    info.klan = min_m
    return min_k1, min_k2


def shuffle_eig(nd, mnd, lam, res, info, kept, locked):
    """trl_shuffle_eig (restart.c:1410-1651).  lam/res are the non-locked
    Ritz values/residuals (sorted ascending); compacted in place.  Returns the
    new `kept`.  Floating-point follows IEEE like the C code (division by
    zero gives inf, sqrt of a negative gives nan)."""
    with np.errstate(all="ignore"):
        return _shuffle_eig(nd, mnd, lam, res, info, kept, locked)


def _shuffle_eig(nd, mnd, lam, res, info, kept, locked):
    if nd <= 5:
        dsort2(nd, res, lam)
        k = 2 if nd > 3 else (1 if nd > 0 else 0)
        if k > 1:
            dsort2(k, lam, res)
        return k
    bnd = min(info.tol, DBL_EPSILON * info.anrm)
    ncr = 0
    ncl = nd - 1
    i = nd - 1
    while i >= 0:
        if res[i] <= bnd:
            i -= 1
        else:
            ncr = i + 1
            i = -1
    i = 0
    while i < nd:
        if res[i] <= bnd:
            i += 1
        else:
            ncl = i - 1
            i = nd
    kl, kr = ncl, ncr
    if ncr > ncl:
        tind = _idiv(kl + kr, 2)
        while lam[tind] != info.trgt and kr > kl:
            if lam[tind] < info.trgt:
                kl = tind + 1
                tind = _idiv(kl + kr, 2)
            elif lam[tind] > info.trgt:
                kr = tind - 1
                tind = _idiv(kl + kr, 2)
            else:
                kl = tind
                kr = tind
        if lam[tind] == info.trgt:
            kl = tind - 1
            while kl >= 0 and lam[kl] == info.trgt:
                kl -= 1
            kr = tind + 1
            while kr < nd and lam[kr] == info.trgt:
                kr += 1
        else:
            kl = tind - 1
            kr = tind + 1
        if info.lohi > 0:
            kr = kl
            kl = min(ncl, max(0, nd - info.ned) - 1)
        elif info.lohi < 0:
            kl = kr
            kr = max(ncr, min(nd - info.nec, info.ned + 1) - 1)
        elif ncr - tind > tind - ncl:
            kl = kr
            kr = max(ncr, min(nd - info.nec, info.ned + 1) - 1)
        else:
            kr = kl
            kl = min(ncl, max(0, nd - info.ned) - 1)
    else:
        return nd
    r = info.restart
    if r == 0:
        if info.restart <= -info.ned:
            # unreachable for restart == 0 with ned > 0 (restart.c:1546-1558)
            pass
        kl, kr = _fixed(nd, mnd, tind, lam, res, info, kl, kr)
    elif r == 1:
        kl, kr = _fixed(nd, mnd, tind, lam, res, info, kl, kr)
    elif r == 2:
        kl, kr = _small_res(nd, mnd, tind, lam, res, info, kl, kr)
    elif r == 3:
        if info.nloop > 0:
            kl, kr = _max_gap_ratio(nd, tind, kept, lam, res, info, kl, kr)
        else:
            kl, kr = _fixed(nd, mnd, tind, lam, res, info, kl, kr)
    elif r == 4:
        if info.nloop > 0:
            kl, kr = _max_progress(nd, tind, kept, lam, res, info, kl, kr)
        else:
            kl, kr = _fixed(nd, mnd, tind, lam, res, info, kl, kr)
    elif r == 5:
        if info.nloop > 0:
            kl, kr = _max_reduction(nd, tind, kept, lam, res, info, kl, kr)
        else:
            kl, kr = _fixed(nd, mnd, tind, lam, res, info, kl, kr)
    elif r == 6:
        kl, kr = _scan(nd, res, info, kept, kl, kr)
    elif r == 8:
        if info.nloop > 0:
            kl, kr = _max_gap_cost_ratio_static(nd, tind, info, lam, res, kl, kr)
        else:
            kl, kr = _fixed(nd, mnd, tind, lam, res, info, kl, kr)
    else:  # 7 and any other value (restart.c:1603-1611)
        if info.nloop > 0:
            kl, kr = _max_gap_cost_ratio(nd, tind, info, lam, res, kl, kr)
        else:
            kl, kr = _fixed(nd, mnd, tind, lam, res, info, kl, kr)
    minsep = max(3, max(nd // 6, nd - 6 * info.ned))
    if kr <= kl + minsep or kl + nd - kr + minsep > mnd:
        if ncl < kl < kr < ncr:
            kl -= 1
            kr += 1
        elif info.lohi > 0:
            kr = max(minsep, min(nd // 3, ncr)) - 1
            kl = -1
        elif info.lohi < 0:
            kl = min(nd - minsep, max((nd + nd) // 3, ncl + 2)) - 1
            kr = nd
        else:
            kl = _idiv(nd - minsep, 2) - 2
            kr = _idiv(nd - minsep + 1, 2)
    kr -= 1
    for i in range(1, nd - kr):
        lam[kl + i] = lam[kr + i]
        res[kl + i] = res[kr + i]
    kept = kl + max(1, nd - kr)
    if info.lohi < -1:
        dsort2(kept, lam, res)
    return kept
