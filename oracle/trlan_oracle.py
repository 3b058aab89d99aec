"""CPU restatement of the reference thick-restart Lanczos (real symmetric, FP64).

TEST INFRASTRUCTURE ONLY.  Only tests/, the smoke check and the bench's CPU
baseline leg may import this module, and only as the checker.  The product
(nutrlan_amd + its HIP library) never calls into it.

This module restates, step by step, the serial reference code path
  trlan()                      SRC/trlan.c:21-301
  trl_clear_counter()          SRC/trlan.c:1217-1349
  trl_init_info()              SRC/trl_comm_none.c:27-192
  trlanczos()                  SRC/trlcore.c:723-1532
  trl_initial_guess()          SRC/trlcore.c:1577-1763
  trl_orth() / trl_cgs()       SRC/trlcore.c:1766-1975 / 2913-3117
  trl_tridiag()                SRC/trlcore.c:1978-2074
  trl_get_eval() (+dstqrb)     SRC/trlcore.c:2077-2146, SRC/dstqrb.c:22-602
  trl_convergence_test()       SRC/trlcore.c:2149-2305
  trl_sort_eig()               SRC/trlcore.c:2308-2378
  trl_get_tvec(_a)()           SRC/trlcore.c:2381-2673
  trl_set_locking()            SRC/trlcore.c:2676-2758
  trl_ritz_vectors()           SRC/trlcore.c:2761-2910
  trl_shuffle_eig() + schemes  SRC/restart.c (see oracle/restart_oracle.py)
with numpy for the O(n) vector work and the reference's own LAPACK calls
(dsytrd/dorgtr/dstein/dsyev, and dsteqr for the reference's one-row dsteqr
variant dstqrb) bound through oracle/lapack.py.

The Lanczos basis is one n x (maxlan+1) array: logical column k is evec[:,k]
for k < mev and base[:,k-mev] beyond, exactly how the reference splits it.

The complex Hermitian path (ztrlan, SRC/ztrlan.c:26-221, SRC/ztrlcore.c)
runs through the same functions with complex_=True: conjugated dots,
ztrl_cgs's ntot-scaled randomization threshold and perturbation retries,
and optionally (gdot_quirk) the reference's two-column ztrl_g_dot_ result.

Pinning (tests/test_oracle.py): the integer trajectory (matvec, nloop,
north, nrand, locked) and eigenvalues agree exactly / to 1e-12 ||A|| with
the compiled reference's recorded outputs (tests/golden/ref_cases.json) for
the real cases; complex eigenvalues agree to 1e-12 ||A||; closed forms
(diagonal operators) and LAPACK eigh of small dense matrices pin the rest;
a planted mistake must break the agreement.  Random-vector paths (guess <= 0
with perturbation, trl_cgs random injection) use numpy's generator instead
of the reference's clock()-seeded drand48 and are parity unpinned.
"""
import math

import numpy as np

from . import lapack
from .restart_oracle import Strip, shuffle_eig
from .sorts import dsort2, dsort2a, dsort2s, dsort2su, dsort2sd

DBL_EPSILON = np.finfo(np.float64).eps
DBL_MIN = np.finfo(np.float64).tiny
DBL_MAX = np.finfo(np.float64).max


class Info:
    """Python mirror of the reference trl_info fields the algorithm uses."""

    def __init__(self):
        self.stat = 0
        self.lohi = 0
        self.ned = 0
        self.nec = 0
        self.tol = 0.0
        self.maxlan = 0
        self.klan = 0
        self.maxmv = 0
        self.restart = 0
        self.locked = 0
        self.guess = 0
        self.matvec = 0
        self.nloop = 0
        self.north = 0
        self.nrand = 0
        self.anrm = 0.0
        self.nloc = 0
        self.ntot = 0
        self.crat = 1.0
        self.trgt = -DBL_MAX
        self.tres = -1.0
        self.tmv = -1
        self.avgm = 0.0
        self.cfac = 0.0
        self.ptres = 0.0
        self.predicted_crate = 0.0
        self.mgap = 0.0
        self.mgamma = 0.0
        self.gamma0 = 0.0
        self.old_target = 0.0
        self.target_id = 0
        self.old_locked = 0
        self.k1 = 0
        self.k2 = 0
        self.k = 0
        self.rfact = 0.0
        self.ref = 0.0
        self.cpflag = 0
        # timing fields used only by restart scheme 3 (SRC/restart.c:1052-1054)
        self.clk_op = 0.0
        self.tick_o = 0.0
        self.clk_orth = 0.0
        self.tick_h = 0.0
        self.clk_res = 0.0
        self.tick_r = 0.0


def init_info(nrow, maxlan, lohi, ned, tol=1.4901e-8, restart=1, maxmv=0):
    """trl_init_info (SRC/trl_comm_none.c:27-192).

    The reference reads info->ntot before assigning it when maxmv <= 0
    (trl_comm_none.c:104); here the freshly known nrow is used instead.
    """
    info = Info()
    if tol > 0:
        info.tol = tol
        if info.tol <= DBL_MIN:
            info.tol = DBL_EPSILON
        elif info.tol > 1.0:
            info.tol = min(0.1, 1.0 / info.tol)
    else:
        info.tol = math.sqrt(DBL_EPSILON)
    info.restart = restart if restart > 0 else 0
    if maxmv > 0:
        info.maxmv = maxmv
    else:
        info.maxmv = min(max(nrow, 1000), 1000 * ned)
    info.maxlan = maxlan
    if maxlan <= ned:
        info.maxlan = ned + max(ned, 6)
    info.lohi = lohi
    info.ned = ned
    info.nloc = nrow
    info.ntot = nrow
    return info


def set_iguess(info, nec, iguess):
    """trl_set_iguess (SRC/trlan.c:405-470) without checkpoint files."""
    info.nec = nec
    info.guess = iguess
    info.stat = 0


def clear_counter(info, nrow, mev, lde):
    """trl_clear_counter (SRC/trlan.c:1217-1349)."""
    info.stat = 0
    if nrow != info.nloc or nrow > info.ntot:
        info.stat = -1
    if info.nec < 0:
        info.nec = 0
    if lde < nrow:
        info.stat = -2
    if info.tol >= 1.0:
        info.tol = math.sqrt(DBL_EPSILON)
    elif info.tol <= DBL_MIN:
        info.tol = DBL_EPSILON
    if info.ned + info.ned >= info.ntot:
        if info.ned > info.ntot:
            info.ned = min(info.ntot - 1, info.maxlan - 3)
    if mev < info.ned:
        info.stat = -3
    if info.ntot < 10:
        info.stat = -4
    if info.maxlan < info.ned + 3:
        info.maxlan = info.ned + min(info.ned, 20) + int(math.log(float(info.ntot)))
        info.maxlan = min(info.maxlan, info.ntot)
    if info.maxlan < mev:
        ntmp = min(info.ntot, max(100 + info.ned, 10 * info.ned))
        info.maxlan = mev if mev < ntmp else ntmp
    if info.maxlan < 5:
        info.stat = -5
    info.tmv = -1
    info.klan = min(info.maxlan, info.ntot)
    if info.restart >= 7:
        info.klan = min(info.maxlan, max(100, min(info.klan, 2 * info.ned)))
    info.locked = info.nec
    info.matvec = 0
    info.nloop = 0
    info.north = 0
    info.nrand = 0
    info.avgm = 0.0


def smooth_rr(rr):
    """trl_smooth_rr (SRC/trlcore.c:1534-1575), end-case quirks included."""
    n = len(rr)
    if n <= 0:
        return
    rr2 = rr[0]
    rr[0] = 2 * rr[0] + rr[2] + rr[n - 1]
    for i in range(1, n - 1):
        tmp = rr[i]
        rr[i] = 2 * rr[i] + rr[i + 1] + rr2
        rr2 = tmp
    rr[n - 1] = 2 * rr[n - 1] + rr[1] + rr2


class _State:
    """Per-solve arrays of trlanczos (SRC/trlcore.c:756-769); a complex
    basis for ztrlanczos (SRC/ztrlcore.c:556-584)."""

    def __init__(self, n, mev, maxlan, rng, complex_=False, gdot_quirk=True):
        self.n = n
        self.mev = mev
        self.complex = complex_
        self.gdot_quirk = gdot_quirk
        self.nbas = max(1, maxlan - mev + 1)
        self.V = np.zeros((n, mev + self.nbas), order="F",
                          dtype=np.complex128 if complex_ else np.float64)
        self.alpha = np.zeros(maxlan)
        self.beta = np.zeros(maxlan)
        self.alfrot = np.zeros(maxlan)
        self.betrot = np.zeros(maxlan)
        self.iwrk = np.zeros(4 * maxlan, dtype=np.int32)
        self.rng = rng

    def cols(self, j1, j2):
        """Logical basis columns [evec(:,0:j1), base(:,0:j2)]."""
        return list(range(j1)) + [self.mev + k for k in range(j2)]

    def random_vector(self):
        if self.complex:
            v = self.rng.random(self.n) + 1j * self.rng.random(self.n)
            smooth_zz(v)
        else:
            v = self.rng.random(self.n)
            smooth_rr(v)
        return v


def smooth_zz(z):
    """trl_smooth_zz (SRC/ztrlcore.c:1289-1332), including the last
    element's imaginary part taking twice its own updated real part."""
    n = len(z)
    if n <= 0:
        return
    r2, i2 = z[0].real, z[0].imag
    z0 = complex(2 * z[0].real + z[2 % n].real + z[n - 1].real,
                 2 * z[0].imag + z[2 % n].imag + z[n - 1].imag)
    z[0] = z0
    for i in range(1, n - 1):
        r1, i1 = z[i].real, z[i].imag
        z[i] = complex(2 * z[i].real + z[i + 1].real + r2, 2 * z[i].imag + z[i + 1].imag + i2)
        r2, i2 = r1, i1
    re = 2 * z[n - 1].real + z[1 % n].real + r2
    # the imaginary part adds twice the real part as just updated
    # (ztrlcore.c:1327-1328)
    z[n - 1] = complex(re, 2 * re + z[1 % n].imag + i2)


def _gdot(st, c1, c2, rr):
    """[V1 V2]^H r.  With st.gdot_quirk the reference's two-column branches
    of ztrl_g_dot_ are reproduced (SRC/ztrl_comm_none.c:88-101 drop the
    imaginary part of the second value; 109-120 add it into the real part
    of the last value); otherwise the exact product."""
    V = st.V
    h = V[:, c1 + c2].conj().T @ rr
    if st.complex and st.gdot_quirk:
        m1 = len(c1)
        if m1 == 2:
            h[1] = h[1].real
        if len(c2) == 2:
            v = V[:, c2[1]]
            t = float(np.sum(v.real * rr.real + v.imag * rr.imag)
                      + np.sum(v.real * rr.imag - v.imag * rr.real))
            h[m1 + 1] = t
    return h


def cgs(info, st, cols, rr_col, rnrm, counter):
    """trl_cgs (SRC/trlcore.c:2913-3117) / ztrl_cgs (SRC/ztrlcore.c:
    1977-2203).  Returns (stat, rnrm).

    ``counter`` is a one-element list standing in for the `north` pointer.
    The reference subtracts the evec part and the base part with two separate
    gemv calls (2996-3011); here the two parts are subtracted in that order.
    """
    V = st.V
    n = st.n
    rr = V[:, rr_col]
    nold = len(cols)
    maxorth = 3
    irnd = 0
    ierr = 0
    normrr2 = 0.0
    normDiff2 = 0.0
    if nold > 0:
        c1 = [c for c in cols if c < st.mev]
        c2 = [c for c in cols if c >= st.mev]
        cnt = 0
        while cnt <= maxorth:
            h = _gdot(st, c1, c2, rr)
            h1 = h[: len(c1)]
            h2 = h[len(c1):]
            if len(c1) > 0:
                rr -= V[:, c1] @ h1
            if len(c2) > 0:
                rr -= V[:, c2] @ h2
            counter[0] += 1
            cnt += 1
            normDiff2 = float(np.vdot(h, h).real)
            normrr2 = float(np.vdot(rr, rr).real)
            rnrm = math.sqrt(normrr2)
            if st.complex:
                old_rnrm = math.sqrt(normrr2 + normDiff2)
                if DBL_EPSILON * normrr2 > normDiff2:
                    break
                if (((cnt > 1 and not (normDiff2 <= info.ntot * DBL_EPSILON * DBL_EPSILON
                                       * (normrr2 + normDiff2)))
                     or not (normrr2 > DBL_MIN)) and irnd < maxorth):
                    cnt = 0
                    irnd += 1
                    info.nrand += 1
                    _perturb_complex(st, rr, irnd, rnrm, old_rnrm)
                continue
            if DBL_EPSILON * normrr2 > normDiff2:
                break
            elif (cnt > 1 and nold * DBL_EPSILON * DBL_EPSILON * (normrr2 + normDiff2)
                  < normDiff2) or not (normrr2 > DBL_MIN):
                if irnd > 0:
                    break
                cnt = 0
                irnd += 1
                info.nrand += 1
                rr[:] = st.random_vector()
        if st.complex:
            if not (normrr2 >= normDiff2):
                ierr = -203
        elif normrr2 < normDiff2:
            return -203, rnrm
    if ierr == 0:
        if rnrm > DBL_MIN:
            rr *= 1.0 / rnrm
        else:
            return -204, rnrm
    if irnd > 0:
        rnrm = 0.0
    return ierr, rnrm


def _perturb_complex(st, rr, irnd, rnrm, old_rnrm):
    """Random retry of ztrl_cgs (SRC/ztrlcore.c:2117-2178), numpy draws
    (parity unpinned: the reference draws from drand48)."""
    n = st.n
    rng = st.rng
    if irnd == 1 and rnrm > 0.0 and rnrm > DBL_EPSILON * old_rnrm:
        i = int(n * rng.random())
        k = i + int(max(1.0, n * math.sqrt(DBL_EPSILON * old_rnrm / rnrm)))
        for j in range(i, min(k, n)):
            t = rng.random()
            rr[j] += rnrm * (t - 0.5) * (1 + 1j)
    else:
        i = int(n * rng.random())
        j = int(n * rng.random())
        lo, hi = (0, n - 1) if i == j else (min(i, j), max(i, j))
        rr[lo:hi + 1] = rng.random(hi + 1 - lo) + 1j * rng.random(hi + 1 - lo)
    smooth_zz(rr)


def guess_cap(info, nrow, mev, lwrk, complex_):
    """The workspace bound on the number of random perturbations of a
    guess <= 0 start.  trlan() hands trlanczos `nmis` doubles of misc
    workspace (SRC/trlan.c:172-226: maxlan*(maxlan+10), or what the
    caller's wrk holds beyond the basis), and trl_initial_guess gets what
    follows alpha/beta/alfrot/betrot in it (trlcore.c:805-806).  ztrlan()
    passes the caller's nmis whole (ztrlcore.c:617-618)."""
    if complex_:
        return lwrk
    maxlan = info.maxlan
    ldb = ((nrow + 3) // 4) * 4
    if ldb % 4096 == 0:
        ldb += 4
    ii = max(1, maxlan - mev + 1) * ldb
    nmis = maxlan * (maxlan + 10)
    if lwrk >= min(ii, nmis):
        if lwrk >= ii + nmis:
            nmis = lwrk - ii
        elif lwrk >= max(ii, nmis):
            if ii < nmis:
                nmis = lwrk
        elif ii > nmis:
            nmis = lwrk
    return nmis - 4 * maxlan


def initial_guess(info, st, start):
    """trl_initial_guess (SRC/trlcore.c:1577-1763) / ztrl_initial_guess
    (SRC/ztrlcore.c:1335-1567) for guess <= 1.

    Returns (j1, j2).  guess == 1 takes evec[:,nec] as given (the path every
    parity test uses); guess <= 0 uses [1,...,1] ((1,0) or (1,1) for complex)
    plus the reference's perturbation count (numpy draws, parity unpinned).
    """
    n = st.n
    j = info.nec
    if info.guess > 1:
        raise NotImplementedError("checkpoint restart (guess > 1)")
    col = st.V[:, j]
    if info.guess == 1:
        col[:] = start
    elif st.complex:
        col[:] = 1.0 if info.guess == 0 else 1.0 + 1.0j
        if info.guess < 0:
            nran = 2 * (min(-info.guess, st.guess_cap) // 2)
            if nran < n:
                w = st.rng.random(nran) + 1j * st.rng.random(nran)
                for i in range(0, nran - 1, 2):
                    ii = int(n * w[i].real)
                    col[ii] += w[i + 1] - (0.5 + 0.5j)
            else:
                col[:] = st.random_vector()
                info.nrand += 1
    else:
        col[:] = 1.0
        nran = min(1 - info.guess, st.guess_cap)
        nran = 2 * (nran // 2)
        if 0 < nran < n:
            w = st.rng.random(nran)
            for i in range(0, nran - 1, 2):
                ii = int(n * w[i])
                col[ii] += w[i + 1] - 0.5
        elif nran >= n:
            col[:] = st.random_vector()
            info.nrand += 1
    j1, j2 = info.nec, 0
    normr2 = float(np.vdot(col, col).real)
    rnrm = 0.0
    if DBL_MIN <= normr2 <= DBL_MAX:
        rnrm = math.sqrt(normr2)
    else:
        col[:] = st.random_vector()
        info.nrand += 1
    local = [0]
    if j1 < st.mev:
        info.stat, _ = cgs(info, st, st.cols(j1, 0), j1, rnrm, local)
    else:
        raise NotImplementedError("nec >= mev on entry")
    return j1, j2


def orth(info, st, j1, j2, rr_col, kept):
    """trl_orth (SRC/trlcore.c:1766-1975) / ztrl_orth (SRC/ztrlcore.c:
    1754-1974)."""
    V = st.V
    alpha, beta = st.alpha, st.beta
    jnd = j1 + j2
    jm1 = jnd - 1
    rr = V[:, rr_col]
    normrr2 = float(np.vdot(rr, rr).real)
    if not (normrr2 >= 0.0) or not (normrr2 <= DBL_MAX):
        info.stat = -102
        return
    info.stat = 0
    beta[jnd - 1] = math.sqrt(normrr2)
    tmp = alpha[jnd - 1] * alpha[jnd - 1]
    if jm1 > kept:
        tmp += beta[jm1 - 1] * beta[jm1 - 1]
    elif kept > 0:
        tmp += float(beta[:jm1] @ beta[:jm1])
    if jm1 == kept:
        usecgs = 1
    elif jnd >= info.ntot:
        usecgs = 0
    elif DBL_EPSILON * normrr2 >= tmp:
        anorm = 0.0
        for i in range(jnd):
            d = abs(alpha[i]) + abs(beta[i])
            if d > anorm:
                anorm = d
        usecgs = 1 if beta[jm1] < DBL_EPSILON * anorm * info.ntot else 0
    else:
        usecgs = 1
    cols = st.cols(j1, j2)
    if usecgs:
        ctr = [info.north]
        info.stat, beta[jnd - 1] = cgs(info, st, cols, rr_col, beta[jnd - 1], ctr)
        info.north = ctr[0]
    elif jnd > 1:
        qa = V[:, cols[-1]]
        qb = V[:, cols[-2]]
        w0 = np.vdot(qa, rr)
        w1 = np.vdot(qb, rr)
        alpha[jnd - 1] += w0.real
        rr -= w0 * qa
        rr -= w1 * qb
        rr *= 1.0 / beta[jnd - 1]
    else:
        qa = V[:, cols[-1]]
        w0 = np.vdot(qa, rr)
        alpha[jnd - 1] += w0.real
        rr -= w0 * qa
        rr *= 1.0 / beta[jnd - 1]
    if info.stat == 0:
        if beta[jnd - 1] <= DBL_EPSILON * abs(alpha[jnd - 1]):
            beta[jnd - 1] = 0.0
        elif jnd >= info.ntot:
            beta[jnd - 1] = 0.0


def tridiag(nd, alpha, beta):
    """trl_tridiag (SRC/trlcore.c:1978-2074): arrowhead -> tridiagonal.

    Returns (rot (nd x nd), alfrot (nd), betrot (nd), ierr).
    """
    if nd == 0:
        return np.zeros((0, 0)), np.zeros(0), np.zeros(0), 0
    if nd <= 1:
        return np.ones((1, 1)), np.array([alpha[0]]), np.array([beta[0]]), 0
    A = np.zeros((nd, nd), order="F")
    for i in range(nd):
        A[i, i] = alpha[i]
    for i in range(nd - 1):
        A[i, nd - 1] = beta[i]
        A[nd - 1, i] = beta[i]
    Ar, d, e, tau, ierr = lapack.dsytrd_upper(A)
    if ierr != 0:
        return None, None, None, -112
    betrot = np.zeros(nd)
    betrot[: nd - 1] = e[: nd - 1]
    betrot[nd - 1] = beta[nd - 1]
    Q, ierr = lapack.dorgtr_upper(Ar, tau)
    if ierr != 0:
        return None, None, None, -113
    return Q, d.copy(), betrot, 0


def get_eval(nd, locked, alpha, beta):
    """trl_get_eval (SRC/trlcore.c:2077-2146).

    dstqrb (SRC/dstqrb.c) is LAPACK dsteqr applying its rotations to the last
    row only; dsteqr(compz='I') gives the same eigenvalues and the last row of
    its eigenvector matrix is that row.
    """
    lam = np.array(alpha[:nd], dtype=np.float64)
    res = np.zeros(nd)
    m = nd - locked
    if m > 0:
        if m == 1:
            w = np.array([lam[locked]])
            z = np.array([1.0])
            ierr = 0
        else:
            w, Z, ierr = lapack.dsteqr_vectors(lam[locked:nd], beta[locked:nd - 1])
            z = Z[m - 1, :]
        if ierr != 0:
            return lam, res, -122
        lam[locked:] = w
        res[locked:] = beta[nd - 1] * np.abs(z)
    return lam, res, 0


def convergence_test(nd, lam, res, info):
    """trl_convergence_test (SRC/trlcore.c:2149-2305); sorts lam/res in place
    when lohi < -1."""
    if info.lohi < -1:
        dsort2s(nd, info.ref, lam, res)
    wl = lam[:nd].copy()
    wr = np.abs(res[:nd])
    if info.lohi == -2:
        dsort2s(nd, info.ref, wl, wr)
    elif info.lohi == -3:
        dsort2su(nd, info.ref, wl, wr)
    elif info.lohi == -4:
        dsort2sd(nd, info.ref, wl, wr)
    else:
        dsort2(nd, wl, wr)
    if info.tmv > 0 and info.matvec > info.tmv:
        j = 0
        bnd = abs(lam[0] - info.trgt)
        for i in range(nd):
            if abs(lam[i] - info.trgt) < bnd:
                bnd = abs(lam[i] - info.trgt)
                j = i
        if info.tres > res[j]:
            bnd = res[j] / info.tres
            if bnd > 0.0:
                info.crat = math.exp(math.log(bnd) / float(info.matvec - info.tmv))
                info.cfac = bnd
            else:
                info.crat = 1.0
                info.cfac = bnd
        else:
            info.crat = 1.0
            info.cfac = bnd
    # the second-smallest Ritz value, not the smallest (trlcore.c:2230-2231)
    info.anrm = max(info.anrm, max(abs(wl[1]), abs(wl[nd - 1])))
    bnd = DBL_MIN + info.tol * info.anrm
    ncl = 0
    ncr = nd
    if info.lohi <= 0:
        ncl = nd - 1
        i = 0
        while i < nd:
            if wr[i] < bnd:
                if info.lohi == -3 and wl[i] < info.ref:
                    ncl = i - 1
                    i = nd
                elif info.lohi == -4 and wl[i] > info.ref:
                    ncl = i - 1
                    i = nd
                else:
                    i += 1
            else:
                ncl = i - 1
                i = nd
    if info.lohi >= 0:
        ncr = 0
        i = nd - 1
        while i >= 0:
            if wr[i] < bnd:
                i -= 1
            else:
                ncr = i + 1
                i = -1
    info.tmv = info.matvec
    info.ptres = info.trgt
    if info.lohi < 0:
        info.nec = ncl + 1
        info.trgt = wl[min(nd - 1, ncl + 1)]
        info.tres = wr[min(nd - 1, ncl + 1)]
    elif info.lohi > 0:
        info.nec = nd - ncr
        info.trgt = wl[max(0, ncr - 1)]
        info.tres = wr[max(0, ncr - 1)]
    else:
        if ncr <= ncl:
            ncl = nd // 2
            ncr = ncl + 1
            info.trgt = wl[(nd + 1) // 2 - 1]
            info.tres = wr[(nd + 1) // 2 - 1]
        elif wr[ncl + 1] <= wr[ncr - 1]:
            info.trgt = wl[ncl + 1]
            info.tres = wr[ncl + 1]
        else:
            info.trgt = wl[ncr - 1]
            info.tres = wr[ncr - 1]
        info.nec = ncl + nd - ncr + 1
        for i in range(ncl + 1, ncr):
            if wr[i] < bnd:
                info.nec += 1
    # what the reference leaves in its workspace: |res| then lambda, sorted
    return np.concatenate([wr, wl])


def sort_eig(nd, lohi, nec, ref, lam, res):
    """trl_sort_eig (SRC/trlcore.c:2308-2378)."""
    if lohi == 0:
        dsort2a(nd, res, lam)
        dsort2(nec, lam, res)
    elif lohi == -2:
        dsort2s(nd, ref, lam, res)
        dsort2(nec, lam, res)
    elif lohi == -3:
        dsort2su(nd, ref, lam, res)
        dsort2(nec, lam, res)
    elif lohi == -4:
        dsort2sd(nd, ref, lam, res)
        dsort2(nec, lam, res)
    else:
        dsort2(nd, lam, res)
        if lohi > 0:
            j = nd - nec
            for i in range(nec):
                res[i] = res[j]
                lam[i] = lam[j]
                j += 1


def get_tvec(nd, alpha, beta, irot, nrot, rot, nlam, lam, iwrk):
    """trl_get_tvec (SRC/trlcore.c:2381-2490): dstein + rotation.

    Returns (yy (nd x nlam), ierr).  iwrk is the solver's shared integer
    workspace; dstein's iblock/isplit/iwork/ifail live in it as in the
    reference (the dsyev path's index map reads what is left there).
    """
    if nlam <= 0:
        return np.zeros((nd, 0)), 0
    iwrk[0:nd] = 1
    iwrk[nd:2 * nd] = nd
    Z, iw, ifail, ierr = lapack.dstein(alpha[:nd], beta[:nd], lam[:nlam],
                                       iwrk[0:nd].copy(), iwrk[nd:2 * nd].copy())
    iwrk[2 * nd:3 * nd] = iw[:nd]
    iwrk[3 * nd:3 * nd + min(nd, nlam)] = ifail[: min(nd, nlam)]
    if ierr != 0:
        return None, -132
    yy = np.asfortranarray(Z.copy())
    if nrot > 1:
        yy[irot:irot + nrot, :] = rot[:nrot, :nrot] @ yy[irot:irot + nrot, :]
    return yy, 0


def get_tvec_a(nd, kept, alpha, beta, nlam, lam, iwrk):
    """trl_get_tvec_a (SRC/trlcore.c:2493-2673): dsyev of the arrowhead.

    Overwrites alpha[:nd] with the (reordered) eigenvalues like the
    reference; the index-map loops keep the reference's `< i2` bounds
    (2639-2653).  Returns (yy (nd x nd, first nlam columns valid), ierr).
    """
    if nlam <= 0:
        return np.zeros((nd, 0)), 0
    A = np.zeros((nd, nd), order="F")
    for i in range(nd):
        A[i, i] = alpha[i]
    if kept > 0:
        A[0:kept, kept] = beta[0:kept]
    for i in range(kept, nd - 1):
        A[i, i + 1] = beta[i]
    w, Z, ierr = lapack.dsyev_vectors(A)
    if ierr != 0:
        return None, -142
    alpha[:nd] = w
    yy = np.asfortranarray(Z.copy())
    if nlam >= nd:
        return yy, 0
    tmp = max(alpha[nd - 1] - alpha[0], max(abs(alpha[nd - 1]), abs(alpha[0])))
    tmp = DBL_EPSILON * tmp * nd
    j = 0
    i = 0
    while i < nlam:
        ii = j
        j = nd - 1
        while ii < nd:
            if alpha[ii] < lam[i] - tmp:
                ii += 1
            else:
                j = ii
                ii = nd
        if alpha[j] > lam[i] + tmp:
            return None, -143
        ii = i + 1
        i2 = nlam - 1
        while ii < nlam:
            if lam[ii] <= lam[i] + tmp:
                ii += 1
            else:
                i2 = ii - 1
                ii = nd
        ii = j + 1
        j2 = nd - 1
        while ii < nd:
            if alpha[ii] <= lam[i] + tmp:
                ii += 1
            else:
                j2 = ii - 1
                ii = nd
        if j2 == j and i2 == i:
            iwrk[i] = j
        elif j2 - j == i2 - i:
            for ii in range(i, i2):
                iwrk[ii] = j + ii - i
        elif j2 - j > i2 - i:
            for ii in range(i, i2):
                iwrk[ii] = j + ii - i
        elif j2 < nd:
            i2 = i + j2 - j
            for ii in range(i, i2):
                iwrk[ii] = j + ii - i
        else:
            return None, -144
        i = i2 + 1
        j = j2 + 1
    for i in range(nlam):
        j = int(iwrk[i])
        if j > i:
            alpha[i] = alpha[j]
            yy[:, i] = yy[:, j]
    return yy, 0


def set_locking(jnd, nlam, lam, res, yy, anrm):
    """trl_set_locking (SRC/trlcore.c:2676-2758); returns the lock count."""
    eps = DBL_EPSILON

    def small(t):
        return eps * abs(t) if abs(t) >= eps else eps * eps * anrm

    i = 0
    j = nlam - 1
    ti = abs(res[i]) <= small(lam[i])
    tj = abs(res[j]) <= small(lam[j])
    while i < j:
        if ti:
            res[i] = 0.0
            i += 1
            ti = (abs(res[i]) <= small(lam[i])) if i <= j else False
        else:
            if tj:
                lam[i], lam[j] = lam[j], lam[i]
                res[j] = res[i]
                res[i] = 0.0
                tmpc = yy[:jnd, i].copy()
                yy[:jnd, i] = yy[:jnd, j]
                yy[:jnd, j] = tmpc
                i += 1
                ti = (abs(res[i]) <= small(lam[i])) if i <= j else False
            j -= 1
            tj = (abs(res[j]) <= small(lam[j])) if j > i else False
    return i + 1 if ti else i


def ritz_vectors(st, lck, ny, yy, jl):
    """trl_ritz_vectors (SRC/trlcore.c:2761-2910):
    basis[:, lck:lck+ny] <- basis[:, lck:lck+jl] @ yy[:jl, :ny] (in place).
    Logical basis column k is column k of the single n x (mev+nbas) array."""
    if jl <= 0 or ny <= 0:
        return
    out = st.V[:, lck:lck + jl] @ yy[:jl, :ny]
    st.V[:, lck:lck + ny] = out


def trlan(matvec, info, nrow, mev, start, seed=12345, complex_=False, gdot_quirk=True, lwrk=0):
    """trlan (SRC/trlan.c:21-301) + trlanczos (SRC/trlcore.c:723-1532), or
    with complex_=True ztrlan (SRC/ztrlan.c:26-221) + ztrlanczos
    (SRC/ztrlcore.c:436-1286).

    matvec(x) -> A x (numpy).  start: the start vector (used when
    info.guess == 1).  gdot_quirk (complex only) reproduces the reference's
    two-column ztrl_g_dot_ branches; False gives the exact V^H r.  lwrk:
    the caller's workspace length (trlan's lwrk; ztrlan's nmis), which
    bounds the perturbations of a guess <= 0 start.
    Returns dict(eval, evec, wrk, info).
    """
    rng = np.random.default_rng(seed)
    out_eval = np.zeros(mev)
    if info.ned > mev:
        info.ned = mev
    if info.ned <= info.nec or info.ned <= 0:
        return dict(eval=out_eval, evec=np.zeros((nrow, mev)), wrk=np.zeros(mev), info=info)
    info.stat = 0
    clear_counter(info, nrow, mev, nrow)
    if info.stat != 0:
        return dict(eval=out_eval, evec=np.zeros((nrow, mev)), wrk=np.zeros(mev), info=info)
    st = _State(nrow, mev, info.maxlan, rng, complex_, gdot_quirk)
    st.guess_cap = guess_cap(info, nrow, mev, lwrk, complex_)
    wrk_out = _trlanczos(matvec, info, st, mev, out_eval, start)
    evec = st.V[:, :mev].copy()
    return dict(eval=out_eval, evec=evec, wrk=wrk_out, info=info)


def _trlanczos(matvec, info, st, mev, eval_, start):
    V = st.V
    alpha, beta, alfrot, betrot = st.alpha, st.beta, st.alfrot, st.betrot
    maxlan = info.maxlan
    nbas = st.nbas
    locked = info.nec
    if locked > 0:
        alpha[:locked] = eval_[:locked]
        beta[:locked] = 0.0
    j1, j2 = initial_guess(info, st, start)
    jnd = j1 + j2
    kept = jnd
    wrk_out = np.zeros(mev)
    if info.stat != 0:
        return wrk_out
    i1 = info.ned - jnd
    next_test = i1 + min(i1, min(6, info.ned // 2))
    degen = -1
    rot = None
    lam = res = None
    qa = qb = rr = 0
    j1n = j2n = 0
    i2 = 0
    while info.matvec < info.maxmv and (degen != -1 or info.nec < info.ned):
        jnd += 1
        if j1 < mev:
            j1 += 1
            qa = j1 - 1
        else:
            j2 += 1
            qa = mev + j2 - 1
        if j1 < mev:
            j1n = j1 + 1
            j2n = 0
            rr = j1n - 1
        else:
            j1n = mev
            j2n = j2 + 1
            rr = mev + j2n - 1
        V[:, rr] = matvec(V[:, qa])
        info.matvec += 1
        alpha[jnd - 1] = float(np.vdot(V[:, qa], V[:, rr]).real)
        beta[jnd - 1] = alpha[jnd - 1]
        if j1 > 2:
            V[:, rr] -= V[:, 0:j1] @ beta[0:j1]
        elif j1 == 1:
            V[:, rr] -= alpha[0] * V[:, qa]
        elif j1 == 2:
            V[:, rr] -= beta[0] * V[:, 0]
            V[:, rr] -= beta[1] * V[:, 1]
        if j2 > 2:
            V[:, rr] -= V[:, mev:mev + j2] @ beta[j1:j1 + j2]
        elif j2 == 1:
            V[:, rr] -= beta[jnd - 1] * V[:, qa]
        elif j2 == 2:
            V[:, rr] -= beta[j1] * V[:, mev]
            V[:, rr] -= beta[jnd - 1] * V[:, mev + 1]
        orth(info, st, j1, j2, rr, kept)
        if info.stat != 0:
            break
        alfrot[:locked] = alpha[:locked]
        betrot[:locked] = 0.0
        i1 = jnd - locked
        rot, ar, br, ierr = tridiag(i1, alpha[locked:jnd], beta[locked:jnd])
        info.stat = ierr
        if ierr != 0:
            break
        alfrot[locked:jnd] = ar
        betrot[locked:jnd] = br
        betrot[jnd - 1] = beta[jnd - 1]
        while jnd < info.klan and (degen != -1 or info.nec < info.ned):
            qb = qa
            qa = rr
            j1 = j1n
            j2 = j2n
            jnd += 1
            if j1n < mev:
                j1n += 1
                rr = j1n - 1
            else:
                j2n += 1
                if j2n <= nbas:
                    rr = mev + j2n - 1
                else:
                    info.stat = -1111
                    break
            V[:, rr] = matvec(V[:, qa])
            info.matvec += 1
            alpha[jnd - 1] = float(np.vdot(V[:, qa], V[:, rr]).real)
            V[:, rr] -= alpha[jnd - 1] * V[:, qa]
            V[:, rr] -= beta[jnd - 2] * V[:, qb]
            orth(info, st, j1, j2, rr, kept)
            alfrot[jnd - 1] = alpha[jnd - 1]
            betrot[jnd - 1] = beta[jnd - 1]
            if info.stat != 0:
                break
            if info.matvec >= next_test:
                lam, res, ierr = get_eval(jnd, locked, alfrot, betrot)
                info.stat = ierr
                if ierr != 0:
                    break
                k = min(mev, jnd)
                eval_[:k] = lam[:k]
                convergence_test(jnd, lam, res, info)
                degen = -1
                if (degen != -1 or info.nec < info.ned) and info.nec > 0:
                    next_test = int(float(info.ned * info.matvec) / float(info.nec))
                elif info.nec == 0:
                    next_test = next_test + next_test
                    if info.maxlan == info.ntot:
                        next_test = int(math.ceil(0.5 * (info.maxlan + info.matvec)))
        if info.stat != 0:
            break
        # ---- restart block (trlcore.c:1245-1494)
        prek = kept
        jml = jnd - locked
        i2 = kept - locked + 1
        if degen != -1 or info.nec < info.ned:
            lam, res, ierr = get_eval(jnd, locked, alfrot, betrot)
            info.stat = ierr
            if ierr != 0:
                break
            ctw = convergence_test(jnd, lam, res, info)
            degen = -1
        if (degen != -1 or info.nec < info.ned) and info.matvec < info.maxmv:
            # The reference's workspace around lambda and res when it
            # restarts (trlcore.c:868, 1200-1201, 1256, 1283): betrot's last
            # element, lambda[0:jnd), res[0:jnd), then the convergence
            # test's |res| and lambda.  The restart schemes read one element
            # beyond either end of their windows, so they get windows of
            # this one array with C indexing (Strip).
            w = np.concatenate([[betrot[info.maxlan - 1]], lam[:jnd], res[:jnd], ctw])
            lam_s = Strip(w, 1 + locked)
            res_s = Strip(w, 1 + jnd + locked)
            kept = shuffle_eig(jml, info.klan - locked, lam_s, res_s, info, kept, locked)
            lam[locked:jnd] = w[1 + locked:1 + jnd]
            res[locked:jnd] = w[1 + jnd + locked:1 + 2 * jnd]
            yy = None
            if kept * 3 < jml:
                yy, ierr = get_tvec(jml, alfrot[locked:], betrot[locked:], 0, i2, rot,
                                    kept, lam[locked:], st.iwrk)
                info.stat = ierr
                if ierr == 0 and locked + kept > 0:
                    alpha[:locked + kept] = lam[:locked + kept]
            if kept * 3 >= jml or info.stat != 0:
                if locked + kept > 0:
                    alfrot[:locked + kept] = lam[:locked + kept]
                a_sub = alpha[locked:locked + jml].copy()
                yy, ierr = get_tvec_a(jml, prek - locked, a_sub, beta[locked:], kept,
                                      alfrot[locked:], st.iwrk)
                alpha[locked:locked + jml] = a_sub
                info.stat = ierr
            if info.stat != 0:
                break
            for i in range(kept):
                beta[locked + i] = yy[jml - 1, i] * betrot[jnd - 1]
            if jml > info.ned + (info.ned // 5 + 6):
                a_sub = alpha[locked:locked + kept].copy()
                b_sub = beta[locked:locked + kept].copy()
                i2 = set_locking(jml, kept, a_sub, b_sub, yy, info.anrm)
                alpha[locked:locked + kept] = a_sub
                beta[locked:locked + kept] = b_sub
            else:
                i2 = 0
            ritz_vectors(st, locked, kept, yy, jml)
            kept += locked
            locked += i2
            info.locked = locked
            jnd = kept
            if jnd <= mev:
                j1 = jnd
                j2 = 0
            else:
                j1 = mev
                j2 = jnd - mev
                if j2 >= nbas - 1:
                    info.stat = -1111
                    break
            if info.nec > 0:
                next_test = int(float(info.matvec * info.ned) / float(info.nec))
            else:
                next_test = next_test + info.maxlan
            k = min(mev, jnd)
            if k > 0:
                eval_[:k] = lam[:k]
            if jnd < mev:
                j1n = j1 + 1
                j2n = 0
                V[:, j1n - 1] = V[:, rr]
                rr = j1n - 1
            else:
                j1n = mev
                j2n = j2 + 1
                V[:, mev + j2n - 1] = V[:, rr]
                rr = mev + j2n - 1
        else:
            kept = min(info.nec, max(info.ned, mev - 1))
            info.nec = kept
            if kept == 0:
                kept = min(mev - 1, info.ned)
            sort_eig(jnd, info.lohi, kept, info.ref, lam, res)
            eval_[:kept] = lam[:kept]
            yy = None
            if kept * 3 < jnd:
                yy, ierr = get_tvec(jnd, alfrot, betrot, locked, i2, rot, kept, eval_,
                                    st.iwrk)
                info.stat = ierr
            if kept * 3 >= jnd or info.stat != 0:
                a_sub = alpha[:jnd].copy()
                ev = eval_.copy()
                yy, ierr = get_tvec_a(jnd, prek, a_sub, beta, kept, ev, st.iwrk)
                alpha[:jnd] = a_sub
                info.stat = ierr
                if ierr != 0:
                    break
            if kept > 0:
                alpha[:kept] = eval_[:kept]
            for i in range(kept):
                beta[i] = betrot[jnd - 1] * yy[jnd - 1, i]
            ritz_vectors(st, 0, kept, yy, jnd)
            jnd = kept
            j1 = kept
            j2 = 0
            if j1 < mev:
                V[:, j1] = V[:, rr]
            else:
                V[:, mev] = V[:, rr]
        info.nloop += 1
    for i in range(min(j1, mev)):
        wrk_out[i] = abs(beta[i])
    return wrk_out
