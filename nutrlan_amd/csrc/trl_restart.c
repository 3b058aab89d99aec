/*
 * trl_restart.c -- which Ritz pairs survive a thick restart.
 *
 * Restatement of the reference's restart policy (SRC/restart.c): the kept
 * set is [0..kl] from the left and [kr..nd) from the right of the
 * ascending Ritz values; schemes 0-8 choose (kl, kr).  Matching matvec
 * counts with the reference needs the same choices, so the decision logic
 * follows the reference scheme by scheme (cited per function); an integer
 * tag array rides along with lambda/res so the driver knows which Ritz
 * vector each kept value belongs to.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>

#include "trl_kit.h"

#define RMIN(a, b) ((a) < (b) ? (a) : (b))
#define RMAX(a, b) ((a) > (b) ? (a) : (b))

static int round_half(double x) { return (int)(x + 0.5); }

/* restart.c:26-91 */
static void search_range(int nd, const double *lam, const double *res, trl_info *info,
                         int ncl, int ncr, int *lohi, int tind, int *klm, int *krm)
{
    int lo = RMAX(ncl, 0), hi = RMIN(ncr, nd - 1);
    const double bnd = info->tol * info->anrm;
    *lohi = info->lohi;
    if (info->lohi > 0) {
        hi = RMIN(RMAX(info->maxlan - info->ned, (info->maxlan + info->nec) / 2) - 1,
                  RMIN(hi, tind - 1));
        while (hi + hi > ncl + ncr && res[hi] < bnd) --hi;
    } else if (info->lohi < 0) {
        lo = RMAX(RMIN(info->ned, (info->maxlan + info->nec) / 2) - 1, RMAX(tind + 1, lo));
        while (lo + lo < ncl + ncr && res[lo] < bnd) ++lo;
    } else {
        if (tind - lo < hi - tind) {
            *lohi = -1;
            lo = tind + 1;
        } else {
            *lohi = 1;
            hi = tind - 1;
        }
        int j = info->locked + lo + nd - hi + 1;
        if (j > 0) {
            j /= 2;
            lo = RMAX(0, lo - j);
            hi = RMIN(hi + j, nd - 1);
        }
    }
    *klm = lo;
    *krm = hi;
}

/* restart.c:94-117 */
static double min_gap_ratio(const trl_info *info, int tind, const double *res)
{
    double gamma = info->maxmv * (info->nec + 1.0) / info->ned - info->matvec;
    if (gamma < info->klan)
        gamma = RMAX((info->maxmv - info->matvec) / ((double)(info->ned - info->nec)), 2.0);
    return RMIN(log(res[tind] / (info->tol * info->anrm)) / gamma, 0.5);
}

static double gap_ratio(int i, int j, int tind, const double *lam)
{
    return (lam[i] - lam[tind]) / (lam[j] - lam[tind]);
}

static int extra_fixed(int mnd, const trl_info *info)
{
    int extra = round_half((mnd - info->nec) * (0.4 + 0.1 * info->ned / (double)mnd));
    if (extra > info->ned + info->ned && extra > 5) {
        const double g = (double)mnd / (double)info->ned;
        extra = round_half((extra + (log(g) * info->ned) * g) / (1.0 + g));
    }
    return extra;
}

/* Balance a two-sided kept window toward the side with smaller residuals
 * (shared by schemes 1 and 6, restart.c:628-650 and 772-796). */
static void balance_window(const double *res, int kl0, int kr0, int *kl, int *kr)
{
    int kli = *kl, kri = *kr, stop = 0;
    while (kli > kl0 && kri < kr0 && !stop) {
        if (10.0 * res[kli] < res[kri]) {
            if (res[kli + 1] < res[kri + 1]) {
                ++kli;
                ++kri;
            } else stop = 1;
        } else if (10.0 * res[kri] < res[kli]) {
            if (res[kri - 1] < res[kli - 1]) {
                --kri;
                --kli;
            } else stop = 1;
        } else stop = 1;
    }
    *kl = kli;
    *kr = kri;
}

/* Scheme 1 (and first restart of 3/4/5/7/8): restart.c:560-689. */
static void fixed_extra(int nd, int mnd, int tind, const double *lam, const double *res,
                        trl_info *info, int *kl, int *kr)
{
    int kli = *kl, kri = *kr;
    const int kl0 = kli, kr0 = kri;
    int extra = extra_fixed(mnd, info);
    const int minsep = RMAX(3, RMAX(nd / 5, nd - 4 * info->ned));
    const double gmin = min_gap_ratio(info, tind, res);
    if (info->lohi > 0) {
        kri = RMIN(tind - 1, kri) - extra;
        kli = -1;
    } else if (info->lohi < 0) {
        kli = RMAX(tind + 1, kli) + extra;
        kri = nd;
    } else {
        ++extra;
        kli += extra / 2;
        kri -= extra / 2;
        balance_window(res, kl0, kr0, &kli, &kri);
    }
    while (kli + minsep < kri &&
           gap_ratio(RMAX(0, kli), RMIN(kri, nd - 1), tind, lam) < gmin) {
        if (info->lohi < 0) ++kli;
        else if (info->lohi > 0) --kri;
        else if (res[kli] < res[kri]) ++kli;
        else ++kri;
    }
    if (info->lohi > 0) {
        int i = kri - 1;
        while (i > kli + minsep && lam[kri] - res[kri] < lam[i] &&
               lam[kri] + res[kri] < lam[i] + res[i])
            --i;
        kri = i + 1;
    } else {
        int i = kli + 1;
        while (i < kri - minsep && lam[kli] + res[kli] > lam[i] &&
               lam[kli] - res[kli] > lam[i] - res[i])
            ++i;
        kli = i - 1;
    }
    *kl = kli;
    *kr = kri;
}

/* Scheme 6: restart.c:692-804. */
static void scan_thickness(int nd, const double *res, const trl_info *info, int kept,
                           int *kl, int *kr)
{
    int kli = *kl, kri = *kr;
    if (info->lohi < 0) {
        kri = nd;
        kli = kept + RMIN(RMAX(info->nec, 1), (nd - kept) / 2) - 1;
        if (kli <= 0) {
            if (nd > 6) kli = nd / 2 - 1;
            else if (nd > 2) kli = 1;
        } else if (kli + 3 > nd) {
            kli = info->nec + RMIN(info->ned, RMIN(10, (nd - info->ned) / 2)) - 1;
        }
    } else if (info->lohi > 0) {
        kli = -1;
        kri = kept + RMIN(RMAX(info->nec, 1), (nd - kept) / 2) - 1;
        if (kri <= 0) {
            if (nd > 6) kri = nd / 2 - 1;
            else if (nd > 2) kri = 1;
        } else if (kri + 4 > nd) {
            kri = info->nec + RMIN(info->ned, RMIN(10, (nd - info->ned) / 2));
        }
        kri = nd - kri - 1;
    } else {
        const int kl0 = kli, kr0 = kri;
        int extra = kept + RMIN(info->nec, (nd - kept) / 2) + 1;
        if (extra <= 0) {
            if (nd > 6) extra = nd / 2;
            else if (nd > 2) extra = 2;
        } else if (extra + 3 > nd) {
            extra = info->nec + RMIN(info->ned, RMIN(10, (nd - info->ned) / 2));
        }
        kli = RMAX(kli, (extra / 2) - 1);
        kri = RMIN(kri, (nd - extra / 2) - 1);
        balance_window(res, kl0, kr0, &kli, &kri);
    }
    *kl = kli;
    *kr = kri;
}

/* Scheme 2: restart.c:840-988. */
static void small_residuals(int nd, int mnd, int tind, const double *lam, const double *res,
                            trl_info *info, int *kl, int *kr)
{
    const int minsep = RMAX(3, RMAX(nd / 5, nd - 4 * info->ned));
    const int extra = extra_fixed(mnd, info);
    int kli = *kl, kri = *kr;
    int kl0 = kli, kr0 = kri;
    double resmax = 0.0;
    if (nd > 0) {
        resmax = res[0];
        for (int i = 1; i < nd; ++i) resmax = RMAX(resmax, res[i]);
    }
    double acpt = resmax / res[tind];
    if (info->lohi > 0) {
        if (acpt < 0.999 && acpt >= 0.0) {
            const int ii = tind - 1;
            acpt = RMAX(sqrt(acpt) * res[tind], res[ii] + res[ii]);
            acpt = RMIN(acpt, resmax);
            kri = ii - 1;
            while (res[kri] < acpt && kri > kli + 3) --kri;
        } else {
            kri = kr0 - extra;
        }
        kri = RMAX(2, kri);
        kli = RMIN(kli, kri - 2);
    } else if (info->lohi < 0) {
        if (acpt < 0.999 && acpt >= 0.0) {
            const int ii = tind + 1;
            acpt = RMAX(sqrt(acpt) * res[tind], res[ii] + res[ii]);
            acpt = RMIN(acpt, resmax);
            kli = ii + 1;
            while (res[kli] < acpt && kli < kri - 3) ++kli;
        } else {
            kli = kl0 + extra;
        }
        kli = RMIN(nd - 4, kli);
        kri = RMAX(kri, kli + 2);
    } else {
        int i = kli + 1, j = kri - 1;
        for (int t = 1; t <= extra; ++t) {
            if (res[i] < res[j]) kli = i++;
            else if (res[i] > res[j]) kri = j--;
            else if (i < nd - j) kli = i++;
            else kri = j--;
        }
    }
    kl0 = kli;
    kr0 = kri;
    const double gmin = min_gap_ratio(info, tind, res);
    int done = 0;
    while (kli + minsep < kri &&
           gap_ratio(RMAX(0, kli), RMIN(nd - 1, kri), tind, lam) < gmin && !done) {
        if (info->lohi < 0) ++kli;
        else if (info->lohi > 0) --kri;
        else if (res[kli] < res[kri]) ++kli;
        else if (kri < nd - 1) ++kri;
        else done = 1;
    }
    if (info->lohi > 0) {
        int i = kr0 - 1;
        while (i > kli + minsep && lam[kri] - res[kri] < lam[i] &&
               lam[kri] + res[kri] < lam[i] + res[i])
            --i;
        kri = RMIN(kri, i + 1);
    } else {
        int i = kl0 + 1;
        while (i < kri - minsep && lam[kli] + res[kli] > lam[i] &&
               lam[kli] - res[kli] > lam[i] - res[i])
            ++i;
        kli = RMAX(kli, i - 1);
    }
    *kl = kli;
    *kr = kri;
}

/* Scheme 3: restart.c:991-1094 (reads operator vs. orthogonalization
 * time, so its choices depend on the clock). */
static void max_gap_ratio(int nd, int tind, int kept, const double *lam, const double *res,
                          trl_info *info, int *kl, int *kr)
{
#define GRATIO(i, j) (lam[(j)] == info->trgt ? DBL_MAX : (lam[(i)] - info->trgt) / (lam[(j)] - info->trgt))
    int lohi, klm, krm;
    search_range(nd, lam, res, info, *kl, *kr, &lohi, tind, &klm, &krm);
    int kli = klm, kri = krm;
    int igap = RMAX(RMIN(nd - info->ned, round_half((krm - klm) * 0.4)), 2);
    if (igap > 2 && info->matvec > info->maxlan) {
        if (info->clk_op + info->tick_o >
            10.0 * (info->clk_orth + info->tick_h + info->clk_res + info->tick_r)) {
            igap = RMAX(2, nd - kept - 1);
        } else if (info->crat < min_gap_ratio(info, tind, res)) {
            igap = RMAX(2, nd - kept - 1);
        }
    }
    if (kli + igap > kri) {
        *kl = kli;
        *kr = kri;
        return;
    }
    double bnd = lohi > 0 ? GRATIO(kri, kli) : GRATIO(kli, kri);
    for (int i = klm; i <= krm - igap; ++i) {
        const int j = i + igap;
        const double t = lohi > 0 ? GRATIO(j, i) : GRATIO(i, j);
        if (t > bnd) {
            kli = i;
            kri = j;
            bnd = t;
        }
    }
    *kl = kli;
    *kr = kri;
#undef GRATIO
}

/* Scheme 4: restart.c:1097-1185. */
static void max_progress(int nd, int tind, int kept, const double *lam, const double *res,
                         trl_info *info, int *kl, int *kr)
{
#define MERIT(i, j) ((lam[(i)] - info->trgt) * abs((j) - (i)) / (lam[(j)] - info->trgt))
    int lohi, klm, krm;
    search_range(nd, lam, res, info, *kl, *kr, &lohi, tind, &klm, &krm);
    int kli = klm, kri = krm;
    int igap = RMAX(RMIN(nd - info->ned, round_half((kri - kli) * 0.4)), 2);
    if (igap > 2 && igap + kept > nd && info->crat > 0.0) {
        if (min_gap_ratio(info, tind, res) > info->crat) igap = RMAX(2, nd - kept - 1);
    }
    double ss = lohi > 0 ? MERIT(kri, kli) : MERIT(kli, kri);
    for (int i = klm; i <= krm - igap; ++i)
        for (int j = i + igap; j <= krm; ++j) {
            const double t = lohi > 0 ? MERIT(j, i) : MERIT(i, j);
            if (t > ss) {
                ss = t;
                kli = i;
                kri = j;
            }
        }
    *kl = kli;
    *kr = kri;
#undef MERIT
}

/* Dynamic minimum-gap fraction shared by schemes 5 and 7
 * (restart.c:1318-1343 and 199-226); updates info->avgm. */
static int dynamic_gap(trl_info *info, int span)
{
    const double def1 = 1.0, def2 = 0.7, up = 1.0, dw = 0.7;
    info->avgm = (info->avgm * ((double)info->nloop - 1) + (double)(info->klan - info->k + 1)) /
                 ((double)info->nloop);
    const double z1 = info->ptres / (info->tol * info->anrm);
    const double z2 = 1.0 / info->cfac;
    if (z2 < 1.0) return (int)(span * def2);
    if (z1 < 1.0) return (int)(span * def1);
    const double t3 = log(z1 + sqrt(z1 - 1.0) * sqrt(z1 + 1.0)) / (2.0 * info->avgm);
    const double t2 = log(z2 + sqrt(z2 - 1.0) * sqrt(z2 + 1.0)) / (info->klan - info->k + 1);
    double t = t2 / t3;
    t = pow(t, 2.0);
    t = atan(t) * (2.0 / 3.141592654);
    t = dw + (up - dw) * t;
    return (int)(span * t);
}

/* Scheme 5: restart.c:1280-1408. */
static void max_reduction(int nd, int tind, int kept, const double *lam, const double *res,
                          trl_info *info, int *kl, int *kr)
{
#define MERIT5(i, j) (sqrt((lam[(i)] - info->trgt) / (lam[(j)] - lam[(i)])) * abs((j) - (i)))
    int lohi, klm, krm;
    search_range(nd, lam, res, info, *kl, *kr, &lohi, tind, &klm, &krm);
    int kli = klm, kri = krm;
    int igap = dynamic_gap(info, abs(krm - klm));
    if (igap > 2 && igap + kept > nd && info->crat > 0.0) {
        if (min_gap_ratio(info, tind, res) > info->crat) igap = RMAX(2, nd - kept - 1);
    }
    double ss = lohi > 0 ? MERIT5(kri, kli) : MERIT5(kli, kri);
    for (int i = klm; i <= krm - igap; ++i)
        for (int j = i + igap; j <= krm; ++j) {
            const double t = lohi > 0 ? MERIT5(j, i) : MERIT5(i, j);
            if (t > ss) {
                ss = t;
                kli = i;
                kri = j;
            }
        }
    *kl = kli;
    *kr = kri;
#undef MERIT5
}

/* Scheme 7 (default): restart.c:121-299.  Also chooses the next basis
 * size info->klan; writes the two-sided choice into info->lohi like the
 * reference does. */
static void max_gap_cost_ratio(int n, int tind, trl_info *info, const double *lam,
                               const double *res, int *kl, int *kr)
{
    const int nd = info->ned, mn = info->maxlan;
    int k1, k2;
    search_range(n, lam, res, info, *kl, *kr, &info->lohi, tind, &k1, &k2);
    const int t = dynamic_gap(info, abs(k2 - k1));
    double min_val = 0.0, min_gamma = 0.0, min_gamma0 = 0.0;
    int min_k1 = k1, min_k2 = k2, min_m = info->klan, min_l = 0;
    for (int i = k1; i <= k2 - t; ++i)
        for (int j = i + t; j <= k2; ++j) {
            const double gamma0 = sqrt((lam[i + 1] - lam[tind]) / (lam[j - 1] - lam[i + 1]));
            const int l1 = (i + 1) + (n - j);
            const int l2 = l1 + info->locked;
            int mn2 = l2 + 1;
            if (mn2 > mn) mn2 = mn;
            if (mn2 < nd) mn2 = nd;
            for (int k = mn2; k <= mn; ++k) {
                const double gamma = (k - l2) * gamma0;
                const double val = ((k - l2) * (k + l2 - 1) + k * l1) / gamma;
                if (min_val == 0.0 || val < min_val) {
                    min_k1 = i;
                    min_k2 = j;
                    min_m = k;
                    min_val = val;
                    min_gamma = gamma;
                    min_gamma0 = gamma0;
                    min_l = l2;
                }
            }
        }
    *kl = min_k1;
    *kr = min_k2;
    info->klan = min_m;
    info->k1 = min_k1;
    info->k2 = min_k2;
    info->k = min_l;
    info->predicted_crate = exp(-min_gamma);
    info->mgamma = min_gamma;
    info->gamma0 = min_gamma0;
    info->old_target = lam[tind];
    info->target_id = tind;
    info->old_locked = info->locked;
}

/* Scheme 8: restart.c:302-432. */
static void max_gap_cost_ratio_static(int n, int tind, trl_info *info, const double *lam,
                                      const double *res, int *kl, int *kr)
{
    const int nd = info->ned, mn = info->maxlan;
    int k1, k2, t;
    search_range(n, lam, res, info, *kl, *kr, &info->lohi, tind, &k1, &k2);
    if (info->crat > 0.0) {
        const double t2 = -log(info->crat);
        const double t3 = -(t2 * (info->maxlan)) / log(info->tol * info->anrm);
        if (t3 >= 0.5 || info->klan < info->maxlan) {
            t = (int)(abs(k2 - k1) * 0.8);
        } else {
            double g = t2 / min_gap_ratio(info, tind, res);
            g = pow(g, 0.25);
            g = atan(g) * (4.0 / 3.141592654);
            t = (int)(abs(k2 - k1) * g);
        }
    } else {
        t = (int)(abs(k2 - k1) * 0.8);
    }
    const int mn1 = round_half(2.0 * (info->klan) / 5.0);
    double min_val = 0.0;
    int min_k1 = k1, min_k2 = k2, min_m = info->klan;
    for (int i = k1; i <= k2 - t; ++i)
        for (int j = i + t; j <= k2; ++j) {
            const double gamma0 = sqrt((lam[i + 1] - lam[tind]) / (lam[j - 1] - lam[i + 1]));
            const int l1 = (i + 1) + (n - j);
            const int l2 = l1 + info->locked;
            int mn2 = mn1;
            if (mn1 < info->rfact * l2) mn2 = (int)(info->rfact * l2);
            if (mn2 > mn) mn2 = mn;
            if (mn2 < nd) mn2 = nd;
            const double gamma = (mn2 - l2 + 1) * gamma0;
            const double val = (mn2 * mn2 - l2 * l2 + 2 * mn2 * l1) / gamma;
            if (min_val == 0.0 || val < min_val) {
                min_k1 = i;
                min_k2 = j;
                min_m = mn2;
                min_val = val;
            }
        }
    *kl = min_k1;
    *kr = min_k2;
    info->klan = min_m;
}

/* The reference's trl_shuffle_eig (restart.c:1410-1651).  lambda/res hold
 * the nd non-locked Ritz values (ascending) and residuals; lambda[nd] must
 * be readable (the driver places res[0] there, as the reference's
 * workspace layout does).  The kept pairs are compacted to the front,
 * carrying tag.  Returns the number kept. */
int trlkit_shuffle_eig(int nd, int mnd, double *lambda, double *res, int *tag,
                       trl_info *info, int kept)
{
    if (nd <= 5) {
        trlkit_sort2(nd, TRLKIT_BY_VALUE, 0.0, res, lambda, tag);
        const int k = nd > 3 ? 2 : (nd > 0 ? 1 : 0);
        if (k > 1) trlkit_sort2(k, TRLKIT_BY_VALUE, 0.0, lambda, res, tag);
        return k;
    }
    const double bnd = RMIN(info->tol, DBL_EPSILON * info->anrm);
    int ncr = 0, ncl = nd - 1;
    for (int i = nd - 1; i >= 0; --i)
        if (!(res[i] <= bnd)) {
            ncr = i + 1;
            break;
        }
    for (int i = 0; i < nd; ++i)
        if (!(res[i] <= bnd)) {
            ncl = i - 1;
            break;
        }
    int kl = ncl, kr = ncr, tind;
    if (ncr <= ncl) return nd; /* everything converged: keep all */
    /* bisection for the target Ritz value */
    tind = (kl + kr) / 2;
    while (lambda[tind] != info->trgt && kr > kl) {
        if (lambda[tind] < info->trgt) {
            kl = tind + 1;
            tind = (kl + kr) / 2;
        } else if (lambda[tind] > info->trgt) {
            kr = tind - 1;
            tind = (kl + kr) / 2;
        } else {
            kl = tind;
            kr = tind;
        }
    }
    if (lambda[tind] == info->trgt) {
        kl = tind - 1;
        while (kl >= 0 && lambda[kl] == info->trgt) --kl;
        kr = tind + 1;
        while (kr < nd && lambda[kr] == info->trgt) ++kr;
    } else {
        kl = tind - 1;
        kr = tind + 1;
    }
    if (info->lohi > 0) {
        kr = kl;
        kl = RMIN(ncl, RMAX(0, nd - info->ned) - 1);
    } else if (info->lohi < 0) {
        kl = kr;
        kr = RMAX(ncr, RMIN(nd - info->nec, info->ned + 1) - 1);
    } else if (ncr - tind > tind - ncl) {
        kl = kr;
        kr = RMAX(ncr, RMIN(nd - info->nec, info->ned + 1) - 1);
    } else {
        kr = kl;
        kl = RMIN(ncl, RMAX(0, nd - info->ned) - 1);
    }
    switch (info->restart) {
    case 0:
    case 1:
        fixed_extra(nd, mnd, tind, lambda, res, info, &kl, &kr);
        break;
    case 2:
        small_residuals(nd, mnd, tind, lambda, res, info, &kl, &kr);
        break;
    case 3:
        if (info->nloop > 0) max_gap_ratio(nd, tind, kept, lambda, res, info, &kl, &kr);
        else fixed_extra(nd, mnd, tind, lambda, res, info, &kl, &kr);
        break;
    case 4:
        if (info->nloop > 0) max_progress(nd, tind, kept, lambda, res, info, &kl, &kr);
        else fixed_extra(nd, mnd, tind, lambda, res, info, &kl, &kr);
        break;
    case 5:
        if (info->nloop > 0) max_reduction(nd, tind, kept, lambda, res, info, &kl, &kr);
        else fixed_extra(nd, mnd, tind, lambda, res, info, &kl, &kr);
        break;
    case 6:
        scan_thickness(nd, res, info, kept, &kl, &kr);
        break;
    case 8:
        if (info->nloop > 0) max_gap_cost_ratio_static(nd, tind, info, lambda, res, &kl, &kr);
        else fixed_extra(nd, mnd, tind, lambda, res, info, &kl, &kr);
        break;
    default: /* 7 and anything unknown */
        if (info->nloop > 0) max_gap_cost_ratio(nd, tind, info, lambda, res, &kl, &kr);
        else fixed_extra(nd, mnd, tind, lambda, res, info, &kl, &kr);
        break;
    }
    /* keep a minimal separation between the two ends (restart.c:1623-1638) */
    const int minsep = RMAX(3, RMAX(nd / 6, nd - 6 * info->ned));
    if (kr <= kl + minsep || kl + nd - kr + minsep > mnd) {
        if (ncl < kl && kl < kr && kr < ncr) {
            --kl;
            ++kr;
        } else if (info->lohi > 0) {
            kr = RMAX(minsep, RMIN(nd / 3, ncr)) - 1;
            kl = -1;
        } else if (info->lohi < 0) {
            kl = RMIN(nd - minsep, RMAX((nd + nd) / 3, ncl + 2)) - 1;
            kr = nd;
        } else {
            kl = (nd - minsep) / 2 - 2;
            kr = (nd - minsep + 1) / 2;
        }
    }
    /* compact [kr, nd) to follow [0, kl] */
    --kr;
    for (int i = 1; i < nd - kr; ++i) {
        lambda[kl + i] = lambda[kr + i];
        res[kl + i] = res[kr + i];
        if (tag) tag[kl + i] = tag[kr + i];
    }
    kept = kl + RMAX(1, nd - kr);
    if (info->lohi < -1) trlkit_sort2(kept, TRLKIT_BY_VALUE, 0.0, lambda, res, tag);
    return kept;
}
