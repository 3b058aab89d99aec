/*
 * trlan.c -- host driver of the MI355X thick-restart Lanczos solver.
 *
 * trlan() keeps the reference's contract (SRC/trlan.c:21-301) and makes the
 * same decisions as the reference iteration (SRC/trlcore.c:723-1532) so the
 * two produce the same matvec/restart counts, but its state is split
 * differently: the Lanczos basis, the residual vector and every O(n)
 * operation live on the GPU (trlhip.h); the host keeps only the projected
 * problem (alpha, beta, the arrowhead/tridiagonal and its eigenvectors,
 * at most maxlan^2 doubles) and the scalars read back from device
 * reductions.  No MPI: the reference's trl_g_sum/trl_g_dot_ reductions are
 * the device's block/last-block reductions.
 *
 * Basis columns: logical column k is device column k (the reference splits
 * them between evec and base, SRC/trlcore.c:883-910).
 */
#define _GNU_SOURCE
#include <float.h>
#include <stdio.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <pthread.h>
#include <sched.h>
#include <ucontext.h>
#include <unistd.h>

#include "trl_kit.h"
#include "trlan.h"
#include "trlan_batch.h"
#include "trlan_internal.h"
#include "ztrlan.h"
#include "trlhip.h"

#define DMIN(a, b) ((a) < (b) ? (a) : (b))
#define DMAX(a, b) ((a) > (b) ? (a) : (b))

/* ------------------------------------------------------------- helpers */

double trl_rand_uniform(uint64_t *state)
{
    /* xorshift64*: 53 random bits -> [0, 1) */
    uint64_t x = *state;
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    *state = x;
    return (double)((x * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
}

/* Clock ticks (CLOCKS_PER_SEC per second) of wall time for the trl_info
 * timers.  The reference reads clock() (process CPU time); here the work
 * runs on the device, and clock() is a system call on the per-step path,
 * so the timers count monotonic wall time (a vDSO read) in the same unit. */
/* Host-time counters of batched solves (ns; TRLHIP_BATCH_PROFILE=1 only):
 * restarts that continue the iteration, the convergence checks inside a
 * cycle, and the arrowhead reduction at a cycle's start. */
static int g_prof;
static long g_ns_restart, g_ns_check, g_ns_tridiag, g_n_restart;
static long g_ns_alloc, g_ns_ctx; /* job setup: host arrays, device context */

static long lz_ns(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (long)ts.tv_sec * 1000000000L + ts.tv_nsec;
}

static clock_t lz_clock(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (clock_t)((double)ts.tv_sec * CLOCKS_PER_SEC + (double)ts.tv_nsec * (CLOCKS_PER_SEC / 1e9));
}

static void add_ticks(trl_info *info, clock_t *acc, double *spill, clock_t t0)
{
    const clock_t t1 = lz_clock();
    const clock_t d = t1 >= t0 ? t1 - t0 : 0;
    if (*acc + d >= *acc) {
        *acc += d;
    } else {
        *spill += (*acc + (double)d) / (double)info->clk_rate;
        *acc = 0;
    }
}

typedef struct {
    trl_info *info;
    trlhip_ctx *dev;
    const trlhip_csr *A;   /* device operator, or NULL for a host op */
    trl_matvec op;
    ztrl_matvec zop;       /* complex Hermitian operator (ztrlan) */
    trl_uo uo;
    int cplx, E;           /* complex basis?  doubles per element */
    int mixed;             /* FP32 shadow basis for Gram-Schmidt (TRLHIP_R32 operator) */
    int n, mev, nbas, ncols, maxlan;
    double *alpha, *beta, *alfrot, *betrot; /* maxlan each */
    double *rot;                            /* maxlan^2 */
    double *lam, *res;                      /* maxlan + 1 */
    int *tag;                               /* maxlan + 1 */
    double *ctw;                            /* 2 (maxlan + 2): the convergence test's workspace */
    double *shw;                            /* 4 (maxlan + 2) + 1: lambda and res as laid out
                                               for the restart schemes */
    double *yy;                             /* maxlan^2 */
    double *zq;                             /* maxlan^2 scratch */
    double *dsc, *esc;                      /* maxlan scratch */
    int *iwrk;                              /* 4 maxlan, index map leftovers */
    double *hx, *hy;                        /* host staging for host ops */
    const float *dstart;                    /* device FP32 start vector (batch jobs), or NULL */
    int chain;                              /* batched real solve: most steps per round (0: 1) */
    uint64_t rng;
    int rot_n;                              /* order of rot */
} lz_t;

#define DEV(call)                                 \
    do {                                          \
        if ((call) != 0) {                        \
            z->info->stat = TRLHIP_STAT_DEVICE;   \
            return TRLHIP_STAT_DEVICE;            \
        }                                         \
    } while (0)

/* y column <- A * x column, counting time in clk_op. */
static int lz_apply_host_op(lz_t *z, int xcol, int ycol)
{
    DEV(trlhip_col_get(z->dev, xcol, z->hx));
    if (z->cplx)
        z->zop(z->n, 1, (const trl_dcomplex *)z->hx, z->n, (trl_dcomplex *)z->hy, z->n,
               z->info->mvparam);
    else
        z->op(z->n, 1, z->hx, z->n, z->hy, z->n, z->info->mvparam);
    DEV(trlhip_col_set(z->dev, ycol, z->hy));
    return 0;
}

/* Hand the residual to the caller's extra orthogonalization (host). */
static int lz_user_ortho(lz_t *z, int rcol, double *change)
{
    DEV(trlhip_col_get(z->dev, rcol, z->hx));
    z->uo(z->n, z->hx, change);
    DEV(trlhip_col_set(z->dev, rcol, z->hx));
    return 0;
}

static int lz_random_column(lz_t *z, int col)
{
    for (int i = 0; i < z->E * z->n; ++i) z->hx[i] = trl_rand_uniform(&z->rng);
    if (z->cplx) trlkit_smooth_zz(z->n, z->hx);
    else trlkit_smooth_rr(z->n, z->hx);
    DEV(trlhip_col_set(z->dev, col, z->hx));
    return 0;
}

/* Complex retry of the reference's ztrl_cgs (SRC/ztrlcore.c:2117-2178):
 * first a segment perturbation scaled by the residual, later a random
 * segment (or a full random vector). */
static int lz_perturb_complex(lz_t *z, int rcol, int irnd, double rnrm, double old_rnrm)
{
    const int n = z->n;
    double *x = z->hx;
    DEV(trlhip_col_get(z->dev, rcol, x));
    if (irnd == 1 && rnrm > 0.0 && rnrm > DBL_EPSILON * old_rnrm) {
        const int i = (int)(n * trl_rand_uniform(&z->rng));
        const double w = n * sqrt(DBL_EPSILON * old_rnrm / rnrm);
        const int k = i + (int)(w > 1.0 ? w : 1.0);
        for (int j = i; j < k && j < n; ++j) {
            double t = trl_rand_uniform(&z->rng);
            while (fabs(t - 0.5) <= DBL_EPSILON) t = trl_rand_uniform(&z->rng);
            x[2 * j] += rnrm * (t - 0.5);
            x[2 * j + 1] += rnrm * (t - 0.5);
        }
    } else {
        const int i = (int)(n * trl_rand_uniform(&z->rng));
        const int j = (int)(n * trl_rand_uniform(&z->rng));
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        const int a = (i == j) ? 0 : lo, b = (i == j) ? n - 1 : hi;
        for (int k = a; k <= b; ++k) {
            x[2 * k] = trl_rand_uniform(&z->rng);
            x[2 * k + 1] = trl_rand_uniform(&z->rng);
        }
    }
    trlkit_smooth_zz(n, x);
    DEV(trlhip_col_set(z->dev, rcol, x));
    return 0;
}

/* Iterated classical Gram-Schmidt of column rcol against columns
 * [0, nold): the decision logic of the reference's trl_cgs
 * (SRC/trlcore.c:2913-3117) around device passes.  When pass1 is given
 * ({||h||^2, ||s||^2}), the first pass has already been applied into the
 * device scratch column (trlhip_step_cgs) and the loop starts from its
 * result.  Returns 0, -203, -204 or the device status. */
static int lz_cgs_complex(lz_t *z, int nold, int rcol, double *rnrm, int *north,
                          const double *pass1);

/* Make column rcol hold the current vector (it may still be in scratch). */
static int lz_adopt(lz_t *z, int rcol, int *src)
{
    if (*src != rcol) {
        DEV(trlhip_col_scale_copy(z->dev, rcol, *src, 1.0));
        *src = rcol;
    }
    return 0;
}

/* The reference's trl_cgs (SRC/trlcore.c:2913-3117) on the device basis.
 * Mixed precision (FP32 shadow basis): the projections are accurate to
 * FP32 rounding of the basis, so the pass/stop test uses FLT_EPSILON and
 * at most two passes run (CGS2). */
static int lz_cgs(lz_t *z, int nold, int rcol, double *rnrm, int *north, const double *pass1)
{
    const double eps = z->mixed ? FLT_EPSILON : DBL_EPSILON;
    const int maxpass = z->mixed ? 1 : 3;
    int irnd = 0, src = rcol;
    double rr2 = 0.0, hh = 0.0;
    if (z->cplx) return lz_cgs_complex(z, nold, rcol, rnrm, north, pass1);
    if (nold > 0) {
        int cnt = 0;
        while (cnt <= maxpass) {
            if (pass1) {
                hh = pass1[0];
                rr2 = pass1[1];
                src = z->ncols;
                pass1 = NULL;
            } else {
                if (lz_adopt(z, rcol, &src)) return TRLHIP_STAT_DEVICE;
                DEV(trlhip_cgs_pass(z->dev, 0, nold, rcol, &hh, &rr2, NULL));
            }
            if (z->uo) {
                double change = 0.0;
                if (lz_adopt(z, rcol, &src)) return TRLHIP_STAT_DEVICE;
                if (lz_user_ortho(z, rcol, &change)) return TRLHIP_STAT_DEVICE;
                DEV(trlhip_update(z->dev, 0, 0, &change, rcol, &rr2));
                hh += change * change;
            }
            ++*north;
            ++cnt;
            *rnrm = sqrt(rr2);
            if (eps * rr2 > hh) break;
            if ((cnt > 1 && nold * eps * eps * (rr2 + hh) < hh) || !(rr2 > DBL_MIN)) {
                /* no independent direction left: one fresh random vector */
                if (irnd > 0) break;
                cnt = 0;
                ++irnd;
                z->info->nrand++;
                if (lz_random_column(z, rcol)) return TRLHIP_STAT_DEVICE;
                src = rcol;
            }
        }
        if (rr2 < hh && z->uo == NULL) return -203;
    }
    if (*rnrm > DBL_MIN) {
        DEV(trlhip_col_scale_copy(z->dev, rcol, src, 1.0 / *rnrm));
    } else {
        return -204;
    }
    if (irnd > 0) *rnrm = 0.0;
    return 0;
}

/* Complex Hermitian variant: the reference's ztrl_cgs
 * (SRC/ztrlcore.c:1977-2203).  Its randomization test scales with ntot
 * (not nold) and allows up to three retries.  The reference's
 * ztrl_g_dot_ mis-computes V^H r when a basis segment has exactly two
 * columns (SRC/ztrl_comm_none.c:100, 117-119); the device projection here
 * is the exact V^H r. */
static int lz_cgs_complex(lz_t *z, int nold, int rcol, double *rnrm, int *north,
                          const double *pass1)
{
    const double eps = DBL_EPSILON;
    const int maxorth = 3;
    int irnd = 0, ierr = 0, src = rcol;
    double rr2 = 0.0, hh = 0.0;
    if (nold > 0) {
        int cnt = 0;
        while (cnt <= maxorth) {
            if (pass1) {
                hh = pass1[0];
                rr2 = pass1[1];
                src = z->ncols;
                pass1 = NULL;
            } else {
                if (lz_adopt(z, rcol, &src)) return TRLHIP_STAT_DEVICE;
                DEV(trlhip_cgs_pass(z->dev, 0, nold, rcol, &hh, &rr2, NULL));
            }
            ++*north;
            ++cnt;
            *rnrm = sqrt(rr2);
            const double old_rnrm = sqrt(rr2 + hh);
            if (eps * rr2 > hh) break;
            if (((cnt > 1 && !(hh <= z->info->ntot * eps * eps * (rr2 + hh))) || !(rr2 > DBL_MIN)) &&
                irnd < maxorth) {
                cnt = 0;
                ++irnd;
                z->info->nrand++;
                if (lz_adopt(z, rcol, &src)) return TRLHIP_STAT_DEVICE;
                if (lz_perturb_complex(z, rcol, irnd, *rnrm, old_rnrm)) return TRLHIP_STAT_DEVICE;
            }
        }
        if (!(rr2 >= hh)) ierr = -203;
    }
    if (ierr == 0) {
        if (*rnrm > DBL_MIN) {
            DEV(trlhip_col_scale_copy(z->dev, rcol, src, 1.0 / *rnrm));
        } else {
            return -204;
        }
    }
    if (irnd > 0) *rnrm = 0.0;
    return ierr;
}

/* Re-orthogonalize and normalize the new residual in column rcol (basis
 * size jnd): the reference's trl_orth (SRC/trlcore.c:1766-1975). */
static void lz_orth(lz_t *z, int jnd, int kept, int rcol, double rr2, const double *pass1)
{
    trl_info *info = z->info;
    double *alpha = z->alpha, *beta = z->beta;
    const double eps = DBL_EPSILON;
    const int jm1 = jnd - 1;
    if (!(rr2 >= 0.0) || !(rr2 <= DBL_MAX)) {
        info->stat = -102;
        return;
    }
    info->stat = 0;
    beta[jnd - 1] = sqrt(rr2);
    double tmp = alpha[jnd - 1] * alpha[jnd - 1];
    if (jm1 > kept) {
        tmp += beta[jm1 - 1] * beta[jm1 - 1];
    } else if (kept > 0) {
        for (int i = 0; i < jm1; ++i) tmp += beta[i] * beta[i];
    }
    int usecgs;
    if (jm1 == kept) {
        usecgs = 1;
    } else if (jnd >= info->ntot) {
        usecgs = 0;
    } else if (eps * rr2 >= tmp) {
        double anorm = 0.0;
        for (int i = 0; i < jnd; ++i) anorm = DMAX(anorm, fabs(alpha[i]) + fabs(beta[i]));
        usecgs = beta[jm1] < eps * anorm * info->ntot;
    } else {
        usecgs = 1;
    }
    if (usecgs) {
        const int rc = lz_cgs(z, jnd, rcol, &beta[jnd - 1], &info->north, pass1);
        if (rc == TRLHIP_STAT_DEVICE) return;
        info->stat = rc;
        info->flop += 4 * z->n * jnd + z->n;
    } else {
        /* local re-orthogonalization against the last one or two vectors */
        double h[4];
        const int c0 = jnd > 1 ? jnd - 2 : 0, m = jnd > 1 ? 2 : 1;
        if (trlhip_proj(z->dev, c0, m, rcol, h) != 0) {
            info->stat = TRLHIP_STAT_DEVICE;
            return;
        }
        alpha[jnd - 1] += h[(m - 1) * z->E]; /* real part for complex */
        if (trlhip_update(z->dev, c0, m, h, rcol, NULL) != 0) {
            info->stat = TRLHIP_STAT_DEVICE;
            return;
        }
        if (z->uo && lz_user_ortho(z, rcol, NULL)) return;
        if (trlhip_col_scale(z->dev, rcol, 1.0 / beta[jnd - 1]) != 0) {
            info->stat = TRLHIP_STAT_DEVICE;
            return;
        }
        info->flop += 9 * z->n;
    }
    if (info->stat == 0) {
        if (beta[jnd - 1] <= eps * fabs(alpha[jnd - 1])) beta[jnd - 1] = 0.0;
        else if (jnd >= info->ntot) beta[jnd - 1] = 0.0;
    }
}

/* Arrowhead (alpha, beta spike in the last row/column) of order nd to
 * tridiagonal form (alfrot, betrot) with the orthogonal rot:
 * reference trl_tridiag (SRC/trlcore.c:1978-2074). */
static void lz_tridiag(lz_t *z, int nd, const double *alpha, const double *beta,
                       double *alfrot, double *betrot)
{
    z->rot_n = nd;
    if (nd == 0) return;
    if (nd == 1) {
        z->rot[0] = 1.0;
        alfrot[0] = alpha[0];
        betrot[0] = beta[0];
        return;
    }
    double *a = z->rot;
    memset(a, 0, sizeof(double) * (size_t)nd * nd);
    for (int i = 0; i < nd; ++i) a[(size_t)i * nd + i] = alpha[i];
    for (int i = 0; i < nd - 1; ++i) a[(size_t)(nd - 1) * nd + i] = beta[i];
    trlkit_householder(nd, a, alfrot, betrot);
    betrot[nd - 1] = beta[nd - 1];
}

/* Ritz values and residual estimates of the current tridiagonal
 * (reference trl_get_eval, SRC/trlcore.c:2077-2146): lam ascending per
 * block, res = beta_jnd * |last eigenvector component|, tag = index of the
 * Ritz vector in the QL ordering (-1-i for locked i). */
static int lz_get_eval(lz_t *z, int jnd, int locked)
{
    const int m = jnd - locked;
    for (int i = 0; i < locked; ++i) {
        z->lam[i] = z->alfrot[i];
        z->res[i] = 0.0;
        z->tag[i] = -1 - i;
    }
    memcpy(z->lam + locked, z->alfrot + locked, sizeof(double) * (size_t)m);
    memcpy(z->esc, z->betrot + locked, sizeof(double) * (size_t)(m > 1 ? m - 1 : 0));
    if (trlkit_tridiag_ql(m, z->lam + locked, z->esc, z->dsc, 1) != 0) return -122;
    for (int i = 0; i < m; ++i) {
        z->res[locked + i] = z->betrot[jnd - 1] * fabs(z->dsc[i]);
        z->tag[locked + i] = i;
    }
    return 0;
}

/* Eigenvectors (columns of yy, order nd) of the tridiagonal alf/bet for
 * the nlam selected Ritz values (tags), then the rotation rot applied to
 * rows irot..irot+nrot-1: the role of the reference's trl_get_tvec
 * (dstein + rot, SRC/trlcore.c:2381-2490).  Here the vectors come from
 * the same QL iteration that produced the Ritz values. */
static int lz_get_tvec(lz_t *z, int nd, int irot, int nrot, const double *alf,
                       const double *bet, int nlam, const int *tag, double *yy)
{
    const int m = nd - irot; /* active block order (rows irot..nd-1) */
    double *Zq = z->zq;
    memset(Zq, 0, sizeof(double) * (size_t)m * m);
    for (int i = 0; i < m; ++i) Zq[(size_t)i * m + i] = 1.0;
    memcpy(z->dsc, alf + irot, sizeof(double) * (size_t)m);
    memcpy(z->esc, bet + irot, sizeof(double) * (size_t)(m > 1 ? m - 1 : 0));
    if (trlkit_tridiag_ql(m, z->dsc, z->esc, Zq, 2) != 0) return -132;
    for (int j = 0; j < nlam; ++j) {
        double *y = yy + (size_t)j * nd;
        memset(y, 0, sizeof(double) * (size_t)nd);
        if (tag[j] < 0) {
            y[-1 - tag[j]] = 1.0;
        } else {
            memcpy(y + irot, Zq + (size_t)tag[j] * m, sizeof(double) * (size_t)m);
        }
    }
    /* dstein bookkeeping the reference leaves in iwrk */
    for (int i = 0; i < nd && i < 4 * z->maxlan; ++i) z->iwrk[i] = 1;
    for (int i = nd; i < 2 * nd && i < 4 * z->maxlan; ++i) z->iwrk[i] = nd;
    if (nrot > 1) {
        double *tmp = z->dsc;
        for (int j = 0; j < nlam; ++j) {
            double *y = yy + (size_t)j * nd + irot;
            for (int r = 0; r < nrot; ++r) {
                double s = 0.0;
                for (int q = 0; q < nrot; ++q) s += z->rot[(size_t)q * nrot + r] * y[q];
                tmp[r] = s;
            }
            memcpy(y, tmp, sizeof(double) * (size_t)nrot);
        }
    }
    return 0;
}

/* Eigen-decomposition of the projected arrowhead matrix (diag alpha, the
 * first `kept` betas in column `kept`, the rest on the superdiagonal) and
 * the mapping of the nlam wanted values lambda onto its eigenvalues: the
 * reference's trl_get_tvec_a (SRC/trlcore.c:2493-2673), whose index loops
 * stop one short inside multiplicity groups (`< i2`, 2639-2653); that
 * behaviour is kept.  Overwrites alpha[0:nd) with the eigenvalues. */
static int lz_get_tvec_a(lz_t *z, int nd, int kept, double *alpha, const double *beta,
                         int nlam, const double *lambda, double *yy)
{
    if (nlam <= 0) return 0;
    memset(yy, 0, sizeof(double) * (size_t)nd * nd);
    for (int i = 0; i < nd; ++i) yy[(size_t)i * nd + i] = alpha[i];
    for (int i = 0; i < kept; ++i) yy[(size_t)kept * nd + i] = beta[i];
    for (int i = kept; i < nd - 1; ++i) yy[(size_t)(i + 1) * nd + i] = beta[i];
    if (trlkit_syev(nd, yy, alpha) != 0) return -142;
    if (nlam >= nd) return 0;
    double tol = DMAX(alpha[nd - 1] - alpha[0], DMAX(fabs(alpha[nd - 1]), fabs(alpha[0])));
    tol = DBL_EPSILON * tol * nd;
    int *map = z->iwrk;
    int i = 0, j = 0;
    while (i < nlam) {
        int ii = j;
        j = nd - 1;
        for (; ii < nd; ++ii)
            if (!(alpha[ii] < lambda[i] - tol)) {
                j = ii;
                break;
            }
        if (alpha[j] > lambda[i] + tol) return -143;
        int i2 = nlam - 1;
        for (ii = i + 1; ii < nlam; ++ii)
            if (!(lambda[ii] <= lambda[i] + tol)) {
                i2 = ii - 1;
                break;
            }
        int j2 = nd - 1;
        for (ii = j + 1; ii < nd; ++ii)
            if (!(alpha[ii] <= lambda[i] + tol)) {
                j2 = ii - 1;
                break;
            }
        if (j2 == j && i2 == i) {
            map[i] = j;
        } else if (j2 - j >= i2 - i) {
            for (ii = i; ii < i2; ++ii) map[ii] = j + ii - i;
        } else if (j2 < nd) {
            i2 = i + j2 - j;
            for (ii = i; ii < i2; ++ii) map[ii] = j + ii - i;
        } else {
            return -144;
        }
        i = i2 + 1;
        j = j2 + 1;
    }
    for (i = 0; i < nlam; ++i) {
        j = map[i];
        if (j > i) {
            alpha[i] = alpha[j];
            memcpy(yy + (size_t)i * nd, yy + (size_t)j * nd, sizeof(double) * (size_t)nd);
        }
    }
    return 0;
}

/* Initial vector in column nec (reference trl_initial_guess,
 * SRC/trlcore.c:1577-1763, guess <= 1).  Random numbers come from this
 * library's own generator (the reference seeds drand48 from clock()). */
/* Column streaming between checkpoint files and the device basis. */
typedef struct {
    lz_t *z;
    int base;   /* read: first basis column to fill */
    int jnd;    /* write: columns 0..jnd-1, then column rcol */
    int rcol;
} lz_colio_t;

static int lz_put_col(void *u, int i, double *buf)
{
    lz_colio_t *io = (lz_colio_t *)u;
    return trlhip_col_set(io->z->dev, io->base + i, buf);
}

static int lz_get_col(void *u, int i, double *buf)
{
    lz_colio_t *io = (lz_colio_t *)u;
    return trlhip_col_get(io->z->dev, i < io->jnd ? i : io->rcol, buf);
}

/* The reference's write_checkpoint (SRC/trlcore.c:434-460) /
 * zwrite_checkpoint (SRC/ztrlcore.c:360-430): alpha/beta of the jnd basis
 * vectors, the vectors and the residual (column rcol) to cpfile + process
 * number; complex columns as interleaved (re, im) pairs. */
static void lz_write_checkpoint(lz_t *z, int jnd, int rcol)
{
    trl_info *info = z->info;
    char file[160];
    trlkit_pe_filename(file, sizeof file, info->cpfile);
    lz_colio_t io = {z, 0, jnd, rcol};
    const clock_t c1 = lz_clock();
    const int ii = trlkit_write_checkpoint_e(file, z->n, z->E, jnd, z->alpha, z->beta, jnd + 1,
                                             lz_get_col, &io, z->hx);
    const clock_t c2 = lz_clock();
    info->clk_out += c2 >= c1 ? c2 - c1 : 0;
    info->wrds_out += jnd * (z->n + z->n + 2) + z->n + 2;
    info->stat = ii;
}

/* The workspace bound on the random perturbations of a guess <= 0 start:
 * trlan() hands trlanczos nmis doubles of misc workspace (SRC/trlan.c:
 * 172-226: maxlan (maxlan + 10), or what the caller's wrk holds beyond the
 * basis) and trl_initial_guess gets what follows alpha, beta, alfrot and
 * betrot in it (SRC/trlcore.c:805-806); ztrlan() passes the caller's nmis
 * whole (SRC/ztrlcore.c:617-618). */
/* The reference's misc workspace (trlanczos' wrk, of nmis doubles) for a
 * caller's lwrk, as trlan() partitions it (SRC/trlan.c:172-226). */
static long lz_ref_nmis(const lz_t *z, int lwrk)
{
    const long maxlan = z->info->maxlan;
    long ldb = ((long)(z->n + 3) / 4) * 4;
    if (ldb % 4096 == 0) ldb += 4;
    const long ii = DMAX(1L, maxlan - z->mev + 1) * ldb;
    long nmis = maxlan * (maxlan + 10);
    if (lwrk >= DMIN(ii, nmis)) {
        if (lwrk >= ii + nmis)
            nmis = lwrk - ii;
        else if (lwrk >= DMAX(ii, nmis)) {
            if (ii < nmis) nmis = lwrk;
        } else if (ii > nmis)
            nmis = lwrk;
    }
    return nmis;
}

static long lz_guess_cap(const lz_t *z, int lwrk)
{
    if (z->cplx) return lwrk;
    return lz_ref_nmis(z, lwrk) - 4L * z->info->maxlan;
}

static int lz_initial_guess(lz_t *z, const double *evec, int lde, int lwrk, int *j1)
{
    trl_info *info = z->info;
    const int j = info->nec;
    double *x = z->hx;
    const int E = z->E;
    if (info->guess > 1) {
        /* restore a checkpoint (SRC/trlcore.c:1653-1679, complex
         * ztrlcore.c:1413-1440): oldcpf if set, else cpfile; the basis
         * continues from its saved state */
        char file[160];
        trlkit_pe_filename(file, sizeof file, strlen(info->oldcpf) > 0 ? info->oldcpf : info->cpfile);
        lz_colio_t io = {z, j, 0, 0};
        int jnd = 0;
        const clock_t c1 = lz_clock();
        const int rc = trlkit_read_checkpoint_e(file, z->n, E, z->mev - j, z->ncols - 1 - j, &jnd,
                                                z->alpha + j, z->beta + j, lz_put_col, &io, x);
        const clock_t c2 = lz_clock();
        info->clk_in = c2 >= c1 ? c2 - c1 : 0;
        info->wrds_in = jnd * (z->n + z->n + 2) + z->n + 2;
        info->stat = rc;
        if (rc) return rc;
        /* The residual follows the restored basis, in column j + jnd.  The
         * reference reads it only when the basis ends inside evec
         * (SRC/trlaux.c:982-1007); otherwise its slot in base keeps the
         * zeros trlan() filled it with, and trl_cgs below replaces a zero
         * vector by a random one. */
        const int jr = j + jnd;
        if (DMIN(jnd, z->mev - j) >= z->mev - j) {
            memset(x, 0, sizeof(double) * (size_t)z->n * E);
            DEV(trlhip_col_set(z->dev, jr, x));
        }
        /* SRC/trlcore.c:1710-1741: the norm check looks at evec(:, nec)
         * (the first restored vector), then trl_cgs orthogonalizes the
         * residual against all jr restored columns and normalizes it */
        double nr2 = 0.0;
        DEV(trlhip_update(z->dev, 0, 0, &nr2, j, &nr2));
        if (!(nr2 >= DBL_MIN && nr2 <= DBL_MAX)) {
            if (lz_random_column(z, j)) return TRLHIP_STAT_DEVICE;
            info->nrand++;
        }
        *j1 = jr;
        double rnrm = 0.0;
        int local_north = 0;
        const int rcg = lz_cgs(z, jr, jr, &rnrm, &local_north, NULL);
        if (rcg == TRLHIP_STAT_DEVICE) return rcg;
        info->stat = rcg;
        return rcg;
    }
    if (info->guess == 1 && z->dstart && !z->cplx) {
        x = NULL;  /* the start vector is already on the device */
    } else if (info->guess == 1) {
        memcpy(x, evec + (size_t)j * lde * E, sizeof(double) * E * (size_t)z->n);
    } else if (z->cplx) {
        /* ztrl_initial_guess (SRC/ztrlcore.c:1421-1490): (1,0), or (1,1)
         * plus -guess random perturbations */
        for (int k = 0; k < z->n; ++k) {
            x[2 * k] = 1.0;
            x[2 * k + 1] = info->guess == 0 ? 0.0 : 1.0;
        }
        if (info->guess < 0) {
            long nran = DMIN(-(long)info->guess, lz_guess_cap(z, lwrk));
            nran = 2 * (nran / 2);
            if (nran < z->n) {
                for (int k = 0; k < nran - 1; k += 2) {
                    const int ii = (int)(z->n * trl_rand_uniform(&z->rng));
                    x[2 * ii] += trl_rand_uniform(&z->rng) - 0.5;
                    x[2 * ii + 1] += trl_rand_uniform(&z->rng) - 0.5;
                }
            } else {
                for (int k = 0; k < 2 * z->n; ++k) x[k] = trl_rand_uniform(&z->rng);
                trlkit_smooth_zz(z->n, x);
                info->nrand++;
            }
        }
    } else {
        for (int k = 0; k < z->n; ++k) x[k] = 1.0;
        long nran = DMIN(1 - (long)info->guess, lz_guess_cap(z, lwrk));
        nran = 2 * (nran / 2);
        if (nran > 0 && nran < z->n) {
            for (int k = 0; k < nran - 1; k += 2) {
                const int ii = (int)(z->n * trl_rand_uniform(&z->rng));
                x[ii] += trl_rand_uniform(&z->rng) - 0.5;
            }
        } else if (nran >= z->n) {
            for (int k = 0; k < z->n; ++k) x[k] = trl_rand_uniform(&z->rng);
            trlkit_smooth_rr(z->n, x);
            info->nrand++;
        }
    }
    if (x)
        DEV(trlhip_col_set(z->dev, j, x));
    else
        DEV(trlhip_col_set_device_f32(z->dev, j, z->dstart));
    /* ||x||^2 on the device; a zero/inf/nan start is replaced by random */
    double nr2 = 0.0;
    DEV(trlhip_update(z->dev, 0, 0, &nr2, j, &nr2));
    double rnrm = 0.0;
    if (nr2 >= DBL_MIN && nr2 <= DBL_MAX) {
        rnrm = sqrt(nr2);
    } else {
        if (lz_random_column(z, j)) return TRLHIP_STAT_DEVICE;
        info->nrand++;
        DEV(trlhip_update(z->dev, 0, 0, &nr2, j, &nr2));
        rnrm = sqrt(nr2);
    }
    *j1 = info->nec;
    int local_north = 0;
    const int rc = lz_cgs(z, *j1, j, &rnrm, &local_north, NULL);
    if (rc == TRLHIP_STAT_DEVICE) return rc;
    info->stat = rc;
    return rc;
}

/* ------------------------------------------------------------- driver */

static int lz_alloc(lz_t *z, int maxlan)
{
    const size_t m = (size_t)maxlan + 2;
    z->alpha = (double *)calloc(m, sizeof(double));
    z->beta = (double *)calloc(m, sizeof(double));
    z->alfrot = (double *)calloc(m, sizeof(double));
    z->betrot = (double *)calloc(m, sizeof(double));
    z->rot = (double *)calloc(m * m, sizeof(double));
    z->lam = (double *)calloc(m, sizeof(double));
    z->res = (double *)calloc(m, sizeof(double));
    z->tag = (int *)calloc(m, sizeof(int));
    z->ctw = (double *)calloc(2 * m, sizeof(double));
    z->shw = (double *)calloc(4 * m + 1, sizeof(double));
    z->yy = (double *)calloc(m * m, sizeof(double));
    z->zq = (double *)calloc(m * m, sizeof(double));
    z->dsc = (double *)calloc(m, sizeof(double));
    z->esc = (double *)calloc(m, sizeof(double));
    z->iwrk = (int *)calloc(4 * m, sizeof(int));
    /* host staging columns: every use writes all n entries before reading
     * them, so no zeroing (calloc would clear 16n bytes per solve, most of
     * a batch's setup time) */
    z->hx = (double *)malloc(sizeof(double) * (size_t)z->n * z->E);
    z->hy = (double *)malloc(sizeof(double) * (size_t)z->n * z->E);
    return (z->alpha && z->beta && z->alfrot && z->betrot && z->rot && z->lam && z->res &&
            z->tag && z->ctw && z->shw && z->yy && z->zq && z->dsc && z->esc && z->iwrk && z->hx && z->hy)
               ? 0
               : -1;
}

static void lz_free(lz_t *z)
{
    free(z->alpha);
    free(z->beta);
    free(z->alfrot);
    free(z->betrot);
    free(z->rot);
    free(z->lam);
    free(z->res);
    free(z->tag);
    free(z->ctw);
    free(z->shw);
    free(z->yy);
    free(z->zq);
    free(z->dsc);
    free(z->esc);
    free(z->iwrk);
    free(z->hx);
    free(z->hy);
}

/* One operator application and recurrence, with the first Gram-Schmidt
 * pass done speculatively on the device in the same round trip
 * (trlhip_step_cgs); pass1 receives {||h||^2, ||s||^2} of that pass. */
static int lz_matvec_step(lz_t *z, int first, int qa, int qb, double beta_prev, int jnd,
                          int rcol, double *alpha, double *rr2, double *pass1)
{
    clock_t t0 = lz_clock();
    if (!z->A) {
        if (lz_apply_host_op(z, qa, rcol)) return TRLHIP_STAT_DEVICE;
    }
    double out[4];
    const int rc = trlhip_step_cgs(z->dev, z->A, first, qa, qb, beta_prev, z->beta, jnd, rcol, out);
    add_ticks(z->info, &z->info->clk_op, &z->info->tick_o, t0);
    if (rc) {
        z->info->stat = TRLHIP_STAT_DEVICE;
        return TRLHIP_STAT_DEVICE;
    }
    *alpha = out[0];
    *rr2 = out[1];
    pass1[0] = out[2];
    pass1[1] = out[3];
    z->info->matvec++;
    return 0;
}

/* Eigen-decomposition of a symmetric-definite pencil (H, G) of order m
 * (column-major, overwritten): Cholesky G = L L', C = L^-1 H L^-T, cyclic
 * Jacobi on C, Z = L^-T W.  theta (m) ascending is not enforced; Z (m x m)
 * holds the G-orthonormal eigenvectors.  Returns 0, or -1 when G is not
 * positive definite or Jacobi does not converge. */
static int lz_sygv(int m, double *H, double *G, double *theta, double *Z)
{
    for (int j = 0; j < m; ++j) { /* Cholesky, lower, in G */
        double d = G[j + (size_t)j * m];
        for (int q = 0; q < j; ++q) d -= G[j + (size_t)q * m] * G[j + (size_t)q * m];
        if (!(d > 0.0)) return -1;
        d = sqrt(d);
        G[j + (size_t)j * m] = d;
        for (int i = j + 1; i < m; ++i) {
            double t = G[i + (size_t)j * m];
            for (int q = 0; q < j; ++q) t -= G[i + (size_t)q * m] * G[j + (size_t)q * m];
            G[i + (size_t)j * m] = t / d;
        }
    }
    /* C = L^-1 H L^-T: forward substitution on the columns, then the rows */
    for (int c = 0; c < m; ++c)
        for (int i = 0; i < m; ++i) {
            double t = H[i + (size_t)c * m];
            for (int q = 0; q < i; ++q) t -= G[i + (size_t)q * m] * H[q + (size_t)c * m];
            H[i + (size_t)c * m] = t / G[i + (size_t)i * m];
        }
    for (int r = 0; r < m; ++r)
        for (int i = 0; i < m; ++i) {
            double t = H[r + (size_t)i * m];
            for (int q = 0; q < i; ++q) t -= G[i + (size_t)q * m] * H[r + (size_t)q * m];
            H[r + (size_t)i * m] = t / G[i + (size_t)i * m];
        }
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) Z[i + (size_t)j * m] = i == j ? 1.0 : 0.0;
    int conv = 0;
    for (int sweep = 0; sweep < 60 && !conv; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                const double v = H[i + (size_t)j * m];
                tot += v * v;
                if (i != j) off += v * v;
            }
        /* converged when the off-diagonal part is at rounding level: below
         * 1e-24 of the total, or when no pair needs a rotation */
        if (off <= 1e-24 * tot) {
            conv = 1;
            break;
        }
        int rotations = 0;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = 0.5 * (H[p + (size_t)q * m] + H[q + (size_t)p * m]);
                const double app = H[p + (size_t)p * m], aqq = H[q + (size_t)q * m];
                if (fabs(apq) <= DBL_EPSILON * sqrt(fabs(app * aqq)) || apq == 0.0) {
                    H[p + (size_t)q * m] = H[q + (size_t)p * m] = 0.0;
                    continue;
                }
                ++rotations;
                const double tau = (aqq - app) / (2.0 * apq);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                const double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
                for (int k = 0; k < m; ++k) { /* columns p, q */
                    const double hp = H[k + (size_t)p * m], hq = H[k + (size_t)q * m];
                    H[k + (size_t)p * m] = cs * hp - sn * hq;
                    H[k + (size_t)q * m] = sn * hp + cs * hq;
                }
                for (int k = 0; k < m; ++k) { /* rows p, q */
                    const double hp = H[p + (size_t)k * m], hq = H[q + (size_t)k * m];
                    H[p + (size_t)k * m] = cs * hp - sn * hq;
                    H[q + (size_t)k * m] = sn * hp + cs * hq;
                }
                for (int k = 0; k < m; ++k) {
                    const double zp = Z[k + (size_t)p * m], zq = Z[k + (size_t)q * m];
                    Z[k + (size_t)p * m] = cs * zp - sn * zq;
                    Z[k + (size_t)q * m] = sn * zp + cs * zq;
                }
            }
        if (rotations == 0) conv = 1;
    }
    if (!conv) return -1;
    for (int i = 0; i < m; ++i) theta[i] = H[i + (size_t)i * m];
    /* Z <- L^-T Z (back substitution on each column) */
    for (int c = 0; c < m; ++c)
        for (int i = m - 1; i >= 0; --i) {
            double t = Z[i + (size_t)c * m];
            for (int q = i + 1; q < m; ++q) t -= G[q + (size_t)i * m] * Z[q + (size_t)c * m];
            Z[i + (size_t)c * m] = t / G[i + (size_t)i * m];
        }
    return 0;
}

/* Rayleigh-Ritz polish of a mixed-precision solve's k Ritz vectors
 * (columns 0..k-1, Rayleigh quotients rho): over Q = [X, R] with R the
 * normalized FP64 residuals A x_i - rho_i x_i (columns k..2k-1), solve
 * Q'AQ z = theta Q'Q z and keep, for each x_i in turn, the unused Ritz
 * pair nearest rho_i; the new vectors overwrite columns 0..k-1 (restart
 * GEMM).  The FP32 shadow's projection errors leave components outside
 * span(X) that no rotation inside it removes; one step over [X, R]
 * removes most of them.  Column `avoid` is preserved.  Returns 1 when
 * the columns were replaced, 0 when skipped (too few columns, a zero
 * residual, an indefinite Gram matrix), or a device error. */
static int lz_polish(lz_t *z, int k, int avoid, const double *rho)
{
    const int m = 2 * k;
    int save = -1, tmp = m;
    if (avoid >= k && avoid <= m) { /* keep the last residual out of the way */
        save = m + 1;
        tmp = m;
    } else if (tmp == avoid) {
        tmp = m + 1;
    }
    if ((save >= 0 ? save : tmp) >= z->ncols) return 0;
    /* G, H, Z, Gs, Hs (m x m each), theta, rnorm (m each), Y (m x k) */
    double *G = (double *)calloc((size_t)m * m * 5 + 2 * (size_t)m + (size_t)m * k, sizeof(double));
    int *sel = (int *)calloc((size_t)m, sizeof(int));
    if (!G || !sel) {
        free(G);
        free(sel);
        return 0;
    }
    double *H = G + (size_t)m * m, *Z = H + (size_t)m * m, *Gs = Z + (size_t)m * m;
    double *Hs = Gs + (size_t)m * m, *theta = Hs + (size_t)m * m, *rnorm = theta + m;
    double *Y = rnorm + m;
    int rc = 0, used = 0;
    if (save >= 0 && trlhip_col_copy(z->dev, save, avoid)) rc = TRLHIP_STAT_DEVICE;
    /* the k residuals and their squared norms in one device pass (rnorm
     * takes the norms in place), then G and H in another */
    if (!rc && trlhip_residuals(z->dev, z->A, k, rho, rnorm)) rc = TRLHIP_STAT_DEVICE;
    for (int i = 0; i < k && !rc; ++i) {
        const double rr2 = rnorm[i];
        if (!(rr2 > DBL_MIN))
            break;
        else if (trlhip_col_scale(z->dev, k + i, 1.0 / sqrt(rr2)))
            rc = TRLHIP_STAT_DEVICE;
        else
            rnorm[used++] = sqrt(rr2);
    }
    if (!rc && used == k && trlhip_gram(z->dev, z->A, m, tmp, G, H)) rc = TRLHIP_STAT_DEVICE;
    z->info->matvec += used + (used == k ? m : 0);
    int done = 0;
    if (!rc && used == k) {
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < i; ++j) {
                const double g = 0.5 * (G[i + (size_t)j * m] + G[j + (size_t)i * m]);
                const double h = 0.5 * (H[i + (size_t)j * m] + H[j + (size_t)i * m]);
                G[i + (size_t)j * m] = G[j + (size_t)i * m] = g;
                H[i + (size_t)j * m] = H[j + (size_t)i * m] = h;
            }
        /* columns of Q to use: X, then the residuals by decreasing norm,
         * each kept only when it adds a direction (Cholesky pivot above
         * 1e-10 of its diagonal): residuals of a cluster of pairs often
         * share one dominant direction, which makes [X, R] singular */
        int ns = 0;
        for (int i = 0; i < k; ++i) sel[ns++] = i;
        int order[256];
        const int kk = k < 256 ? k : 256;
        for (int i = 0; i < kk; ++i) order[i] = i;
        for (int i = 0; i < kk; ++i)
            for (int j = i + 1; j < kk; ++j)
                if (rnorm[order[j]] > rnorm[order[i]]) {
                    const int t = order[i];
                    order[i] = order[j];
                    order[j] = t;
                }
        double *L = Z;  /* scratch: Cholesky rows of the kept columns */
        for (int a = 0; a < ns; ++a) {
            double d = G[sel[a] + (size_t)sel[a] * m];
            for (int q = 0; q < a; ++q) d -= L[a + (size_t)q * m] * L[a + (size_t)q * m];
            d = d > 0.0 ? sqrt(d) : 0.0;
            L[a + (size_t)a * m] = d;
            for (int b = a + 1; b < ns; ++b) {
                double t = G[sel[b] + (size_t)sel[a] * m];
                for (int q = 0; q < a; ++q) t -= L[b + (size_t)q * m] * L[a + (size_t)q * m];
                L[b + (size_t)a * m] = d > 0.0 ? t / d : 0.0;
            }
        }
        for (int ii = 0; ii < kk; ++ii) {
            const int c = k + order[ii];
            double row[512];
            double d = G[c + (size_t)c * m];
            for (int q = 0; q < ns; ++q) {
                double t = G[c + (size_t)sel[q] * m];
                for (int r = 0; r < q; ++r) t -= row[r] * L[q + (size_t)r * m];
                row[q] = L[q + (size_t)q * m] > 0.0 ? t / L[q + (size_t)q * m] : 0.0;
                d -= row[q] * row[q];
            }
            if (!(d > 1e-10 * G[c + (size_t)c * m]) || ns >= 512) continue;
            for (int q = 0; q < ns; ++q) L[ns + (size_t)q * m] = row[q];
            L[ns + (size_t)ns * m] = sqrt(d);
            sel[ns++] = c;
        }
        for (int a = 0; a < ns; ++a)
            for (int b = 0; b < ns; ++b) {
                Gs[a + (size_t)b * ns] = G[sel[a] + (size_t)sel[b] * m];
                Hs[a + (size_t)b * ns] = H[sel[a] + (size_t)sel[b] * m];
            }
        if (lz_sygv(ns, Hs, Gs, theta, Z) == 0) {
            int *taken = (int *)calloc((size_t)ns, sizeof(int));
            if (taken) {
                memset(Y, 0, sizeof(double) * (size_t)m * k);
                for (int i = 0; i < k; ++i) {
                    int best = -1;
                    for (int j = 0; j < ns; ++j)
                        if (!taken[j] && (best < 0 || fabs(theta[j] - rho[i]) < fabs(theta[best] - rho[i])))
                            best = j;
                    taken[best] = 1;
                    for (int a = 0; a < ns; ++a) Y[(size_t)i * m + sel[a]] = Z[a + (size_t)best * ns];
                }
                free(taken);
                if (trlhip_ritz(z->dev, 0, m, k, Y, m)) rc = TRLHIP_STAT_DEVICE;
                else done = 1;
            }
        }
    }
    if (!rc && save >= 0 && trlhip_col_copy(z->dev, avoid, save)) rc = TRLHIP_STAT_DEVICE;
    free(G);
    free(sel);
    if (rc) {
        z->info->stat = TRLHIP_STAT_DEVICE;
        return rc;
    }
    return done;
}

/* Final FP64 Rayleigh-quotient refinement of a mixed-precision solve:
 * for each of the k Ritz vectors v in device columns 0..k-1,
 * eval = v'Av / v'v with the FP64 operator, and beta = ||A v - eval v|| /
 * ||v|| (the true residual, reported in wrk instead of the estimate), all
 * k pairs in one device call (trlhip_refine); column `avoid` (the last
 * residual) is kept.  When a pair's true residual exceeds tol * ||A||,
 * the vectors get one Rayleigh-Ritz polish over [X, residuals]
 * (lz_polish) and are refined again.  Every operator application counts
 * in matvec. */
static int lz_refine(lz_t *z, int k, int avoid, double *eval, double *beta)
{
    if (k < 1) return 0;
    if (!z->A || 2 * k > z->maxlan * z->maxlan) {
        z->info->stat = TRLHIP_STAT_DEVICE;
        return TRLHIP_STAT_DEVICE;
    }
    double *rho = z->zq, *res = z->zq + k;  /* zq: free scratch after the last restart */
    DEV(trlhip_refine(z->dev, z->A, k, avoid, rho, res));
    z->info->matvec += k;
    const double bnd = z->info->tol * z->info->anrm;
    int over = 0, nan = 0;
    for (int i = 0; i < k; ++i) {
        over += res[i] > bnd;
        nan += isnan(rho[i]) != 0;
    }
    if (over > 0 && nan == 0) {
        const int p = lz_polish(z, k, avoid, rho);
        if (p < 0 || z->info->stat != 0) return z->info->stat ? z->info->stat : p;
        if (p == 1) {
            DEV(trlhip_refine(z->dev, z->A, k, avoid, rho, res));
            z->info->matvec += k;
        }
    }
    for (int i = 0; i < k; ++i) {
        if (isnan(rho[i])) continue;
        eval[i] = rho[i];
        beta[i] = res[i];
    }
    return 0;
}

/* The reference's convergence test on the current Ritz values.  In a
 * mixed-precision solve the Gram-Schmidt correction is taken against the
 * FP32 shadow basis, and the part of it the three-term recurrence leaves
 * out (V32 h, of order FLT_EPSILON * beta) moves the true residual of a
 * Ritz pair off the recurrence's estimate by up to about a tenth of
 * tol * ||A|| at tol = 1e-8; there pairs are accepted at half the
 * tolerance so that the refined (true) residual still meets tol. */
static void lz_convergence_test(lz_t *z, int jnd)
{
    trl_info *info = z->info;
    const double tol = info->tol;
    if (z->mixed) info->tol = 0.5 * tol;
    trlkit_convergence_test_w(jnd, z->lam, z->res, z->tag, info, z->ctw);
    info->tol = tol;
}

/* The thick-restart iteration; mirrors the decision sequence of the
 * reference's trlanczos (SRC/trlcore.c:723-1532). */
/* The reference's debug checks of the basis (SRC/trlcore.c:1025-1029,
 * 1199-1210, 1497-1505, 1748-1757): vectors 0..tot-1 with the last one
 * in column `last`, split into the reference's evec and base parts
 * (j1 = min(tot, mev)); the recurrence check at verbose > 7.  avail is
 * the reference's workspace there; the complex solver's checks always run
 * (its workspace is not modelled). */
static void lz_debug_basis(lz_t *z, int tot, int last, int kept, int recurrence, long avail)
{
    const int j1 = DMIN(tot, z->mev), j2 = tot - j1;
    trl_dbg_check_orth(z->info, z->dev, z->E, j1, j2, tot - 1, last, avail);
    if (recurrence && z->info->verbose > 7)
        trl_dbg_check_recurrence(z->info, z->dev, z->n, z->E, z->op, z->zop, j1, j2, kept,
                                 z->alpha, z->beta, tot - 1, last, avail);
}

static void lz_iterate(lz_t *z, double *eval, double *evec, int lde, double *wrk, int lwrk)
{
    trl_info *info = z->info;
    const int mev = z->mev;
    double *alpha = z->alpha, *beta = z->beta, *alfrot = z->alfrot, *betrot = z->betrot;
    int locked = info->nec;
    if (locked > 0) {
        memcpy(alpha, eval, sizeof(double) * (size_t)locked);
        memset(beta, 0, sizeof(double) * (size_t)locked);
        for (int i = 0; i < locked; ++i)
            if (trlhip_col_set(z->dev, i, evec + (size_t)i * lde * z->E)) {
                info->stat = TRLHIP_STAT_DEVICE;
                return;
            }
    }
    int j1 = 0;
    if (lz_initial_guess(z, evec, lde, lwrk, &j1) != 0) return;
    /* the reference's workspace (trlanczos' lwrk) for the debug checks */
    const long ws = z->cplx ? LONG_MAX / 4 : lz_ref_nmis(z, lwrk);
    const long ws4 = ws - 4L * info->maxlan;
    if (info->verbose > 6) {
        /* the reference checks its j1 = nec vectors, plus the start vector
         * only when it lies in base (trlcore.c:1748-1757) */
        const int tot = j1 < mev ? j1 : j1 + 1;
        if (tot > 0) lz_debug_basis(z, tot, tot - 1, 0, 0, ws4);
    }
    int jnd = j1, kept = jnd;
    long ws_cycle = ws4;  /* lwrk2 of the current cycle (trlcore.c:860-869) */
    int i1 = info->ned - jnd;
    int next_test = i1 + DMIN(i1, DMIN(6, info->ned / 2));
    /* checkpoints every maxmv/cpflag operator applications */
    const int ckpt = info->cpflag > 0;
    int chkpnt = ckpt ? info->maxmv / info->cpflag : info->maxmv + info->maxlan;
    int rr = jnd, qa = -1, qb = -1, i2 = 0;
    while (info->matvec < info->maxmv && info->nec < info->ned) {
        /* ---- first step of a cycle: r = A q - sum beta_i q_i */
        ++jnd;
        qa = jnd - 1;
        rr = jnd;
        if (rr >= z->ncols) {
            info->stat = -1111;
            return;
        }
        double a, rr2, pass1[2];
        ws_cycle = ws4 - (long)(jnd - locked) * (jnd - locked);
        if (lz_matvec_step(z, 1, qa, -1, 0.0, jnd, rr, &a, &rr2, pass1)) return;
        alpha[jnd - 1] = a;
        beta[jnd - 1] = a;
        clock_t t0 = lz_clock();
        lz_orth(z, jnd, kept, rr, rr2, pass1);
        if (info->verbose > 8 && info->stat == 0) lz_debug_basis(z, jnd + 1, rr, kept, 0, ws_cycle);
        add_ticks(info, &info->clk_orth, &info->tick_h, t0);
        if (info->stat != 0) return;
        memcpy(alfrot, alpha, sizeof(double) * (size_t)locked);
        memset(betrot, 0, sizeof(double) * (size_t)locked);
        const long tt0 = g_prof ? lz_ns() : 0;
        lz_tridiag(z, jnd - locked, alpha + locked, beta + locked, alfrot + locked, betrot + locked);
        if (g_prof) __atomic_add_fetch(&g_ns_tridiag, lz_ns() - tt0, __ATOMIC_RELAXED);
        info->flop += 8 * (jnd - locked) * (jnd - locked) * (jnd - locked) / 3;
        betrot[jnd - 1] = beta[jnd - 1];
        /* ---- three-term recurrence until the basis is full */
        while (jnd < info->klan && info->nec < info->ned) {
            qb = qa;
            qa = rr;
            ++jnd;
            rr = jnd;
            if (rr >= z->ncols) {
                info->stat = -1111;
                return;
            }
            /* batched: the steps up to the next convergence test or the
             * end of the cycle may run in this round, as long as the device
             * sees each accepted as the solver will (trlhip_step_chain) */
            if (z->chain && !z->uo && info->verbose <= 6) {
                const int left = DMIN(DMIN(z->chain, info->klan - jnd + 1),
                                      DMAX(1, next_test - info->matvec));
                (void)trlhip_step_chain(z->dev, left, kept, info->ntot);
            }
            if (lz_matvec_step(z, 0, qa, qb, beta[jnd - 2], jnd, rr, &a, &rr2, pass1)) return;
            alpha[jnd - 1] = a;
            info->flop += 6 * z->n;
            t0 = lz_clock();
            lz_orth(z, jnd, kept, rr, rr2, pass1);
            add_ticks(info, &info->clk_orth, &info->tick_h, t0);
            alfrot[jnd - 1] = alpha[jnd - 1];
            betrot[jnd - 1] = beta[jnd - 1];
            if (info->stat != 0) return;
            if (info->matvec >= next_test) {
                const long tc0 = g_prof ? lz_ns() : 0;
                if ((info->stat = lz_get_eval(z, jnd, locked)) != 0) return;
                memcpy(eval, z->lam, sizeof(double) * (size_t)DMIN(mev, jnd));
                lz_convergence_test(z, jnd);
                if (info->nec < info->ned && info->nec > 0) {
                    next_test = (int)((double)(info->ned * info->matvec) / (double)info->nec);
                } else if (info->nec == 0) {
                    next_test += next_test;
                    if (info->maxlan == info->ntot)
                        next_test = (int)ceil(0.5 * (info->maxlan + info->matvec));
                }
                if (g_prof) __atomic_add_fetch(&g_ns_check, lz_ns() - tc0, __ATOMIC_RELAXED);
            }
        }
        /* ---- restart or finish */
        if (info->verbose > 6) lz_debug_basis(z, jnd + 1, rr, kept, 1, ws_cycle - 2L * jnd);
        clock_t tr = lz_clock();
        long ws_end = ws4;  /* lwrk2 after the restart (trlcore.c:1354, 1460) */
        const long tr0 = g_prof ? lz_ns() : 0;
        const int prek = kept;
        const int jml = jnd - locked;
        i2 = kept - locked + 1;
        if (info->nec < info->ned) {
            if ((info->stat = lz_get_eval(z, jnd, locked)) != 0) return;
            lz_convergence_test(z, jnd);
        }
        if (info->nec < info->ned && info->matvec < info->maxmv) {
            /* lambda and res as the reference's workspace holds them when
             * it restarts (SRC/trlcore.c:868, 1200-1201, 1256, 1283): the
             * last element of betrot, lambda[0:jnd), res[0:jnd), then the
             * convergence test's |res| and lambda; the restart schemes
             * read one element beyond either end of their windows
             * (lambda[-1], res[-1], lambda[jml], res[jml]) */
            double *w = z->shw;
            w[0] = betrot[info->maxlan - 1];
            memcpy(w + 1, z->lam, sizeof(double) * (size_t)jnd);
            memcpy(w + 1 + jnd, z->res, sizeof(double) * (size_t)jnd);
            memcpy(w + 1 + 2 * jnd, z->ctw, sizeof(double) * (size_t)(2 * jnd));
            kept = trlkit_shuffle_eig(jml, info->klan - locked, w + 1 + locked, w + 1 + jnd + locked,
                                      z->tag + locked, info, kept);
            memcpy(z->lam + locked, w + 1 + locked, sizeof(double) * (size_t)jml);
            memcpy(z->res + locked, w + 1 + jnd + locked, sizeof(double) * (size_t)jml);
            int st = 0;
            if (kept * 3 < jml) {
                st = lz_get_tvec(z, jml, 0, i2, alfrot + locked, betrot + locked, kept,
                                 z->tag + locked, z->yy);
                if (st == 0 && locked + kept > 0)
                    memcpy(alpha, z->lam, sizeof(double) * (size_t)(locked + kept));
            }
            if (kept * 3 >= jml || st != 0) {
                if (locked + kept > 0) memcpy(alfrot, z->lam, sizeof(double) * (size_t)(locked + kept));
                st = lz_get_tvec_a(z, jml, prek - locked, alpha + locked, beta + locked, kept,
                                   alfrot + locked, z->yy);
            }
            if (st != 0) {
                info->stat = st;
                return;
            }
            for (int i = 0; i < kept; ++i)
                beta[locked + i] = z->yy[(size_t)(i + 1) * jml - 1] * betrot[jnd - 1];
            /* a mixed-precision basis locks nothing: its residual estimate
             * reaches the locking level (DBL_EPSILON |lambda|) while the
             * true residual still sits near FLT_EPSILON-level noise, and a
             * locked column is never improved by later restarts */
            if (!z->mixed && jml > info->ned + (info->ned / 5 + 6))
                i2 = trlkit_set_locking(jml, kept, alpha + locked, beta + locked, z->yy, info->anrm);
            else
                i2 = 0;
            if (trlhip_ritz(z->dev, locked, jml, kept, z->yy, jml) != 0) {
                info->stat = TRLHIP_STAT_DEVICE;
                return;
            }
            info->flop += 2 * z->n * jml * kept;
            ws_end = ws4 - (long)kept * jml - (jml + locked);
            kept += locked;
            locked += i2;
            info->locked = locked;
            jnd = kept;
            if (jnd > mev && jnd - mev >= z->nbas - 1) {
                info->stat = -1111;
                return;
            }
            if (info->nec > 0)
                next_test = (int)(((double)(info->matvec * info->ned)) / ((double)info->nec));
            else
                next_test += info->maxlan;
            if (DMIN(mev, jnd) > 0) memcpy(eval, z->lam, sizeof(double) * (size_t)DMIN(mev, jnd));
            /* the residual becomes the vector after the kept Ritz vectors */
            if (trlhip_col_copy(z->dev, jnd, rr) != 0) {
                info->stat = TRLHIP_STAT_DEVICE;
                return;
            }
            rr = jnd;
            if (g_prof) {
                __atomic_add_fetch(&g_ns_restart, lz_ns() - tr0, __ATOMIC_RELAXED);
                __atomic_add_fetch(&g_n_restart, 1, __ATOMIC_RELAXED);
            }
            if (ckpt && info->matvec >= chkpnt) {
                lz_write_checkpoint(z, jnd, rr);
                if (info->stat != 0) return;
                chkpnt += info->maxmv / info->cpflag;
            }
        } else {
            kept = DMIN(info->nec, DMAX(info->ned, mev - 1));
            info->nec = kept;
            if (kept == 0) kept = DMIN(mev - 1, info->ned);
            trlkit_sort_eig(jnd, info->lohi, kept, info->ref, z->lam, z->res, z->tag);
            memcpy(eval, z->lam, sizeof(double) * (size_t)kept);
            int st = 0;
            if (kept * 3 < jnd)
                st = lz_get_tvec(z, jnd, locked, i2, alfrot, betrot, kept, z->tag, z->yy);
            if (kept * 3 >= jnd || st != 0) {
                st = lz_get_tvec_a(z, jnd, prek, alpha, beta, kept, eval, z->yy);
                if (st != 0) {
                    info->stat = st;
                    return;
                }
            }
            if (kept > 0) memcpy(alpha, eval, sizeof(double) * (size_t)kept);
            ws_end = ws4 - (long)kept * jnd;
            for (int i = 0; i < kept; ++i) beta[i] = betrot[jnd - 1] * z->yy[(size_t)(i + 1) * jnd - 1];
            if (kept > 0 && trlhip_ritz(z->dev, 0, jnd, kept, z->yy, jnd) != 0) {
                info->stat = TRLHIP_STAT_DEVICE;
                return;
            }
            info->flop += 2 * z->n * jml * kept;
            if (z->mixed && lz_refine(z, kept, rr, eval, beta) != 0) return;
            jnd = kept;
            j1 = kept;
            /* Ritz vectors (and the last residual after them) to the caller,
             * the vectors in one transfer */
            if (DMIN(kept, mev) > 0 && trlhip_cols_get(z->dev, 0, DMIN(kept, mev), evec, lde)) {
                info->stat = TRLHIP_STAT_DEVICE;
                return;
            }
            if (kept < mev && trlhip_col_get(z->dev, rr, evec + (size_t)kept * lde * z->E)) {
                info->stat = TRLHIP_STAT_DEVICE;
                return;
            }
            if (ckpt) {
                lz_write_checkpoint(z, kept, rr);
                if (info->stat != 0) return;
            }
        }
        if (info->verbose > 6) lz_debug_basis(z, jnd + 1, rr, kept, 1, ws_end);
        add_ticks(info, &info->clk_res, &info->tick_r, tr);
        info->nloop++;
    }
    /* residual-norm estimates, then what the reference's workspace holds;
     * complex: misc[i].r (stride 2) like ztrlanczos (ztrlcore.c:1166-1168) */
    if (!wrk) return;
    const int nw = DMIN(mev, lwrk);
    for (int i = 0; i < nw; ++i) wrk[(size_t)i * z->E] = i < j1 ? fabs(beta[i]) : alpha[i];
}

/* Validation and derived sizes: the reference's trl_clear_counter
 * (SRC/trlan.c:1217-1349). */
static void check_and_clear(trl_info *info, int nrow, int mev, int lde)
{
    info->stat = 0;
    if (nrow != info->nloc || nrow > info->ntot) {
        printf("TRLAN: info not setup for this problem.\n");
        info->stat = -1;
    }
    if (info->nec < 0) info->nec = 0;
    if (lde < nrow) {
        printf("TRLAN: leading dimension of EVEC to small.\n");
        info->stat = -2;
    }
    if (info->tol >= 1.0) info->tol = sqrt(DBL_EPSILON);
    else if (info->tol <= DBL_MIN) info->tol = DBL_EPSILON;
    if (info->ned + info->ned >= info->ntot && info->ned > info->ntot) {
        info->ned = DMIN(info->ntot - 1, info->maxlan - 3);
        printf("TRLAN: ** reduced ned to %d **\n", info->ned);
    }
    if (mev < info->ned) {
        printf("TRLAN: array EVAL and EVEC can not hold wanted number of eigenpairs.\n");
        info->stat = -3;
    }
    if (info->ntot < 10) {
        printf("TRLAN is not designed to work with such a small matrix(%dx%d).\n", info->ntot,
               info->ntot);
        info->stat = -4;
    }
    if (info->maxlan < info->ned + 3) {
        info->maxlan = info->ned + DMIN(info->ned, 20) + (int)log((double)info->ntot);
        info->maxlan = DMIN(info->maxlan, info->ntot);
        printf("TRLAN: ** reset maxlan to %d! **\n", info->maxlan);
    }
    if (info->maxlan < mev) {
        const int ntmp = DMIN(info->ntot, DMAX(100 + info->ned, 10 * info->ned));
        info->maxlan = mev < ntmp ? mev : ntmp;
    }
    if (info->maxlan < 5) {
        printf("TRLAN must have at least 5 basis vectors, it is currently %d.\n", info->maxlan);
        info->stat = -5;
    }
    info->tmv = -1;
    info->klan = DMIN(info->maxlan, info->ntot);
    if (info->restart >= 7)
        info->klan = DMIN(info->maxlan, DMAX(100, DMIN(info->klan, 2 * info->ned)));
    info->locked = info->nec;
    info->matvec = 0;
    info->nloop = 0;
    info->north = 0;
    info->nrand = 0;
    info->flop = 0;
    info->rflp = 0;
    info->flop_h = 0;
    info->rflp_h = 0;
    info->flop_r = 0;
    info->rflp_r = 0;
    info->tick_t = 0.0;
    info->clk_op = 0;
    info->tick_o = 0.0;
    info->clk_orth = 0;
    info->tick_h = 0.0;
    info->clk_res = 0;
    info->tick_r = 0.0;
    info->clk_in = 0;
    info->clk_out = 0;
    info->wrds_in = 0;
    info->wrds_out = 0;
    info->avgm = 0.0;
}

/* Shared entry of trlan()/ztrlan(): the reference's trlan()
 * (SRC/trlan.c:21-301) / ztrlan() (SRC/ztrlan.c:26-221) bookkeeping around
 * one solve on a fresh device basis. */
/* The reference's trlan() bookkeeping before trlanczos (SRC/trlan.c:
 * 21-301): argument checks, sizes, host state and a fresh device basis
 * (in batch b when given).  Returns 1 when there is an iteration to run;
 * 0 when done early or on error (info->stat). */
static int lz_setup(lz_t *z, trl_matvec op, ztrl_matvec zop, trl_uo user_ortho, trl_info *info,
                    int nrow, int mev, int lde, int cplx, trlhip_batch *b)
{
    memset(z, 0, sizeof *z);
    info->clk_tot = lz_clock();
    if (info->ned > mev) {
        printf("info->ned (%d) is larger than mev (%d) reducing info->ned to %d\n", info->ned,
               mev, mev);
        info->ned = mev;
    }
    if (info->ned <= info->nec || info->ned <= 0) return 0;
    info->stat = 0;
    check_and_clear(info, nrow, mev, lde);
    if (info->stat != 0) return 0;
    z->info = info;
    z->op = op;
    z->zop = zop;
    z->uo = user_ortho;
    z->cplx = cplx;
    z->E = cplx ? 2 : 1;
    z->n = nrow;
    z->mev = mev;
    z->maxlan = info->maxlan;
    z->nbas = DMAX(1, info->maxlan - mev + 1);
    z->ncols = DMAX(info->maxlan, mev) + 1;
    z->rng = 0x9E3779B97F4A7C15ull;
    z->A = trlhip_csr_from_op(cplx ? (void (*)(void))zop : (void (*)(void))op, info->mvparam,
                              cplx ? TRLHIP_C128 : TRLHIP_R64);
    z->mixed = !cplx && trlhip_csr_is_mixed(z->A);
    const int vt = cplx ? TRLHIP_C128 : (z->mixed ? TRLHIP_R32 : TRLHIP_R64);
    const int prof = g_prof && b; /* batched setups (several threads, atomic sums) */
    const long ta0 = prof ? lz_ns() : 0;
    if (lz_alloc(z, DMAX(info->maxlan, mev)) != 0) {
        info->stat = -4;
        lz_free(z);
        return 0;
    }
    const long ta1 = prof ? lz_ns() : 0;
    if (trlhip_ctx_create_batched(&z->dev, nrow, z->ncols, vt, b) != 0) {
        info->stat = TRLHIP_STAT_DEVICE;
        lz_free(z);
        return 0;
    }
    if (prof) {
        __atomic_add_fetch(&g_ns_alloc, ta1 - ta0, __ATOMIC_RELAXED);
        __atomic_add_fetch(&g_ns_ctx, lz_ns() - ta1, __ATOMIC_RELAXED);
    }
    /* batched real solves chain up to 8 inner steps per round (measured
     * on the batched 2D Anderson workload: 4 steps 8,344, 6: 8,560,
     * 8: 8,667, 12: 8,408, 16: 8,445 eigenpairs/s; longer chains make the
     * host/device overlap coarser); TRLHIP_BATCH_CHAIN=k: at most k, 0 or
     * 1: one step per round */
    const char *ec = getenv("TRLHIP_BATCH_CHAIN");
    z->chain = (b != NULL && z->A != NULL && !cplx) ? (ec ? atoi(ec) : 8) : 0;
    if (z->chain < 2) z->chain = 0;
    return 1;
}

static void lz_teardown(lz_t *z)
{
    trl_info *info = z->info;
    trlhip_ctx_destroy(z->dev);
    lz_free(z);
    const clock_t t_end = lz_clock();
    info->tick_t += (double)(t_end - info->clk_tot) / (double)info->clk_rate;
    info->clk_tot = 0;
}

/* Shared entry of trlan()/ztrlan(): the reference's trlan()
 * (SRC/trlan.c:21-301) / ztrlan() (SRC/ztrlan.c:26-221) bookkeeping around
 * one solve on a fresh device basis. */
void trl_solve_device(trl_matvec op, ztrl_matvec zop, trl_uo user_ortho, trl_info *info,
                      int nrow, int mev, double *eval, double *evec, int lde, int lwrk,
                      double *wrk, int cplx)
{
    lz_t z;
    if (!lz_setup(&z, op, zop, user_ortho, info, nrow, mev, lde, cplx, NULL)) return;
    /* the debug checks write to the log file named by trl_set_debug, with
     * the PE number appended, as the reference's trl_open_logfile does
     * (SRC/trlaux.c:53-84; opened for the solve, closed after it) */
    FILE *log = NULL;
    if (info->verbose > 6 && !info->log_fp && info->log_file[0]) {
        char name[160];
        snprintf(name, sizeof name, "%.128s%d", info->log_file, info->my_pe);
        log = fopen(name, "w");
        info->log_fp = log;
    }
    lz_iterate(&z, eval, evec, lde, wrk, lwrk);
    lz_teardown(&z);
    if (log) {
        fclose(log);
        info->log_fp = NULL;
    }
}

/* ------------------------------------------------------------- batches
 *
 * trlan_batch runs many trlan() solves together.  Each solve is the same
 * lz_iterate as above, run as a coroutine (ucontext) on its own stack;
 * wherever it must wait for the device it yields to the scheduler below.
 * Once every unfinished solve has yielded, the scheduler flushes the
 * batch: the recorded Lanczos steps of all solves run as one launch per
 * phase, the shared stream is synchronized once, and every solve resumes
 * with its results.  Each solve makes exactly the decisions it would make
 * alone, so its eigenpairs and counters match a single trlan() call.
 *
 * The solves are split into lanes (two by default), batches that queue on
 * one stream and wait on their own events: while the device runs one
 * lane's launch, the host threads step the other lane's solves, so host
 * and device work overlap. */

enum { LZ_STACK = 1024 * 1024 };  /* per coroutine; HIP runtime calls run on it */

typedef struct {
    ucontext_t uc;
    void *stack;
    lz_t z;
    trl_batch_job *job;
    int done;     /* lz_iterate returned */
    int ended;    /* torn down */
    int lane;     /* the batch (lane) it steps in */
    int home;     /* the scheduler thread that resumes it first */
} lz_co_t;

/* One host thread of the scheduler: it owns the coroutines i with
 * i % nthr == tid and resumes them every round. */
typedef struct lz_worker {
    ucontext_t sched;
    lz_co_t *cur;
    struct lz_pool *pool;
    int tid;
    double busy;  /* seconds spent resuming its coroutines */
    long resumes;
    double idle;  /* seconds between finishing its part of a round and the round's end */
    double rmax, tlong;  /* longest resume; time in resumes over 100 us */
    long nlong;
    int cpu;  /* CPU the thread is pinned to, or -1 */
} lz_worker_t;

/* Barrier for the scheduler threads.  Rounds come ~0.1-1 ms apart and the
 * threads are dedicated, so a waiting thread spins (pause) instead of
 * sleeping in the kernel (futex wake-ups cost ~0.1 ms a round with 16
 * threads); after ~10 ms of spinning it yields its core instead. */
typedef struct {
    int n;      /* participating threads */
    int count;  /* arrivals in the current generation */
    int gen;
} lz_spin_t;

static void lz_spin_init(lz_spin_t *b, int n)
{
    b->n = n;
    b->count = 0;
    b->gen = 0;
}

static void lz_spin_wait(lz_spin_t *b)
{
    const int gen = __atomic_load_n(&b->gen, __ATOMIC_ACQUIRE);
    if (__atomic_add_fetch(&b->count, 1, __ATOMIC_ACQ_REL) == b->n) {
        __atomic_store_n(&b->count, 0, __ATOMIC_RELAXED);
        __atomic_store_n(&b->gen, gen + 1, __ATOMIC_RELEASE);
        return;
    }
    for (long spins = 0; __atomic_load_n(&b->gen, __ATOMIC_ACQUIRE) == gen; ++spins) {
        if (spins < 200000)
            __builtin_ia32_pause();
        else
            sched_yield();
    }
}

/* A thread's share of the running lane's list: [start, start + len),
 * taken through the atomic index next (own cache line). */
typedef struct {
    int next, start, len;
    char pad[64 - 3 * sizeof(int)];
} lz_share_t;

typedef struct lz_pool {
    lz_co_t **lanes[4];  /* per lane: its unfinished coroutines, ordered by home */
    int nlane[4];
    lz_share_t *share;   /* per thread */
    lz_co_t *all;
    int njobs, nthr;
    lz_worker_t *w;
    lz_spin_t start, stop;
    volatile int quit;
    volatile int lane;  /* lane of the round in progress */
} lz_pool_t;

static __thread lz_worker_t *tl_worker;  /* the scheduler thread running a coroutine */

static double lz_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static __attribute__((noinline, noipa)) lz_worker_t *lz_cur_worker(void);

static void lz_co_yield(void *user)
{
    (void)user;
    lz_worker_t *w = lz_cur_worker();
    swapcontext(&w->cur->uc, &w->sched);
}

/* The scheduler thread running the calling coroutine.  A coroutine may be
 * resumed by another thread after each yield, so this reads the
 * thread-local afresh on every call (never a cached TLS address). */
static __attribute__((noinline, noipa)) lz_worker_t *lz_cur_worker(void) { return tl_worker; }

static void lz_co_main(void)
{
    lz_co_t *co = lz_cur_worker()->cur;
    trl_batch_job *j = co->job;
    lz_iterate(&co->z, j->eval, j->evec, j->lde, j->wrk, j->lwrk);
    co->done = 1;
    /* back to the scheduler for good (a finished coroutine is never
     * resumed; returning would follow uc_link instead) */
    for (;;) swapcontext(&co->uc, &lz_cur_worker()->sched);
}

/* Resume every unfinished coroutine of the current lane once (each runs
 * until it waits on the device or finishes).  A worker first takes the
 * coroutines whose home it is, so a coroutine's state stays in the caches
 * of the core that stepped it last round, then takes what is left of the
 * other workers' shares: a round ends when the work, not one thread's
 * share of it, is done. */
static void lz_round(lz_worker_t *w)
{
    tl_worker = w;
    const double t0 = lz_now();
    lz_pool_t *p = w->pool;
    lz_co_t **list = p->lanes[p->lane];
    for (int v = 0; v < p->nthr; ++v) {
        lz_share_t *sh = &p->share[(w->tid + v) % p->nthr];
        for (;;) {
            const int k = __atomic_fetch_add(&sh->next, 1, __ATOMIC_RELAXED);
            if (k >= sh->len) break;
            lz_co_t *co = list[sh->start + k];
            if (co->done) continue;
            w->cur = co;
            const double r0 = lz_now();
            swapcontext(&w->sched, &co->uc);
            const double dr = lz_now() - r0;
            ++w->resumes;
            if (dr > w->rmax) w->rmax = dr;
            if (dr > 1e-4) {
                w->tlong += dr;
                ++w->nlong;
            }
        }
    }
    w->busy += lz_now() - t0;
}

/* Shares of lane l's list (ordered by home) for the next round. */
static void lz_shares(lz_pool_t *p, int l)
{
    for (int t = 0; t < p->nthr; ++t) {
        p->share[t].next = 0;
        p->share[t].len = 0;
    }
    const int cnt = p->nlane[l];
    for (int k = 0; k < cnt; ++k) ++p->share[p->lanes[l][k]->home].len;
    int at = 0;
    for (int t = 0; t < p->nthr; ++t) {
        p->share[t].start = at;
        at += p->share[t].len;
    }
}

/* The CPUs this process may run on, in order (at most cap of them). */
static int lz_cpus(int *cpu, int cap)
{
    cpu_set_t set;
    CPU_ZERO(&set);
    if (sched_getaffinity(0, sizeof set, &set) != 0) return 0;
    int n = 0;
    for (int c = 0; c < CPU_SETSIZE && n < cap; ++c)
        if (CPU_ISSET(c, &set)) cpu[n++] = c;
    return n;
}

/* Pin the calling thread to one CPU (no-op on failure). */
static void lz_pin(int c)
{
    cpu_set_t set;
    CPU_ZERO(&set);
    CPU_SET(c, &set);
    pthread_setaffinity_np(pthread_self(), sizeof set, &set);
}

static void *lz_worker_main(void *arg)
{
    lz_worker_t *w = (lz_worker_t *)arg;
    if (w->cpu >= 0) lz_pin(w->cpu);
    for (;;) {
        lz_spin_wait(&w->pool->start);
        if (w->pool->quit) break;
        lz_round(w);
        const double tf = lz_now();
        lz_spin_wait(&w->pool->stop);
        w->idle += lz_now() - tf;
    }
    return NULL;
}

/* Lanes: TRLHIP_BATCH_LANES (1..4), else 2 (1 for a single job). */
static int lz_batch_lanes(int njobs)
{
    const char *e = getenv("TRLHIP_BATCH_LANES");
    long n = e ? strtol(e, NULL, 10) : 2;
    if (n < 1) n = 1;
    if (n > 4) n = 4;
    if (n > njobs) n = njobs;
    return (int)n;
}

/* Host threads for the coroutines: TRLHIP_BATCH_THREADS, else up to 16
 * (and no more than the online CPUs or the jobs). */
/* A coroutine's context on its stack, starting at lz_co_main (the context
 * getcontext saves is only a template: nothing returns to it). */
static __attribute__((noinline)) int lz_co_make(lz_co_t *co)
{
    if (getcontext(&co->uc) != 0) return -1;
    co->uc.uc_stack.ss_sp = co->stack;
    co->uc.uc_stack.ss_size = LZ_STACK;
    makecontext(&co->uc, lz_co_main, 0);
    return 0;
}

/* One thread's share of a batch's job setup: jobs t, t + nt, ... */
typedef struct {
    trl_batch_job *jobs;
    lz_co_t *cos;
    trlhip_batch **bl;
    int njobs, nlanes, t, nt;
    double t_su, t_st; /* profile: job setup, coroutine stacks */
} lz_setup_share_t;

static void *lz_setup_share(void *arg)
{
    lz_setup_share_t *a = (lz_setup_share_t *)arg;
    for (int i = a->t; i < a->njobs; i += a->nt) {
        trl_batch_job *j = &a->jobs[i];
        lz_co_t *co = &a->cos[i];
        co->job = j;
        co->done = 1;
        co->ended = 1;
        co->lane = i % a->nlanes;
        j->info->mvparam = j->A;
        const double ts0 = lz_now();
        const int ok = lz_setup(&co->z, trlhip_csr_matvec, NULL, NULL, j->info, j->nrow, j->mev,
                                j->lde, 0, a->bl[co->lane]);
        a->t_su += lz_now() - ts0;
        if (!ok) continue;
        co->z.dstart = j->dstart;
        co->ended = 0;
        const double ts1 = lz_now();
        co->stack = malloc(LZ_STACK);
        a->t_st += lz_now() - ts1;
        if (!co->stack || lz_co_make(co) != 0) {
            j->info->stat = -4;
            continue;
        }
        co->done = 0;
    }
    return NULL;
}

/* One thread's share of a batch's teardown: jobs t, t + nt, ... */
static void *lz_teardown_share(void *arg)
{
    lz_setup_share_t *a = (lz_setup_share_t *)arg;
    for (int i = a->t; i < a->njobs; i += a->nt) {
        lz_co_t *co = &a->cos[i];
        if (!co->ended) lz_teardown(&co->z);
        free(co->stack);
    }
    return NULL;
}

/* Runs fn over jobs 0..njobs-1 on nt threads (shares t, t + nt, ...); a
 * share whose thread cannot start runs on the calling thread.  Returns the
 * summed profile times in *t_su, *t_st (may be NULL). */
static void lz_run_shares(void *(*fn)(void *), trl_batch_job *jobs, lz_co_t *cos,
                          trlhip_batch **bl, int njobs, int nlanes, int nt, double *t_su,
                          double *t_st)
{
    enum { LZ_SHARES_MAX = 64 };
    if (nt > LZ_SHARES_MAX) nt = LZ_SHARES_MAX;
    if (nt < 1) nt = 1;
    lz_setup_share_t sh[LZ_SHARES_MAX];
    pthread_t sth[LZ_SHARES_MAX];
    int started[LZ_SHARES_MAX];
    for (int t = 0; t < nt; ++t) {
        sh[t] = (lz_setup_share_t){jobs, cos, bl, njobs, nlanes, t, nt, 0.0, 0.0};
        started[t] = t > 0 && pthread_create(&sth[t], NULL, fn, &sh[t]) == 0;
    }
    for (int t = 0; t < nt; ++t)
        if (!started[t]) fn(&sh[t]);
    for (int t = 1; t < nt; ++t)
        if (started[t]) pthread_join(sth[t], NULL);
    for (int t = 0; t < nt; ++t) {
        if (t_su) *t_su += sh[t].t_su;
        if (t_st) *t_st += sh[t].t_st;
    }
}

static int lz_batch_threads(int njobs)
{
    const char *e = getenv("TRLHIP_BATCH_THREADS");
    long n = e ? strtol(e, NULL, 10) : 0;
    if (n <= 0) {
        n = sysconf(_SC_NPROCESSORS_ONLN);
        if (n > 16) n = 16;
    }
    if (n < 1) n = 1;
    if (n > njobs) n = njobs;
    return (int)n;
}

int trlan_batch(trl_batch_job *jobs, int njobs)
{
    if (njobs <= 0) return 0;
    if (!jobs) return TRLHIP_EINVAL;
    const double t_begin = lz_now();
    /* TRLHIP_BATCH_PROFILE=1: host time in the setup, in the solves and
     * waiting for the device */
    const char *prof = getenv("TRLHIP_BATCH_PROFILE");
    g_prof = prof && prof[0] == '1';
    g_ns_restart = g_ns_check = g_ns_tridiag = g_n_restart = 0;
    g_ns_alloc = g_ns_ctx = 0;
    const int nlanes = lz_batch_lanes(njobs);
    trlhip_batch *bl[4] = {NULL, NULL, NULL, NULL};
    int rc = 0;
    /* TRLHIP_BATCH_STREAMS: unset or 2, a stream per lane with each launch
     * ordered after the other lane's (the device runs one lane at a time,
     * so every kernel's timing is its own); 1, streams that may also run
     * two lanes' kernels at once; 0, one shared stream */
    const char *es = getenv("TRLHIP_BATCH_STREAMS");
    const int mode = es && es[0] >= '0' && es[0] <= '2' ? es[0] - '0' : 2;
    const int own = mode != 0;
    volatile int serial = mode == 2;
    for (int l = 0; l < nlanes && !rc; ++l) {
        rc = trlhip_batch_create_on(&bl[l], (l > 0 && !own) ? bl[0] : NULL);
        if (!rc) trlhip_batch_set_yield(bl[l], lz_co_yield, NULL);
    }
    lz_co_t *cos = rc ? NULL : (lz_co_t *)calloc((size_t)njobs, sizeof(lz_co_t));
    if (!cos) {
        for (int i = 0; i < njobs; ++i) jobs[i].info->stat = TRLHIP_STAT_DEVICE;
        for (int l = nlanes - 1; l >= 0; --l) trlhip_batch_destroy(bl[l]); /* lane 0 owns the stream */
        return rc ? rc : TRLHIP_ENOMEM;
    }
    int live[4] = {0, 0, 0, 0};
    double t_su = 0.0, t_st = 0.0;  /* profile: job setup, coroutine stacks */
    const double t_jobs = lz_now();
    /* the jobs' setups (host arrays, device contexts from the lanes'
     * arenas, coroutine stacks) on as many threads as the scheduler's */
    lz_run_shares(lz_setup_share, jobs, cos, bl, njobs, nlanes, lz_batch_threads(njobs), &t_su,
                  &t_st);
    for (int i = 0; i < njobs; ++i)
        if (!cos[i].done) ++live[cos[i].lane];
    const double t_pool = lz_now();
    lz_pool_t pool;
    memset(&pool, 0, sizeof pool);
    pool.all = cos;
    pool.njobs = njobs;
    pool.nthr = lz_batch_threads(njobs);
    pool.w = (lz_worker_t *)calloc((size_t)pool.nthr, sizeof(lz_worker_t));
    pool.share = (lz_share_t *)aligned_alloc(64, sizeof(lz_share_t) * (size_t)pool.nthr);
    pthread_t *th = (pthread_t *)calloc((size_t)pool.nthr, sizeof(pthread_t));
    /* TRLHIP_BATCH_PIN (default 1): pin the scheduler threads, this one
     * included until the batch returns, to distinct CPUs of the process's
     * set when it has enough of them */
    const char *epin = getenv("TRLHIP_BATCH_PIN");
    int cpus[64];
    const int ncpu = (epin && epin[0] == '0') ? 0 : lz_cpus(cpus, 64);
    const int pin = pool.nthr > 1 && ncpu >= pool.nthr;
    cpu_set_t main_mask;
    const int main_saved = pin && sched_getaffinity(0, sizeof main_mask, &main_mask) == 0;
    int nstarted = 0;
    if (pool.w && th && pool.share) {
        lz_spin_init(&pool.start, pool.nthr);
        lz_spin_init(&pool.stop, pool.nthr);
        for (int t = 0; t < pool.nthr; ++t) {
            pool.w[t].pool = &pool;
            pool.w[t].tid = t;
            pool.w[t].cpu = pin ? cpus[t] : -1;
        }
        if (main_saved) lz_pin(cpus[0]);
        for (int t = 1; t < pool.nthr; ++t)
            if (pthread_create(&th[t], NULL, lz_worker_main, &pool.w[t]) == 0) ++nstarted;
    }
    if (!pool.w || !th || !pool.share || nstarted != pool.nthr - 1) {
        /* could not start the threads: stop any started, run on this one */
        if (nstarted > 0) {
            pool.quit = 1;
            lz_spin_init(&pool.start, nstarted + 1);  /* only those threads wait */
            lz_spin_wait(&pool.start);
            for (int t = 1; t <= nstarted; ++t) pthread_join(th[t], NULL);
        }
        free(th);
        th = NULL;
        if (!pool.w) pool.w = (lz_worker_t *)calloc(1, sizeof(lz_worker_t));
        if (!pool.share) pool.share = (lz_share_t *)aligned_alloc(64, sizeof(lz_share_t));
        pool.nthr = 1;
        pool.w[0].pool = &pool;
        pool.w[0].tid = 0;
        pool.w[0].cpu = -1;
    }
    /* each lane's list of unfinished coroutines */
    int lists_ok = 1;
    for (int l = 0; l < nlanes; ++l) {
        pool.lanes[l] = (lz_co_t **)calloc((size_t)njobs, sizeof(lz_co_t *));
        if (!pool.lanes[l]) lists_ok = 0;
    }
    if (!lists_ok) {  /* no solve runs */
        for (int i = 0; i < njobs; ++i)
            if (!cos[i].done) {
                cos[i].done = 1;
                jobs[i].info->stat = -4;
            }
        for (int l = 0; l < nlanes; ++l) live[l] = 0;
        rc = TRLHIP_ENOMEM;
    } else {
        for (int i = 0; i < njobs; ++i)
            if (!cos[i].done) pool.lanes[cos[i].lane][pool.nlane[cos[i].lane]++] = &cos[i];
        /* homes: contiguous runs of each lane's list, so it is ordered by
         * home and stays so as finished coroutines drop out */
        for (int l = 0; l < nlanes; ++l)
            for (int k = 0; k < pool.nlane[l]; ++k)
                pool.lanes[l][k]->home = (int)((long)k * pool.nthr / pool.nlane[l]);
    }
    double t_co = 0.0, t_wait = 0.0;
    long rounds = 0;
    const double t_loop = lz_now();
    int pending[4] = {0, 0, 0, 0};
    /* profile: wall time of lane rounds by the number of solves they step */
    double t_band[4] = {0, 0, 0, 0}, w_band[4] = {0, 0, 0, 0}, l_band[4] = {0, 0, 0, 0};
    long n_band[4] = {0, 0, 0, 0};
    int last_band[4] = {0, 0, 0, 0};
    /* When the host is the bottleneck (the next lane's last launch has
     * already finished), a lane's recorded work is launched while the
     * workers already step the next lane: this thread issues the launch
     * (about a dozen HIP calls) after the round's start barrier and then
     * joins the round.  When the device is (that launch still runs), the
     * launch is issued before waiting, so the device never runs dry.
     * Launches stay in lane order, each after the one before it. */
    int deferred = -1;
#define LZ_LAUNCH(L)                                                                   \
    do {                                                                               \
        const double tl_ = lz_now();                                                   \
        if (serial && nlanes > 1) {                                                    \
            const int orc_ = trlhip_batch_after(bl[L], bl[((L) + nlanes - 1) % nlanes]); \
            if (orc_ && !rc) rc = orc_;                                                \
        }                                                                              \
        const int lrc_ = trlhip_batch_launch(bl[L]);                                   \
        if (lrc_ && !rc) rc = lrc_;                                                    \
        pending[L] = 1;                                                                \
        l_band[last_band[L]] += lz_now() - tl_;                                        \
    } while (0)
    for (;;) {
        int any = 0;
        for (int l = 0; l < nlanes; ++l) {
            if (deferred == l || (deferred >= 0 && pending[l] && !trlhip_batch_ready(bl[l]))) {
                LZ_LAUNCH(deferred);
                deferred = -1;
            }
            if (pending[l]) {
                /* the lane's last launch must finish before its solves resume */
                const double tw = lz_now();
                const int wrc = trlhip_batch_wait(bl[l]);
                if (wrc && !rc) rc = wrc;
                pending[l] = 0;
                t_wait += lz_now() - tw;
                w_band[last_band[l]] += lz_now() - tw;
            }
            if (live[l] == 0) continue;
            any = 1;
            const int active = live[l];
            const double t0 = lz_now();
            pool.lane = l;
            lz_shares(&pool, l);
            if (pool.nthr > 1) lz_spin_wait(&pool.start);
            if (deferred >= 0) {
                LZ_LAUNCH(deferred);
                deferred = -1;
            }
            lz_round(&pool.w[0]);
            const double tf = lz_now();
            if (pool.nthr > 1) lz_spin_wait(&pool.stop);
            pool.w[0].idle += lz_now() - tf;
            /* drop the finished ones from the lane's list */
            int m = 0;
            for (int k = 0; k < pool.nlane[l]; ++k)
                if (!pool.lanes[l][k]->done) pool.lanes[l][m++] = pool.lanes[l][k];
            pool.nlane[l] = m;
            live[l] = m;
            t_co += lz_now() - t0;
            ++rounds;
            const int band = active >= 256 ? 0 : active >= 32 ? 1 : active >= 4 ? 2 : 3;
            if (live[l] > 0) deferred = l;
            t_band[band] += lz_now() - t0;
            ++n_band[band];
            last_band[l] = band;
        }
        if (!any) break;
    }
#undef LZ_LAUNCH
    if (pool.nthr > 1) {
        pool.quit = 1;
        lz_spin_wait(&pool.start);
        for (int t = 1; t < pool.nthr; ++t) pthread_join(th[t], NULL);
    }
    if (prof && prof[0] == '1') {
        double bsum = 0.0, bmax = 0.0, isum = 0.0, rmax = 0.0, tlong = 0.0;
        long res = 0, nlong = 0;
        for (int t = 0; t < pool.nthr; ++t) {
            bsum += pool.w[t].busy;
            bmax = pool.w[t].busy > bmax ? pool.w[t].busy : bmax;
            res += pool.w[t].resumes;
            isum += pool.w[t].idle;
            rmax = pool.w[t].rmax > rmax ? pool.w[t].rmax : rmax;
            tlong += pool.w[t].tlong;
            nlong += pool.w[t].nlong;
        }
        fprintf(stderr, "trlan_batch: lane rounds by solves stepped (count, round s, its lane's "
                "launch s, wait s): >=256 %ld %.3f %.3f %.3f; 32-255 %ld %.3f %.3f %.3f; 4-31 %ld "
                "%.3f %.3f %.3f; 1-3 %ld %.3f %.3f %.3f\n",
                n_band[0], t_band[0], l_band[0], w_band[0], n_band[1], t_band[1], l_band[1],
                w_band[1], n_band[2], t_band[2], l_band[2], w_band[2], n_band[3], t_band[3],
                l_band[3], w_band[3]);
        fprintf(stderr, "trlan_batch: threads pinned: %s (%d CPUs in the process's set)\n",
                pin ? "yes" : "no", ncpu);
        /* the solves' own timers: restarts (host eigenproblems, recorded
         * GEMMs, final refinement) and orthogonalization decisions */
        double sres = 0.0, sorth = 0.0;
        for (int i = 0; i < njobs; ++i) {
            const trl_info *in = jobs[i].info;
            const double r = in->clk_rate > 0 ? (double)in->clk_rate : (double)CLOCKS_PER_SEC;
            sres += (double)in->clk_res / r + in->tick_r;
            sorth += (double)in->clk_orth / r + in->tick_h;
        }
        fprintf(stderr, "trlan_batch: summed over solves: restart %.3f s, orthogonalization %.3f s\n",
                sres, sorth);
        fprintf(stderr, "trlan_batch: host work summed over threads: %ld restarts %.3f s "
                "(%.1f us each), convergence checks %.3f s, arrowhead reductions %.3f s\n",
                g_n_restart, 1e-9 * g_ns_restart, g_n_restart ? 1e-3 * g_ns_restart / g_n_restart : 0.0,
                1e-9 * g_ns_check, 1e-9 * g_ns_tridiag);
        fprintf(stderr, "trlan_batch: idle at round ends mean %.3f s; longest resume %.1f us; "
                "%ld resumes over 100 us take %.3f s\n", isum / pool.nthr, 1e6 * rmax, nlong, tlong);
        fprintf(stderr,
                "trlan_batch: %d host threads, %d lanes, %ld lane rounds, %.3f s in solves, "
                "%.3f s waiting for the device; %ld resumes, thread busy mean %.3f max %.3f s "
                "(%.1f us per resume)\n",
                pool.nthr, nlanes, rounds, t_co, t_wait, res, bsum / pool.nthr, bmax,
                res ? 1e6 * bsum / res : 0.0);
    }
    const double t_end_loop = lz_now();
    /* teardown after the last round (the jobs' host arrays and stacks;
     * their device memory is the lanes' arenas), on the same threads */
    lz_run_shares(lz_teardown_share, jobs, cos, bl, njobs, nlanes, lz_batch_threads(njobs), NULL,
                  NULL);
    free(th);
    free(pool.w);
    free(pool.share);
    if (main_saved) sched_setaffinity(0, sizeof main_mask, &main_mask);
    for (int l = 0; l < nlanes; ++l) free(pool.lanes[l]);
    free(cos);
    for (int l = nlanes - 1; l >= 0; --l) trlhip_batch_destroy(bl[l]); /* lane 0 owns the stream */
    if (prof && prof[0] == '1')
        fprintf(stderr, "trlan_batch: setup %.3f s (lanes %.3f, jobs %.3f of which job setup %.3f "
                "[host arrays %.3f, device contexts %.3f] and stacks %.3f), rounds %.3f s, "
                "teardown %.3f s\n",
                t_loop - t_begin, t_jobs - t_begin, t_pool - t_jobs, t_su, 1e-9 * g_ns_alloc,
                1e-9 * g_ns_ctx, t_st,
                t_end_loop - t_loop, lz_now() - t_end_loop);
    return rc;
}

void trlan(trl_matvec op, trl_uo user_ortho, trl_info *info, int nrow, int mev, double *eval,
           double *evec, int lde, int lwrk, double *wrk)
{
    trl_solve_device(op, NULL, user_ortho, info, nrow, mev, eval, evec, lde, lwrk, wrk, 0);
}
