/*
 * trl_kit.c -- host FP64 kit of the thick-restart Lanczos driver.
 *
 * Dense eigen-problems of size <= maxlan (implicit QL, Householder),
 * paired Shell sorts, and the reference's convergence / restart / locking
 * decisions.  The decisions are restated from the reference so the solver
 * makes the same choices (same kept sets, same matvec counts); the dense
 * linear algebra is this project's own (no LAPACK dependency).
 */
#include "trl_kit.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define KMIN(a, b) ((a) < (b) ? (a) : (b))
#define KMAX(a, b) ((a) > (b) ? (a) : (b))

/* ------------------------------------------------------------ sorting */

static double sort_key(int key, double s, double x, double maxd)
{
    double v, d;
    switch (key) {
    case TRLKIT_BY_ABS:
        return fabs(x);
    case TRLKIT_BY_DIST:
        return fabs(x - s);
    case TRLKIT_ABOVE: /* reference dsort2su_, SRC/dsort2.c:179-249 */
        v = fabs(x);
        d = v - s;
        return d < 0.0 ? maxd + v : d;
    case TRLKIT_BELOW: /* reference dsort2sd, SRC/dsort2.c:252-323 */
        v = fabs(x);
        d = s - v;
        return d < 0.0 ? s + (maxd - v) : d;
    default:
        return x;
    }
}

/* Shell sort with gaps n/2, n/4, ..., 1 (the reference's dsort2 family,
 * SRC/dsort2.c): tie order depends on the gap sequence, so it is kept. */
void trlkit_sort2(int n, int key, double s, double *a1, double *a2, int *tag)
{
    double maxd = 0.0;
    if (n <= 0) return;
    if (key == TRLKIT_ABOVE || key == TRLKIT_BELOW) {
        maxd = fabs(a1[0]);
        for (int i = 1; i < n; ++i)
            if (maxd < fabs(a1[i])) maxd = fabs(a1[i]);
        if (key == TRLKIT_BELOW) maxd += 1.0;
    }
    for (int gap = n / 2; gap > 0; gap /= 2) {
        for (int i = gap; i < n; ++i) {
            for (int j = i - gap; j >= 0; j -= gap) {
                if (!(sort_key(key, s, a1[j], maxd) > sort_key(key, s, a1[j + gap], maxd)))
                    break;
                double t = a1[j];
                a1[j] = a1[j + gap];
                a1[j + gap] = t;
                t = a2[j];
                a2[j] = a2[j + gap];
                a2[j + gap] = t;
                if (tag) {
                    int u = tag[j];
                    tag[j] = tag[j + gap];
                    tag[j + gap] = u;
                }
            }
        }
    }
}

/* ----------------------------------------------------- dense eigen kit */

/* Implicit QL, Wilkinson shift, on (d, e); see trl_kit.h.  The eigenvalue
 * iteration is independent of `want`, so eigenvalues computed with and
 * without vectors are bitwise identical. */
int trlkit_tridiag_ql(int n, double *d, double *e, double *z, int want)
{
    if (n <= 0) return 0;
    if (n == 1) {
        if (want == 1) z[0] = 1.0;
        return 0;
    }
    /* e[i] couples d[i], d[i+1]; e[n-1] is a zero sentinel */
    double *ee = (double *)malloc(sizeof(double) * (size_t)n);
    if (!ee) return n;
    memcpy(ee, e, sizeof(double) * (size_t)(n - 1));
    ee[n - 1] = 0.0;
    if (want == 1) {
        for (int i = 0; i < n; ++i) z[i] = 0.0;
        z[n - 1] = 1.0;
    }
    const double eps = DBL_EPSILON;
    int fail = 0;
    for (int l = 0; l < n && !fail; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m < n - 1; ++m) {
                const double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(ee[m]) <= eps * dd) break;
            }
            if (m == l) break;
            if (++iter > 30 * n) {
                fail = l + 1;
                break;
            }
            /* shift from the leading 2x2 block, Wilkinson style */
            double g = (d[l + 1] - d[l]) / (2.0 * ee[l]);
            double r = hypot(g, 1.0);
            g = d[m] - d[l] + ee[l] / (g + (g >= 0.0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = m - 1;
            int early = 0;
            for (; i >= l; --i) {
                const double f = s * ee[i];
                const double b = c * ee[i];
                r = hypot(f, g);
                ee[i + 1] = r;
                if (r == 0.0) { /* underflow: deflate and retry */
                    d[i + 1] -= p;
                    ee[m] = 0.0;
                    early = 1;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                if (want == 1) {
                    const double t = z[i + 1];
                    z[i + 1] = s * z[i] + c * t;
                    z[i] = c * z[i] - s * t;
                } else if (want == 2) {
                    double *zi = z + (size_t)i * n, *zj = z + (size_t)(i + 1) * n;
                    for (int k = 0; k < n; ++k) {
                        const double t = zj[k];
                        zj[k] = s * zi[k] + c * t;
                        zi[k] = c * zi[k] - s * t;
                    }
                }
            }
            if (early) continue;
            d[l] -= p;
            ee[l] = g;
            ee[m] = 0.0;
        }
    }
    free(ee);
    if (fail) return fail;
    /* selection sort ascending, carrying vectors */
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        double p = d[i];
        for (int j = i + 1; j < n; ++j)
            if (d[j] < p) {
                k = j;
                p = d[j];
            }
        if (k != i) {
            d[k] = d[i];
            d[i] = p;
            if (want == 1) {
                const double t = z[i];
                z[i] = z[k];
                z[k] = t;
            } else if (want == 2) {
                double *zi = z + (size_t)i * n, *zk = z + (size_t)k * n;
                for (int q = 0; q < n; ++q) {
                    const double t = zi[q];
                    zi[q] = zk[q];
                    zk[q] = t;
                }
            }
        }
    }
    return 0;
}

/* Bottom-up Householder tridiagonalization (see trl_kit.h). */
void trlkit_householder(int n, double *a, double *d, double *e)
{
#define A_(i, j) a[(size_t)(j) * n + (i)]
    if (n <= 0) return;
    double *Q = (double *)calloc((size_t)n * n, sizeof(double));
    double *v = (double *)malloc(sizeof(double) * (size_t)n);
    double *p = (double *)malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; ++i) Q[(size_t)i * n + i] = 1.0;
    /* symmetrize from the upper triangle */
    for (int j = 0; j < n; ++j)
        for (int i = j + 1; i < n; ++i) A_(i, j) = A_(j, i);
    for (int k = n - 1; k >= 2; --k) {
        /* annihilate A(0:k-2, k), keep A(k-1, k) */
        const double alpha = A_(k - 1, k);
        double sigma = 0.0;
        for (int i = 0; i < k - 1; ++i) sigma += A_(i, k) * A_(i, k);
        if (sigma == 0.0) {
            e[k - 1] = alpha;
            continue;
        }
        const double mu = sqrt(alpha * alpha + sigma);
        const double beta = (alpha <= 0.0) ? alpha - mu : -sigma / (alpha + mu);
        const double tau = 2.0 * beta * beta / (sigma + beta * beta);
        for (int i = 0; i < k - 1; ++i) v[i] = A_(i, k) / beta;
        v[k - 1] = 1.0;
        /* H x = mu e_{k-1} with H = I - tau v v^T */
        e[k - 1] = mu;
        A_(k - 1, k) = mu;
        A_(k, k - 1) = mu;
        for (int i = 0; i < k - 1; ++i) {
            A_(i, k) = 0.0;
            A_(k, i) = 0.0;
        }
        /* leading k x k block: A <- H A H = A - v w^T - w v^T */
        double K = 0.0;
        for (int i = 0; i < k; ++i) {
            double s = 0.0;
            for (int j = 0; j < k; ++j) s += A_(i, j) * v[j];
            p[i] = tau * s;
            K += v[i] * p[i];
        }
        K *= 0.5 * tau;
        for (int i = 0; i < k; ++i) p[i] -= K * v[i];
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < k; ++i) A_(i, j) -= v[i] * p[j] + p[i] * v[j];
        /* Q <- Q H on columns 0..k-1 */
        for (int r = 0; r < n; ++r) {
            double s = 0.0;
            for (int j = 0; j < k; ++j) s += Q[(size_t)j * n + r] * v[j];
            s *= tau;
            for (int j = 0; j < k; ++j) Q[(size_t)j * n + r] -= s * v[j];
        }
    }
    if (n >= 2) e[0] = A_(0, 1);
    for (int i = 0; i < n; ++i) d[i] = A_(i, i);
    memcpy(a, Q, sizeof(double) * (size_t)n * n);
    free(Q);
    free(v);
    free(p);
#undef A_
}

int trlkit_syev(int n, double *a, double *w)
{
    if (n <= 0) return 0;
    double *e = (double *)malloc(sizeof(double) * (size_t)(n > 1 ? n : 1));
    trlkit_householder(n, a, w, e);
    const int rc = trlkit_tridiag_ql(n, w, e, a, 2);
    free(e);
    return rc;
}

/* --------------------------------------------- convergence and restart */

/* Convergence bookkeeping after a Ritz-value update: the reference's
 * trl_convergence_test (SRC/trlcore.c:2149-2305). */
void trlkit_convergence_test(int nd, double *lambda, double *res, int *tag,
                             trl_info *info)
{
    double *w = (double *)malloc(sizeof(double) * (size_t)(2 * nd + 2));
    trlkit_convergence_test_w(nd, lambda, res, tag, info, w);
    free(w);
}

void trlkit_convergence_test_w(int nd, double *lambda, double *res, int *tag,
                               trl_info *info, double *wrk)
{
    /* the reference's layout of its workspace: |res| then lambda, both in
     * the sorted order (SRC/trlcore.c:2181-2201) */
    double *wr = wrk, *wl = wrk + nd;
    if (info->lohi < -1) trlkit_sort2(nd, TRLKIT_BY_DIST, info->ref, lambda, res, tag);
    for (int i = 0; i < nd; ++i) {
        wl[i] = lambda[i];
        wr[i] = fabs(res[i]);
    }
    int key = TRLKIT_BY_VALUE;
    if (info->lohi == -2) key = TRLKIT_BY_DIST;
    else if (info->lohi == -3) key = TRLKIT_ABOVE;
    else if (info->lohi == -4) key = TRLKIT_BELOW;
    trlkit_sort2(nd, key, info->ref, wl, wr, NULL);

    /* convergence rate of the previous target */
    if (info->tmv > 0 && info->matvec > info->tmv) {
        int j = 0;
        double bnd = fabs(lambda[0] - info->trgt);
        for (int i = 0; i < nd; ++i)
            if (fabs(lambda[i] - info->trgt) < bnd) {
                bnd = fabs(lambda[i] - info->trgt);
                j = i;
            }
        if (info->tres > res[j]) {
            bnd = res[j] / info->tres;
            info->crat = bnd > 0.0 ? exp(log(bnd) / (double)(info->matvec - info->tmv)) : 1.0;
            info->cfac = bnd;
        } else {
            info->crat = 1.0;
            info->cfac = bnd;
        }
    }
    /* ||A|| estimate uses the second-smallest and the largest Ritz value
     * (reference trlcore.c:2230-2231) */
    info->anrm = KMAX(info->anrm, KMAX(fabs(wl[1]), fabs(wl[nd - 1])));
    const double bnd = DBL_MIN + info->tol * info->anrm;
    int ncl = 0, ncr = nd;
    if (info->lohi <= 0) {
        ncl = nd - 1;
        for (int i = 0; i < nd; ++i) {
            if (wr[i] < bnd) {
                if ((info->lohi == -3 && wl[i] < info->ref) ||
                    (info->lohi == -4 && wl[i] > info->ref)) {
                    ncl = i - 1;
                    break;
                }
            } else {
                ncl = i - 1;
                break;
            }
        }
    }
    if (info->lohi >= 0) {
        ncr = 0;
        for (int i = nd - 1; i >= 0; --i)
            if (!(wr[i] < bnd)) {
                ncr = i + 1;
                break;
            }
    }
    info->tmv = info->matvec;
    info->ptres = info->trgt;
    if (info->lohi < 0) {
        info->nec = ncl + 1;
        info->trgt = wl[KMIN(nd - 1, ncl + 1)];
        info->tres = wr[KMIN(nd - 1, ncl + 1)];
    } else if (info->lohi > 0) {
        info->nec = nd - ncr;
        info->trgt = wl[KMAX(0, ncr - 1)];
        info->tres = wr[KMAX(0, ncr - 1)];
    } else {
        if (ncr <= ncl) {
            ncl = nd / 2;
            ncr = ncl + 1;
            info->trgt = wl[(nd + 1) / 2 - 1];
            info->tres = wr[(nd + 1) / 2 - 1];
        } else if (wr[ncl + 1] <= wr[ncr - 1]) {
            info->trgt = wl[ncl + 1];
            info->tres = wr[ncl + 1];
        } else {
            info->trgt = wl[ncr - 1];
            info->tres = wr[ncr - 1];
        }
        info->nec = ncl + nd - ncr + 1;
        for (int i = ncl + 1; i < ncr; ++i)
            if (wr[i] < bnd) info->nec++;
    }
}

/* Final output order: the reference's trl_sort_eig (SRC/trlcore.c:2308). */
void trlkit_sort_eig(int nd, int lohi, int nec, double ref, double *lambda,
                     double *res, int *tag)
{
    if (lohi == 0) {
        trlkit_sort2(nd, TRLKIT_BY_ABS, 0.0, res, lambda, tag);
        trlkit_sort2(nec, TRLKIT_BY_VALUE, 0.0, lambda, res, tag);
    } else if (lohi == -2 || lohi == -3 || lohi == -4) {
        const int key = lohi == -2 ? TRLKIT_BY_DIST : (lohi == -3 ? TRLKIT_ABOVE : TRLKIT_BELOW);
        trlkit_sort2(nd, key, ref, lambda, res, tag);
        trlkit_sort2(nec, TRLKIT_BY_VALUE, 0.0, lambda, res, tag);
    } else {
        trlkit_sort2(nd, TRLKIT_BY_VALUE, 0.0, lambda, res, tag);
        if (lohi > 0) { /* largest nec to the front, still ascending */
            for (int i = 0, j = nd - nec; i < nec; ++i, ++j) {
                res[i] = res[j];
                lambda[i] = lambda[j];
                if (tag) tag[i] = tag[j];
            }
        }
    }
}

/* Move pairs with tiny residuals to the front for locking: the reference's
 * trl_set_locking (SRC/trlcore.c:2676-2758).  yy is jnd x nlam. */
int trlkit_set_locking(int jnd, int nlam, double *lambda, double *res,
                       double *yy, double anrm)
{
    const double eps = DBL_EPSILON;
#define TINY(x) (fabs(x) >= eps ? eps * fabs(x) : eps * eps * anrm)
    int i = 0, j = nlam - 1;
    int ti = fabs(res[i]) <= TINY(lambda[i]);
    int tj = fabs(res[j]) <= TINY(lambda[j]);
    while (i < j) {
        if (ti) {
            res[i] = 0.0;
            ++i;
            ti = (i <= j) ? (fabs(res[i]) <= TINY(lambda[i])) : 0;
        } else {
            if (tj) {
                double t = lambda[i];
                lambda[i] = lambda[j];
                lambda[j] = t;
                res[j] = res[i];
                res[i] = 0.0;
                double *ci = yy + (size_t)i * jnd, *cj = yy + (size_t)j * jnd;
                for (int q = 0; q < jnd; ++q) {
                    t = ci[q];
                    ci[q] = cj[q];
                    cj[q] = t;
                }
                ++i;
                ti = (i <= j) ? (fabs(res[i]) <= TINY(lambda[i])) : 0;
            }
            --j;
            tj = (j > i) ? (fabs(res[j]) <= TINY(lambda[j])) : 0;
        }
    }
#undef TINY
    return ti ? i + 1 : i;
}

void trlkit_smooth_rr(int n, double *rr)
{
    if (n <= 0) return;
    double prev = rr[0];
    rr[0] = 2 * rr[0] + rr[2 < n ? 2 : 0] + rr[n - 1];
    for (int i = 1; i < n - 1; ++i) {
        const double t = rr[i];
        rr[i] = 2 * rr[i] + rr[i + 1] + prev;
        prev = t;
    }
    rr[n - 1] = 2 * rr[n - 1] + rr[1 < n ? 1 : 0] + prev;
}

/* The reference's trl_smooth_zz (SRC/ztrlcore.c:1289-1332) on interleaved
 * complex z[0:2n): as trl_smooth_rr per part, except that the last
 * element's imaginary part adds twice its real part as just updated
 * (ztrlcore.c:1328) instead of twice its own imaginary part.  Indices past
 * the end of a vector shorter than 3 wrap (the reference reads beyond it). */
void trlkit_smooth_zz(int n, double *z)
{
    if (n <= 0) return;
    double r2 = z[0], i2 = z[1];
    z[0] = 2 * z[0] + z[2 * (2 % n)] + z[2 * (n - 1)];
    z[1] = 2 * z[1] + z[2 * (2 % n) + 1] + z[2 * (n - 1) + 1];
    for (int i = 1; i < n - 1; ++i) {
        const double r1 = z[2 * i], i1 = z[2 * i + 1];
        z[2 * i] = 2 * z[2 * i] + z[2 * (i + 1)] + r2;
        z[2 * i + 1] = 2 * z[2 * i + 1] + z[2 * (i + 1) + 1] + i2;
        r2 = r1;
        i2 = i1;
    }
    const int l = n - 1;
    z[2 * l] = 2 * z[2 * l] + z[2 * (1 % n)] + r2;
    z[2 * l + 1] = 2 * z[2 * l] + z[2 * (1 % n) + 1] + i2;
}

/* --------------------------------------------------------- checkpoints */

static int ck_close(FILE *f, int ok_code, int err_code)
{
    return fclose(f) == 0 ? ok_code : err_code;
}

int trlkit_write_checkpoint(const char *file, int nrow, int jnd, const double *alpha,
                            const double *beta, int ncol,
                            int (*col)(void *u, int i, double *buf), void *u, double *buf)
{
    return trlkit_write_checkpoint_e(file, nrow, 1, jnd, alpha, beta, ncol, col, u, buf);
}

int trlkit_write_checkpoint_e(const char *file, int nrow, int E, int jnd, const double *alpha,
                              const double *beta, int ncol,
                              int (*col)(void *u, int i, double *buf), void *u, double *buf)
{
    const size_t cw = (size_t)nrow * E;
    FILE *f = fopen(file, "w");
    if (!f) {
        printf("TRL_WRITE_CHECKPOINT: failed to open file: %s.\n", file);
        return -221;
    }
    if (fwrite(&nrow, sizeof nrow, 1, f) != 1 || fwrite(&jnd, sizeof jnd, 1, f) != 1 ||
        (jnd > 0 && fwrite(alpha, sizeof(double), (size_t)jnd, f) != (size_t)jnd) ||
        (jnd > 0 && fwrite(beta, sizeof(double), (size_t)jnd, f) != (size_t)jnd))
        return ck_close(f, -223, -222);
    for (int i = 0; i < ncol; ++i) {
        if (col(u, i, buf) != 0 || fwrite(buf, sizeof(double), cw, f) != cw)
            return ck_close(f, -223, -222);
    }
    return ck_close(f, 0, -223);
}

int trlkit_read_checkpoint(const char *file, int nrow, int mev, int maxjnd, int *jnd_out,
                           double *alpha, double *beta,
                           int (*col)(void *u, int i, double *buf), void *u, double *buf)
{
    return trlkit_read_checkpoint_e(file, nrow, 1, mev, maxjnd, jnd_out, alpha, beta, col, u, buf);
}

int trlkit_read_checkpoint_e(const char *file, int nrow, int E, int mev, int maxjnd, int *jnd_out,
                             double *alpha, double *beta,
                             int (*col)(void *u, int i, double *buf), void *u, double *buf)
{
    const size_t cw = (size_t)nrow * E;
    FILE *f = fopen(file, "r");
    if (!f) {
        printf("TRL_READ_CHECKPOINT: failed to open check-point file %s.\n", file);
        return -212;
    }
    int fn = 0, jnd = 0;
    if (fread(&fn, sizeof fn, 1, f) != 1 || fread(&jnd, sizeof jnd, 1, f) != 1)
        return ck_close(f, -215, -216);
    if (fn != nrow) {
        printf("TRL_READ_CHECKPOINT: Nrow mismatch.\n");
        (void)fclose(f);
        return -213;
    }
    if (jnd > maxjnd || jnd < 0) {
        printf("TRL_READ_CHECKPOINT: MAXLAN too small.");
        (void)fclose(f);
        return -214;
    }
    if (jnd > 0 && (fread(alpha, sizeof(double), (size_t)jnd, f) != (size_t)jnd ||
                    fread(beta, sizeof(double), (size_t)jnd, f) != (size_t)jnd))
        return ck_close(f, -215, -216);
    const int j1 = jnd < mev ? jnd : mev;
    const int ncol = j1 < mev ? j1 + 1 : jnd;
    for (int i = 0; i < ncol; ++i) {
        if (fread(buf, sizeof(double), cw, f) != cw || col(u, i, buf) != 0)
            return ck_close(f, -215, -216);
    }
    *jnd_out = jnd;
    return ck_close(f, 0, -215);
}

void trlkit_pe_filename(char *out, size_t cap, const char *base)
{
    size_t len = strcspn(base, " ");
    if (len + 2 > cap) len = cap > 2 ? cap - 2 : 0;
    memcpy(out, base, len);
    out[len] = '0';
    out[len + 1] = '\0';
}
