/*
 * trl_debug.c -- the reference's debugging checks of the Lanczos basis, on
 * the device-resident basis:
 *
 *   trl_check_orth / ztrl_check_orth (SRC/trlaux.c:331-460,
 *   SRC/ztrlaux.c:308-431): the orthogonality levels v_k^H v_i - delta_ki
 *   of every basis vector, their Frobenius norm and largest value;
 *   trl_check_recurrence / ztrl_check_recurrence (SRC/trlaux.c:462-778,
 *   SRC/ztrlaux.c:437-830): alpha, beta and the Lanczos relations
 *   recomputed with the operator.
 *
 * The solver calls them where the reference does (verbose > 6 after an
 * inner loop and at the end of a restart, > 7 adds the recurrence check,
 * > 8 after a cycle's first step, > 6 after the initial guess), with the
 * same message text and number formats, written to info->log_fp (stdout
 * when it is unset).  The dot products come from the device projection
 * kernel in FP64 (trlhip_proj: V^H r); the recurrence check copies the
 * basis vectors to the host and applies the caller's operator there (the
 * device CSR operator accepts host vectors too), as the reference does.
 * Basis vector i of a check is column i of the context for i < k and
 * column `last` for i = k.  `avail` is what the reference's workspace
 * holds at that point of its solve (the solver works it out): below what a
 * check needs, the check prints the reference's "workspace too small" /
 * "not enough workspace" line instead.  Nothing here changes the solve: a projection
 * first materializes a deferred column copy or scale, which stores the
 * same values the fused path would.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "trlan_internal.h"

static FILE *dbg_fp(const trl_info *info) { return info->log_fp ? info->log_fp : stdout; }

static int dbg_col(int i, int k, int last) { return i < k ? i : last; }

/* w[0:E*(i+1)) = v_0..v_i ^H v_i for check vector i. */
static int dbg_dots(trlhip_ctx *dev, int E, int i, int k, int last, double *w)
{
    if (i < k) return trlhip_proj(dev, 0, i + 1, i, w);
    if (k > 0 && trlhip_proj(dev, 0, k, last, w)) return -1;
    return trlhip_proj(dev, last, 1, last, w + (size_t)E * k);
}

void trl_dbg_check_orth(trl_info *info, trlhip_ctx *dev, int E, int j1, int j2, int k, int last,
                        long avail)
{
    const int jnd = j1 + j2;
    if (jnd <= 0) return;
    FILE *fp = dbg_fp(info);
    double *w = avail < 2L * jnd ? NULL : (double *)calloc((size_t)E * (jnd + 1), sizeof(double));
    if (!w) {
        fprintf(fp, "TRL_CHECK_ORTH: workspace too small.\n");
        return;
    }
    fprintf(fp, "TRL_CHECK_ORTH: check orthogonality of %d basis vectors.\n", jnd);
    double nrmfro = 0.0, nrminf = 0.0;
    for (int i = 0; i < jnd; ++i) {
        if (dbg_dots(dev, E, i, k, last, w)) {
            info->stat = TRLHIP_STAT_DEVICE;
            break;
        }
        w[(size_t)E * i] -= 1.0;
        if (info->verbose > 7) {
            fprintf(fp, "Orthogonality level of v(%d) ..\n", i + 1);
            const int per = E == 1 ? 8 : 5;
            for (int j = 0; j <= i; ++j) {
                if (E == 1)
                    fprintf(fp, " %10.3e", w[j]);
                else
                    fprintf(fp, " %10.3e + i * %10.3e, ", w[2 * j], w[2 * j + 1]);
                if (j % per == per - 1) fprintf(fp, "\n");
            }
            if (i % per != per - 1) fprintf(fp, "\n");
        }
        double off = 0.0;  /* sum over k < i of |w_k|^2 (ddot / Re zdotc) */
        for (int j = 0; j < E * i; ++j) off += w[j] * w[j];
        double diag = w[(size_t)E * i] * w[(size_t)E * i];
        if (E == 2) diag += w[2 * i + 1] * w[2 * i + 1];
        nrmfro += 2 * off + diag;
        if (i < j1) {
            /* the reference's running value here compares signed levels
             * (trlaux.c:417-421; ztrlaux.c:386-390) */
            const double t = i == 0 ? fabs(w[0]) : fmax(w[(size_t)E * i], w[(size_t)E * (i - 1)]);
            nrminf = fmax(nrminf, t);
        } else {
            nrminf = fmax(nrminf, fabs(w[(size_t)E * i]));
        }
    }
    fprintf(fp, "Frobenius norm of orthogonality level %10i %4i  %14.5e\n", info->matvec, jnd,
            sqrt(nrmfro));
    fprintf(fp, "Maximum abs. value of orthogonality level is  %14.5e\n", nrminf);
    free(w);
}

static void dbg_print_real(trl_info *info, const char *title, int n, const double *a)
{
    FILE *fp = dbg_fp(info);
    fprintf(fp, "PE %d : %s", info->my_pe, title);
    if (n > 1) fprintf(fp, "\n");
    for (int i = 0; i < n; ++i) {
        fprintf(fp, " %10.7e", a[i]);
        if (i % 8 == 7) fprintf(fp, "\n");
    }
    if ((n - 1) % 8 != 7) fprintf(fp, "\n");
}

static void dbg_print_complex(trl_info *info, const char *title, int n, const double *re,
                              const double *im)
{
    FILE *fp = dbg_fp(info);
    fprintf(fp, "PE %d : %s", info->my_pe, title);
    if (n > 1) fprintf(fp, "\n");
    for (int i = 0; i < n; ++i) {
        fprintf(fp, " %10.7e+%10.7ei", re[i], im[i]);
        if (i % 4 == 3) fprintf(fp, "\n");
    }
    if ((n - 1) % 4 != 3) fprintf(fp, "\n");
}

/* The recurrence check over m1 + m2 basis vectors (the reference's evec and
 * base parts) of which `kept` came from the last restart. */
void trl_dbg_check_recurrence(trl_info *info, trlhip_ctx *dev, int n, int E, trl_matvec op,
                              ztrl_matvec zop, int m1, int m2, int kept, const double *alpha,
                              const double *beta, int k, int last, long avail)
{
    const int jnd = m1 + m2 - 1;  /* trlaux.c:531-538 */
    if (jnd <= 0) return;
    if (avail < 4L * jnd + (4L * jnd > n ? 4L * jnd : n)) {
        fprintf(dbg_fp(info), "TRL_CHECK_RECURRENCE: not enough workspace.\n");
        return;
    }
    const size_t ne = (size_t)E * n;
    double *V = (double *)malloc(sizeof(double) * ne * (jnd + 1));
    double *aq = (double *)malloc(sizeof(double) * ne);
    /* alf, bet, cs, wrk: real and imaginary parts */
    double *acc = (double *)calloc((size_t)8 * jnd, sizeof(double));
    if (!V || !aq || !acc) {
        fprintf(dbg_fp(info), "TRL_CHECK_RECURRENCE: failed to allcoate workspace.\n");
        free(V);
        free(aq);
        free(acc);
        return;
    }
    double *wr = acc, *wi = acc + jnd, *csr = acc + 2 * jnd, *csi = acc + 3 * jnd;
    double *alr = acc + 4 * jnd, *ali = acc + 5 * jnd, *btr = acc + 6 * jnd, *bti = acc + 7 * jnd;
    for (int i = 0; i <= jnd; ++i)
        if (trlhip_col_get(dev, dbg_col(i, k, last), V + ne * i)) {
            info->stat = TRLHIP_STAT_DEVICE;
            goto out;
        }
    /* aq <- A v */
#define APPLY(v)                                                                            \
    do {                                                                                    \
        if (E == 1)                                                                         \
            op(n, 1, (v), n, aq, n, info->mvparam);                                         \
        else                                                                                \
            zop(n, 1, (const trl_dcomplex *)(v), n, (trl_dcomplex *)aq, n, info->mvparam); \
    } while (0)
    const double *qk = V + ne * kept;
    /* A q_i = alpha_i q_i + beta_i q_{kept}, i < kept (trlaux.c:564-606) */
    for (int i = 0; i < kept && i < jnd; ++i) {
        const double *v = V + ne * i;
        APPLY(v);
        for (int ii = 0; ii < n; ++ii) {
            if (E == 1) {
                alr[i] += aq[ii] * v[ii];
                aq[ii] -= alpha[i] * v[ii];
                btr[i] += aq[ii] * aq[ii];
                csr[i] += aq[ii] * qk[ii];
                aq[ii] -= beta[i] * qk[ii];
                wr[i] += aq[ii] * aq[ii];
            } else {
                double *a = aq + 2 * ii;
                const double *x = v + 2 * ii, *q = qk + 2 * ii;
                alr[i] += a[0] * x[0] + a[1] * x[1];
                ali[i] += a[1] * x[0] - a[0] * x[1];
                a[0] -= alpha[i] * x[0];
                a[1] -= alpha[i] * x[1];
                btr[i] += a[0] * a[0] + a[1] * a[1];
                csr[i] += a[0] * q[0] + a[1] * q[1];
                csi[i] += a[0] * q[1] - a[1] * q[0];
                a[0] -= beta[i] * q[0];
                a[1] -= beta[i] * q[1];
                wr[i] += a[0] * a[0] + a[1] * a[1];
            }
        }
    }
    /* q_{kept+1} against every earlier vector (trlaux.c:611-643) */
    if (jnd > kept) {
        APPLY(qk);
        const double *q1 = V + ne * (kept + 1);
        double s0 = 0.0, s1 = 0.0;
        for (int ii = 0; ii < n; ++ii) { /* zdotc(aq, qk) / ddot */
            if (E == 1) {
                s0 += aq[ii] * qk[ii];
            } else {
                s0 += aq[2 * ii] * qk[2 * ii] + aq[2 * ii + 1] * qk[2 * ii + 1];
                s1 += aq[2 * ii] * qk[2 * ii + 1] - aq[2 * ii + 1] * qk[2 * ii];
            }
        }
        alr[kept] = s0;
        ali[kept] = s1;
        for (size_t t = 0; t < ne; ++t) aq[t] += -alpha[kept] * qk[t];
        for (int i = 0; i < kept; ++i)
            for (size_t t = 0; t < ne; ++t) aq[t] += -beta[i] * V[ne * i + t];
        s0 = 0.0;
        for (size_t t = 0; t < ne; ++t) s0 += aq[t] * aq[t];
        btr[kept] = s0;
        s0 = s1 = 0.0;
        for (int ii = 0; ii < n; ++ii) {
            if (E == 1) {
                s0 += aq[ii] * q1[ii];
            } else {
                s0 += aq[2 * ii] * q1[2 * ii] + aq[2 * ii + 1] * q1[2 * ii + 1];
                s1 += aq[2 * ii] * q1[2 * ii + 1] - aq[2 * ii + 1] * q1[2 * ii];
            }
        }
        csr[kept] = s0;
        csi[kept] = s1;
        for (size_t t = 0; t < ne; ++t) aq[t] += -beta[kept] * q1[t];
        s0 = 0.0;
        for (size_t t = 0; t < ne; ++t) s0 += aq[t] * aq[t];
        wr[kept] = s0;
    }
    /* the three-term recurrence (trlaux.c:648-737) */
    for (int i = kept + 1; i < jnd; ++i) {
        const double *v = V + ne * i, *vp = V + ne * (i - 1), *vn = V + ne * (i + 1);
        APPLY(v);
        for (int ii = 0; ii < n; ++ii) {
            if (E == 1) {
                alr[i] += aq[ii] * v[ii];
                aq[ii] -= alpha[i] * v[ii] + beta[i - 1] * vp[ii];
                btr[i] += aq[ii] * aq[ii];
                csr[i] += aq[ii] * vn[ii];
                aq[ii] -= beta[i] * vn[ii];
                wr[i] += aq[ii] * aq[ii];
            } else {
                double *a = aq + 2 * ii;
                const double *x = v + 2 * ii, *p = vp + 2 * ii, *q = vn + 2 * ii;
                alr[i] += a[0] * x[0] + a[1] * x[1];
                ali[i] += a[1] * x[0] - a[0] * x[1];
                a[0] -= alpha[i] * x[0] + beta[i - 1] * p[0];
                a[1] -= alpha[i] * x[1] + beta[i - 1] * p[1];
                btr[i] += a[0] * a[0] + a[1] * a[1];
                csr[i] += a[0] * q[0] + a[1] * q[1];
                csi[i] += a[0] * q[1] - a[1] * q[0];
                a[0] -= beta[i] * q[0];
                a[1] -= beta[i] * q[1];
                wr[i] += a[0] * a[0] + a[1] * a[1];
            }
        }
    }
#undef APPLY
    {
        double tot = 0.0;
        for (int i = 0; i < jnd; ++i) tot += wr[i];
        tot = sqrt(tot);
        for (int i = 0; i < jnd; ++i) wr[i] = sqrt(wr[i]);
        for (int i = 0; i < jnd; ++i) {
            if (btr[i] > 0.0) {
                btr[i] = beta[i] < 0.0 ? -sqrt(btr[i]) : sqrt(btr[i]);
                csr[i] /= btr[i];
                csi[i] /= btr[i];
            } else {
                btr[i] = 0.0;
            }
        }
        /* the overall error takes aq[0]'s place; the complex form prints
         * its imaginary part as it stands (ztrlaux.c:766-772) */
        const double toti = E == 2 ? aq[1] : 0.0;
        dbg_print_real(info, "Alpha computed by TRLAN ..", jnd, alpha);
        if (E == 1) {
            dbg_print_real(info, "Alpha computed explicitly in TRL_CHECK_RECURRENCE ..", jnd, alr);
            for (int i = 0; i < jnd; ++i) alr[i] += -1.0 * alpha[i];
            dbg_print_real(info, "Differences in alpha ..", jnd, alr);
            dbg_print_real(info, "Beta computed by TRLAN ..", jnd, beta);
            dbg_print_real(info, "Beta computed explicitly in TRL_CHECK_RECURRENCE ..", jnd, btr);
            for (int i = 0; i < jnd; ++i) btr[i] += -1.0 * beta[i];
            dbg_print_real(info, "Differences in beta ..", jnd, btr);
            dbg_print_real(info, "Error in Lanczos recurrence (overall) =", 1, &tot);
        } else {
            dbg_print_complex(info, "Alpha computed explicitly in TRL_CHECK_RECURRENCE ..", jnd,
                              alr, ali);
            for (int i = 0; i < jnd; ++i) alr[i] -= alpha[i];
            dbg_print_complex(info, "Differences in alpha ..", jnd, alr, ali);
            dbg_print_real(info, "Beta computed by TRLAN ..", jnd, beta);
            dbg_print_complex(info, "Beta computed explicitly in TRL_CHECK_RECURRENCE ..", jnd, btr,
                              bti);
            for (int i = 0; i < jnd; ++i) btr[i] -= beta[i];
            dbg_print_complex(info, "Differences in beta ..", jnd, btr, bti);
            dbg_print_complex(info, "Error in Lanczos recurrence (overall) =", 1, &tot, &toti);
        }
        /* the per-vector detail: verbose > 7 for real (trlaux.c:759-775);
         * always for complex, whose guard is commented out (ztrlaux.c:809) */
        if (E == 2 || info->verbose > 7) {
            const char *t1 = "|| A q_i - alpha_i q_i - beta_{i-1} q_{i-1} - beta_i q_{i+1} ||..";
            const char *t2 = "(A q_i - alpha_i q_i - beta_{i-1} q_{i-1})*q_{i+1}/beta_i ..";
            if (E == 1) {
                dbg_print_real(info, t1, jnd, wr);
                dbg_print_real(info, t2, jnd, csr);
            } else {
                dbg_print_complex(info, t1, jnd, wr, wi);
                dbg_print_complex(info, t2, jnd, csr, csi);
            }
            for (int i = 0; i < jnd; ++i) {
                const double c2 = E == 1 ? csr[i] * csr[i] : csr[i] * csr[i] + csi[i] * csi[i];
                csr[i] = c2 < 1.0 ? sqrt(1.0 - c2) : -1.0;
            }
            if (E == 1)
                dbg_print_real(info, "Sine of the angles ..", jnd, csr);
            else
                dbg_print_complex(info, "Sine of the angles ..", jnd, csr, csi);
        }
    }
out:
    free(V);
    free(aq);
    free(acc);
}
