/*
 * trl_info.c -- trl_info set-up, setters, summaries and the Ritz-pair
 * checks of the public API (include/trlan.h).
 *
 * The checks (trl_check_ritz, trl_rayleigh_quotients, trl_ritz_projection)
 * apply the operator and take inner products on the GPU through a small
 * device context, like the solver itself.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "trl_kit.h"
#include "trlan.h"
#include "trlan_internal.h"
#include "trlhip.h"

#define IMIN(a, b) ((a) < (b) ? (a) : (b))
#define IMAX(a, b) ((a) > (b) ? (a) : (b))

/* Defaults and clamps: the reference's trl_init_info
 * (SRC/trl_comm_none.c:27-192).  When maxmv <= 0 the reference reads ntot
 * and ned before assigning them (trl_comm_none.c:104); here the default is
 * computed from the arguments, min(max(nrow, 1000), 1000*ned). */
void trl_init_info(trl_info *info, int nrow, int mxlan, int lohi, int ned, double tol,
                   int restart, int maxmv, void *mpicomp)
{
    memset(info, 0, sizeof *info);
    if (tol > 0) {
        info->tol = tol;
        if (info->tol <= DBL_MIN) info->tol = DBL_EPSILON;
        else if (info->tol > 1.0) info->tol = fmin(0.1, 1.0 / info->tol);
    } else {
        info->tol = sqrt(DBL_EPSILON);
    }
    info->restart = restart > 0 ? restart : 0;
    info->maxmv = maxmv > 0 ? maxmv : IMIN(IMAX(nrow, 1000), 1000 * ned);
    if (mpicomp != NULL) fprintf(stderr, "TRL_INIT_INFO:  ignoring MPI Comm.\n");
    info->mpicomp = NULL;
    info->maxlan = mxlan <= ned ? ned + IMAX(ned, 6) : mxlan;
    info->lohi = lohi;
    info->ned = ned;
    info->nloc = nrow;
    info->ntot = nrow;
    info->clk_rate = CLOCKS_PER_SEC;
    info->clk_max = (clock_t)9223372036854775807LL;
    info->tmv = -1;
    info->trgt = -DBL_MAX;
    info->tres = -1.0;
    info->crat = 1.0;
    info->my_pe = 0;
    info->npes = 1;
    info->log_io = 99;
    info->cpio = 98;
}

void trl_set_restart(trl_info *info, double rfact) { info->rfact = rfact; }

void trl_set_debug(trl_info *info, int msglvl, char *filename)
{
    info->verbose = msglvl;
    if (filename != NULL) {
        strncpy(info->log_file, filename, sizeof info->log_file - 1);
        info->log_file[sizeof info->log_file - 1] = '\0';
        if (msglvl >= 0 && info->my_pe == 0)
            printf("TRLan will write diagnostic messages to files with prefix %s.\n",
                   info->log_file);
    }
}

void trl_set_checkpoint(trl_info *info, int cpflag, char *file)
{
    info->cpflag = cpflag;
    if (file != NULL) {
        strncpy(info->cpfile, file, sizeof info->cpfile - 1);
        info->cpfile[sizeof info->cpfile - 1] = '\0';
        if (info->verbose >= 0 && info->my_pe == 0)
            printf("TRLan will write checkpoint to files with prefix %s.\n", info->cpfile);
    }
}

/* Reference SRC/trlan.c:405-470: record nec/guess; for guess > 1 the
 * checkpoint file named by oldcpf (+ PE number) must exist. */
void trl_set_iguess(trl_info *info, int nec, int iguess, int ncps, char *cpf)
{
    (void)ncps;
    info->nec = nec;
    info->guess = iguess;
    info->stat = 0;
    if (info->guess > 1 && strlen(info->oldcpf) > 0) {
        char name[128];
        snprintf(name, sizeof name, "%.118s%d", info->oldcpf, info->my_pe);
        if (cpf) memcpy(cpf, name, sizeof name);
        FILE *f = fopen(name, "r");
        if (f == NULL) info->stat = -8;
        else if (fclose(f) != 0) info->stat = -9;
    }
}

static double total_seconds(clock_t ticks, double spill, clock_t rate)
{
    return spill + (double)ticks / (double)(rate > 0 ? rate : CLOCKS_PER_SEC);
}

void trl_print_info(trl_info *info, int mvflop)
{
    FILE *fp = info->log_fp ? info->log_fp : stdout;
    const double t_op = total_seconds(info->clk_op, info->tick_o, info->clk_rate);
    const double t_or = total_seconds(info->clk_orth, info->tick_h, info->clk_rate);
    const double t_rs = total_seconds(info->clk_res, info->tick_r, info->clk_rate);
    const double t_tot = total_seconds(info->clk_tot, info->tick_t, info->clk_rate);
    fprintf(fp, "TRLAN (MI355X device basis) summary\n");
    fprintf(fp, "  Problem size: %d   wanted: %d %s   converged: %d   locked: %d\n", info->ntot,
            info->ned, info->lohi > 0 ? "largest" : (info->lohi < 0 ? "smallest" : "either end"),
            info->nec, info->locked);
    fprintf(fp, "  Restart scheme: %d   max basis: %d   tol: %.4e (abs %.4e, ||A|| ~ %.4e)\n",
            info->restart, info->maxlan, info->tol, info->tol * info->anrm, info->anrm);
    fprintf(fp, "  MATVEC: %d (max %d)   re-orthogonalizations: %d   restarts: %d   random: %d\n",
            info->matvec, info->maxmv, info->north, info->nloop, info->nrand);
    fprintf(fp, "  Host CPU seconds: total %.3e  op %.3e  orth %.3e  restart %.3e\n", t_tot,
            t_op, t_or, t_rs);
    if (mvflop > 0)
        fprintf(fp, "  Operator flops (caller estimate): %.4e\n", (double)mvflop * info->matvec);
    fprintf(fp, "  Status: %d\n", info->stat);
}

void trl_terse_info(trl_info *info, FILE *iou)
{
    FILE *fp = iou ? iou : stdout;
    fprintf(fp,
            "MAXLAN:%6d, Restart:%3d, NED:%3d%s, NEC:%3d, MATVEC:%7d, Reorth:%7d, Nloop:%5d, "
            "Nlocked:%4d, Stat:%d\n",
            info->maxlan, info->restart, info->ned,
            info->lohi > 0 ? "+" : (info->lohi < 0 ? "-" : "0"), info->nec, info->matvec,
            info->north, info->nloop, info->locked, info->stat);
}

/* ------------------------------------------------------ device checks */

typedef struct {
    trlhip_ctx *dev;
    const trlhip_csr *A;
    trl_matvec op;
    void *mvparam;
    int n;
    double *hy;
} chk_t;

static int chk_open(chk_t *c, trl_matvec op, trl_info *info, int n, int ncols)
{
    memset(c, 0, sizeof *c);
    c->op = op;
    c->mvparam = info->mvparam;
    c->n = n;
    c->A = trlhip_csr_from_op((void (*)(void))op, info->mvparam, TRLHIP_R64);
    c->hy = (double *)malloc(sizeof(double) * (size_t)n);
    if (!c->hy) return -1;
    return trlhip_ctx_create(&c->dev, n, ncols, TRLHIP_R64);
}

static void chk_close(chk_t *c)
{
    trlhip_ctx_destroy(c->dev);
    free(c->hy);
}

/* column ycol <- A * (host vector x, also uploaded to column xcol) */
static int chk_apply(chk_t *c, const double *x, int ldx, int xcol, int ycol)
{
    int rc = trlhip_col_set(c->dev, xcol, x);
    if (rc) return rc;
    if (c->A) return trlhip_spmv(c->dev, c->A, xcol, ycol, NULL);
    c->op(c->n, 1, x, ldx, c->hy, c->n, c->mvparam);
    return trlhip_col_set(c->dev, ycol, c->hy);
}

/* Reference SRC/trlan.c:1111-1215. */
void trl_rayleigh_quotients(trl_matvec op, trl_info *info, int ncol, double *evec, int lde,
                            double *eres, double *base)
{
    (void)base;
    const int n = info->nloc;
    if (ncol <= 0) return;
    chk_t c;
    if (chk_open(&c, op, info, n, 2) != 0) {
        info->stat = TRLHIP_STAT_DEVICE;
        chk_close(&c);
        return;
    }
    for (int i = 0; i < ncol; ++i) {
        double *v = evec + (size_t)i * lde;
        double vv = 0.0, vav = 0.0, rr2 = 0.0;
        if (chk_apply(&c, v, lde, 0, 1) || trlhip_update(c.dev, 0, 0, &vv, 0, &vv) ||
            trlhip_proj(c.dev, 0, 1, 1, &vav)) {
            info->stat = TRLHIP_STAT_DEVICE;
            break;
        }
        info->matvec++;
        if (vv > 0.0) {
            eres[i] = vav / vv;
            const double h = eres[i];
            if (trlhip_update(c.dev, 0, 1, &h, 1, &rr2) || trlhip_col_scale(c.dev, 0, 1.0 / sqrt(vv)) ||
                trlhip_col_get(c.dev, 0, v)) {
                info->stat = TRLHIP_STAT_DEVICE;
                break;
            }
            eres[ncol + i] = sqrt(rr2) / sqrt(vv);
        } else {
            eres[i] = -DBL_MAX;
            eres[ncol + i] = -DBL_MAX;
        }
    }
    chk_close(&c);
}

/* Reference SRC/trlan.c:805-1108: residual norms, error estimates and the
 * pass/fail count in *check (the thresholds at trlan.c:1015-1047). */
void trl_check_ritz(trl_matvec op, trl_info *info, int nrow, int ncol, double *rvec, int ldrvec,
                    double *alpha, int *check, double *beta, double *eval, double *wrk, int lwrk)
{
    (void)wrk;
    (void)lwrk;
    if (ncol <= 0) return;
    *check = 0;
    double *rq = (double *)calloc((size_t)3 * ncol, sizeof(double));
    double *res = rq + ncol, *err = rq + 2 * ncol;
    chk_t c;
    memset(&c, 0, sizeof c);
    if (!rq || chk_open(&c, op, info, nrow, 2) != 0) {
        info->stat = TRLHIP_STAT_DEVICE;
        free(rq);
        chk_close(&c);
        return;
    }
    for (int i = 0; i < ncol; ++i) {
        double r2 = 0.0;
        if (chk_apply(&c, rvec + (size_t)i * ldrvec, ldrvec, 0, 1) ||
            trlhip_proj(c.dev, 0, 1, 1, &rq[i]) || trlhip_update(c.dev, 0, 1, &rq[i], 1, &r2)) {
            info->stat = TRLHIP_STAT_DEVICE;
            break;
        }
        res[i] = sqrt(r2);
    }
    chk_close(&c);
    double gapl = alpha[ncol - 1] - alpha[0];
    for (int i = 0; i < ncol - 1; ++i) {
        const double gapr = alpha[i + 1] - alpha[i];
        gapl = fmin(gapl, gapr);
        err[i] = res[i] >= gapl ? res[i] : res[i] * res[i] / gapl;
        gapl = gapr;
    }
    err[ncol - 1] = res[ncol - 1] >= gapl ? res[ncol - 1] : res[ncol - 1] * res[ncol - 1] / gapl;
    FILE *fp = info->log_fp ? info->log_fp : stdout;
    if (info->stat != 0) *check = -4;
    fprintf(fp, "TRL_CHECK_RITZ: \n");
    fprintf(fp, "           Ritz value       res norm   res diff  est error  diff w rq  act. error\n");
    for (int i = 0; i < ncol; ++i) {
        if (beta && eval) {
            fprintf(fp, "%21.14f    %11.3e%11.3e%11.3e%11.3e %11.3e%11.3e\n", alpha[i], res[i],
                    beta[i] - res[i], err[i], rq[i] - alpha[i], eval[i] - alpha[i], eval[i]);
        } else if (beta) {
            fprintf(fp, "%21.14f    %11.3e%11.3e%11.3e%11.3e\n", alpha[i], res[i],
                    beta[i] - res[i], err[i], rq[i] - alpha[i]);
        } else {
            fprintf(fp, "%21.14f    %11.5e           %11.3e%11.3e\n", alpha[i], res[i], err[i],
                    rq[i] - alpha[i]);
        }
        if (beta) {
            if (fabs(beta[i] - res[i]) > 0.00001) --*check;
            if (fabs(rq[i] - alpha[i]) > (double)nrow * nrow * info->tol) --*check;
        }
        if (beta && eval) {
            if (fabs(eval[i] - alpha[i]) > 10.0 * nrow * nrow * info->tol ||
                fabs(eval[i] - alpha[i]) > 10.0 * err[i])
                --*check;
        }
    }
    if (info->nec < info->ned) *check = 1;
    free(rq);
}

/* Rayleigh-Ritz on the first mev columns of evec (reference
 * SRC/trlan.c:1438-1678): G = V'V = R'R, H = R^-T (V'AV) R^-1 = Y D Y',
 * V <- V R^-1 Y, eres[0:mev) = D, eres[mev:2mev) = ||A v_i - d_i v_i||.
 * (The reference adds eres to the shifted vector with the wrong sign at
 * trlan.c:1644; the residual here is the mathematical one.) */
void trl_ritz_projection(trl_matvec op, trl_info *info, int mev, double *evec, int lde,
                         double *eres, double *wrk, int lwrk, double *base)
{
    (void)wrk;
    (void)lwrk;
    (void)base;
    const int n = info->nloc, m = mev;
    if (m <= 0) return;
    chk_t c;
    memset(&c, 0, sizeof c);
    double *G = (double *)calloc((size_t)m * m, sizeof(double));
    double *H = (double *)calloc((size_t)m * m, sizeof(double));
    double *T = (double *)calloc((size_t)m * m, sizeof(double));
    double *w = (double *)calloc((size_t)m, sizeof(double));
    if (!G || !H || !T || !w || chk_open(&c, op, info, n, 2 * m) != 0) {
        info->stat = TRLHIP_STAT_DEVICE;
        goto done_dev;
    }
    /* columns 0..m-1: V ; m..2m-1: A V */
    for (int i = 0; i < m; ++i)
        if (chk_apply(&c, evec + (size_t)i * lde, lde, i, m + i)) {
            info->stat = TRLHIP_STAT_DEVICE;
            goto done_dev;
        }
    info->matvec += m;
    for (int j = 0; j < m; ++j)
        if (trlhip_proj(c.dev, 0, m, j, G + (size_t)j * m) ||
            trlhip_proj(c.dev, 0, m, m + j, H + (size_t)j * m)) {
            info->stat = TRLHIP_STAT_DEVICE;
            goto done_dev;
        }
    /* Cholesky G = R'R (upper R stored in G) */
    for (int j = 0; j < m; ++j) {
        double s = G[(size_t)j * m + j];
        for (int k = 0; k < j; ++k) s -= G[(size_t)j * m + k] * G[(size_t)j * m + k];
        if (!(s > 0.0)) {
            info->stat = -221;
            goto done_dev;
        }
        const double r = sqrt(s);
        G[(size_t)j * m + j] = r;
        for (int i = j + 1; i < m; ++i) {
            double t = G[(size_t)i * m + j];
            for (int k = 0; k < j; ++k) t -= G[(size_t)j * m + k] * G[(size_t)i * m + k];
            G[(size_t)i * m + j] = t / r;
        }
    }
    /* R^-1 in T (upper triangular) */
    for (int j = 0; j < m; ++j) {
        T[(size_t)j * m + j] = 1.0 / G[(size_t)j * m + j];
        for (int i = j - 1; i >= 0; --i) {
            double s = 0.0;
            for (int k = i + 1; k <= j; ++k) s += G[(size_t)k * m + i] * T[(size_t)j * m + k];
            T[(size_t)j * m + i] = -s / G[(size_t)i * m + i];
        }
    }
    {
        /* H <- T' H T */
        double *tmp = (double *)calloc((size_t)m * m, sizeof(double));
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) {
                double s = 0.0;
                for (int k = 0; k < m; ++k) s += H[(size_t)k * m + i] * T[(size_t)j * m + k];
                tmp[(size_t)j * m + i] = s;
            }
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) {
                double s = 0.0;
                for (int k = 0; k < m; ++k) s += T[(size_t)i * m + k] * tmp[(size_t)j * m + k];
                H[(size_t)j * m + i] = s;
            }
        if (trlkit_syev(m, H, w) != 0) {
            free(tmp);
            info->stat = -222;
            goto done_dev;
        }
        /* Y' = T * Y  (V <- V T Y) */
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) {
                double s = 0.0;
                for (int k = 0; k < m; ++k) s += T[(size_t)k * m + i] * H[(size_t)j * m + k];
                tmp[(size_t)j * m + i] = s;
            }
        if (trlhip_ritz(c.dev, 0, m, m, tmp, m) || trlhip_ritz(c.dev, m, m, m, tmp, m)) {
            free(tmp);
            info->stat = TRLHIP_STAT_DEVICE;
            goto done_dev;
        }
        free(tmp);
    }
    for (int i = 0; i < m; ++i) {
        double r2 = 0.0;
        const double h = w[i];
        eres[i] = w[i];
        if (trlhip_col_get(c.dev, i, evec + (size_t)i * lde)) {
            info->stat = TRLHIP_STAT_DEVICE;
            break;
        }
        /* ||A v - d v|| = ||(A V Y)_i - d_i (V Y)_i||, in place on column m+i */
        if (trlhip_update(c.dev, i, 1, &h, m + i, &r2)) {
            info->stat = TRLHIP_STAT_DEVICE;
            break;
        }
        eres[m + i] = sqrt(r2);
    }
done_dev:
    chk_close(&c);
    free(G);
    free(H);
    free(T);
    free(w);
}
