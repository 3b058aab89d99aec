/*
 * ztrlan.c -- complex Hermitian entry points (include/ztrlan.h).
 *
 * ztrlan() runs the same device-resident thick-restart driver as trlan()
 * with a complex basis (interleaved re/im doubles on the GPU) and a real
 * tridiagonal on the host, like the reference's ztrlanczos
 * (SRC/ztrlcore.c:436-1286).
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "trlan_internal.h"
#include "ztrlan.h"

void ztrlan(ztrl_matvec op, trl_info *info, int nrow, int mev, double *eval,
            trl_dcomplex *evec, int lde, trl_dcomplex *misc, int nmis, double *dwrk, int ldwrk)
{
    (void)dwrk;
    (void)ldwrk;
    if (misc && nmis > 0) memset(misc, 0, sizeof(trl_dcomplex) * (size_t)nmis);
    trl_solve_device(NULL, op, NULL, info, nrow, mev, eval, (double *)evec, lde, nmis,
                     (double *)misc, 1);
}

/* Reference SRC/ztrlan.c:225-500: Rayleigh quotients Re(v^H A v),
 * residual norms, error estimates and the pass/fail count. */
void ztrl_check_ritz(ztrl_matvec op, trl_info *info, int nrow, int ncol, trl_dcomplex *rvec,
                     int ldrvec, double *alpha, int *check, double *beta, double *eval,
                     trl_dcomplex *wrk, int lwrk)
{
    (void)wrk;
    (void)lwrk;
    if (ncol <= 0) return;
    *check = 0;
    double *rq = (double *)calloc((size_t)3 * ncol, sizeof(double));
    double *res = rq + ncol, *err = rq + 2 * ncol;
    double *hy = (double *)malloc(sizeof(double) * 2 * (size_t)nrow);
    trlhip_ctx *dev = NULL;
    const trlhip_csr *A = trlhip_csr_from_op((void (*)(void))op, info->mvparam, TRLHIP_C128);
    if (!rq || !hy || trlhip_ctx_create(&dev, nrow, 2, TRLHIP_C128) != 0) {
        info->stat = TRLHIP_STAT_DEVICE;
        goto done;
    }
    for (int i = 0; i < ncol; ++i) {
        const trl_dcomplex *v = rvec + (size_t)i * ldrvec;
        double h[2], r2 = 0.0;
        int rc = trlhip_col_set(dev, 0, v);
        if (!rc) {
            if (A) {
                rc = trlhip_spmv(dev, A, 0, 1, NULL);
            } else {
                op(nrow, 1, v, ldrvec, (trl_dcomplex *)hy, nrow, info->mvparam);
                rc = trlhip_col_set(dev, 1, hy);
            }
        }
        if (!rc) rc = trlhip_proj(dev, 0, 1, 1, h);
        if (!rc) {
            rq[i] = h[0];
            h[1] = 0.0; /* subtract the real Rayleigh quotient times v */
            rc = trlhip_update(dev, 0, 1, h, 1, &r2);
        }
        if (rc) {
            info->stat = TRLHIP_STAT_DEVICE;
            break;
        }
        res[i] = sqrt(r2);
    }
    {
        double gapl = alpha[ncol - 1] - alpha[0];
        for (int i = 0; i < ncol - 1; ++i) {
            const double gapr = alpha[i + 1] - alpha[i];
            gapl = fmin(gapl, gapr);
            err[i] = res[i] >= gapl ? res[i] : res[i] * res[i] / gapl;
            gapl = gapr;
        }
        err[ncol - 1] = res[ncol - 1] >= gapl ? res[ncol - 1] : res[ncol - 1] * res[ncol - 1] / gapl;
    }
    {
        FILE *fp = info->log_fp ? info->log_fp : stdout;
        if (info->stat != 0) *check = -4;
        fprintf(fp, "ZTRL_CHECK_RITZ: \n");
        fprintf(fp, "           Ritz value       res norm   res diff  est error  diff w rq  act. error\n");
        for (int i = 0; i < ncol; ++i) {
            if (beta && eval)
                fprintf(fp, "%21.14f    %11.3e%11.3e%11.3e%11.3e %11.3e%11.3e\n", alpha[i], res[i],
                        beta[i] - res[i], err[i], rq[i] - alpha[i], eval[i] - alpha[i], eval[i]);
            else
                fprintf(fp, "%21.14f    %11.3e           %11.3e%11.3e\n", alpha[i], res[i], err[i],
                        rq[i] - alpha[i]);
            if (beta) {
                if (fabs(beta[i] - res[i]) > 0.00001) --*check;
                if (fabs(rq[i] - alpha[i]) > (double)nrow * nrow * info->tol) --*check;
            }
            if (beta && eval &&
                (fabs(eval[i] - alpha[i]) > 10.0 * nrow * nrow * info->tol ||
                 fabs(eval[i] - alpha[i]) > 10.0 * err[i]))
                --*check;
        }
        if (info->nec < info->ned) *check = 1;
    }
done:
    trlhip_ctx_destroy(dev);
    free(rq);
    free(hy);
}
