/*
 * ztrlan.h -- complex Hermitian entry points of the MI355X thick-restart
 * Lanczos solver (drop-in for reference INC/ztrlan.h; TRL_FORTRAN_COMPATIBLE
 * selects the pointer-argument callback, as there).
 *
 * The basis is complex (interleaved re/im doubles, like trl_dcomplex) and
 * lives on the device; the projected tridiagonal stays real FP64 on the host
 * because alpha and beta are real for a Hermitian operator.
 */
#ifndef NUTRLAN_AMD_ZTRLAN_H
#define NUTRLAN_AMD_ZTRLAN_H

#include "trlan.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Complex scalar with the reference's layout {double r, i}
 * (reference INC/trl_map.h:26-28). */
typedef struct {
    double r, i;
} trl_dcomplex;

#ifdef TRL_FORTRAN_COMPATIBLE
/* Pointer-argument callback of the reference's TRL_FORTRAN_COMPATIBLE
 * build (INC/ztrlan.h:30-31); see trlan.h. */
typedef void (*ztrl_matvec)(int *pnrow, int *pncol, trl_dcomplex *x, int *pldx,
                            trl_dcomplex *y, int *pldy);
#define ztrlan ztrlan_fc
#define ztrl_check_ritz ztrl_check_ritz_fc
#else
/* Replaces: reference INC/ztrlan.h:62-64. */
typedef void (*ztrl_matvec)(const int nrow, const int ncol,
                            const trl_dcomplex *x, const int ldx,
                            trl_dcomplex *y, const int ldy, void *mvparam);
#endif

/* Thick-restart Lanczos for a complex Hermitian operator.  misc/nmis and
 * dwrk/ldwrk are accepted for ABI compatibility; on return misc[i].r holds
 * the residual-norm estimate of eigenpair i (i < nec), as in the reference.
 * Replaces: reference INC/ztrlan.h:139-142 / SRC/ztrlan.c:26. */
void ztrlan(ztrl_matvec op, trl_info *info, int nrow, int mev, double *eval,
            trl_dcomplex *evec, int lde, trl_dcomplex *misc, int nmis,
            double *dwrk, int ldwrk);

/* Replaces: reference INC/ztrlan.h:173-176 / SRC/ztrlan.c:225. */
void ztrl_check_ritz(ztrl_matvec op, trl_info *info, int nrow, int ncol,
                     trl_dcomplex *rvec, int ldrvec, double *alpha,
                     int *check, double *beta, double *eval,
                     trl_dcomplex *wrk, int lwrk);

#ifdef __cplusplus
}
#endif
#endif /* NUTRLAN_AMD_ZTRLAN_H */
