/*
 * trl_fortran.c -- the pointer-argument ("Fortran compatible") callback
 * convention of nu-TRLan (reference INC/trlan.h:9-36, INC/ztrlan.h:30-31,
 * selected there by compiling with TRL_FORTRAN_COMPATIBLE).
 *
 * In that build the reference's operator takes every scalar by address and
 * no mvparam: op(&nrow, &ncol, x, &ldx, y, &ldy) (SRC/trlcore.c:938-942,
 * 1098-1102; SRC/ztrlcore.c:701-705, 862-866; SRC/trlan.c:950-954,
 * 1183-1187, 1580-1584, 1639-1643; SRC/ztrlan.c:341-345), and user_ortho
 * takes &nrow.  This library is built once, so the two conventions are two
 * sets of entry points: the *_fc functions below take pointer-argument
 * callbacks and run the ordinary entry points through an adapter, which
 * rides in info->mvparam for the duration of the call (the caller's
 * mvparam is restored before returning; with this convention the operator
 * never sees it anyway).  include/trlan.h and include/ztrlan.h map the
 * reference's names onto them when TRL_FORTRAN_COMPATIBLE is defined, so
 * code written for the reference's Fortran-compatible build compiles
 * unchanged against this library.  Results are those of the ordinary entry
 * points with an equivalent C callback, bit for bit.
 *
 * user_ortho has no parameter of its own, so the adapter's pointer to it is
 * thread-local: a thread runs one solve at a time, and solves on different
 * threads do not share it.
 */
#include <stddef.h>

#include "trlan.h"
#include "ztrlan.h"

typedef void (*trl_matvec_fc)(int *pnrow, int *pncol, double *x, int *pldx, double *y, int *pldy);
typedef void (*trl_uo_fc)(int *pnrow, double *x, double *y);
typedef void (*ztrl_matvec_fc)(int *pnrow, int *pncol, trl_dcomplex *x, int *pldx, trl_dcomplex *y,
                               int *pldy);

struct fc_op {
    trl_matvec_fc op;
    ztrl_matvec_fc zop;
};

static _Thread_local trl_uo_fc fc_user_ortho;

static void fc_matvec(const int nrow, const int ncol, const double *x, const int ldx, double *y,
                      const int ldy, void *mvparam)
{
    const struct fc_op *a = (const struct fc_op *)mvparam;
    int nr = nrow, nc = ncol, lx = ldx, ly = ldy;
    a->op(&nr, &nc, (double *)x, &lx, y, &ly);
}

static void fc_zmatvec(const int nrow, const int ncol, const trl_dcomplex *x, const int ldx,
                       trl_dcomplex *y, const int ldy, void *mvparam)
{
    const struct fc_op *a = (const struct fc_op *)mvparam;
    int nr = nrow, nc = ncol, lx = ldx, ly = ldy;
    a->zop(&nr, &nc, (trl_dcomplex *)x, &lx, y, &ly);
}

static void fc_uo(const int nrow, double *x, double *y)
{
    int nr = nrow;
    fc_user_ortho(&nr, x, y);
}

/* The adapter in place of the caller's mvparam for one call. */
#define FC_ENTER(info, a)                        \
    void *const fc_saved_ = (info)->mvparam;     \
    (info)->mvparam = (a)
#define FC_LEAVE(info) (info)->mvparam = fc_saved_

/* Replaces: reference trlan() built with TRL_FORTRAN_COMPATIBLE
 * (INC/trlan.h:9-36, 308-310). */
void trlan_fc(trl_matvec_fc op, trl_uo_fc user_ortho, trl_info *info, int nrow, int mev,
              double *eval, double *evec, int lde, int lwrk, double *wrk)
{
    struct fc_op a = {op, NULL};
    const trl_uo_fc prev = fc_user_ortho;
    fc_user_ortho = user_ortho;
    FC_ENTER(info, &a);
    trlan(fc_matvec, user_ortho ? fc_uo : NULL, info, nrow, mev, eval, evec, lde, lwrk, wrk);
    FC_LEAVE(info);
    fc_user_ortho = prev;
}

/* Replaces: reference trl_check_ritz() in that build (INC/trlan.h:346-349). */
void trl_check_ritz_fc(trl_matvec_fc op, trl_info *info, int nrow, int ncol, double *rvec,
                       int ldrvec, double *alpha, int *check, double *beta, double *eval,
                       double *wrk, int lwrk)
{
    struct fc_op a = {op, NULL};
    FC_ENTER(info, &a);
    trl_check_ritz(fc_matvec, info, nrow, ncol, rvec, ldrvec, alpha, check, beta, eval, wrk, lwrk);
    FC_LEAVE(info);
}

/* Replaces: reference trl_ritz_projection() in that build
 * (INC/trlan.h:557-560). */
void trl_ritz_projection_fc(trl_matvec_fc op, trl_info *info, int mev, double *evec, int lde,
                            double *eres, double *wrk, int lwrk, double *base)
{
    struct fc_op a = {op, NULL};
    FC_ENTER(info, &a);
    trl_ritz_projection(fc_matvec, info, mev, evec, lde, eres, wrk, lwrk, base);
    FC_LEAVE(info);
}

/* Replaces: reference trl_rayleigh_quotients() in that build
 * (INC/trlan.h:591-593). */
void trl_rayleigh_quotients_fc(trl_matvec_fc op, trl_info *info, int ncol, double *evec, int lde,
                               double *eres, double *base)
{
    struct fc_op a = {op, NULL};
    FC_ENTER(info, &a);
    trl_rayleigh_quotients(fc_matvec, info, ncol, evec, lde, eres, base);
    FC_LEAVE(info);
}

/* Replaces: reference ztrlan() in that build (INC/ztrlan.h:30-31, 139-142). */
void ztrlan_fc(ztrl_matvec_fc op, trl_info *info, int nrow, int mev, double *eval,
               trl_dcomplex *evec, int lde, trl_dcomplex *misc, int nmis, double *dwrk, int ldwrk)
{
    struct fc_op a = {NULL, op};
    FC_ENTER(info, &a);
    ztrlan(fc_zmatvec, info, nrow, mev, eval, evec, lde, misc, nmis, dwrk, ldwrk);
    FC_LEAVE(info);
}

/* Replaces: reference ztrl_check_ritz() in that build (INC/ztrlan.h:173-176). */
void ztrl_check_ritz_fc(ztrl_matvec_fc op, trl_info *info, int nrow, int ncol, trl_dcomplex *rvec,
                        int ldrvec, double *alpha, int *check, double *beta, double *eval,
                        trl_dcomplex *wrk, int lwrk)
{
    struct fc_op a = {NULL, op};
    FC_ENTER(info, &a);
    ztrl_check_ritz(fc_zmatvec, info, nrow, ncol, rvec, ldrvec, alpha, check, beta, eval, wrk,
                    lwrk);
    FC_LEAVE(info);
}
