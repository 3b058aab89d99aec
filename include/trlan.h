/*
 * trlan.h -- public C API of the MI355X thick-restart Lanczos solver.
 *
 * Drop-in for nu-TRLan's real symmetric entry points (reference
 * INC/trlan.h).  Every declaration here keeps the reference's name,
 * signature and C linkage, and `trl_info` keeps the reference's field order
 * and types so the struct is byte-for-byte the same on LP64 (800 bytes,
 * `guess` at offset 52, `mvparam` at offset 792).  Callers built against the
 * reference header link against libnutrlan_amd.so unchanged.
 *
 * What differs inside: the Lanczos basis lives in device memory (HBM) and
 * all O(n) vector work runs as HIP kernels on the GPU; the operator callback
 * is called with host pointers (staged through pinned buffers) unless it is
 * the library's device CSR operator (see trlhip.h), which runs on the GPU.
 */
#ifndef NUTRLAN_AMD_TRLAN_H
#define NUTRLAN_AMD_TRLAN_H

#include <stdio.h>
#include <time.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifdef TRL_FORTRAN_COMPATIBLE
/* The reference's pointer-argument callbacks (INC/trlan.h:9-36): every
 * scalar by address, no mvparam.  The library exports them as separate
 * entry points (trlan_fc ...; the adapters pass the library's own C
 * callbacks through info->mvparam for the duration of a call and restore
 * it), and the reference's names map onto them here, so sources written
 * for the reference's TRL_FORTRAN_COMPATIBLE build compile unchanged. */
typedef void (*trl_matvec)(int *pnrow, int *pncol, double *x, int *pldx,
                           double *y, int *pldy);
typedef void (*trl_uo)(int *pnrow, double *x, double *y);
#define trlan trlan_fc
#define trl_check_ritz trl_check_ritz_fc
#define trl_ritz_projection trl_ritz_projection_fc
#define trl_rayleigh_quotients trl_rayleigh_quotients_fc
#else
/* y(:,j) = A x(:,j) for j < ncol; column j of x starts at x + j*ldx.
 * Replaces: reference INC/trlan.h:67-69 (non-Fortran callback).
 * The solver always calls it with ncol == 1. */
typedef void (*trl_matvec)(const int nrow, const int ncol,
                           const double *x, const int ldx,
                           double *y, const int ldy, void *mvparam);

/* Optional extra orthogonalization of a host vector; y (may be NULL)
 * receives the 2-norm of the change.  Replaces: INC/trlan.h:70. */
typedef void (*trl_uo)(const int nrow, double *x, double *y);
#endif

/* Problem description, progress counters and bookkeeping of one solve.
 * Field order and types are the ABI (reference INC/trlan.h:77-245). */
typedef struct strct_trl_info {
    int stat;          /* 0 or a negative error code                     */
    int lohi;          /* <0 smallest, >0 largest, 0 whichever converges */
    int ned;           /* eigenpairs wanted                              */
    int nec;           /* eigenpairs converged (in: already known)       */
    double tol;        /* residual tolerance relative to ||A||           */

    void *mpicomp;     /* unused (single process); kept for layout       */
    int maxlan;        /* maximum basis size                             */
    int klan;          /* current basis size limit                       */
    int maxmv;         /* maximum operator applications                  */

    int restart;       /* thick-restart scheme 0..8                      */
    int locked;        /* eigenpairs locked                              */
    int guess;         /* <=0 ones(+perturbation), 1 user start vector   */

    int matvec;        /* operator applications                          */
    int nloop;         /* restarts                                       */
    int north;         /* full reorthogonalization passes                */
    int nrand;         /* random vectors injected                        */
    int flop;          /* flop counters (excluding the operator)         */
    int flop_h;
    int flop_r;
    double rflp;
    double rflp_h;
    double rflp_r;

    clock_t clk_rate;  /* timers, in clock ticks + spill seconds         */
    clock_t clk_max;
    clock_t clk_tot;
    clock_t clk_op;
    clock_t clk_orth;
    clock_t clk_res;
    double tick_t;
    double tick_o;
    double tick_h;
    double tick_r;
    int clk_in;
    int wrds_in;
    int clk_out;
    int wrds_out;

    double anrm;       /* running estimate of ||A||                      */

    int my_pe;
    int npes;
    int nloc;          /* local (== global) problem size                 */
    int ntot;

    int verbose;
    double crat;       /* convergence-rate estimate                      */
    double trgt;       /* next target Ritz value                         */
    double tres;       /* its residual                                   */
    int tmv;           /* matvec count when trgt/tres were recorded      */
    double avgm;

    double cfac;       /* history used by the adaptive restart schemes   */
    double ptres;
    double predicted_crate;
    double mgap;
    double mgamma, gamma0;
    double old_target;
    int target_id;
    int old_locked;
    int k1, k2, k;
    double rfact;

    double ref;        /* shift for lohi < -1                            */

    int log_io;
    FILE *log_fp;
    char log_file[128];

    int cpflag, cpio;
    FILE *cpt_fp;
    char cpfile[128], oldcpf[128];

    void *mvparam;     /* passed to the operator untouched               */
} trl_info;

/* Thick-restart Lanczos for a real symmetric operator.
 * Replaces: reference INC/trlan.h:308-310 / SRC/trlan.c:21. */
void trlan(trl_matvec op, trl_uo user_ortho, trl_info *info, int nrow,
           int mev, double *eval, double *evec, int lde, int lwrk,
           double *wrk);

/* Residual/accuracy check of computed Ritz pairs.
 * Replaces: INC/trlan.h:346-349 / SRC/trlan.c:805. */
void trl_check_ritz(trl_matvec op, trl_info *info, int nrow, int ncol,
                    double *rvec, int ldrvec, double *alpha, int *check,
                    double *beta, double *eval, double *wrk, int lwrk);

/* Replaces: INC/trlan.h:395-397 / SRC/trl_comm_none.c:27. */
void trl_init_info(trl_info *info, int nrow, int mxlan, int lohi, int ned,
                   double tol, int restart, int maxmv, void *mpicomp);

/* Replaces: INC/trlan.h:411 / SRC/trlan.c:304. */
void trl_set_restart(trl_info *info, double rfact);

/* Replaces: INC/trlan.h:434 / SRC/trlan.c:325. */
void trl_set_debug(trl_info *info, int msglvl, char *filename);

/* Replaces: INC/trlan.h:452 / SRC/trlan.c:367. */
void trl_set_checkpoint(trl_info *info, int cpflag, char *file);

/* Replaces: INC/trlan.h:484-485 / SRC/trlan.c:405. */
void trl_set_iguess(trl_info *info, int nec, int iguess, int ncps,
                    char *cpf);

/* Replaces: INC/trlan.h:503 / SRC/trlan.c:472. */
void trl_print_info(trl_info *info, int mvflop);

/* Replaces: INC/trlan.h:517 / SRC/trlan.c:727. */
void trl_terse_info(trl_info *info, FILE *iou);

/* Rayleigh-Ritz projection on a set of approximate eigenvectors.
 * Replaces: INC/trlan.h:557-560 / SRC/trlan.c:1438. */
void trl_ritz_projection(trl_matvec op, trl_info *info, int mev,
                         double *evec, int lde, double *eres, double *wrk,
                         int lwrk, double *base);

/* Normalize Ritz vectors and replace Ritz values by Rayleigh quotients;
 * eres[0:ncol) = quotients, eres[ncol:2ncol) = residual norms.
 * Replaces: INC/trlan.h:591-593 / SRC/trlan.c:1111. */
void trl_rayleigh_quotients(trl_matvec op, trl_info *info, int ncol,
                            double *evec, int lde, double *eres,
                            double *base);

#ifdef __cplusplus
}
#endif
#endif /* NUTRLAN_AMD_TRLAN_H */
