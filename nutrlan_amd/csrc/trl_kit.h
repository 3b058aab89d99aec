/*
 * trl_kit.h -- host FP64 kit of the thick-restart Lanczos driver: the small
 * (at most maxlan x maxlan) dense eigen-problems, paired sorts, convergence
 * test, restart policy and locking.  Everything here runs on the host in
 * FP64; the O(n) work lives on the device (trlhip.h).
 *
 * Exported with the trlkit_ prefix so the CPU tests can check each piece
 * against the oracle; not part of the reference API.
 */
#ifndef NUTRLAN_AMD_TRL_KIT_H
#define NUTRLAN_AMD_TRL_KIT_H

#include "trlan.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Paired Shell sort keys (reference SRC/dsort2.c). */
enum {
    TRLKIT_BY_VALUE = 0,   /* dsort2   */
    TRLKIT_BY_ABS = 1,     /* dsort2a  */
    TRLKIT_BY_DIST = 2,    /* dsort2s  */
    TRLKIT_ABOVE = 3,      /* dsort2su */
    TRLKIT_BELOW = 4       /* dsort2sd */
};

/* Sort a1[0:n) ascending by `key` (s = reference point for DIST/ABOVE/
 * BELOW), applying the same swaps to a2 and, if non-NULL, to tag. */
void trlkit_sort2(int n, int key, double s, double *a1, double *a2, int *tag);

/* Implicit QL with Wilkinson shifts on the symmetric tridiagonal (d[0:n),
 * off-diagonal e[0:n-1)); d returns the eigenvalues ascending.
 * want = 0: no vectors; 1: z[0:n) = last row of the eigenvector matrix;
 * 2: z (n x n, ld n, column-major) holds on entry a basis Q and on return
 * Q * eigenvectors.  e is destroyed.  Returns 0 or the index of an
 * eigenvalue that failed to converge (> 0). */
int trlkit_tridiag_ql(int n, double *d, double *e, double *z, int want);

/* Householder reduction of a symmetric n x n matrix (column-major, upper
 * triangle read) to tridiagonal T = Q^T A Q, working from the last column
 * so that Q leaves the last coordinate fixed (the reference's
 * dsytrd('U') + dorgtr convention).  a is overwritten by Q; d[0:n),
 * e[0:n-1) receive T. */
void trlkit_householder(int n, double *a, double *d, double *e);

/* All eigenpairs of a symmetric n x n matrix (upper triangle read):
 * a <- eigenvectors, w <- eigenvalues ascending.  Returns 0 on success. */
int trlkit_syev(int n, double *a, double *w);

/* Restatements of the reference's host decisions, same arguments. */
void trlkit_convergence_test(int nd, double *lambda, double *res, int *tag,
                             trl_info *info);
int trlkit_shuffle_eig(int nd, int mnd, double *lambda, double *res,
                       int *tag, trl_info *info, int kept);
int trlkit_set_locking(int jnd, int nlam, double *lambda, double *res,
                       double *yy, double anrm);
void trlkit_sort_eig(int nd, int lohi, int nec, double ref, double *lambda,
                     double *res, int *tag);

/* The reference's trl_smooth_rr (SRC/trlcore.c:1534-1575), end quirks kept. */
void trlkit_smooth_rr(int n, double *rr);

#ifdef __cplusplus
}
#endif
#endif
