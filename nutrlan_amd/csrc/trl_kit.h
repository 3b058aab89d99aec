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

#include <stddef.h>

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
/* The convergence test leaving in wrk[0:2nd) what the reference's leaves
 * in its workspace: |res| then lambda, in the sorted order. */
void trlkit_convergence_test_w(int nd, double *lambda, double *res, int *tag,
                               trl_info *info, double *wrk);
/* lambda and res are windows of one array laid out as the reference's
 * workspace (SRC/trlcore.c:1200-1201, 1256, 1305-1306): the restart schemes
 * read one element before and after them (lambda[-1], res[-1], lambda[nd],
 * res[nd]), as the reference does. */
int trlkit_shuffle_eig(int nd, int mnd, double *lambda, double *res,
                       int *tag, trl_info *info, int kept);
int trlkit_set_locking(int jnd, int nlam, double *lambda, double *res,
                       double *yy, double anrm);
void trlkit_sort_eig(int nd, int lohi, int nec, double ref, double *lambda,
                     double *res, int *tag);

/* The reference's trl_smooth_rr (SRC/trlcore.c:1534-1575), end quirks kept. */
void trlkit_smooth_rr(int n, double *rr);
/* The reference's trl_smooth_zz (SRC/ztrlcore.c:1289-1332) on interleaved
 * complex z[0:2n), its last-element quirk kept. */
void trlkit_smooth_zz(int n, double *z);

/* Checkpoint files, the reference's format (SRC/trlaux.c:780-1014): int
 * nrow, int jnd, alpha[jnd], beta[jnd], then columns of nrow doubles,
 * native byte order.  Columns are streamed through a callback so they can
 * come from (or go to) device memory: col(u, i, buf) fills buf with column
 * i for writing / stores buf as column i when reading, returning 0.
 *
 * Write ncol columns (the reference writes jnd + 1: the basis and the
 * residual).  Returns 0, -221 (open), -222 (write) or -223 (close). */
int trlkit_write_checkpoint(const char *file, int nrow, int jnd, const double *alpha,
                            const double *beta, int ncol,
                            int (*col)(void *u, int i, double *buf), void *u, double *buf);
/* Read like the reference's trl_read_checkpoint: check nrow (-213) and
 * jnd <= maxjnd (-214), read alpha/beta[jnd]; with j1 = min(mev, jnd) read
 * j1 + 1 columns when j1 < mev, else jnd columns (the residual then is not
 * in the file's prefix that is read).  *jnd_out = jnd.  Returns 0, -212
 * (open), -213, -214, -215/-216 (short read). */
int trlkit_read_checkpoint(const char *file, int nrow, int mev, int maxjnd, int *jnd_out,
                           double *alpha, double *beta,
                           int (*col)(void *u, int i, double *buf), void *u, double *buf);
/* The same with E doubles per element: E = 2 is the complex format of the
 * reference's ztrl_write_checkpoint / ztrl_read_checkpoint
 * (SRC/ztrlaux.c:69-307): the same header and real alpha/beta, then
 * columns of nrow interleaved (re, im) pairs. */
int trlkit_write_checkpoint_e(const char *file, int nrow, int E, int jnd, const double *alpha,
                              const double *beta, int ncol,
                              int (*col)(void *u, int i, double *buf), void *u, double *buf);
int trlkit_read_checkpoint_e(const char *file, int nrow, int E, int mev, int maxjnd, int *jnd_out,
                             double *alpha, double *beta,
                             int (*col)(void *u, int i, double *buf), void *u, double *buf);
/* The reference's per-process file name (SRC/trlaux.c:1067-1104) for one
 * process: base up to its first blank, then rank 0 as one digit. */
void trlkit_pe_filename(char *out, size_t cap, const char *base);

#ifdef __cplusplus
}
#endif
#endif
