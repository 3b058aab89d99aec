/*
 * Test-only harness around the compiled reference nu-TRLan (oracle/_ref).
 *
 * Exposes plain C entry points (ctypes-friendly) that run the reference's
 * trlan()/ztrlan() on a CSR operator with an explicit start vector
 * (info.guess = 1, so the reference's clock()-seeded RNG is never used on the
 * start), and report eigenvalues, Ritz vectors, residual estimates and the
 * integer trajectory (matvec, nloop, north, nrand, locked).  Wall time is taken
 * with CLOCK_MONOTONIC around the solver call (trl_info's clk_* fields are
 * process CPU time).
 *
 * The CSR operator is this harness's own code (the reference ships operators
 * only as example diagonals): y = A x, OpenMP over rows when nthreads > 1.
 */
#include <stdlib.h>
#include <string.h>
#include <stddef.h>
#include <time.h>
#include <stdio.h>
#include <omp.h>
#include "trl_map.h"
#include "trlan.h"
#include "ztrlan.h"

typedef struct {
    int n;
    const int *rp, *ci;
    const double *v;      /* real values, or interleaved (re,im) pairs */
    int nthreads;
} ref_csr;

static void op_csr_d(const int nrow, const int ncol, const double *x,
                     const int ldx, double *y, const int ldy, void *p)
{
    const ref_csr *A = (const ref_csr *)p;
    for (int c = 0; c < ncol; ++c) {
        const double *xc = x + (size_t)c * ldx;
        double *yc = y + (size_t)c * ldy;
#pragma omp parallel for schedule(static) num_threads(A->nthreads) if (A->nthreads > 1)
        for (int i = 0; i < nrow; ++i) {
            double s = 0.0;
            for (int k = A->rp[i]; k < A->rp[i + 1]; ++k)
                s += A->v[k] * xc[A->ci[k]];
            yc[i] = s;
        }
    }
}

static void op_csr_z(const int nrow, const int ncol, const trl_dcomplex *x,
                     const int ldx, trl_dcomplex *y, const int ldy, void *p)
{
    const ref_csr *A = (const ref_csr *)p;
    for (int c = 0; c < ncol; ++c) {
        const trl_dcomplex *xc = x + (size_t)c * ldx;
        trl_dcomplex *yc = y + (size_t)c * ldy;
#pragma omp parallel for schedule(static) num_threads(A->nthreads) if (A->nthreads > 1)
        for (int i = 0; i < nrow; ++i) {
            double sr = 0.0, si = 0.0;
            for (int k = A->rp[i]; k < A->rp[i + 1]; ++k) {
                double ar = A->v[2 * k], ai = A->v[2 * k + 1];
                double xr = xc[A->ci[k]].r, xi = xc[A->ci[k]].i;
                sr += ar * xr - ai * xi;
                si += ar * xi + ai * xr;
            }
            yc[i].r = sr;
            yc[i].i = si;
        }
    }
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* info->guess for the solve: 1 (the start vector in evec(:,0)) unless the
 * environment variable TRL_REF_GUESS says otherwise (0: the reference's
 * all-ones start; 2: resume from a checkpoint). */
static int ref_guess(void)
{
    const char *g = getenv("TRL_REF_GUESS");
    return g ? atoi(g) : 1;
}

/* Checkpoint settings from the environment: TRL_REF_CPFLAG and
 * TRL_REF_CPFILE (write, trl_set_checkpoint), TRL_REF_OLDCPF (the prefix a
 * guess = 2 solve resumes from), then info->guess as ref_guess() says;
 * TRL_REF_VERBOSE with TRL_REF_LOGFILE: trl_set_debug (the reference
 * writes its log to that prefix + the PE number). */
static void ref_setup(trl_info *info)
{
    static char cpf[160];
    const char *v = getenv("TRL_REF_VERBOSE"), *lf = getenv("TRL_REF_LOGFILE");
    if (v && lf) trl_set_debug(info, atoi(v), (char *)lf);
    const char *f = getenv("TRL_REF_CPFILE"), *fl = getenv("TRL_REF_CPFLAG");
    if (f && fl) trl_set_checkpoint(info, atoi(fl), (char *)f);
    const char *old = getenv("TRL_REF_OLDCPF");
    if (old) strncpy(info->oldcpf, old, sizeof info->oldcpf - 1);
    trl_set_iguess(info, 0, ref_guess(), 0, cpf);
}

static void pack_counters(const trl_info *info, int *iout, double *dout)
{
    iout[0] = info->stat;
    iout[1] = info->nec;
    iout[2] = info->matvec;
    iout[3] = info->nloop;
    iout[4] = info->north;
    iout[5] = info->nrand;
    iout[6] = info->locked;
    iout[7] = info->klan;
    iout[8] = info->maxlan;
    iout[9] = info->ned;
    dout[0] = info->anrm;
    dout[2] = info->crat;
    dout[3] = info->tol;
}

/* iout[10]: stat nec matvec nloop north nrand locked klan maxlan ned
   dout[4]: anrm wall_s crat tol
   evec: n*mev (column c at c*n); column 0 holds the start vector on entry.
   resid: mev residual-norm estimates (the reference's wrk[0:mev]). */
int ref_trlan_csr(int n, const int *rp, const int *ci, const double *val,
                  int lohi, int ned, int maxlan, double tol, int restart,
                  int maxmv, int mev, const double *start, int nthreads,
                  double *eval, double *evec, double *resid,
                  int *iout, double *dout)
{
    ref_csr A = {n, rp, ci, val, nthreads};
    trl_info info;
    memset(&info, 0, sizeof info);
    trl_init_info(&info, n, maxlan, lohi, ned, tol, restart, maxmv, NULL);
    ref_setup(&info);
    info.mvparam = &A;
    if (getenv("TRL_REF_VERBOSE")) info.verbose = atoi(getenv("TRL_REF_VERBOSE"));
    int ldb = ((n + 3) / 4) * 4 + 4;
    int nbas = maxlan - mev + 1 > 1 ? maxlan - mev + 1 : 1;
    int lwrk = nbas * ldb + maxlan * (maxlan + 10);
    double *wrk = (double *)calloc((size_t)lwrk, sizeof(double));
    memset(eval, 0, sizeof(double) * mev);
    memset(evec, 0, sizeof(double) * (size_t)n * mev);
    memcpy(evec, start, sizeof(double) * n);
    double t0 = now_s();
    trlan(op_csr_d, NULL, &info, n, mev, eval, evec, n, lwrk, wrk);
    double t1 = now_s();
    memcpy(resid, wrk, sizeof(double) * mev);
    free(wrk);
    pack_counters(&info, iout, dout);
    dout[1] = t1 - t0;
    return info.stat;
}

/* Complex Hermitian: val holds nnz interleaved (re,im) pairs, start/evec are
   interleaved complex (n*mev pairs). */
int ref_ztrlan_csr(int n, const int *rp, const int *ci, const double *val,
                   int lohi, int ned, int maxlan, double tol, int restart,
                   int maxmv, int mev, const double *start, int nthreads,
                   double *eval, double *evec, double *resid,
                   int *iout, double *dout)
{
    ref_csr A = {n, rp, ci, val, nthreads};
    trl_info info;
    memset(&info, 0, sizeof info);
    trl_init_info(&info, n, maxlan, lohi, ned, tol, restart, maxmv, NULL);
    ref_setup(&info);
    info.mvparam = &A;
    int nmis = 10 * n > maxlan * (maxlan + 10) ? 10 * n : maxlan * (maxlan + 10);
    int ldwrk = maxlan * (maxlan + 10);
    trl_dcomplex *misc = (trl_dcomplex *)calloc((size_t)nmis, sizeof(trl_dcomplex));
    double *dwrk = (double *)calloc((size_t)ldwrk, sizeof(double));
    trl_dcomplex *z = (trl_dcomplex *)evec;
    memset(eval, 0, sizeof(double) * mev);
    memset(z, 0, sizeof(trl_dcomplex) * (size_t)n * mev);
    memcpy(z, start, sizeof(trl_dcomplex) * n);
    double t0 = now_s();
    ztrlan(op_csr_z, &info, n, mev, eval, z, n, misc, nmis, dwrk, ldwrk);
    double t1 = now_s();
    for (int i = 0; i < mev; ++i) resid[i] = misc[i].r;
    free(misc);
    free(dwrk);
    pack_counters(&info, iout, dout);
    dout[1] = t1 - t0;
    return info.stat;
}

/* Layout probe of the reference's trl_info (LP64): sizeof, then the offsets
   of stat lohi ned nec tol mpicomp maxlan klan maxmv restart locked guess
   matvec nloop north nrand anrm nloc ntot verbose trgt tres ref log_fp
   log_file cpfile oldcpf mvparam. */
int ref_info_layout(long *out, int nout)
{
    long v[] = {
        (long)sizeof(trl_info),
        offsetof(trl_info, stat), offsetof(trl_info, lohi),
        offsetof(trl_info, ned), offsetof(trl_info, nec),
        offsetof(trl_info, tol), offsetof(trl_info, mpicomp),
        offsetof(trl_info, maxlan), offsetof(trl_info, klan),
        offsetof(trl_info, maxmv), offsetof(trl_info, restart),
        offsetof(trl_info, locked), offsetof(trl_info, guess),
        offsetof(trl_info, matvec), offsetof(trl_info, nloop),
        offsetof(trl_info, north), offsetof(trl_info, nrand),
        offsetof(trl_info, anrm), offsetof(trl_info, nloc),
        offsetof(trl_info, ntot), offsetof(trl_info, verbose),
        offsetof(trl_info, trgt), offsetof(trl_info, tres),
        offsetof(trl_info, ref), offsetof(trl_info, log_fp),
        offsetof(trl_info, log_file), offsetof(trl_info, cpfile),
        offsetof(trl_info, oldcpf), offsetof(trl_info, mvparam),
    };
    int k = (int)(sizeof v / sizeof v[0]);
    for (int i = 0; i < k && i < nout; ++i) out[i] = v[i];
    return k;
}
