/*
 * trlhip.h -- device layer of the MI355X thick-restart Lanczos solver.
 *
 * Plain C ABI (pointers, sizes, status codes; no HIP or torch types), used
 * by the host driver behind trlan()/ztrlan() and callable directly.  These
 * entry points are additive: the reference has no device layer.  Each one
 * replaces a group of reference BLAS/communication calls on the hot path,
 * cited per function.
 *
 * Data layout in HBM: one Lanczos basis slab per solve, column-major,
 * `ld` elements per column (n rounded up to 64 elements = 512 B for FP64),
 * complex bases store interleaved (re, im) doubles.  Column k of the slab is
 * logical basis vector k: the reference's evec(:,k) for k < mev, else
 * base(:,k-mev) (SRC/trlcore.c:883-910).
 *
 * Status: 0 on success; TRLHIP_E* (negative) on failure.  The solver maps a
 * failure to info->stat = TRLHIP_STAT_DEVICE.
 */
#ifndef NUTRLAN_AMD_TRLHIP_H
#define NUTRLAN_AMD_TRLHIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    TRLHIP_R64 = 0,   /* real FP64 values / basis       */
    TRLHIP_C128 = 1,  /* complex FP64 (interleaved)     */
    TRLHIP_R32 = 2    /* real, mixed precision: FP64 values, basis and
                         recurrence; FP32 shadow of the basis for
                         Gram-Schmidt                    */
};

enum {
    TRLHIP_OK = 0,
    TRLHIP_EINVAL = -1,   /* bad argument                          */
    TRLHIP_ENOMEM = -2,   /* device or pinned allocation failed    */
    TRLHIP_EHIP = -3,     /* a HIP runtime call failed             */
    TRLHIP_ENODEV = -4,   /* no usable GPU                          */
    TRLHIP_ERANGE = -5    /* a size beyond a fixed maximum          */
};

/* info->stat value for a device failure inside trlan()/ztrlan(). */
#define TRLHIP_STAT_DEVICE (-1201)

typedef struct trlhip_csr trlhip_csr;   /* device CSR operator       */
typedef struct trlhip_ctx trlhip_ctx;   /* device basis + workspace  */

/* ---- device CSR operator ------------------------------------------ */

/* Copy a CSR matrix (row_ptr[n+1], col_idx[nnz] int32 in [0, n)) to the
 * device.  values: nnz doubles for TRLHIP_R64 and TRLHIP_R32, nnz
 * interleaved complex doubles for TRLHIP_C128.  TRLHIP_R32 marks the
 * operator for the mixed-precision solver: trlan() then reorthogonalizes
 * against an FP32 shadow of the basis (half the Gram-Schmidt traffic),
 * keeps the recurrence, basis and restart in FP64, and refines each
 * returned pair by an FP64 Rayleigh quotient whose true residual it
 * reports in wrk.  Invalid structure (decreasing row_ptr, column out of
 * range) is rejected with EINVAL.  A real matrix with at most 255
 * distinct off-diagonal values (compared bit for bit), at most one
 * diagonal entry per row and every column within 2^23 rows of its row
 * block also gets a packed copy of its nonzeros (one 32-bit word each
 * instead of 12 bytes, diagonal values apart), which the CSR-stream kernel
 * then reads with bit-identical results; the environment variable
 * TRLHIP_CSR_PACK=0 turns packing off. */
int trlhip_csr_create(trlhip_csr **A, int n, long nnz, const int *row_ptr,
                      const int *col_idx, const void *values, int vtype);

/* 1 when A's CSR-stream SpMV reads packed nonzeros, else 0.  No
 * counterpart in the reference. */
int trlhip_csr_is_packed(const trlhip_csr *A);

/* 1 when A's SpMV in use is the column-panel kernel (see
 * trlhip_csr_set_spmv), else 0.  No counterpart in the reference. */
int trlhip_csr_is_paneled(const trlhip_csr *A);

/* D = 2 or 3 when A is a periodic nearest-neighbour stencil on an L^D
 * lattice (site x + L y (+ L^2 z), L >= 3, each row holding its 2D
 * neighbours and itself with ascending columns, one off-diagonal value bit
 * for bit), else 0; *L (may be NULL) gets L.  Such an operator's CSR-stream
 * SpMV forms its columns from the lattice instead of reading them, with
 * results equal to the CSR kernel's bit for bit; the environment variable
 * TRLHIP_CSR_STENCIL=0 turns that off.  No counterpart in the reference. */
int trlhip_csr_stencil(const trlhip_csr *A, int *L);

/* SpMV autotuner: times the column-panel kernel (when A has panels), the
 * CSR-stream kernel (row blocks) and the vector kernel at 2, 4, 8, 16 and
 * 32 lanes per row on A itself (reps launches each, 0 = 10, HIP events on
 * A's stream) and keeps the fastest for every later use of A, single and
 * batched solves alike.  *lpr_out = the choice (1: column panels, 0:
 * CSR-stream), *ms_out = its time per SpMV (either may be NULL).
 * Opt-in: a different kernel sums each row in a different order, so
 * results can differ from the default kernel's in the last bits.  No
 * counterpart in the reference, whose operator is the caller's own op. */
int trlhip_csr_autotune(trlhip_csr *A, int reps, int *lpr_out, float *ms_out);
/* Sets A's SpMV kernel as trlhip_csr_autotune would: lpr = 1 (column
 * panels, when A has them), 0 (CSR-stream) or 2, 4, 8, 16, 32 lanes per
 * row; for a cached autotuner choice.  No counterpart in the reference.
 *
 * Column panels: with TRLHIP_CSR_PANEL=1 set at creation, a real unpacked
 * operator also gets its nonzeros split into column panels of 256 K
 * columns (TRLHIP_CSR_PANEL_COLS: another width, a multiple of 64; about 16
 * bytes per nonzero more device memory) and its SpMV reads each panel's
 * share of x through one XCD's L2, combining per-row partials in panel
 * order in a second kernel; this changes the sums' grouping (results
 * differ from the CSR-stream kernel's in the last bits; run to run they
 * are the same).  Opt-in only: on a random n = 2^23 pattern the two
 * kernels take 2.94 ms a product against CSR-stream's 2.62 ms.  Batched
 * solves of such operators use the CSR-stream kernel.  The panel kernel
 * keeps its partials in the operator: one such operator serves one solve
 * at a time. */
int trlhip_csr_set_spmv(trlhip_csr *A, int lpr);

/* Gram-Schmidt projection tiles (h = V^T r): columns per block, 4, 8 or
 * 16, for an FP64 basis (cb_fp64) and for the FP32 shadow of a
 * mixed-precision basis (cb_fp32); 0 restores the default 8.  Process-wide,
 * read by every later solve, single and batched alike (so a batch still
 * equals its single solves bit for bit); set it while no solve runs.  A
 * width other than the default sums each coefficient in another order.
 * No counterpart in the reference. */
int trlhip_proj_tiles(int cb_fp64, int cb_fp32);
int trlhip_proj_tiles_get(int *cb_fp64, int *cb_fp32);
/* Autotuner for the projection tiles: times the projection of an n x m
 * basis (and its FP32 shadow) at 4, 8 and 16 columns per block (reps
 * launches each, 0 = 10) and keeps the fastest of each through
 * trlhip_proj_tiles.  cb_out[0..1] and ms_out[0..1] (FP64, FP32) may be
 * NULL.  No counterpart in the reference. */
int trlhip_proj_autotune(int n, int m, int reps, int *cb_out, float *ms_out);
int trlhip_csr_destroy(trlhip_csr *A);

/* trl_matvec-compatible operator for a trlhip_csr passed as mvparam.
 * Called with host vectors it stages x to the device, multiplies there and
 * copies y back.  trlan() recognises this function and applies the matrix
 * to its device-resident basis directly, with no host staging.
 * Replaces: the user `op` callback (SRC/trlcore.c:938-942, 1098-1102). */
void trlhip_csr_matvec(const int nrow, const int ncol, const double *x,
                       const int ldx, double *y, const int ldy,
                       void *mvparam);
/* Same for complex (ztrl_matvec-compatible; x, y are trl_dcomplex*). */
void trlhip_csr_zmatvec(const int nrow, const int ncol, const void *x,
                        const int ldx, void *y, const int ldy, void *mvparam);

/* ---- basis context -------------------------------------------------- */

/* Allocate an n x ncols basis of `vtype` plus reduction scratch. */
int trlhip_ctx_create(trlhip_ctx **c, int n, int ncols, int vtype);
int trlhip_ctx_destroy(trlhip_ctx *c);
int trlhip_ctx_ld(const trlhip_ctx *c);

/* Host <-> device column transfers (elements of the context's vtype;
 * complex columns are 2n doubles).  Replace the reference's memcpy of
 * vectors between evec/base (SRC/trlcore.c:1395-1403). */
int trlhip_col_set(trlhip_ctx *c, int col, const void *x);
int trlhip_col_get(trlhip_ctx *c, int col, void *x);
/* Columns c0 .. c0+k-1 to x (column i at x + i*ldx elements, ldx >= n) in
 * one device-to-host transfer: the returned Ritz vectors (trlcore.c:
 * 1395-1403 copies them one by one). */
int trlhip_cols_get(trlhip_ctx *c, int c0, int k, void *x, long ldx);

/* Column col = the FP32 device vector dsrc (n entries) converted to FP64,
 * without a host round trip (real bases).  dsrc must stay valid until the
 * context's next synchronisation (in a batch: until the solve resumes).
 * No counterpart in the reference, whose start vector is the host
 * evec(:, nec); used for start vectors computed on the device. */
int trlhip_col_set_device_f32(trlhip_ctx *c, int col, const float *dsrc);
int trlhip_col_copy(trlhip_ctx *c, int dst, int src);
int trlhip_col_scale(trlhip_ctx *c, int col, double s);

/* y = A x for basis columns x, y; *xdoty = Re(x^H y) (reduced on device).
 * Replaces: op + trl_ddot + trl_g_sum (SRC/trlcore.c:1098-1108). */
int trlhip_spmv(trlhip_ctx *c, const trlhip_csr *A, int xcol, int ycol,
                double *xdoty);

/* One Lanczos three-term step after r = A q_a is in column rcol:
 *   alpha = Re(q_a^H r);  r -= alpha q_a + beta_prev q_b;  *rr2 = ||r||^2.
 * If A != NULL the product is formed on the device first (y = A q_a into
 * rcol); q_b < 0 means no q_b term.
 * Replaces: SRC/trlcore.c:1098-1117 (op, ddot, g_sum, 2 daxpy) and the
 * ||r||^2 of trl_orth (1865-1871). */
int trlhip_lanczos_step(trlhip_ctx *c, const trlhip_csr *A, int qa, int qb,
                        double beta_prev, int rcol, double *alpha,
                        double *rr2);

/* First step after a restart: r = A q_a (if A != NULL),
 * alpha = Re(q_a^H r), then r -= sum_{i<jnd} coef_i V(:,i) with
 * coef[0:jnd-1) = beta[0:jnd-1) and coef[jnd-1] = alpha; *rr2 = ||r||^2.
 * Replaces: SRC/trlcore.c:938-1017. */
int trlhip_restart_step(trlhip_ctx *c, const trlhip_csr *A, int qa,
                        const double *beta, int jnd, int rcol,
                        double *alpha, double *rr2);

/* One Lanczos step with the first Gram-Schmidt pass done speculatively,
 * all in one device round trip:
 *   r = A q_a (if A != NULL) in column rcol; alpha = Re(q_a^H r);
 *   r -= alpha q_a + beta_prev q_b        (first == 0; q_b < 0: no term)
 *   r -= sum_{i<jnd} coef_i V(:,i)        (first != 0; coef = beta, alpha last)
 *   h = V(:,0:jnd)^H r; s = r - V(:,0:jnd) h into the context's scratch
 *   column (column ncols), leaving r itself in rcol.
 * out = {alpha, ||r||^2, ||h||^2, ||s||^2}.  The caller either adopts s
 * (trlhip_col_scale_copy from column ncols) or keeps working on r.
 * Replaces: SRC/trlcore.c:938-1017 / 1098-1117 plus the first trip of
 * trl_cgs (2993-3032). */
int trlhip_step_cgs(trlhip_ctx *c, const trlhip_csr *A, int first, int qa, int qb,
                    double beta_prev, const double *beta, int jnd, int rcol,
                    double out[4]);

/* dst = s * src; src may be the scratch column (index ncols).  From the
 * scratch column the copy is deferred: the next trlhip_step_cgs whose q_a
 * is dst forms s * src inside its SpMV and stores it to dst; any other
 * entry point on the context applies the copy first.  Replaces the final
 * dscal of trl_cgs (SRC/trlcore.c:3104-3109). */
int trlhip_col_scale_copy(trlhip_ctx *c, int dst, int src, double s);

/* One classical Gram-Schmidt pass against columns [c0, c0+m):
 *   h = V^H r;  r -= V h;  *hh = ||h||^2;  *rr2 = ||r||^2.
 * h is kept on the device; if h_out != NULL it is also returned (m
 * doubles, 2m for complex).
 * Replaces: one trip of the trl_cgs loop (SRC/trlcore.c:2993-3032:
 * trl_g_dot_, two dgemv, two ddot, g_sum). */
int trlhip_cgs_pass(trlhip_ctx *c, int c0, int m, int rcol, double *hh,
                    double *rr2, double *h_out);

/* The same pass with no allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.
 * V: m device columns of n rows, leading dimension ld elements (FP64 for
 * TRLHIP_R64, interleaved complex FP64 for TRLHIP_C128, FP32 for
 * TRLHIP_R32: the FP32 shadow a mixed-precision solve projects against);
 * r: n device FP64 elements (2n for complex), updated in place.  Real
 * bases need ld a multiple of 4; V and r 16-byte aligned.  stream: a
 * hipStream_t, NULL for the default stream; the call synchronizes it.
 * Outputs as trlhip_cgs_pass (h_out: E*m doubles, may be NULL); results
 * equal trlhip_cgs_pass's bit for bit on the same data.  At most
 * TRLHIP_NOALLOC_MAX_COLS columns (else TRLHIP_ERANGE): the reduction
 * partials live in a fixed device array of about 8.4 MiB that each GPU
 * the library runs on holds until the process ends.  Every call on one
 * GPU uses that one array: calls that overlap on one GPU race. */
#define TRLHIP_NOALLOC_MAX_COLS 256
int trlhip_cgs_pass_noalloc(int n, int m, int vtype, const void *V, long ld, void *r,
                            double *hh, double *rr2, double *h_out, void *stream);

/* h = V(:, c0:c0+m)^H r (no update).  Replaces trl_g_dot_
 * (SRC/trl_comm_none.c:238-329). */
int trlhip_proj(trlhip_ctx *c, int c0, int m, int rcol, double *h);

/* r -= V(:, c0:c0+m) h for a host h; *rr2 = ||r||^2 (may be NULL). */
int trlhip_update(trlhip_ctx *c, int c0, int m, const double *h, int rcol,
                  double *rr2);

/* In-place restart update V(:, lck:lck+k) <- V(:, lck:lck+jl) Y with
 * Y jl x k column-major (ldy >= jl), real FP64; complex bases are
 * multiplied as 2n x jl real matrices.  FP64 MFMA.
 * Replaces: trl_ritz_vectors (SRC/trlcore.c:2761-2910). */
int trlhip_ritz(trlhip_ctx *c, int lck, int jl, int k, const double *Y,
                int ldy);

/* FP64 Rayleigh-quotient refinement of the Ritz vectors in columns
 * [0, k) with operator A (real bases): rho[i] = v'Av / v'v and res[i] =
 * ||A v - rho[i] v|| / ||v||, computed on the device (A v goes to free
 * columns after column k-1, never to column `avoid`, which is kept);
 * NaN for a zero column.  One device round trip for all k pairs; in a
 * batch the pairs of every context run together.  No counterpart in the
 * reference: it is the final residual check of a mixed-precision solve,
 * whose pairs the reference's convergence test (trl_convergence_test,
 * SRC/trlcore.c:2149-2305) judges from the recurrence's estimate. */
int trlhip_refine(trlhip_ctx *c, const trlhip_csr *A, int k, int avoid,
                  double *rho, double *res);

/* The two device passes of the Rayleigh-Ritz polish of a mixed-precision
 * solve's Ritz vectors (real bases, FP64 basis columns and operator), one
 * device round trip each.  trlhip_residuals: column k+i = A v_i - rho[i]
 * v_i for the columns v_i, i < k, and rr2[i] = its squared norm (the same
 * kernels and sums as trlhip_spmv then trlhip_update per column).
 * trlhip_gram: G = Q'Q and H = Q'AQ (m x m, column-major) over the
 * columns Q = [0, m), with A Q_j formed in column tmp >= m: H by the
 * same sums as trlhip_proj per column, G (for m <= 32) in one pass over Q
 * with fixed-order per-block sums, exactly symmetric, else as H).  No
 * counterpart in the reference: the polish
 * follows the final refinement (trlhip_refine). */
int trlhip_residuals(trlhip_ctx *c, const trlhip_csr *A, int k, const double *rho,
                     double *rr2);
int trlhip_gram(trlhip_ctx *c, const trlhip_csr *A, int m, int tmp, double *G, double *H);

/* ---- measurement ---------------------------------------------------- */

/* Kernel classes timed with HIP events on the solver stream when enabled. */
enum {
    TRLHIP_K_SPMV = 0,
    TRLHIP_K_RECUR = 1,
    TRLHIP_K_PROJ = 2,
    TRLHIP_K_UPDATE = 3,
    TRLHIP_K_RITZ = 4,
    TRLHIP_K_MISC = 5,
    /* the batched kernels of trlan_batch (one launch for many solves) */
    TRLHIP_K_BSPMV = 6,
    TRLHIP_K_BRECUR = 7,
    TRLHIP_K_BPROJ = 8,
    TRLHIP_K_BUPDATE = 9,
    TRLHIP_K_BRITZ = 10,
    TRLHIP_K_COUNT = 11
};

/* Per-kernel HIP-event timing: on = 0 disables, on = N > 0 times every
 * N-th launch of each kernel kind (1: every launch).  Resets counters. */
void trlhip_stats_enable(int on);
/* out[k*3 + 0] = timed launches, out[k*3+1] = their total ms,
 * out[k*3+2] = their algorithmic bytes, for k < TRLHIP_K_COUNT.  Returns
 * TRLHIP_K_COUNT. */
int trlhip_stats_get(double *out, int nout);
/* out[k] = floating-point operations of the timed launches of kernel class
 * k (the restart GEMMs: 2 n jl k per launch, complex bases 4 n jl k as a
 * 2n-row real product; 0 for the bandwidth-bound classes).  Returns
 * TRLHIP_K_COUNT. */
int trlhip_stats_flops(double *out, int nout);

/* Number of visible GPUs (0 when none); never aborts. */
int trlhip_device_count(void);

#ifdef __cplusplus
}
#endif
#endif /* NUTRLAN_AMD_TRLHIP_H */
