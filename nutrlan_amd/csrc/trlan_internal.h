/*
 * trlan_internal.h -- declarations shared between the host driver files
 * and the device layer; not part of the public API.
 */
#ifndef NUTRLAN_AMD_TRLAN_INTERNAL_H
#define NUTRLAN_AMD_TRLAN_INTERNAL_H

#include <stdint.h>

#include "trlan.h"
#include "trlhip.h"
#include "ztrlan.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The device CSR handle behind an operator callback, or NULL when `op` is
 * not this library's device operator for `vtype` or mvparam is not a live
 * handle of that type. */
const trlhip_csr *trlhip_csr_from_op(void (*op)(void), void *mvparam, int vtype);

/* 1 when A was created for the mixed-precision solver (TRLHIP_R32). */
int trlhip_csr_is_mixed(const trlhip_csr *A);

/* Uniform [0,1) from a 64-bit xorshift state (the library's own generator;
 * the reference uses drand48 seeded from clock()). */
double trl_rand_uniform(uint64_t *state);

/* A batch of contexts that step together (trlhip.hip, batch section):
 * their device waits yield through `yield(user)` to a scheduler, which
 * calls trlhip_batch_flush once every waiting context has yielded. */
typedef struct trlhip_batch trlhip_batch;
int trlhip_batch_create(trlhip_batch **b);
/* A batch on the stream of `share` (a lane that interleaves with it on
 * one stream), or on a stream of its own when share is NULL. */
int trlhip_batch_create_on(trlhip_batch **b, trlhip_batch *share);
void trlhip_batch_destroy(trlhip_batch *b);
/* Frees the arena chunks that destroyed batches left for reuse. */
void trlhip_batch_trim(void);
void trlhip_batch_set_yield(trlhip_batch *b, void (*yield)(void *user), void *user);
int trlhip_batch_flush(trlhip_batch *b);
/* trlhip_batch_flush in two halves: launch the recorded work (no wait), and
 * wait for it; the batch's contexts may resume only after the wait. */
int trlhip_batch_launch(trlhip_batch *b);
int trlhip_batch_wait(trlhip_batch *b);
/* 1 when the batch's last launch has finished on the device (or failed),
 * 0 while it runs; never blocks. */
int trlhip_batch_ready(trlhip_batch *b);
/* Work b queues from now on starts on the device after prev's last launch
 * (lanes on their own streams that still run one at a time). */
int trlhip_batch_after(trlhip_batch *b, trlhip_batch *prev);
int trlhip_ctx_create_batched(trlhip_ctx **c, int n, int ncols, int vtype, trlhip_batch *b);
/* The solver allows its next (non-first) batched step to bring up to
 * `steps` - 1 followers (steps jnd + 1, ... on consecutive columns) into the
 * same round: each runs only if the device finds the solver would accept
 * the step before it after its one speculative Gram-Schmidt pass (kept and
 * ntot as lz_orth sees them), and the solver then takes its results from
 * trlhip_step_cgs without a round of its own.  No effect outside a batch. */
int trlhip_step_chain(trlhip_ctx *c, int steps, int kept, int ntot);

/* The reference's basis debug checks on a device basis (trl_debug.c):
 * check vector i is context column i for i < k and column `last` for
 * i == k.  trl_check_orth / ztrl_check_orth over j1 + j2 vectors (j1 in
 * the place of the reference's evec part), and trl_check_recurrence /
 * ztrl_check_recurrence over m1 + m2 vectors with `kept` from the last
 * restart, applying op (E == 1) or zop (E == 2) to host copies.  avail:
 * the reference's workspace at that point (too little: its message). */
void trl_dbg_check_orth(trl_info *info, trlhip_ctx *dev, int E, int j1, int j2, int k, int last,
                        long avail);
void trl_dbg_check_recurrence(trl_info *info, trlhip_ctx *dev, int n, int E, trl_matvec op,
                              ztrl_matvec zop, int m1, int m2, int kept, const double *alpha,
                              const double *beta, int k, int last, long avail);

/* One thick-restart solve on a device basis; cplx selects ztrlan (zop,
 * evec/wrk interleaved complex) or trlan (op).  Implemented in trlan.c. */
void trl_solve_device(trl_matvec op, ztrl_matvec zop, trl_uo user_ortho, trl_info *info,
                      int nrow, int mev, double *eval, double *evec, int lde, int lwrk,
                      double *wrk, int cplx);

#ifdef __cplusplus
}
#endif
#endif
