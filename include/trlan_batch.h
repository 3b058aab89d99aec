/*
 * trlan_batch.h -- many independent trlan() solves advancing together on
 * one GPU.  Additive: the reference has no batched entry point; each job is
 * exactly one trlan() call (SRC/trlan.c:21-301) on a device CSR operator.
 */
#ifndef NUTRLAN_AMD_TRLAN_BATCH_H
#define NUTRLAN_AMD_TRLAN_BATCH_H

#include "trlan.h"
#include "trlhip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One solve: the arguments trlan(trlhip_csr_matvec, NULL, info, nrow, mev,
 * eval, evec, lde, lwrk, wrk) would take, with info->mvparam = A (set by
 * trlan_batch).  info is prepared as for trlan() (trl_init_info,
 * trl_set_iguess: guess = 1 takes evec(:, nec) as the start vector). */
typedef struct {
    trlhip_csr *A;     /* real device operator: TRLHIP_R64, or TRLHIP_R32 for
                          the mixed-precision solver */
    trl_info *info;
    int nrow, mev;
    double *eval, *evec;
    int lde, lwrk;
    double *wrk;
    const float *dstart; /* NULL, or an FP32 start vector (nrow entries) in
                            device memory, used instead of evec(:, nec) when
                            info->guess == 1; it must stay valid until
                            trlan_batch returns */
} trl_batch_job;

/* Runs every job to completion.  Each job makes the decisions a lone
 * trlan() call would make (same eigenpairs and counters); their Lanczos
 * steps, restart GEMMs, column copies and final refinements run as one
 * batched launch per phase across the jobs of a lane (two lanes by default,
 * pipelined so the host steps one lane while the device runs the other).  Per-job status in info->stat; returns
 * 0, or a TRLHIP_E* code when the batch itself failed (then every unfinished
 * job's info->stat is TRLHIP_STAT_DEVICE).  Not re-entrant: one batch at a
 * time per process. */
int trlan_batch(trl_batch_job *jobs, int njobs);

#ifdef __cplusplus
}
#endif
#endif
