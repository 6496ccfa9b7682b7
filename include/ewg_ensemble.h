/* ewg_ensemble.h -- ensembles of realizations (uneven grids, own permeability fields and
 * injection rates) time-stepped together on one GPU.
 *
 * No counterpart in the reference: eWoms runs one simulation per process
 * (Simulator::run, utils/simulator.hh:660-831). Each member here is an ordinary context
 * (ewg_create, ewg_set_static, ewg_set_boundary, ewg_set_solution), created with
 * params.ilu_schedule = 1 (the ensemble solver has its own ILU(0) schedule and needs no
 * tile buffers) and, for the FP32 preconditioner, params.precond_fp32 = 1. Every member
 * keeps its own parameters (tolerances, limits).
 *
 * ewg_ensemble_run advances all members by `steps` implicit-Euler steps of dt in one
 * persistent launch: one workgroup per member runs that member's whole time loop
 * (Newton with the reference's rules, BiCGStab with its criterion, ILU(0) in natural
 * order by hyperplanes), so Newton and Krylov convergence are decided per member on the
 * device, and members run independently to the end. The launch is captured once in a
 * HIP graph per (steps, dt, launch shape) and replayed. A member whose Newton solve fails
 * keeps its last converged state and stops (its status says why); the reference would
 * retry with half the time step, which a lock-step ensemble does not do.
 *
 * ewg_ensemble_tune times the candidate launch shapes on one step from the current state
 * (restoring it afterwards) and keeps the fastest for the following runs.
 */
#ifndef EWG_ENSEMBLE_H
#define EWG_ENSEMBLE_H

#include <stdint.h>

#include "ewg_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ewg_ensemble ewg_ensemble;

typedef struct ewg_ensemble_member_stats {
    int64_t steps;             /* time steps completed */
    int64_t newton_iterations; /* linearizations */
    int64_t linear_iterations; /* BiCGStab iterations */
    int64_t linear_solves;
    int32_t status;            /* 0 running; else the EWG_* code of the failed solve */
    int32_t pad_;
    double last_error;         /* Newton error of the last solve */
} ewg_ensemble_member_stats;

#define EWG_ENSEMBLE_MAX_CANDIDATES 8

typedef struct ewg_ensemble_tuning {
    int32_t threads;           /* workgroup size chosen (threads per member) */
    int32_t largest_first;     /* 1: members ordered by decreasing cell count */
    int32_t ncandidates;
    int32_t pad_;
    int32_t cand_threads[EWG_ENSEMBLE_MAX_CANDIDATES];
    int32_t cand_largest_first[EWG_ENSEMBLE_MAX_CANDIDATES];
    double cand_ms[EWG_ENSEMBLE_MAX_CANDIDATES];  /* one step; < 0: candidate failed */
} ewg_ensemble_tuning;

/* An ensemble over n contexts (not owned; they must outlive it). */
int ewg_ensemble_create(ewg_ctx *const *members, int32_t n, ewg_ensemble **out);
void ewg_ensemble_destroy(ewg_ensemble *ens);
const char *ewg_ensemble_last_error(const ewg_ensemble *ens);

int ewg_ensemble_tune(ewg_ensemble *ens, double dt, ewg_ensemble_tuning *out);
int ewg_ensemble_run(ewg_ensemble *ens, double dt, int32_t steps, double *seconds);
int ewg_ensemble_stats(const ewg_ensemble *ens, int32_t i,
                       ewg_ensemble_member_stats *out);

#ifdef __cplusplus
}
#endif

#endif /* EWG_ENSEMBLE_H */
