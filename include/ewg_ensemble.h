/* ewg_ensemble.h -- ensembles of realizations (uneven grids, own permeability fields and
 * injection rates) time-stepped together on one GPU.
 *
 * No counterpart in the reference: eWoms runs one simulation per process
 * (Simulator::run, utils/simulator.hh:660-831). Each member here is an ordinary context
 * (ewg_create, ewg_set_static, ewg_set_boundary, ewg_set_solution), with
 * params.precond_fp32 = 1 for the FP32 preconditioner. When every member is a 3D context
 * created with params.ilu_schedule = 0 (the tiles' wavefront layout) and all use one
 * preconditioner precision, the batched mode applies the preconditioner as tile-wavefront
 * sweeps over all members at once (two launches per application); otherwise (e.g.
 * ilu_schedule = 1, no tile buffers) hyperplane by hyperplane. Every member keeps its own
 * parameters (tolerances, limits).
 *
 * ewg_ensemble_run advances all members by `steps` implicit-Euler steps of dt (Newton
 * with the reference's rules, BiCGStab with its criterion, ILU(0) in natural order by
 * hyperplanes), with each member's Newton and Krylov convergence its own, in one of two
 * modes (below): in the persistent mode one launch, captured once in a HIP graph per
 * (steps, dt, launch shape) and replayed, runs the whole time loop of every member. A
 * failed solve is retried from the step's start with 2, 4, ... substeps (up to
 * params.max_time_step_divisions halvings), as the reference halves a failed step, so the
 * member stays at the same time level as the others; a member whose step still fails
 * keeps its last converged state and stops (its status says why).
 *
 * ewg_ensemble_tune times the candidates on one step from the current state, restoring it
 * afterwards, and keeps the fastest for the following runs (default: batched, replay
 * depth 8): the batched mode with BiCGStab replay depths 4, 8, 16 and 32 (the most
 * iterations of all members queued between two reads of their Krylov states), and the
 * persistent mode with 256, 512 and 1024 threads per member, members largest first, and
 * with 512 threads in creation order. ewg_ensemble_tune_modes restricts the search to the
 * modes in `modes` (bit EWG_ENSEMBLE_TUNE_BATCHED, bit EWG_ENSEMBLE_TUNE_PERSISTENT; a
 * persistent trial runs one whole step of every member in one workgroup each).
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
    int32_t step_divisions;    /* time steps halved after a failed solve */
    double last_error;         /* Newton error of the last solve */
    int32_t precond_retries;   /* batched mode, params.precond_fp32 = 2: solves redone with
                                  FP64 factors (the member's later solves use them too) */
} ewg_ensemble_member_stats;

#define EWG_ENSEMBLE_MAX_CANDIDATES 8

/* Execution modes: EWG_ENSEMBLE_PERSISTENT runs every member's whole time loop in one
 * workgroup of one launch; EWG_ENSEMBLE_BATCHED runs each phase (linearization, norms,
 * ILU(0) hyperplanes, each BiCGStab step) as one launch over the cells of all members,
 * with per-member reductions and per-member Newton and Krylov states, one BiCGStab
 * iteration of all members replayed from a HIP graph. */
#define EWG_ENSEMBLE_PERSISTENT 0
#define EWG_ENSEMBLE_BATCHED 1

typedef struct ewg_ensemble_tuning {
    int32_t mode;              /* EWG_ENSEMBLE_* chosen */
    int32_t threads;           /* persistent mode: workgroup size (threads per member) */
    int32_t largest_first;     /* persistent mode: members ordered by decreasing size */
    int32_t replay;            /* batched mode: BiCGStab replay depth */
    int32_t ncandidates;
    int32_t cand_mode[EWG_ENSEMBLE_MAX_CANDIDATES];
    int32_t cand_threads[EWG_ENSEMBLE_MAX_CANDIDATES];
    int32_t cand_largest_first[EWG_ENSEMBLE_MAX_CANDIDATES];
    int32_t cand_replay[EWG_ENSEMBLE_MAX_CANDIDATES];
    double cand_ms[EWG_ENSEMBLE_MAX_CANDIDATES];  /* one step; < 0: candidate failed or not tried */
} ewg_ensemble_tuning;

#define EWG_ENSEMBLE_TUNE_BATCHED 1
#define EWG_ENSEMBLE_TUNE_PERSISTENT 2

/* An ensemble over n contexts (not owned; they must outlive it). Members linearize by
 * automatic differentiation: a context with params.local_linearizer = 1 is rejected
 * (EWG_INVALID_ARGUMENT). */
int ewg_ensemble_create(ewg_ctx *const *members, int32_t n, ewg_ensemble **out);
void ewg_ensemble_destroy(ewg_ensemble *ens);
const char *ewg_ensemble_last_error(const ewg_ensemble *ens);

int ewg_ensemble_tune(ewg_ensemble *ens, double dt, ewg_ensemble_tuning *out);
int ewg_ensemble_tune_modes(ewg_ensemble *ens, double dt, int32_t modes,
                            ewg_ensemble_tuning *out);
/* Choose the mode without tuning (threads: 256, 512 or 1024 for the persistent mode). */
int ewg_ensemble_set_mode(ewg_ensemble *ens, int32_t mode, int32_t threads);
/* The batched mode's BiCGStab replay depth (1 to 64; default 4): the iterations queued
 * between two host reads of the members' states. Results do not depend on it. */
int ewg_ensemble_set_replay(ewg_ensemble *ens, int32_t depth);
/* The batched mode replays its BiCGStab iteration from a HIP graph (on = 1, the default)
 * or launches the iteration's kernels each time (on = 0). Results do not depend on it;
 * the second form is for tools that follow each launch. */
int ewg_ensemble_set_graph(ewg_ensemble *ens, int32_t on);
/* Compact rounds of the batched mode (members with the tile-wavefront layout): with
 * on = 1 (the default), while the members whose linear solve runs hold at most half of the
 * ensemble's cells, a round's BiCGStab iterations launch over those members' cells and
 * tiles only (direct launches) instead of over every member's (graph replays); on = 2:
 * every round compact; on = 0: none. Results do not depend on it. */
int ewg_ensemble_set_compact(ewg_ensemble *ens, int32_t on);
int ewg_ensemble_run(ewg_ensemble *ens, double dt, int32_t steps, double *seconds);
int ewg_ensemble_stats(const ewg_ensemble *ens, int32_t i,
                       ewg_ensemble_member_stats *out);

/* Test hook of the failure path (no counterpart in the reference): the next `solves`
 * linear solves of member i in the batched mode start their preconditioner sweeps with
 * the member's wavefront-wait error flag already set, as if one of its waits had run
 * out. The member's solve then breaks down (EWG_NUMERICAL_ISSUE; redone with FP64 factors
 * when the member has params.precond_fp32 = 2, else retried with a halved step), its tile
 * edge buffers are reset, and the other members are unaffected. */
int ewg_ensemble_force_sweep_failure(ewg_ensemble *ens, int32_t i, int32_t solves);

#ifdef __cplusplus
}
#endif

#endif /* EWG_ENSEMBLE_H */
