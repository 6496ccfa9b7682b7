/*
 * ewg_capi.h -- C ABI of the MI355X-native ECFV two-phase Newton step.
 *
 * The library (libewg.so) replaces two compile-time plugins of eWoms'
 * ewoms-numerics for the immiscible two-phase model on a structured grid:
 *
 *   - the linearizer  FvBaseLinearizer<TypeTag>
 *       (ewoms/numerics/discretizations/common/fvbaselinearizer.hh:62-571)
 *     together with the AD local linearizer it drives
 *       (ewoms/numerics/discretizations/common/fvbaseadlocallinearizer.hh:91-308);
 *   - the linear-solver backend ParallelBiCGStabSolverBackend<TypeTag>
 *       (ewoms/numerics/linear/parallelbicgstabbackend.hh:81-163,
 *        ewoms/numerics/linear/parallelbasebackend.hh:156-440)
 *     with BiCGStab (ewoms/numerics/linear/bicgstabsolver.hh:116-326),
 *     CombinedCriterion (ewoms/numerics/linear/combinedcriterion.hh:51-232)
 *     and block ILU(0) (ewoms/numerics/linear/istlpreconditionerwrappers.hh:155-193).
 *
 * Conventions
 *   - Plain C types only; no pointer passed in is retained after the call.
 *   - Status: 0 = OK, > 0 = numerical issue (the host adapter maps it to
 *     Ewoms::NumericalIssue), < 0 = API misuse. ewg_last_error() explains.
 *   - Solution and residual vectors are AoS [p_w, S_w] per cell, cells in
 *     lexicographic order (x fastest), like Dune::BlockVector<FieldVector<double,2>>.
 *   - Jacobian blocks are 2x2 row-major [eq][pv], eq k = mass balance of phase k,
 *     pv 0 = p_w, pv 1 = S_w (Dune::FieldMatrix<double,2,2>).
 *   - One context per host thread; the context owns every device buffer.
 */
#ifndef EWG_CAPI_H
#define EWG_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EWG_ABI_VERSION 1

enum {
    EWG_OK = 0,
    EWG_NUMERICAL_ISSUE = 1,     /* breakdown, non-finite values, singular block */
    EWG_NOT_CONVERGED = 2,       /* linear solver or Newton did not converge */
    EWG_INVALID_ARGUMENT = -1,
    EWG_DEVICE_ERROR = -2,
    EWG_INVALID_STATE = -3
};

enum { EWG_LAW_VAN_GENUCHTEN = 0, EWG_LAW_BROOKS_COREY = 1 };

enum { EWG_BC_NO_FLOW = 0, EWG_BC_FREE_FLOW = 1, EWG_BC_MASS_RATE = 2 };

/* Structured Cartesian grid. dim = 2 uses (nx, ny) with nz = 1 and extrusion 1.
 * Replaces the StructuredGridVanguard INTERVAL grid
 * (ewoms/numerics/io/structuredgridvanguard.hh:136-175) with YaspGrid ordering. */
typedef struct ewg_grid_desc {
    int32_t dim;
    int32_t nx, ny, nz;
    double lx, ly, lz;
} ewg_grid_desc;

/* Two incompressible liquid phases: index 0 = wetting (water), 1 = non-wetting.
 * Mirrors TwoPhaseImmiscibleFluidSystem<LiquidPhase, LiquidPhase>
 * (tests/problems/lensproblem.hh:87-104) and the problem gravity
 * (ewoms/numerics/common/multiphasebaseproblem.hh:393-398). */
typedef struct ewg_fluid_desc {
    double rho[2];
    double mu[2];
    double gravity[3];
} ewg_fluid_desc;

/* Material law EffToAbsLaw<Regularized{VanGenuchten,BrooksCorey}>.
 * region_params[r*4 + 0..3]:
 *   van Genuchten: alpha [1/Pa], n, Swr, Snr
 *   Brooks-Corey:  pe [Pa] (used when no per-cell pe is given), lambda, Swr, Snr */
typedef struct ewg_material_desc {
    int32_t law;
    int32_t nregions;
    const double *region_params;
} ewg_material_desc;

/* Run-time parameters with the reference defaults
 * (newtonmethod.hh:150-155, fvbasediscretization.hh:128,207,210,249,252,
 *  parallelbasebackend.hh:449-514, parallelbicgstabbackend.hh:49). */
typedef struct ewg_params {
    double newton_tolerance;           /* NewtonTolerance = 1e-8 */
    double newton_max_error;           /* NewtonMaxError = 1e100 */
    int32_t newton_target_iterations;  /* NewtonTargetIterations = 10 */
    int32_t newton_max_iterations;     /* NewtonMaxIterations = 18 */
    double linear_tolerance;           /* LinearSolverTolerance = 1e-3 (relative) */
    double linear_abs_tolerance;       /* LinearSolverAbsTolerance = -1 -> newton_tolerance/100 */
    double linear_max_error;           /* LinearSolverMaxError = 1e7 */
    int32_t linear_max_iterations;     /* LinearSolverMaxIterations = 1000 */
    int32_t enable_storage_cache;      /* EnableStorageCache (lens problem: true) */
    int32_t recycle_first_iteration_storage; /* FvBaseProblem::recycleFirstIterationStorage = true */
    int32_t max_time_step_divisions;   /* MaxTimeStepDivisions = 10 */
    double min_time_step_size;         /* MinTimeStepSize = 0 */
    double max_time_step_size;         /* MaxTimeStepSize = inf */
    int32_t precond_fp32;              /* 0: FP64 ILU0 (library default); 1: FP32 ILU0 factors
                                          and preconditioner inside FP64 BiCGStab; 2: the same,
                                          and a linear solve that fails with them (not
                                          converged, breakdown) is redone with FP64 factors,
                                          which the context then keeps for its later solves
                                          (until the next ewg_set_params) */
    int32_t ilu_schedule;              /* ILU(0) schedule on 3D grids (2D grids always use
                                          hyperplane launches); both give the same results:
                                          0 tile wavefront, 8x8 tiles with an edge wavefront
                                          each, factors in the tiles' wavefront layout
                                          (default); 1 one launch per hyperplane
                                          x+y+z = const. Schedule 0 needs its buffers, which
                                          ewg_create allocates unless it is given 1 */
    int32_t kernel_events;             /* 1: bracket sampled launches with device events (the
                                          first preconditioner apply and SpMV of each batch
                                          of BiCGStab iterations, every factorization and
                                          linearization); read with ewg_kernel_samples */
    int32_t linearize_fp32;            /* 0: FP64 linearization (library default); 1: face
                                          fluxes and their derivatives in FP32, accumulated
                                          into the FP64 residual and Jacobian (intensive
                                          quantities, storage terms and pressure differences
                                          stay FP64) */
    int32_t warm_start;                /* 1: start each Newton solve from the prediction of
                                          the network attached with ewg_set_warm_start
                                          (ewg_cnn.h); 0: from the previous time level */
    int32_t local_linearizer;          /* LocalLinearizerSplice: 0 automatic differentiation
                                          (AutoDiffLocalLinearizer, default); 1 finite
                                          differences (FiniteDifferenceLocalLinearizer,
                                          fvbasefdlocallinearizer.hh:131-515) in FP64 */
    int32_t numeric_difference_method; /* NumericDifferenceMethod of the finite differences:
                                          +1 forward (default), 0 central, -1 backward
                                          (fvbasefdlocallinearizer.hh:83) */
    int32_t operator_fp32;             /* 0: BiCGStab's operator products v = J y read the FP64
                                          Jacobian (library default); 1: they read a copy of
                                          the Jacobian's blocks rounded to FP32 once per linear
                                          solve (products and sums in FP64, the recurrences
                                          unchanged): the FP32 SpMV of the mixed-precision
                                          form. Not available to ensemble members */
} ewg_params;

/* Counters and timers of the last calls; the split follows
 * FvBaseProblem::finalize (ewoms/numerics/discretizations/common/fvbaseproblem.hh:452-500). */
typedef struct ewg_stats {
    int64_t newton_iterations;      /* linearizations (NewtonMethod::numIterations summed) */
    int64_t linear_iterations;      /* BiCGStab iterations summed */
    int64_t linear_solves;
    int64_t time_steps;             /* accepted time steps */
    int64_t failed_steps;           /* Newton failures that triggered a dt division */
    int32_t last_newton_iterations;
    int32_t last_linear_iterations;
    double last_newton_error;
    double time_linearize_s;
    double time_solve_s;
    double time_update_s;
    int64_t precond_retries;        /* linear solves redone with FP64 factors (precond_fp32 = 2) */
} ewg_stats;

typedef struct ewg_ctx ewg_ctx;

void ewg_default_params(ewg_params *params);

/* FvBaseLinearizer::init + ParallelBaseBackend ctor (fvbaselinearizer.hh:127-131,
 * parallelbasebackend.hh:170-200). Allocates all device memory for the grid. */
int ewg_create(const ewg_grid_desc *grid, const ewg_fluid_desc *fluid,
               const ewg_material_desc *material, const ewg_params *params,
               ewg_ctx **ctx_out);
void ewg_destroy(ewg_ctx *ctx);
const char *ewg_last_error(const ewg_ctx *ctx);
int ewg_set_params(ewg_ctx *ctx, const ewg_params *params);

int64_t ewg_num_cells(const ewg_ctx *ctx);
int64_t ewg_num_blocks(const ewg_ctx *ctx);          /* nnzb of the BCRS Jacobian */
int64_t ewg_num_boundary_faces(const ewg_ctx *ctx);

/* Static per-cell fields, the problem callbacks intrinsicPermeability / porosity /
 * materialLawParams (tests/problems/lensproblem.hh:357-398). K is isotropic [m^2];
 * region indexes material->region_params; pe (Brooks-Corey only) may be NULL. */
int ewg_set_static(ewg_ctx *ctx, const double *K, const double *phi,
                   const int32_t *region, const double *pe);

/* Boundary conditions per boundary face, the problem callback boundary()
 * (tests/problems/lensproblem.hh:464-533). Faces are grouped by side
 * (x-, x+, y-, y+, z-, z+); within a side they are ordered lexicographically over the
 * two remaining axes. values[2f+0..1] = (p_w, S_w) for EWG_BC_FREE_FLOW and
 * (rate_w, rate_n) [kg/(m^2 s), negative = into the domain] for EWG_BC_MASS_RATE. */
int ewg_set_boundary(ewg_ctx *ctx, const int32_t *type, const double *values);

/* FvBaseDiscretization::solution(timeIdx) (fvbasediscretization.hh:1110-1117):
 * timeIdx 0 = current iterate, 1 = previous time level. */
int ewg_set_solution(ewg_ctx *ctx, int time_idx, const double *x_aos);
int ewg_get_solution(ewg_ctx *ctx, int time_idx, double *x_aos);
/* FvBaseDiscretization::advanceTimeLevel (fvbasediscretization.hh:1396-1407). */
int ewg_advance_time_level(ewg_ctx *ctx);

/* IstlSparseMatrixAdapter::reserve pattern (fvbaselinearizer.hh:303-340):
 * rowptr[N+1], colidx[nnzb], rows in cell order, columns ascending. */
int ewg_get_pattern(const ewg_ctx *ctx, int32_t *rowptr, int32_t *colidx);

/* FvBaseLinearizer::linearizeDomain (fvbaselinearizer.hh:166-197): residual and
 * Jacobian at solution(0) for time-step size dt. newton_iter selects the storage-cache
 * semantics of fvbaselocalresidual.hh:538-575 (iteration 0 refreshes the cached
 * storage of the previous time level). The results stay on the device; resid
 * (2N doubles) and jac (nnzb*4 doubles, BCRS order) are optional host copies that
 * reproduce FvBaseLinearizer::residual() / jacobian(). */
int ewg_linearize(ewg_ctx *ctx, double dt, int newton_iter, double *resid, double *jac);

/* ParallelBaseBackend::prepare/setResidual/setMatrix (parallelbasebackend.hh:231-312)
 * for a system built elsewhere: uploads a residual (2N doubles, AoS) and Jacobian values
 * (nnzb*4 doubles in the BCRS order of ewg_get_pattern) as the system ewg_solve solves.
 * Either may be NULL to keep the current one. */
int ewg_set_system(ewg_ctx *ctx, const double *resid, const double *jac);

/* ParallelBaseBackend::setMatrix + solve (parallelbasebackend.hh:290-338): solves
 * J dx = R for the last linearization with ILU0-preconditioned BiCGStab from x0 = 0.
 * Returns EWG_NOT_CONVERGED when the criterion is not met, EWG_NUMERICAL_ISSUE on
 * breakdown. dx (2N doubles) is an optional host copy; iterations may be NULL. */
int ewg_solve(ewg_ctx *ctx, double *dx, int32_t *iterations);

/* PreconditionerWrapperILU::prepare (ewoms/numerics/linear/istlpreconditionerwrappers.hh:177-183,
 * which builds Dune::SeqILU<.., 0> on the solver's matrix): factorizes the Jacobian of the
 * last linearization (or ewg_set_system) exactly as ewg_solve does (FP64, or FP32 factors
 * with params.precond_fp32) and returns the factors in the BCRS order of ewg_get_pattern:
 * nnzb*4 doubles, L_ij = A_ij inv(D_j) at the lower blocks, inv(D_i) at the diagonal and
 * the upper blocks U_ij = A_ij as the sweeps read them (FP32 factors widened exactly). */
int ewg_get_preconditioner(ewg_ctx *ctx, double *lu);

/* SeqILU::apply as the BiCGStab solver calls it through its preconditioner
 * (ewoms/numerics/linear/bicgstabsolver.hh:221, :276): v = (LU)^-1 d with the factors of the last factorization
 * (a solve or ewg_get_preconditioner); d and v are 2N doubles, AoS. */
int ewg_apply_preconditioner(ewg_ctx *ctx, const double *d, double *v);

/* No counterpart in the reference: ewg_apply_preconditioner for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation). d and v are DEVICE pointers (2N doubles, AoS) on the context's device; the
 * sweeps' scratch (tile tickets, tile edge buffers, the forward intermediate) is a
 * fixed 48 MiB array per device set up by ewg_create, so the call allocates nothing.
 * Grids whose tile edges exceed 16 MiB or whose intermediate exceeds 32 MiB are rejected
 * with EWG_INVALID_ARGUMENT (128x128x64 fits; ilu_schedule 2 and 3 are rejected). The
 * call runs on the context's stream and synchronizes it before returning: the caller
 * orders its own writes of d before the call. Calls on one device share the scratch and
 * take a per-device lock; work of other processes or unlocked code on that scratch
 * would race. */
int ewg_apply_preconditioner_noalloc(ewg_ctx *ctx, const double *d, double *v);

/* OverlappingOperator::apply (ewoms/numerics/linear/overlappingoperator.hh:59-63),
 * i.e. BCRSMatrix::mv: y = J x with the Jacobian of the last linearization; x and y are
 * 2N doubles, AoS. */
int ewg_apply_operator(ewg_ctx *ctx, const double *x, double *y);

/* NewtonMethod::apply (nonlinear/newtonmethod.hh:308-517) for one time step of size
 * dt on the device, starting from solution(0). Returns 0 when converged,
 * EWG_NOT_CONVERGED / EWG_NUMERICAL_ISSUE on failure (solution(0) then holds the
 * last iterate; the caller restores it as FvBaseDiscretization::updateFailed does). */
int ewg_newton(ewg_ctx *ctx, double dt, int32_t *iterations, double *final_error);

/* FvBaseProblem::timeIntegration + Simulator::run time loop
 * (fvbaseproblem.hh:507-600, utils/simulator.hh:660-831) from time t0 to t_end with
 * initial step dt0; dt is capped by params.max_time_step_size. steps_out may be NULL. */
int ewg_run(ewg_ctx *ctx, double t0, double t_end, double dt0, int32_t *steps_out);

/* The simulator clock a restart file carries (Simulator::serialize,
 * ewoms/numerics/utils/simulator.hh:947-958). */
typedef struct ewg_restart_clock {
    int32_t episode_idx;
    double episode_start_time;
    double episode_length;
    double start_time;
    double time;
    int32_t time_step_idx;
} ewg_restart_clock;

/* Restart::restartFileName_ (ewoms/numerics/io/restart.hh:70-85):
 * "<output_dir>/<sim_name>_time=<t>_rank=<rank>.ers" into buf (len bytes). */
int ewg_restart_file_name(const char *output_dir, const char *sim_name, double t, int32_t rank,
                          char *buf, int32_t len);
/* Simulator::serialize (utils/simulator.hh:923-937) of one process with VTK output off:
 * Restart::serializeBegin (restart.hh:91-106), the clock, then solution(0) of every
 * element (EcfvDiscretization::serialize, ecfvdiscretization.hh:186-188;
 * FvBaseDiscretization::serializeEntity, fvbasediscretization.hh:1445-1460). */
int ewg_write_restart(ewg_ctx *ctx, const char *path, const ewg_restart_clock *clock);
/* The inverse (Simulator::run's restart branch, utils/simulator.hh:670-681;
 * EcfvDiscretization::deserialize, ecfvdiscretization.hh:197-202): checks the cookie
 * against the context's grid, loads solution(0) and sets solution(1) = solution(0). */
int ewg_read_restart(ewg_ctx *ctx, const char *path, ewg_restart_clock *clock);

/* VtkMultiWriter::beginWrite/endWrite (ewoms/numerics/io/vtkmultiwriter.hh:180-260) with
 * the default fields of VtkMultiPhaseModule (io/vtkmultiphasemodule.hh:61-72, :311-337):
 * "<output_dir>/<sim_name>-<5-digit index>.vtu" (the index counts this context's calls)
 * with cell data pressure_w/n, density_w/n, saturation_w/n, relativePerm_w/n and
 * porosity of solution(0), and "<output_dir>/<sim_name>.pvd" listing every file written
 * with its time t. ASCII Float64, 17 significant digits. */
int ewg_write_vtk(ewg_ctx *ctx, const char *output_dir, const char *sim_name, double t);

int ewg_get_stats(const ewg_ctx *ctx, ewg_stats *stats);
int ewg_reset_stats(ewg_ctx *ctx);

/* Measurement hook (no counterpart in the reference): launches one of the hot
 * kernels `reps` times on the context's stream at the current state and returns the
 * mean duration per launch measured with HIP events on that stream. EWG_KERNEL_COPY is a
 * plain device-to-device copy of EWG_COPY_BYTES between two buffers allocated on its
 * first use (reads and writes EWG_COPY_BYTES each: the bandwidth reference). */
enum { EWG_KERNEL_LINEARIZE = 0, EWG_KERNEL_SPMV = 1, EWG_KERNEL_ILU_APPLY = 2,
       EWG_KERNEL_ILU_FACTOR = 3, EWG_KERNEL_COPY = 5 };
/* or-ed into `which` (not with EWG_KERNEL_COPY): before every launch a read-only pass over
 * EWG_COPY_BYTES fills the caches with other data, so that each launch starts with none of
 * its operands cached; only the launches are timed */
#define EWG_KERNEL_COLD 0x100
#define EWG_COPY_BYTES (1ull << 30)
int ewg_time_kernel(ewg_ctx *ctx, int which, int reps, double *mean_ms);

/* Launches timed with device events while params.kernel_events is set (no counterpart in
 * the reference): count and summed milliseconds of the sampled launches of
 * which = EWG_KERNEL_LINEARIZE, _SPMV, _ILU_APPLY or _ILU_FACTOR since the last
 * ewg_reset_stats. */
int ewg_kernel_samples(ewg_ctx *ctx, int which, int64_t *count, double *total_ms);

/* Device description for run records (no counterpart in the reference). */
int ewg_device_info(char *buf, int len);

#ifdef __cplusplus
}
#endif

#endif /* EWG_CAPI_H */
