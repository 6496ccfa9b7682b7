/*
 * ewo_oracle.h -- CPU oracle of the ECFV two-phase Newton step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for libewg: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. Nothing in
 * the product links or calls it.
 *
 * It is a from-scratch, element-by-element restatement of the reference algorithm
 * (the reference itself cannot be compiled here: it needs dune-common/geometry/grid/
 * istl, ewoms-common and ewoms-material, none of which exist in this image).
 */
#ifndef EWO_ORACLE_H
#define EWO_ORACLE_H

#include "../include/ewg_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ewo_ctx ewo_ctx;

int ewo_create(const ewg_grid_desc *grid, const ewg_fluid_desc *fluid,
               const ewg_material_desc *material, const ewg_params *params,
               ewo_ctx **out);
void ewo_destroy(ewo_ctx *ctx);
const char *ewo_last_error(const ewo_ctx *ctx);
int ewo_set_params(ewo_ctx *ctx, const ewg_params *params);
int ewo_set_static(ewo_ctx *ctx, const double *K, const double *phi,
                   const int32_t *region, const double *pe);
int ewo_set_boundary(ewo_ctx *ctx, const int32_t *type, const double *values);
int ewo_set_solution(ewo_ctx *ctx, int time_idx, const double *x);
int ewo_get_solution(ewo_ctx *ctx, int time_idx, double *x);
int ewo_advance_time_level(ewo_ctx *ctx);
int64_t ewo_num_cells(const ewo_ctx *ctx);
int64_t ewo_num_blocks(const ewo_ctx *ctx);
int64_t ewo_num_boundary_faces(const ewo_ctx *ctx);
int ewo_get_pattern(const ewo_ctx *ctx, int32_t *rowptr, int32_t *colidx);

/* worker threads of the element loop (ThreadsPerProcess, fvbasediscretization.hh:198,
 * default 1); the linear solver stays serial as in the reference */
int ewo_set_threads(ewo_ctx *ctx, int threads);
/* residual + Jacobian at solution(0); storage cache semantics as in ewg_linearize */
int ewo_linearize(ewo_ctx *ctx, double dt, int newton_iter, double *resid, double *jac);
/* storage M(x) [2N] (value only) of an arbitrary state */
int ewo_storage(const ewo_ctx *ctx, const double *x, double *M);

/* linear algebra on BCRS values in the oracle's pattern */
int ewo_spmv(const ewo_ctx *ctx, const double *A, const double *x, double *y);
int ewo_ilu0_factor(const ewo_ctx *ctx, const double *A, double *LU);
int ewo_ilu0_apply(const ewo_ctx *ctx, const double *LU, const double *d, double *v);
/* the same apply in float: LU and d rounded to float, the result widened */
int ewo_ilu0_apply32(const ewo_ctx *ctx, const float *LU, const double *d, double *v);
int ewo_bicgstab(ewo_ctx *ctx, const double *A, const double *b, double *x,
                 int32_t *iterations);

int ewo_set_system(ewo_ctx *ctx, const double *resid, const double *jac);
int ewo_solve(ewo_ctx *ctx, double *dx, int32_t *iterations);   /* last linearization */
int ewo_newton(ewo_ctx *ctx, double dt, int32_t *iterations, double *final_error);
int ewo_run(ewo_ctx *ctx, double t0, double t_end, double dt0, int32_t *steps_out);
/* per accepted step: newton iterations, dt; returns count written (<= cap) */
int ewo_step_log(const ewo_ctx *ctx, int32_t *newton_its, double *dts, int32_t cap);
int ewo_get_stats(const ewo_ctx *ctx, ewg_stats *stats);
/* max-norm residual of every linearization of the last Newton solve */
int ewo_newton_history(const ewo_ctx *ctx, double *errs, int32_t cap);

/* pieces pinned by closed-form tests */
double ewo_log(double x);
double ewo_exp(double x);
double ewo_pow(double x, double y);
/* material law of region r (pe <= 0: region default) at absolute Sw;
 * out[0..8] = pc, dpc/dSw, krw, dkrw/dSw, krn, dkrn/dSw, Se, (unused), (unused) */
int ewo_material(const ewo_ctx *ctx, int region, double pe, double Sw, double *out);

#ifdef __cplusplus
}
#endif

#endif
