// ewg_internal.h -- device data layout and kernel entry points of libewg.
//
// Layout in HBM (N cells, lexicographic, x fastest):
//   vectors      AoS double2 per cell [p_w, S_w] / [eq0, eq1]          16 B/cell
//   Jacobian     structured BCRS, slot-major: J[s][cell][4], s = 0..2*dim in ascending
//                column order (3D: -z,-y,-x,diag,+x,+y,+z), 2x2 blocks row-major
//                [eq][pv]; blocks of absent neighbours are zero and never read.
//                Column indices are implicit (cell + offset[s]).        224 B/cell (3D)
//   ILU(0)       L[s][cell][4] for the dim lower slots, Dinv[cell][4] (inverted
//                pivots); U is the Jacobian's upper slots (FP64) or an FP32 copy.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#include "ewg_physics.h"

namespace ewg {

// Workgroups are dealt to the 8 XCDs in turn (workgroup b to XCD b % 8), each XCD with
// its own L2. Renumbered so that XCD k takes the k-th eighth of the logical blocks in
// order, neighbouring blocks, which share cache lines, run on one XCD (with nb not a
// multiple of 8: no renumbering).
__device__ __forceinline__ int xcd_block(int b, int nb) {
    return nb % 8 == 0 ? (b % 8) * (nb / 8) + b / 8 : b;
}

struct Grid {
    int nx, ny, nz, dim;
    int N, nxny;
    int nslots, diag;   // 2*dim+1, dim
    int nlevels;        // nx+ny+nz-2 (3D) / nx+ny-1 (2D): hyperplanes x+y+z = const
    double h[3];
    double V;
    double area[3];
};

struct NewtonState;

// Geometry of the interior faces of one direction as the flux evaluates it per face: the
// axis component of (x_ex - x_in) / |x_ex - x_in|^2, and, per phase, the hydrostatic term
// added to the gradient (as the dual number the flux adds). When every grid spacing is a
// power of two every position, difference and square of the evaluation is exact, so the
// values are the same for every face of the direction and are evaluated once.
template <class S>
struct FaceGeo {
    S gscale;
    Dual<S> cterm[2];
};
struct FaceGeoSet {
    FaceGeo<double> d[6];
    FaceGeo<float> f[6];
};

struct LinArgs {
    Grid g;
    double grav[3];
    double rho[2], mu[2];
    const double *x;      // solution(0)      [2N]
    const double *x1;     // solution(1)      [2N]
    double *sto;          // storage cache of the previous time level [2N]
    const double *K;
    const double *phi;
    const int *region;    // nullptr when there is a single region
    const double *pe;     // nullptr unless per-cell Brooks-Corey entry pressure
    const MatReg *regs;
    const int *bc_type;
    const double *bc_val;
    const double *bc_aux; // per boundary face: p_c, lambda_w, lambda_n of the free-flow state
                          // (k_bc_aux; nullptr: evaluated in the flux)
    long long side_off[7];
    double dt;
    int newton_iter;
    int enable_cache;
    int recycle;
    double *R;            // [2N]
    double *J;            // [nslots][N][4]
    double *iqv;          // [N][4] intensive quantities of each cell (values)
    double *iqd;          // [N][6] their derivatives w.r.t. the cell's primary variables
    int flux_fp32;        // 1: face fluxes and their derivatives in FP32
    int fd;               // 1: finite-difference local linearizer (FP64)
    int fd_method;        // its NumericDifferenceMethod: +1 forward, 0 central, -1 backward
    const double *pref;   // its reference pressure (the pressure's weight is 10 / pref)
    const FaceGeoSet *fgeo; // per-direction face geometry (nullptr: evaluated per face)
    const NewtonState *ns;  // device Newton state (nullptr: dt and newton_iter above): the
                            // kernels return unless ns->active, and read dt, newton_iter
                            // from it
};

// The state of one Newton solve (NewtonMethod::apply, newtonmethod.hh:308-517) on the
// device: the kernels of each iteration evaluate the proceed_ rules (:866-884) and the
// error checks, and gate the iteration's launches, so that the host reads the state once
// per iteration together with the linear solver's.
struct NewtonState {
    double error, last_error, first_error;
    double upd;                // sum |dx| of the last update (its finiteness is checked)
    double dt, tol, max_error;
    int it;                    // numIterations
    int target, max_it;
    int status;                // NS_RUNNING, NS_DONE, NS_FAILED
    int active;                // this iteration linearizes
    int solve;                 // this iteration solves the linear system
    int has_update;            // an update awaits its finiteness check
    int code;                  // NS_FAILED: NS_FAIL_MAX_ERROR or NS_FAIL_UPDATE
};
enum { NS_RUNNING = 0, NS_DONE = 1, NS_FAILED = 2 };
enum { NS_FAIL_MAX_ERROR = 1, NS_FAIL_UPDATE = 2 };

// BiCGStab scalars, kept on the device so that one host read per iteration suffices
struct KrylovState {
    double rho, alpha, omega, beta;
    double rho_i, denom;
    double resid_err, init_err, last_err;
    double rel_tol, abs_tol, max_resid;
    int status;       // 0 running, 1 converged, 2 failed (criterion), 3 breakdown
    int stagnates;
    int iterations;
    int pad;
};
enum { KS_RUNNING = 0, KS_CONVERGED = 1, KS_FAILED = 2, KS_BREAKDOWN = 3, KS_IDLE = 4 };

constexpr int kRedBlocks = 1024;   // fixed grid of the reducing vector kernels
constexpr int kRedThreads = 256;

// linearization
void launch_linearize(const LinArgs &a, hipStream_t s);
void launch_storage(const LinArgs &a, const double *x, double *M, hipStream_t s);
void launch_face_geo(const LinArgs &a, FaceGeoSet *out, hipStream_t s);  // one-time face geometry
void launch_bc_aux(const LinArgs &a, double *aux, hipStream_t s);  // per-face boundary values

// ILU(0). gate (optional): a device flag; the launches return at once while it is 0
void launch_ilu_factor64(const Grid &g, const double *A, double *L, double *Dinv, int *err,
                         hipStream_t s, const int *gate = nullptr);
// FP32 preconditioner: the FP64 factors and the Jacobian's upper blocks rounded to FP32
void launch_round_factors(const Grid &g, const double *L, const double *Dinv, const double *A,
                          float *L32, float *D32, float *U32, hipStream_t s,
                          const int *gate = nullptr);
void launch_ilu_apply64(const Grid &g, const double *L, const double *Dinv, const double *A,
                        const double *d, double *v, const int *status, hipStream_t s);
void launch_ilu_apply32(const Grid &g, const float *L, const float *Dinv, const float *U,
                        const double *d, float *w, double *v, const int *status,
                        hipStream_t s);

// SpMV y = A x; optional fused partial dots (dotA = <a, y>, dotB = <b, y> with a,b
// given; nullptr to skip), partials per block into part[0..gridDim)
void launch_spmv(const Grid &g, const double *A, const double *x, double *y,
                 const double *a, double *partA, const double *b, double *partB,
                 const int *status, hipStream_t s);
// the same with the blocks read from their FP32 copy (widened to FP64 on load)
void launch_spmv32(const Grid &g, const float *A32, const double *x, double *y,
                   const double *a, double *partA, const double *b, double *partB,
                   const int *status, hipStream_t s);
// A32 = every block of A rounded to FP32 (gate optional: nothing while *gate == 0)
void launch_round_operator(const Grid &g, const double *A, float *A32, hipStream_t s,
                           const int *gate = nullptr);

// Krylov helpers (n = 2N doubles)
// gate (optional): while *gate == 0 no solve starts (ks_init leaves the state KS_IDLE,
// which every iteration kernel skips)
void launch_bicg_init(int n, const double *b, double *r, double *x, double *p, double *v,
                      double *partMax, double *partDot, hipStream_t s,
                      const int *gate = nullptr);
void launch_ks_init(KrylovState *ks, const double *partMax, const double *partDot,
                    double rel_tol, double abs_tol, double max_resid, hipStream_t s,
                    const int *gate = nullptr);
void launch_ks_beta(KrylovState *ks, const double *partRho, hipStream_t s);
// tickets (optional): two tile tickets the kernel clears for the preconditioner apply
// that follows it
void launch_p_update(int n, const KrylovState *ks, const double *r, const double *v, double *p,
                     hipStream_t s, unsigned *tickets = nullptr);
void launch_ks_alpha(KrylovState *ks, const double *partDen, hipStream_t s);
void launch_hs_update(int n, const KrylovState *ks, const double *y, const double *v, double *x,
                      double *r, double *partMax, int *anyNonzero, hipStream_t s,
                      unsigned *tickets = nullptr);
void launch_ks_crit1(KrylovState *ks, const double *partMax, const int *anyNonzero, hipStream_t s);
void launch_ks_omega(KrylovState *ks, const double *partTT, const double *partTS, hipStream_t s);
void launch_xr_update(int n, const KrylovState *ks, const double *z, const double *t, double *x,
                      double *r, const double *r0hat, double *partRho, int *anyNonzero,
                      hipStream_t s);
void launch_ks_crit2(KrylovState *ks, const int *anyNonzero, hipStream_t s);

// Newton helpers
void launch_max_abs(int n, const double *x, double *part, hipStream_t s);
void launch_reduce_max(const double *part, double *out, hipStream_t s);
void launch_newton_update(int n, const double *cur, const double *dx, double *x, double *part,
                          hipStream_t s);
void launch_reduce_sum(const double *part, double *out, hipStream_t s);
// the device Newton loop: begin of an iteration (proceed_ at the loop head), after the
// linearization's max-norm partials, after an update's sum |dx| partials
void launch_newton_begin(NewtonState *ns, hipStream_t s);
void launch_newton_linearized(NewtonState *ns, const double *partMax, hipStream_t s);
void launch_newton_updated(NewtonState *ns, const double *partSum, hipStream_t s);
// dst = src (n doubles) while *gate
void launch_copy_gated(int n, const double *src, double *dst, const int *gate, hipStream_t s);
void launch_flag_clear(int *flag, hipStream_t s);

// dst = src (n doubles, n even)
void launch_copy(size_t n, const double *src, double *dst, hipStream_t s);

}  // namespace ewg
