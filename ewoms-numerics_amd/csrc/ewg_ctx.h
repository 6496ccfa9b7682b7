// ewg_ctx.h -- the context behind the C ABI's ewg_ctx handle (device state of one
// realization), shared by the C ABI (ewg_capi.hip) and the ensemble solver (ewg_batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/ewg_capi.h"
#include "ewg_internal.h"
#include "ewg_wavefront.h"

struct ewg_cnn;

struct ewg_ctx {
    using Grid = ewg::Grid;
    using WfSync = ewg::WfSync;
    using KrylovState = ewg::KrylovState;
    using MatReg = ewg::MatReg;
    template <class T>
    using WlPrecond = ewg::WlPrecond<T>;
    Grid g{};
    ewg_grid_desc grid{};
    ewg_fluid_desc fluid{};
    ewg_params prm{};
    int law = 0;
    int nregions = 0;
    long long B = 0;
    long long side_off[7] = {0};
    hipStream_t stream = nullptr;
    int device = 0;  // the device current at ewg_create (every buffer lives there)
    // device state
    double *x0 = nullptr, *x1 = nullptr, *xcur = nullptr, *sto = nullptr;
    double *R = nullptr, *J = nullptr;
    double *K = nullptr, *phi = nullptr, *pe = nullptr;
    int *region = nullptr;
    MatReg *regs = nullptr;
    std::vector<MatReg> regs_host;     // the same regions on the host (output fields)
    // VTK output (ewg_write_vtk): the steps written so far, for the collection file
    std::vector<std::pair<double, std::string>> vtk_steps;
    int *bc_type = nullptr;
    double *bc_val = nullptr;
    double *bc_aux = nullptr;  // material-law values of the free-flow boundary states
    bool bc_set = false;
    int src_box[6] = {0, -1, 0, -1, 0, -1};  // cells next to injecting rate boundaries
    // intensive quantities of the linearization's first pass
    double *iqv = nullptr, *iqd = nullptr;
    // preconditioner
    double *L64 = nullptr, *D64 = nullptr;
    float *L32 = nullptr, *D32 = nullptr, *U32 = nullptr, *w32 = nullptr;
    // params.precond_fp32 = 2: a solve failed with the FP32 factors, and the context's
    // solves use FP64 factors from then on (until the next ewg_set_params)
    bool fp64_latched = false;
    // the Jacobian's blocks rounded to FP32 for BiCGStab's operator products
    // (params.operator_fp32), slot-major like J
    float *J32 = nullptr;
    // Krylov vectors
    double *kr = nullptr, *kp = nullptr, *kv = nullptr, *ky = nullptr, *kz = nullptr,
           *kx = nullptr;
    double *part0 = nullptr, *part1 = nullptr, *part2 = nullptr;
    double *dscalar = nullptr;
    int *ilu_err = nullptr;
    int *partAny = nullptr;
    // tile-wavefront synchronisation (3D grids)
    unsigned *wf_ticket = nullptr;
    unsigned long long *wf_ex = nullptr, *wf_ey = nullptr;
    int2 *wf_order = nullptr;
    WfSync wf{};
    // ILU(0) factors in the tiles' wavefront layout (schedule 0)
    WlPrecond<double> wl64{};
    WlPrecond<float> wl32{};
    double *wl_af = nullptr;  // the factorization's operands in the forward layout
    // event-timed samples of the hot launches (params.kernel_events), by EWG_KERNEL_* id
    struct Sample {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        bool armed = false, pending = false;
        int64_t count = 0;
        double total_ms = 0.0;
    } samples[4];
    double *copy_src = nullptr, *copy_dst = nullptr;  // the bandwidth-reference copy's buffers
    KrylovState *ks = nullptr;
    ewg::NewtonState *ns = nullptr;   // the device Newton loop's state
    ewg::FaceGeoSet *fgeo = nullptr;  // per-direction interior face geometry
    bool fgeo_const = false;          // fgeo holds it (every spacing a power of two)
    double *pref = nullptr;           // reference pressure of the finite-difference weights
    // pinned host mirrors
    KrylovState *ks_host = nullptr;
    ewg::NewtonState *ns_host = nullptr;
    double *scal_host = nullptr;
    int *err_host = nullptr;
    // host copies of the BCRS pattern (rows ascending columns) and slot mapping
    std::vector<int32_t> rowptr, colidx;
    std::vector<int64_t> bcrs_src;   // BCRS block -> slot * N + row
    // Newton state (newtonmethod.hh:197 starts error_ at 1e100)
    double error = 1e100, last_error = 1e100;
    int num_iterations = 0;
    int last_linear_iterations = 0;
    bool static_set = false;
    bool have_system = false;
    bool have_precond = false;   // factors of the current Jacobian exist (a solve or
                                 // ewg_get_preconditioner made them)
    ewg_stats stats{};
    // learned warm start (params.warm_start with an attached network)
    ewg_cnn *cnn = nullptr;
    bool no_recycle = false;   // storage of the previous level from solution(1)
    int64_t warm_accepted = 0, warm_rejected = 0;
    double warm_predict_s = 0.0;  // predictions whose events have been read
    // events around the last prediction on the stream; pending: not yet in warm_predict_s
    hipEvent_t warm_e0 = nullptr, warm_e1 = nullptr;
    bool warm_pending = false;
    double first_error = 0.0;  // Newton error at the first linearization of the last solve
    std::string err;
};

namespace ewg {
// the linearization arguments of a context's current state
LinArgs ctx_lin_args(ewg_ctx *c, double dt, int newton_iter);
}  // namespace ewg
