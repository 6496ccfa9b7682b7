// ewg_capi.hip -- host side of libewg: context, C ABI, BiCGStab / Newton / time-loop
// drivers. The numerical work runs in the kernels of ewg_linearize.hip and
// ewg_linalg.hip on the context's stream; the host keeps only the control flow that
// NewtonMethod::apply, BiCGStabSolver::apply and FvBaseProblem::timeIntegration define.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ewg_capi.h"
#include "../../include/ewg_cnn.h"
#include "ewg_cnn_internal.h"
#include "ewg_ctx.h"
#include "ewg_internal.h"
#include "ewg_wavefront.h"

using namespace ewg;


namespace {

int set_err(ewg_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return set_err(ctx, EWG_DEVICE_ERROR,                                         \
                           std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

// ---------------------------------------------------------------------------------
// Scratch of ewg_apply_preconditioner_noalloc: one array per device (each device that
// loads the library holds its own copy until the process ends), split into fixed regions
// so that the edge words always live in the edge region:
//   [0, 256)                 the two tile tickets of the sweeps
//   [256, 256 + 16 MiB)      both edge buffers of the tile sweeps (the sentinel between
//                            calls; the sweeps restore it word by word)
//   [.., .. + 32 MiB)        the forward sweep's intermediate vector
// Host code never passes `scratch` itself to a kernel: it is a host-side stand-in there;
// the device address comes from hipGetSymbolAddress once per device (noalloc_setup).
// ---------------------------------------------------------------------------------
constexpr size_t kScratchTickets = 256;
constexpr size_t kScratchEdges = 16ull << 20;
constexpr size_t kScratchVec = 32ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchTickets + kScratchEdges + kScratchVec];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex[kMaxDevices];   // one call at a time per device
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// the edge region as one pair of edge buffers (for the sentinel fill)
WfSync scratch_edges(unsigned char *base) {
    WfSync w{};
    w.ex = reinterpret_cast<unsigned long long *>(base + kScratchTickets);
    w.edge_words = kScratchEdges / 16;
    w.ey = w.ex + w.edge_words;
    return w;
}

// Once per device, with that device current: its scratch array's device address, the edge
// region filled with the sentinel.
int noalloc_setup(int dev, hipStream_t s) {
    if (dev < 0 || dev >= kMaxDevices) return EWG_DEVICE_ERROR;
    std::lock_guard<std::mutex> lock(g_scratch_mutex[dev]);
    if (g_scratch_ptr[dev]) return EWG_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return EWG_DEVICE_ERROR;
    unsigned char *base = static_cast<unsigned char *>(p);
    wf_init_edges(scratch_edges(base), s);
    if (hipStreamSynchronize(s) != hipSuccess) return EWG_DEVICE_ERROR;
    g_scratch_ptr[dev] = base;
    return EWG_OK;
}

template <class T>
hipError_t dalloc(T **p, size_t count) {
    return hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(count, 1) * sizeof(T));
}

void free_all(ewg_ctx *c) {
    void *ptrs[] = {c->x0, c->x1, c->xcur, c->sto, c->R, c->J, c->iqv, c->iqd, c->K, c->phi,
                    c->pe, c->region,
                    c->regs, c->bc_type, c->bc_val, c->bc_aux, c->L64, c->D64, c->L32, c->D32, c->U32,
                    c->w32, c->kr, c->kp, c->kv, c->ky, c->kz, c->kx, c->part0, c->part1,
                    c->part2, c->dscalar, c->ilu_err, c->ks, c->ns, c->pref, c->fgeo, c->partAny,
                    c->wf_ticket, c->wf_ex, c->wf_ey, c->wf_order, c->wl64.L, c->wl64.U,
                    c->wl64.D, c->wl64.w, c->wl32.L, c->wl32.U, c->wl32.D, c->wl32.w,
                    c->wl_af, c->J32};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (c->copy_src) (void)hipFree(c->copy_src);
    if (c->copy_dst) (void)hipFree(c->copy_dst);
    for (auto &sm : c->samples) {
        if (sm.e0) (void)hipEventDestroy(sm.e0);
        if (sm.e1) (void)hipEventDestroy(sm.e1);
    }
    if (c->warm_e0) (void)hipEventDestroy(c->warm_e0);
    if (c->warm_e1) (void)hipEventDestroy(c->warm_e1);
    if (c->ks_host) (void)hipHostFree(c->ks_host);
    if (c->ns_host) (void)hipHostFree(c->ns_host);
    if (c->scal_host) (void)hipHostFree(c->scal_host);
    if (c->err_host) (void)hipHostFree(c->err_host);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

LinArgs lin_args(ewg_ctx *c, double dt, int newton_iter) {
    LinArgs a{};
    a.g = c->g;
    for (int d = 0; d < 3; ++d) a.grav[d] = c->fluid.gravity[d];
    a.rho[0] = c->fluid.rho[0];
    a.rho[1] = c->fluid.rho[1];
    a.mu[0] = c->fluid.mu[0];
    a.mu[1] = c->fluid.mu[1];
    a.x = c->x0;
    a.x1 = c->x1;
    a.sto = c->sto;
    a.K = c->K;
    a.phi = c->phi;
    a.region = c->nregions > 1 ? c->region : nullptr;
    a.pe = c->pe;
    a.regs = c->regs;
    a.bc_type = c->bc_type;
    a.bc_val = c->bc_val;
    a.bc_aux = c->static_set && c->bc_set ? c->bc_aux : nullptr;
    for (int s = 0; s < 7; ++s) a.side_off[s] = c->side_off[s];
    a.dt = dt;
    a.newton_iter = newton_iter;
    a.enable_cache = c->prm.enable_storage_cache;
    a.recycle = c->prm.recycle_first_iteration_storage && !c->no_recycle;
    a.R = c->R;
    a.J = c->J;
    a.iqv = c->iqv;
    a.iqd = c->iqd;
    a.flux_fp32 = c->prm.linearize_fp32;
    a.fd = c->prm.local_linearizer;
    a.fd_method = c->prm.numeric_difference_method;
    a.pref = c->pref;
    a.fgeo = c->fgeo_const ? c->fgeo : nullptr;
    return a;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Sampled launches: sample_begin/sample_end bracket one launch with events on the
// context's stream (one sample per kind in flight); sample_collect, called after a
// stream synchronisation, adds the finished ones to the totals.
void sample_begin(ewg_ctx *c, int which) {
    if (!c->prm.kernel_events) return;
    auto &sm = c->samples[which];
    if (sm.pending) return;
    if (!sm.e0 && (hipEventCreate(&sm.e0) != hipSuccess || hipEventCreate(&sm.e1) != hipSuccess))
        return;
    if (hipEventRecord(sm.e0, c->stream) == hipSuccess) sm.armed = true;
}
void sample_end(ewg_ctx *c, int which) {
    auto &sm = c->samples[which];
    if (!sm.armed) return;
    sm.armed = false;
    if (hipEventRecord(sm.e1, c->stream) == hipSuccess) sm.pending = true;
}
// a sample whose launches were gated off on the device: not a duration of the kernel
void sample_drop(ewg_ctx *c, int which) {
    c->samples[which].armed = false;
    c->samples[which].pending = false;
}
void sample_collect(ewg_ctx *c) {
    for (auto &sm : c->samples) {
        if (!sm.pending) continue;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, sm.e0, sm.e1) == hipSuccess) {
            sm.count += 1;
            sm.total_ms += ms;
        }
        sm.pending = false;
    }
}

// ilu_schedule 0 on a 3D grid: the factors and sweeps in the tiles' wavefront layout
// (its buffers exist when the layout fits 32-bit byte offsets); otherwise the hyperplane
// launches (2D grids, ilu_schedule 1, grids beyond the layout), with the same results
bool use_wavefront_layout(const ewg_ctx *c) {
    return c->g.dim == 3 && c->prm.ilu_schedule == 0 && c->g.N <= kWfMaxCells &&
           c->wf.ticket != nullptr && c->wl64.L && (!c->prm.precond_fp32 || c->wl32.L);
}

// The factorization always runs in FP64; the FP32 preconditioner stores the factors and
// the upper blocks rounded to FP32 (one rounding of the exact-arithmetic-order FP64
// factors, the same for both schedules).
// gate (optional): device flag; the launches return at once while it is 0
// the parameter values the library accepts beyond the reference's own ranges
const char *check_params(const ewg_params *p) {
    if (p->ilu_schedule < 0 || p->ilu_schedule > 1) return "ilu_schedule must be 0 or 1";
    if (p->precond_fp32 < 0 || p->precond_fp32 > 2) return "precond_fp32 must be 0, 1 or 2";
    if (p->linearize_fp32 != 0 && p->linearize_fp32 != 1) return "linearize_fp32 must be 0 or 1";
    if (p->operator_fp32 != 0 && p->operator_fp32 != 1) return "operator_fp32 must be 0 or 1";
    if (p->local_linearizer != 0 && p->local_linearizer != 1)
        return "local_linearizer must be 0 (automatic differentiation) or 1 (finite differences)";
    if (p->numeric_difference_method < -1 || p->numeric_difference_method > 1)
        return "numeric_difference_method must be -1, 0 or +1";
    return nullptr;
}

// FvBaseDiscretization::updateBegin of the immiscible model (immisciblemodel.hh:278-293):
// the reference pressure of the finite-difference weights is the first cell's pressure
// at the start of a solve
int capture_reference_pressure(ewg_ctx *c) {
    HIP_TRY(c, hipMemcpyAsync(c->pref, c->x0, sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return EWG_OK;
}

// With params.operator_fp32 the solve's FP32 copy of the Jacobian is made here too: by the
// wavefront layout's operand packing, which reads every block anyway, or by a rounding pass.
// the current solve's preconditioner precision
bool solve_fp32(const ewg_ctx *c) { return c->prm.precond_fp32 != 0 && !c->fp64_latched; }

void precond_factor(ewg_ctx *c, const int *gate = nullptr) {
    const bool fp32 = solve_fp32(c);
    float *op32 = c->prm.operator_fp32 ? c->J32 : nullptr;
    if (use_wavefront_layout(c)) {
        // straight into the sweeps' layout (no natural-order factors)
        if (fp32)
            wl_factor32(c->g, c->wf, c->J, c->wl_af, c->wl32, c->ilu_err, c->stream, gate, op32);
        else
            wl_factor64(c->g, c->wf, c->J, c->wl_af, c->wl64, c->ilu_err, c->stream, gate, op32);
        return;
    }
    launch_ilu_factor64(c->g, c->J, c->L64, c->D64, c->ilu_err, c->stream, gate);
    if (fp32)
        launch_round_factors(c->g, c->L64, c->D64, c->J, c->L32, c->D32, c->U32, c->stream, gate);
    if (op32) launch_round_operator(c->g, c->J, op32, c->stream, gate);
}

// BiCGStab's operator product y = J x (fused dots as launch_spmv): the FP64 Jacobian, or
// its FP32 copy of the current solve with params.operator_fp32
void operator_apply(ewg_ctx *c, const double *x, double *y, const double *a, double *partA,
                    const double *b, double *partB, const int *status) {
    if (c->prm.operator_fp32)
        launch_spmv32(c->g, c->J32, x, y, a, partA, b, partB, status, c->stream);
    else
        launch_spmv(c->g, c->J, x, y, a, partA, b, partB, status, c->stream);
}

// tickets_clear: a kernel queued on the stream since the last apply has cleared the
// layout sweeps' tickets
void precond_apply(ewg_ctx *c, const double *d, double *v, const int *status,
                   bool tickets_clear) {
    if (use_wavefront_layout(c)) {
        if (solve_fp32(c))
            wl_apply32(c->g, c->wf, c->wl32, d, v, status, c->stream, tickets_clear);
        else
            wl_apply64(c->g, c->wf, c->wl64, d, v, status, c->stream, tickets_clear);
        return;
    }
    if (solve_fp32(c))
        launch_ilu_apply32(c->g, c->L32, c->D32, c->U32, d, c->w32, v, status, c->stream);
    else
        launch_ilu_apply64(c->g, c->L64, c->D64, c->J, d, v, status, c->stream);
}

// One BiCGStab iteration (bicgstabsolver.hh:191-322) of J x = b on the device: every
// kernel returns at once when the device-side status has left KS_RUNNING. tk: the layout
// sweeps' tickets, cleared by the vector update before each apply; sampled: the
// iteration's first apply and SpMV are event-timed samples.
void bicg_iteration(ewg_ctx *c, const double *b, double *x, unsigned *tk, bool sampled) {
    const int n = 2 * c->g.N;
    hipStream_t s = c->stream;
    const int *st = &c->ks->status;
    launch_ks_beta(c->ks, c->part2, s);
    launch_p_update(n, c->ks, c->kr, c->kv, c->kp, s, tk);
    if (sampled) sample_begin(c, EWG_KERNEL_ILU_APPLY);
    precond_apply(c, c->kp, c->ky, st, tk != nullptr);
    if (sampled) {
        sample_end(c, EWG_KERNEL_ILU_APPLY);
        sample_begin(c, EWG_KERNEL_SPMV);
    }
    operator_apply(c, c->ky, c->kv, b, c->part0, nullptr, nullptr, st);
    if (sampled) sample_end(c, EWG_KERNEL_SPMV);
    launch_ks_alpha(c->ks, c->part0, s);
    launch_hs_update(n, c->ks, c->ky, c->kv, x, c->kr, c->part1, c->partAny, s, tk);
    launch_ks_crit1(c->ks, c->part1, c->partAny, s);
    precond_apply(c, c->kr, c->kz, st, tk != nullptr);
    operator_apply(c, c->kz, c->ky, c->ky, c->part0, c->kr, c->part1, st);
    launch_ks_omega(c->ks, c->part0, c->part1, s);
    launch_xr_update(n, c->ks, c->kz, c->ky, x, c->kr, b, c->part2, c->partAny, s);
    launch_ks_crit2(c->ks, c->partAny, s);
}

unsigned *layout_tickets(ewg_ctx *c) { return use_wavefront_layout(c) ? c->wf.ticket : nullptr; }

// BiCGStabSolver::apply on J dx = b (b = residual on the device), first part: queues the
// factorization, the initial state (x0 = 0, r = b), the first iteration and the copies the
// host reads after its next wait (err_host[0]: factorization error, err_host[1]: initial
// status, ks_host and err_host[2]: state and sweep error after the first iteration).
// gate (optional): device flag; nothing of it runs while 0 (the state stays KS_IDLE).
void bicgstab_begin(ewg_ctx *c, const double *b, double *x, const int *gate) {
    const int n = 2 * c->g.N;
    hipStream_t s = c->stream;
    double absTol = c->prm.linear_abs_tolerance;
    if (absTol < 0.0) absTol = c->prm.newton_tolerance / 100.0;
    launch_flag_clear(c->ilu_err, s);
    sample_begin(c, EWG_KERNEL_ILU_FACTOR);
    precond_factor(c, gate);
    sample_end(c, EWG_KERNEL_ILU_FACTOR);
    launch_bicg_init(n, b, c->kr, x, c->kp, c->kv, c->part0, c->part1, s, gate);
    launch_ks_init(c->ks, c->part0, c->part1, c->prm.linear_tolerance, absTol,
                   c->prm.linear_max_error, s, gate);
    (void)hipMemcpyAsync(c->err_host, c->ilu_err, sizeof(int), hipMemcpyDeviceToHost, s);
    (void)hipMemcpyAsync(c->err_host + 1, &c->ks->status, sizeof(int), hipMemcpyDeviceToHost, s);
    if (c->prm.linear_max_iterations > 0) bicg_iteration(c, b, x, layout_tickets(c), false);
    (void)hipMemcpyAsync(c->ks_host, c->ks, sizeof(KrylovState), hipMemcpyDeviceToHost, s);
    (void)hipMemcpyAsync(c->err_host + 2, c->ilu_err, sizeof(int), hipMemcpyDeviceToHost, s);
}

// Second part, after the host has waited for the first: the remaining iterations are
// queued speculatively and the state read back after 1, 2, 4, then every 8 of them (the
// iterations queued past convergence are no-ops, so the count kept on the device is the
// one of a solver that stops exactly; most solves of a time step after the first end
// within their second iteration, so the first batch is a single one). Returns EWG_OK, EWG_NOT_CONVERGED or
// EWG_NUMERICAL_ISSUE; *iters as BiCGStab's report.
int bicgstab_end(ewg_ctx *c, const double *b, double *x, int *iters) {
    hipStream_t s = c->stream;
    sample_collect(c);
    c->have_precond = true;
    *iters = 0;
    if (c->err_host[0]) {
        // the iterations after the failed factorization ran for nothing; a wavefront wait
        // that ran out may have left words in the edge buffers
        if (c->wf.ex) wf_init_edges(c->wf, s);
        HIP_TRY(c, hipStreamSynchronize(s));
        return set_err(c, EWG_NUMERICAL_ISSUE, "Creating the preconditioner failed");
    }
    if (c->err_host[1] == KS_CONVERGED) return EWG_OK;
    if (c->prm.linear_max_iterations <= 0)
        return set_err(c, EWG_NOT_CONVERGED, "linear solver did not converge");
    unsigned *tk = layout_tickets(c);
    int it = 1, batch = 1;
    while (true) {
        if (c->err_host[2]) {
            // a tile-wavefront wait ran out: its edge buffers may hold unconsumed words
            if (c->wf.ex) wf_init_edges(c->wf, s);
            HIP_TRY(c, hipStreamSynchronize(s));
            *iters = c->ks_host->iterations;
            return set_err(c, EWG_NUMERICAL_ISSUE, "ILU(0) sweep failed");
        }
        if (c->ks_host->status != KS_RUNNING || it >= c->prm.linear_max_iterations) break;
        const int nb = std::min(batch, c->prm.linear_max_iterations - it);
        for (int q = 0; q < nb; ++q) bicg_iteration(c, b, x, tk, q == 0);
        it += nb;
        HIP_TRY(c, hipMemcpyAsync(c->ks_host, c->ks, sizeof(KrylovState), hipMemcpyDeviceToHost,
                                  s));
        HIP_TRY(c, hipMemcpyAsync(c->err_host + 2, c->ilu_err, sizeof(int), hipMemcpyDeviceToHost,
                                  s));
        HIP_TRY(c, hipStreamSynchronize(s));
        sample_collect(c);
        batch = std::min(batch * 2, 8);
    }
    *iters = c->ks_host->iterations;
    switch (c->ks_host->status) {
        case KS_CONVERGED: return EWG_OK;
        case KS_BREAKDOWN: return set_err(c, EWG_NUMERICAL_ISSUE, "Breakdown of the BiCGStab solver");
        default: return set_err(c, EWG_NOT_CONVERGED, "linear solver did not converge");
    }
}

// BiCGStabSolver::apply on J dx = b for ewg_solve: both parts with the one wait between.
int bicgstab(ewg_ctx *c, const double *b, double *x, int *iters) {
    bicgstab_begin(c, b, x, nullptr);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return bicgstab_end(c, b, x, iters);
}

// precond_fp32 = 2: a solve that failed with the FP32 factors (rc > 0: not converged or a
// numerical issue) is redone from x0 = 0 with FP64 factors, which the context keeps for its
// later solves; the redone solve's iterations are added to *iters. The FP32 factors'
// rounding can stall BiCGStab on strongly heterogeneous permeability fields where the FP64
// ones converge.
int fp64_retry(ewg_ctx *c, const double *b, double *x, int rc, int *iters) {
    if (rc <= 0 || c->prm.precond_fp32 != 2 || c->fp64_latched) return rc;
    c->fp64_latched = true;
    c->stats.precond_retries++;
    bicgstab_begin(c, b, x, nullptr);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    int its = 0;
    rc = bicgstab_end(c, b, x, &its);
    *iters += its;
    return rc;
}

int linearize_device(ewg_ctx *c, double dt, int newton_iter) {
    LinArgs a = lin_args(c, dt, newton_iter);
    sample_begin(c, EWG_KERNEL_LINEARIZE);
    launch_linearize(a, c->stream);
    sample_end(c, EWG_KERNEL_LINEARIZE);
    HIP_TRY(c, hipGetLastError());
    c->have_system = true;
    return EWG_OK;
}

// FP32 preconditioner storage, allocated on first use
int ensure_fp32(ewg_ctx *c) {
    if (c->L32) return EWG_OK;
    size_t N = c->g.N;
    HIP_TRY(c, dalloc(&c->L32, (size_t)c->g.dim * N * 4));
    HIP_TRY(c, dalloc(&c->D32, N * 4));
    HIP_TRY(c, dalloc(&c->U32, (size_t)c->g.dim * N * 4));
    HIP_TRY(c, dalloc(&c->w32, 2 * N));
    if (c->wl64.L) {
        const size_t E = wl_entries(c->g, c->wf);
        HIP_TRY(c, dalloc(&c->wl32.L, 3 * E * 4));
        HIP_TRY(c, dalloc(&c->wl32.U, 3 * E * 4));
        HIP_TRY(c, dalloc(&c->wl32.D, E * 4));
        HIP_TRY(c, dalloc(&c->wl32.w, E * 2));
        // the layout's entries without a cell stay zero (the factorization writes cells)
        HIP_TRY(c, hipMemset(c->wl32.L, 0, 3 * E * 4 * sizeof(float)));
        HIP_TRY(c, hipMemset(c->wl32.U, 0, 3 * E * 4 * sizeof(float)));
        HIP_TRY(c, hipMemset(c->wl32.D, 0, E * 4 * sizeof(float)));
        HIP_TRY(c, hipMemset(c->wl32.w, 0, E * 2 * sizeof(float)));
    }
    return EWG_OK;
}

// FP32 copy of the Jacobian for BiCGStab's operator, allocated on first use
int ensure_op32(ewg_ctx *c) {
    if (c->J32) return EWG_OK;
    HIP_TRY(c, dalloc(&c->J32, (size_t)c->g.nslots * c->g.N * 4));
    return EWG_OK;
}

// NewtonMethod::apply (newtonmethod.hh:308-517) with the loop's decisions on the device
// (k_newton_begin / k_newton_linearized / k_newton_updated): per iteration the host queues
// the loop head, the linearization, its max-norm and proceed_, gated by the device state,
// and reads the state back. In the first iteration of a solve, which nearly always goes
// on to a linear solve, the start of that solve (factorization, initial state, first
// BiCGStab iteration) is queued behind it, gated by the device's decision, so that one
// wait covers both the Newton and the Krylov state. Later iterations converge more often
// than not; there the solve is queued only once the state says so (a converged
// iteration then costs no launches of a solve that does not happen).
int newton(ewg_ctx *c, double dt) {
    hipStream_t s = c->stream;
    const int n2 = 2 * c->g.N;
    NewtonState &h = *c->ns_host;
    h = NewtonState{};
    h.error = 1e100;  // newtonmethod.hh:197
    h.last_error = 1e100;
    h.dt = dt;
    h.tol = c->prm.newton_tolerance;
    h.max_error = c->prm.newton_max_error;
    h.target = c->prm.newton_target_iterations;
    h.max_it = c->prm.newton_max_iterations;
    h.status = NS_RUNNING;
    HIP_TRY(c, hipMemcpyAsync(c->ns, c->ns_host, sizeof(NewtonState), hipMemcpyHostToDevice, s));
    if (int rc = capture_reference_pressure(c)) return rc;
    const int *active = &c->ns->active, *solve = &c->ns->solve;
    auto mirror = [&]() {
        c->error = h.error;
        c->last_error = h.last_error;
        c->first_error = h.first_error;
        c->num_iterations = h.it;
    };
    for (int round = 0;; ++round) {
        const bool merged = round == 0;
        const double t0 = now_s();
        launch_newton_begin(c->ns, s);
        launch_copy_gated(n2, c->x0, c->xcur, active, s);
        LinArgs a = lin_args(c, dt, 0);
        a.ns = c->ns;
        sample_begin(c, EWG_KERNEL_LINEARIZE);
        launch_linearize(a, s);
        sample_end(c, EWG_KERNEL_LINEARIZE);
        launch_max_abs(n2, c->R, c->part0, s);
        launch_newton_linearized(c->ns, c->part0, s);
        if (merged) bicgstab_begin(c, c->R, c->kx, solve);
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(c->ns_host, c->ns, sizeof(NewtonState), hipMemcpyDeviceToHost, s));
        HIP_TRY(c, hipStreamSynchronize(s));
        if (!h.active) {
            // the loop ended at its head: nothing of this round ran
            sample_drop(c, EWG_KERNEL_LINEARIZE);
            sample_drop(c, EWG_KERNEL_ILU_FACTOR);
            break;
        }
        c->have_system = true;
        c->stats.newton_iterations++;
        const double t1 = now_s();
        c->stats.time_linearize_s += t1 - t0;
        if (!h.solve) {  // converged, or failed, at this linearization
            sample_drop(c, EWG_KERNEL_ILU_FACTOR);
            sample_collect(c);
            break;
        }
        if (!merged) {
            bicgstab_begin(c, c->R, c->kx, nullptr);
            HIP_TRY(c, hipGetLastError());
            HIP_TRY(c, hipStreamSynchronize(s));
        }
        int its = 0;
        const int rc = fp64_retry(c, c->R, c->kx, bicgstab_end(c, c->R, c->kx, &its), &its);
        const double t2 = now_s();
        c->stats.time_solve_s += t2 - t1;
        c->last_linear_iterations = its;
        c->stats.linear_iterations += its;
        c->stats.linear_solves++;
        if (rc != EWG_OK) {
            mirror();
            c->num_iterations = 2 * c->prm.newton_target_iterations;
            return rc;
        }
        launch_newton_update(n2, c->xcur, c->kx, c->x0, c->part0, s);
        launch_newton_updated(c->ns, c->part0, s);
        c->stats.time_update_s += now_s() - t2;
    }
    mirror();
    if (h.status == NS_FAILED)
        return set_err(c, EWG_NUMERICAL_ISSUE, h.code == NS_FAIL_MAX_ERROR
                                                   ? "Newton: error above NewtonMaxError"
                                                   : "Non-finite update!");
    if (!(c->error <= c->prm.newton_tolerance)) {
        c->num_iterations = 2 * c->prm.newton_target_iterations;
        return set_err(c, EWG_NOT_CONVERGED, "Newton did not converge");
    }
    return EWG_OK;
}

// One Newton solve of a time step. With a warm start, the solve starts from
// solution(1) + the network's prediction; if it fails to converge from there, it is redone
// from the plain guess solution(1).
// Duration of the last prediction (device clock between the events around it), 0 when
// none is pending. Its events have completed once a Newton solve after it has returned.
double warm_pending_s(const ewg_ctx *c) {
    if (!c->warm_pending) return 0.0;
    float ms = 0.0f;
    if (hipEventSynchronize(c->warm_e1) != hipSuccess ||
        hipEventElapsedTime(&ms, c->warm_e0, c->warm_e1) != hipSuccess)
        return 0.0;
    return 1e-3 * ms;
}

int newton_step(ewg_ctx *c, double dt) {
    if (!(c->cnn && c->prm.warm_start)) return newton(c, dt);
    // the prediction is timed by events on the stream, so that the solve's launches
    // queue behind it without a wait (the host-timed first linearization then also
    // covers the end of the prediction)
    c->warm_predict_s += warm_pending_s(c);
    c->warm_pending = false;
    if (!c->warm_e0) {
        HIP_TRY(c, hipEventCreate(&c->warm_e0));
        HIP_TRY(c, hipEventCreate(&c->warm_e1));
    }
    HIP_TRY(c, hipEventRecord(c->warm_e0, c->stream));
    int rc = cnn::predict_guess(c->cnn, c->x1, c->K, c->phi, dt, c->x0, c->stream, c->src_box);
    if (rc != EWG_OK) return set_err(c, rc, "warm start: prediction failed");
    HIP_TRY(c, hipEventRecord(c->warm_e1, c->stream));
    c->warm_pending = true;
    c->no_recycle = true;
    rc = newton(c, dt);
    c->no_recycle = false;
    if (rc == EWG_OK) {
        c->warm_accepted++;
        return rc;
    }
    if (rc < 0) return rc;
    c->warm_rejected++;
    HIP_TRY(c, hipMemcpyAsync(c->x0, c->x1, 2 * (size_t)c->g.N * sizeof(double),
                              hipMemcpyDeviceToDevice, c->stream));
    return newton(c, dt);
}

// Entry of cell (x, y, z) in the wavefront layout of the 8x8 tiles (the inverse of the
// mapping the packing and factorization kernels write: lane lx + 8 ly of tile
// tx + ntx ty reaches z at local step z + lx + ly). The backward layout holds cell
// (x, y, z) at the forward entry of the mirrored cell.
size_t wl_entry(const Grid &g, const WfSync &w, int x, int y, int z) {
    const int lx = x & 7, ly = y & 7;
    const size_t tile = (size_t)(x >> 3) + (size_t)w.ntx * (y >> 3);
    return (tile * (size_t)wl_nsteps(g.nz) + (size_t)(z + lx + ly)) * 64 + (size_t)(lx + 8 * ly);
}

// The factors the sweeps read, widened to double, as natural-order slot arrays:
// L[dim][N][4] (lower slots in ascending column order), Dinv[N][4], U[dim][N][4] (upper
// slots in ascending column order).
template <class T>
int gather_layout(ewg_ctx *c, const WlPrecond<T> &wl, std::vector<double> &L,
                  std::vector<double> &D, std::vector<double> &U) {
    const Grid &g = c->g;
    const size_t E = wl_entries(g, c->wf), N = g.N;
    std::vector<T> hL(3 * E * 4), hU(3 * E * 4), hD(E * 4);
    HIP_TRY(c, hipMemcpy(hL.data(), wl.L, hL.size() * sizeof(T), hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(hU.data(), wl.U, hU.size() * sizeof(T), hipMemcpyDeviceToHost));
    HIP_TRY(c, hipMemcpy(hD.data(), wl.D, hD.size() * sizeof(T), hipMemcpyDeviceToHost));
    for (int z = 0; z < g.nz; ++z)
        for (int y = 0; y < g.ny; ++y)
            for (int x = 0; x < g.nx; ++x) {
                const size_t i = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z);
                const size_t pf = wl_entry(g, c->wf, x, y, z);
                const size_t pb = wl_entry(g, c->wf, g.nx - 1 - x, g.ny - 1 - y, g.nz - 1 - z);
                for (int e = 0; e < 4; ++e) {
                    for (int s = 0; s < 3; ++s) {
                        L[(s * N + i) * 4 + e] = (double)hL[(s * E + pf) * 4 + e];
                        U[(s * N + i) * 4 + e] = (double)hU[(s * E + pb) * 4 + e];
                    }
                    D[i * 4 + e] = (double)hD[pb * 4 + e];
                }
            }
    return EWG_OK;
}

template <class T>
int gather_natural(ewg_ctx *c, const T *dL, const T *dD, std::vector<double> &v, size_t off,
                   size_t count) {
    std::vector<T> h(count);
    HIP_TRY(c, hipMemcpy(h.data(), dL ? dL : dD, count * sizeof(T), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < count; ++k) v[off + k] = (double)h[k];
    return EWG_OK;
}

double suggest_dt(const ewg_ctx *c, double oldDt) {  // newtonmethod.hh:527-544
    int tgt = c->prm.newton_target_iterations;
    if (c->num_iterations > tgt) {
        double percent = static_cast<double>(c->num_iterations - tgt) / tgt;
        return std::max(c->prm.min_time_step_size, oldDt / (1.0 + percent));
    }
    double percent = static_cast<double>(tgt - c->num_iterations) / tgt;
    return std::max(c->prm.min_time_step_size, oldDt * (1.0 + percent / 1.2));
}

}  // namespace

namespace ewg {
LinArgs ctx_lin_args(ewg_ctx *c, double dt, int newton_iter) {
    return lin_args(c, dt, newton_iter);
}
}  // namespace ewg

extern "C" {

void ewg_default_params(ewg_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->newton_tolerance = 1e-8;
    p->newton_max_error = 1e100;
    p->newton_target_iterations = 10;
    p->newton_max_iterations = 18;
    p->linear_tolerance = 1e-3;
    p->linear_abs_tolerance = -1.0;
    p->linear_max_error = 1e7;
    p->linear_max_iterations = 1000;
    p->enable_storage_cache = 1;
    p->recycle_first_iteration_storage = 1;
    p->max_time_step_divisions = 10;
    p->min_time_step_size = 0.0;
    p->max_time_step_size = std::numeric_limits<double>::infinity();
    p->precond_fp32 = 0;
    p->numeric_difference_method = 1;
}

int ewg_create(const ewg_grid_desc *grid, const ewg_fluid_desc *fluid,
               const ewg_material_desc *material, const ewg_params *params, ewg_ctx **out) {
    if (!grid || !fluid || !material || !params || !out) return EWG_INVALID_ARGUMENT;
    *out = nullptr;
    if (grid->dim != 2 && grid->dim != 3) return EWG_INVALID_ARGUMENT;
    if (grid->dim == 2 && grid->nz != 1) return EWG_INVALID_ARGUMENT;
    // the structured ILU(0) relies on nx >= 3 (and ny >= 3 in 3D): then the only fill
    // position two pattern offsets can reach is the diagonal
    if (grid->nx < 3 || grid->ny < (grid->dim == 3 ? 3 : 1) || grid->nz < 1)
        return EWG_INVALID_ARGUMENT;
    long long N = (long long)grid->nx * grid->ny * grid->nz;
    if (N >= (1LL << 31) / 8) return EWG_INVALID_ARGUMENT;
    if (material->nregions < 1 || !material->region_params) return EWG_INVALID_ARGUMENT;
    if (material->law != EWG_LAW_VAN_GENUCHTEN && material->law != EWG_LAW_BROOKS_COREY)
        return EWG_INVALID_ARGUMENT;
    if (check_params(params)) return EWG_INVALID_ARGUMENT;
    auto *c = new ewg_ctx();
    c->grid = *grid;
    c->fluid = *fluid;
    c->prm = *params;
    c->law = material->law;
    c->nregions = material->nregions;
    Grid &g = c->g;
    g.dim = grid->dim;
    g.nx = grid->nx;
    g.ny = grid->ny;
    g.nz = grid->nz;
    g.N = static_cast<int>(N);
    g.nxny = g.nx * g.ny;
    g.nslots = 2 * g.dim + 1;
    g.diag = g.dim;
    g.nlevels = g.dim == 3 ? g.nx + g.ny + g.nz - 2 : g.nx + g.ny - 1;
    g.h[0] = grid->lx / g.nx;
    g.h[1] = grid->ly / g.ny;
    g.h[2] = g.dim == 3 ? grid->lz / g.nz : 1.0;
    if (g.dim == 2) {
        g.V = g.h[0] * g.h[1];
        g.area[0] = g.h[1];
        g.area[1] = g.h[0];
        g.area[2] = 0.0;
    } else {
        g.V = g.h[0] * g.h[1] * g.h[2];
        g.area[0] = g.h[1] * g.h[2];
        g.area[1] = g.h[0] * g.h[2];
        g.area[2] = g.h[0] * g.h[1];
    }
    long long side[6];
    side[0] = side[1] = (long long)g.ny * g.nz;
    side[2] = side[3] = (long long)g.nx * g.nz;
    side[4] = side[5] = g.dim == 3 ? (long long)g.nx * g.ny : 0;
    c->side_off[0] = 0;
    for (int s = 0; s < 6; ++s) c->side_off[s + 1] = c->side_off[s] + side[s];
    c->B = c->side_off[6];

    std::vector<MatReg> regs;
    for (int r = 0; r < material->nregions; ++r)
        regs.push_back(make_region(material->law, material->region_params + 4 * r));

    auto fail = [&](const char *what, hipError_t e) {
        std::string msg = std::string(what) + ": " + hipGetErrorString(e);
        free_all(c);
        delete c;
        std::fprintf(stderr, "ewg_create: %s\n", msg.c_str());
        return EWG_DEVICE_ERROR;
    };
    hipError_t e;
    if ((e = hipGetDevice(&c->device)) != hipSuccess) return fail("device", e);
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail("stream", e);
    if (noalloc_setup(c->device, c->stream) != EWG_OK)
        return fail("no-allocation scratch", hipErrorInvalidSymbol);
    const size_t n2 = 2 * (size_t)N;
    const size_t nb = (size_t)g.nslots * N * 4;
#define ALLOC(p, cnt) \
    if ((e = dalloc(&(p), (cnt))) != hipSuccess) return fail(#p, e)
    ALLOC(c->x0, n2);
    ALLOC(c->x1, n2);
    ALLOC(c->xcur, n2);
    ALLOC(c->sto, n2);
    ALLOC(c->R, n2);
    ALLOC(c->J, nb);
    ALLOC(c->iqv, 4 * (size_t)N);
    ALLOC(c->iqd, 6 * (size_t)N);
    ALLOC(c->K, N);
    ALLOC(c->phi, N);
    ALLOC(c->region, N);
    ALLOC(c->regs, regs.size());
    ALLOC(c->bc_type, c->B);
    ALLOC(c->bc_val, 2 * c->B);
    ALLOC(c->bc_aux, 3 * c->B);
    ALLOC(c->L64, (size_t)g.dim * N * 4);
    ALLOC(c->D64, (size_t)N * 4);
    ALLOC(c->kr, n2);
    ALLOC(c->kp, n2);
    ALLOC(c->kv, n2);
    ALLOC(c->ky, n2);
    ALLOC(c->kz, n2);
    ALLOC(c->kx, n2);
    ALLOC(c->part0, kRedBlocks);
    ALLOC(c->part1, kRedBlocks);
    ALLOC(c->part2, kRedBlocks);
    ALLOC(c->dscalar, 8);
    ALLOC(c->ilu_err, 1);
    ALLOC(c->ks, 1);
    ALLOC(c->ns, 1);
    ALLOC(c->pref, 1);
    ALLOC(c->fgeo, 1);
    ALLOC(c->partAny, kRedBlocks);
    // the tile-wavefront buffers (3D; every schedule but the hyperplane launches)
    if (g.dim == 3 && N <= kWfMaxCells && params->ilu_schedule != 1) {
        auto setup = [&](WfSync &w, int T, unsigned *&ticket, int2 *&dorder,
                         unsigned long long *&ex, unsigned long long *&ey) -> int {
            w.ntx = (g.nx + T - 1) / T;
            w.nty = (g.ny + T - 1) / T;
            w.ntiles = w.ntx * w.nty;
            ALLOC(ticket, 2);
            ALLOC(dorder, w.ntiles);
            w.edge_words = (size_t)w.ntiles * T * g.nz * kWfMaxWords;
            ALLOC(ex, w.edge_words);
            ALLOC(ey, w.edge_words);
            std::vector<int2> order;
            for (int d = 0; d <= w.ntx + w.nty - 2; ++d)
                for (int tx = 0; tx < w.ntx; ++tx) {
                    int ty = d - tx;
                    if (ty >= 0 && ty < w.nty) order.push_back(make_int2(tx, ty));
                }
            if ((e = hipMemcpy(dorder, order.data(), order.size() * sizeof(int2),
                               hipMemcpyHostToDevice)) != hipSuccess)
                return fail("wf_order", e);
            w.ticket = ticket;
            w.order = dorder;
            w.ex = ex;
            w.ey = ey;
            w.err = c->ilu_err;
            wf_init_edges(w, c->stream);
            if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return fail("wf_init", e);
            return EWG_OK;
        };
        int rc = setup(c->wf, kWfTX, c->wf_ticket, c->wf_order, c->wf_ex, c->wf_ey);
        if (rc != EWG_OK) return rc;
        // the wavefront layout must fit 32-bit byte offsets (32 bytes per FP64 block)
        const size_t E = wl_entries(g, c->wf);
        if (E * 32 <= 0x7FFFFF00ull) {
            ALLOC(c->wl64.L, 3 * E * 4);
            ALLOC(c->wl64.U, 3 * E * 4);
            ALLOC(c->wl64.D, E * 4);
            ALLOC(c->wl64.w, E * 2);
            ALLOC(c->wl_af, 7 * E * 4);
            // the layout's entries without a cell stay zero (the factorization writes cells)
            if ((e = hipMemset(c->wl64.L, 0, 3 * E * 4 * sizeof(double))) != hipSuccess ||
                (e = hipMemset(c->wl64.U, 0, 3 * E * 4 * sizeof(double))) != hipSuccess ||
                (e = hipMemset(c->wl64.D, 0, E * 4 * sizeof(double))) != hipSuccess ||
                (e = hipMemset(c->wl64.w, 0, E * 2 * sizeof(double))) != hipSuccess ||
                (e = hipMemset(c->wl_af, 0, 7 * E * 4 * sizeof(double))) != hipSuccess)
                return fail("wl64", e);
        }
    }
#undef ALLOC
    if ((e = hipHostMalloc(reinterpret_cast<void **>(&c->ks_host), sizeof(KrylovState))) != hipSuccess)
        return fail("ks_host", e);
    if ((e = hipHostMalloc(reinterpret_cast<void **>(&c->ns_host), sizeof(NewtonState))) != hipSuccess)
        return fail("ns_host", e);
    if ((e = hipHostMalloc(reinterpret_cast<void **>(&c->scal_host), 8 * sizeof(double))) != hipSuccess)
        return fail("scal_host", e);
    if ((e = hipHostMalloc(reinterpret_cast<void **>(&c->err_host), 4 * sizeof(int))) != hipSuccess)
        return fail("err_host", e);
    // blocks of absent neighbours stay zero for the whole life of the context
    if ((e = hipMemset(c->J, 0, nb * sizeof(double))) != hipSuccess) return fail("J", e);
    if ((e = hipMemset(c->x0, 0, n2 * sizeof(double))) != hipSuccess) return fail("x0", e);
    if ((e = hipMemset(c->x1, 0, n2 * sizeof(double))) != hipSuccess) return fail("x1", e);
    if ((e = hipMemset(c->sto, 0, n2 * sizeof(double))) != hipSuccess) return fail("sto", e);
    if ((e = hipMemset(c->bc_type, 0, c->B * sizeof(int))) != hipSuccess) return fail("bc", e);
    if ((e = hipMemset(c->bc_val, 0, 2 * c->B * sizeof(double))) != hipSuccess) return fail("bcv", e);
    if ((e = hipMemset(c->bc_aux, 0, 3 * c->B * sizeof(double))) != hipSuccess) return fail("bca", e);
    c->regs_host = regs;
    if ((e = hipMemcpy(c->regs, regs.data(), regs.size() * sizeof(MatReg), hipMemcpyHostToDevice)) !=
        hipSuccess)
        return fail("regs", e);

    // BCRS pattern on the host (IstlSparseMatrixAdapter::reserve order)
    c->rowptr.assign(N + 1, 0);
    c->colidx.reserve((size_t)N * g.nslots);
    c->bcrs_src.reserve((size_t)N * g.nslots);
    const int nn[3] = {g.nx, g.ny, g.nz};
    for (int i = 0; i < g.N; ++i) {
        int ijk[3] = {i % g.nx, (i / g.nx) % g.ny, i / g.nxny};
        for (int s = 0; s < g.nslots; ++s) {
            int col;
            if (s == g.diag) {
                col = i;
            } else if (s < g.diag) {
                int ax = g.dim - 1 - s;
                if (ijk[ax] == 0) continue;
                col = i - (ax == 0 ? 1 : (ax == 1 ? g.nx : g.nxny));
            } else {
                int ax = s - g.diag - 1;
                if (ijk[ax] + 1 >= nn[ax]) continue;
                col = i + (ax == 0 ? 1 : (ax == 1 ? g.nx : g.nxny));
            }
            c->colidx.push_back(col);
            c->bcrs_src.push_back((int64_t)s * N + i);
        }
        c->rowptr[i + 1] = static_cast<int32_t>(c->colidx.size());
    }
    if ((c->prm.precond_fp32 && ensure_fp32(c) != EWG_OK) ||
        (c->prm.operator_fp32 && ensure_op32(c) != EWG_OK)) {
        std::fprintf(stderr, "ewg_create: %s\n", c->err.c_str());
        free_all(c);
        delete c;
        return EWG_DEVICE_ERROR;
    }
    // the interior faces' geometry once per direction when every spacing is a power of two
    // (every position, difference and square the flux evaluates is then exact, the same for
    // every face of a direction, and the per-face evaluation gives these bits)
    {
        bool pow2 = true;
        for (int d = 0; d < g.dim; ++d) {
            int ex = 0;
            const double m = std::frexp(g.h[d], &ex);
            const double n = d == 0 ? g.nx : (d == 1 ? g.ny : g.nz);
            pow2 = pow2 && m == 0.5 && (n + 1.0) * g.h[d] < 0x1p40;
        }
        if (pow2) {
            launch_face_geo(lin_args(c, 1.0, 0), c->fgeo, c->stream);
            if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return fail("face geometry", e);
            c->fgeo_const = true;
        }
    }
    *out = c;
    return EWG_OK;
}

void ewg_destroy(ewg_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
}

const char *ewg_last_error(const ewg_ctx *c) { return c ? c->err.c_str() : "null context"; }

int ewg_set_params(ewg_ctx *c, const ewg_params *p) {
    if (!c || !p) return EWG_INVALID_ARGUMENT;
    if (const char *msg = check_params(p)) return set_err(c, EWG_INVALID_ARGUMENT, msg);
    if (p->ilu_schedule != 1 && c->g.dim == 3 && c->g.N <= kWfMaxCells && !c->wf.ticket)
        return set_err(c, EWG_INVALID_ARGUMENT,
                       "a tile-wavefront schedule needs a context created with one");
    if (p->precond_fp32) {
        int rc = ensure_fp32(c);
        if (rc) return rc;
    }
    if (p->operator_fp32) {
        int rc = ensure_op32(c);
        if (rc) return rc;
    }
    c->prm = *p;
    c->fp64_latched = false;
    return EWG_OK;
}

int64_t ewg_num_cells(const ewg_ctx *c) { return c ? c->g.N : 0; }
int64_t ewg_num_blocks(const ewg_ctx *c) { return c ? (int64_t)c->colidx.size() : 0; }
int64_t ewg_num_boundary_faces(const ewg_ctx *c) { return c ? c->B : 0; }

int ewg_set_static(ewg_ctx *c, const double *K, const double *phi, const int32_t *region,
                   const double *pe) {
    if (!c || !K || !phi || !region) return EWG_INVALID_ARGUMENT;
    const size_t N = c->g.N;
    for (size_t i = 0; i < N; ++i)
        if (region[i] < 0 || region[i] >= c->nregions)
            return set_err(c, EWG_INVALID_ARGUMENT, "region index out of range");
    if (pe && c->law != EWG_LAW_BROOKS_COREY)
        return set_err(c, EWG_INVALID_ARGUMENT, "per-cell entry pressure needs Brooks-Corey");
    HIP_TRY(c, hipMemcpy(c->K, K, N * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->phi, phi, N * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->region, region, N * sizeof(int), hipMemcpyHostToDevice));
    if (pe) {
        if (!c->pe) HIP_TRY(c, dalloc(&c->pe, N));
        HIP_TRY(c, hipMemcpy(c->pe, pe, N * sizeof(double), hipMemcpyHostToDevice));
    } else if (c->pe) {
        HIP_TRY(c, hipFree(c->pe));
        c->pe = nullptr;
    }
    c->static_set = true;
    if (c->bc_set) {
        launch_bc_aux(lin_args(c, 1.0, 0), c->bc_aux, c->stream);
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return EWG_OK;
}

int ewg_set_boundary(ewg_ctx *c, const int32_t *type, const double *values) {
    if (!c || !type || !values) return EWG_INVALID_ARGUMENT;
    for (long long f = 0; f < c->B; ++f)
        if (type[f] < EWG_BC_NO_FLOW || type[f] > EWG_BC_MASS_RATE)
            return set_err(c, EWG_INVALID_ARGUMENT, "unknown boundary type");
    HIP_TRY(c, hipMemcpy(c->bc_type, type, c->B * sizeof(int), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->bc_val, values, 2 * c->B * sizeof(double), hipMemcpyHostToDevice));
    c->bc_set = true;
    // the cells next to boundaries that inject the non-wetting phase (a rate, or a
    // free-flow state with S_n > 0), for the warm start's active region (side s of the
    // boundary faces: x-, x+, y-, y+, z-, z+, each in the order of the other two axes)
    {
        int bx[6] = {0x7fffffff, -1, 0x7fffffff, -1, 0x7fffffff, -1};
        const Grid &g = c->g;
        for (int s = 0; s < 2 * g.dim; ++s) {
            const int ax = s >> 1;
            for (long long bf = c->side_off[s]; bf < c->side_off[s + 1]; ++bf) {
                const bool inject = (type[bf] == 2 && values[2 * bf + 1] != 0.0) ||
                                    (type[bf] == 1 && 1.0 - values[2 * bf + 1] > 1e-8);
                if (!inject) continue;
                const long long loc = bf - c->side_off[s];
                int ijk[3];
                const int hi[3] = {g.nx - 1, g.ny - 1, g.nz - 1};
                if (ax == 0) { ijk[1] = (int)(loc % g.ny); ijk[2] = (int)(loc / g.ny); }
                else if (ax == 1) { ijk[0] = (int)(loc % g.nx); ijk[2] = (int)(loc / g.nx); }
                else { ijk[0] = (int)(loc % g.nx); ijk[1] = (int)(loc / g.nx); }
                ijk[ax] = (s & 1) ? hi[ax] : 0;
                for (int d = 0; d < 3; ++d) {
                    bx[2 * d] = std::min(bx[2 * d], ijk[d]);
                    bx[2 * d + 1] = std::max(bx[2 * d + 1], ijk[d]);
                }
            }
        }
        for (int k = 0; k < 6; ++k) c->src_box[k] = bx[k];
    }
    if (c->static_set) {
        launch_bc_aux(lin_args(c, 1.0, 0), c->bc_aux, c->stream);
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    return EWG_OK;
}

int ewg_set_solution(ewg_ctx *c, int time_idx, const double *x) {
    if (!c || !x || (time_idx != 0 && time_idx != 1)) return EWG_INVALID_ARGUMENT;
    HIP_TRY(c, hipMemcpy(time_idx == 0 ? c->x0 : c->x1, x, 2 * (size_t)c->g.N * sizeof(double),
                         hipMemcpyHostToDevice));
    return EWG_OK;
}

int ewg_get_solution(ewg_ctx *c, int time_idx, double *x) {
    if (!c || !x || (time_idx != 0 && time_idx != 1)) return EWG_INVALID_ARGUMENT;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(x, time_idx == 0 ? c->x0 : c->x1, 2 * (size_t)c->g.N * sizeof(double),
                         hipMemcpyDeviceToHost));
    return EWG_OK;
}

int ewg_advance_time_level(ewg_ctx *c) {
    if (!c) return EWG_INVALID_ARGUMENT;
    HIP_TRY(c, hipMemcpyAsync(c->x1, c->x0, 2 * (size_t)c->g.N * sizeof(double),
                              hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return EWG_OK;
}

int ewg_get_pattern(const ewg_ctx *c, int32_t *rowptr, int32_t *colidx) {
    if (!c || !rowptr || !colidx) return EWG_INVALID_ARGUMENT;
    std::copy(c->rowptr.begin(), c->rowptr.end(), rowptr);
    std::copy(c->colidx.begin(), c->colidx.end(), colidx);
    return EWG_OK;
}

int ewg_linearize(ewg_ctx *c, double dt, int newton_iter, double *resid, double *jac) {
    if (!c || !(dt > 0.0) || newton_iter < 0) return EWG_INVALID_ARGUMENT;
    if (!c->static_set) return set_err(c, EWG_INVALID_STATE, "static fields not set");
    if (int rc = capture_reference_pressure(c)) return rc;
    int rc = linearize_device(c, dt, newton_iter);
    if (rc) return rc;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const size_t N = c->g.N;
    if (resid) HIP_TRY(c, hipMemcpy(resid, c->R, 2 * N * sizeof(double), hipMemcpyDeviceToHost));
    if (jac) {
        std::vector<double> slots((size_t)c->g.nslots * N * 4);
        HIP_TRY(c, hipMemcpy(slots.data(), c->J, slots.size() * sizeof(double),
                             hipMemcpyDeviceToHost));
        for (size_t b = 0; b < c->bcrs_src.size(); ++b)
            std::memcpy(jac + 4 * b, slots.data() + 4 * c->bcrs_src[b], 4 * sizeof(double));
    }
    return EWG_OK;
}

int ewg_set_system(ewg_ctx *c, const double *resid, const double *jac) {
    if (!c) return EWG_INVALID_ARGUMENT;
    const size_t N = c->g.N;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (resid) HIP_TRY(c, hipMemcpy(c->R, resid, 2 * N * sizeof(double), hipMemcpyHostToDevice));
    if (jac) {
        // absent neighbours keep zero blocks in the slot layout
        std::vector<double> slots((size_t)c->g.nslots * N * 4, 0.0);
        for (size_t b = 0; b < c->bcrs_src.size(); ++b)
            std::memcpy(slots.data() + 4 * c->bcrs_src[b], jac + 4 * b, 4 * sizeof(double));
        HIP_TRY(c, hipMemcpy(c->J, slots.data(), slots.size() * sizeof(double),
                             hipMemcpyHostToDevice));
    }
    if (resid || jac) c->have_system = true;
    return EWG_OK;
}

int ewg_solve(ewg_ctx *c, double *dx, int32_t *iterations) {
    if (!c) return EWG_INVALID_ARGUMENT;
    if (!c->have_system) return set_err(c, EWG_INVALID_STATE, "no linearization");
    int its = 0;
    int rc = bicgstab(c, c->R, c->kx, &its);
    rc = fp64_retry(c, c->R, c->kx, rc, &its);
    c->last_linear_iterations = its;
    if (iterations) *iterations = its;
    if (dx && rc >= 0) {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        HIP_TRY(c, hipMemcpy(dx, c->kx, 2 * (size_t)c->g.N * sizeof(double), hipMemcpyDeviceToHost));
    }
    return rc;
}

int ewg_get_preconditioner(ewg_ctx *c, double *lu) {
    if (!c || !lu) return EWG_INVALID_ARGUMENT;
    if (!c->have_system) return set_err(c, EWG_INVALID_STATE, "no linearization");
    hipStream_t s = c->stream;
    HIP_TRY(c, hipMemsetAsync(c->ilu_err, 0, sizeof(int), s));
    precond_factor(c);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(c->err_host, c->ilu_err, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    c->have_precond = false;
    if (c->err_host[0]) return set_err(c, EWG_NUMERICAL_ISSUE, "Creating the preconditioner failed");
    c->have_precond = true;
    const Grid &g = c->g;
    const size_t N = g.N, nd = (size_t)g.dim;
    std::vector<double> L(nd * N * 4), D(N * 4), U(nd * N * 4);
    int rc = EWG_OK;
    const bool fp32 = c->prm.precond_fp32;
    if (use_wavefront_layout(c)) {
        rc = fp32 ? gather_layout(c, c->wl32, L, D, U) : gather_layout(c, c->wl64, L, D, U);
    } else if (fp32) {
        if ((rc = gather_natural<float>(c, c->L32, nullptr, L, 0, L.size())) == EWG_OK &&
            (rc = gather_natural<float>(c, nullptr, c->D32, D, 0, D.size())) == EWG_OK)
            rc = gather_natural<float>(c, c->U32, nullptr, U, 0, U.size());
    } else {
        if ((rc = gather_natural<double>(c, c->L64, nullptr, L, 0, L.size())) == EWG_OK &&
            (rc = gather_natural<double>(c, nullptr, c->D64, D, 0, D.size())) == EWG_OK)
            rc = gather_natural<double>(c, c->J + (size_t)(g.diag + 1) * N * 4, nullptr, U, 0,
                                        U.size());
    }
    if (rc != EWG_OK) return rc;
    for (size_t b = 0; b < c->bcrs_src.size(); ++b) {
        const size_t slot = (size_t)(c->bcrs_src[b] / (int64_t)N);
        const size_t i = (size_t)(c->bcrs_src[b] % (int64_t)N);
        const double *src = slot < (size_t)g.diag ? &L[(slot * N + i) * 4]
                            : slot == (size_t)g.diag ? &D[i * 4]
                                                     : &U[((slot - g.diag - 1) * N + i) * 4];
        std::memcpy(lu + 4 * b, src, 4 * sizeof(double));
    }
    return EWG_OK;
}

int ewg_apply_preconditioner(ewg_ctx *c, const double *d, double *v) {
    if (!c || !d || !v) return EWG_INVALID_ARGUMENT;
    if (!c->have_precond) return set_err(c, EWG_INVALID_STATE, "no preconditioner");
    hipStream_t s = c->stream;
    const size_t bytes = 2 * (size_t)c->g.N * sizeof(double);
    HIP_TRY(c, hipMemcpyAsync(c->kp, d, bytes, hipMemcpyHostToDevice, s));
    HIP_TRY(c, hipMemsetAsync(c->ilu_err, 0, sizeof(int), s));
    precond_apply(c, c->kp, c->ky, nullptr, false);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipMemcpyAsync(c->err_host, c->ilu_err, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(c, hipStreamSynchronize(s));
    if (c->err_host[0]) {
        // a tile-wavefront wait ran out: its edge buffers may hold unconsumed words
        if (c->wf.ex) wf_init_edges(c->wf, s);
        HIP_TRY(c, hipStreamSynchronize(s));
        return set_err(c, EWG_NUMERICAL_ISSUE, "ILU(0) sweep failed");
    }
    HIP_TRY(c, hipMemcpy(v, c->ky, bytes, hipMemcpyDeviceToHost));
    return EWG_OK;
}

int ewg_apply_preconditioner_noalloc(ewg_ctx *c, const double *d, double *v) {
    if (!c || !d || !v) return EWG_INVALID_ARGUMENT;
    const Grid &g = c->g;
    const bool fp32 = c->prm.precond_fp32;
    const bool layout = use_wavefront_layout(c);
    const size_t vec_bytes = layout ? wl_entries(g, c->wf) * 2 * (fp32 ? sizeof(float) : sizeof(double))
                                    : (fp32 ? 2 * (size_t)g.N * sizeof(float) : 0);
    if ((layout && 2 * c->wf.edge_words * 8 > kScratchEdges) || vec_bytes > kScratchVec)
        return set_err(c, EWG_INVALID_ARGUMENT,
                       "grid too large for the no-allocation scratch (16 MiB of tile edges, "
                       "32 MiB of intermediate)");
    if (!c->have_precond) return set_err(c, EWG_INVALID_STATE, "no preconditioner");
    if (c->device < 0 || c->device >= kMaxDevices || !g_scratch_ptr[c->device])
        return set_err(c, EWG_INVALID_STATE, "no-allocation scratch not set up on this device");
    std::lock_guard<std::mutex> lock(g_scratch_mutex[c->device]);
    unsigned char *base = g_scratch_ptr[c->device];
    hipStream_t s = c->stream;
    if (layout) {
        WfSync w = c->wf;
        w.ticket = reinterpret_cast<unsigned *>(base);
        w.ex = reinterpret_cast<unsigned long long *>(base + kScratchTickets);
        w.ey = w.ex + c->wf.edge_words;
        w.err = c->ilu_err;
        launch_flag_clear(c->ilu_err, s);
        if (fp32) {
            WlPrecond<float> wl = c->wl32;
            wl.w = reinterpret_cast<float *>(base + kScratchTickets + kScratchEdges);
            wl_apply32(g, w, wl, d, v, nullptr, s, false);
        } else {
            WlPrecond<double> wl = c->wl64;
            wl.w = reinterpret_cast<double *>(base + kScratchTickets + kScratchEdges);
            wl_apply64(g, w, wl, d, v, nullptr, s, false);
        }
        HIP_TRY(c, hipGetLastError());
        HIP_TRY(c, hipMemcpyAsync(c->err_host, c->ilu_err, sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_TRY(c, hipStreamSynchronize(s));
        if (c->err_host[0]) {
            // a wait ran out: the edge region may hold unconsumed words
            wf_init_edges(scratch_edges(base), s);
            HIP_TRY(c, hipStreamSynchronize(s));
            return set_err(c, EWG_NUMERICAL_ISSUE, "ILU(0) sweep failed");
        }
        return EWG_OK;
    }
    if (fp32) {
        float *w = reinterpret_cast<float *>(base + kScratchTickets + kScratchEdges);
        launch_ilu_apply32(g, c->L32, c->D32, c->U32, d, w, v, nullptr, s);
    } else {
        launch_ilu_apply64(g, c->L64, c->D64, c->J, d, v, nullptr, s);
    }
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(s));
    return EWG_OK;
}

int ewg_apply_operator(ewg_ctx *c, const double *x, double *y) {
    if (!c || !x || !y) return EWG_INVALID_ARGUMENT;
    if (!c->have_system) return set_err(c, EWG_INVALID_STATE, "no linearization");
    hipStream_t s = c->stream;
    const size_t bytes = 2 * (size_t)c->g.N * sizeof(double);
    HIP_TRY(c, hipMemcpyAsync(c->kp, x, bytes, hipMemcpyHostToDevice, s));
    if (c->prm.operator_fp32) launch_round_operator(c->g, c->J, c->J32, s);
    operator_apply(c, c->kp, c->ky, nullptr, nullptr, nullptr, nullptr, nullptr);
    HIP_TRY(c, hipGetLastError());
    HIP_TRY(c, hipStreamSynchronize(s));
    HIP_TRY(c, hipMemcpy(y, c->ky, bytes, hipMemcpyDeviceToHost));
    return EWG_OK;
}

int ewg_newton(ewg_ctx *c, double dt, int32_t *iterations, double *final_error) {
    if (!c || !(dt > 0.0)) return EWG_INVALID_ARGUMENT;
    if (!c->static_set) return set_err(c, EWG_INVALID_STATE, "static fields not set");
    int rc = newton_step(c, dt);
    if (iterations) *iterations = c->num_iterations;
    if (final_error) *final_error = c->error;
    c->stats.last_newton_iterations = c->num_iterations;
    c->stats.last_newton_error = c->error;
    return rc;
}

int ewg_run(ewg_ctx *c, double t0, double t_end, double dt0, int32_t *steps_out) {
    if (!c || !(dt0 > 0.0)) return EWG_INVALID_ARGUMENT;
    if (!c->static_set) return set_err(c, EWG_INVALID_STATE, "static fields not set");
    const size_t vbytes = 2 * (size_t)c->g.N * sizeof(double);
    double t = t0;
    auto remaining = [&](double tt) { return std::max(0.0, t_end - tt); };
    double dt = std::min({dt0, c->prm.max_time_step_size, remaining(t)});
    int steps = 0;
    auto finished = [&](double tt, double dtt) {  // simulator.hh:489-496
        double eps = std::max(std::fabs(tt), dtt) * std::numeric_limits<double>::epsilon() * 1e3;
        return tt * (1.0 + eps) >= t_end;
    };
    while (!finished(t, dt)) {
        bool ok = false;
        for (int fails = 0; fails < c->prm.max_time_step_divisions; ++fails) {
            int rc = newton_step(c, dt);
            if (rc == EWG_OK) {
                ok = true;
                break;
            }
            if (rc < 0) return rc;  // device / API error: not a convergence failure
            HIP_TRY(c, hipMemcpyAsync(c->x0, c->x1, vbytes, hipMemcpyDeviceToDevice, c->stream));
            c->stats.failed_steps++;
            double nextDt = dt / 2.0;
            if (dt < c->prm.min_time_step_size * (1 + 1e-9))
                return set_err(c, EWG_NOT_CONVERGED, "time integration failed at minimum step");
            if (nextDt < c->prm.min_time_step_size) nextDt = c->prm.min_time_step_size;
            dt = nextDt;
        }
        if (!ok) return set_err(c, EWG_NOT_CONVERGED, "Newton did not converge after dt divisions");
        HIP_TRY(c, hipMemcpyAsync(c->x1, c->x0, vbytes, hipMemcpyDeviceToDevice, c->stream));
        c->stats.time_steps++;
        t += dt;
        ++steps;
        double dtNext = std::min(c->prm.max_time_step_size, suggest_dt(c, dt));
        double simMax = remaining(t);
        if (dtNext < simMax && simMax < dtNext * 2) dtNext = simMax / 2 * 1.01;
        dt = std::min(simMax, dtNext);
        if (finished(t, dt)) break;
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (steps_out) *steps_out = steps;
    return EWG_OK;
}

int ewg_get_stats(const ewg_ctx *c, ewg_stats *s) {
    if (!c || !s) return EWG_INVALID_ARGUMENT;
    *s = c->stats;
    s->last_newton_iterations = c->num_iterations;
    s->last_linear_iterations = c->last_linear_iterations;
    s->last_newton_error = c->error;
    return EWG_OK;
}

int ewg_reset_stats(ewg_ctx *c) {
    if (!c) return EWG_INVALID_ARGUMENT;
    c->stats = ewg_stats{};
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &sm : c->samples) {
        sm.armed = sm.pending = false;
        sm.count = 0;
        sm.total_ms = 0.0;
    }
    c->warm_accepted = c->warm_rejected = 0;
    c->warm_predict_s = 0.0;
    c->warm_pending = false;
    return EWG_OK;
}

int ewg_kernel_samples(ewg_ctx *c, int which, int64_t *count, double *total_ms) {
    if (!c || !count || !total_ms || which < 0 || which > 3) return EWG_INVALID_ARGUMENT;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    sample_collect(c);
    *count = c->samples[which].count;
    *total_ms = c->samples[which].total_ms;
    return EWG_OK;
}

int ewg_time_kernel(ewg_ctx *c, int which, int reps, double *mean_ms) {
    if (!c || reps < 1 || !mean_ms) return EWG_INVALID_ARGUMENT;
    if (!c->static_set) return set_err(c, EWG_INVALID_STATE, "static fields not set");
    hipStream_t s = c->stream;
    const bool cold = (which & EWG_KERNEL_COLD) != 0;
    which &= ~EWG_KERNEL_COLD;
    if (cold && which == EWG_KERNEL_COPY)
        return set_err(c, EWG_INVALID_ARGUMENT, "EWG_KERNEL_COLD does not apply to the copy");
    // the copy moves EWG_COPY_BYTES from one buffer to another, both far larger than the
    // 256 MB last-level cache
    constexpr size_t kCopyDoubles = EWG_COPY_BYTES / sizeof(double);
    if ((which == EWG_KERNEL_COPY || cold) && !c->copy_dst) {
        HIP_TRY(c, dalloc(&c->copy_src, kCopyDoubles));
        HIP_TRY(c, dalloc(&c->copy_dst, kCopyDoubles));
        HIP_TRY(c, hipMemsetAsync(c->copy_src, 0, EWG_COPY_BYTES, s));
    }
    if (which != EWG_KERNEL_LINEARIZE && which != EWG_KERNEL_COPY && !c->have_system) {
        int rc = linearize_device(c, 1.0, 0);
        if (rc) return rc;
    }
    if (which == EWG_KERNEL_ILU_APPLY) {
        HIP_TRY(c, hipMemsetAsync(c->ilu_err, 0, sizeof(int), s));
        precond_factor(c);
    }
    if (which == EWG_KERNEL_SPMV && c->prm.operator_fp32)
        launch_round_operator(c->g, c->J, c->J32, s);
    hipEvent_t e0, e1;
    HIP_TRY(c, hipEventCreate(&e0));
    HIP_TRY(c, hipEventCreate(&e1));
    if (which == EWG_KERNEL_LINEARIZE)
        if (int rc = capture_reference_pressure(c)) return rc;
    LinArgs a = lin_args(c, 1.0, 1);
    float cold_ms = 0.f;
    if (!cold) HIP_TRY(c, hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) {
        if (cold) {
            // a read-only pass over EWG_COPY_BYTES: the caches then hold clean lines of
            // other data (a copy would also leave them full of dirty lines to write back)
            launch_max_abs((int)kCopyDoubles, c->copy_src, c->part0, s);
            HIP_TRY(c, hipEventRecord(e0, s));
        }
        switch (which) {
            case EWG_KERNEL_LINEARIZE: launch_linearize(a, s); break;
            case EWG_KERNEL_SPMV:
                operator_apply(c, c->x0, c->ky, nullptr, nullptr, nullptr, nullptr, nullptr);
                break;
            case EWG_KERNEL_ILU_APPLY: precond_apply(c, c->R, c->ky, nullptr, false); break;
            case EWG_KERNEL_ILU_FACTOR: precond_factor(c); break;
            case EWG_KERNEL_COPY: launch_copy(kCopyDoubles, c->copy_src, c->copy_dst, s); break;
            default:
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                return set_err(c, EWG_INVALID_ARGUMENT, "unknown kernel id");
        }
        if (cold) {
            HIP_TRY(c, hipEventRecord(e1, s));
            HIP_TRY(c, hipEventSynchronize(e1));
            float one = 0.f;
            HIP_TRY(c, hipEventElapsedTime(&one, e0, e1));
            cold_ms += one;
        }
    }
    float ms = cold_ms;
    if (!cold) {
        HIP_TRY(c, hipEventRecord(e1, s));
        HIP_TRY(c, hipEventSynchronize(e1));
        HIP_TRY(c, hipEventElapsedTime(&ms, e0, e1));
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIP_TRY(c, hipGetLastError());
    *mean_ms = ms / reps;
    return EWG_OK;
}

// device name / CU count for run records (no counterpart in the reference)
int ewg_device_info(char *buf, int len) {
    if (!buf || len < 1) return EWG_INVALID_ARGUMENT;
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) {
        std::snprintf(buf, len, "unknown");
        return EWG_DEVICE_ERROR;
    }
    std::snprintf(buf, len, "%s (%s), %d CUs, %.1f GiB", p.name, p.gcnArchName,
                  p.multiProcessorCount, p.totalGlobalMem / 1073741824.0);
    return EWG_OK;
}

}  // extern "C"

extern "C" {

int ewg_set_warm_start(ewg_ctx *c, ewg_cnn *net) {
    if (!c) return EWG_INVALID_ARGUMENT;
    if (net && !cnn::fits_grid(net, c->grid.nx, c->grid.ny, c->grid.nz))
        return set_err(c, EWG_INVALID_ARGUMENT, "warm start: network made for another grid");
    c->cnn = net;
    return EWG_OK;
}

int ewg_warm_start_stats(const ewg_ctx *c, int64_t *accepted, int64_t *rejected,
                         double *predict_s, double *last_initial_error) {
    if (!c) return EWG_INVALID_ARGUMENT;
    if (last_initial_error) *last_initial_error = c->first_error;
    if (accepted) *accepted = c->warm_accepted;
    if (rejected) *rejected = c->warm_rejected;
    if (predict_s) *predict_s = c->warm_predict_s + warm_pending_s(c);
    return EWG_OK;
}

}  // extern "C"
