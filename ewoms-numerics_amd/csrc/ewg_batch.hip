// ewg_batch.hip -- ensembles of realizations in one persistent launch.
//
// Workgroup b runs member order[b] through all requested time steps: per step the Newton
// loop of newtonmethod.hh:308-517 (the rules of the single-context driver), per Newton
// iteration the linearization (ewg_lin_device.h, one cell per thread and stride), the
// max-norm, the ILU(0) factorization and BiCGStab (bicgstabsolver.hh:116-326 with the
// criterion of combinedcriterion.hh) with the hyperplane-ordered ILU(0) sweeps of
// ewg_cell_ops.h, and the update. A barrier of the workgroup separates hyperplanes and
// passes; the workgroup's own stores are visible to it after the barrier (one compute
// unit, workgroup-scope fences). Members never wait for each other: each workgroup ends
// when its member is done, so a member's convergence (Newton and Krylov) is its own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/ewg_ensemble.h"
#include "ewg_cell_ops.h"
#include "ewg_ctx.h"
#include "ewg_lin_device.h"

namespace ewg {
namespace batch {

// one member's device arrays and parameters
struct Member {
    LinArgs lin;  // lin.x is solution(0)
    double *x0, *x1, *xcur;
    double *L, *D;                     // FP64 factors
    float *L32, *D32, *U32, *w32;      // FP32 preconditioner (fp32 = 1)
    double *r, *p, *v, *y, *z, *dx;    // BiCGStab vectors
    double newton_tol, newton_max_error, lin_rel_tol, lin_abs_tol, lin_max_error;
    int newton_max_it, lin_max_it, fp32, pad_;
};

struct Outcome {
    long long steps, newton, linear, solves;
    int status, pad_;
    double last_error;
};

constexpr double kBreakdownEps = 2.2250738585072014e-308 * 1e10;  // DBL_MIN * 1e10
constexpr int KS_RUN = 0, KS_CONV = 1, KS_FAIL = 2, KS_BREAK = 3;

template <int NT>
__device__ __forceinline__ double wg_sum(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < NT / 64; ++k) t += red[k];
    return t;
}
template <int NT>
__device__ __forceinline__ double wg_max(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < NT / 64; ++k) t = fmax(t, red[k]);
    return t;
}
__device__ __forceinline__ bool wg_any(bool v) { return __syncthreads_or(v) != 0; }

__device__ __forceinline__ int level_count(const Grid &g) { return g.dim == 3 ? g.ny * g.nz : g.ny; }

template <int NT>
struct Solver {
    const Member &m;
    const LinArgs &la;  // the workgroup's copy in LDS (dt, newton_iter set per iteration)
    double *red;

    __device__ __forceinline__ int n2() const { return 2 * la.g.N; }

    // cells of every hyperplane x + y + z = l in order (reverse: descending l)
    template <class F>
    __device__ __forceinline__ void levels(bool reverse, F &&f) const {
        const Grid &g = la.g;
        const int cnt = level_count(g);
        for (int q = 0; q < g.nlevels; ++q) {
            const int l = reverse ? g.nlevels - 1 - q : q;
            for (int t = threadIdx.x; t < cnt; t += NT) {
                int ijk[3];
                const int i = cell::level_cell(g, l, t, ijk);
                if (i >= 0) f(i, ijk);
            }
            __syncthreads();
        }
    }

    // out = (LU)^-1 d
    __device__ void apply(const double *d, double *out) const {
        const Grid &g = la.g;
        if (m.fp32) {
            levels(false, [&](int i, const int *ijk) { cell::ilu_fwd_cell<float>(g, i, ijk, m.L32, d, m.w32); });
            levels(true, [&](int i, const int *ijk) {
                cell::ilu_bwd_cell<float, float>(g, i, ijk, m.U32, 0, m.D32, m.w32, out);
            });
        } else {
            levels(false, [&](int i, const int *ijk) { cell::ilu_fwd_cell<double>(g, i, ijk, m.L, d, out); });
            levels(true, [&](int i, const int *ijk) {
                cell::ilu_bwd_cell<double, double>(g, i, ijk, la.J, g.diag + 1, m.D, out, out);
            });
        }
    }

    // y = A x; returns sum_i a_i . (A x)_i and, if b, sum_i b_i . (A x)_i
    __device__ void spmv(const double *x, double *y, const double *a, double &da,
                         const double *b, double &db) const {
        const Grid &g = la.g;
        double sa = 0.0, sb = 0.0;
        for (int i = threadIdx.x; i < g.N; i += NT) {
            const double2 yv = cell::spmv_cell(g, i, la.J, x);
            reinterpret_cast<double2 *>(y)[i] = yv;
            const double2 av = reinterpret_cast<const double2 *>(a)[i];
            sa += av.x * yv.x + av.y * yv.y;
            if (b) {
                const double2 bv = reinterpret_cast<const double2 *>(b)[i];
                sb += bv.x * yv.x + bv.y * yv.y;
            }
        }
        da = wg_sum<NT>(sa, red);
        if (b) db = wg_sum<NT>(sb, red);
    }

    // J dx = R (right-preconditioned BiCGStab as the single-context driver runs it);
    // returns 0, or EWG_NUMERICAL_ISSUE / EWG_NOT_CONVERGED
    __device__ int bicgstab(int &iterations) const {
        const Grid &g = la.g;
        const int n = n2();
        const double *b = la.R;
        iterations = 0;
        // factorization (FP64; the FP32 preconditioner keeps it rounded)
        bool singular = false;
        levels(false, [&](int i, const int *ijk) {
            if (!cell::ilu_factor_cell<double>(g, i, ijk, la.J, m.L, m.D, static_cast<double *>(nullptr)))
                singular = true;
        });
        if (m.fp32) {
            const int nl = g.dim * g.N * 4, nd = g.N * 4;
            const double *Aup = la.J + (size_t)(g.diag + 1) * g.N * 4;
            for (int e = threadIdx.x; e < nl; e += NT) {
                m.L32[e] = static_cast<float>(m.L[e]);
                m.U32[e] = static_cast<float>(Aup[e]);
            }
            for (int e = threadIdx.x; e < nd; e += NT) m.D32[e] = static_cast<float>(m.D[e]);
        }
        if (wg_any(singular)) return EWG_NUMERICAL_ISSUE;
        double mx = 0.0, dd = 0.0;
        for (int i = threadIdx.x; i < n; i += NT) {
            const double bi = b[i];
            m.r[i] = bi;
            m.dx[i] = 0.0;
            m.p[i] = 0.0;
            m.v[i] = 0.0;
            mx = fmax(mx, fabs(bi));
            dd += bi * bi;
        }
        mx = wg_max<NT>(mx, red);
        dd = wg_sum<NT>(dd, red);
        const double e0 = fmax(mx, kBreakdownEps);
        double resid = e0, init = e0;
        if (e0 / fmax(1e-20, e0) <= m.lin_rel_tol || e0 <= m.lin_abs_tol) return 0;
        double rho = 1.0, alpha = 1.0, omega = 1.0, rho_next = dd;
        int status = KS_RUN;
        auto criterion = [&](double err, bool stagnates) {
            resid = err;
            const double red_ = err / fmax(1e-20, init);
            const bool conv = red_ <= m.lin_rel_tol || err <= m.lin_abs_tol;
            const bool failed = !conv && (stagnates || err > m.lin_max_error);
            if (conv) status = KS_CONV;
            else if (failed) status = KS_FAIL;
        };
        for (int it = 0; it < m.lin_max_it && status == KS_RUN; ++it) {
            // beta
            if (fabs(rho) <= kBreakdownEps || fabs(omega) <= kBreakdownEps) {
                status = KS_BREAK;
                break;
            }
            const double beta = (rho_next / rho) * (alpha / omega);
            rho = rho_next;
            for (int i = threadIdx.x; i < n; i += NT) {
                double tmp = m.v[i];
                tmp *= omega;
                tmp -= m.p[i];
                tmp *= -beta;
                double pi = m.r[i];
                pi += tmp;
                m.p[i] = pi;
            }
            __syncthreads();
            apply(m.p, m.y);
            double den, unused = 0.0;
            spmv(m.y, m.v, b, den, nullptr, unused);
            if (fabs(den) <= kBreakdownEps) {
                status = KS_BREAK;
                break;
            }
            alpha = rho / den;
            if (fabs(alpha) <= kBreakdownEps) {
                status = KS_BREAK;
                break;
            }
            double rm = 0.0;
            bool nz = false;
            for (int i = threadIdx.x; i < n; i += NT) {
                const double yi = m.y[i];
                double tmp = yi;
                tmp *= alpha;
                tmp += m.dx[i];
                m.dx[i] = tmp;
                tmp = m.v[i];
                tmp *= alpha;
                const double ri = m.r[i] - tmp;
                m.r[i] = ri;
                rm = fmax(rm, fabs(ri));
                nz |= (yi != 0.0);
            }
            rm = wg_max<NT>(rm, red);
            criterion(rm, !wg_any(nz));
            if (status != KS_RUN) break;
            apply(m.r, m.z);
            double tt, ts;
            spmv(m.z, m.y, m.y, tt, m.r, ts);
            if (fabs(tt) <= kBreakdownEps) {
                status = KS_BREAK;
                break;
            }
            omega = ts / tt;
            if (fabs(omega) <= kBreakdownEps) {
                status = KS_BREAK;
                break;
            }
            double d = 0.0;
            nz = false;
            for (int i = threadIdx.x; i < n; i += NT) {
                const double zi = m.z[i];
                m.dx[i] += omega * zi;
                const double ri = m.r[i] + -omega * m.y[i];
                m.r[i] = ri;
                d += b[i] * ri;
                nz |= (zi != 0.0);
            }
            rho_next = wg_sum<NT>(d, red);
            // the residual checked here is still s (the object r is updated afterwards)
            criterion(resid, !wg_any(nz));
            if (status == KS_RUN) iterations += 1;
        }
        if (status == KS_CONV) return 0;
        if (status == KS_BREAK) return EWG_NUMERICAL_ISSUE;
        return EWG_NOT_CONVERGED;
    }
};

// NewtonMethod::proceed_ (newtonmethod.hh:866-884)
__device__ __forceinline__ bool proceed(int it, double err, double lerr, const Member &m) {
    if (it < 1) return true;
    if (err <= m.newton_tol) return false;
    if (it >= m.newton_max_it) return err * 4.0 < lerr;
    return true;
}

// commit = 0: a trial (the tuner's): solve the step, then return to the state before it
template <int NT>
__global__ void __launch_bounds__(NT)
k_ensemble(const Member *members, const int *order, int nsteps, double dt, int commit,
           Outcome *out) {
    __shared__ double red[NT / 64];
    __shared__ LinArgs la;
    const int id = order[blockIdx.x];
    const Member &m = members[id];
    Outcome o = out[id];  // every thread holds the same copy; thread 0 writes it back
    if (o.status != 0) return;
    if (threadIdx.x == 0) {
        la = m.lin;
        la.dt = dt;
    }
    __syncthreads();
    const Solver<NT> S{m, la, red};
    const Grid &g = la.g;
    const int n = 2 * g.N;
    for (int step = 0; step < nsteps && o.status == 0; ++step) {
        int it = 0, rc = 0;
        double err = 1e100, lerr = 1e100;
        while (proceed(it, err, lerr, m)) {
            lerr = err;  // beginIteration_
            for (int i = threadIdx.x; i < n; i += NT) m.xcur[i] = m.x0[i];
            if (threadIdx.x == 0) la.newton_iter = it;
            __syncthreads();
            for (int i = threadIdx.x; i < g.N; i += NT) lin::iq_cell(la, i);
            __syncthreads();
            if (g.dim == 3) {
                for (int i = threadIdx.x; i < g.N; i += NT) lin::linearize_cell<3, double>(la, i);
            } else {
                for (int i = threadIdx.x; i < g.N; i += NT) lin::linearize_cell<2, double>(la, i);
            }
            __syncthreads();
            double e = 0.0;
            for (int i = threadIdx.x; i < n; i += NT) e = fmax(e, fabs(la.R[i]));
            e = wg_max<NT>(e, red);
            o.newton++;
            lerr = err;  // preSolve_
            err = e;
            if (err > m.newton_max_error) {
                rc = EWG_NUMERICAL_ISSUE;
                break;
            }
            if (!proceed(it, err, lerr, m)) {
                ++it;
                break;
            }
            int lits = 0;
            rc = S.bicgstab(lits);
            o.linear += lits;
            o.solves++;
            if (rc != 0) break;
            double s = 0.0;
            for (int i = threadIdx.x; i < n; i += NT) {
                const double d = m.dx[i];
                double nv = m.xcur[i];
                nv -= d;
                m.x0[i] = nv;
                s += fabs(d);
            }
            s = wg_sum<NT>(s, red);
            if (!isfinite(s)) {
                rc = EWG_NUMERICAL_ISSUE;
                break;
            }
            ++it;
        }
        if (rc == 0 && !(err <= m.newton_tol)) rc = EWG_NOT_CONVERGED;
        o.last_error = err;
        if (rc != 0 || !commit) {
            // keep the last converged state
            for (int i = threadIdx.x; i < n; i += NT) m.x0[i] = m.x1[i];
            if (rc != 0) o.status = rc;
        } else {
            for (int i = threadIdx.x; i < n; i += NT) m.x1[i] = m.x0[i];
            o.steps++;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[id] = o;
}

}  // namespace batch
}  // namespace ewg

using namespace ewg;
using namespace ewg::batch;

struct ewg_ensemble {
    std::vector<ewg_ctx *> members;
    std::vector<Outcome> outcome;
    Member *dmem = nullptr;
    int *dorder = nullptr;
    Outcome *dout = nullptr;
    hipStream_t stream = nullptr;
    int threads = 512;
    bool largest_first = true;
    hipGraphExec_t graph = nullptr;
    int graph_steps = -1, graph_threads = 0;
    bool graph_largest = true;
    double graph_dt = 0.0;
    std::string err;
};

namespace {

int set_err(ewg_ensemble *e, int code, const std::string &msg) {
    if (e) e->err = msg;
    return code;
}

#define ENS_TRY(e, expr)                                                                  \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return set_err(e, EWG_DEVICE_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int upload_order(ewg_ensemble *e, bool largest_first) {
    const int n = static_cast<int>(e->members.size());
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    if (largest_first)
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return e->members[a]->g.N > e->members[b]->g.N;
        });
    ENS_TRY(e, hipMemcpy(e->dorder, order.data(), n * sizeof(int), hipMemcpyHostToDevice));
    return EWG_OK;
}

int launch(ewg_ensemble *e, int threads, int steps, double dt, int commit) {
    const int n = static_cast<int>(e->members.size());
    switch (threads) {
        case 256:
            hipLaunchKernelGGL(k_ensemble<256>, dim3(n), dim3(256), 0, e->stream, e->dmem,
                               e->dorder, steps, dt, commit, e->dout);
            break;
        case 512:
            hipLaunchKernelGGL(k_ensemble<512>, dim3(n), dim3(512), 0, e->stream, e->dmem,
                               e->dorder, steps, dt, commit, e->dout);
            break;
        case 1024:
            hipLaunchKernelGGL(k_ensemble<1024>, dim3(n), dim3(1024), 0, e->stream, e->dmem,
                               e->dorder, steps, dt, commit, e->dout);
            break;
        default: return set_err(e, EWG_INVALID_ARGUMENT, "launch shape");
    }
    ENS_TRY(e, hipGetLastError());
    return EWG_OK;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

}  // namespace

extern "C" {

int ewg_ensemble_create(ewg_ctx *const *members, int32_t n, ewg_ensemble **out) {
    if (!members || n < 1 || !out) return EWG_INVALID_ARGUMENT;
    *out = nullptr;
    std::vector<Member> host(n);
    for (int i = 0; i < n; ++i) {
        ewg_ctx *c = members[i];
        if (!c || !c->static_set) return EWG_INVALID_STATE;
        if (c->prm.precond_fp32 && !c->L32) return EWG_INVALID_STATE;
        Member &m = host[i];
        m.lin = ctx_lin_args(c, 1.0, 0);
        m.lin.recycle = c->prm.recycle_first_iteration_storage;
        m.x0 = c->x0;
        m.x1 = c->x1;
        m.xcur = c->xcur;
        m.L = c->L64;
        m.D = c->D64;
        m.L32 = c->L32;
        m.D32 = c->D32;
        m.U32 = c->U32;
        m.w32 = c->w32;
        m.r = c->kr;
        m.p = c->kp;
        m.v = c->kv;
        m.y = c->ky;
        m.z = c->kz;
        m.dx = c->kx;
        m.newton_tol = c->prm.newton_tolerance;
        m.newton_max_error = c->prm.newton_max_error;
        m.newton_max_it = c->prm.newton_max_iterations;
        m.lin_rel_tol = c->prm.linear_tolerance;
        m.lin_abs_tol = c->prm.linear_abs_tolerance < 0.0 ? c->prm.newton_tolerance / 100.0
                                                          : c->prm.linear_abs_tolerance;
        m.lin_max_error = c->prm.linear_max_error;
        m.lin_max_it = c->prm.linear_max_iterations;
        m.fp32 = c->prm.precond_fp32 ? 1 : 0;
    }
    auto *e = new ewg_ensemble();
    e->members.assign(members, members + n);
    e->outcome.assign(n, Outcome{});
    auto fail = [&](hipError_t) {
        ewg_ensemble_destroy(e);
        return EWG_DEVICE_ERROR;
    };
    hipError_t h;
    if ((h = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking)) != hipSuccess) return fail(h);
    if ((h = hipMalloc(reinterpret_cast<void **>(&e->dmem), n * sizeof(Member))) != hipSuccess) return fail(h);
    if ((h = hipMalloc(reinterpret_cast<void **>(&e->dorder), n * sizeof(int))) != hipSuccess) return fail(h);
    if ((h = hipMalloc(reinterpret_cast<void **>(&e->dout), n * sizeof(Outcome))) != hipSuccess) return fail(h);
    if ((h = hipMemcpy(e->dmem, host.data(), n * sizeof(Member), hipMemcpyHostToDevice)) != hipSuccess) return fail(h);
    if ((h = hipMemcpy(e->dout, e->outcome.data(), n * sizeof(Outcome), hipMemcpyHostToDevice)) != hipSuccess) return fail(h);
    // the members' queued work (uploads) is done before the ensemble's stream reads it
    for (ewg_ctx *c : e->members)
        if ((h = hipStreamSynchronize(c->stream)) != hipSuccess) return fail(h);
    if (upload_order(e, true) != EWG_OK) return fail(hipErrorUnknown);
    *out = e;
    return EWG_OK;
}

void ewg_ensemble_destroy(ewg_ensemble *e) {
    if (!e) return;
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->graph) (void)hipGraphExecDestroy(e->graph);
    if (e->dmem) (void)hipFree(e->dmem);
    if (e->dorder) (void)hipFree(e->dorder);
    if (e->dout) (void)hipFree(e->dout);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char *ewg_ensemble_last_error(const ewg_ensemble *e) { return e ? e->err.c_str() : "null ensemble"; }

int ewg_ensemble_tune(ewg_ensemble *e, double dt, ewg_ensemble_tuning *out) {
    if (!e || !(dt > 0.0)) return EWG_INVALID_ARGUMENT;
    const int n = static_cast<int>(e->members.size());
    ewg_ensemble_tuning t{};
    const int shapes[3] = {256, 512, 1024};
    double best = 1e300;
    hipEvent_t e0, e1;
    ENS_TRY(e, hipEventCreate(&e0));
    ENS_TRY(e, hipEventCreate(&e1));
    for (int lf = 1; lf >= 0; --lf)
        for (int s = 0; s < 3; ++s) {
            const int k = t.ncandidates++;
            t.cand_threads[k] = shapes[s];
            t.cand_largest_first[k] = lf;
            if (int rc = upload_order(e, lf != 0)) return rc;
            ENS_TRY(e, hipMemcpyAsync(e->dout, e->outcome.data(), n * sizeof(Outcome),
                                      hipMemcpyHostToDevice, e->stream));
            ENS_TRY(e, hipEventRecord(e0, e->stream));
            if (int rc = launch(e, shapes[s], 1, dt, 0)) return rc;
            ENS_TRY(e, hipEventRecord(e1, e->stream));
            std::vector<Outcome> trial(n);
            ENS_TRY(e, hipMemcpyAsync(trial.data(), e->dout, n * sizeof(Outcome),
                                      hipMemcpyDeviceToHost, e->stream));
            ENS_TRY(e, hipStreamSynchronize(e->stream));
            float ms = 0.0f;
            ENS_TRY(e, hipEventElapsedTime(&ms, e0, e1));
            bool ok = true;
            for (int i = 0; i < n; ++i) ok &= trial[i].status == 0 || e->outcome[i].status != 0;
            t.cand_ms[k] = ok ? ms : -1.0;
            if (ok && ms < best) {
                best = ms;
                t.threads = shapes[s];
                t.largest_first = lf;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ENS_TRY(e, hipMemcpy(e->dout, e->outcome.data(), n * sizeof(Outcome), hipMemcpyHostToDevice));
    if (best >= 1e300) return set_err(e, EWG_NUMERICAL_ISSUE, "every candidate failed");
    e->threads = t.threads;
    e->largest_first = t.largest_first != 0;
    if (int rc = upload_order(e, e->largest_first)) return rc;
    if (out) *out = t;
    return EWG_OK;
}

int ewg_ensemble_run(ewg_ensemble *e, double dt, int32_t steps, double *seconds) {
    if (!e || !(dt > 0.0) || steps < 1) return EWG_INVALID_ARGUMENT;
    const int n = static_cast<int>(e->members.size());
    if (!e->graph || e->graph_steps != steps || e->graph_dt != dt ||
        e->graph_threads != e->threads || e->graph_largest != e->largest_first) {
        if (e->graph) {
            (void)hipGraphExecDestroy(e->graph);
            e->graph = nullptr;
        }
        hipGraph_t gr;
        ENS_TRY(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        int rc = launch(e, e->threads, steps, dt, 1);
        hipError_t h = hipStreamEndCapture(e->stream, &gr);
        if (rc) return rc;
        ENS_TRY(e, h);
        h = hipGraphInstantiate(&e->graph, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        ENS_TRY(e, h);
        e->graph_steps = steps;
        e->graph_dt = dt;
        e->graph_threads = e->threads;
        e->graph_largest = e->largest_first;
    }
    ENS_TRY(e, hipStreamSynchronize(e->stream));
    const double t0 = now_s();
    ENS_TRY(e, hipGraphLaunch(e->graph, e->stream));
    ENS_TRY(e, hipStreamSynchronize(e->stream));
    if (seconds) *seconds = now_s() - t0;
    ENS_TRY(e, hipMemcpy(e->outcome.data(), e->dout, n * sizeof(Outcome), hipMemcpyDeviceToHost));
    return EWG_OK;
}

int ewg_ensemble_stats(const ewg_ensemble *e, int32_t i, ewg_ensemble_member_stats *out) {
    if (!e || !out || i < 0 || i >= static_cast<int>(e->members.size())) return EWG_INVALID_ARGUMENT;
    const Outcome &o = e->outcome[i];
    out->steps = o.steps;
    out->newton_iterations = o.newton;
    out->linear_iterations = o.linear;
    out->linear_solves = o.solves;
    out->status = o.status;
    out->pad_ = 0;
    out->last_error = o.last_error;
    return EWG_OK;
}

}  // extern "C"
