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
    double *xs;                        // the time step's start state (retries)
    double newton_tol, newton_max_error, lin_rel_tol, lin_abs_tol, lin_max_error;
    int newton_max_it, lin_max_it, fp32, max_div;
};

struct Outcome {
    long long steps, newton, linear, solves;
    int status, divisions;
    double last_error;
    long long retries;  // solves redone with FP64 factors (batched mode, precond_fp32 = 2)
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
    if (threadIdx.x == 0) la = m.lin;
    __syncthreads();
    const Solver<NT> S{m, la, red};
    const Grid &g = la.g;
    const int n = 2 * g.N;
    // one Newton solve of dts from x0 = x1 (newtonmethod.hh:308-517)
    auto solve = [&](double dts) -> int {
        if (threadIdx.x == 0) la.dt = dts;
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
        return rc;
    };
    for (int step = 0; step < nsteps && o.status == 0; ++step) {
        // a failed solve is retried from the step's start with 2, 4, ... substeps
        // (FvBaseProblem::timeIntegration halves the step, fvbaseproblem.hh:507-553)
        for (int i = threadIdx.x; i < n; i += NT) m.xs[i] = m.x1[i];
        __syncthreads();
        int rc = 0;
        for (int level = 0; level <= m.max_div; ++level) {
            const int nsub = 1 << level;
            rc = 0;
            for (int j = 0; j < nsub && rc == 0; ++j) {
                rc = solve(dt / nsub);
                if (rc == 0)
                    for (int i = threadIdx.x; i < n; i += NT) m.x1[i] = m.x0[i];
                __syncthreads();
            }
            if (rc == 0) break;
            for (int i = threadIdx.x; i < n; i += NT) m.x0[i] = m.x1[i] = m.xs[i];
            __syncthreads();
            if (level < m.max_div) o.divisions++;
        }
        if (rc != 0 || !commit) {
            // the state the step started from
            for (int i = threadIdx.x; i < n; i += NT) m.x0[i] = m.x1[i] = m.xs[i];
            if (rc != 0) o.status = rc;
        } else {
            o.steps++;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[id] = o;
}


// ---------------------------------------------------------------------------------------
// Batched mode: every launch covers the cells of all members (block -> member maps), so
// the whole GPU works on each phase; reductions are per member (block partials summed in
// a fixed order by one block per member); the ILU(0) sweeps run hyperplane by
// hyperplane for all members at once. Kernels return at once for members that do not
// take part (act[m] = 0, or a Krylov status other than running).
// ---------------------------------------------------------------------------------------
constexpr int BT = 256;
constexpr int KS_IDLE = 4, KS_MAXIT = 5;

struct Maps {
    const int *cmem, *coff;  // cell launches: block -> member, member -> first block (n+1)
    const int *lmem, *loff;  // hyperplane launches
};

enum VecId { V_R = 0, V_P, V_V, V_Y, V_Z, V_DX, V_X0, V_X1, V_XCUR, V_RES, V_XS };
__device__ __forceinline__ double *vec(const Member &m, int k) {
    switch (k) {
        case V_R: return m.r;
        case V_P: return m.p;
        case V_V: return m.v;
        case V_Y: return m.y;
        case V_Z: return m.z;
        case V_DX: return m.dx;
        case V_X0: return m.x0;
        case V_X1: return m.x1;
        case V_XCUR: return m.xcur;
        case V_XS: return m.xs;
        default: return m.lin.R;
    }
}

__device__ __forceinline__ double bt_sum(double v) {
    __shared__ double r[BT / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) r[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < BT / 64; ++k) t += r[k];
    return t;
}
__device__ __forceinline__ double bt_max(double v) {
    __shared__ double r[BT / 64];
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) r[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    for (int k = 0; k < BT / 64; ++k) t = fmax(t, r[k]);
    return t;
}
// one block per member: reductions of the member's block partials
__device__ __forceinline__ double range_sum(const double *part, int b0, int b1) {
    double s = 0.0;
    for (int b = b0 + threadIdx.x; b < b1; b += BT) s += part[b];
    return bt_sum(s);
}
__device__ __forceinline__ double range_max(const double *part, int b0, int b1) {
    double s = 0.0;
    for (int b = b0 + threadIdx.x; b < b1; b += BT) s = fmax(s, part[b]);
    return bt_max(s);
}
__device__ __forceinline__ int range_any(const int *part, int b0, int b1) {
    int v = 0;
    for (int b = b0 + threadIdx.x; b < b1; b += BT) v |= part[b];
    return __syncthreads_or(v);
}

__global__ void __launch_bounds__(BT) k_b_copy(const Member *mem, Maps mp, const int *act, int dst, int src) {
    const int id = mp.cmem[blockIdx.x];
    if (!act[id]) return;
    const Member &m = mem[id];
    const int i = (blockIdx.x - mp.coff[id]) * BT + threadIdx.x;
    if (i >= m.lin.g.N) return;
    const double *a = vec(m, src);
    double *b = vec(m, dst);
    b[2 * i] = a[2 * i];
    b[2 * i + 1] = a[2 * i + 1];
}

// the two passes of the linearization (ewg_lin_device.h) with the member's arguments
template <int PASS>
__global__ void __launch_bounds__(BT) k_b_lin(const Member *mem, Maps mp, const int *act, const int *niter,
                                              const double *dts) {
    __shared__ LinArgs la;
    const int id = mp.cmem[blockIdx.x];
    if (!act[id]) return;
    const Member &m = mem[id];
    if (threadIdx.x == 0) {
        la = m.lin;
        la.dt = dts[id];
        la.newton_iter = niter[id];
    }
    __syncthreads();
    const int i = (blockIdx.x - mp.coff[id]) * BT + threadIdx.x;
    if (i >= la.g.N) return;
    if (PASS == 0) lin::iq_cell(la, i);
    else if (la.g.dim == 3) lin::linearize_cell<3, double>(la, i);
    else lin::linearize_cell<2, double>(la, i);
}

__global__ void __launch_bounds__(BT) k_b_maxabs(const Member *mem, Maps mp, const int *act, int v, double *part) {
    const int id = mp.cmem[blockIdx.x];
    if (!act[id]) return;
    const Member &m = mem[id];
    const int i = (blockIdx.x - mp.coff[id]) * BT + threadIdx.x;
    double e = 0.0;
    if (i < m.lin.g.N) {
        const double *x = vec(m, v);
        e = fmax(fabs(x[2 * i]), fabs(x[2 * i + 1]));
    }
    e = bt_max(e);
    if (threadIdx.x == 0) part[blockIdx.x] = e;
}

__global__ void __launch_bounds__(BT) k_b_reduce(Maps mp, const int *act, const double *part, int op, double *out) {
    const int id = blockIdx.x;
    if (!act[id]) return;
    const double r = op == 0 ? range_sum(part, mp.coff[id], mp.coff[id + 1])
                             : range_max(part, mp.coff[id], mp.coff[id + 1]);
    if (threadIdx.x == 0) out[id] = r;
}

__global__ void __launch_bounds__(BT) k_b_factor(const Member *mem, Maps mp, const int *act, int level, int *err) {
    const int id = mp.lmem[blockIdx.x];
    if (!act[id]) return;
    const Member &m = mem[id];
    const Grid &g = m.lin.g;
    if (level >= g.nlevels) return;
    int ijk[3];
    const int i = cell::level_cell(g, level, (blockIdx.x - mp.loff[id]) * BT + threadIdx.x, ijk);
    if (i < 0) return;
    if (!cell::ilu_factor_cell<double>(g, i, ijk, m.lin.J, m.L, m.D, static_cast<double *>(nullptr)))
        atomicExch(&err[id], 1);
}

__global__ void __launch_bounds__(BT) k_b_round(const Member *mem, Maps mp, const int *act) {
    const int id = mp.cmem[blockIdx.x];
    if (!act[id]) return;
    const Member &m = mem[id];
    if (!m.fp32) return;
    const Grid &g = m.lin.g;
    const int i = (blockIdx.x - mp.coff[id]) * BT + threadIdx.x;
    if (i >= g.N) return;
    const double *Aup = m.lin.J + (size_t)(g.diag + 1) * g.N * 4;
    for (int s = 0; s < g.dim; ++s)
        for (int e = 0; e < 4; ++e) {
            const size_t k = ((size_t)s * g.N + i) * 4 + e;
            m.L32[k] = static_cast<float>(m.L[k]);
            m.U32[k] = static_cast<float>(Aup[k]);
        }
    for (int e = 0; e < 4; ++e) m.D32[(size_t)i * 4 + e] = static_cast<float>(m.D[(size_t)i * 4 + e]);
}

__global__ void __launch_bounds__(BT) k_b_fwd(const Member *mem, Maps mp, const KrylovState *ks, int level, int src, int dst) {
    const int id = mp.lmem[blockIdx.x];
    if (ks[id].status != KS_RUN) return;
    const Member &m = mem[id];
    const Grid &g = m.lin.g;
    if (level >= g.nlevels) return;
    int ijk[3];
    const int i = cell::level_cell(g, level, (blockIdx.x - mp.loff[id]) * BT + threadIdx.x, ijk);
    if (i < 0) return;
    if (m.fp32) cell::ilu_fwd_cell<float>(g, i, ijk, m.L32, vec(m, src), m.w32);
    else cell::ilu_fwd_cell<double>(g, i, ijk, m.L, vec(m, src), vec(m, dst));
}

__global__ void __launch_bounds__(BT) k_b_bwd(const Member *mem, Maps mp, const KrylovState *ks, int level, int dst) {
    const int id = mp.lmem[blockIdx.x];
    if (ks[id].status != KS_RUN) return;
    const Member &m = mem[id];
    const Grid &g = m.lin.g;
    if (level >= g.nlevels) return;
    int ijk[3];
    const int i = cell::level_cell(g, level, (blockIdx.x - mp.loff[id]) * BT + threadIdx.x, ijk);
    if (i < 0) return;
    double *out = vec(m, dst);
    if (m.fp32) cell::ilu_bwd_cell<float, float>(g, i, ijk, m.U32, 0, m.D32, m.w32, out);
    else cell::ilu_bwd_cell<double, double>(g, i, ijk, m.lin.J, g.diag + 1, m.D, out, out);
}

// y = A x with sum_i a_i (A x)_i and sum_i b_i (A x)_i per block (b < 0: none)
__global__ void __launch_bounds__(BT) k_b_spmv(const Member *mem, Maps mp, const KrylovState *ks, int x, int y, int a, int b,
                                               double *partA, double *partB) {
    const int id = mp.cmem[blockIdx.x];
    if (ks[id].status != KS_RUN) return;
    const Member &m = mem[id];
    const Grid &g = m.lin.g;
    const int i = (blockIdx.x - mp.coff[id]) * BT + threadIdx.x;
    double sa = 0.0, sb = 0.0;
    if (i < g.N) {
        const double2 yv = cell::spmv_cell(g, i, m.lin.J, vec(m, x));
        reinterpret_cast<double2 *>(vec(m, y))[i] = yv;
        const double2 av = reinterpret_cast<const double2 *>(vec(m, a))[i];
        sa += av.x * yv.x + av.y * yv.y;
        if (b >= 0) {
            const double2 bv = reinterpret_cast<const double2 *>(vec(m, b))[i];
            sb += bv.x * yv.x + bv.y * yv.y;
        }
    }
    sa = bt_sum(sa);
    if (b >= 0) sb = bt_sum(sb);
    if (threadIdx.x == 0) {
        partA[blockIdx.x] = sa;
        if (b >= 0) partB[blockIdx.x] = sb;
    }
}

__global__ void __launch_bounds__(BT) k_b_init(const Member *mem, Maps mp, const int *act, double *partMax, double *partDot) {
    const int id = mp.cmem[blockIdx.x];
    if (!act[id]) return;
    const Member &m = mem[id];
    const int i = (blockIdx.x - mp.coff[id]) * BT + threadIdx.x;
    double mx = 0.0, d = 0.0;
    if (i < m.lin.g.N) {
        for (int c = 2 * i; c < 2 * i + 2; ++c) {
            const double bi = m.lin.R[c];
            m.r[c] = bi;
            m.dx[c] = 0.0;
            m.p[c] = 0.0;
            m.v[c] = 0.0;
            mx = fmax(mx, fabs(bi));
            d += bi * bi;
        }
    }
    mx = bt_max(mx);
    d = bt_sum(d);
    if (threadIdx.x == 0) {
        partMax[blockIdx.x] = mx;
        partDot[blockIdx.x] = d;
    }
}

// err: singular pivot of the member's factorization; serr: one of its tile-wavefront
// waits ran out (either is a breakdown of the member's solve)
__global__ void __launch_bounds__(BT) k_b_ks_init(const Member *mem, Maps mp, const int *act, const int *err,
                                                  const int *serr, KrylovState *ks,
                                                  const double *partMax, const double *partDot) {
    const int id = blockIdx.x;
    if (!act[id]) return;  // another member's solve may be running: left as it is
    const Member &m = mem[id];
    const double mx = range_max(partMax, mp.coff[id], mp.coff[id + 1]);
    const double d = range_sum(partDot, mp.coff[id], mp.coff[id + 1]);
    if (threadIdx.x == 0) {
        KrylovState &k = ks[id];
        k.rel_tol = m.lin_rel_tol;
        k.abs_tol = m.lin_abs_tol;
        k.max_resid = m.lin_max_error;
        const double e = fmax(mx, kBreakdownEps);
        k.resid_err = e;
        k.init_err = e;
        k.last_err = e;
        k.stagnates = 0;
        k.rho = 1.0;
        k.alpha = 1.0;
        k.omega = 1.0;
        k.rho_i = d;
        k.iterations = 0;
        const double red = e / fmax(1e-20, e);
        k.status = (err[id] || serr[id]) ? KS_BREAK
                   : ((red <= m.lin_rel_tol || e <= m.lin_abs_tol) ? KS_CONV : KS_RUN);
    }
}

__device__ __forceinline__ void b_criterion(KrylovState &k, double err, int stagnates) {
    k.last_err = k.resid_err;
    k.resid_err = err;
    k.stagnates = stagnates;
    const double red = err / fmax(1e-20, k.init_err);
    const bool conv = red <= k.rel_tol || err <= k.abs_tol;
    const bool failed = !conv && (stagnates || err > k.max_resid);
    if (conv) k.status = KS_CONV;
    else if (failed) k.status = KS_FAIL;
}

// rho_i: the solve start's r.r (k_b_ks_init), then the r_hat.r of the iteration before
// (k_b_ks_crit2), each reduced in the launches that produced its partials, so that
// successive iterations may run through different block maps
__global__ void __launch_bounds__(BT) k_b_ks_beta(const Member *mem, KrylovState *ks) {
    const int id = blockIdx.x;
    KrylovState &k = ks[id];
    if (k.status != KS_RUN) return;
    const double rho_i = k.rho_i;
    if (threadIdx.x == 0) {
        if (k.iterations >= mem[id].lin_max_it) {
            k.status = KS_MAXIT;
            return;
        }
        if (fabs(k.rho) <= kBreakdownEps || fabs(k.omega) <= kBreakdownEps) {
            k.status = KS_BREAK;
            return;
        }
        k.beta = (rho_i / k.rho) * (k.alpha / k.omega);
        k.rho = rho_i;
        k.rho_i = rho_i;
    }
}

__global__ void __launch_bounds__(BT) k_b_p_update(const Member *mem, Maps mp, const KrylovState *ks) {
    const int id = mp.cmem[blockIdx.x];
    const KrylovState &k = ks[id];
    if (k.status != KS_RUN) return;
    const Member &m = mem[id];
    const int i = (blockIdx.x - mp.coff[id]) * BT + threadIdx.x;
    if (i >= m.lin.g.N) return;
    const double omega = k.omega, beta = k.beta;
    for (int c = 2 * i; c < 2 * i + 2; ++c) {
        double tmp = m.v[c];
        tmp *= omega;
        tmp -= m.p[c];
        tmp *= -beta;
        double pi = m.r[c];
        pi += tmp;
        m.p[c] = pi;
    }
}

__global__ void __launch_bounds__(BT) k_b_ks_alpha(Maps mp, KrylovState *ks, const double *partDen) {
    const int id = blockIdx.x;
    KrylovState &k = ks[id];
    if (k.status != KS_RUN) return;
    const double den = range_sum(partDen, mp.coff[id], mp.coff[id + 1]);
    if (threadIdx.x == 0) {
        k.denom = den;
        if (fabs(den) <= kBreakdownEps) {
            k.status = KS_BREAK;
            return;
        }
        const double alpha = k.rho / den;
        k.alpha = alpha;
        if (fabs(alpha) <= kBreakdownEps) k.status = KS_BREAK;
    }
}

__global__ void __launch_bounds__(BT) k_b_hs_update(const Member *mem, Maps mp, const KrylovState *ks, double *partMax, int *partAny) {
    const int id = mp.cmem[blockIdx.x];
    const KrylovState &k = ks[id];
    if (k.status != KS_RUN) return;
    const Member &m = mem[id];
    const int i = (blockIdx.x - mp.coff[id]) * BT + threadIdx.x;
    const double alpha = k.alpha;
    double mx = 0.0;
    bool nz = false;
    if (i < m.lin.g.N) {
        for (int c = 2 * i; c < 2 * i + 2; ++c) {
            const double yi = m.y[c];
            double tmp = yi;
            tmp *= alpha;
            tmp += m.dx[c];
            m.dx[c] = tmp;
            tmp = m.v[c];
            tmp *= alpha;
            const double ri = m.r[c] - tmp;
            m.r[c] = ri;
            mx = fmax(mx, fabs(ri));
            nz |= (yi != 0.0);
        }
    }
    mx = bt_max(mx);
    const int any = __syncthreads_or(nz);
    if (threadIdx.x == 0) {
        partMax[blockIdx.x] = mx;
        partAny[blockIdx.x] = any;
    }
}

// serr[id] set: a wait of the member's preconditioner sweeps ran out, its result is not
// the preconditioner's (a breakdown of the member's solve)
__global__ void __launch_bounds__(BT) k_b_ks_crit1(Maps mp, KrylovState *ks, const double *partMax, const int *partAny,
                                                   const int *serr) {
    const int id = blockIdx.x;
    KrylovState &k = ks[id];
    if (k.status != KS_RUN) return;
    if (serr[id]) {
        if (threadIdx.x == 0) k.status = KS_BREAK;
        return;
    }
    const double mx = range_max(partMax, mp.coff[id], mp.coff[id + 1]);
    const int any = range_any(partAny, mp.coff[id], mp.coff[id + 1]);
    if (threadIdx.x == 0) b_criterion(k, mx, any == 0);
}

__global__ void __launch_bounds__(BT) k_b_ks_omega(Maps mp, KrylovState *ks, const double *partTT, const double *partTS) {
    const int id = blockIdx.x;
    KrylovState &k = ks[id];
    if (k.status != KS_RUN) return;
    const double tt = range_sum(partTT, mp.coff[id], mp.coff[id + 1]);
    const double ts = range_sum(partTS, mp.coff[id], mp.coff[id + 1]);
    if (threadIdx.x == 0) {
        if (fabs(tt) <= kBreakdownEps) {
            k.status = KS_BREAK;
            return;
        }
        const double omega = ts / tt;
        k.omega = omega;
        if (fabs(omega) <= kBreakdownEps) k.status = KS_BREAK;
    }
}

__global__ void __launch_bounds__(BT) k_b_xr_update(const Member *mem, Maps mp, const KrylovState *ks, double *partRho, int *partAny) {
    const int id = mp.cmem[blockIdx.x];
    const KrylovState &k = ks[id];
    if (k.status != KS_RUN) return;
    const Member &m = mem[id];
    const int i = (blockIdx.x - mp.coff[id]) * BT + threadIdx.x;
    const double omega = k.omega;
    double d = 0.0;
    bool nz = false;
    if (i < m.lin.g.N) {
        for (int c = 2 * i; c < 2 * i + 2; ++c) {
            const double zi = m.z[c];
            m.dx[c] += omega * zi;
            const double ri = m.r[c] + -omega * m.y[c];
            m.r[c] = ri;
            d += m.lin.R[c] * ri;
            nz |= (zi != 0.0);
        }
    }
    d = bt_sum(d);
    const int any = __syncthreads_or(nz);
    if (threadIdx.x == 0) {
        partRho[blockIdx.x] = d;
        partAny[blockIdx.x] = any;
    }
}

__global__ void __launch_bounds__(BT) k_b_ks_crit2(Maps mp, KrylovState *ks, const double *partRho,
                                                   const int *partAny, const int *serr) {
    const int id = blockIdx.x;
    KrylovState &k = ks[id];
    if (k.status != KS_RUN) return;
    if (serr[id]) {
        if (threadIdx.x == 0) k.status = KS_BREAK;
        return;
    }
    const int any = range_any(partAny, mp.coff[id], mp.coff[id + 1]);
    const double rho_i = range_sum(partRho, mp.coff[id], mp.coff[id + 1]);
    if (threadIdx.x == 0) {
        // the residual checked here is still s (the object r is updated afterwards)
        b_criterion(k, k.resid_err, any == 0);
        if (k.status == KS_RUN) {
            k.iterations += 1;
            k.rho_i = rho_i;  // the next iteration's r_hat.r
        }
    }
}

// x0 = xcur - dx, with sum |dx| per block (finiteness of the update)
__global__ void __launch_bounds__(BT) k_b_update(const Member *mem, Maps mp, const int *act, double *part) {
    const int id = mp.cmem[blockIdx.x];
    if (!act[id]) return;
    const Member &m = mem[id];
    const int i = (blockIdx.x - mp.coff[id]) * BT + threadIdx.x;
    double s = 0.0;
    if (i < m.lin.g.N) {
        for (int c = 2 * i; c < 2 * i + 2; ++c) {
            const double d = m.dx[c];
            double nv = m.xcur[c];
            nv -= d;
            m.x0[c] = nv;
            s += fabs(d);
        }
    }
    s = bt_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(BT) k_b_zero(int n, int *v) {
    const int i = blockIdx.x * BT + threadIdx.x;
    if (i < n) v[i] = 0;
}

// v[i] = 1 where mask[i] (the host's forced sweep failures)
__global__ void __launch_bounds__(BT) k_b_set(int n, int *v, const int *mask) {
    const int i = blockIdx.x * BT + threadIdx.x;
    if (i < n && mask[i]) v[i] = 1;
}

// v[i] = 0 where mask[i]
__global__ void __launch_bounds__(BT) k_b_zero_masked(int n, int *v, const int *mask) {
    const int i = blockIdx.x * BT + threadIdx.x;
    if (i < n && mask[i]) v[i] = 0;
}

// no member in a solve (the start of a run)
__global__ void __launch_bounds__(BT) k_b_ks_idle(int n, KrylovState *ks) {
    const int i = blockIdx.x * BT + threadIdx.x;
    if (i < n) ks[i].status = KS_IDLE;
}

// the rows of the batched run's per-member flags (ewg_ensemble::dmask)
enum MaskSlot {
    M_ACCEPT,   // solution(1) <- solution(0): a (sub)step converged
    M_RESTORE,  // solution(0), solution(1) <- the step's start: a solve failed
    M_SAVE,     // the step's start <- solution(1): a new step
    M_EDGE,     // tile edge buffers back to the sentinel (a wavefront wait ran out)
    M_UPD,      // Newton update
    M_LIN,      // linearization
    M_FAC,      // factorization and solve start
    M_FAC32,    // ... of the members whose solves use FP32 factors
    M_FAC64,    // ... of the members whose solves use FP64 factors
    M_PREC,     // each member's preconditioner precision (0 FP32, 1 FP64): the wavefront
                // batches' sweeps read it in place
    M_FORCE,    // sweep error flag set by the host (test hook)
    kMaskSlots
};

}  // namespace batch
}  // namespace ewg

using namespace ewg;
using namespace ewg::batch;

struct ewg_ensemble {
    std::vector<ewg_ctx *> members;
    std::vector<Member> host;
    std::vector<Outcome> outcome;
    Member *dmem = nullptr;
    int *dorder = nullptr;
    Outcome *dout = nullptr;
    hipStream_t stream = nullptr;
    int mode = 1;  // 0: persistent (one workgroup per member), 1: batched launches
    int threads = 512;
    bool largest_first = true;
    int replay = 4;  // batched mode: most BiCGStab iterations queued between state reads
    hipGraphExec_t graph = nullptr;
    int graph_steps = -1, graph_threads = 0;
    bool graph_largest = true;
    double graph_dt = 0.0;
    // batched mode
    int nblkC = 0, nblkL = 0, lmax = 0;
    int *cmem = nullptr, *coff = nullptr, *lmem = nullptr, *loff = nullptr;
    int *act = nullptr, *niter = nullptr, *ilu_err = nullptr;
    double *dts = nullptr;
    std::vector<double *> xs;  // per member: the time step's start state
    KrylovState *ks = nullptr;
    double *part0 = nullptr, *part1 = nullptr, *part2 = nullptr, *pres = nullptr;
    int *partAny = nullptr;
    hipGraphExec_t iter_graph = nullptr;  // one BiCGStab iteration of every member
    // members with the tiles' wavefront layout (params.ilu_schedule = 0, 3D, all of one
    // precision): the preconditioner applications run as batched tile-wavefront sweeps
    void *wl_py = nullptr, *wl_rz = nullptr;
    // FP32 preconditioner with FP64 retries (members with precond_fp32 = 2): a member whose
    // FP32-preconditioned solve fails redoes it, and makes its later solves, with FP64
    // factors (use64); the wavefront path then also launches these FP64 batches over the
    // members' FP64 layouts (any64), factorizing those with act64 set
    std::vector<int> retry_ok, use64;
    void *wl_py64 = nullptr, *wl_rz64 = nullptr;
    int *act64 = nullptr;
    bool any64 = false;
    bool regraph = false;   // the iteration graph must be captured again (any64 changed)
    bool use_graph = true;  // replay the iteration graph (else launch the iteration's kernels)
    int *h_fp32 = nullptr;  // pinned 0 and 1: a member's Member::fp32 on the hyperplane path
    // per member: set when one of its tile-wavefront waits (sweeps or factorization) ran
    // out; turned into a Krylov breakdown of that member
    int *sweep_err = nullptr;
    std::vector<int> force_left;         // per member: solves still to start with the flag set
    // the batched run's per-member flags, kMaskSlots rows of n (device copy of h_mask)
    int *dmask = nullptr;
    double *pres_upd = nullptr;          // per member: the Newton update's norm
    // pinned host staging of every host<->device transfer of the batched mode, n entries
    // each (h_mask: kMaskSlots x n); each is rewritten only after a stream synchronisation
    // that follows its last transfer
    int *h_mask = nullptr, *h_it = nullptr, *h_serr = nullptr;
    double *h_dts = nullptr, *h_res = nullptr, *h_res_upd = nullptr;
    KrylovState *h_ks = nullptr;
    // Compact rounds (tile-wavefront path): when the members whose solve runs hold at most
    // half of the cell blocks, a round's BiCGStab iterations launch over their blocks and
    // tiles only, through maps of these members (device cc_*, pinned host h_cc_*, rebuilt
    // when the set changes) instead of the full maps and the iteration graph, whose
    // launches also dispatch every block of the members that are done.
    int compact = 1;                     // 0: never, 1: at most half of the blocks, 2: always
    std::vector<int> nblk_of, ntile_of;  // per member: cell blocks, tiles
    int ntile_all = 0;
    int *cc_mem = nullptr, *cc_off = nullptr, *cc_t32 = nullptr, *cc_t64 = nullptr;
    int *h_cc_mem = nullptr, *h_cc_off = nullptr, *h_cc_t32 = nullptr, *h_cc_t64 = nullptr;
    std::vector<int> cc_set;              // the members of the current compact maps
    std::vector<int> cc_use64;            // their precisions when the maps were built
    int cc_nblk = 0, cc_n32 = 0, cc_n64 = 0;
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

// ------------------------------------------------------------------------ batched mode
template <class T>
hipError_t dal(T **p, size_t count) {
    return hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(count, 1) * sizeof(T));
}

int batched_setup(ewg_ensemble *e) {
    const int n = static_cast<int>(e->members.size());
    std::vector<int> cmem, coff(n + 1), lmem, loff(n + 1);
    for (int m = 0; m < n; ++m) {
        const Grid &g = e->host[m].lin.g;
        coff[m] = static_cast<int>(cmem.size());
        cmem.insert(cmem.end(), (g.N + BT - 1) / BT, m);
        loff[m] = static_cast<int>(lmem.size());
        const int lc = g.dim == 3 ? g.ny * g.nz : g.ny;
        lmem.insert(lmem.end(), (lc + BT - 1) / BT, m);
        e->lmax = std::max(e->lmax, g.nlevels);
    }
    coff[n] = static_cast<int>(cmem.size());
    loff[n] = static_cast<int>(lmem.size());
    e->nblkC = coff[n];
    e->nblkL = loff[n];
    ENS_TRY(e, dal(&e->cmem, cmem.size()));
    ENS_TRY(e, dal(&e->coff, coff.size()));
    ENS_TRY(e, dal(&e->lmem, lmem.size()));
    ENS_TRY(e, dal(&e->loff, loff.size()));
    ENS_TRY(e, dal(&e->act, n));
    ENS_TRY(e, dal(&e->niter, n));
    ENS_TRY(e, dal(&e->ilu_err, n));
    ENS_TRY(e, dal(&e->ks, n));
    ENS_TRY(e, dal(&e->part0, e->nblkC));
    ENS_TRY(e, dal(&e->part1, e->nblkC));
    ENS_TRY(e, dal(&e->part2, e->nblkC));
    ENS_TRY(e, dal(&e->partAny, e->nblkC));
    ENS_TRY(e, dal(&e->pres, n));
    ENS_TRY(e, dal(&e->dts, n));
    ENS_TRY(e, hipMemcpy(e->cmem, cmem.data(), cmem.size() * sizeof(int), hipMemcpyHostToDevice));
    ENS_TRY(e, hipMemcpy(e->coff, coff.data(), coff.size() * sizeof(int), hipMemcpyHostToDevice));
    ENS_TRY(e, hipMemcpy(e->lmem, lmem.data(), lmem.size() * sizeof(int), hipMemcpyHostToDevice));
    ENS_TRY(e, hipMemcpy(e->loff, loff.data(), loff.size() * sizeof(int), hipMemcpyHostToDevice));
    ENS_TRY(e, dal(&e->sweep_err, n));
    ENS_TRY(e, dal(&e->act64, n));
    ENS_TRY(e, dal(&e->dmask, kMaskSlots * n));
    ENS_TRY(e, hipMemset(e->dmask, 0, kMaskSlots * n * sizeof(int)));
    ENS_TRY(e, hipHostMalloc(reinterpret_cast<void **>(&e->h_fp32), 2 * sizeof(int)));
    e->h_fp32[0] = 0;
    e->h_fp32[1] = 1;
    e->retry_ok.assign(n, 0);
    e->use64.assign(n, 0);
    for (int m = 0; m < n; ++m) e->retry_ok[m] = e->members[m]->prm.precond_fp32 == 2;
    ENS_TRY(e, dal(&e->pres_upd, n));
    ENS_TRY(e, hipMemset(e->sweep_err, 0, n * sizeof(int)));
    e->force_left.assign(n, 0);
    ENS_TRY(e, hipHostMalloc(reinterpret_cast<void **>(&e->h_mask), kMaskSlots * n * sizeof(int)));
    ENS_TRY(e, hipHostMalloc(reinterpret_cast<void **>(&e->h_it), n * sizeof(int)));
    ENS_TRY(e, hipHostMalloc(reinterpret_cast<void **>(&e->h_serr), n * sizeof(int)));
    ENS_TRY(e, hipHostMalloc(reinterpret_cast<void **>(&e->h_dts), n * sizeof(double)));
    ENS_TRY(e, hipHostMalloc(reinterpret_cast<void **>(&e->h_res), n * sizeof(double)));
    ENS_TRY(e, hipHostMalloc(reinterpret_cast<void **>(&e->h_res_upd), n * sizeof(double)));
    ENS_TRY(e, hipHostMalloc(reinterpret_cast<void **>(&e->h_ks), n * sizeof(KrylovState)));
    e->nblk_of.assign(n, 0);
    e->ntile_of.assign(n, 0);
    e->ntile_all = 0;
    for (int m = 0; m < n; ++m) {
        e->nblk_of[m] = coff[m + 1] - coff[m];
        e->ntile_of[m] = e->members[m]->wf.ntiles;
        e->ntile_all += e->ntile_of[m];
    }
    ENS_TRY(e, dal(&e->cc_mem, e->nblkC));
    ENS_TRY(e, dal(&e->cc_off, n + 1));
    ENS_TRY(e, dal(&e->cc_t32, e->ntile_all));
    ENS_TRY(e, dal(&e->cc_t64, e->ntile_all));
    ENS_TRY(e, hipHostMalloc(reinterpret_cast<void **>(&e->h_cc_mem), std::max(e->nblkC, 1) * sizeof(int)));
    ENS_TRY(e, hipHostMalloc(reinterpret_cast<void **>(&e->h_cc_off), (n + 1) * sizeof(int)));
    ENS_TRY(e, hipHostMalloc(reinterpret_cast<void **>(&e->h_cc_t32), std::max(e->ntile_all, 1) * sizeof(int)));
    ENS_TRY(e, hipHostMalloc(reinterpret_cast<void **>(&e->h_cc_t64), std::max(e->ntile_all, 1) * sizeof(int)));
    // batched tile-wavefront sweeps when every member carries the layout
    bool wl = true;
    for (int m = 0; m < n; ++m) {
        const ewg_ctx *c = e->members[m];
        wl &= c->g.dim == 3 && c->prm.ilu_schedule == 0 && c->wf.ticket != nullptr &&
              c->g.N <= kWfMaxCells && c->wl64.L != nullptr &&
              (!c->prm.precond_fp32 || c->wl32.L != nullptr) &&
              (c->prm.precond_fp32 != 0) == (e->members[0]->prm.precond_fp32 != 0);
    }
    if (wl) {
        const bool fp32 = e->members[0]->prm.precond_fp32 != 0;
        std::vector<WlMember> py(n), rz(n);
        for (int m = 0; m < n; ++m) {
            const ewg_ctx *c = e->members[m];
            WlMember w{};
            w.g = c->g;
            w.sync = c->wf;
            w.sync.err = e->sweep_err + m;  // the member's own flag, checked per solve
            if (fp32) {
                w.L = c->wl32.L; w.U = c->wl32.U; w.D = c->wl32.D; w.w = c->wl32.w;
            } else {
                w.L = c->wl64.L; w.U = c->wl64.U; w.D = c->wl64.D; w.w = c->wl64.w;
            }
            w.status = &e->ks[m].status;  // device address of the member's Krylov status
            w.prec = e->dmask + M_PREC * n + m;
            w.want = 0;
            py[m] = w;
            py[m].A = c->J;               // the first batch also factorizes
            py[m].Af = c->wl_af;
            py[m].err = e->ilu_err + m;
            py[m].active = e->act + m;
            py[m].d = e->host[m].p;
            py[m].out = e->host[m].y;
            rz[m] = w;
            rz[m].d = e->host[m].r;
            rz[m].out = e->host[m].z;
        }
        e->wl_py = wl_batch_create(py, fp32, e->stream);
        e->wl_rz = wl_batch_create(rz, fp32, e->stream);
        if (!e->wl_py || !e->wl_rz) return set_err(e, EWG_DEVICE_ERROR, "wavefront batch");
        bool retry = false;
        for (int m = 0; m < n; ++m) retry |= e->retry_ok[m] != 0;
        if (fp32 && retry) {
            // the same members over their FP64 layouts, for the members that fall back
            for (int m = 0; m < n; ++m) {
                const ewg_ctx *c = e->members[m];
                for (WlMember *w : {&py[m], &rz[m]}) {
                    w->L = c->wl64.L; w->U = c->wl64.U; w->D = c->wl64.D; w->w = c->wl64.w;
                    w->want = 1;
                }
                py[m].active = e->act64 + m;
            }
            e->wl_py64 = wl_batch_create(py, false, e->stream);
            e->wl_rz64 = wl_batch_create(rz, false, e->stream);
            if (!e->wl_py64 || !e->wl_rz64) return set_err(e, EWG_DEVICE_ERROR, "FP64 wavefront batch");
        }
    }
    return EWG_OK;
}

Maps maps(const ewg_ensemble *e) { return Maps{e->cmem, e->coff, e->lmem, e->loff}; }

// out = (LU)^-1 in for all members whose Krylov status runs: tile-wavefront sweeps over
// the members' wavefront layouts, else hyperplane by hyperplane
// (compact: over the tiles of the compact maps' members only)
void launch_apply(ewg_ensemble *e, int src, int dst, bool compact) {
    if (e->wl_py) {
        if (compact) {
            wl_batch_apply_tiles(src == V_P ? e->wl_py : e->wl_rz, e->cc_t32, e->cc_n32, e->stream);
            if (e->any64)
                wl_batch_apply_tiles(src == V_P ? e->wl_py64 : e->wl_rz64, e->cc_t64, e->cc_n64,
                                     e->stream);
            return;
        }
        wl_batch_apply(src == V_P ? e->wl_py : e->wl_rz, e->stream);
        if (e->any64) wl_batch_apply(src == V_P ? e->wl_py64 : e->wl_rz64, e->stream);
        return;
    }
    const Maps mp = maps(e);
    for (int l = 0; l < e->lmax; ++l)
        hipLaunchKernelGGL(k_b_fwd, dim3(e->nblkL), dim3(BT), 0, e->stream, e->dmem, mp, e->ks, l, src, dst);
    for (int l = e->lmax - 1; l >= 0; --l)
        hipLaunchKernelGGL(k_b_bwd, dim3(e->nblkL), dim3(BT), 0, e->stream, e->dmem, mp, e->ks, l, dst);
}

// one BiCGStab iteration (bicgstabsolver.hh:191-322) of every member whose solve runs
// (compact: through the compact maps, over their members' blocks and tiles; the other
// members' solves are not running then)
void enqueue_iteration(ewg_ensemble *e, bool compact = false) {
    const int n = static_cast<int>(e->members.size());
    const Maps mp = compact ? Maps{e->cc_mem, e->cc_off, e->lmem, e->loff} : maps(e);
    const dim3 gC(compact ? e->cc_nblk : e->nblkC);
    if (gC.x == 0) return;
    hipLaunchKernelGGL(k_b_ks_beta, dim3(n), dim3(BT), 0, e->stream, e->dmem, e->ks);
    hipLaunchKernelGGL(k_b_p_update, gC, dim3(BT), 0, e->stream, e->dmem, mp, e->ks);
    launch_apply(e, V_P, V_Y, compact);
    hipLaunchKernelGGL(k_b_spmv, gC, dim3(BT), 0, e->stream, e->dmem, mp, e->ks,
                       (int)V_Y, (int)V_V, (int)V_RES, -1, e->part0, static_cast<double *>(nullptr));
    hipLaunchKernelGGL(k_b_ks_alpha, dim3(n), dim3(BT), 0, e->stream, mp, e->ks, e->part0);
    hipLaunchKernelGGL(k_b_hs_update, gC, dim3(BT), 0, e->stream, e->dmem, mp, e->ks, e->part1, e->partAny);
    hipLaunchKernelGGL(k_b_ks_crit1, dim3(n), dim3(BT), 0, e->stream, mp, e->ks, e->part1, e->partAny,
                       e->sweep_err);
    launch_apply(e, V_R, V_Z, compact);
    hipLaunchKernelGGL(k_b_spmv, gC, dim3(BT), 0, e->stream, e->dmem, mp, e->ks,
                       (int)V_Z, (int)V_Y, (int)V_Y, (int)V_R, e->part0, e->part1);
    hipLaunchKernelGGL(k_b_ks_omega, dim3(n), dim3(BT), 0, e->stream, mp, e->ks, e->part0, e->part1);
    hipLaunchKernelGGL(k_b_xr_update, gC, dim3(BT), 0, e->stream, e->dmem, mp, e->ks, e->part2, e->partAny);
    hipLaunchKernelGGL(k_b_ks_crit2, dim3(n), dim3(BT), 0, e->stream, mp, e->ks, e->part2, e->partAny,
                       e->sweep_err);
}

// The compact maps of the members in `set` (ascending): block -> member and each member's
// first block (members outside the set: empty ranges), and the FP32 and FP64 batches'
// tile lists of the set's members of that precision. Queued on the stream when the set or
// a member's precision changed (the staging is rewritten only after the round's
// synchronisation, which follows its copies).
int compact_maps(ewg_ensemble *e, const std::vector<int> &set) {
    const int n = static_cast<int>(e->members.size());
    std::vector<int> p64(set.size());
    for (size_t k = 0; k < set.size(); ++k) p64[k] = e->use64[set[k]];
    if (set == e->cc_set && p64 == e->cc_use64) return EWG_OK;
    int nb = 0, n32 = 0, n64 = 0, k = 0;
    for (int m = 0; m < n; ++m) {
        e->h_cc_off[m] = nb;
        if (k < (int)set.size() && set[k] == m) {
            ++k;
            if (nb + e->nblk_of[m] > e->nblkC) return set_err(e, EWG_INVALID_STATE, "compact maps: blocks");
            for (int b = 0; b < e->nblk_of[m]; ++b) e->h_cc_mem[nb++] = m;
            int *t = e->use64[m] ? e->h_cc_t64 : e->h_cc_t32;
            int &nt = e->use64[m] ? n64 : n32;
            if (nt + e->ntile_of[m] > e->ntile_all) return set_err(e, EWG_INVALID_STATE, "compact maps: tiles");
            for (int q = 0; q < e->ntile_of[m]; ++q) t[nt++] = m;
        }
    }
    e->h_cc_off[n] = nb;
    hipStream_t s = e->stream;
    if (nb > 0) ENS_TRY(e, hipMemcpyAsync(e->cc_mem, e->h_cc_mem, nb * sizeof(int), hipMemcpyHostToDevice, s));
    ENS_TRY(e, hipMemcpyAsync(e->cc_off, e->h_cc_off, (n + 1) * sizeof(int), hipMemcpyHostToDevice, s));
    if (n32 > 0) ENS_TRY(e, hipMemcpyAsync(e->cc_t32, e->h_cc_t32, n32 * sizeof(int), hipMemcpyHostToDevice, s));
    if (n64 > 0) ENS_TRY(e, hipMemcpyAsync(e->cc_t64, e->h_cc_t64, n64 * sizeof(int), hipMemcpyHostToDevice, s));
    e->cc_set = set;
    e->cc_use64 = p64;
    e->cc_nblk = nb;
    e->cc_n32 = n32;
    e->cc_n64 = n64;
    return EWG_OK;
}

// the same, captured once in a HIP graph
int iteration_graph(ewg_ensemble *e) {
    if (e->iter_graph) return EWG_OK;
    hipGraph_t gr;
    ENS_TRY(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    enqueue_iteration(e);
    hipError_t h = hipStreamEndCapture(e->stream, &gr);
    ENS_TRY(e, h);
    h = hipGraphInstantiate(&e->iter_graph, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    ENS_TRY(e, h);
    return EWG_OK;
}

bool proceed_h(int it, double err, double lerr, const Member &m) {
    if (it < 1) return true;
    if (err <= m.newton_tol) return false;
    if (it >= m.newton_max_it) return err * 4.0 < lerr;
    return true;
}

// The batched run: every member advances through its own time steps, Newton iterations
// (newtonmethod.hh:308-517) and BiCGStab solves, independently of the others. Each round
// of the host (one stream synchronisation) queues, for the members that need them: the
// step bookkeeping copies, the Newton update, the linearization and its max-norm, the
// factorization and start of a solve; then `replay` BiCGStab iterations of every member
// whose solve runs (members in other states skip them); then reads back the norms and
// Krylov states. The host applies each member's Newton rules to what it reads. A member's
// operations, and their order, are those of one context stepped alone, so its results
// are too; the rounds only decide which members' work shares a launch (a member that
// converges early goes on to its next iteration or step instead of waiting for the
// slowest). A failed solve is retried from the step's start with 2, 4, ... substeps
// (fvbaseproblem.hh:507-553 halves a failed step); a member whose step still fails keeps
// its last converged state and stops. commit = false (one step: the tuner's trial): the
// members return to their start state afterwards.
// After a Newton update the member's next linearization is queued in the same round,
// before the update's finiteness is known; if the update turns out non-finite the solve
// fails, and that linearization is neither counted nor used.
// Every transfer goes through the ensemble's pinned staging buffers (h_*), each rewritten
// only after the synchronisation that ends the round of its previous transfer.
int batched_run(ewg_ensemble *e, double dt, int steps, bool commit) {
    const int n = static_cast<int>(e->members.size());
    if ((int)e->force_left.size() != n || !e->h_mask || !e->h_ks || !e->h_res_upd || !e->dmask)
        return set_err(e, EWG_INVALID_STATE, "batched run: per-member buffers missing");
    if (!commit && steps != 1) return set_err(e, EWG_INVALID_ARGUMENT, "trial runs take one step");
    const Maps mp = maps(e);
    hipStream_t s = e->stream;
    // Per-member phases. Queued in a round: NEED_LIN (linearization), NEED_FAC
    // (factorization and solve start), NEED_UPD (update, with the next linearization when
    // the loop goes on). Read at the round's end: LIN_PEND, SOLVING, UPD_PEND. Host-only,
    // passed through within a round: STEP (a new time step), NSTART (a new Newton solve),
    // HEAD (the loop head), NDONE (a solve ended with rc). IDLE: done or out.
    enum { IDLE, STEP, NSTART, HEAD, NDONE, NEED_LIN, LIN_PEND, NEED_FAC, SOLVING, NEED_UPD, UPD_PEND };
    struct Prog {
        int phase = IDLE, it = 0, left = 0, level = 0, sub = 0, rc = 0;
        bool lin = false;  // a linearization is queued beside the update
        double err = 1e100, lerr = 1e100, dt = 0.0;
    };
    std::vector<Prog> P(n);
    int *mask = e->h_mask;  // kMaskSlots rows of n flags
    auto mark = [&](int slot, int m) { mask[slot * n + m] = 1; };
    auto marked = [&](int slot) {
        for (int m = 0; m < n; ++m)
            if (mask[slot * n + m]) return true;
        return false;
    };
    auto dm = [&](int slot) -> const int * { return e->dmask + slot * n; };
    // the host's transitions of member m up to its next device work (or its end)
    auto advance = [&](int m) {
        Prog &p = P[m];
        const Member &h = e->host[m];
        for (;;) {
            switch (p.phase) {
                case STEP:  // a new time step: its start state is kept for retries
                    mark(M_SAVE, m);
                    p.level = 0;
                    p.sub = 0;
                    p.phase = NSTART;
                    break;
                case NSTART:  // a Newton solve of the (sub)step from solution(1)
                    p.it = 0;
                    p.err = 1e100;  // newtonmethod.hh:197
                    p.lerr = 1e100;
                    p.dt = dt / (double)(1 << p.level);
                    p.phase = HEAD;
                    break;
                case HEAD:  // proceed_ (newtonmethod.hh:866-884)
                    if (proceed_h(p.it, p.err, p.lerr, h)) {
                        p.lerr = p.err;  // beginIteration_
                        p.phase = NEED_LIN;
                        return;
                    }
                    p.rc = p.err <= h.newton_tol ? 0 : EWG_NOT_CONVERGED;
                    p.phase = NDONE;
                    break;
                case NDONE:
                    e->outcome[m].last_error = p.err;
                    if (p.rc == 0) {
                        mark(M_ACCEPT, m);  // solution(1) <- solution(0)
                        if (++p.sub < (1 << p.level)) {
                            p.phase = NSTART;
                            break;
                        }
                        if (commit) e->outcome[m].steps++;
                        if (--p.left > 0) {
                            p.phase = STEP;
                            break;
                        }
                        p.phase = IDLE;
                        return;
                    }
                    mark(M_RESTORE, m);  // back to the step's start
                    p.sub = 0;
                    if (p.level < h.max_div) {
                        ++p.level;
                        e->outcome[m].divisions++;
                        p.phase = NSTART;
                        break;
                    }
                    e->outcome[m].status = p.rc;
                    p.phase = IDLE;
                    return;
                default:
                    return;
            }
        }
    };
    // the members' precisions at the start (a trial restores them)
    const std::vector<int> use64_in = e->use64;
    const bool any64_in = e->any64;
    std::vector<int> flip(n, 0);  // hyperplane path: Member::fp32 of these to 0 next round
    std::fill(mask, mask + kMaskSlots * n, 0);
    bool any_member = false;
    for (int m = 0; m < n; ++m) {
        if (e->outcome[m].status != 0) continue;
        P[m].left = steps;
        P[m].phase = STEP;
        advance(m);
        any_member = true;
    }
    std::vector<int> took(n, 0);  // members of this run (the trial restores them)
    for (int m = 0; m < n; ++m) took[m] = P[m].phase != IDLE;
    // no member starts in a solve
    hipLaunchKernelGGL(k_b_ks_idle, dim3((n + BT - 1) / BT), dim3(BT), 0, s, n, e->ks);
    while (any_member) {
        // ------------------------------------------------------------- queue this round
        if (e->use_graph && (e->regraph || !e->iter_graph)) {
            // one BiCGStab iteration of every member, now with (or without) the FP64
            // batches' sweeps
            if (e->iter_graph) (void)hipGraphExecDestroy(e->iter_graph);
            e->iter_graph = nullptr;
            e->regraph = false;
            if (int r = iteration_graph(e)) return r;
        }
        for (int m = 0; m < n; ++m) {
            if (!flip[m]) continue;
            flip[m] = 0;
            ENS_TRY(e, hipMemcpyAsync(&e->dmem[m].fp32, &e->h_fp32[0], sizeof(int), hipMemcpyHostToDevice, s));
        }
        bool lin = false, fac = false, fac32 = false, fac64 = false, upd = false, forced = false,
             solving = false;
        for (int m = 0; m < n; ++m) {
            Prog &p = P[m];
            e->h_it[m] = p.it;
            e->h_dts[m] = p.dt;
            mask[M_PREC * n + m] = e->use64[m];
            switch (p.phase) {
                case NEED_UPD:
                    mark(M_UPD, m);
                    upd = true;
                    p.phase = UPD_PEND;
                    // the loop head after a finite update
                    if (proceed_h(p.it + 1, p.err, p.lerr, e->host[m])) {
                        mark(M_LIN, m);
                        e->h_it[m] = p.it + 1;
                        p.lin = lin = true;
                    }
                    break;
                case NEED_LIN:
                    mark(M_LIN, m);
                    lin = true;
                    p.phase = LIN_PEND;
                    break;
                case NEED_FAC:
                    mark(M_FAC, m);
                    mark(e->use64[m] ? M_FAC64 : M_FAC32, m);
                    fac = true;
                    (e->use64[m] ? fac64 : fac32) = true;
                    if (e->force_left[m] > 0) {
                        --e->force_left[m];
                        mark(M_FORCE, m);
                        forced = true;
                    }
                    p.phase = SOLVING;
                    break;
                default:
                    break;
            }
            solving |= p.phase == SOLVING;
        }
        const bool copies = marked(M_ACCEPT) || marked(M_RESTORE) || marked(M_SAVE) || marked(M_EDGE);
        if (!lin && !fac && !upd && !solving && !copies) break;
        ENS_TRY(e, hipMemcpyAsync(e->dmask, mask, kMaskSlots * n * sizeof(int), hipMemcpyHostToDevice, s));
        ENS_TRY(e, hipMemcpyAsync(e->niter, e->h_it, n * sizeof(int), hipMemcpyHostToDevice, s));
        ENS_TRY(e, hipMemcpyAsync(e->dts, e->h_dts, n * sizeof(double), hipMemcpyHostToDevice, s));
        const dim3 gC(e->nblkC), gN((n + BT - 1) / BT), bT(BT);
        if (marked(M_ACCEPT))
            hipLaunchKernelGGL(k_b_copy, gC, bT, 0, s, e->dmem, mp, dm(M_ACCEPT), (int)V_X1, (int)V_X0);
        if (marked(M_RESTORE)) {
            hipLaunchKernelGGL(k_b_copy, gC, bT, 0, s, e->dmem, mp, dm(M_RESTORE), (int)V_X0, (int)V_XS);
            hipLaunchKernelGGL(k_b_copy, gC, bT, 0, s, e->dmem, mp, dm(M_RESTORE), (int)V_X1, (int)V_XS);
        }
        if (marked(M_SAVE))
            hipLaunchKernelGGL(k_b_copy, gC, bT, 0, s, e->dmem, mp, dm(M_SAVE), (int)V_XS, (int)V_X1);
        // a member whose wavefront wait ran out may have left unconsumed words in its tile
        // edge buffers: back to the sentinel before its next sweep
        if (e->wl_py)
            for (int m = 0; m < n; ++m)
                if (mask[M_EDGE * n + m]) wf_init_edges(e->members[m]->wf, s);
        if (upd) {
            hipLaunchKernelGGL(k_b_update, gC, bT, 0, s, e->dmem, mp, dm(M_UPD), e->part0);
            hipLaunchKernelGGL(k_b_reduce, dim3(n), bT, 0, s, mp, dm(M_UPD), e->part0, 0, e->pres_upd);
        }
        if (lin) {
            hipLaunchKernelGGL(k_b_copy, gC, bT, 0, s, e->dmem, mp, dm(M_LIN), (int)V_XCUR, (int)V_X0);
            hipLaunchKernelGGL(k_b_lin<0>, gC, bT, 0, s, e->dmem, mp, dm(M_LIN), e->niter, e->dts);
            hipLaunchKernelGGL(k_b_lin<1>, gC, bT, 0, s, e->dmem, mp, dm(M_LIN), e->niter, e->dts);
            hipLaunchKernelGGL(k_b_maxabs, gC, bT, 0, s, e->dmem, mp, dm(M_LIN), (int)V_RES, e->part0);
            hipLaunchKernelGGL(k_b_reduce, dim3(n), bT, 0, s, mp, dm(M_LIN), e->part0, 1, e->pres);
        }
        if (fac) {
            hipLaunchKernelGGL(k_b_zero_masked, gN, bT, 0, s, n, e->ilu_err, dm(M_FAC));
            hipLaunchKernelGGL(k_b_zero_masked, gN, bT, 0, s, n, e->sweep_err, dm(M_FAC));
            if (e->wl_py) {
                // the factorizations in one tile-wavefront launch per precision, into the
                // layouts (each batch reads its members' flags in act / act64)
                if (fac32) {
                    ENS_TRY(e, hipMemcpyAsync(e->act, dm(M_FAC32), n * sizeof(int), hipMemcpyDeviceToDevice, s));
                    wl_batch_factor(e->wl_py, s);
                }
                if (fac64) {
                    ENS_TRY(e, hipMemcpyAsync(e->act64, dm(M_FAC64), n * sizeof(int), hipMemcpyDeviceToDevice, s));
                    wl_batch_factor(e->wl_py64, s);
                }
            } else {
                for (int l = 0; l < e->lmax; ++l)
                    hipLaunchKernelGGL(k_b_factor, dim3(e->nblkL), bT, 0, s, e->dmem, mp, dm(M_FAC), l, e->ilu_err);
                hipLaunchKernelGGL(k_b_round, gC, bT, 0, s, e->dmem, mp, dm(M_FAC));
            }
            hipLaunchKernelGGL(k_b_init, gC, bT, 0, s, e->dmem, mp, dm(M_FAC), e->part0, e->part1);
            hipLaunchKernelGGL(k_b_ks_init, dim3(n), bT, 0, s, e->dmem, mp, dm(M_FAC), e->ilu_err,
                               e->sweep_err, e->ks, e->part0, e->part1);
            // members whose sweeps the host makes fail in this solve (a test of the
            // failure path): the flag is set after the factorization, before the first
            // sweep
            if (forced) hipLaunchKernelGGL(k_b_set, gN, bT, 0, s, n, e->sweep_err, dm(M_FORCE));
        }
        ENS_TRY(e, hipGetLastError());
        if (solving) {
            std::vector<int> set;
            int blocks = 0;
            for (int m = 0; m < n; ++m)
                if (P[m].phase == SOLVING) {
                    set.push_back(m);
                    blocks += e->nblk_of[m];
                }
            const bool compact =
                e->wl_py && (e->compact == 2 || (e->compact == 1 && 2 * blocks <= e->nblkC));
            if (compact)
                if (int r = compact_maps(e, set)) return r;
            for (int q = 0; q < e->replay; ++q) {
                if (compact)
                    enqueue_iteration(e, true);
                else if (e->use_graph)
                    ENS_TRY(e, hipGraphLaunch(e->iter_graph, s));
                else
                    enqueue_iteration(e);
            }
        }
        if (lin) ENS_TRY(e, hipMemcpyAsync(e->h_res, e->pres, n * sizeof(double), hipMemcpyDeviceToHost, s));
        if (upd)
            ENS_TRY(e, hipMemcpyAsync(e->h_res_upd, e->pres_upd, n * sizeof(double), hipMemcpyDeviceToHost, s));
        if (solving) {
            ENS_TRY(e, hipMemcpyAsync(e->h_ks, e->ks, n * sizeof(KrylovState), hipMemcpyDeviceToHost, s));
            ENS_TRY(e, hipMemcpyAsync(e->h_serr, e->sweep_err, n * sizeof(int), hipMemcpyDeviceToHost, s));
        }
        ENS_TRY(e, hipStreamSynchronize(s));
        // ------------------------------------------------ the round's results, per member
        std::fill(mask, mask + kMaskSlots * n, 0);
        for (int m = 0; m < n; ++m) {
            Prog &p = P[m];
            const Member &h = e->host[m];
            switch (p.phase) {
                case UPD_PEND: {
                    const bool lin_queued = p.lin;
                    p.lin = false;
                    if (!std::isfinite(e->h_res_upd[m])) {
                        p.rc = EWG_NUMERICAL_ISSUE;  // "Non-finite update!"
                        p.phase = NDONE;
                        break;
                    }
                    ++p.it;
                    if (!lin_queued) {
                        p.phase = HEAD;
                        break;
                    }
                    p.lerr = p.err;  // the loop head went on: beginIteration_
                    [[fallthrough]];
                }
                case LIN_PEND:
                    e->outcome[m].newton++;
                    p.lerr = p.err;  // preSolve_
                    p.err = e->h_res[m];
                    if (p.err > h.newton_max_error) {
                        p.rc = EWG_NUMERICAL_ISSUE;
                        p.phase = NDONE;
                    } else if (!proceed_h(p.it, p.err, p.lerr, h)) {
                        ++p.it;
                        p.phase = HEAD;
                    } else {
                        p.phase = NEED_FAC;
                    }
                    break;
                case SOLVING: {
                    const KrylovState &k = e->h_ks[m];
                    if (k.status == KS_RUN) break;
                    e->outcome[m].linear += k.iterations;
                    if (e->h_serr[m]) mark(M_EDGE, m);
                    if (k.status == KS_CONV) {
                        e->outcome[m].solves++;
                        p.phase = NEED_UPD;
                    } else if (e->retry_ok[m] && !e->use64[m]) {
                        // the FP32 factors failed: this solve again, and the member's later
                        // ones, with FP64 factors
                        e->use64[m] = 1;
                        e->outcome[m].retries++;
                        if (e->wl_py) {
                            e->regraph |= !e->any64;
                            e->any64 = true;
                        } else {
                            flip[m] = 1;
                        }
                        p.phase = NEED_FAC;
                    } else {
                        e->outcome[m].solves++;
                        p.rc = k.status == KS_BREAK ? EWG_NUMERICAL_ISSUE : EWG_NOT_CONVERGED;
                        p.phase = NDONE;
                    }
                    break;
                }
                default:
                    break;
            }
            advance(m);
        }
    }
    if (!commit) {
        // the trial's members back to their start state (the step's start, kept in xs) and
        // their preconditioner precisions
        std::copy(took.begin(), took.end(), mask);
        ENS_TRY(e, hipMemcpyAsync(e->dmask, mask, n * sizeof(int), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_b_copy, dim3(e->nblkC), dim3(BT), 0, s, e->dmem, mp, dm(0), (int)V_X0, (int)V_XS);
        hipLaunchKernelGGL(k_b_copy, dim3(e->nblkC), dim3(BT), 0, s, e->dmem, mp, dm(0), (int)V_X1, (int)V_XS);
        for (int m = 0; m < n; ++m) {
            if (e->use64[m] == use64_in[m]) continue;
            e->use64[m] = use64_in[m];
            if (!e->wl_py)
                ENS_TRY(e, hipMemcpyAsync(&e->dmem[m].fp32, &e->h_fp32[use64_in[m] ? 0 : 1],
                                          sizeof(int), hipMemcpyHostToDevice, s));
        }
        e->regraph |= e->any64 != any64_in;
        e->any64 = any64_in;
        ENS_TRY(e, hipGetLastError());
        ENS_TRY(e, hipStreamSynchronize(s));
        std::fill(mask, mask + kMaskSlots * n, 0);
    }
    return EWG_OK;
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
        // the batched and persistent kernels linearize by automatic differentiation, and
        // their operator products read the FP64 Jacobian
        if (c->prm.local_linearizer != 0 || c->prm.operator_fp32 != 0) return EWG_INVALID_ARGUMENT;
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
        m.max_div = std::max(0, std::min(c->prm.max_time_step_divisions, 16));
    }
    auto *e = new ewg_ensemble();
    e->members.assign(members, members + n);
    e->xs.assign(n, nullptr);
    for (int i = 0; i < n; ++i) {
        if (hipMalloc(reinterpret_cast<void **>(&e->xs[i]), 2 * (size_t)host[i].lin.g.N * sizeof(double)) !=
            hipSuccess) {
            ewg_ensemble_destroy(e);
            return EWG_DEVICE_ERROR;
        }
        host[i].xs = e->xs[i];
    }
    e->host = host;
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
    if (batched_setup(e) != EWG_OK) return fail(hipErrorUnknown);
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
    if (e->iter_graph) (void)hipGraphExecDestroy(e->iter_graph);
    wl_batch_free(e->wl_py);
    wl_batch_free(e->wl_rz);
    wl_batch_free(e->wl_py64);
    wl_batch_free(e->wl_rz64);
    void *bufs[] = {e->cmem, e->coff, e->lmem, e->loff, e->act, e->niter, e->ilu_err, e->ks,
                    e->part0, e->part1, e->part2, e->pres, e->partAny, e->dts, e->sweep_err,
                    e->dmask, e->pres_upd, e->act64, e->cc_mem, e->cc_off, e->cc_t32, e->cc_t64};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    void *pinned[] = {e->h_mask, e->h_it, e->h_serr, e->h_dts, e->h_res, e->h_res_upd, e->h_ks,
                      e->h_fp32, e->h_cc_mem, e->h_cc_off, e->h_cc_t32, e->h_cc_t64};
    for (void *b : pinned)
        if (b) (void)hipHostFree(b);
    for (double *b : e->xs)
        if (b) (void)hipFree(b);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char *ewg_ensemble_last_error(const ewg_ensemble *e) { return e ? e->err.c_str() : "null ensemble"; }

int ewg_ensemble_force_sweep_failure(ewg_ensemble *e, int32_t member, int32_t solves) {
    if (!e || member < 0 || member >= (int32_t)e->members.size() || solves < 0)
        return EWG_INVALID_ARGUMENT;
    if ((int)e->force_left.size() != (int)e->members.size())
        return set_err(e, EWG_INVALID_STATE, "ensemble not set up");
    e->force_left[member] = solves;
    return EWG_OK;
}

int ewg_ensemble_tune_modes(ewg_ensemble *e, double dt, int32_t modes, ewg_ensemble_tuning *out) {
    if (!e || !(dt > 0.0) || (modes & ~(EWG_ENSEMBLE_TUNE_BATCHED | EWG_ENSEMBLE_TUNE_PERSISTENT)) ||
        modes == 0)
        return EWG_INVALID_ARGUMENT;
    const int n = static_cast<int>(e->members.size());
    ewg_ensemble_tuning t{};
    double best = 1e300;
    hipEvent_t e0, e1;
    ENS_TRY(e, hipEventCreate(&e0));
    ENS_TRY(e, hipEventCreate(&e1));
    const std::vector<Outcome> saved = e->outcome;
    const int keep_replay = e->replay;
    struct Cand {
        int mode, threads, largest_first, replay;
    };
    const Cand cand[] = {{EWG_ENSEMBLE_BATCHED, 0, 1, 4},      {EWG_ENSEMBLE_BATCHED, 0, 1, 8},
                         {EWG_ENSEMBLE_BATCHED, 0, 1, 16},     {EWG_ENSEMBLE_BATCHED, 0, 1, 32},
                         {EWG_ENSEMBLE_PERSISTENT, 256, 1, 0}, {EWG_ENSEMBLE_PERSISTENT, 512, 1, 0},
                         {EWG_ENSEMBLE_PERSISTENT, 1024, 1, 0}, {EWG_ENSEMBLE_PERSISTENT, 512, 0, 0}};
    static_assert(sizeof(cand) / sizeof(cand[0]) <= EWG_ENSEMBLE_MAX_CANDIDATES, "candidates");
    if (modes & EWG_ENSEMBLE_TUNE_BATCHED) {
        // an untimed trial first: the first batched step also captures the iteration graph
        if (int rc = batched_run(e, dt, 1, false)) return rc;
        e->outcome = saved;
    }
    for (const Cand &c : cand) {
        const int k = t.ncandidates++;
        t.cand_mode[k] = c.mode;
        t.cand_threads[k] = c.threads;
        t.cand_largest_first[k] = c.largest_first;
        t.cand_replay[k] = c.replay;
        t.cand_ms[k] = -1.0;
        const bool batched = c.mode == EWG_ENSEMBLE_BATCHED;
        if (!(modes & (batched ? EWG_ENSEMBLE_TUNE_BATCHED : EWG_ENSEMBLE_TUNE_PERSISTENT))) continue;
        bool ok = true;
        double ms = 0.0;
        if (batched) {
            e->replay = c.replay;
            const double t0 = now_s();
            const int rc = batched_run(e, dt, 1, false);
            e->replay = keep_replay;
            if (rc) return rc;
            ms = 1e3 * (now_s() - t0);
            for (int i = 0; i < n; ++i) ok &= e->outcome[i].status == 0 || saved[i].status != 0;
            e->outcome = saved;
        } else {
            if (int rc = upload_order(e, c.largest_first != 0)) return rc;
            ENS_TRY(e, hipMemcpyAsync(e->dout, saved.data(), n * sizeof(Outcome),
                                      hipMemcpyHostToDevice, e->stream));
            ENS_TRY(e, hipEventRecord(e0, e->stream));
            if (int rc = launch(e, c.threads, 1, dt, 0)) return rc;
            ENS_TRY(e, hipEventRecord(e1, e->stream));
            std::vector<Outcome> trial(n);
            ENS_TRY(e, hipMemcpyAsync(trial.data(), e->dout, n * sizeof(Outcome),
                                      hipMemcpyDeviceToHost, e->stream));
            ENS_TRY(e, hipStreamSynchronize(e->stream));
            float f = 0.0f;
            ENS_TRY(e, hipEventElapsedTime(&f, e0, e1));
            ms = f;
            for (int i = 0; i < n; ++i) ok &= trial[i].status == 0 || saved[i].status != 0;
        }
        t.cand_ms[k] = ok ? ms : -1.0;
        if (ok && ms < best) {
            best = ms;
            t.mode = c.mode;
            t.threads = c.threads;
            t.largest_first = c.largest_first;
            t.replay = c.replay;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ENS_TRY(e, hipMemcpy(e->dout, saved.data(), n * sizeof(Outcome), hipMemcpyHostToDevice));
    if (best >= 1e300) return set_err(e, EWG_NUMERICAL_ISSUE, "every candidate failed");
    e->mode = t.mode;
    if (t.mode == EWG_ENSEMBLE_PERSISTENT) {
        e->threads = t.threads;
        e->largest_first = t.largest_first != 0;
    } else {
        e->replay = t.replay;
    }
    if (out) *out = t;
    return EWG_OK;
}

int ewg_ensemble_tune(ewg_ensemble *e, double dt, ewg_ensemble_tuning *out) {
    return ewg_ensemble_tune_modes(e, dt, EWG_ENSEMBLE_TUNE_BATCHED | EWG_ENSEMBLE_TUNE_PERSISTENT,
                                   out);
}

int ewg_ensemble_run(ewg_ensemble *e, double dt, int32_t steps, double *seconds) {
    if (!e || !(dt > 0.0) || steps < 1) return EWG_INVALID_ARGUMENT;
    const int n = static_cast<int>(e->members.size());
    if (e->mode == EWG_ENSEMBLE_BATCHED) {
        ENS_TRY(e, hipStreamSynchronize(e->stream));
        const double t0 = now_s();
        if (int rc = batched_run(e, dt, steps, true)) return rc;
        if (seconds) *seconds = now_s() - t0;
        ENS_TRY(e, hipMemcpy(e->dout, e->outcome.data(), n * sizeof(Outcome), hipMemcpyHostToDevice));
        return EWG_OK;
    }
    ENS_TRY(e, hipMemcpy(e->dout, e->outcome.data(), n * sizeof(Outcome), hipMemcpyHostToDevice));
    if (int rc = upload_order(e, e->largest_first)) return rc;
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

/* choose the execution mode without tuning (EWG_ENSEMBLE_*; threads for the persistent
   mode: 256, 512 or 1024) */
int ewg_ensemble_set_mode(ewg_ensemble *e, int32_t mode, int32_t threads) {
    if (!e) return EWG_INVALID_ARGUMENT;
    if (mode == EWG_ENSEMBLE_BATCHED) {
        e->mode = mode;
        return EWG_OK;
    }
    if (mode != EWG_ENSEMBLE_PERSISTENT || (threads != 256 && threads != 512 && threads != 1024))
        return set_err(e, EWG_INVALID_ARGUMENT, "mode / threads");
    e->mode = mode;
    e->threads = threads;
    return EWG_OK;
}

int ewg_ensemble_set_graph(ewg_ensemble *e, int32_t on) {
    if (!e || (on != 0 && on != 1)) return EWG_INVALID_ARGUMENT;
    e->use_graph = on != 0;
    return EWG_OK;
}

int ewg_ensemble_set_compact(ewg_ensemble *e, int32_t on) {
    if (!e || on < 0 || on > 2) return EWG_INVALID_ARGUMENT;
    e->compact = on;
    return EWG_OK;
}

int ewg_ensemble_set_replay(ewg_ensemble *e, int32_t depth) {
    if (!e) return EWG_INVALID_ARGUMENT;
    if (depth < 1 || depth > 64) return set_err(e, EWG_INVALID_ARGUMENT, "replay depth must be 1 to 64");
    e->replay = depth;
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
    out->step_divisions = o.divisions;
    out->last_error = o.last_error;
    out->precond_retries = static_cast<int32_t>(o.retries);
    return EWG_OK;
}

}  // extern "C"
