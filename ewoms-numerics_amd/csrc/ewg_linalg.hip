// ewg_linalg.hip -- block ILU(0), SpMV and the BiCGStab vector/scalar kernels.
//
// ILU(0) in natural (lexicographic) order on the 7-point (5-point in 2D) pattern, as
// dune-istl's blockILU0Decomposition computes it: L_ij = A_ij inv(U_jj), A_ik -= L_ij A_jk
// on the pattern, pivots stored inverted. For nx, ny >= 3 the only pattern entry two
// lower/upper offsets can reach is the diagonal, so U_ij = A_ij (j > i) and
//   D_i = A_ii - sum_{j in {i-nxny, i-nx, i-1}} L_ij A_ji   (ascending j).
// Row i depends on rows i-1, i-nx, i-nxny only, so all cells on the hyperplane
// x + y + z = l are independent: factorization and both triangular sweeps run one
// hyperplane per launch (exact natural-order ILU(0); the summation order inside each row
// is the sequential one).
#include "ewg_cell_ops.h"

namespace ewg {

namespace {

using namespace cell;

template <class T>
__global__ void __launch_bounds__(256)
k_ilu_factor_level(Grid g, int level, const double *A, T *L, T *Dinv, T *U, int *err,
                   const int *gate) {
    if (gate && *gate == 0) return;
    int ijk[3];
    int i = level_cell(g, level, blockIdx.x * blockDim.x + threadIdx.x, ijk);
    if (i < 0) return;
    if (!ilu_factor_cell<T>(g, i, ijk, A, L, Dinv, U)) atomicExch(err, 1);
}

// forward sweep: w_i = d_i - sum_j L_ij w_j
template <class T>
__global__ void __launch_bounds__(256)
k_ilu_fwd_level(Grid g, int level, const T *L, const double *d, T *w, const int *status) {
    if (status && *status != KS_RUNNING) return;
    int ijk[3];
    int i = level_cell(g, level, blockIdx.x * blockDim.x + threadIdx.x, ijk);
    if (i < 0) return;
    ilu_fwd_cell<T>(g, i, ijk, L, d, w);
}

// backward sweep: w_i = inv(D_i) (w_i - sum_k U_ik w_k), out = (double) w
template <class T, class UT>
__global__ void __launch_bounds__(256)
k_ilu_bwd_level(Grid g, int level, const UT *U, int uslot0, const T *Dinv, T *w, double *out,
                const int *status) {
    if (status && *status != KS_RUNNING) return;
    int ijk[3];
    int i = level_cell(g, level, blockIdx.x * blockDim.x + threadIdx.x, ijk);
    if (i < 0) return;
    ilu_bwd_cell<T, UT>(g, i, ijk, U, uslot0, Dinv, w, out);
}

inline int level_threads(const Grid &g) { return g.dim == 3 ? g.ny * g.nz : g.ny; }

// ----------------------------------------------------------------------------------
// reductions
// ----------------------------------------------------------------------------------
__device__ __forceinline__ double warp_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ __forceinline__ double warp_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
    return v;
}
// block of kRedThreads (4 wavefronts); result valid in thread 0
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sh[kRedThreads / 64];
    v = warp_sum(v);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kRedThreads / 64; ++w) r += sh[w];
    return r;
}
__device__ __forceinline__ double block_max(double v) {
    __shared__ double shm[kRedThreads / 64];
    v = warp_max(v);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) shm[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kRedThreads / 64; ++w) r = fmax(r, shm[w]);
    return r;
}
// sum / max of kRedBlocks partials by one block (fixed order: deterministic)
__device__ __forceinline__ double reduce_partials_sum(const double *part) {
    double v = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += kRedThreads) v += part[i];
    return block_sum(v);
}
__device__ __forceinline__ double reduce_partials_max(const double *part) {
    double v = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += kRedThreads) v = fmax(v, part[i]);
    return block_max(v);
}

// SpMV over the present slots in ascending column order (BCRSMatrix::mv semantics); TB:
// the type the Jacobian's blocks are stored in
template <class TB>
__global__ void __launch_bounds__(256)
k_spmv(Grid g, const TB *A, const double *x, double *y, const double *a, double *partA,
       const double *b, double *partB, const int *status) {
    if (status && *status != KS_RUNNING) return;
    double accA = 0.0, accB = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.N; i += gridDim.x * blockDim.x) {
        const double2 yv = spmv_cell(g, i, A, x);
        const double s0 = yv.x, s1 = yv.y;
        reinterpret_cast<double2 *>(y)[i] = make_double2(s0, s1);
        if (partA) {
            double2 av = reinterpret_cast<const double2 *>(a)[i];
            accA += av.x * s0 + av.y * s1;
        }
        if (partB) {
            double2 bv = reinterpret_cast<const double2 *>(b)[i];
            accB += bv.x * s0 + bv.y * s1;
        }
    }
    if (partA) {
        double t = block_sum(accA);
        if (threadIdx.x == 0) partA[blockIdx.x] = t;
    }
    if (partB) {
        double t = block_sum(accB);
        if (threadIdx.x == 0) partB[blockIdx.x] = t;
    }
}

// every 2x2 block of the Jacobian rounded to nearest FP32 (the operator of the
// mixed-precision BiCGStab); gate: nothing while *gate == 0
__global__ void __launch_bounds__(256)
k_round_operator(size_t nblk, const double2 *A, float4 *A32, const int *gate) {
    if (gate && *gate == 0) return;
    for (size_t b = blockIdx.x * (size_t)blockDim.x + threadIdx.x; b < nblk;
         b += (size_t)gridDim.x * blockDim.x) {
        const double2 u = A[2 * b], w = A[2 * b + 1];
        A32[b] = make_float4(static_cast<float>(u.x), static_cast<float>(u.y),
                             static_cast<float>(w.x), static_cast<float>(w.y));
    }
}

// ----------------------------------------------------------------------------------
// BiCGStab (bicgstabsolver.hh:116-326) with CombinedCriterion (combinedcriterion.hh)
// ----------------------------------------------------------------------------------
constexpr double kBreakdownEps = 2.2250738585072014e-308 * 1e10;  // DBL_MIN * 1e10

__global__ void __launch_bounds__(kRedThreads)
k_bicg_init(int n, const double *b, double *r, double *x, double *p, double *v, double *partMax,
            double *partDot, const int *gate) {
    if (gate && *gate == 0) return;
    double m = 0.0, d = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double bi = b[i];
        r[i] = bi;
        x[i] = 0.0;
        p[i] = 0.0;
        v[i] = 0.0;
        m = fmax(m, fabs(bi));
        d += bi * bi;
    }
    double tm = block_max(m);
    double td = block_sum(d);
    if (threadIdx.x == 0) {
        partMax[blockIdx.x] = tm;
        partDot[blockIdx.x] = td;
    }
}

__global__ void __launch_bounds__(kRedThreads)
k_ks_init(KrylovState *ks, const double *partMax, const double *partDot, double rel_tol,
          double abs_tol, double max_resid, const int *gate) {
    if (gate && *gate == 0) {
        if (threadIdx.x == 0) {
            ks->iterations = 0;
            ks->status = KS_IDLE;
        }
        return;
    }
    double m = reduce_partials_max(partMax);
    __syncthreads();
    double d = reduce_partials_sum(partDot);
    if (threadIdx.x == 0) {
        ks->rel_tol = rel_tol;
        ks->abs_tol = abs_tol;
        ks->max_resid = max_resid;
        double e = fmax(m, 2.2250738585072014e-308 * 1e10);
        ks->resid_err = e;
        ks->init_err = e;
        ks->last_err = e;
        ks->stagnates = 0;
        ks->rho = 1.0;
        ks->alpha = 1.0;
        ks->omega = 1.0;
        ks->rho_i = d;
        ks->iterations = 0;
        double red = e / fmax(1e-20, e);
        ks->status = (red <= rel_tol || e <= abs_tol) ? KS_CONVERGED : KS_RUNNING;
    }
}

__global__ void __launch_bounds__(kRedThreads) k_ks_beta(KrylovState *ks, const double *partRho) {
    if (ks->status != KS_RUNNING) return;
    double rho_i = ks->iterations == 0 ? ks->rho_i : reduce_partials_sum(partRho);
    if (threadIdx.x == 0) {
        if (fabs(ks->rho) <= kBreakdownEps || fabs(ks->omega) <= kBreakdownEps) {
            ks->status = KS_BREAKDOWN;
            return;
        }
        ks->beta = (rho_i / ks->rho) * (ks->alpha / ks->omega);
        ks->rho = rho_i;
        ks->rho_i = rho_i;
    }
}

__global__ void __launch_bounds__(kRedThreads)
k_p_update(int n, const KrylovState *ks, const double *r, const double *v, double *p,
           unsigned *tickets) {
    // the tile tickets of the preconditioner apply that follows (its two sweeps)
    if (tickets && blockIdx.x == 0 && threadIdx.x == 0) {
        tickets[0] = 0u;
        tickets[1] = 0u;
    }
    if (ks->status != KS_RUNNING) return;
    const double omega = ks->omega, beta = ks->beta;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double tmp = v[i];
        tmp *= omega;
        tmp -= p[i];
        tmp *= -beta;
        double pi = r[i];
        pi += tmp;
        p[i] = pi;
    }
}

__global__ void __launch_bounds__(kRedThreads) k_ks_alpha(KrylovState *ks, const double *partDen) {
    if (ks->status != KS_RUNNING) return;
    double den = reduce_partials_sum(partDen);
    if (threadIdx.x == 0) {
        ks->denom = den;
        if (fabs(den) <= kBreakdownEps) {
            ks->status = KS_BREAKDOWN;
            return;
        }
        double alpha = ks->rho / den;
        ks->alpha = alpha;
        if (fabs(alpha) <= kBreakdownEps) ks->status = KS_BREAKDOWN;
    }
}

__global__ void __launch_bounds__(kRedThreads)
k_hs_update(int n, const KrylovState *ks, const double *y, const double *v, double *x, double *r,
            double *partMax, int *anyNonzero, unsigned *tickets) {
    // the tile tickets of the preconditioner apply that follows (its two sweeps)
    if (tickets && blockIdx.x == 0 && threadIdx.x == 0) {
        tickets[0] = 0u;
        tickets[1] = 0u;
    }
    if (ks->status != KS_RUNNING) return;
    const double alpha = ks->alpha;
    double m = 0.0;
    bool nz = false;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double yi = y[i];
        double tmp = yi;
        tmp *= alpha;
        tmp += x[i];
        x[i] = tmp;
        tmp = v[i];
        tmp *= alpha;
        double ri = r[i] - tmp;
        r[i] = ri;
        m = fmax(m, fabs(ri));
        nz |= (yi != 0.0);
    }
    double tm = block_max(m);
    int bnz = __syncthreads_or(nz);
    if (threadIdx.x == 0) {
        partMax[blockIdx.x] = tm;
        anyNonzero[blockIdx.x] = bnz;
    }
}

__device__ __forceinline__ void criterion(KrylovState *ks, double err, int stagnates) {
    ks->last_err = ks->resid_err;
    ks->resid_err = err;
    ks->stagnates = stagnates;
    double red = err / fmax(1e-20, ks->init_err);
    bool conv = red <= ks->rel_tol || err <= ks->abs_tol;
    bool failed = !conv && (stagnates || err > ks->max_resid);
    if (conv) ks->status = KS_CONVERGED;
    else if (failed) ks->status = KS_FAILED;
}

__device__ __forceinline__ int reduce_partials_any(const int *part) {
    int v = 0;
    for (int i = threadIdx.x; i < kRedBlocks; i += kRedThreads) v |= part[i];
    return __syncthreads_or(v);
}

__global__ void __launch_bounds__(kRedThreads)
k_ks_crit1(KrylovState *ks, const double *partMax, const int *anyNonzero) {
    if (ks->status != KS_RUNNING) return;
    double m = reduce_partials_max(partMax);
    int any = reduce_partials_any(anyNonzero);
    if (threadIdx.x == 0) criterion(ks, m, any == 0);
}

__global__ void __launch_bounds__(kRedThreads)
k_ks_omega(KrylovState *ks, const double *partTT, const double *partTS) {
    if (ks->status != KS_RUNNING) return;
    double tt = reduce_partials_sum(partTT);
    __syncthreads();
    double ts = reduce_partials_sum(partTS);
    if (threadIdx.x == 0) {
        if (fabs(tt) <= kBreakdownEps) {
            ks->status = KS_BREAKDOWN;
            return;
        }
        double omega = ts / tt;
        ks->omega = omega;
        if (fabs(omega) <= kBreakdownEps) ks->status = KS_BREAKDOWN;
    }
}

__global__ void __launch_bounds__(kRedThreads)
k_xr_update(int n, const KrylovState *ks, const double *z, const double *t, double *x, double *r,
            const double *r0hat, double *partRho, int *anyNonzero) {
    if (ks->status != KS_RUNNING) return;
    const double omega = ks->omega;
    double d = 0.0;
    bool nz = false;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double zi = z[i];
        x[i] += omega * zi;
        double ri = r[i] + -omega * t[i];
        r[i] = ri;
        d += r0hat[i] * ri;
        nz |= (zi != 0.0);
    }
    double td = block_sum(d);
    int bnz = __syncthreads_or(nz);
    if (threadIdx.x == 0) {
        partRho[blockIdx.x] = td;
        anyNonzero[blockIdx.x] = bnz;
    }
}

__global__ void __launch_bounds__(kRedThreads) k_ks_crit2(KrylovState *ks, const int *anyNonzero) {
    if (ks->status != KS_RUNNING) return;
    int any = reduce_partials_any(anyNonzero);
    if (threadIdx.x == 0) {
        // the residual checked here is still s (the object r is updated afterwards)
        criterion(ks, ks->resid_err, any == 0);
        if (ks->status == KS_RUNNING) ks->iterations += 1;
    }
}

// Newton helpers
// plain streaming copy, 32 bytes per lane and iteration: the bandwidth reference
// dst = src over n16 16-byte words: the bandwidth reference. Workgroup b streams the
// contiguous chunk [b chunk, (b + 1) chunk) with four coalesced 16-byte loads in flight
// per thread (contiguous chunks per workgroup measured 5.4 TB/s on 1 GiB against 4.5 TB/s
// for a grid-stride loop).
__global__ void __launch_bounds__(256)
k_copy(size_t n16, size_t chunk, const uint4 *__restrict__ src, uint4 *__restrict__ dst) {
    const size_t b0 = blockIdx.x * chunk;
    const size_t b1 = b0 + chunk < n16 ? b0 + chunk : n16;
    size_t i = b0 + threadIdx.x;
    for (; i + 768 < b1; i += 1024) {
        const uint4 a = src[i], b = src[i + 256], c = src[i + 512], d = src[i + 768];
        dst[i] = a;
        dst[i + 256] = b;
        dst[i + 512] = c;
        dst[i + 768] = d;
    }
    for (; i < b1; i += 256) dst[i] = src[i];
}

__global__ void __launch_bounds__(kRedThreads) k_max_abs(int n, const double *x, double *part) {
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        m = fmax(m, fabs(x[i]));
    double t = block_max(m);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}
__global__ void __launch_bounds__(kRedThreads) k_reduce_max(const double *part, double *out) {
    double m = reduce_partials_max(part);
    if (threadIdx.x == 0) *out = m;
}
__global__ void __launch_bounds__(kRedThreads) k_reduce_sum(const double *part, double *out) {
    double m = reduce_partials_sum(part);
    if (threadIdx.x == 0) *out = m;
}
__global__ void __launch_bounds__(kRedThreads)
k_newton_update(int n, const double *cur, const double *dx, double *x, double *part) {
    double s = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double d = dx[i];
        double nv = cur[i];
        nv -= d;
        x[i] = nv;
        s += fabs(d);
    }
    double t = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ----------------------------------------------------------------------------------
// The Newton loop's decisions on the device (newtonmethod.hh:308-517): one thread of one
// block evaluates the rules with the host driver's arithmetic; the reductions are the
// ones of k_reduce_max / k_reduce_sum.
// ----------------------------------------------------------------------------------
__device__ __forceinline__ bool ns_proceed(const NewtonState &n) {  // newtonmethod.hh:866-884
    if (n.it < 1) return true;
    if (n.error <= n.tol) return false;
    if (n.it >= n.max_it) return n.error * 4.0 < n.last_error;
    return true;
}

// head of an iteration: the loop condition, then beginIteration_ (lastError_ = error_)
__global__ void k_newton_begin(NewtonState *ns) {
    NewtonState &n = *ns;
    n.solve = 0;
    if (n.status != NS_RUNNING) {
        n.active = 0;
        return;
    }
    if (!ns_proceed(n)) {
        n.active = 0;
        if (n.has_update && !isfinite(n.upd)) {  // the last update, checked at the loop's end
            n.status = NS_FAILED;
            n.code = NS_FAIL_UPDATE;
            n.it = 2 * n.target;
        } else {
            n.status = NS_DONE;
        }
        return;
    }
    n.last_error = n.error;
    n.active = 1;
}

// after the linearization: preSolve_ (lastError_ = error_, error_ = max |R|), the check of
// the previous update, NewtonMaxError, and proceed_ before the linear solve
__global__ void __launch_bounds__(kRedThreads)
k_newton_linearized(NewtonState *ns, const double *partMax) {
    if (!ns->active) return;
    const double e = reduce_partials_max(partMax);
    if (threadIdx.x != 0) return;
    NewtonState &n = *ns;
    n.last_error = n.error;
    if (n.has_update && !isfinite(n.upd)) {
        n.status = NS_FAILED;
        n.code = NS_FAIL_UPDATE;
        n.it = 2 * n.target;
        return;
    }
    n.has_update = 0;
    n.error = e;
    if (n.it == 0) n.first_error = e;
    if (n.error > n.max_error) {
        n.status = NS_FAILED;
        n.code = NS_FAIL_MAX_ERROR;
        n.it = 2 * n.target;
        return;
    }
    if (!ns_proceed(n)) {
        ++n.it;
        n.status = NS_DONE;
        return;
    }
    n.solve = 1;
}

// after an update x = xcur - dx: sum |dx| for the finiteness check, ++numIterations
__global__ void __launch_bounds__(kRedThreads)
k_newton_updated(NewtonState *ns, const double *partSum) {
    const double s = reduce_partials_sum(partSum);
    if (threadIdx.x != 0) return;
    ns->upd = s;
    ns->has_update = 1;
    ++ns->it;
}

__global__ void __launch_bounds__(256)
k_copy_gated(int n, const double *src, double *dst, const int *gate) {
    if (!*gate) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_flag_clear(int *flag) { *flag = 0; }

}  // namespace

// ------------------------------------------------------------------------------------
void launch_ilu_factor64(const Grid &g, const double *A, double *L, double *Dinv, int *err,
                         hipStream_t s, const int *gate) {
    int threads = level_threads(g), blocks = (threads + 255) / 256;
    for (int l = 0; l < g.nlevels; ++l)
        hipLaunchKernelGGL(k_ilu_factor_level<double>, dim3(blocks), dim3(256), 0, s, g, l, A, L,
                           Dinv, static_cast<double *>(nullptr), err, gate);
}
// FP32 copies of FP64 factors and of the Jacobian's upper blocks (rounded to nearest)
__global__ void k_round_factors(size_t nl, size_t nd, const double *L, const double *D,
                                const double *Aup, float *L32, float *D32, float *U32,
                                const int *gate) {
    if (gate && *gate == 0) return;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nl) {
        L32[e] = static_cast<float>(L[e]);
        U32[e] = static_cast<float>(Aup[e]);
    }
    if (e < nd) D32[e] = static_cast<float>(D[e]);
}
void launch_round_factors(const Grid &g, const double *L, const double *Dinv, const double *A,
                          float *L32, float *D32, float *U32, hipStream_t s, const int *gate) {
    const size_t nl = (size_t)g.dim * g.N * 4, nd = (size_t)g.N * 4;
    hipLaunchKernelGGL(k_round_factors, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, s, nl,
                       nd, L, Dinv, A + (size_t)(g.diag + 1) * g.N * 4, L32, D32, U32, gate);
}
void launch_ilu_apply64(const Grid &g, const double *L, const double *Dinv, const double *A,
                        const double *d, double *v, const int *status, hipStream_t s) {
    int threads = level_threads(g), blocks = (threads + 255) / 256;
    for (int l = 0; l < g.nlevels; ++l)
        hipLaunchKernelGGL(k_ilu_fwd_level<double>, dim3(blocks), dim3(256), 0, s, g, l, L, d, v,
                           status);
    for (int l = g.nlevels - 1; l >= 0; --l)
        hipLaunchKernelGGL((k_ilu_bwd_level<double, double>), dim3(blocks), dim3(256), 0, s, g, l,
                           A, g.diag + 1, Dinv, v, v, status);
}
void launch_ilu_apply32(const Grid &g, const float *L, const float *Dinv, const float *U,
                        const double *d, float *w, double *v, const int *status,
                        hipStream_t s) {
    int threads = level_threads(g), blocks = (threads + 255) / 256;
    for (int l = 0; l < g.nlevels; ++l)
        hipLaunchKernelGGL(k_ilu_fwd_level<float>, dim3(blocks), dim3(256), 0, s, g, l, L, d, w,
                           status);
    for (int l = g.nlevels - 1; l >= 0; --l)
        hipLaunchKernelGGL((k_ilu_bwd_level<float, float>), dim3(blocks), dim3(256), 0, s, g, l, U,
                           0, Dinv, w, v, status);
}

void launch_spmv(const Grid &g, const double *A, const double *x, double *y, const double *a,
                 double *partA, const double *b, double *partB, const int *status,
                 hipStream_t s) {
    hipLaunchKernelGGL(k_spmv<double>, dim3(kRedBlocks), dim3(kRedThreads), 0, s, g, A, x, y, a,
                       partA, b, partB, status);
}
void launch_spmv32(const Grid &g, const float *A32, const double *x, double *y, const double *a,
                   double *partA, const double *b, double *partB, const int *status,
                   hipStream_t s) {
    hipLaunchKernelGGL(k_spmv<float>, dim3(kRedBlocks), dim3(kRedThreads), 0, s, g, A32, x, y, a,
                       partA, b, partB, status);
}
void launch_round_operator(const Grid &g, const double *A, float *A32, hipStream_t s,
                           const int *gate) {
    const size_t nblk = (size_t)g.nslots * g.N;
    hipLaunchKernelGGL(k_round_operator, dim3(2048), dim3(256), 0, s, nblk,
                       reinterpret_cast<const double2 *>(A), reinterpret_cast<float4 *>(A32), gate);
}

void launch_bicg_init(int n, const double *b, double *r, double *x, double *p, double *v,
                      double *partMax, double *partDot, hipStream_t s, const int *gate) {
    hipLaunchKernelGGL(k_bicg_init, dim3(kRedBlocks), dim3(kRedThreads), 0, s, n, b, r, x, p, v,
                       partMax, partDot, gate);
}
void launch_ks_init(KrylovState *ks, const double *partMax, const double *partDot,
                    double rel_tol, double abs_tol, double max_resid, hipStream_t s,
                    const int *gate) {
    hipLaunchKernelGGL(k_ks_init, dim3(1), dim3(kRedThreads), 0, s, ks, partMax, partDot, rel_tol,
                       abs_tol, max_resid, gate);
}
void launch_ks_beta(KrylovState *ks, const double *partRho, hipStream_t s) {
    hipLaunchKernelGGL(k_ks_beta, dim3(1), dim3(kRedThreads), 0, s, ks, partRho);
}
void launch_p_update(int n, const KrylovState *ks, const double *r, const double *v, double *p,
                     hipStream_t s, unsigned *tickets) {
    hipLaunchKernelGGL(k_p_update, dim3(kRedBlocks), dim3(kRedThreads), 0, s, n, ks, r, v, p,
                       tickets);
}
void launch_ks_alpha(KrylovState *ks, const double *partDen, hipStream_t s) {
    hipLaunchKernelGGL(k_ks_alpha, dim3(1), dim3(kRedThreads), 0, s, ks, partDen);
}
void launch_hs_update(int n, const KrylovState *ks, const double *y, const double *v, double *x,
                      double *r, double *partMax, int *anyNonzero, hipStream_t s,
                      unsigned *tickets) {
    hipLaunchKernelGGL(k_hs_update, dim3(kRedBlocks), dim3(kRedThreads), 0, s, n, ks, y, v, x, r,
                       partMax, anyNonzero, tickets);
}
void launch_ks_crit1(KrylovState *ks, const double *partMax, const int *anyNonzero, hipStream_t s) {
    hipLaunchKernelGGL(k_ks_crit1, dim3(1), dim3(kRedThreads), 0, s, ks, partMax, anyNonzero);
}
void launch_ks_omega(KrylovState *ks, const double *partTT, const double *partTS, hipStream_t s) {
    hipLaunchKernelGGL(k_ks_omega, dim3(1), dim3(kRedThreads), 0, s, ks, partTT, partTS);
}
void launch_xr_update(int n, const KrylovState *ks, const double *z, const double *t, double *x,
                      double *r, const double *r0hat, double *partRho, int *anyNonzero,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_xr_update, dim3(kRedBlocks), dim3(kRedThreads), 0, s, n, ks, z, t, x, r,
                       r0hat, partRho, anyNonzero);
}
void launch_ks_crit2(KrylovState *ks, const int *anyNonzero, hipStream_t s) {
    hipLaunchKernelGGL(k_ks_crit2, dim3(1), dim3(kRedThreads), 0, s, ks, anyNonzero);
}
void launch_max_abs(int n, const double *x, double *part, hipStream_t s) {
    hipLaunchKernelGGL(k_max_abs, dim3(kRedBlocks), dim3(kRedThreads), 0, s, n, x, part);
}
void launch_reduce_max(const double *part, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(kRedThreads), 0, s, part, out);
}
void launch_copy(size_t n, const double *src, double *dst, hipStream_t s) {
    const size_t n16 = n / 2, blocks = 2048, chunk = (n16 + blocks - 1) / blocks;
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, s, n16, chunk,
                       reinterpret_cast<const uint4 *>(src), reinterpret_cast<uint4 *>(dst));
}

void launch_reduce_sum(const double *part, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kRedThreads), 0, s, part, out);
}
void launch_newton_update(int n, const double *cur, const double *dx, double *x, double *part,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_newton_update, dim3(kRedBlocks), dim3(kRedThreads), 0, s, n, cur, dx, x,
                       part);
}

void launch_newton_begin(NewtonState *ns, hipStream_t s) {
    hipLaunchKernelGGL(k_newton_begin, dim3(1), dim3(1), 0, s, ns);
}
void launch_newton_linearized(NewtonState *ns, const double *partMax, hipStream_t s) {
    hipLaunchKernelGGL(k_newton_linearized, dim3(1), dim3(kRedThreads), 0, s, ns, partMax);
}
void launch_newton_updated(NewtonState *ns, const double *partSum, hipStream_t s) {
    hipLaunchKernelGGL(k_newton_updated, dim3(1), dim3(kRedThreads), 0, s, ns, partSum);
}
void launch_copy_gated(int n, const double *src, double *dst, const int *gate, hipStream_t s) {
    hipLaunchKernelGGL(k_copy_gated, dim3(2048), dim3(256), 0, s, n, src, dst, gate);
}
void launch_flag_clear(int *flag, hipStream_t s) {
    hipLaunchKernelGGL(k_flag_clear, dim3(1), dim3(1), 0, s, flag);
}

}  // namespace ewg
