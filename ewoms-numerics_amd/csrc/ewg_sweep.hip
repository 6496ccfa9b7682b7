// ewg_sweep.hip -- the two ILU(0) triangular sweeps as persistent tile-wavefront launches,
// tuned for instruction count (the sweeps run twice per BiCGStab half-iteration).
//
// Same schedule as k_wavefront in ewg_wavefront.hip (tile columns through z, skewed
// local steps, tickets in anti-diagonal order, edge buffers with a sentinel), with:
//   - per-lane 32-bit byte offsets that advance by a constant per step; a lane whose
//     cell is outside the grid (z < 0, z >= nz, or outside the plane) lands out of the
//     buffer range, so the hardware drops its stores and zeroes its loads: no
//     predication of memory operations (this needs N * 32 bytes plus the skew margin
//     below 2^31, which the kWfMaxCells limit guarantees);
//   - no predication of the block products either: blocks of absent neighbours are
//     stored as zero, and the neighbour values that enter them are selected to zero;
//   - the in-tile exchange of step results through a double-buffered LDS array
//     (one store and two loads per step).
// The arithmetic per row is the one of the hyperplane kernels (ascending columns, no
// fused multiply-add), so both schedules agree to the last bit up to the sign of zeros.
#include "ewg_internal.h"
#include "ewg_wavefront.h"

namespace ewg {

namespace {

constexpr unsigned kSpinLimit = 1u << 22;
constexpr unsigned long long kSentinel = 0x7FF4DEAD7FF4DEADull;
constexpr unsigned kFar = 0x80000000u;  // offset base of lanes without a column

using u32x2 = __attribute__((ext_vector_type(2))) unsigned;
using u32x4 = __attribute__((ext_vector_type(4))) unsigned;
constexpr int kCoh = 16;  // sc1: cache policy of agent-scope atomics on this target

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p, size_t bytes) {
    unsigned n = bytes > 0x7FFFFFF0u ? 0x7FFFFFF0u : static_cast<unsigned>(bytes);
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, static_cast<int>(n),
                                             0x00020000);
}

template <class T>
struct P2 {
    T x, y;
};

// 2x2 block of T: load (16 or 32 bytes)
template <class T>
__device__ __forceinline__ void ldb(__amdgpu_buffer_rsrc_t r, unsigned off, T *o) {
    if constexpr (sizeof(T) == 8) {
        double2 u = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
        double2 v = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off + 16, 0, 0));
        o[0] = u.x; o[1] = u.y; o[2] = v.x; o[3] = v.y;
    } else {
        float4 u = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
        o[0] = u.x; o[1] = u.y; o[2] = u.z; o[3] = u.w;
    }
}
template <class T>
__device__ __forceinline__ P2<T> ldp(__amdgpu_buffer_rsrc_t r, unsigned off) {
    if constexpr (sizeof(T) == 8) {
        double2 u = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
        return P2<T>{u.x, u.y};
    } else {
        float2 u = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
        return P2<T>{u.x, u.y};
    }
}
template <class T>
__device__ __forceinline__ void stp(__amdgpu_buffer_rsrc_t r, unsigned off, P2<T> v) {
    if constexpr (sizeof(T) == 8)
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, make_double2(v.x, v.y)),
                                               r, off, 0, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, make_float2(v.x, v.y)), r,
                                              off, 0, 0);
}

// edge words of one P2<T> value: 1 (float) or 2 (double) 64-bit words, coherent access
template <class T>
struct EW {
    unsigned long long w[sizeof(T) == 8 ? 2 : 1];
};
template <class T>
__device__ __forceinline__ EW<T> lde(__amdgpu_buffer_rsrc_t r, unsigned off) {
    EW<T> e;
    if constexpr (sizeof(T) == 8) {
        u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, kCoh);
        e.w[0] = (unsigned long long)v.x | ((unsigned long long)v.y << 32);
        e.w[1] = (unsigned long long)v.z | ((unsigned long long)v.w << 32);
    } else {
        u32x2 v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, kCoh);
        e.w[0] = (unsigned long long)v.x | ((unsigned long long)v.y << 32);
    }
    return e;
}
template <class T>
__device__ __forceinline__ void ste(__amdgpu_buffer_rsrc_t r, unsigned off, const EW<T> &e) {
    if constexpr (sizeof(T) == 8) {
        u32x4 v;
        v.x = (unsigned)e.w[0]; v.y = (unsigned)(e.w[0] >> 32);
        v.z = (unsigned)e.w[1]; v.w = (unsigned)(e.w[1] >> 32);
        __builtin_amdgcn_raw_buffer_store_b128(v, r, off, 0, kCoh);
    } else {
        u32x2 v;
        v.x = (unsigned)e.w[0]; v.y = (unsigned)(e.w[0] >> 32);
        __builtin_amdgcn_raw_buffer_store_b64(v, r, off, 0, kCoh);
    }
}
template <class T>
__device__ __forceinline__ bool is_sentinel(const EW<T> &e) {
    bool s = e.w[0] == kSentinel;
    if constexpr (sizeof(T) == 8) s = s || e.w[1] == kSentinel;
    return s;
}
template <class T>
__device__ __forceinline__ P2<T> ew_val(const EW<T> &e) {
    P2<T> v;
    __builtin_memcpy(&v, e.w, sizeof(v));
    return v;
}
template <class T>
__device__ __forceinline__ EW<T> val_ew(const P2<T> &v) {
    EW<T> e;
    __builtin_memcpy(e.w, &v, sizeof(v));
    return e;
}

// r -= m v (m 2x2 row-major)
template <class T>
__device__ __forceinline__ void mmv(const T *m, P2<T> v, T &r0, T &r1) {
    r0 -= m[0] * v.x;
    r0 -= m[1] * v.y;
    r1 -= m[2] * v.x;
    r1 -= m[3] * v.y;
}

template <class T>
struct SweepArgs {
    int nx, ny, nz, nxny, N;
    // forward: rhs = d (FP64), blocks = L slots 0..2 (-z,-y,-x), w = result (T)
    // backward: rhs = w (T, forward result), blocks = U slots (+x,+y,+z), Dinv;
    //           result -> w (T) and out (FP64; may alias w when T is double)
    const double *d;
    const void *blk[3];
    unsigned blk_bytes;     // bytes per block of the backward sweep's U arrays
    const T *Dinv;
    T *w;
    double *out;
    const int *status;
    WfSync sync;
};

template <class T, bool FWD>
struct In {
    P2<T> rhs;
    T b[3][4];
    T di[4];
};

template <class T, bool FWD, int TX, int TY, int PD, int PE>
__global__ void __launch_bounds__(TX *TY) k_sweep(SweepArgs<T> a) {
    constexpr int NT = TX * TY;
    constexpr int NWAVE = NT / 64;
    static_assert(NT % 64 == 0, "whole wavefronts");
    static_assert(PD % PE == 0 && PE <= PD, "edge ring must divide the operand ring");
    __shared__ P2<T> xch[2][NT];
    __shared__ unsigned s_tk;
    if (a.status && *a.status != KS_RUNNING) return;
    const int tid = threadIdx.x;
    if (tid == 0) s_tk = atomicAdd(a.sync.ticket, 1u);
    __syncthreads();
    const unsigned tk = s_tk;
    if (tk >= (unsigned)a.sync.ntiles) return;
    const int2 tile = a.sync.order[tk];
    const int tx = tile.x, ty = tile.y;
    const int lx = tid % TX, ly = tid / TX;
    const int xs = tx * TX + lx, ys = ty * TY + ly;
    const bool inplane = xs < a.nx && ys < a.ny;
    const int nz = a.nz;
    const int tileId = ty * a.sync.ntx + tx;
    const bool consX = inplane && lx == 0 && tx > 0;
    const bool consY = inplane && ly == 0 && ty > 0;
    const bool prodX = inplane && lx == TX - 1 && tx + 1 < a.sync.ntx;
    const bool prodY = inplane && ly == TY - 1 && ty + 1 < a.sync.nty;

    // cell of (xs, ys, z) = c0 + z*dz in sweep coordinates
    int c0, dz;
    if (FWD) {
        c0 = xs + a.nx * ys;
        dz = a.nxny;
    } else {
        c0 = (a.nx - 1 - xs) + a.nx * (a.ny - 1 - ys) + a.nxny * (nz - 1);
        dz = -a.nxny;
    }
    // byte offsets at local step 0 (z = -lx-ly), advanced by a constant per step
    const int cs0 = c0 - (lx + ly) * dz;  // cell of local step 0
    const unsigned bB = FWD ? 4 * sizeof(T) : a.blk_bytes;  // block array stride per cell
    const unsigned vB = 2 * sizeof(T);                     // result vector stride per cell
    const unsigned diB = 4 * sizeof(T);
    // current-step offsets (stores) and prefetch offsets (loads, PD steps ahead)
    unsigned ovec = inplane ? (unsigned)(cs0 * (int)vB) : kFar;
    unsigned oout = inplane ? (unsigned)(cs0 * 16) : kFar;
    unsigned pblk = inplane ? (unsigned)(cs0 * (int)bB) : kFar;
    unsigned pdi = inplane ? (unsigned)(cs0 * (int)diB) : kFar;
    unsigned pvec = ovec, pd = oout;
    // lanes outside the plane stay at kFar (their offsets never move into range)
    const int dzl = inplane ? dz : 0;
    const unsigned sblk = (unsigned)(dzl * (int)bB), sdi = (unsigned)(dzl * (int)diB),
                   svec = (unsigned)(dzl * (int)vB), sd = (unsigned)(dzl * 16);

    const size_t N = (size_t)a.N;
    const __amdgpu_buffer_rsrc_t rb0 = rsrc(a.blk[0], N * bB), rb1 = rsrc(a.blk[1], N * bB),
                                 rb2 = rsrc(a.blk[2], N * bB);
    const __amdgpu_buffer_rsrc_t rdi = rsrc(a.Dinv, FWD ? 0 : N * diB);
    const __amdgpu_buffer_rsrc_t rw = rsrc(a.w, N * vB);
    const __amdgpu_buffer_rsrc_t rd = rsrc(a.d, FWD ? N * 16 : 0);
    const __amdgpu_buffer_rsrc_t rout = rsrc(a.out, a.out ? N * 16 : 0);
    const __amdgpu_buffer_rsrc_t rex = rsrc(a.sync.ex, a.sync.edge_words * 8);
    const __amdgpu_buffer_rsrc_t rey = rsrc(a.sync.ey, a.sync.edge_words * 8);
    constexpr unsigned zw = kWfMaxWords * 8;  // bytes per z of one edge lane
    const unsigned exIn = consX ? (unsigned)((((size_t)(tileId - 1) * TY + ly) * nz) * zw) : 0u;
    const unsigned eyIn = consY ? (unsigned)((((size_t)(tileId - a.sync.ntx) * TX + lx) * nz) * zw) : 0u;
    const unsigned exOut = prodX ? (unsigned)((((size_t)tileId * TY + ly) * nz) * zw) : 0u;
    const unsigned eyOut = prodY ? (unsigned)((((size_t)tileId * TX + lx) * nz) * zw) : 0u;
    // edge-word load offsets of local step 0, advanced by zw per step: a consumer lane
    // reads its column's word of z = step - lx - ly (steps outside 0..nz-1 read words
    // that are never used), every other lane stays out of range
    unsigned lex = consX ? exIn - (unsigned)(lx + ly) * zw : kFar;
    unsigned ley = consY ? eyIn - (unsigned)(lx + ly) * zw : kFar;
    const unsigned sex = consX ? zw : 0u, sey = consY ? zw : 0u;
    const int nsteps = (TX - 1) + (TY - 1) + nz;

    In<T, FWD> ring[PD];
    EW<T> ex[PE], ey[PE];

    // operands of the step the p* offsets point at; steps past the end read out of
    // range and are never used
    auto load = [&](In<T, FWD> &in) {
        if (FWD) {
            P2<double> dv = ldp<double>(rd, pd);
            in.rhs = P2<T>{static_cast<T>(dv.x), static_cast<T>(dv.y)};
        } else {
            in.rhs = ldp<T>(rw, pvec);
        }
        ldb<T>(rb0, pblk, in.b[0]);
        ldb<T>(rb1, pblk, in.b[1]);
        ldb<T>(rb2, pblk, in.b[2]);
        if (!FWD) ldb<T>(rdi, pdi, in.di);
        pblk += sblk;
        pvec += svec;
        pdi += sdi;
        pd += sd;
    };
    // cross-tile words of the step the le* offsets point at. They are loaded only PE
    // steps ahead: a consumer runs just behind its producer, and a word read earlier
    // would still hold the sentinel and cost a round trip to poll.
    auto load_edge = [&](EW<T> &e1, EW<T> &e2) {
        e1 = lde<T>(rex, lex);
        e2 = lde<T>(rey, ley);
        lex += sex;
        ley += sey;
    };

    // The scheduling barriers keep the prefetches in step order: the wait for a slot
    // then counts every later step's memory operations, on the first pass through the
    // loop as on the others (a reordered prologue makes the waitcnt insertion assume
    // the operands of the first steps are young and wait for nearly all loads).
#pragma unroll
    for (int k = 0; k < PD; ++k) {
        load(ring[k]);
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int k = 0; k < PE; ++k) {
        load_edge(ex[k], ey[k]);
        __builtin_amdgcn_sched_barrier(0);
    }

    P2<T> prev = P2<T>{T(0), T(0)};
    EW<T> sent;
    sent.w[0] = kSentinel;
    if constexpr (sizeof(T) == 8) sent.w[1] = kSentinel;
    const bool hasL = lx > 0, hasD = ly > 0;
    const int ixl = hasL ? tid - 1 : tid, ixd = hasD ? tid - TX : tid;

    for (int tau0 = 0; tau0 < nsteps; tau0 += PD) {
#pragma unroll
        for (int k = 0; k < PD; ++k) {
            const int tau = tau0 + k;
            const int z = tau - lx - ly;
            const bool active = inplane && z >= 0 && z < nz;
            const bool needX = active && consX, needY = active && consY;
            const unsigned offX = needX ? exIn + (unsigned)z * zw : kFar;
            const unsigned offY = needY ? eyIn + (unsigned)z * zw : kFar;
            // cross-tile values of this step; poll while not yet published
            EW<T> cx = ex[k % PE], cy = ey[k % PE];
            bool wait = (needX && is_sentinel(cx)) || (needY && is_sentinel(cy));
            if (__builtin_amdgcn_ballot_w64(wait) != 0ull) {
                unsigned spins = 0;
                while (__builtin_amdgcn_ballot_w64(wait) != 0ull && spins < kSpinLimit) {
                    __builtin_amdgcn_s_sleep(1);
                    const EW<T> nx2 = lde<T>(rex, offX), ny2 = lde<T>(rey, offY);
                    if (needX) cx = nx2;
                    if (needY) cy = ny2;
                    wait = (needX && is_sentinel(cx)) || (needY && is_sentinel(cy));
                    ++spins;
                    if ((spins & 63) == 0 &&
                        __hip_atomic_load(a.sync.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                        break;
                }
                if (wait) atomicExch(a.sync.err, 1);
            }
            // recycle the consumed edge words
            ste<T>(rex, offX, sent);
            ste<T>(rey, offY, sent);
            // neighbour values in sweep order: -z (own previous step), -y, -x. Lanes
            // without an in-tile neighbour take the edge word, which reads as zero
            // where there is no cross-tile neighbour either (the matching block is
            // zero then); the z selection keeps values of steps before z = 0 out.
            const int rbuf = (tau + 1) & 1, wbuf = tau & 1;
            const P2<T> fx = xch[rbuf][ixl];
            const P2<T> fy = xch[rbuf][ixd];
            const P2<T> evx = ew_val(cx), evy = ew_val(cy);
            P2<T> nx_, ny_, nz_;
            nx_.x = hasL ? fx.x : evx.x;
            nx_.y = hasL ? fx.y : evx.y;
            ny_.x = hasD ? fy.x : evy.x;
            ny_.y = hasD ? fy.y : evy.y;
            nz_.x = z > 0 ? prev.x : T(0);
            nz_.y = z > 0 ? prev.y : T(0);
            const In<T, FWD> &c = ring[k];
            T r0 = c.rhs.x, r1 = c.rhs.y;
            P2<T> res;
            if (FWD) {
                // ascending columns: -z, -y, -x
                mmv(c.b[0], nz_, r0, r1);
                mmv(c.b[1], ny_, r0, r1);
                mmv(c.b[2], nx_, r0, r1);
                res = P2<T>{r0, r1};
            } else {
                // ascending columns: +x, +y, +z (sweep -x, -y, -z)
                mmv(c.b[0], nx_, r0, r1);
                mmv(c.b[1], ny_, r0, r1);
                mmv(c.b[2], nz_, r0, r1);
                T s0 = T(0), s1 = T(0);
                s0 += c.di[0] * r0; s0 += c.di[1] * r1;
                s1 += c.di[2] * r0; s1 += c.di[3] * r1;
                res = P2<T>{s0, s1};
            }
            stp<T>(rw, ovec, res);
            if (!FWD && sizeof(T) == 4)
                stp<double>(rout, oout,
                            P2<double>{static_cast<double>(res.x), static_cast<double>(res.y)});
            ste<T>(rex, prodX && active ? exOut + (unsigned)z * zw : kFar, val_ew(res));
            ste<T>(rey, prodY && active ? eyOut + (unsigned)z * zw : kFar, val_ew(res));
            prev = res;
            xch[wbuf][tid] = res;
            if (NWAVE > 1) {
                __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the LDS store is done
                __builtin_amdgcn_s_barrier();
            }
            __builtin_amdgcn_wave_barrier();
            ovec += svec;
            oout += sd;
            // operands of step tau + PD and edge words of step tau + PE into the slots
            // just consumed
            // (edge words first: the wait for them then covers only older operands)
            load_edge(ex[k % PE], ey[k % PE]);
            __builtin_amdgcn_sched_barrier(0);
            load(ring[k]);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

__global__ void k_sweep_reset_ticket(unsigned *ticket) { *ticket = 0u; }

}  // namespace

template <class T, bool FWD>
static void launch_sweep(const SweepArgs<T> &a, int tiles_x, hipStream_t s) {
    hipLaunchKernelGGL(k_sweep_reset_ticket, dim3(1), dim3(1), 0, s, a.sync.ticket);
    if (tiles_x == 16)
        hipLaunchKernelGGL((k_sweep<T, FWD, 16, 16, kSweepPD, kSweepPE>), dim3(a.sync.ntiles), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_sweep<T, FWD, 8, 8, kSweepPD, kSweepPE>), dim3(a.sync.ntiles), dim3(64), 0, s, a);
}

void sweep_apply64(const Grid &g, const double *L, const double *Dinv, const double *A,
                   const double *d, double *v, const int *status, const WfSync &sync,
                   int tile, hipStream_t s) {
    SweepArgs<double> f{};
    f.nx = g.nx; f.ny = g.ny; f.nz = g.nz; f.nxny = g.nxny; f.N = g.N;
    f.d = d;
    f.blk[0] = L; f.blk[1] = L + (size_t)g.N * 4; f.blk[2] = L + (size_t)2 * g.N * 4;
    f.blk_bytes = 32;
    f.Dinv = nullptr; f.w = v; f.out = nullptr; f.status = status; f.sync = sync;
    launch_sweep<double, true>(f, tile, s);
    SweepArgs<double> b = f;
    b.d = nullptr;
    b.blk[0] = A + (size_t)(g.diag + 1) * g.N * 4;
    b.blk[1] = A + (size_t)(g.diag + 2) * g.N * 4;
    b.blk[2] = A + (size_t)(g.diag + 3) * g.N * 4;
    b.Dinv = Dinv;
    launch_sweep<double, false>(b, tile, s);
}

void sweep_apply32(const Grid &g, const float *L, const float *Dinv, const float *U,
                   const double *d, float *w, double *v, const int *status, const WfSync &sync,
                   int tile, hipStream_t s) {
    SweepArgs<float> f{};
    f.nx = g.nx; f.ny = g.ny; f.nz = g.nz; f.nxny = g.nxny; f.N = g.N;
    f.d = d;
    f.blk[0] = L; f.blk[1] = L + (size_t)g.N * 4; f.blk[2] = L + (size_t)2 * g.N * 4;
    f.blk_bytes = 16;
    f.Dinv = nullptr; f.w = w; f.out = nullptr; f.status = status; f.sync = sync;
    launch_sweep<float, true>(f, tile, s);
    SweepArgs<float> b = f;
    b.d = nullptr;
    b.blk[0] = U; b.blk[1] = U + (size_t)g.N * 4; b.blk[2] = U + (size_t)2 * g.N * 4;
    b.Dinv = Dinv;
    b.out = v;
    launch_sweep<float, false>(b, tile, s);
}

}  // namespace ewg
