// ewg_sweep.hip -- the two ILU(0) triangular sweeps as persistent tile-wavefront launches,
// tuned for instruction count (the sweeps run twice per BiCGStab half-iteration).
//
// Schedule: 8x8 tiles of z-columns, each tile's lanes skewed so that lane (lx, ly) does
// z = step - lx - ly, tiles taken by ticket in anti-diagonal order (every producer is
// resident before its consumers), cross-tile values through edge buffers with a
// sentinel, with:
//   - per-lane 32-bit byte offsets that advance by a constant per step; a lane whose
//     cell is outside the grid (z < 0, z >= nz, or outside the plane) lands out of the
//     buffer range, so the hardware drops its stores and zeroes its loads: no
//     predication of memory operations (this needs N * 32 bytes plus the skew margin
//     below 2^31, which the kWfMaxCells limit guarantees);
//   - no predication of the block products either: blocks of absent neighbours are
//     stored as zero, and the neighbour values that enter them are finite (the backward
//     sweep's -z value of a column's first cell is selected to zero);
//   - the in-tile exchange of step results through LDS (one store and two loads per
//     step), and one comparison per edge direction to see whether a step's upstream
//     values are in place.
// The arithmetic per row is the one of the hyperplane kernels (ascending columns, no
// fused multiply-add), so both schedules agree to the last bit up to the sign of zeros.
#include <vector>

#include "ewg_cell_ops.h"
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

// arguments of the warp-specialized sweep, whose factors live in the wavefront layout:
// entry ((tile * nsteps + step) * 64 + lane) holds the cell that lane of that tile
// reaches at that local step (forward order for L and the forward result, backward
// order for U and Dinv; positions without a cell hold zeros)
template <class T>
struct WsArgs {
    int nx, ny, nz, nxny, N;
    int nsteps;           // local steps per tile: wl_nsteps(nz), a multiple of kRing
    unsigned wl_bytes_blk, wl_bytes_vec;  // byte sizes of one layout array of blocks / pairs
    const double *d;      // forward: right-hand side, natural order
    const T *blk[3];      // forward: L (-z,-y,-x); backward: U (+x,+y,+z)
    const T *Dinv;        // backward
    T *w;                 // forward result (forward layout): written forward, read backward
    double *out;          // backward: preconditioned vector, natural order
    const int *status;
    const int *prec;      // null, or skip the launch unless *prec == want (the ensemble's
    int want;             // per-member preconditioner precision: 0 FP32, 1 FP64)
    WfSync sync;
};

template <class T, bool FWD>
struct In {
    P2<T> rhs;
    T b[3][4];
    T di[4];
};

// ----------------------------------------------------------------------------------
// Warp-specialized sweep: one 8x8 tile per workgroup of two wavefronts.
//   wave 0 computes the tile's columns (lane (lx, ly) one z-column, the in-tile
//          neighbour values through a double-buffered LDS array) and reads the values of
//          the upstream tiles from an LDS ring instead of global memory, so that its vector
//          memory counter only holds operand prefetches issued PD steps earlier and its
//          own stores (a late cross-tile load would make every wait on it also wait for
//          all older loads and stores). It stores its last column and row straight to
//          the edge buffers with coherent stores: each is issued PD steps before the
//          wait on any load issued after it, which covers its completion.
//   wave 1 polls the upstream tiles' edge words (16 edge lanes x 4 z-values in flight),
//          copies each value into its ring slot once the compute wave has moved past the
//          value the slot held, publishes how many values of each edge lane are in place,
//          and recycles the global words to the sentinel.
// The compute wave checks each step against counts it read one step earlier (so the
// check does not expose the LDS latency) and announces its progress once per PD steps.
// (A third wave that took the edge values from an LDS outbox and stored them made every
// tile hand-over 1-2 us slower: the outbox round trip and that wave's polling.)
// ----------------------------------------------------------------------------------
constexpr int kRing = 16;  // ring slots per edge lane (z modulo kRing)
constexpr int kSweepThreads = 128;  // compute wave and edge wave
constexpr unsigned kLgkm0 = 0xc07f;  // s_waitcnt lgkmcnt(0), other counters untouched

__device__ __forceinline__ int lds_ld(const int *p) {
    return __hip_atomic_load(const_cast<int *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_st(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Buffer descriptors kept opaque to the compiler: every descriptor of a launch has the
// same flags word and several share their size word, and the compiler otherwise keeps
// only one copy of such words and rebuilds the four-register descriptors before each use
// (several scalar moves per step of the sweeps and the factorization).
__device__ __forceinline__ void pin(__amdgpu_buffer_rsrc_t &r) { asm volatile("" : "+s"(r)); }

// value of the edge wave's counts that never holds up a step
constexpr int kReady = 1 << 30;

template <class T, bool FWD, int PD>
__device__ __forceinline__ void sweep_ws_body(const WsArgs<T> &a) {
    constexpr int TX = 8, TY = 8, NT = 64;
    // the compute loop is unrolled over kRing steps, and the neighbour values a lane uses at
    // a step are kept in slot (step mod kRing) of its own column of xv (-x neighbour) and
    // yv (-y neighbour): every read is one LDS load at a fixed offset. The lanes of the
    // tile write the next step's slot of their +x / +y neighbour (the last column and row
    // write into spare columns), the edge wave the slots of the first column / row.
    static_assert(kRing % PD == 0 && PD < kRing, "prefetch depth must divide the ring");
    static_assert(kRing == 16, "wl_nsteps rounds the tile's steps to the unroll of 16");
    __shared__ P2<T> xv[kRing][NT + 1];
    __shared__ P2<T> yv[kRing][NT + TX];
    // per edge lane e (x edge: lane (0, e); y edge: lane (e - 8, 0)): the count of its
    // values in place plus e & 7 (the skew of the lane that uses them), kReady once all
    // are; [16]: kReady, read by the lanes off the first column / row
    __shared__ int filled[17];
    __shared__ int progress;              // compute steps completed (per PD steps)
    __shared__ unsigned s_tk;
    if (a.status && *a.status != KS_RUNNING) return;
    if (a.prec && *a.prec != a.want) return;
    const int tid = threadIdx.x;
    if (tid == 0) s_tk = atomicAdd(a.sync.ticket, 1u);
    __syncthreads();
    const unsigned tk = s_tk;
    if (tk >= (unsigned)a.sync.ntiles) return;
    const int2 tile = a.sync.order[tk];
    const int tx = tile.x, ty = tile.y;
    const int nz = a.nz;
    const int tileId = ty * a.sync.ntx + tx;
    // slots without an upstream tile read zeros for good (the matching blocks are zero as
    // well) and their edge lanes count as complete
    for (int i = tid; i < kRing * (NT + 1); i += kSweepThreads) (&xv[0][0])[i] = P2<T>{T(0), T(0)};
    for (int i = tid; i < kRing * (NT + TX); i += kSweepThreads) (&yv[0][0])[i] = P2<T>{T(0), T(0)};
    if (tid < 16) {
        const int dir = tid >> 3, e = tid & 7;
        const bool valid = dir == 0 ? (tx > 0 && ty * TY + e < a.ny) : (ty > 0 && tx * TX + e < a.nx);
        filled[tid] = valid ? e : kReady;
    }
    if (tid == 16) filled[16] = kReady;
    if (tid == 0) progress = 0;
    __syncthreads();
    constexpr unsigned zw = kWfMaxWords * 8;  // bytes per z of one edge lane
    __amdgpu_buffer_rsrc_t rex = rsrc(a.sync.ex, a.sync.edge_words * 8);
    __amdgpu_buffer_rsrc_t rey = rsrc(a.sync.ey, a.sync.edge_words * 8);
    pin(rex);
    pin(rey);

    if (tid >= NT) {
        // ---------------------------------------------------------------- edge wave
        // lanes e + 16 j (j = 0..3) serve edge lane e: e < 8 the x edge (the -x tile's
        // last column), e >= 8 the y edge (the -y tile's last row)
        const int l = tid - NT, e = l & 15, j = l >> 4;
        const int dir = e >> 3, idx = e & 7;
        const bool valid = dir == 0 ? (tx > 0 && ty * TY + idx < a.ny) : (ty > 0 && tx * TX + idx < a.nx);
        const unsigned gbase =
            !valid ? 0u
            : dir == 0 ? (unsigned)((((size_t)(tileId - 1) * TY + idx) * nz) * zw)
                       : (unsigned)((((size_t)(tileId - a.sync.ntx) * TX + idx) * nz) * zw);
        // value z of this edge lane is used by lane (0, idx) (x edge) or (idx, 0) (y edge)
        // at its step z + idx
        P2<T> *const ring = dir == 0 ? &xv[0][TX * idx] : &yv[0][idx];
        const int sstride = dir == 0 ? NT + 1 : NT + TX;  // elements from one slot to the next
        EW<T> sent;
        sent.w[0] = kSentinel;
        if constexpr (sizeof(T) == 8) sent.w[1] = kSentinel;
        // sub-lane j delivers the values z = j (mod 4) of its edge lane in order; the
        // count of values in place is the smallest z any of the four still lacks
        int z = valid ? j : nz;
        unsigned idle = 0;
        int published = valid ? 0 : nz;
        while (__builtin_amdgcn_ballot_w64(z < nz) != 0ull) {
            const bool mine = z < nz;
            // the slot held the value of step z + idx - kRing, consumed once the compute
            // wave has completed that step
            const bool room = z - kRing + idx < lds_ld(&progress);
            const unsigned go = mine ? gbase + (unsigned)z * zw : kFar;
            const EW<T> vx = lde<T>(rex, dir == 0 ? go : kFar);
            const EW<T> vy = lde<T>(rey, dir == 1 ? go : kFar);
            const EW<T> v = dir == 0 ? vx : vy;
            const bool deliver = mine && room && !is_sentinel(v);
            if (deliver) {
                ring[((z + idx) & (kRing - 1)) * sstride] = ew_val(v);
                ste<T>(rex, dir == 0 ? go : kFar, sent);
                ste<T>(rey, dir == 1 ? go : kFar, sent);
                z += 4;
            }
            // smallest missing z over the four sub-lanes (lanes e, e+16, e+32, e+48)
            int zmin = z < nz ? z : nz;
            zmin = min(zmin, __shfl_xor(zmin, 16));
            zmin = min(zmin, __shfl_xor(zmin, 32));
            const bool advance = j == 0 && zmin > published;
            if (__builtin_amdgcn_ballot_w64(advance) != 0ull) {
                // the ring values land before the count (each wave's LDS operations
                // execute in order; the fences keep the compiler from reordering them)
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __builtin_amdgcn_s_waitcnt(kLgkm0);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                if (advance) lds_st(&filled[e], zmin >= nz ? kReady : zmin + idx);
                published = zmin > published ? zmin : published;
            }
            if (__builtin_amdgcn_ballot_w64(deliver) != 0ull) {
                idle = 0;
            } else {
                if (++idle >= kSpinLimit) {
                    atomicExch(a.sync.err, 1);
                    break;
                }
                if ((idle & 63) == 0 &&
                    __hip_atomic_load(a.sync.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    break;
                __builtin_amdgcn_s_sleep(1);
            }
        }
        return;
    }

    // -------------------------------------------------------------------- compute wave
    const int lx = tid % TX, ly = tid / TX;
    const int xs = tx * TX + lx, ys = ty * TY + ly;
    const bool inplane = xs < a.nx && ys < a.ny;
    const bool prodX = inplane && lx == TX - 1 && tx + 1 < a.sync.ntx;
    const bool prodY = inplane && ly == TY - 1 && ty + 1 < a.sync.nty;

    // natural-order cell of (xs, ys, z) = c0 + z * dz in sweep coordinates
    int c0, dz;
    if (FWD) {
        c0 = xs + a.nx * ys;
        dz = a.nxny;
    } else {
        c0 = (a.nx - 1 - xs) + a.nx * (a.ny - 1 - ys) + a.nxny * (nz - 1);
        dz = -a.nxny;
    }
    const int cs0 = c0 - (lx + ly) * dz;  // cell of local step 0
    const int nsteps = a.nsteps;          // a multiple of kRing
    constexpr unsigned bB = 4 * sizeof(T), vB = 2 * sizeof(T);
    // wavefront-layout arrays: this lane's entry of local step 0, then 64 entries on
    const unsigned wl0 = (unsigned)((tileId * nsteps) * 64 + tid);
    unsigned pblk = wl0 * bB;
    constexpr unsigned sblk = 64 * bB;
    // natural-order vectors (forward right-hand side, backward result)
    unsigned pnat = inplane ? (unsigned)(cs0 * 16) : kFar;
    const unsigned snat = inplane ? (unsigned)(dz * 16) : 0u;
    unsigned onat = pnat;
    // forward result: written in the forward layout by the forward sweep (every step the
    // loop runs has its own entry); the backward sweep reads entry (forward tile,
    // z + lxf + lyf, forward lane) of its cell, which moves back by one row of 64 per step
    unsigned ow = wl0 * vB;  // forward: store offset of the chunk's first step
    unsigned pw = kFar;      // backward: load offset
    if (!FWD && inplane) {
        const int x = a.nx - 1 - xs, y = a.ny - 1 - ys;  // natural coordinates
        const int z0 = nz - 1 + lx + ly;                  // natural z at local step 0
        const int lxf = x & 7, lyf = y & 7;
        const int tf = (y >> 3) * a.sync.ntx + (x >> 3);
        pw = (unsigned)(((tf * nsteps + z0 + lxf + lyf) * 64 + lxf + 8 * lyf) * (int)vB);
    }
    constexpr unsigned sw = 64 * vB;  // the forward entry moves by sw per step
    // this tile's words in the edge buffers: its last column (edge lane ly) and last row
    // (edge lane lx), z-value words consecutive
    const unsigned gX = (unsigned)((((size_t)tileId * 8 + ly) * nz) * zw);
    const unsigned gY = (unsigned)((((size_t)tileId * 8 + lx) * nz) * zw);
    __amdgpu_buffer_rsrc_t rb0 = rsrc(a.blk[0], a.wl_bytes_blk), rb1 = rsrc(a.blk[1], a.wl_bytes_blk),
                           rb2 = rsrc(a.blk[2], a.wl_bytes_blk);
    __amdgpu_buffer_rsrc_t rdi = rsrc(a.Dinv, FWD ? 0 : a.wl_bytes_blk);
    __amdgpu_buffer_rsrc_t rw = rsrc(a.w, a.wl_bytes_vec);
    const size_t N = (size_t)a.N;
    __amdgpu_buffer_rsrc_t rd = rsrc(a.d, FWD ? N * 16 : 0);
    __amdgpu_buffer_rsrc_t rout = rsrc(a.out, FWD ? 0 : N * 16);
    pin(rb0);
    pin(rb1);
    pin(rb2);
    pin(rdi);
    pin(rw);
    pin(rd);
    pin(rout);

    In<T, FWD> ring[PD];
    auto load = [&](In<T, FWD> &in) {
        if (FWD) {
            P2<double> dv = ldp<double>(rd, pnat);
            in.rhs = P2<T>{static_cast<T>(dv.x), static_cast<T>(dv.y)};
            pnat += snat;
        } else {
            in.rhs = ldp<T>(rw, pw);
            pw -= sw;
        }
        ldb<T>(rb0, pblk, in.b[0]);
        ldb<T>(rb1, pblk, in.b[1]);
        ldb<T>(rb2, pblk, in.b[2]);
        if (!FWD) ldb<T>(rdi, pblk, in.di);
        pblk += sblk;
    };
    // the scheduling barriers keep the prefetches in step order: the in-order vector
    // memory counter then lets each step wait for its own operands only
#pragma unroll
    for (int k = 0; k < PD; ++k) {
        load(ring[k]);
        __builtin_amdgcn_sched_barrier(0);
    }

    P2<T> prev = P2<T>{T(0), T(0)};
    // this lane's exchange columns, and the columns it writes (its +x / +y neighbour's, or
    // a spare column)
    const P2<T> *const rx = &xv[0][tid];
    const P2<T> *const ry = &yv[0][tid];
    P2<T> *const wx = lx < TX - 1 ? &xv[0][tid + 1] : &xv[0][NT];
    P2<T> *const wy = ly < TY - 1 ? &yv[0][tid + TX] : &yv[0][NT + lx];
    const int *const fX = lx == 0 ? &filled[ly] : &filled[16];
    const int *const fY = ly == 0 ? &filled[8 + lx] : &filled[16];

    // The counts and the neighbour values of a step are read at the end of the step
    // before, so that their LDS latency overlaps that step's tail and the check; the
    // counts come first (a ring value read after a count that covers it is in place),
    // and the rare wait re-reads the values. Value z = tau - lx - ly of a first-column
    // lane's edge lane is in place once that count exceeds tau (z < 0: from the start).
    int cntX = lds_ld(fX), cntY = lds_ld(fY);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    P2<T> nx_ = rx[0], ny_ = ry[0];
    for (int tau0 = 0; tau0 < nsteps; tau0 += kRing) {
#pragma unroll
        for (int k = 0; k < kRing; ++k) {
            const int tau = tau0 + k;
            if (k % PD == 0) {
                // steps before tau are complete: their slots may be refilled (the
                // compiler fence keeps the earlier steps' reads ahead of the
                // announcement; LDS executes them in that order)
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                if (tid == 0) lds_st(&progress, tau);
            }
            if ((__builtin_amdgcn_ballot_w64(cntX <= tau) | __builtin_amdgcn_ballot_w64(cntY <= tau)) !=
                0ull) {
                unsigned spins = 0;
                bool wait;
                do {
                    __builtin_amdgcn_s_sleep(1);
                    cntX = lds_ld(fX);
                    cntY = lds_ld(fY);
                    wait = cntX <= tau || cntY <= tau;
                    ++spins;
                    if ((spins & 63) == 0 &&
                        __hip_atomic_load(a.sync.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                        break;
                } while (__builtin_amdgcn_ballot_w64(wait) != 0ull && spins < kSpinLimit);
                if (wait) atomicExch(a.sync.err, 1);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                nx_ = rx[k * (NT + 1)];
                ny_ = ry[k * (NT + TX)];
            }
            // the -z neighbour: the lane's previous result. Forward: the steps before the
            // column's first cell load a zero right-hand side (out of range) and zero
            // blocks (entries without a cell), so their results are zeros and the first
            // cell gets a zero -z value without a test (the sign of a zero aside).
            // Backward: those steps read other entries of the forward result, so the
            // first cell (z = 0: tau = lx + ly) takes an explicit zero.
            P2<T> nz_ = prev;
            if (!FWD) {
                const bool zpos = tau > lx + ly;
                nz_.x = zpos ? prev.x : T(0);
                nz_.y = zpos ? prev.y : T(0);
            }
            const In<T, FWD> &c = ring[k % PD];
            T r0 = c.rhs.x, r1 = c.rhs.y;
            P2<T> res;
            if (FWD) {
                mmv(c.b[0], nz_, r0, r1);
                mmv(c.b[1], ny_, r0, r1);
                mmv(c.b[2], nx_, r0, r1);
                res = P2<T>{r0, r1};
            } else {
                mmv(c.b[0], nx_, r0, r1);
                mmv(c.b[1], ny_, r0, r1);
                mmv(c.b[2], nz_, r0, r1);
                T s0 = c.di[0] * r0, s1 = c.di[2] * r0;
                s0 += c.di[1] * r1;
                s1 += c.di[3] * r1;
                res = P2<T>{s0, s1};
            }
            // the downstream tiles' values (the cells of the column only: unconsumed words
            // would stay in the buffers), then the result vector
            const int z = tau - lx - ly;
            const bool zin = (unsigned)z < (unsigned)nz;
            ste<T>(rex, prodX && zin ? gX + (unsigned)z * zw : kFar, val_ew(res));
            ste<T>(rey, prodY && zin ? gY + (unsigned)z * zw : kFar, val_ew(res));
            if (FWD) {
                stp<T>(rw, ow + (unsigned)k * sw, res);
            } else {
                stp<double>(rout, onat,
                            P2<double>{static_cast<double>(res.x), static_cast<double>(res.y)});
                onat += snat;
            }
            prev = res;
            const int kn = (k + 1) & (kRing - 1);
            wx[kn * (NT + 1)] = res;
            wy[kn * (NT + TX)] = res;
            __builtin_amdgcn_wave_barrier();
            // counts and neighbour values for the next step
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            cntX = lds_ld(fX);
            cntY = lds_ld(fY);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            nx_ = rx[kn * (NT + 1)];
            ny_ = ry[kn * (NT + TX)];
            load(ring[k % PD]);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (FWD) ow += kRing * sw;
    }
}

template <class T, bool FWD, int PD>
__global__ void __launch_bounds__(kSweepThreads) k_sweep_ws(WsArgs<T> a) {
    sweep_ws_body<T, FWD, PD>(a);
}

// the same sweep for several realizations in one launch: block -> realization by tmem,
// each block then takes a tile ticket of its realization as the single launch does
template <class T, bool FWD, int PD>
__global__ void __launch_bounds__(kSweepThreads) k_sweep_ws_batch(const WsArgs<T> *args, const int *tmem) {
    const WsArgs<T> a = args[tmem[blockIdx.x]];
    sweep_ws_body<T, FWD, PD>(a);
}

__global__ void k_reset_tickets(unsigned *const *tickets, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) *tickets[i] = 0u;
}

__global__ void k_sweep_reset_ticket(unsigned *ticket) { *ticket = 0u; }
__global__ void k_reset_ticket_pair(unsigned *ticket) {
    ticket[0] = 0u;
    ticket[1] = 0u;
}

// ----------------------------------------------------------------------------------
// The factorization straight into the wavefront layout, warp-specialized like the
// sweeps (natural order, i.e. the forward sweep's schedule):
//   wave 0 computes Dinv_i = inv(A_ii - sum_j (A_ij Dinv_j) A_ji) over the -z, -y, -x
//          neighbours j (Dinv_j from its own register, the LDS exchange or the edge
//          ring), stores L_ij = A_ij Dinv_j in the forward layout (zeros where there is
//          no cell or no neighbour) and Dinv_i at the cell's backward-layout entry,
//          and stores its last column's and row's Dinv to the edge buffers;
//   wave 1 polls the upstream tiles' edge words into the LDS ring as in the sweeps
//          (four 64-bit words per value).
// The compute wave reads its operands from a copy of the Jacobian's blocks in the
// forward layout (k_wl_pack_factor, a pass over the cells before the launch), so that
// its loads are contiguous like the sweeps' (the natural-order blocks of a skewed tile
// column touch a separate cache line per lane and load). The same pass copies the upper
// blocks A_i,+x/+y/+z into the backward layout.
// Stored as T: FP64, or FP32 rounded to nearest from the FP64 factors. The arithmetic
// per row is ilu_factor_cell's (ascending columns, no fused multiply-add), so every
// schedule produces the same factors bit for bit. The backward layout's entries without
// a cell are never written here (they are zero from allocation on).
// ----------------------------------------------------------------------------------
template <class T>
struct FwsArgs {
    int nx, ny, nz, nxny, N, diag;
    int nsteps;            // local steps per tile: wl_nsteps(nz)
    const double *A;       // Jacobian, slot-major [2 diag + 1][N][4]
    double *Af;            // forward layout, 7 arrays of entries x 4: A_ii, A_ij (-z, -y,
                           // -x), A_ji (the -z, -y, -x neighbour's block towards the cell)
    T *L;                  // forward layout: 3 arrays (-z, -y, -x) of entries x 4
    T *U;                  // backward layout: 3 arrays (+x, +y, +z)
    T *D;                  // backward layout: Dinv
    size_t entries;
    int *err;              // singular pivot (|det| < 1e-30)
    const int *active;     // null, or skip the launch unless *active
    float *A32;            // null, or the packing pass also stores every block of A rounded
                           // to FP32 here, slot-major like A (BiCGStab's FP32 operator)
    WfSync sync;
};

struct B4 {
    double a[4];
};

// one exchanged Dinv: 4 coherent 64-bit words
__device__ __forceinline__ void lde4(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned long long *w) {
    const u32x4 u = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, kCoh);
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16u, 0, kCoh);
    w[0] = (unsigned long long)u.x | ((unsigned long long)u.y << 32);
    w[1] = (unsigned long long)u.z | ((unsigned long long)u.w << 32);
    w[2] = (unsigned long long)v.x | ((unsigned long long)v.y << 32);
    w[3] = (unsigned long long)v.z | ((unsigned long long)v.w << 32);
}
__device__ __forceinline__ void ste4(__amdgpu_buffer_rsrc_t r, unsigned off, const unsigned long long *w) {
    u32x4 u, v;
    u.x = (unsigned)w[0]; u.y = (unsigned)(w[0] >> 32); u.z = (unsigned)w[1]; u.w = (unsigned)(w[1] >> 32);
    v.x = (unsigned)w[2]; v.y = (unsigned)(w[2] >> 32); v.z = (unsigned)w[3]; v.w = (unsigned)(w[3] >> 32);
    __builtin_amdgcn_raw_buffer_store_b128(u, r, off, 0, kCoh);
    __builtin_amdgcn_raw_buffer_store_b128(v, r, off + 16u, 0, kCoh);
}

// a 2x2 FP64 block stored as T
template <class T>
__device__ __forceinline__ void stb(__amdgpu_buffer_rsrc_t r, unsigned off, const double *v) {
    if constexpr (sizeof(T) == 8) {
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, make_double2(v[0], v[1])), r, off, 0, 0);
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, make_double2(v[2], v[3])), r, off + 16u, 0, 0);
    } else {
        __builtin_amdgcn_raw_buffer_store_b128(
            __builtin_bit_cast(u32x4, make_float4(static_cast<float>(v[0]), static_cast<float>(v[1]),
                                                  static_cast<float>(v[2]), static_cast<float>(v[3]))),
            r, off, 0, 0);
    }
}

struct FIn {
    double Aii[4];
    double Aij[3][4];  // -z, -y, -x
    double Aji[3][4];  // the neighbour's block towards this cell
};

template <class T, int PD>
__device__ __forceinline__ void factor_ws_body(const FwsArgs<T> &a) {
    constexpr int TX = 8, TY = 8, NT = 64;
    static_assert(PD < kRing, "the edge wave must be able to run ahead of the compute wave");
    __shared__ B4 xch[2][NT];
    __shared__ B4 ering[2][8][kRing];
    __shared__ int filled[16];
    __shared__ int progress;
    __shared__ unsigned s_tk;
    if (a.active && *a.active == 0) return;
    const int tid = threadIdx.x;
    if (tid == 0) s_tk = atomicAdd(a.sync.ticket, 1u);
    __syncthreads();
    const unsigned tk = s_tk;
    if (tk >= (unsigned)a.sync.ntiles) return;
    const int2 tile = a.sync.order[tk];
    const int tx = tile.x, ty = tile.y;
    const int nz = a.nz, nsteps = a.nsteps;
    const int ntx = a.sync.ntx;
    const int tileId = ty * ntx + tx;
    for (int i = tid; i < 2 * 8 * kRing; i += 2 * NT) (&ering[0][0][0])[i] = B4{{0.0, 0.0, 0.0, 0.0}};
    for (int i = tid; i < 2 * NT; i += 2 * NT) (&xch[0][0])[i] = B4{{0.0, 0.0, 0.0, 0.0}};
    if (tid < 16) {
        const int dir = tid >> 3, e = tid & 7;
        const bool valid = dir == 0 ? (tx > 0 && ty * TY + e < a.ny) : (ty > 0 && tx * TX + e < a.nx);
        filled[tid] = valid ? 0 : nz;
    }
    if (tid == 0) progress = 0;
    __syncthreads();
    constexpr unsigned zw = kWfMaxWords * 8;
    static_assert(kWfMaxWords == 4, "a Dinv block takes four edge words");
    __amdgpu_buffer_rsrc_t rex = rsrc(a.sync.ex, a.sync.edge_words * 8);
    __amdgpu_buffer_rsrc_t rey = rsrc(a.sync.ey, a.sync.edge_words * 8);
    pin(rex);
    pin(rey);
    constexpr unsigned bB = 4 * sizeof(T);
    const unsigned blk_bytes = (unsigned)(a.entries * bB);
    const int wave = tid >> 6, l = tid & 63;

    // the column (lx, ly) of the tile and its entry in the backward layout: the cell of
    // local step tau sits at backward entry eb0 - 64 tau
    const int lx = l % TX, ly = l / TX;
    const int xs = tx * TX + lx, ys = ty * TY + ly;
    const bool inplane = xs < a.nx && ys < a.ny;
    int eb0 = 0;
    if (inplane) {
        const int xb = a.nx - 1 - xs, yb = a.ny - 1 - ys;
        const int lxb = xb & 7, lyb = yb & 7;
        const int tb = (yb >> 3) * ntx + (xb >> 3);
        eb0 = (tb * nsteps + nz - 1 + lx + ly + lxb + lyb) * 64 + lxb + 8 * lyb;
    }

    if (wave == 1) {
        // ---------------------------------------------------------------- edge wave
        const int e = l & 15, j = l >> 4;
        const int dir = e >> 3, idx = e & 7;
        const bool valid = dir == 0 ? (tx > 0 && ty * TY + idx < a.ny) : (ty > 0 && tx * TX + idx < a.nx);
        const unsigned gbase =
            !valid ? 0u
            : dir == 0 ? (unsigned)((((size_t)(tileId - 1) * TY + idx) * nz) * zw)
                       : (unsigned)((((size_t)(tileId - ntx) * TX + idx) * nz) * zw);
        B4 *ring = &ering[dir][idx][0];
        const unsigned long long sent[4] = {kSentinel, kSentinel, kSentinel, kSentinel};
        int z = valid ? j : nz;
        unsigned idle = 0;
        int published = valid ? 0 : nz;
        while (__builtin_amdgcn_ballot_w64(z < nz) != 0ull) {
            const bool mine = z < nz;
            const bool room = z - kRing + idx < lds_ld(&progress);
            const unsigned go = mine ? gbase + (unsigned)z * zw : kFar;
            unsigned long long vx[4], vy[4];
            lde4(rex, dir == 0 ? go : kFar, vx);
            lde4(rey, dir == 1 ? go : kFar, vy);
            unsigned long long v[4];
            bool got = true;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                v[k] = dir == 0 ? vx[k] : vy[k];
                got &= v[k] != kSentinel;
            }
            const bool deliver = mine && room && got;
            if (deliver) {
                B4 b;
                __builtin_memcpy(b.a, v, sizeof(v));
                ring[z & (kRing - 1)] = b;
                ste4(rex, dir == 0 ? go : kFar, sent);
                ste4(rey, dir == 1 ? go : kFar, sent);
                z += 4;
            }
            int zmin = z < nz ? z : nz;
            zmin = min(zmin, __shfl_xor(zmin, 16));
            zmin = min(zmin, __shfl_xor(zmin, 32));
            const bool advance = j == 0 && zmin > published;
            if (__builtin_amdgcn_ballot_w64(advance) != 0ull) {
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __builtin_amdgcn_s_waitcnt(kLgkm0);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                if (advance) lds_st(&filled[e], zmin);
                published = zmin > published ? zmin : published;
            }
            if (__builtin_amdgcn_ballot_w64(deliver) != 0ull) {
                idle = 0;
            } else {
                if (++idle >= kSpinLimit) {
                    atomicExch(a.sync.err, 1);
                    break;
                }
                if ((idle & 63) == 0 &&
                    __hip_atomic_load(a.sync.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    break;
                __builtin_amdgcn_s_sleep(1);
            }
        }
        return;
    }

    // -------------------------------------------------------------------- compute wave
    const bool prodX = inplane && lx == TX - 1 && tx + 1 < ntx;
    const bool prodY = inplane && ly == TY - 1 && ty + 1 < a.sync.nty;
    const size_t E4 = a.entries * 4;
    const unsigned af_bytes = (unsigned)(a.entries * 32);
    __amdgpu_buffer_rsrc_t rAii = rsrc(a.Af, af_bytes);
    __amdgpu_buffer_rsrc_t rAij0 = rsrc(a.Af + E4, af_bytes), rAij1 = rsrc(a.Af + 2 * E4, af_bytes),
                           rAij2 = rsrc(a.Af + 3 * E4, af_bytes);
    __amdgpu_buffer_rsrc_t rAji0 = rsrc(a.Af + 4 * E4, af_bytes), rAji1 = rsrc(a.Af + 5 * E4, af_bytes),
                           rAji2 = rsrc(a.Af + 6 * E4, af_bytes);
    __amdgpu_buffer_rsrc_t rL0 = rsrc(a.L, blk_bytes), rL1 = rsrc(a.L + a.entries * 4, blk_bytes),
                           rL2 = rsrc(a.L + 2 * a.entries * 4, blk_bytes);
    __amdgpu_buffer_rsrc_t rD = rsrc(a.D, blk_bytes);
    pin(rAii);
    pin(rAij0);
    pin(rAij1);
    pin(rAij2);
    pin(rAji0);
    pin(rAji1);
    pin(rAji2);
    pin(rL0);
    pin(rL1);
    pin(rL2);
    pin(rD);
    const unsigned wl0 = (unsigned)((tileId * nsteps) * 64 + l);

    FIn ring[PD];
    // operands of local step tau: this lane's forward-layout entry of that step (steps
    // past the tile's last read out of range)
    auto load = [&](FIn &in, int tau) {
        const unsigned o = tau < nsteps ? (wl0 + (unsigned)tau * 64u) * 32u : kFar;
        ldb<double>(rAii, o, in.Aii);
        ldb<double>(rAij0, o, in.Aij[0]);
        ldb<double>(rAij1, o, in.Aij[1]);
        ldb<double>(rAij2, o, in.Aij[2]);
        ldb<double>(rAji0, o, in.Aji[0]);
        ldb<double>(rAji1, o, in.Aji[1]);
        ldb<double>(rAji2, o, in.Aji[2]);
    };
#pragma unroll
    for (int k = 0; k < PD; ++k) {
        load(ring[k], k);
        __builtin_amdgcn_sched_barrier(0);
    }

    B4 prev = B4{{0.0, 0.0, 0.0, 0.0}};
    const bool hasL = lx > 0, hasD = ly > 0;
    // this tile's words in the edge buffers (last column: edge lane ly; last row: lx)
    const unsigned gX = (unsigned)((((size_t)tileId * 8 + ly) * nz) * zw);
    const unsigned gY = (unsigned)((((size_t)tileId * 8 + lx) * nz) * zw);
    const B4 *const eX = &ering[0][ly][0];
    const B4 *const eY = &ering[1][lx][0];
    const int *const fX = &filled[ly];
    const int *const fY = &filled[8 + lx];
    const bool edgeX = lx == 0, edgeY = ly == 0;
    auto nbr = [&](int tau, B4 &vx, B4 &vy) {
        const int z = tau - lx - ly;
        const int rbuf = (tau + 1) & 1;
        const int slot = z & (kRing - 1);
        vx = hasL ? xch[rbuf][l - 1] : eX[slot];
        vy = hasD ? xch[rbuf][l - TX] : eY[slot];
    };
    int cntX = lds_ld(fX), cntY = lds_ld(fY);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    B4 nx_, ny_;
    nbr(0, nx_, ny_);
    for (int tau0 = 0; tau0 < nsteps; tau0 += PD) {
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (l == 0) lds_st(&progress, tau0);
#pragma unroll
        for (int k = 0; k < PD; ++k) {
            const int tau = tau0 + k;
            const int z = tau - lx - ly;
            const bool zin = (unsigned)z < (unsigned)nz;
            const bool active = inplane && zin && tau < nsteps;
            const int wbuf = tau & 1;
            bool wait = (edgeX & zin & (cntX <= z)) | (edgeY & zin & (cntY <= z));
            if (__builtin_amdgcn_ballot_w64(wait) != 0ull) {
                unsigned spins = 0;
                do {
                    __builtin_amdgcn_s_sleep(1);
                    cntX = lds_ld(fX);
                    cntY = lds_ld(fY);
                    wait = (edgeX & zin & (cntX <= z)) | (edgeY & zin & (cntY <= z));
                    ++spins;
                    if ((spins & 63) == 0 &&
                        __hip_atomic_load(a.sync.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                        break;
                } while (__builtin_amdgcn_ballot_w64(wait) != 0ull && spins < kSpinLimit);
                if (wait) atomicExch(a.sync.err, 1);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                nbr(tau, nx_, ny_);
            }
            const FIn &c = ring[k];
            const double *nb[3] = {prev.a, ny_.a, nx_.a};
            // Branch-free: an absent neighbour's blocks are zero in the packed operands
            // (and so are all blocks of an entry without a cell), and every neighbour
            // value is finite, so its L_ij is a zero and D minus it is D (a zero entry of D
            // may change its sign): the same numbers as skipping the neighbour.
            double D[4] = {c.Aii[0], c.Aii[1], c.Aii[2], c.Aii[3]};
            double Lv[3][4];
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                double Bm[4];
                cell::bmul(c.Aij[s], nb[s], Lv[s]);
                cell::bmul(Lv[s], c.Aji[s], Bm);
#pragma unroll
                for (int e = 0; e < 4; ++e) D[e] -= Bm[e];
            }
            double det = D[0] * D[3] - D[1] * D[2];
            const bool sing = !(fabs(det) >= 1e-30);
            if (active && sing) atomicExch(a.err, 1);
            if (sing) det = 1.0;
            const double det_1 = 1.0 / det;
            B4 res;
            res.a[0] = D[3] * det_1;
            res.a[1] = -D[1] * det_1;
            res.a[2] = -D[2] * det_1;
            res.a[3] = D[0] * det_1;
            {
                // the downstream tiles' edge words (coherent stores)
                unsigned long long w[4];
                __builtin_memcpy(w, res.a, sizeof(w));
                ste4(rex, prodX && zin ? gX + (unsigned)z * zw : kFar, w);
                ste4(rey, prodY && zin ? gY + (unsigned)z * zw : kFar, w);
            }
            const unsigned oL = tau < nsteps ? (wl0 + (unsigned)tau * 64u) * bB : kFar;
            stb<T>(rL0, oL, Lv[0]);
            stb<T>(rL1, oL, Lv[1]);
            stb<T>(rL2, oL, Lv[2]);
            stb<T>(rD, active ? (unsigned)(eb0 - tau * 64) * bB : kFar, res.a);
            prev = res;
            xch[wbuf][l] = res;
            __builtin_amdgcn_wave_barrier();
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            cntX = lds_ld(fX);
            cntY = lds_ld(fY);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            nbr(tau + 1, nx_, ny_);
            load(ring[k], tau + PD);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

template <class T, int PD>
__global__ void __launch_bounds__(128) k_factor_ws(FwsArgs<T> a) {
    factor_ws_body<T, PD>(a);
}

template <class T, int PD>
__global__ void __launch_bounds__(128) k_factor_ws_batch(const FwsArgs<T> *args, const int *tmem) {
    const FwsArgs<T> a = args[tmem[blockIdx.x]];
    factor_ws_body<T, PD>(a);
}

// The factorization's operands in the forward layout and the upper blocks in the
// backward layout, one thread per cell (see pack_cell_of for the order): each cell's
// entries are written where its tile column reaches it (32 bytes at a time). Entries
// without a cell are not written (zero from allocation).
template <class T>
__device__ __forceinline__ void pack_factor_cell(const FwsArgs<T> &a, int c) {
    if (c >= a.N) return;
    const int nsteps = a.nsteps, ntx = a.sync.ntx;
    const int x = c % a.nx, r = c / a.nx, y = r % a.ny, z = r / a.ny;
    const size_t N = (size_t)a.N, E4 = a.entries * 4;
    const double2 *__restrict__ A = reinterpret_cast<const double2 *>(a.A);
    // all ten blocks first (the loads overlap), then the stores
    const long long str[3] = {(long long)a.nxny, (long long)a.nx, 1};
    const bool has[3] = {z > 0, y > 0, x > 0};
    double2 v[10][2];
    auto blk = [&](bool h, size_t slot, long long cell, double2 *o) {
        const size_t k = (slot * N + (size_t)(h ? cell : 0)) * 2;
        const double2 u = A[k], w = A[k + 1];
        o[0] = h ? u : make_double2(0.0, 0.0);
        o[1] = h ? w : make_double2(0.0, 0.0);
    };
    blk(true, a.diag, c, v[0]);
    for (int s = 0; s < 3; ++s) {
        blk(true, s, c, v[1 + s]);
        // the lower neighbour's block towards this cell: its +z, +y, +x slot
        blk(has[s], a.diag + 3 - s, c - str[s], v[4 + s]);
        blk(true, a.diag + 1 + s, c, v[7 + s]);
    }
    // forward entry: tile (x / 8, y / 8), local step z + lx + ly, lane lx + 8 ly
    const int lx = x & 7, ly = y & 7;
    const size_t ef = ((size_t)((y >> 3) * ntx + (x >> 3)) * nsteps + z + lx + ly) * 64 + lx + 8 * ly;
    double *__restrict__ Af = a.Af;
    for (int k = 0; k < 7; ++k) {
        double2 *dst = reinterpret_cast<double2 *>(Af + k * E4 + ef * 4);
        dst[0] = v[k][0];
        dst[1] = v[k][1];
    }
    // backward entry of the cell: the same in the reflected grid
    const int xb = a.nx - 1 - x, yb = a.ny - 1 - y, zb = a.nz - 1 - z;
    const int lxb = xb & 7, lyb = yb & 7;
    const size_t eb = ((size_t)((yb >> 3) * ntx + (xb >> 3)) * nsteps + zb + lxb + lyb) * 64 + lxb + 8 * lyb;
    if (a.A32) {
        // the cell's own seven blocks (diagonal, lower and upper slots) rounded to FP32
        auto st32 = [&](int slot, const double2 *b) {
            *reinterpret_cast<float4 *>(a.A32 + ((size_t)slot * N + c) * 4) =
                make_float4(static_cast<float>(b[0].x), static_cast<float>(b[0].y),
                            static_cast<float>(b[1].x), static_cast<float>(b[1].y));
        };
        st32(a.diag, v[0]);
        for (int s = 0; s < 3; ++s) {
            st32(s, v[1 + s]);
            st32(a.diag + 1 + s, v[7 + s]);
        }
    }
    T *__restrict__ U = a.U;
    for (int s = 0; s < 3; ++s) {
        T *dst = U + s * E4 + eb * 4;
        const double2 p0 = v[7 + s][0], p1 = v[7 + s][1];
        if constexpr (sizeof(T) == 8) {
            reinterpret_cast<double2 *>(dst)[0] = p0;
            reinterpret_cast<double2 *>(dst)[1] = p1;
        } else {
            *reinterpret_cast<float4 *>(dst) =
                make_float4(static_cast<float>(p0.x), static_cast<float>(p0.y),
                            static_cast<float>(p1.x), static_cast<float>(p1.y));
        }
    }
}

// block b: 4 z-planes of one tile column (b = tile * ceil(nz / 4) + z-chunk), thread t:
// column (t % 8, (t / 8) % 8) at z-plane t / 64, so that a wave reads one plane of the
// tile (eight 8-cell runs) and a block's writes fill rows of its tile's entries together
template <class T>
__device__ __forceinline__ int pack_cell_of(const FwsArgs<T> &a, int b, int t) {
    const int nzc = (a.nz + 3) / 4;
    const int tile = b / nzc, zc = b % nzc;
    const int x = (tile % a.sync.ntx) * 8 + (t & 7), y = (tile / a.sync.ntx) * 8 + ((t >> 3) & 7);
    const int z = 4 * zc + (t >> 6);
    return x < a.nx && y < a.ny && z < a.nz ? x + a.nx * (y + a.ny * z) : a.N;
}

// the z-chunks of a tile column write parts of the same rows of entries: renumbered for
// the XCDs, they run together on one XCD and their partial cache lines meet in its L2
template <class T>
__global__ void __launch_bounds__(256) k_wl_pack_factor(FwsArgs<T> a) {
    pack_factor_cell<T>(a, pack_cell_of(a, xcd_block(blockIdx.x, gridDim.x), threadIdx.x));
}

// several realizations: block -> realization by pmem, its first block at poff (blocks
// renumbered for the XCDs over the whole grid, as in the single launch)
template <class T>
__global__ void __launch_bounds__(256)
k_wl_pack_factor_batch(const FwsArgs<T> *args, const int *pmem, const int *poff) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int m = pmem[b];
    const FwsArgs<T> &a = args[m];
    if (a.active && *a.active == 0) return;
    pack_factor_cell<T>(a, pack_cell_of(a, b - poff[m], threadIdx.x));
}
}  // namespace

size_t wl_entries(const Grid &g, const WfSync &sync) {
    return (size_t)sync.ntiles * (size_t)wl_nsteps(g.nz) * 64;
}

template <class T>
static void wl_apply(const Grid &g, const WfSync &sync, const WlPrecond<T> &wl, const double *d,
                     double *v, const int *status, hipStream_t s, bool tickets_clear) {
    const size_t entries = wl_entries(g, sync);
    WsArgs<T> f{};
    f.nx = g.nx; f.ny = g.ny; f.nz = g.nz; f.nxny = g.nxny; f.N = g.N;
    f.nsteps = wl_nsteps(g.nz);
    f.wl_bytes_blk = (unsigned)(entries * 4 * sizeof(T));
    f.wl_bytes_vec = (unsigned)(entries * 2 * sizeof(T));
    f.d = d;
    f.blk[0] = wl.L; f.blk[1] = wl.L + entries * 4; f.blk[2] = wl.L + 2 * entries * 4;
    f.Dinv = nullptr; f.w = wl.w; f.out = nullptr; f.status = status; f.sync = sync;
    if (!tickets_clear) hipLaunchKernelGGL(k_reset_ticket_pair, dim3(1), dim3(1), 0, s, sync.ticket);
    hipLaunchKernelGGL((k_sweep_ws<T, true, kSweepPD>), dim3(sync.ntiles), dim3(kSweepThreads), 0, s, f);
    WsArgs<T> b = f;
    b.d = nullptr;
    b.blk[0] = wl.U; b.blk[1] = wl.U + entries * 4; b.blk[2] = wl.U + 2 * entries * 4;
    b.Dinv = wl.D;
    b.out = v;
    b.sync.ticket = sync.ticket + 1;
    hipLaunchKernelGGL((k_sweep_ws<T, false, kSweepPD>), dim3(sync.ntiles), dim3(kSweepThreads), 0, s, b);
}

void wl_apply64(const Grid &g, const WfSync &sync, const WlPrecond<double> &wl, const double *d,
                double *v, const int *status, hipStream_t s, bool tickets_clear) {
    wl_apply<double>(g, sync, wl, d, v, status, s, tickets_clear);
}
void wl_apply32(const Grid &g, const WfSync &sync, const WlPrecond<float> &wl, const double *d,
                double *v, const int *status, hipStream_t s, bool tickets_clear) {
    wl_apply<float>(g, sync, wl, d, v, status, s, tickets_clear);
}

template <class T>
static FwsArgs<T> fws_args(const Grid &g, const WfSync &sync, const double *A, double *Af, T *L,
                           T *U, T *D, int *err) {
    FwsArgs<T> f{};
    f.nx = g.nx; f.ny = g.ny; f.nz = g.nz; f.nxny = g.nxny; f.N = g.N; f.diag = g.diag;
    f.nsteps = wl_nsteps(g.nz);
    f.A = A; f.Af = Af; f.L = L; f.U = U; f.D = D;
    f.entries = wl_entries(g, sync);
    f.err = err;
    f.active = nullptr;
    f.A32 = nullptr;
    f.sync = sync;
    return f;
}

template <class T>
static void wl_factor(const Grid &g, const WfSync &sync, const double *A, double *Af,
                      const WlPrecond<T> &wl, int *err, hipStream_t s, const int *gate,
                      float *A32) {
    FwsArgs<T> f = fws_args<T>(g, sync, A, Af, wl.L, wl.U, wl.D, err);
    f.active = gate;
    f.A32 = A32;
    hipLaunchKernelGGL(k_wl_pack_factor<T>, dim3((unsigned)(sync.ntiles * ((g.nz + 3) / 4))), dim3(256), 0,
                       s, f);
    hipLaunchKernelGGL(k_sweep_reset_ticket, dim3(1), dim3(1), 0, s, sync.ticket);
    hipLaunchKernelGGL((k_factor_ws<T, kFactorPD>), dim3(sync.ntiles), dim3(128), 0, s, f);
}

void wl_factor64(const Grid &g, const WfSync &sync, const double *A, double *Af,
                 const WlPrecond<double> &wl, int *err, hipStream_t s, const int *gate,
                 float *A32) {
    wl_factor<double>(g, sync, A, Af, wl, err, s, gate, A32);
}

void wl_factor32(const Grid &g, const WfSync &sync, const double *A, double *Af,
                 const WlPrecond<float> &wl, int *err, hipStream_t s, const int *gate,
                 float *A32) {
    wl_factor<float>(g, sync, A, Af, wl, err, s, gate, A32);
}

struct WlBatch {
    int n = 0, ntiles = 0;
    bool fp32 = false;
    void *fwd = nullptr, *bwd = nullptr;  // WsArgs<T>[n]
    void *fac = nullptr;                  // FwsArgs<T>[n] (members given a Jacobian)
    int npack = 0;                        // blocks of the operand packing
    int *pmem = nullptr, *poff = nullptr; // packing block -> realization, its first block
    int *tmem = nullptr;                  // block -> realization
    unsigned **tickets = nullptr;         // [n]
};

template <class T>
static int wl_batch_fill(WlBatch &b, const std::vector<WlMember> &m, hipStream_t s) {
    std::vector<WsArgs<T>> f(m.size()), bw(m.size());
    std::vector<FwsArgs<T>> fa(m.size());
    const bool factor = !m.empty() && m[0].A != nullptr;
    std::vector<int> tmem, pmem, poff(m.size());
    std::vector<unsigned *> tk(m.size());
    for (size_t i = 0; i < m.size(); ++i) {
        const Grid &g = m[i].g;
        const size_t entries = wl_entries(g, m[i].sync);
        WsArgs<T> a{};
        a.nx = g.nx; a.ny = g.ny; a.nz = g.nz; a.nxny = g.nxny; a.N = g.N;
        a.nsteps = wl_nsteps(g.nz);
        a.wl_bytes_blk = (unsigned)(entries * 4 * sizeof(T));
        a.wl_bytes_vec = (unsigned)(entries * 2 * sizeof(T));
        a.d = m[i].d;
        const T *L = static_cast<const T *>(m[i].L), *U = static_cast<const T *>(m[i].U);
        a.blk[0] = L; a.blk[1] = L + entries * 4; a.blk[2] = L + 2 * entries * 4;
        a.Dinv = nullptr; a.w = static_cast<T *>(m[i].w); a.out = nullptr;
        a.status = m[i].status; a.sync = m[i].sync;
        a.prec = m[i].prec; a.want = m[i].want;
        f[i] = a;
        WsArgs<T> c = a;
        c.d = nullptr;
        c.blk[0] = U; c.blk[1] = U + entries * 4; c.blk[2] = U + 2 * entries * 4;
        c.Dinv = static_cast<const T *>(m[i].D);
        c.out = m[i].out;
        bw[i] = c;
        if (factor) {
            WfSync fs = m[i].sync;
            fs.err = m[i].err;
            fa[i] = fws_args<T>(g, fs, m[i].A, m[i].Af, const_cast<T *>(L), const_cast<T *>(U),
                                const_cast<T *>(static_cast<const T *>(m[i].D)), m[i].err);
            fa[i].active = m[i].active;
            poff[i] = (int)pmem.size();
            pmem.insert(pmem.end(), (size_t)m[i].sync.ntiles * ((g.nz + 3) / 4), (int)i);
        }
        tmem.insert(tmem.end(), m[i].sync.ntiles, (int)i);
        tk[i] = m[i].sync.ticket;
    }
    b.ntiles = (int)tmem.size();
    const size_t bytes = m.size() * sizeof(WsArgs<T>);
    if (hipMalloc(&b.fwd, bytes) != hipSuccess || hipMalloc(&b.bwd, bytes) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&b.tmem), tmem.size() * sizeof(int)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void **>(&b.tickets), m.size() * sizeof(unsigned *)) != hipSuccess)
        return -1;
    if (factor) {
        const size_t fb = m.size() * sizeof(FwsArgs<T>);
        b.npack = (int)pmem.size();
        if (hipMalloc(&b.fac, fb) != hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&b.pmem), pmem.size() * sizeof(int)) != hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&b.poff), poff.size() * sizeof(int)) != hipSuccess ||
            hipMemcpyAsync(b.fac, fa.data(), fb, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(b.pmem, pmem.data(), pmem.size() * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(b.poff, poff.data(), poff.size() * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess)
            return -1;
    }
    if (hipMemcpyAsync(b.fwd, f.data(), bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(b.bwd, bw.data(), bytes, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(b.tmem, tmem.data(), tmem.size() * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(b.tickets, tk.data(), m.size() * sizeof(unsigned *), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return -1;
    return 0;
}

void *wl_batch_create(const std::vector<WlMember> &members, bool fp32, hipStream_t s) {
    auto *b = new WlBatch();
    b->n = (int)members.size();
    b->fp32 = fp32;
    const int rc = fp32 ? wl_batch_fill<float>(*b, members, s) : wl_batch_fill<double>(*b, members, s);
    if (rc) {
        wl_batch_free(b);
        return nullptr;
    }
    return b;
}

void wl_batch_apply_tiles(void *h, const int *tmem, int ntiles, hipStream_t s) {
    auto *b = static_cast<WlBatch *>(h);
    if (ntiles <= 0) return;
    const dim3 rg((b->n + 255) / 256), rb(256);
    hipLaunchKernelGGL(k_reset_tickets, rg, rb, 0, s, b->tickets, b->n);
    if (b->fp32)
        hipLaunchKernelGGL((k_sweep_ws_batch<float, true, kSweepPD>), dim3(ntiles), dim3(kSweepThreads), 0, s,
                           static_cast<const WsArgs<float> *>(b->fwd), tmem);
    else
        hipLaunchKernelGGL((k_sweep_ws_batch<double, true, kSweepPD>), dim3(ntiles), dim3(kSweepThreads), 0, s,
                           static_cast<const WsArgs<double> *>(b->fwd), tmem);
    hipLaunchKernelGGL(k_reset_tickets, rg, rb, 0, s, b->tickets, b->n);
    if (b->fp32)
        hipLaunchKernelGGL((k_sweep_ws_batch<float, false, kSweepPD>), dim3(ntiles), dim3(kSweepThreads), 0, s,
                           static_cast<const WsArgs<float> *>(b->bwd), tmem);
    else
        hipLaunchKernelGGL((k_sweep_ws_batch<double, false, kSweepPD>), dim3(ntiles), dim3(kSweepThreads), 0, s,
                           static_cast<const WsArgs<double> *>(b->bwd), tmem);
}

void wl_batch_apply(void *h, hipStream_t s) {
    auto *b = static_cast<WlBatch *>(h);
    wl_batch_apply_tiles(h, b->tmem, b->ntiles, s);
}

void wl_batch_factor(void *h, hipStream_t s) {
    auto *b = static_cast<WlBatch *>(h);
    const dim3 rg((b->n + 255) / 256), rb(256);
    if (b->fp32)
        hipLaunchKernelGGL(k_wl_pack_factor_batch<float>, dim3(b->npack), dim3(256), 0, s,
                           static_cast<const FwsArgs<float> *>(b->fac), b->pmem, b->poff);
    else
        hipLaunchKernelGGL(k_wl_pack_factor_batch<double>, dim3(b->npack), dim3(256), 0, s,
                           static_cast<const FwsArgs<double> *>(b->fac), b->pmem, b->poff);
    hipLaunchKernelGGL(k_reset_tickets, rg, rb, 0, s, b->tickets, b->n);
    if (b->fp32)
        hipLaunchKernelGGL((k_factor_ws_batch<float, kFactorPD>), dim3(b->ntiles), dim3(128), 0, s,
                           static_cast<const FwsArgs<float> *>(b->fac), b->tmem);
    else
        hipLaunchKernelGGL((k_factor_ws_batch<double, kFactorPD>), dim3(b->ntiles), dim3(128), 0, s,
                           static_cast<const FwsArgs<double> *>(b->fac), b->tmem);
}

void wl_batch_free(void *h) {
    auto *b = static_cast<WlBatch *>(h);
    if (!b) return;
    void *p[] = {b->fwd, b->bwd, b->fac, b->tmem, b->tickets, b->pmem, b->poff};
    for (void *q : p)
        if (q) (void)hipFree(q);
    delete b;
}

}  // namespace ewg
