// ewg_wavefront.hip -- ILU(0) factorization and triangular sweeps as one persistent
// tile-wavefront launch each (3D structured grids).
//
// The natural-order ILU(0) recurrences have dependencies on the -x, -y, -z neighbours
// (forward sweep, factorization) or on the +x, +y, +z neighbours (backward sweep). In
// "sweep coordinates" (reflected for the backward sweep) every recurrence is a forward
// one. The (x, y) plane is cut into TX x TY tiles; one workgroup owns a tile column
// through all z. Lane (lx, ly) handles z = tau - lx - ly at local step tau, so its -x and
// -y neighbours were computed by lanes (lx-1, ly) and (lx, ly-1) one step earlier
// (exchanged by cross-lane shuffles) and its -z neighbour by itself (kept in registers).
// Values across a tile edge come from the neighbouring tile through small edge buffers
// of 64-bit words (see k_wavefront). Workgroups take tiles by ticket in anti-diagonal
// order, so a workgroup only ever waits for tiles already held by running workgroups:
// no assumption about residency, and every wait is bounded (a spin limit raises the
// error flag and lets the launch drain). No step needs a memory fence, so the
// prefetched loads stay in flight across steps.
//
// Each lane's own operands (matrix blocks, right-hand side) do not depend on the
// recurrence; they are loaded PD steps ahead into a register ring, and the cross-tile
// values are fetched PD steps ahead as well, so the critical path per step is the LDS
// exchange and a few 2x2 block products.
//
// The arithmetic inside each row is the one of the level-scheduled kernels (ascending
// column order), so both produce bit-identical factors and sweeps.
#include "ewg_internal.h"
#include "ewg_wavefront.h"

namespace ewg {

namespace {

constexpr unsigned kSpinLimit = 1u << 22;

template <class T>
struct V2 {
    T a, b;
};
template <class T>
struct V4 {
    T a[4];
};

template <class T>
__device__ __forceinline__ void bmmv2(const T *m, const V2<T> &x, T &y0, T &y1) {
    y0 -= m[0] * x.a;
    y0 -= m[1] * x.b;
    y1 -= m[2] * x.a;
    y1 -= m[3] * x.b;
}
template <class T>
__device__ __forceinline__ void bmul4(const T *a, const T *b, T *out) {
    T r0 = T(0), r1 = T(0), r2 = T(0), r3 = T(0);
    r0 += a[0] * b[0]; r0 += a[1] * b[2];
    r1 += a[0] * b[1]; r1 += a[1] * b[3];
    r2 += a[2] * b[0]; r2 += a[3] * b[2];
    r3 += a[2] * b[1]; r3 += a[3] * b[3];
    out[0] = r0; out[1] = r1; out[2] = r2; out[3] = r3;
}

// ----------------------------------------------------------------------------------
// Branch-free memory access: raw buffer loads/stores with hardware range checking. A
// lane without work uses an out-of-range offset (its load returns 0, its store is
// dropped), so every step issues the same instruction stream and the compiler can keep
// the prefetched loads of the following steps in flight (on gfx9 loads and stores share
// one counter, and a conditional memory operation would make it wait for all of them).
// ----------------------------------------------------------------------------------
using u32x2 = __attribute__((ext_vector_type(2))) unsigned;
using u32x4 = __attribute__((ext_vector_type(4))) unsigned;
constexpr unsigned kNoOff = 0xFFFFFF00u;   // beyond every buffer used here
constexpr int kCoherent = 16;             // sc1: the cache policy of agent-scope atomics

struct Buf {
    __amdgpu_buffer_rsrc_t r;
};
__device__ __forceinline__ Buf make_buf(const void *p, size_t bytes) {
    unsigned n = bytes > 0xFFFFFFF0u ? 0xFFFFFFF0u : static_cast<unsigned>(bytes);
    return Buf{__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, static_cast<int>(n),
                                                 0x00020000)};
}
template <int AUX = 0>
__device__ __forceinline__ u32x4 bl16(const Buf &b, unsigned off) {
    return __builtin_amdgcn_raw_buffer_load_b128(b.r, off, 0, AUX);
}
template <int AUX = 0>
__device__ __forceinline__ u32x2 bl8(const Buf &b, unsigned off) {
    return __builtin_amdgcn_raw_buffer_load_b64(b.r, off, 0, AUX);
}
template <int AUX = 0>
__device__ __forceinline__ void bs16(const Buf &b, unsigned off, u32x4 v) {
    __builtin_amdgcn_raw_buffer_store_b128(v, b.r, off, 0, AUX);
}
template <int AUX = 0>
__device__ __forceinline__ void bs8(const Buf &b, unsigned off, u32x2 v) {
    __builtin_amdgcn_raw_buffer_store_b64(v, b.r, off, 0, AUX);
}

// a 2x2 block of T (row-major) at byte offset off
template <class T>
__device__ __forceinline__ void ldblk(const Buf &b, unsigned off, T *o) {
    if constexpr (sizeof(T) == 8) {
        double2 u = __builtin_bit_cast(double2, bl16(b, off));
        double2 v = __builtin_bit_cast(double2, bl16(b, off == kNoOff ? kNoOff : off + 16));
        o[0] = u.x; o[1] = u.y; o[2] = v.x; o[3] = v.y;
    } else {
        float4 u = __builtin_bit_cast(float4, bl16(b, off));
        o[0] = u.x; o[1] = u.y; o[2] = u.z; o[3] = u.w;
    }
}
// a 2x2 FP64 block converted to T
template <class T>
__device__ __forceinline__ void ldblk64(const Buf &b, unsigned off, T *o) {
    double2 u = __builtin_bit_cast(double2, bl16(b, off));
    double2 v = __builtin_bit_cast(double2, bl16(b, off == kNoOff ? kNoOff : off + 16));
    o[0] = static_cast<T>(u.x); o[1] = static_cast<T>(u.y);
    o[2] = static_cast<T>(v.x); o[3] = static_cast<T>(v.y);
}
template <class T>
__device__ __forceinline__ void stblk(const Buf &b, unsigned off, const T *o) {
    if constexpr (sizeof(T) == 8) {
        bs16(b, off, __builtin_bit_cast(u32x4, make_double2(o[0], o[1])));
        bs16(b, off == kNoOff ? kNoOff : off + 16, __builtin_bit_cast(u32x4, make_double2(o[2], o[3])));
    } else {
        bs16(b, off, __builtin_bit_cast(u32x4, make_float4(o[0], o[1], o[2], o[3])));
    }
}
// a pair of T (one cell of a 2-vector)
template <class T>
__device__ __forceinline__ void ldpair(const Buf &b, unsigned off, T &x, T &y) {
    if constexpr (sizeof(T) == 8) {
        double2 u = __builtin_bit_cast(double2, bl16(b, off));
        x = u.x; y = u.y;
    } else {
        float2 u = __builtin_bit_cast(float2, bl8(b, off));
        x = u.x; y = u.y;
    }
}
template <class T>
__device__ __forceinline__ void stpair(const Buf &b, unsigned off, T x, T y) {
    if constexpr (sizeof(T) == 8) bs16(b, off, __builtin_bit_cast(u32x4, make_double2(x, y)));
    else bs8(b, off, __builtin_bit_cast(u32x2, make_float2(x, y)));
}
__device__ __forceinline__ unsigned offs(bool ok, size_t byteoff) {
    return ok ? static_cast<unsigned>(byteoff) : kNoOff;
}

// geometry of one sweep: sweep coordinate -> cell
struct SweepGeom {
    int nx, ny, nz, nxny;
    int reflect;  // 1 for the backward sweep
    __device__ __forceinline__ int cell(int xs, int ys, int zs) const {
        int x = reflect ? nx - 1 - xs : xs;
        int y = reflect ? ny - 1 - ys : ys;
        int z = reflect ? nz - 1 - zs : zs;
        return x + nx * y + nxny * z;
    }
};

// ----------------------------------------------------------------------------------
// per-cell operations; `ok` is false for lanes without a cell this step
// ----------------------------------------------------------------------------------
// forward sweep: w_i = d_i - sum_{j in -z,-y,-x} L_ij w_j
template <class T>
struct FwdOp {
    using Val = V2<T>;
    struct In {
        T d0, d1;
        T L[3][4];
    };
    struct Bufs {
        Buf L, d, w;
    };
    const T *L;
    const double *d;
    T *w;
    int N;
    __device__ __forceinline__ Bufs bufs() const {
        return Bufs{make_buf(L, (size_t)3 * N * 4 * sizeof(T)), make_buf(d, (size_t)N * 16),
                    make_buf(w, (size_t)N * 2 * sizeof(T))};
    }
    __device__ __forceinline__ void load(const Bufs &b, bool ok, int cell, In &in) const {
        double dx, dy;
        ldpair<double>(b.d, offs(ok, (size_t)cell * 16), dx, dy);
        in.d0 = static_cast<T>(dx);
        in.d1 = static_cast<T>(dy);
#pragma unroll
        for (int s = 0; s < 3; ++s)
            ldblk<T>(b.L, offs(ok, ((size_t)s * N + cell) * 4 * sizeof(T)), in.L[s]);
    }
    __device__ __forceinline__ Val compute(const Bufs &, bool, int, const In &in, bool hz,
                                           bool hy, bool hx, const Val &vz, const Val &vy,
                                           const Val &vx) const {
        T r0 = in.d0, r1 = in.d1;
        if (hz) bmmv2(in.L[0], vz, r0, r1);
        if (hy) bmmv2(in.L[1], vy, r0, r1);
        if (hx) bmmv2(in.L[2], vx, r0, r1);
        return Val{r0, r1};
    }
    __device__ __forceinline__ void store(const Bufs &b, bool ok, int cell, const Val &v) const {
        stpair<T>(b.w, offs(ok, (size_t)cell * 2 * sizeof(T)), v.a, v.b);
    }
};

// backward sweep (sweep coordinates reflected): w_i = inv(D_i)(w_i - sum_{+x,+y,+z} U_ik w_k)
template <class T, class UT>
struct BwdOp {
    using Val = V2<T>;
    struct In {
        T w0, w1;
        T U[3][4];  // +x, +y, +z
        T Di[4];
    };
    struct Bufs {
        Buf U, Dinv, w, out;
    };
    const UT *U;  // slot s of U at (uslot0 + s)
    int uslot0, uslots;
    const T *Dinv;
    T *w;
    double *out;  // null when w is already the FP64 output
    int N;
    __device__ __forceinline__ Bufs bufs() const {
        return Bufs{make_buf(U, (size_t)uslots * N * 4 * sizeof(UT)),
                    make_buf(Dinv, (size_t)N * 4 * sizeof(T)),
                    make_buf(w, (size_t)N * 2 * sizeof(T)),
                    make_buf(out ? static_cast<const void *>(out) : static_cast<const void *>(w),
                             out ? (size_t)N * 16 : 0)};
    }
    __device__ __forceinline__ void load(const Bufs &b, bool ok, int cell, In &in) const {
        ldpair<T>(b.w, offs(ok, (size_t)cell * 2 * sizeof(T)), in.w0, in.w1);
#pragma unroll
        for (int s = 0; s < 3; ++s)
            ldblk<T>(b.U, offs(ok, ((size_t)(uslot0 + s) * N + cell) * 4 * sizeof(UT)), in.U[s]);
        ldblk<T>(b.Dinv, offs(ok, (size_t)cell * 4 * sizeof(T)), in.Di);
    }
    __device__ __forceinline__ Val compute(const Bufs &, bool, int, const In &in, bool hz,
                                           bool hy, bool hx, const Val &vz, const Val &vy,
                                           const Val &vx) const {
        T r0 = in.w0, r1 = in.w1;
        if (hx) bmmv2(in.U[0], vx, r0, r1);
        if (hy) bmmv2(in.U[1], vy, r0, r1);
        if (hz) bmmv2(in.U[2], vz, r0, r1);
        T s0 = T(0), s1 = T(0);
        s0 += in.Di[0] * r0; s0 += in.Di[1] * r1;
        s1 += in.Di[2] * r0; s1 += in.Di[3] * r1;
        return Val{s0, s1};
    }
    __device__ __forceinline__ void store(const Bufs &b, bool ok, int cell, const Val &v) const {
        stpair<T>(b.w, offs(ok, (size_t)cell * 2 * sizeof(T)), v.a, v.b);
        if (out)
            stpair<double>(b.out, offs(ok, (size_t)cell * 16), static_cast<double>(v.a),
                           static_cast<double>(v.b));
    }
};

// factorization: D_i = A_ii - sum_{j in -z,-y,-x} (A_ij Dinv_j) A_ji; value = Dinv_i
template <class T>
struct FactorOp {
    using Val = V4<T>;
    struct In {
        T Aii[4];
        T Aij[3][4];
        T Aji[3][4];
    };
    struct Bufs {
        Buf A, L, Dinv, U;
    };
    const double *A;
    T *L;
    T *Dinv;
    T *U;  // FP32 copy of the upper blocks or null
    int *err;
    int N, diag, nx, nxny;
    __device__ __forceinline__ Bufs bufs() const {
        return Bufs{make_buf(A, (size_t)(2 * diag + 1) * N * 32), make_buf(L, (size_t)3 * N * 4 * sizeof(T)),
                    make_buf(Dinv, (size_t)N * 4 * sizeof(T)),
                    make_buf(U ? static_cast<const void *>(U) : static_cast<const void *>(L),
                             U ? (size_t)3 * N * 4 * sizeof(T) : 0)};
    }
    __device__ __forceinline__ void load(const Bufs &b, bool ok, int cell, In &in) const {
        ldblk64<T>(b.A, offs(ok, ((size_t)diag * N + cell) * 32), in.Aii);
        const int str[3] = {nxny, nx, 1};
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            ldblk64<T>(b.A, offs(ok, ((size_t)s * N + cell) * 32), in.Aij[s]);
            // row j = cell - stride, its upper block towards cell (absent neighbours read
            // out of range or an unused block; the value is only used when present)
            const int j = cell - str[s];
            ldblk64<T>(b.A, offs(ok && j >= 0, ((size_t)(diag + 3 - s) * N + j) * 32), in.Aji[s]);
        }
    }
    __device__ __forceinline__ Val compute(const Bufs &b, bool ok, int cell, const In &in, bool hz,
                                           bool hy, bool hx, const Val &vz, const Val &vy,
                                           const Val &vx) const {
        T D[4] = {in.Aii[0], in.Aii[1], in.Aii[2], in.Aii[3]};
        const bool has[3] = {hz, hy, hx};
        const Val *nb[3] = {&vz, &vy, &vx};
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            T Lij[4] = {T(0), T(0), T(0), T(0)};
            if (has[s]) {
                T Bm[4];
                bmul4(in.Aij[s], nb[s]->a, Lij);
                bmul4(Lij, in.Aji[s], Bm);
                for (int e = 0; e < 4; ++e) D[e] -= Bm[e];
            }
            stblk<T>(b.L, offs(ok, ((size_t)s * N + cell) * 4 * sizeof(T)), Lij);
        }
        T det = D[0] * D[3] - D[1] * D[2];
        if (ok && !(fabs(static_cast<double>(det)) >= 1e-30)) atomicExch(err, 1);
        if (!(fabs(static_cast<double>(det)) >= 1e-30)) det = T(1);
        T det_1 = T(1) / det;
        Val r;
        r.a[0] = D[3] * det_1;
        r.a[1] = -D[1] * det_1;
        r.a[2] = -D[2] * det_1;
        r.a[3] = D[0] * det_1;
        return r;
    }
    __device__ __forceinline__ void store(const Bufs &b, bool ok, int cell, const Val &v) const {
        stblk<T>(b.Dinv, offs(ok, (size_t)cell * 4 * sizeof(T)), v.a);
        if (U) {
#pragma unroll
            for (int s = 0; s < 3; ++s) {
                T Ui[4];
                ldblk64<T>(b.A, offs(ok, ((size_t)(diag + 1 + s) * N + cell) * 32), Ui);
                stblk<T>(b.U, offs(ok, ((size_t)s * N + cell) * 4 * sizeof(T)), Ui);
            }
        }
    }
};

// ----------------------------------------------------------------------------------
// the wavefront driver: one wavefront (TX x TY = 64 lanes) per tile column
// ----------------------------------------------------------------------------------
constexpr unsigned long long kSentinel = 0x7FF4DEAD7FF4DEADull;  // signalling-NaN pattern

template <class Val>
__device__ __forceinline__ Val shfl_val(const Val &v, int src) {
    constexpr int NW = sizeof(Val) / 4;
    const int *w = reinterpret_cast<const int *>(&v);
    Val r;
    int *o = reinterpret_cast<int *>(&r);
#pragma unroll
    for (int k = 0; k < NW; ++k) o[k] = __shfl(w[k], src, 64);
    return r;
}

// Edge words of one exchanged value: W 64-bit words at byte offset off (W <= 4)
template <int W>
struct Words {
    unsigned long long w[W];
};
template <int W>
__device__ __forceinline__ Words<W> ldwords(const Buf &b, unsigned off) {
    Words<W> r;
    if constexpr (W == 1) {
        r.w[0] = __builtin_bit_cast(unsigned long long, bl8<kCoherent>(b, off));
    } else {
#pragma unroll
        for (int k = 0; k < W; k += 2) {
            u32x4 v = bl16<kCoherent>(b, off == kNoOff ? kNoOff : off + 8 * k);
            r.w[k] = (unsigned long long)v.x | ((unsigned long long)v.y << 32);
            r.w[k + 1] = (unsigned long long)v.z | ((unsigned long long)v.w << 32);
        }
    }
    return r;
}
template <int W>
__device__ __forceinline__ void stwords(const Buf &b, unsigned off, const Words<W> &r) {
    if constexpr (W == 1) {
        bs8<kCoherent>(b, off, __builtin_bit_cast(u32x2, r.w[0]));
    } else {
#pragma unroll
        for (int k = 0; k < W; k += 2) {
            u32x4 v;
            v.x = (unsigned)r.w[k];
            v.y = (unsigned)(r.w[k] >> 32);
            v.z = (unsigned)r.w[k + 1];
            v.w = (unsigned)(r.w[k + 1] >> 32);
            bs16<kCoherent>(b, off == kNoOff ? kNoOff : off + 8 * k, v);
        }
    }
}

// Cross-tile values: a tile's last column (lx = TX-1) and last row (ly = TY-1) are
// published per z in edge buffers of 64-bit words with coherent (sc1) stores. The single
// consumer (the neighbouring tile) reads them with coherent loads, polls until no word
// is the sentinel, then writes the sentinel back, so the buffers are clean for the next
// launch. Every word is written at most once per launch, so no ordering between
// locations (and no fence) is needed.
template <class Op, int TX, int TY, int PD>
__global__ void __launch_bounds__(64)
k_wavefront(SweepGeom sg, Op op, WfSync sync, const int *status) {
    static_assert(TX * TY == 64, "one wavefront per tile");
    using Val = typename Op::Val;
    using In = typename Op::In;
    constexpr int W = sizeof(Val) / 8;
    static_assert(W >= 1 && W <= kWfMaxWords, "edge word count");

    if (status && *status != KS_RUNNING) return;
    const int lane = threadIdx.x;
    unsigned tk = 0;
    if (lane == 0) tk = atomicAdd(sync.ticket, 1u);
    tk = __shfl(tk, 0, 64);
    if (tk >= (unsigned)sync.ntiles) return;
    const int2 tile = sync.order[tk];
    const int tx = tile.x, ty = tile.y;
    const int lx = lane % TX, ly = lane / TX;
    const int xs = tx * TX + lx, ys = ty * TY + ly;
    const bool inplane = xs < sg.nx && ys < sg.ny;
    const int nz = sg.nz;
    const int tileId = ty * sync.ntx + tx;
    const bool consX = inplane && lx == 0 && tx > 0;
    const bool consY = inplane && ly == 0 && ty > 0;
    const bool prodX = inplane && lx == TX - 1 && tx + 1 < sync.ntx;
    const bool prodY = inplane && ly == TY - 1 && ty + 1 < sync.nty;
    // edge buffers ex[tile][ly][z][kWfMaxWords], ey[tile][lx][z][kWfMaxWords] (bytes)
    const Buf ebx = make_buf(sync.ex, sync.edge_words * 8);
    const Buf eby = make_buf(sync.ey, sync.edge_words * 8);
    const size_t zw = (size_t)kWfMaxWords * 8;
    const size_t exIn = consX ? ((size_t)(tileId - 1) * TY + ly) * nz * zw : 0;
    const size_t eyIn = consY ? ((size_t)(tileId - sync.ntx) * TX + lx) * nz * zw : 0;
    const size_t exOut = prodX ? ((size_t)tileId * TY + ly) * nz * zw : 0;
    const size_t eyOut = prodY ? ((size_t)tileId * TX + lx) * nz * zw : 0;
    const int nsteps = (TX - 1) + (TY - 1) + nz;
    const typename Op::Bufs bufs = op.bufs();

    In ring[PD];
    Words<W> bx[PD], by[PD];
    Val prev = Val{}, fromX = Val{}, fromY = Val{};

    auto issue = [&](int tau, int k) {
        const int z = tau - lx - ly;
        const bool ok = inplane && z >= 0 && z < nz;
        const int cell = ok ? sg.cell(xs, ys, z) : 0;
        op.load(bufs, ok, cell, ring[k]);
        bx[k] = ldwords<W>(ebx, offs(ok && consX, exIn + (size_t)z * zw));
        by[k] = ldwords<W>(eby, offs(ok && consY, eyIn + (size_t)z * zw));
    };
    // the prefetched edge value of this step, once published; its words are recycled
    auto take = [&](const Buf &eb, size_t off, Words<W> wd, Val &out) {
        bool ready = true;
#pragma unroll
        for (int w = 0; w < W; ++w) ready &= wd.w[w] != kSentinel;
        if (!ready) {
            unsigned spins = 0;
            while (!ready && spins < kSpinLimit) {
                __builtin_amdgcn_s_sleep(1);
                wd = ldwords<W>(eb, static_cast<unsigned>(off));
                ready = true;
#pragma unroll
                for (int w = 0; w < W; ++w) ready &= wd.w[w] != kSentinel;
                ++spins;
                // another wait already gave up: stop waiting, the launch is failed anyway
                if ((spins & 63) == 0 &&
                    __hip_atomic_load(sync.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    break;
            }
            if (!ready) atomicExch(sync.err, 1);
        }
        __builtin_memcpy(&out, wd.w, sizeof(Val));
    };
    Words<W> sentinel;
#pragma unroll
    for (int w = 0; w < W; ++w) sentinel.w[w] = kSentinel;

#pragma unroll
    for (int k = 0; k < PD; ++k) issue(k, k);
    for (int tau0 = 0; tau0 < nsteps; tau0 += PD) {
#pragma unroll
        for (int k = 0; k < PD; ++k) {
            const int tau = tau0 + k;
            const int z = tau - lx - ly;
            const bool active = inplane && tau < nsteps && z >= 0 && z < nz;
            const In cur = ring[k];
            const Words<W> cx = bx[k], cy = by[k];
            if (tau + PD < nsteps) issue(tau + PD, k);
            Val vx = fromX, vy = fromY;
            if (active && consX) take(ebx, exIn + (size_t)z * zw, cx, vx);
            if (active && consY) take(eby, eyIn + (size_t)z * zw, cy, vy);
            stwords<W>(ebx, offs(active && consX, exIn + (size_t)z * zw), sentinel);
            stwords<W>(eby, offs(active && consY, eyIn + (size_t)z * zw), sentinel);
            const int cell = active ? sg.cell(xs, ys, z) : 0;
            Val res = op.compute(bufs, active, cell, cur, z > 0, ys > 0, xs > 0, prev, vy, vx);
            op.store(bufs, active, cell, res);
            Words<W> rw;
            __builtin_memcpy(rw.w, &res, sizeof(Val));
            stwords<W>(ebx, offs(active && prodX, exOut + (size_t)z * zw), rw);
            stwords<W>(eby, offs(active && prodY, eyOut + (size_t)z * zw), rw);
            if (!active) res = prev;
            prev = res;
            fromX = shfl_val(res, lane - 1);
            fromY = shfl_val(res, lane - TX);
        }
    }
}

__global__ void k_wf_reset_ticket(unsigned *ticket) { *ticket = 0u; }

__global__ void k_wf_fill(unsigned long long *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = kSentinel;
}

template <class Op, int PD = kWfPD>
void launch(const Grid &g, int reflect, const Op &op, const WfSync &sync, const int *status,
            hipStream_t s) {
    SweepGeom sg{g.nx, g.ny, g.nz, g.nxny, reflect};
    hipLaunchKernelGGL(k_wf_reset_ticket, dim3(1), dim3(1), 0, s, sync.ticket);
    hipLaunchKernelGGL((k_wavefront<Op, kWfTX, kWfTY, PD>), dim3(sync.ntiles), dim3(64), 0, s,
                       sg, op, sync, status);
}

}  // namespace

// ------------------------------------------------------------------------------------
void wf_init_edges(const WfSync &sync, hipStream_t s) {
    hipLaunchKernelGGL(k_wf_fill, dim3(256), dim3(256), 0, s, sync.ex, sync.edge_words);
    hipLaunchKernelGGL(k_wf_fill, dim3(256), dim3(256), 0, s, sync.ey, sync.edge_words);
}

void wf_factor64(const Grid &g, const double *A, double *L, double *Dinv, int *err,
                 const WfSync &sync, hipStream_t s) {
    FactorOp<double> op{A, L, Dinv, nullptr, err, g.N, g.diag, g.nx, g.nxny};
    launch<FactorOp<double>, kWfPDFactor>(g, 0, op, sync, nullptr, s);
}
void wf_apply64(const Grid &g, const double *L, const double *Dinv, const double *A,
                const double *d, double *v, const int *status, const WfSync &sync,
                hipStream_t s) {
    FwdOp<double> f{L, d, v, g.N};
    launch(g, 0, f, sync, status, s);
    BwdOp<double, double> b{A, g.diag + 1, 2 * g.diag + 1, Dinv, v, nullptr, g.N};
    launch(g, 1, b, sync, status, s);
}
void wf_apply32(const Grid &g, const float *L, const float *Dinv, const float *U,
                const double *d, float *w, double *v, const int *status, const WfSync &sync,
                hipStream_t s) {
    FwdOp<float> f{L, d, w, g.N};
    launch(g, 0, f, sync, status, s);
    BwdOp<float, float> b{U, 0, 3, Dinv, w, v, g.N};
    launch(g, 1, b, sync, status, s);
}

}  // namespace ewg
