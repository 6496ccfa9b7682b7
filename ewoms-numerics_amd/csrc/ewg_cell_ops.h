// ewg_cell_ops.h -- per-cell steps of the block ILU(0) and the SpMV on the structured
// pattern, shared by the grid-wide kernels (ewg_linalg.hip) and the ensemble solver
// (ewg_batch.hip), so that both do the same arithmetic in the same order.
//
// ILU(0) in natural (lexicographic) order, as dune-istl's blockILU0Decomposition computes
// it: L_ij = A_ij inv(U_jj), A_ik -= L_ij A_jk on the pattern, pivots stored inverted. For
// nx, ny >= 3 the only pattern entry two lower/upper offsets can reach is the diagonal,
// so U_ij = A_ij (j > i) and D_i = A_ii - sum_{j in {i-nxny, i-nx, i-1}} L_ij A_ji
// (ascending j).
#pragma once

#include "ewg_internal.h"

namespace ewg {
namespace cell {

// The products and sums of dune's dense loops (each entry: a_i0 b_0j, then + a_i1 b_1j)
// without their initial zero: 0 + x equals x except that it turns -0 into +0, so the
// results are the same numbers, the sign of a zero aside.
template <class T>
__device__ __forceinline__ void bmul(const T *a, const T *b, T *out) {
    T r0 = a[0] * b[0], r1 = a[0] * b[1], r2 = a[2] * b[0], r3 = a[2] * b[1];
    r0 += a[1] * b[2];
    r1 += a[1] * b[3];
    r2 += a[3] * b[2];
    r3 += a[3] * b[3];
    out[0] = r0; out[1] = r1; out[2] = r2; out[3] = r3;
}

template <class T>
__device__ __forceinline__ void bmmv(const T *a, const T *x, T *y) {  // y -= a x
    y[0] -= a[0] * x[0];
    y[0] -= a[1] * x[1];
    y[1] -= a[2] * x[0];
    y[1] -= a[3] * x[1];
}

template <class T>
__device__ __forceinline__ void bmv(const T *a, const T *x, T *y) {  // y = a x (as bmul)
    T s0 = a[0] * x[0], s1 = a[2] * x[0];
    s0 += a[1] * x[1];
    s1 += a[3] * x[1];
    y[0] = s0; y[1] = s1;
}

template <class T>
__device__ __forceinline__ void load_blk(const T *base, size_t idx, T *o) {
    const T *p = base + idx * 4;
    o[0] = p[0]; o[1] = p[1]; o[2] = p[2]; o[3] = p[3];
}
template <class T>
__device__ __forceinline__ void load_blk_cast(const double *base, size_t idx, T *o) {
    const double *p = base + idx * 4;
    o[0] = static_cast<T>(p[0]); o[1] = static_cast<T>(p[1]);
    o[2] = static_cast<T>(p[2]); o[3] = static_cast<T>(p[3]);
}
template <class T>
__device__ __forceinline__ void store_blk(T *base, size_t idx, const T *o) {
    T *p = base + idx * 4;
    p[0] = o[0]; p[1] = o[1]; p[2] = o[2]; p[3] = o[3];
}

// map thread -> cell of hyperplane `level`; returns -1 if the thread has no cell
__device__ __forceinline__ int level_cell(const Grid &g, int level, int t, int ijk[3]) {
    int y, z;
    if (g.dim == 3) {
        if (t >= g.ny * g.nz) return -1;
        y = t % g.ny;
        z = t / g.ny;
    } else {
        if (t >= g.ny) return -1;
        y = t;
        z = 0;
    }
    int x = level - y - z;
    if (x < 0 || x >= g.nx) return -1;
    ijk[0] = x; ijk[1] = y; ijk[2] = z;
    return x + g.nx * y + g.nxny * z;
}

__device__ __forceinline__ int lower_axis(const Grid &g, int s) { return g.dim - 1 - s; }
__device__ __forceinline__ int axis_stride(const Grid &g, int a) {
    return a == 0 ? 1 : (a == 1 ? g.nx : g.nxny);
}

// factor row i: L_ij for the lower slots, Dinv_i; U (if given) gets the row's upper blocks
// of A in T. Returns false for a (numerically) singular pivot, which is replaced by 1.
template <class T>
__device__ __forceinline__ bool ilu_factor_cell(const Grid &g, int i, const int ijk[3],
                                                const double *A, T *L, T *Dinv, T *U) {
    T D[4];
    load_blk_cast<T>(A, (size_t)g.diag * g.N + i, D);
    for (int s = 0; s < g.dim; ++s) {  // lower slots, ascending column
        int ax = lower_axis(g, s);
        T Lij[4] = {T(0), T(0), T(0), T(0)};
        if (ijk[ax] > 0) {
            int j = i - axis_stride(g, ax);
            T Aij[4], Dj[4], Aji[4], Bm[4];
            load_blk_cast<T>(A, (size_t)s * g.N + i, Aij);
            load_blk<T>(Dinv, j, Dj);
            bmul(Aij, Dj, Lij);
            load_blk_cast<T>(A, (size_t)(g.diag + 1 + ax) * g.N + j, Aji);
            bmul(Lij, Aji, Bm);
            for (int e = 0; e < 4; ++e) D[e] -= Bm[e];
        }
        store_blk(L, (size_t)s * g.N + i, Lij);
    }
    T det = D[0] * D[3] - D[1] * D[2];
    bool ok = true;
    if (!(fabs(static_cast<double>(det)) >= 1e-30)) {
        ok = false;
        det = T(1);
    }
    T det_1 = T(1) / det;
    T inv[4] = {D[3] * det_1, -D[1] * det_1, -D[2] * det_1, D[0] * det_1};
    store_blk(Dinv, i, inv);
    if (U) {
        for (int s = 0; s < g.dim; ++s) {
            T Ui[4];
            load_blk_cast<T>(A, (size_t)(g.diag + 1 + s) * g.N + i, Ui);
            store_blk(U, (size_t)s * g.N + i, Ui);
        }
    }
    return ok;
}

// forward sweep at row i: w_i = d_i - sum_j L_ij w_j
template <class T>
__device__ __forceinline__ void ilu_fwd_cell(const Grid &g, int i, const int ijk[3], const T *L,
                                             const double *d, T *w) {
    T rhs[2] = {static_cast<T>(d[2 * i]), static_cast<T>(d[2 * i + 1])};
    for (int s = 0; s < g.dim; ++s) {
        int ax = lower_axis(g, s);
        if (ijk[ax] == 0) continue;
        int j = i - axis_stride(g, ax);
        T Lij[4];
        load_blk<T>(L, (size_t)s * g.N + i, Lij);
        T wj[2] = {w[2 * j], w[2 * j + 1]};
        bmmv(Lij, wj, rhs);
    }
    w[2 * i] = rhs[0];
    w[2 * i + 1] = rhs[1];
}

// backward sweep at row i: w_i = inv(D_i) (w_i - sum_k U_ik w_k), out_i = (double) w_i
template <class T, class UT>
__device__ __forceinline__ void ilu_bwd_cell(const Grid &g, int i, const int ijk[3], const UT *U,
                                             int uslot0, const T *Dinv, T *w, double *out) {
    int n[3] = {g.nx, g.ny, g.nz};
    T rhs[2] = {w[2 * i], w[2 * i + 1]};
    for (int s = 0; s < g.dim; ++s) {  // upper slots ascending: +x, +y, +z
        if (ijk[s] + 1 >= n[s]) continue;
        int k = i + axis_stride(g, s);
        const UT *up = U + ((size_t)(uslot0 + s) * g.N + i) * 4;
        T Uik[4] = {static_cast<T>(up[0]), static_cast<T>(up[1]), static_cast<T>(up[2]),
                    static_cast<T>(up[3])};
        T wk[2] = {w[2 * k], w[2 * k + 1]};
        bmmv(Uik, wk, rhs);
    }
    T Di[4];
    load_blk<T>(Dinv, i, Di);
    T res[2];
    bmv(Di, rhs, res);
    w[2 * i] = res[0];
    w[2 * i + 1] = res[1];
    if (out != reinterpret_cast<double *>(w)) {
        out[2 * i] = static_cast<double>(res[0]);
        out[2 * i + 1] = static_cast<double>(res[1]);
    }
}

// one 2x2 block of the slot-major Jacobian as two FP64 rows: stored in FP64 (two 16-byte
// loads) or rounded to FP32 (one 16-byte load, widened exactly)
__device__ __forceinline__ void load_row_blk(const double *A, size_t idx, double2 &r0,
                                             double2 &r1) {
    typedef double blk2 __attribute__((ext_vector_type(2)));
    const blk2 *blk = reinterpret_cast<const blk2 *>(A + idx * 4);
    const blk2 q0 = __builtin_nontemporal_load(blk), q1 = __builtin_nontemporal_load(blk + 1);
    r0 = make_double2(q0.x, q0.y);
    r1 = make_double2(q1.x, q1.y);
}
__device__ __forceinline__ void load_row_blk(const float *A, size_t idx, double2 &r0,
                                             double2 &r1) {
    typedef float blk4 __attribute__((ext_vector_type(4)));
    const blk4 q = __builtin_nontemporal_load(reinterpret_cast<const blk4 *>(A + idx * 4));
    r0 = make_double2(static_cast<double>(q.x), static_cast<double>(q.y));
    r1 = make_double2(static_cast<double>(q.z), static_cast<double>(q.w));
}

// (A x)_i on the slot-major structured pattern. The Jacobian's blocks are read once per
// product with non-temporal loads: streamed that way they do not displace the
// preconditioner's factors from the last-level cache between the BiCGStab iterations'
// preconditioner applies (the factors then come from that cache instead of HBM). TB: the
// type the blocks are stored in (the FP32 copy of the mixed-precision operator is widened
// to FP64 on load; products and sums are FP64 either way).
template <class TB>
__device__ __forceinline__ double2 spmv_cell(const Grid &g, int i, const TB *A,
                                             const double *x) {
    int ijk[3];
    ijk[0] = i % g.nx;
    int r = i / g.nx;
    ijk[1] = r % g.ny;
    ijk[2] = r / g.ny;
    int n[3] = {g.nx, g.ny, g.nz};
    double s0 = 0.0, s1 = 0.0;
    for (int slot = 0; slot < g.nslots; ++slot) {
        int off;
        if (slot == g.diag) {
            off = 0;
        } else if (slot < g.diag) {
            int ax = g.dim - 1 - slot;
            if (ijk[ax] == 0) continue;
            off = -(ax == 0 ? 1 : (ax == 1 ? g.nx : g.nxny));
        } else {
            int ax = slot - g.diag - 1;
            if (ijk[ax] + 1 >= n[ax]) continue;
            off = (ax == 0 ? 1 : (ax == 1 ? g.nx : g.nxny));
        }
        double2 r0, r1;
        load_row_blk(A, (size_t)slot * g.N + i, r0, r1);
        double2 xv = reinterpret_cast<const double2 *>(x)[i + off];
        s0 += r0.x * xv.x;
        s0 += r0.y * xv.y;
        s1 += r1.x * xv.x;
        s1 += r1.y * xv.y;
    }
    return make_double2(s0, s1);
}

}  // namespace cell
}  // namespace ewg
