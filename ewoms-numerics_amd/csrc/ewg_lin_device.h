// ewg_lin_device.h -- per-cell device code of the ECFV forward-AD linearization, shared
// by the grid-wide kernels (ewg_linearize.hip) and the ensemble solver (ewg_batch.hip).
//
// One thread per cell evaluates the element context of its cell exactly as
// FvBaseAdLocalLinearizer::linearize does (fvbaseadlocallinearizer.hh:178-204): the cell
// is the focus DOF, its (p_w, S_w) carry the two derivative slots, neighbour quantities
// enter as values only. The thread writes
//   R[cell]                         value of the focus residual,
//   J[diag][cell]                   derivative of the focus residual,
//   J[mirror(dir)][neighbour]       derivative of the neighbour's residual w.r.t. the
//                                   focus PVs, i.e. column `cell` of row `neighbour`
//                                   (fvbaselinearizer.hh:495-508),
// so every block is written by exactly one thread (no atomics, bitwise reproducible),
// and with the slot-major layout each of the 7 block stores is a unit-stride stream.
//
// The intensive quantities are evaluated once per cell by a first pass (k_iq, with the
// cell's two derivative slots) and stored; the flux pass reads the neighbours' values
// (the value part of the dual evaluation is the plain evaluation, operation for
// operation) instead of evaluating the material laws of all six neighbours again.
//
// The FP32 flux form (LinArgs::flux_fp32) evaluates each face flux and its derivatives in
// float duals; the pressure differences are taken in double before the narrowing, and the
// residual, the Jacobian blocks and the storage term stay double.
#pragma once

#include "ewg_internal.h"

namespace ewg {
namespace lin {


template <class T>
struct IQ {
    T p[2];
    T Sw;
    T lam[2];
};

__device__ __forceinline__ void cell_coords(const Grid &g, int i, int ijk[3]) {
    ijk[0] = i % g.nx;
    int r = i / g.nx;
    ijk[1] = r % g.ny;
    ijk[2] = r / g.ny;
}

__device__ __forceinline__ int stride_of(const Grid &g, int axis) {
    return axis == 0 ? 1 : (axis == 1 ? g.nx : g.nxny);
}

__device__ __forceinline__ bool has_neighbour(const Grid &g, const int ijk[3], int dir) {
    int a = dir >> 1;
    int n = a == 0 ? g.nx : (a == 1 ? g.ny : g.nz);
    return (dir & 1) ? (ijk[a] + 1 < n) : (ijk[a] > 0);
}

__device__ __forceinline__ int slot_of_dir(const Grid &g, int dir) {
    int a = dir >> 1;
    return (dir & 1) ? g.diag + 1 + a : g.diag - 1 - a;
}

__device__ __forceinline__ void centre(const Grid &g, const int ijk[3], double pos[3]) {
    pos[0] = (ijk[0] + 0.5) * g.h[0];
    pos[1] = (ijk[1] + 0.5) * g.h[1];
    pos[2] = (ijk[2] + 0.5) * g.h[2];
}

__device__ __forceinline__ void face_centre(const Grid &g, const int ijk[3], int dir,
                                            double pos[3]) {
    centre(g, ijk, pos);
    int a = dir >> 1;
    pos[a] = ((dir & 1) ? (ijk[a] + 1) : ijk[a]) * g.h[a];
}

template <int DIM>
__device__ __forceinline__ double dot_g(const LinArgs &a, const double v[3]) {
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) s += a.grav[d] * v[d];
    return s;
}

__device__ __forceinline__ const MatReg &region_of(const LinArgs &a, int cell) {
    return a.regs[a.region ? a.region[cell] : 0];
}

// Scalar type S of the flux arithmetic: double, or float for the FP32 flux form, where the
// pressure differences are still taken in double (they cancel values of order 1e5 Pa) and
// the face fluxes are accumulated into the double residual and Jacobian.
template <class S>
__device__ __forceinline__ Dual<S> cst(double x) {
    return Dual<S>{static_cast<S>(x), S(0), S(0)};
}
template <class S>
__device__ __forceinline__ Dual<S> narrow(const Ev &x);
template <>
__device__ __forceinline__ Ev narrow<double>(const Ev &x) {
    return x;
}
template <>
__device__ __forceinline__ Evf narrow<float>(const Ev &x) {
    return to_f32(x);
}
__device__ __forceinline__ Ev widen(const Ev &x) { return x; }
__device__ __forceinline__ Ev widen(const Evf &x) { return to_f64(x); }

// x / d where d is usually a power of two (grid spacings, cell volumes): then x times the
// exact reciprocal, which rounds the same real number once as the division does (so the
// bits are the division's); otherwise the division. d is uniform over the launch, so the
// branch does not diverge.
__device__ __forceinline__ bool pow2_recip(double d, double &r) {
    const unsigned long long b = __double_as_longlong(d);
    const unsigned e = (unsigned)(b >> 52) & 0x7FFu;
    if ((b & 0xFFFFFFFFFFFFFull) != 0ull || e < 1u || e > 2045u) return false;
    r = __longlong_as_double((b & 0x8000000000000000ull) | ((unsigned long long)(2046u - e) << 52));
    return true;
}
__device__ __forceinline__ bool pow2_recip(float d, float &r) {
    const unsigned b = __float_as_uint(d);
    const unsigned e = (b >> 23) & 0xFFu;
    if ((b & 0x7FFFFFu) != 0u || e < 1u || e > 253u) return false;
    r = __uint_as_float((b & 0x80000000u) | ((254u - e) << 23));
    return true;
}
template <class T>
__device__ __forceinline__ T div_p2(T x, T d) {
    T r;
    if (pow2_recip(d, r)) return x * r;
    return x / d;
}
template <class T>
__device__ __forceinline__ Dual<T> div_p2(const Dual<T> &x, T d) {
    T r;
    if (pow2_recip(d, r)) return Dual<T>{x.v * r, x.d0 * r, x.d1 * r};
    return x / d;
}

// ImmiscibleIntensiveQuantities::update (immiscibleintensivequantities.hh:87-145)
template <class T>
__device__ __forceinline__ IQ<T> intensive(const LinArgs &a, int cell, T p0, T Sw) {
    const MatReg &r = region_of(a, cell);
    double pe = a.pe ? a.pe[cell] : 0.0;
    IQ<T> q;
    T pc = pcnw(r, pe, Sw);
    T krw, krn;
    relperm(r, Sw, krw, krn);
    q.p[0] = p0;
    q.p[1] = p0 + pc;
    q.Sw = Sw;
    q.lam[0] = krw / lift(krw, a.mu[0]);
    q.lam[1] = krn / lift(krn, a.mu[1]);
    return q;
}

// M_k = phi S_k rho_k (immisciblelocalresidual.hh:69-103)
template <class T>
__device__ __forceinline__ void storage(const LinArgs &a, double phi, T Sw, T M[2]) {
    M[0] = (phi * Sw) * a.rho[0];
    M[1] = (phi * (1.0 - Sw)) * a.rho[1];
}

// Interior face flux from the focus side: fvbasegradientcalculator.hh:195-240,
// darcyfluxmodule.hh:185-324, :461-527, immisciblelocalresidual.hh:123-149.
//
// The reference evaluates the gradient, the Darcy velocity K grad and the normal flux as
// dim-vectors with a permeability tensor. On this structured Cartesian grid the face
// normal and the centre-to-centre vector lie on the face's axis and the permeability is
// K I, so every other component is a sum of products with an exact zero: only the axis
// component is evaluated here, with the same operations in the same order. Its nonzero
// values are therefore the reference's bit for bit (dropping "+ (+-0)" terms can change
// only the sign of a zero result, which no comparison or nonzero value depends on).
template <int DIM, class S>
__device__ __forceinline__ FaceGeo<S> face_geo(const LinArgs &a, const int ijk[3], int dir) {
    using D = Dual<S>;
    const Grid &g = a.g;
    constexpr int dim = DIM;
    const int ax = dir >> 1;
    int q[3] = {ijk[0], ijk[1], ijk[2]};
    q[ax] += (dir & 1) ? 1 : -1;
    double posIn[3], posEx[3], posFace[3];
    centre(g, ijk, posIn);
    centre(g, q, posEx);
    face_centre(g, ijk, dir, posFace);
    double distSq = 0.0;
    for (int d = 0; d < dim; ++d) {
        double t = posEx[d] - posIn[d];
        distSq += t * t;
    }
    double dvIn[3], dvEx[3], dvTot[3];
    for (int d = 0; d < 3; ++d) {
        dvIn[d] = posIn[d] - posFace[d];
        dvEx[d] = posEx[d] - posFace[d];
        dvTot[d] = posEx[d] - posIn[d];
    }
    double absDistTotSq = 0.0;
    for (int d = 0; d < dim; ++d) absDistTotSq += dvTot[d] * dvTot[d];
    double gIn = dot_g<DIM>(a, dvIn), gEx = dot_g<DIM>(a, dvEx);
    FaceGeo<S> fg;
    fg.gscale = static_cast<S>(div_p2(posEx[ax] - posIn[ax], distSq));
    for (int k = 0; k < 2; ++k) {
        const D pStatIn = -cst<S>(a.rho[k]) * static_cast<S>(gIn);
        const D pStatEx = cst<S>(-a.rho[k] * gEx);
        const D qq = div_p2(pStatEx - pStatIn, static_cast<S>(absDistTotSq));
        fg.cterm[k] = cst<S>(dvTot[ax]) * qq;
    }
    return fg;
}

template <class S>
__device__ __forceinline__ const FaceGeo<S> &face_geo_of(const FaceGeoSet *fs, int dir);
template <>
__device__ __forceinline__ const FaceGeo<double> &face_geo_of<double>(const FaceGeoSet *fs, int dir) {
    return fs->d[dir];
}
template <>
__device__ __forceinline__ const FaceGeo<float> &face_geo_of<float>(const FaceGeoSet *fs, int dir) {
    return fs->f[dir];
}

template <int DIM, class S>
__device__ __forceinline__ void interior_flux_geo(const LinArgs &a, const IQ<Ev> &in,
                                                  const IQ<double> &ex, double Kin, double Kex,
                                                  const FaceGeo<S> &fg, int dir, Ev F[2]) {
    using D = Dual<S>;
    const S normal = (dir & 1) ? S(1) : S(-1);  // the axis component of the unit normal
    const S Kf = static_cast<S>(harmonic_mean(Kin, Kex));
    for (int k = 0; k < 2; ++k) {
        // gradient along the axis: two-point difference plus the hydrostatic correction
        const D deltay = narrow<S>(ev(ex.p[k]) - in.p[k]);
        D grad = deltay * fg.gscale;
        grad = grad + fg.cterm[k];
        // upstream mobility by the sign of grad . n, Darcy velocity -lambda K grad
        const bool upIsEx = grad.v * normal > S(0);
        const D lam = upIsEx ? cst<S>(ex.lam[k]) : narrow<S>(in.lam[k]);
        D v = Kf * grad;
        v = v * (-lam);
        const D vol = v * normal;
        F[k] = widen(cst<S>(0.0) + vol * cst<S>(a.rho[k]));
    }
}

// Interior face flux from the focus side: fvbasegradientcalculator.hh:195-240,
// darcyfluxmodule.hh:185-324, :461-527, immisciblelocalresidual.hh:123-149.
//
// The reference evaluates the gradient, the Darcy velocity K grad and the normal flux as
// dim-vectors with a permeability tensor. On this structured Cartesian grid the face
// normal and the centre-to-centre vector lie on the face's axis and the permeability is
// K I, so every other component is a sum of products with an exact zero: only the axis
// component is evaluated here, with the same operations in the same order. Its nonzero
// values are therefore the reference's bit for bit (dropping "+ (+-0)" terms can change
// only the sign of a zero result, which no comparison or nonzero value depends on). The
// geometry (face_geo) and the phases' fluxes (interior_flux_geo) are split so that the
// geometry can be evaluated once per direction where it is the same for every face.
template <int DIM, class S>
__device__ __forceinline__ void interior_flux(const LinArgs &a, const IQ<Ev> &in,
                                              const IQ<double> &ex, double Kin, double Kex,
                                              const int ijk[3], int dir, Ev F[2]) {
    interior_flux_geo<DIM, S>(a, in, ex, Kin, Kex, face_geo<DIM, S>(a, ijk, dir), dir, F);
}

// Dirichlet free-flow boundary: darcyfluxmodule.hh:332-453, :488-509,
// fvbasegradientcalculator.hh:278-312, immiscibleboundaryratevector.hh:91-153. As for the
// interior faces, only the axis component of the gradient and velocity is evaluated
// (the others are sums of products with exact zeros). The boundary state's capillary
// pressure and mobilities come from a.bc_aux when it is set (computed once per boundary
// face by bc_aux_cell with the same functions), else they are evaluated here.
template <int DIM, class S>
__device__ __forceinline__ void freeflow_flux(const LinArgs &a, const IQ<Ev> &in, int cell,
                                              double Kc, const int ijk[3], int dir,
                                              double pwB, double SwB, long long bf, Ev B[2]) {
    using D = Dual<S>;
    const Grid &g = a.g;
    constexpr int dim = DIM;
    const int ax = dir >> 1;
    double pcB, mobB[2];
    if (a.bc_aux) {
        pcB = a.bc_aux[3 * bf];
        mobB[0] = a.bc_aux[3 * bf + 1];
        mobB[1] = a.bc_aux[3 * bf + 2];
    } else {
        const MatReg &r = region_of(a, cell);
        const double pe = a.pe ? a.pe[cell] : 0.0;
        pcB = pcnw(r, pe, SwB);
        double krB[2];
        relperm(r, SwB, krB[0], krB[1]);
        mobB[0] = krB[0] / a.mu[0];
        mobB[1] = krB[1] / a.mu[1];
    }
    double pB[2] = {pwB, (pwB + pcB) - 0.0};
    double posIn[3], posFace[3];
    centre(g, ijk, posIn);
    face_centre(g, ijk, dir, posFace);
    const S normal = (dir & 1) ? S(1) : S(-1);
    double distSq = 0.0;
    for (int d = 0; d < dim; ++d) {
        double t = posFace[d] - posIn[d];
        distSq += t * t;
    }
    double dvIn[3];
    for (int d = 0; d < 3; ++d) dvIn[d] = posIn[d] - posFace[d];
    double absDistSq = 0.0;
    for (int d = 0; d < dim; ++d) absDistSq += dvIn[d] * dvIn[d];
    double gTimesDist = dot_g<DIM>(a, dvIn);
    const S gscale = static_cast<S>(div_p2(posFace[ax] - posIn[ax], distSq));
    const S Ks = static_cast<S>(Kc);
    for (int k = 0; k < 2; ++k) {
        const D deltay = narrow<S>(pB[k] - in.p[k]);
        D grad = deltay * gscale;
        const D pStatIn = static_cast<S>(-gTimesDist) * cst<S>(a.rho[k]);
        const D qq = div_p2(pStatIn, static_cast<S>(absDistSq));
        grad = grad + cst<S>(dvIn[ax]) * qq;
        const bool upIsBoundary = grad.v * normal > S(0);
        const D mob = upIsBoundary ? cst<S>(mobB[k]) : narrow<S>(in.lam[k]);
        D v = Ks * grad;
        v = v * (-mob);
        const D vol = v * normal;
        // density by pressure comparison: both candidates are the same constant here
        double dens = a.rho[k];
        B[k] = widen(cst<S>(0.0) + vol * cst<S>(dens));
    }
}

// the material-law values of the free-flow boundary states next to cell i (per face)
__device__ __forceinline__ void bc_aux_cell(const LinArgs &a, int i, double *aux);

__device__ __forceinline__ long long boundary_face(const LinArgs &a, const int ijk[3], int dir) {
    const Grid &g = a.g;
    int ax = dir >> 1;
    long long loc;
    if (ax == 0) loc = ijk[1] + (long long)g.ny * ijk[2];
    else if (ax == 1) loc = ijk[0] + (long long)g.nx * ijk[2];
    else loc = ijk[0] + (long long)g.nx * ijk[1];
    return a.side_off[dir] + loc;
}

__device__ __forceinline__ void bc_aux_cell(const LinArgs &a, int i, double *aux) {
    const Grid &g = a.g;
    int ijk[3];
    cell_coords(g, i, ijk);
    for (int dir = 0; dir < 2 * g.dim; ++dir) {
        if (has_neighbour(g, ijk, dir)) continue;
        const long long bf = boundary_face(a, ijk, dir);
        if (a.bc_type[bf] != 1) continue;
        const MatReg &r = region_of(a, i);
        const double pe = a.pe ? a.pe[i] : 0.0;
        const double SwB = a.bc_val[2 * bf + 1];
        double krB[2];
        relperm(r, SwB, krB[0], krB[1]);
        aux[3 * bf] = pcnw(r, pe, SwB);
        aux[3 * bf + 1] = krB[0] / a.mu[0];
        aux[3 * bf + 2] = krB[1] / a.mu[1];
    }
}

__device__ __forceinline__ void store_block(double *J, int slot, int N, int cell, const Ev r[2]) {
    double2 *b = reinterpret_cast<double2 *>(J + ((size_t)slot * N + cell) * 4);
    b[0] = make_double2(r[0].d0, r[0].d1);
    b[1] = make_double2(r[1].d0, r[1].d1);
}

// first pass: iqv[cell] = (p_w, p_n, lambda_w, lambda_n) values, iqd[cell] = derivatives
// of (p_n, lambda_w, lambda_n) w.r.t. the cell's (p_w, S_w)
__device__ __forceinline__ void iq_cell(const LinArgs &a, int i) {
    const double2 xi = reinterpret_cast<const double2 *>(a.x)[i];
    const IQ<Ev> q = intensive<Ev>(a, i, Ev{xi.x, 1.0, 0.0}, Ev{xi.y, 0.0, 1.0});
    double4 v;
    v.x = q.p[0].v;
    v.y = q.p[1].v;
    v.z = q.lam[0].v;
    v.w = q.lam[1].v;
    reinterpret_cast<double4 *>(a.iqv)[i] = v;
    double2 *d = reinterpret_cast<double2 *>(a.iqd + (size_t)i * 6);
    d[0] = make_double2(q.p[1].d0, q.p[1].d1);
    d[1] = make_double2(q.lam[0].d0, q.lam[0].d1);
    d[2] = make_double2(q.lam[1].d0, q.lam[1].d1);
}

__device__ __forceinline__ IQ<double> iq_values(const LinArgs &a, int cell) {
    const double4 v = reinterpret_cast<const double4 *>(a.iqv)[cell];
    IQ<double> q;
    q.p[0] = v.x;
    q.p[1] = v.y;
    q.Sw = 0.0;  // not used by the flux
    q.lam[0] = v.z;
    q.lam[1] = v.w;
    return q;
}

// The element context of cell i as focus DOF: `in` holds the cell's intensive quantities
// as dual numbers in its (p_w, S_w), nbf(dir, nb, K) returns the values of the neighbour nb
// across direction dir (and sets K to its permeability). dt and newton_iter are the
// time-step size and Newton iteration of this linearization (a.dt and a.newton_iter
// unless a device Newton state supplies them).
template <int DIM, class S, class NbFn, bool CONST_GEO = false>
__device__ __forceinline__ void linearize_context(const LinArgs &a, int i, const int ijk[3],
                                                  double Ki, const IQ<Ev> &in, NbFn nbf,
                                                  double dt, int newton_iter) {
    const Grid &g = a.g;
    Ev res0[2] = {ev(0.0), ev(0.0)};
    constexpr int nfaces = 2 * DIM;
    // fluxes over the interior faces, intersection order (fvbaselocalresidual.hh:295-367)
#pragma unroll
    for (int dir = 0; dir < nfaces; ++dir) {
        if (!has_neighbour(g, ijk, dir)) continue;
        int nb = i + ((dir & 1) ? 1 : -1) * stride_of(g, dir >> 1);
        double Kex;
        const IQ<double> ex = nbf(dir, nb, Kex);
        Ev F[2];
        if constexpr (CONST_GEO)
            interior_flux_geo<DIM, S>(a, in, ex, Ki, Kex, face_geo_of<S>(a.fgeo, dir), dir, F);
        else
            interior_flux<DIM, S>(a, in, ex, Ki, Kex, ijk, dir, F);
        double alpha = 1.0;
        alpha *= g.area[dir >> 1];
        Ev rj[2];
        for (int k = 0; k < 2; ++k) {
            F[k] = F[k] * alpha;
            res0[k] = res0[k] + F[k];
            rj[k] = div_p2(ev(0.0) - F[k], g.V);
        }
        store_block(a.J, slot_of_dir(g, dir ^ 1), g.N, nb, rj);
    }
    // storage (fvbaselocalresidual.hh:494-629)
    {
        const double phi = a.phi[i];
        Ev M[2];
        storage(a, phi, in.Sw, M);
        double Mold[2];
        double2 *sto = reinterpret_cast<double2 *>(a.sto);
        if (a.enable_cache) {
            if (newton_iter == 0) {
                if (a.recycle) {
                    Mold[0] = M[0].v;
                    Mold[1] = M[1].v;
                } else {
                    const double2 xo = reinterpret_cast<const double2 *>(a.x1)[i];
                    double Mo[2];
                    storage<double>(a, phi, xo.y, Mo);
                    Mold[0] = Mo[0];
                    Mold[1] = Mo[1];
                }
                sto[i] = make_double2(Mold[0], Mold[1]);
            } else {
                double2 s = sto[i];
                Mold[0] = s.x;
                Mold[1] = s.y;
            }
        } else {
            const double2 xo = reinterpret_cast<const double2 *>(a.x1)[i];
            double Mo[2];
            storage<double>(a, phi, xo.y, Mo);
            Mold[0] = Mo[0];
            Mold[1] = Mo[1];
        }
        double scv = g.V * 1.0;
        for (int k = 0; k < 2; ++k) {
            Ev tmp = M[k] - Mold[k];
            tmp = tmp * (scv / dt);
            res0[k] = res0[k] + tmp;
        }
    }
    // boundary segments (fvbaselocalresidual.hh:425-487)
#pragma unroll
    for (int dir = 0; dir < nfaces; ++dir) {
        if (has_neighbour(g, ijk, dir)) continue;
        long long bf = boundary_face(a, ijk, dir);
        int t = a.bc_type[bf];
        Ev Bv[2] = {ev(0.0), ev(0.0)};
        if (t == 1) {
            freeflow_flux<DIM, S>(a, in, i, Ki, ijk, dir, a.bc_val[2 * bf], a.bc_val[2 * bf + 1], bf, Bv);
        } else if (t == 2) {
            Bv[0] = ev(a.bc_val[2 * bf]);
            Bv[1] = ev(a.bc_val[2 * bf + 1]);
        }
        double mult = g.area[dir >> 1] * 1.0;
        for (int k = 0; k < 2; ++k) res0[k] = res0[k] + Bv[k] * mult;
    }
    // volumetric residual (fvbaselocalresidual.hh:170-186)
    for (int k = 0; k < 2; ++k) res0[k] = div_p2(res0[k], g.V);
    reinterpret_cast<double2 *>(a.R)[i] = make_double2(res0[0].v, res0[1].v);
    store_block(a.J, g.diag, g.N, i, res0);
}

// the element context of cell i with its own and its neighbours' intensive quantities
// from the first pass (iq_cell)
template <int DIM, class S>
__device__ __forceinline__ void linearize_cell(const LinArgs &a, int i, double dt,
                                               int newton_iter) {
    const Grid &g = a.g;
    int ijk[3];
    cell_coords(g, i, ijk);
    const double2 xi = reinterpret_cast<const double2 *>(a.x)[i];
    IQ<Ev> in;
    {
        const double4 v = reinterpret_cast<const double4 *>(a.iqv)[i];
        const double2 *d = reinterpret_cast<const double2 *>(a.iqd + (size_t)i * 6);
        const double2 d0 = d[0], d1 = d[1], d2 = d[2];
        in.p[0] = Ev{xi.x, 1.0, 0.0};
        in.Sw = Ev{xi.y, 0.0, 1.0};
        in.p[1] = Ev{v.y, d0.x, d0.y};
        in.lam[0] = Ev{v.z, d1.x, d1.y};
        in.lam[1] = Ev{v.w, d2.x, d2.y};
    }
    auto nbf = [&a](int, int nb, double &K) {
        K = a.K[nb];
        return iq_values(a, nb);
    };
    linearize_context<DIM, S>(a, i, ijk, a.K[i], in, nbf, dt, newton_iter);
}

template <int DIM, class S>
__device__ __forceinline__ void linearize_cell(const LinArgs &a, int i) {
    linearize_cell<DIM, S>(a, i, a.dt, a.newton_iter);
}

// ----------------------------------------------------------------------------------
// Finite-difference local linearizer (FvBaseFdLocalLinearizer,
// fvbasefdlocallinearizer.hh:131-515): the element-local residual of cell i's context is
// evaluated in values at the focus DOF's primary variables and with one of them deflected
// by eps = BaseEpsilon / primaryVarWeight (:260-283; weights 10 / p_ref for the pressure
// and 1 for the saturation, immisciblemodel.hh:298-310), and the differences give column
// i of the Jacobian (:370-470). The neighbours' intensive quantities stay those of the
// undeflected state (the context deflects only its focus DOF, :373-379), and the storage
// term of the previous level is the one cached by the undeflected evaluation
// (fvbaselocalresidual.hh:538-575: no refresh while the focus quantities are stashed).
// The operations are those of linearize_cell, so the undeflected residual is the AD
// path's value.
// ----------------------------------------------------------------------------------
constexpr double kFdBaseEpsilon = 0.9123e-10;  // max(0.9123e-10, 1.23e3 * DBL_EPSILON)

// the element-local residual (values): res[0] of the focus cell, res[1 + dir] the face
// flux's contribution to the neighbour across direction dir (zero without a neighbour)
template <int DIM>
__device__ __forceinline__ void element_residual(const LinArgs &a, int i, const int ijk[3],
                                                 double p, double Sw, const double Mold[2],
                                                 double dt, double res[2 * DIM + 1][2]) {
    const Grid &g = a.g;
    const double Ki = a.K[i];
    const IQ<Ev> in = intensive<Ev>(a, i, ev(p), ev(Sw));
    Ev res0[2] = {ev(0.0), ev(0.0)};
    constexpr int nfaces = 2 * DIM;
#pragma unroll
    for (int dir = 0; dir < nfaces; ++dir) {
        res[1 + dir][0] = res[1 + dir][1] = 0.0;
        if (!has_neighbour(g, ijk, dir)) continue;
        const int nb = i + ((dir & 1) ? 1 : -1) * stride_of(g, dir >> 1);
        const IQ<double> ex = iq_values(a, nb);
        Ev F[2];
        interior_flux<DIM, double>(a, in, ex, Ki, a.K[nb], ijk, dir, F);
        double alpha = 1.0;
        alpha *= g.area[dir >> 1];
        for (int k = 0; k < 2; ++k) {
            F[k] = F[k] * alpha;
            res0[k] = res0[k] + F[k];
            res[1 + dir][k] = div_p2(ev(0.0) - F[k], g.V).v;
        }
    }
    {
        Ev M[2];
        storage(a, a.phi[i], in.Sw, M);
        const double scv = g.V * 1.0;
        for (int k = 0; k < 2; ++k) {
            Ev tmp = M[k] - Mold[k];
            tmp = tmp * (scv / dt);
            res0[k] = res0[k] + tmp;
        }
    }
#pragma unroll
    for (int dir = 0; dir < nfaces; ++dir) {
        if (has_neighbour(g, ijk, dir)) continue;
        const long long bf = boundary_face(a, ijk, dir);
        const int t = a.bc_type[bf];
        Ev Bv[2] = {ev(0.0), ev(0.0)};
        if (t == 1) {
            freeflow_flux<DIM, double>(a, in, i, Ki, ijk, dir, a.bc_val[2 * bf],
                                       a.bc_val[2 * bf + 1], bf, Bv);
        } else if (t == 2) {
            Bv[0] = ev(a.bc_val[2 * bf]);
            Bv[1] = ev(a.bc_val[2 * bf + 1]);
        }
        const double mult = g.area[dir >> 1] * 1.0;
        for (int k = 0; k < 2; ++k) res0[k] = res0[k] + Bv[k] * mult;
    }
    for (int k = 0; k < 2; ++k) res[0][k] = div_p2(res0[k], g.V).v;
}

// method: +1 forward, 0 central, -1 backward differences (NumericDifferenceMethod);
// pref: the reference pressure of the pressure's weight
template <int DIM>
__device__ __forceinline__ void linearize_cell_fd(const LinArgs &a, int i, double dt,
                                                  int newton_iter, double pref, int method) {
    constexpr int ND = 2 * DIM + 1;
    const Grid &g = a.g;
    int ijk[3];
    cell_coords(g, i, ijk);
    const double2 xi = reinterpret_cast<const double2 *>(a.x)[i];
    // storage of the previous level with linearize_cell's cache semantics
    double Mold[2];
    {
        const double phi = a.phi[i];
        double2 *sto = reinterpret_cast<double2 *>(a.sto);
        if (a.enable_cache && newton_iter != 0) {
            const double2 c = sto[i];
            Mold[0] = c.x;
            Mold[1] = c.y;
        } else {
            const double Sold = (a.enable_cache && a.recycle) ? xi.y
                                : reinterpret_cast<const double2 *>(a.x1)[i].y;
            storage<double>(a, phi, Sold, Mold);
            if (a.enable_cache) sto[i] = make_double2(Mold[0], Mold[1]);
        }
    }
    double base[ND][2];
    element_residual<DIM>(a, i, ijk, xi.x, xi.y, Mold, dt, base);
    reinterpret_cast<double2 *>(a.R)[i] = make_double2(base[0][0], base[0][1]);
    double d[2][ND][2];  // [pv][dof][eq]
    for (int pv = 0; pv < 2; ++pv) {
        const double x = pv == 0 ? xi.x : xi.y;
        const double pvWeight = pv == 0 ? 10.0 / pref : 1.0;
        const double eps = kFdBaseEpsilon / pvWeight;
        double hi[ND][2], lo[ND][2];
        double delta = 0.0;
        double xp = x;
        if (method >= 0) {
            xp += eps;
            delta += eps;
            element_residual<DIM>(a, i, ijk, pv == 0 ? xp : xi.x, pv == 1 ? xp : xi.y, Mold, dt, hi);
        } else {
            for (int l = 0; l < ND; ++l) hi[l][0] = base[l][0], hi[l][1] = base[l][1];
        }
        if (method <= 0) {
            xp -= delta + eps;
            delta += eps;
            element_residual<DIM>(a, i, ijk, pv == 0 ? xp : xi.x, pv == 1 ? xp : xi.y, Mold, dt, lo);
        } else {
            for (int l = 0; l < ND; ++l) lo[l][0] = base[l][0], lo[l][1] = base[l][1];
        }
        for (int l = 0; l < ND; ++l)
            for (int k = 0; k < 2; ++k) {
                double v = hi[l][k];
                v -= lo[l][k];
                d[pv][l][k] = v / delta;
            }
    }
    // column i: the focus block and the neighbours' blocks towards the focus
    Ev r[2];
    for (int k = 0; k < 2; ++k) r[k] = Ev{0.0, d[0][0][k], d[1][0][k]};
    store_block(a.J, g.diag, g.N, i, r);
    for (int dir = 0; dir < 2 * DIM; ++dir) {
        if (!has_neighbour(g, ijk, dir)) continue;
        const int nb = i + ((dir & 1) ? 1 : -1) * stride_of(g, dir >> 1);
        for (int k = 0; k < 2; ++k) r[k] = Ev{0.0, d[0][1 + dir][k], d[1][1 + dir][k]};
        store_block(a.J, slot_of_dir(g, dir ^ 1), g.N, nb, r);
    }
}

}  // namespace lin
}  // namespace ewg
