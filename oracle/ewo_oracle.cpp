// ewo_oracle.cpp -- CPU oracle of the ECFV immiscible two-phase Newton step.
//
// TEST INFRASTRUCTURE ONLY: the checker for libewg. Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg load this library; no product code links or calls it.
//
// A from-scratch restatement of the reference's behaviour, element by element, as the
// reference evaluates it (one element context per cell, focus DOF = the element, 2-slot
// forward-mode dual numbers). The reference files each part follows are cited inline.
// Third-party arithmetic the reference delegates to ewoms-common / ewoms-material /
// dune-istl (dual-number rules, harmonic mean, the regularized material laws, block
// ILU(0), 2x2 inversion) is restated from the published formulas; those parts are
// pinned by closed-form tests in tests/, not by the reference (it cannot be built here).
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off: every rounding step is the one
// written here, so the same inputs give the same bits on every host).

#include "ewo_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

struct NumericalIssue : std::runtime_error {
    explicit NumericalIssue(const std::string &s) : std::runtime_error(s) {}
};

// ---------------------------------------------------------------------------------
// Deterministic elementary functions. log via the atanh series on a mantissa reduced
// to [sqrt(1/2), sqrt(2)], exp via round-to-nearest reduction by ln2 (split in a
// 32-bit head and a tail) and a degree-13 Taylor polynomial, pow = exp(y log x).
// Only +,-,*,/ and fused multiply-add are used, so the bits are host independent.
// ---------------------------------------------------------------------------------
const double kLn2Hi = 0x1.62e42fee00000p-1;
const double kLn2Lo = 0x1.a39ef35793c76p-33;
const double kInvLn2 = 0x1.71547652b82fep+0;
const double kSqrt2 = 0x1.6a09e667f3bcdp+0;

double exp2i(int k) {  // exact 2^k, -1022 <= k <= 1023
    uint64_t b = static_cast<uint64_t>(k + 1023) << 52;
    double r;
    std::memcpy(&r, &b, 8);
    return r;
}

double d_exp(double x) {
    if (x != x) return x;
    if (x > 709.78) return std::numeric_limits<double>::infinity();
    if (x < -745.2) return 0.0;
    double kd = std::rint(x * kInvLn2);
    double r = std::fma(-kd, kLn2Hi, x);
    r = std::fma(-kd, kLn2Lo, r);
    double p = 1.0 / 6227020800.0;
    p = std::fma(p, r, 1.0 / 479001600.0);
    p = std::fma(p, r, 1.0 / 39916800.0);
    p = std::fma(p, r, 1.0 / 3628800.0);
    p = std::fma(p, r, 1.0 / 362880.0);
    p = std::fma(p, r, 1.0 / 40320.0);
    p = std::fma(p, r, 1.0 / 5040.0);
    p = std::fma(p, r, 1.0 / 720.0);
    p = std::fma(p, r, 1.0 / 120.0);
    p = std::fma(p, r, 1.0 / 24.0);
    p = std::fma(p, r, 1.0 / 6.0);
    p = std::fma(p, r, 0.5);
    p = std::fma(p, r, 1.0);
    p = std::fma(p, r, 1.0);
    int k = static_cast<int>(kd);
    if (k > 1023) return (p * exp2i(1023)) * exp2i(k - 1023);
    if (k < -1022) return (p * exp2i(-1022)) * exp2i(k + 1022);
    return p * exp2i(k);
}

double d_log(double x) {
    if (x != x || x < 0.0) return std::numeric_limits<double>::quiet_NaN();
    if (x == 0.0) return -std::numeric_limits<double>::infinity();
    if (std::isinf(x)) return x;
    uint64_t b;
    std::memcpy(&b, &x, 8);
    int e = static_cast<int>((b >> 52) & 0x7ff);
    if (e == 0) {  // subnormal: scale into the normal range first
        double xs = x * 0x1p54;
        std::memcpy(&b, &xs, 8);
        e = static_cast<int>((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    uint64_t mb = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    double m;
    std::memcpy(&m, &mb, 8);
    if (m > kSqrt2) {
        m = m * 0.5;
        e += 1;
    }
    double f = m - 1.0;
    double s = f / (m + 1.0);
    double s2 = s * s;
    double p = 2.0 / 23.0;
    p = std::fma(p, s2, 2.0 / 21.0);
    p = std::fma(p, s2, 2.0 / 19.0);
    p = std::fma(p, s2, 2.0 / 17.0);
    p = std::fma(p, s2, 2.0 / 15.0);
    p = std::fma(p, s2, 2.0 / 13.0);
    p = std::fma(p, s2, 2.0 / 11.0);
    p = std::fma(p, s2, 2.0 / 9.0);
    p = std::fma(p, s2, 2.0 / 7.0);
    p = std::fma(p, s2, 2.0 / 5.0);
    p = std::fma(p, s2, 2.0 / 3.0);
    double logm = std::fma(s * s2, p, 2.0 * s);
    double ed = static_cast<double>(e);
    return std::fma(ed, kLn2Hi, std::fma(ed, kLn2Lo, logm));
}

double d_pow(double x, double y) {
    if (x > 0.0) return d_exp(y * d_log(x));
    if (x == 0.0) {
        if (y > 0.0) return 0.0;
        if (y == 0.0) return 1.0;
        return std::numeric_limits<double>::infinity();
    }
    return std::numeric_limits<double>::quiet_NaN();
}

// ---------------------------------------------------------------------------------
// Forward-mode dual number with two derivative slots, the restatement of
// DenseAd::Evaluation<double,2> chosen at fvbaseadlocallinearizer.hh:68-77.
// Slot 0 = d/dp_w, slot 1 = d/dS_w (fvbaseprimaryvariables.hh:86-97).
// ---------------------------------------------------------------------------------
struct Ev {
    double v;
    double d[2];
    Ev() : v(0.0), d{0.0, 0.0} {}
    Ev(double x) : v(x), d{0.0, 0.0} {}  // NOLINT: implicit constant
    static Ev var(double x, int slot) {
        Ev e(x);
        e.d[slot] = 1.0;
        return e;
    }
};
inline Ev operator+(const Ev &a, const Ev &b) {
    Ev r; r.v = a.v + b.v; r.d[0] = a.d[0] + b.d[0]; r.d[1] = a.d[1] + b.d[1]; return r;
}
inline Ev operator-(const Ev &a, const Ev &b) {
    Ev r; r.v = a.v - b.v; r.d[0] = a.d[0] - b.d[0]; r.d[1] = a.d[1] - b.d[1]; return r;
}
inline Ev operator-(const Ev &a) {
    Ev r; r.v = -a.v; r.d[0] = -a.d[0]; r.d[1] = -a.d[1]; return r;
}
inline Ev operator*(const Ev &a, const Ev &b) {  // (uv)' = u'v + v'u
    Ev r;
    r.v = a.v * b.v;
    r.d[0] = a.d[0] * b.v + b.d[0] * a.v;
    r.d[1] = a.d[1] * b.v + b.d[1] * a.v;
    return r;
}
inline Ev operator/(const Ev &a, const Ev &b) {  // (u/v)' = (v u' - u v')/v^2
    Ev r;
    double vv = b.v * b.v;
    r.d[0] = (b.v * a.d[0] - a.v * b.d[0]) / vv;
    r.d[1] = (b.v * a.d[1] - a.v * b.d[1]) / vv;
    r.v = a.v / b.v;
    return r;
}
inline Ev operator+(const Ev &a, double s) { Ev r = a; r.v = a.v + s; return r; }
inline Ev operator+(double s, const Ev &a) { Ev r = a; r.v = s + a.v; return r; }
inline Ev operator-(const Ev &a, double s) { Ev r = a; r.v = a.v - s; return r; }
inline Ev operator-(double s, const Ev &a) {
    Ev r; r.v = s - a.v; r.d[0] = -a.d[0]; r.d[1] = -a.d[1]; return r;
}
inline Ev operator*(const Ev &a, double s) {
    Ev r; r.v = a.v * s; r.d[0] = a.d[0] * s; r.d[1] = a.d[1] * s; return r;
}
inline Ev operator*(double s, const Ev &a) { return a * s; }
inline Ev operator/(const Ev &a, double s) {
    Ev r; r.v = a.v / s; r.d[0] = a.d[0] / s; r.d[1] = a.d[1] / s; return r;
}
inline Ev ev_pow(const Ev &x, double y) {
    double pw = d_pow(x.v, y);
    double df = (x.v == 0.0) ? 0.0 : y * pw / x.v;
    Ev r; r.v = pw; r.d[0] = x.d[0] * df; r.d[1] = x.d[1] * df; return r;
}
inline Ev ev_sqrt(const Ev &x) {
    double s = std::sqrt(x.v);
    double ds = (s == 0.0) ? 0.0 : 0.5 / s;
    Ev r; r.v = s; r.d[0] = x.d[0] * ds; r.d[1] = x.d[1] * ds; return r;
}
inline Ev value_of(const Ev &x) { return Ev(x.v); }  // Toolbox::value / getValue

// harmonic mean of ewoms-common (used at multiphasebaseproblem.hh:123)
double harmonic_mean(double x, double y) {
    if (x * y <= 0.0) return 0.0;
    return (2.0 * x * y) / (x + y);
}

// ---------------------------------------------------------------------------------
// Material laws: EffToAbsLaw over a regularized effective-saturation law
// (lensproblem.hh:107-126; tutorial1problem.hh:99-104 for Brooks-Corey).
// Regularization (ewoms-material is not available, so this is the documented choice):
//   pc: linear extrapolation below Se = 0.01 with the slope there; van Genuchten: cubic
//       Hermite spline on [0.99, 1] from (pc(0.99), pc'(0.99)) to (0, -2 pc(0.99)/0.01),
//       then linear with that end slope above Se = 1; Brooks-Corey: linear with the
//       slope at Se = 1 above 1.
//   kr: clamped outside (0, 1) to their limits (krw 0|1, krn 1|0).
// ---------------------------------------------------------------------------------
const double kSeLow = 0.01;
const double kSeHigh = 0.99;

struct MatReg {
    int law = 0;
    double p1 = 0, p2 = 0, Swr = 0, Snr = 0, denom = 1;
    // van Genuchten
    double alpha = 0, n = 0, m = 0, e_pc1 = 0, e_pc2 = 0, e_kr1 = 0, e_kr2 = 0, e_krn1 = 0,
           e_krn2 = 0;
    double pcLow = 0, dpcLow = 0, pcHigh = 0, dpcHigh = 0, mEnd = 0;
    // Brooks-Corey (unit entry pressure)
    double lambda = 0, e_pc = 0, e_krw = 0, e_krn = 0, cLow = 0, dLow = 0;
};

double vg_pc_raw(const MatReg &r, double Se) {
    return d_pow(d_pow(Se, r.e_pc1) - 1.0, r.e_pc2) / r.alpha;
}
double vg_dpc_raw(const MatReg &r, double Se) {
    double a = d_pow(Se, r.e_pc1);
    double b = a - 1.0;
    double pc = d_pow(b, r.e_pc2) / r.alpha;
    return -pc * a / (r.n * r.m * b * Se);
}

MatReg make_region(int law, const double *p) {
    MatReg r;
    r.law = law;
    r.p1 = p[0]; r.p2 = p[1]; r.Swr = p[2]; r.Snr = p[3];
    r.denom = 1.0 - r.Swr - r.Snr;
    if (law == EWG_LAW_VAN_GENUCHTEN) {
        r.alpha = p[0];
        r.n = p[1];
        r.m = 1.0 - 1.0 / r.n;
        r.e_pc1 = -1.0 / r.m;
        r.e_pc2 = 1.0 / r.n;
        r.e_kr1 = 1.0 / r.m;
        r.e_kr2 = r.m;
        r.e_krn1 = 1.0 / 3.0;
        r.e_krn2 = 2.0 * r.m;
        r.pcLow = vg_pc_raw(r, kSeLow);
        r.dpcLow = vg_dpc_raw(r, kSeLow);
        r.pcHigh = vg_pc_raw(r, kSeHigh);
        r.dpcHigh = vg_dpc_raw(r, kSeHigh);
        r.mEnd = -2.0 * r.pcHigh / (1.0 - kSeHigh);
    } else {
        r.lambda = p[1];
        r.e_pc = -1.0 / r.lambda;
        r.e_krw = (2.0 + 3.0 * r.lambda) / r.lambda;
        r.e_krn = (2.0 + r.lambda) / r.lambda;
        r.cLow = d_pow(kSeLow, r.e_pc);
        r.dLow = r.e_pc * r.cLow / kSeLow;
    }
    return r;
}

Ev eff_sat(const MatReg &r, const Ev &Sw) { return (Sw - r.Swr) / r.denom; }

Ev pcnw(const MatReg &r, double pe, const Ev &Sw) {
    Ev Se = eff_sat(r, Sw);
    if (r.law == EWG_LAW_VAN_GENUCHTEN) {
        if (Se.v < kSeLow) return r.pcLow + r.dpcLow * (Se - kSeLow);
        if (Se.v <= kSeHigh) return ev_pow(ev_pow(Se, r.e_pc1) - 1.0, r.e_pc2) / r.alpha;
        if (Se.v < 1.0) {
            const double h = 1.0 - kSeHigh;
            Ev t = (Se - kSeHigh) / h;
            Ev t2 = t * t;
            Ev t3 = t2 * t;
            Ev h00 = (2.0 * t3 - 3.0 * t2) + 1.0;
            Ev h10 = (t3 - 2.0 * t2) + t;
            Ev h11 = t3 - t2;
            return (h00 * r.pcHigh + h10 * (h * r.dpcHigh)) + h11 * (h * r.mEnd);
        }
        return r.mEnd * (Se - 1.0);
    }
    double peu = pe > 0.0 ? pe : r.p1;
    if (Se.v < kSeLow) return peu * (r.cLow + r.dLow * (Se - kSeLow));
    if (Se.v <= 1.0) return peu * ev_pow(Se, r.e_pc);
    return peu * (1.0 + r.e_pc * (Se - 1.0));
}

void relperm(const MatReg &r, const Ev &Sw, Ev kr[2]) {
    Ev Se = eff_sat(r, Sw);
    if (Se.v <= 0.0) { kr[0] = Ev(0.0); kr[1] = Ev(1.0); return; }
    if (Se.v >= 1.0) { kr[0] = Ev(1.0); kr[1] = Ev(0.0); return; }
    if (r.law == EWG_LAW_VAN_GENUCHTEN) {
        Ev a = 1.0 - ev_pow(1.0 - ev_pow(Se, r.e_kr1), r.e_kr2);
        kr[0] = ev_sqrt(Se) * (a * a);
        kr[1] = ev_pow(1.0 - Se, r.e_krn1) * ev_pow(1.0 - ev_pow(Se, r.e_kr1), r.e_krn2);
    } else {
        kr[0] = ev_pow(Se, r.e_krw);
        Ev t = 1.0 - Se;
        kr[1] = (t * t) * (1.0 - ev_pow(Se, r.e_krn));
    }
}

// intensive quantities of one DOF (immiscibleintensivequantities.hh:87-145)
struct IQ {
    Ev p[2];
    Ev S[2];
    Ev lam[2];
};

}  // namespace

// -------------------------------------------------------------------------------------
// context
// -------------------------------------------------------------------------------------
struct ewo_ctx {
    ewg_grid_desc grid{};
    ewg_fluid_desc fluid{};
    ewg_params prm{};
    int dim = 3, nx = 1, ny = 1, nz = 1;
    int64_t N = 0, B = 0;
    double h[3] = {1, 1, 1};
    double V = 1.0;
    double area[3] = {1, 1, 1};
    std::vector<MatReg> regs;
    std::vector<double> K, phi, pe;
    std::vector<int32_t> region;
    std::vector<int32_t> bc_type;
    std::vector<double> bc_val;
    int64_t side_off[7] = {0};
    // solution history and caches
    std::vector<double> x0, x1, storage_old;
    // BCRS pattern
    std::vector<int32_t> rowptr, colidx, diagpos;
    // last linear system
    std::vector<double> resid, jac;
    bool have_system = false;
    // Newton state (newtonmethod.hh:931-933; error_ starts at 1e100, :197)
    double error = 1e100, last_error = 1e100;
    int num_iterations = 0;
    int last_linear_iterations = 0;
    ewg_stats stats{};
    std::vector<int32_t> step_its;
    std::vector<double> step_dts;
    std::vector<double> newton_errors;  // error of every linearization of the last Newton
    std::string err;
    bool static_set = false;
    int threads = 1;  // worker threads of the linearization (ThreadsPerProcess)
    double pref = 1e5;  // reference pressure of the finite-difference weights
};

namespace {

inline void cell_coords(const ewo_ctx &c, int64_t idx, int ijk[3]) {
    ijk[0] = static_cast<int>(idx % c.nx);
    ijk[1] = static_cast<int>((idx / c.nx) % c.ny);
    ijk[2] = static_cast<int>(idx / (static_cast<int64_t>(c.nx) * c.ny));
}
inline int64_t cell_index(const ewo_ctx &c, const int ijk[3]) {
    return ijk[0] + static_cast<int64_t>(c.nx) * (ijk[1] + static_cast<int64_t>(c.ny) * ijk[2]);
}
// cell centre (YaspGrid equidistant coordinates)
inline void centre(const ewo_ctx &c, const int ijk[3], double pos[3]) {
    for (int d = 0; d < 3; ++d) pos[d] = (ijk[d] + 0.5) * c.h[d];
}
// neighbour across face dir (0:x-,1:x+,2:y-,3:y+,4:z-,5:z+); returns -1 at the boundary
inline int64_t neighbour(const ewo_ctx &c, const int ijk[3], int dir) {
    int a = dir >> 1;
    int s = (dir & 1) ? 1 : -1;
    int n[3] = {c.nx, c.ny, c.nz};
    int q[3] = {ijk[0], ijk[1], ijk[2]};
    q[a] += s;
    if (q[a] < 0 || q[a] >= n[a]) return -1;
    return cell_index(c, q);
}
inline void face_centre(const ewo_ctx &c, const int ijk[3], int dir, double pos[3]) {
    centre(c, ijk, pos);
    int a = dir >> 1;
    pos[a] = ((dir & 1) ? (ijk[a] + 1) : ijk[a]) * c.h[a];
}
inline int64_t boundary_face_index(const ewo_ctx &c, const int ijk[3], int dir) {
    int a = dir >> 1;
    int64_t loc;
    if (a == 0) loc = ijk[1] + static_cast<int64_t>(c.ny) * ijk[2];
    else if (a == 1) loc = ijk[0] + static_cast<int64_t>(c.nx) * ijk[2];
    else loc = ijk[0] + static_cast<int64_t>(c.nx) * ijk[1];
    return c.side_off[dir] + loc;
}

// the intensive quantities of DOF `cell` at primary variables (p_w, S_w), both derivative
// slots seeded
IQ intensive_at(const ewo_ctx &c, int64_t cell, double pw, double sw) {
    IQ q;
    Ev p0 = Ev::var(pw, 0);
    Ev Sw = Ev::var(sw, 1);
    q.S[0] = Sw;
    q.S[1] = 1.0 - Sw;
    const MatReg &r = c.regs[c.region[cell]];
    double pe = c.pe.empty() ? 0.0 : c.pe[cell];
    Ev pc = pcnw(r, pe, Sw);
    Ev kr[2];
    relperm(r, Sw, kr);
    q.p[0] = p0;
    q.p[1] = p0 + pc;
    for (int k = 0; k < 2; ++k) q.lam[k] = kr[k] / Ev(c.fluid.mu[k]);
    return q;
}

IQ intensive(const ewo_ctx &c, int64_t cell, const double *x) {
    return intensive_at(c, cell, x[2 * cell], x[2 * cell + 1]);
}

// storage term M_k = phi S_k rho_k (immisciblelocalresidual.hh:69-103)
void storage_of(const ewo_ctx &c, int64_t cell, const IQ &q, Ev M[2]) {
    for (int k = 0; k < 2; ++k) M[k] = (c.phi[cell] * q.S[k]) * c.fluid.rho[k];
}

inline double dot_g(const ewo_ctx &c, const double v[3]) {
    double s = 0.0;
    for (int d = 0; d < c.dim; ++d) s += c.fluid.gravity[d] * v[d];
    return s;
}

// Interior face flux seen from the focus element (in = focus, ex = neighbour):
// fvbasegradientcalculator.hh:195-240, darcyfluxmodule.hh:185-324 and :461-527,
// immisciblelocalresidual.hh:123-149.
void interior_flux(const ewo_ctx &c, const IQ &in, const IQ &ex, int64_t cin, int64_t cex,
                   const int ijk[3], int dir, Ev F[2]) {
    int qex[3];
    cell_coords(c, cex, qex);
    double posIn[3], posEx[3], posFace[3];
    centre(c, ijk, posIn);
    centre(c, qex, posEx);
    face_centre(c, ijk, dir, posFace);
    double normal[3] = {0, 0, 0};
    normal[dir >> 1] = (dir & 1) ? 1.0 : -1.0;
    const int dim = c.dim;

    double distSq = 0.0;
    for (int d = 0; d < dim; ++d) {
        double t = posEx[d] - posIn[d];
        distSq += t * t;
    }
    Ev grad[2][3];
    for (int k = 0; k < 2; ++k) {
        Ev deltay = value_of(ex.p[k]) - in.p[k];
        for (int d = 0; d < dim; ++d) grad[k][d] = deltay * ((posEx[d] - posIn[d]) / distSq);
    }
    // gravity correction (darcyfluxmodule.hh:217-290)
    double dvIn[3], dvEx[3], dvTot[3];
    for (int d = 0; d < 3; ++d) {
        dvIn[d] = posIn[d] - posFace[d];
        dvEx[d] = posEx[d] - posFace[d];
        dvTot[d] = posEx[d] - posIn[d];
    }
    double absDistTotSq = 0.0;
    for (int d = 0; d < dim; ++d) absDistTotSq += dvTot[d] * dvTot[d];
    double gIn = dot_g(c, dvIn), gEx = dot_g(c, dvEx);
    for (int k = 0; k < 2; ++k) {
        Ev pStatIn = -Ev(c.fluid.rho[k]) * gIn;           // interior is the focus
        Ev pStatEx = Ev(-c.fluid.rho[k] * gEx);           // exterior: value only
        Ev q = (pStatEx - pStatIn) / absDistTotSq;
        for (int d = 0; d < dim; ++d) grad[k][d] = grad[k][d] + Ev(dvTot[d]) * q;
    }
    double Kf = harmonic_mean(c.K[cin], c.K[cex]);
    for (int k = 0; k < 2; ++k) {
        Ev tmp(0.0);
        for (int d = 0; d < dim; ++d) tmp = tmp + grad[k][d] * normal[d];
        bool upIsEx = tmp.v > 0.0;  // strict, darcyfluxmodule.hh:307
        Ev lam = upIsEx ? value_of(ex.lam[k]) : in.lam[k];
        Ev v[3];
        for (int r = 0; r < dim; ++r) {
            v[r] = Ev(0.0);
            for (int cc = 0; cc < dim; ++cc) v[r] = v[r] + (r == cc ? Kf : 0.0) * grad[k][cc];
        }
        Ev mlam = -lam;
        for (int r = 0; r < dim; ++r) v[r] = v[r] * mlam;
        Ev vol(0.0);
        for (int d = 0; d < dim; ++d) vol = vol + v[d] * normal[d];
        F[k] = Ev(0.0) + vol * Ev(c.fluid.rho[k]);  // rho of the upstream DOF (constant)
    }
}

// Dirichlet "free flow" boundary (darcyfluxmodule.hh:332-453, :488-509;
// fvbasegradientcalculator.hh:278-312; immiscibleboundaryratevector.hh:91-153;
// boundary fluid state as built at lensproblem.hh:473-508).
void freeflow_flux(const ewo_ctx &c, const IQ &in, int64_t cin, const int ijk[3], int dir,
                   double pwB, double SwB, Ev B[2]) {
    const MatReg &r = c.regs[c.region[cin]];
    double pe = c.pe.empty() ? 0.0 : c.pe[cin];
    double pcB = pcnw(r, pe, Ev(SwB)).v;
    double pB[2] = {pwB, (pwB + pcB) - 0.0};
    Ev krB[2];
    relperm(r, Ev(SwB), krB);
    double posIn[3], posFace[3];
    centre(c, ijk, posIn);
    face_centre(c, ijk, dir, posFace);
    double normal[3] = {0, 0, 0};
    normal[dir >> 1] = (dir & 1) ? 1.0 : -1.0;
    const int dim = c.dim;
    double distSq = 0.0;
    for (int d = 0; d < dim; ++d) {
        double t = posFace[d] - posIn[d];
        distSq += t * t;
    }
    Ev grad[2][3];
    for (int k = 0; k < 2; ++k) {
        Ev deltay = pB[k] - in.p[k];
        for (int d = 0; d < dim; ++d) grad[k][d] = deltay * ((posFace[d] - posIn[d]) / distSq);
    }
    double dvIn[3];
    for (int d = 0; d < 3; ++d) dvIn[d] = posIn[d] - posFace[d];
    double absDistSq = 0.0;
    for (int d = 0; d < dim; ++d) absDistSq += dvIn[d] * dvIn[d];
    double gTimesDist = dot_g(c, dvIn);
    for (int k = 0; k < 2; ++k) {
        Ev pStatIn = (-gTimesDist) * Ev(c.fluid.rho[k]);
        Ev q = pStatIn / absDistSq;
        for (int d = 0; d < dim; ++d) grad[k][d] = grad[k][d] + Ev(dvIn[d]) * q;
    }
    double Kc = c.K[cin];
    for (int k = 0; k < 2; ++k) {
        Ev tmp(0.0);
        for (int d = 0; d < dim; ++d) tmp = tmp + grad[k][d] * normal[d];
        bool upIsBoundary = tmp.v > 0.0;
        Ev mob = upIsBoundary ? Ev(krB[k].v / c.fluid.mu[k]) : in.lam[k];
        Ev v[3];
        for (int rr = 0; rr < dim; ++rr) {
            v[rr] = Ev(0.0);
            for (int cc = 0; cc < dim; ++cc) v[rr] = v[rr] + (rr == cc ? Kc : 0.0) * grad[k][cc];
        }
        Ev mmob = -mob;
        for (int rr = 0; rr < dim; ++rr) v[rr] = v[rr] * mmob;
        Ev vol(0.0);
        for (int d = 0; d < dim; ++d) vol = vol + v[d] * normal[d];
        // density chosen by pressure comparison (immiscibleboundaryratevector.hh:109-119);
        // both choices are the same constant for incompressible liquids
        double dens = (pB[k] > in.p[k].v) ? c.fluid.rho[k] : c.fluid.rho[k];
        B[k] = Ev(0.0) + vol * Ev(dens);
    }
}

int64_t find_block(const ewo_ctx &c, int64_t row, int64_t col) {
    for (int32_t p = c.rowptr[row]; p < c.rowptr[row + 1]; ++p)
        if (c.colidx[p] == col) return p;
    return -1;
}

// The element context of `cell` (FvBaseLocalResidual::eval, fvbaselocalresidual.hh:140-187):
// the local residual of every stencil DOF with the focus DOF's intensive quantities iq0
// (the neighbours' from iqs, values only where the focus is not upstream) and the storage
// term of the previous level Mold. Stencil: dofs[0] = the element, then one DOF per
// interior intersection in intersection order; res[l] is DOF l's residual.
void element_local(const ewo_ctx &c, int64_t cell, const int ijk[3], const IQ &iq0,
                   const std::vector<IQ> &iqs, const double Mold[2], double dt, int &nd,
                   int64_t dofs[7], Ev res[7][2]) {
    const int nfaces = 2 * c.dim;
    int dirs[6];
    int ni = 0;
    nd = 1;
    dofs[0] = cell;
    for (int dir = 0; dir < nfaces; ++dir) {
        int64_t nb = neighbour(c, ijk, dir);
        if (nb >= 0) {
            dirs[ni++] = dir;
            dofs[nd++] = nb;
        }
    }
    for (int l = 0; l < nd; ++l) res[l][0] = res[l][1] = Ev(0.0);
    // fluxes (fvbaselocalresidual.hh:295-367)
    for (int f = 0; f < ni; ++f) {
        Ev F[2];
        interior_flux(c, iq0, iqs[dofs[f + 1]], cell, dofs[f + 1], ijk, dirs[f], F);
        double alpha = 1.0;
        alpha *= c.area[dirs[f] >> 1];
        for (int k = 0; k < 2; ++k) {
            F[k] = F[k] * alpha;
            res[0][k] = res[0][k] + F[k];
            res[f + 1][k] = res[f + 1][k] - F[k];
        }
    }
    // storage (fvbaselocalresidual.hh:494-629)
    {
        Ev M[2];
        storage_of(c, cell, iq0, M);
        double scv = c.V * 1.0;
        for (int k = 0; k < 2; ++k) {
            Ev tmp = M[k] - Mold[k];
            tmp = tmp * (scv / dt);
            res[0][k] = res[0][k] + tmp;
        }
    }
    // boundary (fvbaselocalresidual.hh:425-487)
    for (int dir = 0; dir < nfaces; ++dir) {
        if (neighbour(c, ijk, dir) >= 0) continue;
        int64_t bf = boundary_face_index(c, ijk, dir);
        int t = c.bc_type[bf];
        Ev Bv[2];
        if (t == EWG_BC_FREE_FLOW) {
            freeflow_flux(c, iq0, cell, ijk, dir, c.bc_val[2 * bf], c.bc_val[2 * bf + 1], Bv);
        } else if (t == EWG_BC_MASS_RATE) {
            Bv[0] = Ev(c.bc_val[2 * bf]);
            Bv[1] = Ev(c.bc_val[2 * bf + 1]);
        } else {
            Bv[0] = Ev(0.0);
            Bv[1] = Ev(0.0);
        }
        double mult = c.area[dir >> 1] * 1.0;
        for (int k = 0; k < 2; ++k) res[0][k] = res[0][k] + Bv[k] * mult;
    }
    // volumetric residual (fvbaselocalresidual.hh:170-186)
    for (int l = 0; l < nd; ++l)
        for (int k = 0; k < 2; ++k) res[l][k] = res[l][k] / c.V;
}

// the storage term of the previous time level for element `cell` with the cache semantics
// of fvbaselocalresidual.hh:538-575 (iteration 0 refreshes the cache from the focus
// quantities iq0, or from solution(1) without first-iteration recycling)
void previous_storage(ewo_ctx &c, int64_t cell, const IQ &iq0, int newton_iter, double Mold[2]) {
    auto from_level1 = [&]() {
        IQ qo = intensive(c, cell, c.x1.data());
        Ev Mo[2];
        storage_of(c, cell, qo, Mo);
        Mold[0] = Mo[0].v;
        Mold[1] = Mo[1].v;
    };
    if (!c.prm.enable_storage_cache) {
        from_level1();
        return;
    }
    if (newton_iter != 0) {
        Mold[0] = c.storage_old[2 * cell];
        Mold[1] = c.storage_old[2 * cell + 1];
        return;
    }
    if (c.prm.recycle_first_iteration_storage) {
        Ev M[2];
        storage_of(c, cell, iq0, M);
        Mold[0] = M[0].v;
        Mold[1] = M[1].v;
    } else {
        from_level1();
    }
    c.storage_old[2 * cell] = Mold[0];
    c.storage_old[2 * cell + 1] = Mold[1];
}

// FvBaseLinearizer::linearize_ + linearizeElement_ (fvbaselinearizer.hh:401-512) with
// FvBaseAdLocalLinearizer::linearize (fvbaseadlocallinearizer.hh:178-204), or with
// FvBaseFdLocalLinearizer::linearize (fvbasefdlocallinearizer.hh:223-245, 370-470) when
// params.local_linearizer = 1: the focus DOF's primary variable pvIdx deflected by
// eps = BaseEpsilon / primaryVarWeight (:260-283), the element-local residual
// re-evaluated with the neighbours' quantities and the cached storage unchanged.
void linearize_all(ewo_ctx &c, double dt, int newton_iter, double *R, double *J) {
    const double *x = c.x0.data();
    std::fill(R, R + 2 * c.N, 0.0);
    std::fill(J, J + 4 * static_cast<int64_t>(c.colidx.size()), 0.0);
    const bool fd = c.prm.local_linearizer == 1;
    const int method = c.prm.numeric_difference_method;
    // immisciblemodel.hh:298-310: weight 10 / p_ref for the pressure, 1 for the saturation
    const double pvWeight[2] = {10.0 / c.pref, 1.0};
    const double baseEpsilon = std::max(0.9123e-10, std::numeric_limits<double>::epsilon() * 1.23e3);
    // intensive-quantity cache (EnableIntensiveQuantityCache, fvbasediscretization.hh:650-681):
    // every DOF is evaluated once per linearization with both derivative slots seeded
    std::vector<IQ> iqs(static_cast<size_t>(c.N));
    // Elements are independent: element `cell` writes R[cell], column `cell` of J and its
    // own storage cache entry only, so a contiguous range of elements per worker thread
    // gives the same bits as one thread (the reference dispatches the element loop over
    // ThreadsPerProcess worker threads, fvbaselinearizer.hh:416-471).
    auto over_cells = [&](auto &&body) {
        const int T = std::max(1, c.threads);
        if (T == 1) {
            body(int64_t(0), c.N);
            return;
        }
        std::vector<std::thread> pool;
        for (int t = 0; t < T; ++t)
            pool.emplace_back(body, c.N * t / T, c.N * (t + 1) / T);
        for (auto &th : pool) th.join();
    };
    over_cells([&](int64_t c0, int64_t c1) {
        for (int64_t cell = c0; cell < c1; ++cell) iqs[cell] = intensive(c, cell, x);
    });
    over_cells([&](int64_t c0, int64_t c1) {
        for (int64_t cell = c0; cell < c1; ++cell) {
            int ijk[3];
            cell_coords(c, cell, ijk);
            double Mold[2];
            previous_storage(c, cell, iqs[cell], newton_iter, Mold);
            int64_t dofs[7];
            int nd = 0;
            Ev res[7][2];
            element_local(c, cell, ijk, iqs[cell], iqs, Mold, dt, nd, dofs, res);
            // scatter (fvbaselinearizer.hh:495-508)
            R[2 * cell] += res[0][0].v;
            R[2 * cell + 1] += res[0][1].v;
            double d[7][2][2];  // [dof][eq][pv]
            for (int l = 0; l < nd; ++l)
                for (int k = 0; k < 2; ++k)
                    for (int pv = 0; pv < 2; ++pv) d[l][k][pv] = res[l][k].d[pv];
            if (fd) {
                for (int pv = 0; pv < 2; ++pv) {
                    const double eps = baseEpsilon / pvWeight[pv];
                    double pvs[2] = {x[2 * cell], x[2 * cell + 1]};
                    double delta = 0.0;
                    Ev hi[7][2], lo[7][2];
                    int nd2 = 0;
                    int64_t dofs2[7];
                    if (method >= 0) {
                        pvs[pv] += eps;
                        delta += eps;
                        element_local(c, cell, ijk, intensive_at(c, cell, pvs[0], pvs[1]), iqs,
                                      Mold, dt, nd2, dofs2, hi);
                    } else {
                        for (int l = 0; l < nd; ++l) hi[l][0] = res[l][0], hi[l][1] = res[l][1];
                    }
                    if (method <= 0) {
                        pvs[pv] -= delta + eps;
                        delta += eps;
                        element_local(c, cell, ijk, intensive_at(c, cell, pvs[0], pvs[1]), iqs,
                                      Mold, dt, nd2, dofs2, lo);
                    } else {
                        for (int l = 0; l < nd; ++l) lo[l][0] = res[l][0], lo[l][1] = res[l][1];
                    }
                    for (int l = 0; l < nd; ++l)
                        for (int k = 0; k < 2; ++k) {
                            double v = hi[l][k].v;
                            v -= lo[l][k].v;
                            d[l][k][pv] = v / delta;
                        }
                }
            }
            for (int l = 0; l < nd; ++l) {
                int64_t p = find_block(c, dofs[l], cell);
                double *blk = J + 4 * p;
                for (int k = 0; k < 2; ++k)
                    for (int pv = 0; pv < 2; ++pv) blk[2 * k + pv] += d[l][k][pv];
            }
        }
    });
}

void build_pattern(ewo_ctx &c) {
    c.rowptr.assign(c.N + 1, 0);
    c.colidx.clear();
    c.diagpos.assign(c.N, 0);
    for (int64_t cell = 0; cell < c.N; ++cell) {
        int ijk[3];
        cell_coords(c, cell, ijk);
        std::vector<int64_t> cols{cell};
        for (int dir = 0; dir < 2 * c.dim; ++dir) {
            int64_t nb = neighbour(c, ijk, dir);
            if (nb >= 0) cols.push_back(nb);
        }
        std::sort(cols.begin(), cols.end());
        for (int64_t col : cols) {
            if (col == cell) c.diagpos[cell] = static_cast<int32_t>(c.colidx.size());
            c.colidx.push_back(static_cast<int32_t>(col));
        }
        c.rowptr[cell + 1] = static_cast<int32_t>(c.colidx.size());
    }
}

// ------------------------------ block linear algebra ------------------------------
template <class T>
inline void blk_mul(const T *a, const T *b, T *out) {  // out = a*b (DenseMatrix loop order)
    T r[4];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            T s = 0;
            for (int k = 0; k < 2; ++k) s += a[2 * i + k] * b[2 * k + j];
            r[2 * i + j] = s;
        }
    std::memcpy(out, r, sizeof(r));
}
template <class T>
inline void blk_invert(T *a) {  // 2x2 inverse via the adjugate (dune-common densematrix)
    T det = a[0] * a[3] - a[1] * a[2];
    if (!(std::fabs(static_cast<double>(det)) >= 1e-30))
        throw NumericalIssue("ILU0: singular diagonal block");
    T det_1 = T(1) / det;
    T i00 = a[3] * det_1, i01 = -a[1] * det_1, i10 = -a[2] * det_1, i11 = a[0] * det_1;
    a[0] = i00; a[1] = i01; a[2] = i10; a[3] = i11;
}
template <class T>
inline void blk_mmv(const T *a, const T *x, T *y) {  // y -= a x
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) y[i] -= a[2 * i + j] * x[j];
}
template <class T>
inline void blk_mv(const T *a, const T *x, T *y) {  // y = a x
    for (int i = 0; i < 2; ++i) {
        T s = 0;
        for (int j = 0; j < 2; ++j) s += a[2 * i + j] * x[j];
        y[i] = s;
    }
}

// block ILU(0) in natural order (dune-istl blockILU0Decomposition semantics:
// L_ij = A_ij A_jj^-1, A_ik -= L_ij A_jk on the pattern, diagonal stored inverted)
template <class T>
void ilu0_factor(const ewo_ctx &c, T *LU) {
    for (int64_t i = 0; i < c.N; ++i) {
        for (int32_t ij = c.rowptr[i]; ij < c.rowptr[i + 1] && c.colidx[ij] < i; ++ij) {
            int64_t j = c.colidx[ij];
            const T *ajj = LU + 4 * static_cast<int64_t>(c.diagpos[j]);
            blk_mul(LU + 4 * static_cast<int64_t>(ij), ajj, LU + 4 * static_cast<int64_t>(ij));
            int32_t ik = ij + 1, jk = c.diagpos[j] + 1;
            int32_t ie = c.rowptr[i + 1], je = c.rowptr[j + 1];
            while (ik < ie && jk < je) {
                if (c.colidx[ik] == c.colidx[jk]) {
                    T Bm[4];
                    blk_mul(LU + 4 * static_cast<int64_t>(ij), LU + 4 * static_cast<int64_t>(jk), Bm);
                    T *aik = LU + 4 * static_cast<int64_t>(ik);
                    for (int e = 0; e < 4; ++e) aik[e] -= Bm[e];
                    ++ik;
                    ++jk;
                } else if (c.colidx[ik] < c.colidx[jk]) {
                    ++ik;
                } else {
                    ++jk;
                }
            }
        }
        blk_invert(LU + 4 * static_cast<int64_t>(c.diagpos[i]));
    }
}

template <class T>
void ilu0_apply(const ewo_ctx &c, const T *LU, const T *d, T *v) {
    for (int64_t i = 0; i < c.N; ++i) {
        T rhs[2] = {d[2 * i], d[2 * i + 1]};
        for (int32_t p = c.rowptr[i]; p < c.diagpos[i]; ++p)
            blk_mmv(LU + 4 * static_cast<int64_t>(p), v + 2 * static_cast<int64_t>(c.colidx[p]), rhs);
        v[2 * i] = rhs[0];
        v[2 * i + 1] = rhs[1];
    }
    for (int64_t i = c.N - 1; i >= 0; --i) {
        T rhs[2] = {v[2 * i], v[2 * i + 1]};
        for (int32_t p = c.diagpos[i] + 1; p < c.rowptr[i + 1]; ++p)
            blk_mmv(LU + 4 * static_cast<int64_t>(p), v + 2 * static_cast<int64_t>(c.colidx[p]), rhs);
        blk_mv(LU + 4 * static_cast<int64_t>(c.diagpos[i]), rhs, v + 2 * i);
    }
}

void spmv(const ewo_ctx &c, const double *A, const double *x, double *y) {
    for (int64_t i = 0; i < c.N; ++i) {
        double s[2] = {0.0, 0.0};
        for (int32_t p = c.rowptr[i]; p < c.rowptr[i + 1]; ++p) {
            const double *a = A + 4 * static_cast<int64_t>(p);
            const double *xv = x + 2 * static_cast<int64_t>(c.colidx[p]);
            for (int r = 0; r < 2; ++r)
                for (int q = 0; q < 2; ++q) s[r] += a[2 * r + q] * xv[q];
        }
        y[2 * i] = s[0];
        y[2 * i + 1] = s[1];
    }
}

double dot(const std::vector<double> &a, const std::vector<double> &b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

// preconditioner: FP64 or FP32 factors (LinearSolverScalar cast of
// overlappingbcrsmatrix.hh:250-254 applied to the preconditioner only)
struct Precond {
    const ewo_ctx *c = nullptr;
    bool fp32 = false;
    std::vector<double> lu64;
    std::vector<float> lu32;
    mutable std::vector<float> d32, v32;
    void prepare(const ewo_ctx &ctx, const double *A) {
        c = &ctx;
        fp32 = ctx.prm.precond_fp32 != 0;
        size_t n = 4 * ctx.colidx.size();
        if (fp32) {
            lu32.resize(n);
            for (size_t i = 0; i < n; ++i) lu32[i] = static_cast<float>(A[i]);
            ilu0_factor(ctx, lu32.data());
        } else {
            lu64.assign(A, A + n);
            ilu0_factor(ctx, lu64.data());
        }
    }
    void apply(const double *d, double *v) const {
        if (fp32) {
            size_t n = 2 * static_cast<size_t>(c->N);
            d32.resize(n);
            v32.resize(n);
            for (size_t i = 0; i < n; ++i) d32[i] = static_cast<float>(d[i]);
            ilu0_apply(*c, lu32.data(), d32.data(), v32.data());
            for (size_t i = 0; i < n; ++i) v[i] = static_cast<double>(v32[i]);
        } else {
            ilu0_apply(*c, lu64.data(), d, v);
        }
    }
};

// CombinedCriterion (combinedcriterion.hh:112-205)
struct Criterion {
    double relTol, absTol, maxResid;
    double residualError = 0, initialResidualError = 0, lastResidualError = 0;
    bool stagnates = false;
    void update(const std::vector<double> &change, const std::vector<double> &resid) {
        lastResidualError = residualError;
        residualError = 0.0;
        stagnates = true;
        for (size_t i = 0; i < resid.size(); ++i) {
            residualError = std::max(residualError, std::fabs(resid[i]));
            if (stagnates && change[i] != 0.0) stagnates = false;
        }
    }
    void set_initial(const std::vector<double> &x, const std::vector<double> &r) {
        update(x, r);
        stagnates = false;
        residualError = std::max(residualError, std::numeric_limits<double>::min() * 1e10);
        initialResidualError = residualError;
        lastResidualError = residualError;
    }
    double reduction() const { return residualError / std::max(1e-20, initialResidualError); }
    bool converged() const { return reduction() <= relTol || residualError <= absTol; }
    bool failed() const { return !converged() && (stagnates || residualError > maxResid); }
};

// BiCGStabSolver::apply (bicgstabsolver.hh:116-326), with its aliasing h == x, s == r,
// t == y. Returns true if converged; throws NumericalIssue on breakdown.
bool bicgstab(ewo_ctx &c, const double *A, const std::vector<double> &b, std::vector<double> &x,
              int &iters) {
    const double breakdownEps = std::numeric_limits<double>::min() * 1e10;
    const size_t n = b.size();
    double absTol = c.prm.linear_abs_tolerance;
    if (absTol < 0.0) absTol = c.prm.newton_tolerance / 100.0;  // parallelbicgstabbackend.hh:129-132
    Criterion crit{c.prm.linear_tolerance, absTol, c.prm.linear_max_error};
    Precond M;
    try {
        M.prepare(c, A);
    } catch (const NumericalIssue &) {
        throw NumericalIssue("Creating the preconditioner failed");
    }
    x.assign(n, 0.0);
    std::vector<double> r(b);
    iters = 0;
    crit.set_initial(x, r);
    if (crit.converged()) return true;
    const std::vector<double> &r0hat = b;
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    std::vector<double> v(n, 0.0), p(n, 0.0), y(n, 0.0), z(n, 0.0);
    for (int it = 0; it < c.prm.linear_max_iterations; ++it) {
        iters = it;
        double rho_i = dot(r0hat, r);
        if (std::fabs(rho) <= breakdownEps || std::fabs(omega) <= breakdownEps)
            throw NumericalIssue("Breakdown of the BiCGStab solver (division by zero)");
        double beta = (rho_i / rho) * (alpha / omega);
        rho = rho_i;
        for (size_t i = 0; i < n; ++i) {
            double tmp = v[i];
            tmp *= omega;
            tmp -= p[i];
            tmp *= -beta;
            p[i] = r[i];
            p[i] += tmp;
        }
        M.apply(p.data(), y.data());
        spmv(c, A, y.data(), v.data());
        double denom = dot(r0hat, v);
        if (std::fabs(denom) <= breakdownEps)
            throw NumericalIssue("Breakdown of the BiCGStab solver (division by zero)");
        alpha = rho_i / denom;
        if (std::fabs(alpha) <= breakdownEps)
            throw NumericalIssue("Breakdown of the BiCGStab solver (stagnation detected)");
        for (size_t i = 0; i < n; ++i) {
            double tmp = y[i];
            tmp *= alpha;
            tmp += x[i];
            x[i] = tmp;
            tmp = v[i];
            tmp *= alpha;
            r[i] -= tmp;
        }
        crit.update(y, r);
        if (crit.converged()) return true;
        if (crit.failed()) return false;
        M.apply(r.data(), z.data());
        spmv(c, A, z.data(), y.data());  // t == y
        denom = dot(y, y);
        if (std::fabs(denom) <= breakdownEps)
            throw NumericalIssue("Breakdown of the BiCGStab solver (division by zero)");
        omega = dot(y, r) / denom;
        if (std::fabs(omega) <= breakdownEps)
            throw NumericalIssue("Breakdown of the BiCGStab solver (stagnation detected)");
        for (size_t i = 0; i < n; ++i) x[i] += omega * z[i];
        crit.update(z, r);  // r is still s here
        if (crit.converged()) return true;
        if (crit.failed()) return false;
        for (size_t i = 0; i < n; ++i) r[i] += -omega * y[i];
    }
    iters = c.prm.linear_max_iterations;
    return false;
}

double max_norm(const std::vector<double> &r) {
    double e = 0.0;
    for (double v : r) e = std::max(std::fabs(v), e);
    return e;
}

bool proceed(const ewo_ctx &c) {  // newtonmethod.hh:866-884
    if (c.num_iterations < 1) return true;
    if (c.error <= c.prm.newton_tolerance) return false;
    if (c.num_iterations >= c.prm.newton_max_iterations) return c.error * 4.0 < c.last_error;
    return true;
}

// NewtonMethod::apply (newtonmethod.hh:308-517); returns EWG_OK or a failure code
int newton(ewo_ctx &c, double dt) {
    c.num_iterations = 0;
    c.newton_errors.clear();
    c.pref = c.x0[0];  // ImmiscibleModel::updateBegin (immisciblemodel.hh:278-293)
    std::vector<double> cur;
    std::vector<double> dx;
    c.resid.assign(2 * c.N, 0.0);
    c.jac.assign(4 * c.colidx.size(), 0.0);
    try {
        while (proceed(c)) {
            c.last_error = c.error;  // beginIteration_
            cur = c.x0;
            linearize_all(c, dt, c.num_iterations, c.resid.data(), c.jac.data());
            c.have_system = true;
            c.stats.newton_iterations++;
            c.last_error = c.error;  // preSolve_
            c.error = max_norm(c.resid);
            c.newton_errors.push_back(c.error);
            if (c.error > c.prm.newton_max_error) throw NumericalIssue("Newton: error too large");
            if (!proceed(c)) {
                ++c.num_iterations;
                break;
            }
            int its = 0;
            bool conv = bicgstab(c, c.jac.data(), c.resid, dx, its);
            c.last_linear_iterations = its;
            c.stats.linear_iterations += its;
            c.stats.linear_solves++;
            if (!conv) {
                c.num_iterations = 2 * c.prm.newton_target_iterations;
                return EWG_NOT_CONVERGED;
            }
            double one = 0.0;
            for (double v : dx) one += std::fabs(v);
            if (!std::isfinite(one)) throw NumericalIssue("Non-finite update!");
            for (size_t i = 0; i < dx.size(); ++i) {
                double nv = cur[i];
                nv -= dx[i];
                c.x0[i] = nv;
            }
            ++c.num_iterations;
        }
    } catch (const NumericalIssue &) {
        c.num_iterations = 2 * c.prm.newton_target_iterations;
        return EWG_NUMERICAL_ISSUE;
    }
    if (!(c.error <= c.prm.newton_tolerance)) {
        c.num_iterations = 2 * c.prm.newton_target_iterations;
        return EWG_NOT_CONVERGED;
    }
    return EWG_OK;
}

double suggest_dt(const ewo_ctx &c, double oldDt) {  // newtonmethod.hh:527-544
    int tgt = c.prm.newton_target_iterations;
    if (c.num_iterations > tgt) {
        double percent = static_cast<double>(c.num_iterations - tgt) / tgt;
        return std::max(c.prm.min_time_step_size, oldDt / (1.0 + percent));
    }
    double percent = static_cast<double>(tgt - c.num_iterations) / tgt;
    return std::max(c.prm.min_time_step_size, oldDt * (1.0 + percent / 1.2));
}

int fail(ewo_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

}  // namespace

// -------------------------------------------------------------------------------------
// C API
// -------------------------------------------------------------------------------------
extern "C" {

int ewo_set_threads(ewo_ctx *c, int threads) {
    if (!c || threads < 1) return EWG_INVALID_ARGUMENT;
    c->threads = threads;
    return EWG_OK;
}

double ewo_log(double x) { return d_log(x); }
double ewo_exp(double x) { return d_exp(x); }
double ewo_pow(double x, double y) { return d_pow(x, y); }

int ewo_create(const ewg_grid_desc *grid, const ewg_fluid_desc *fluid,
               const ewg_material_desc *material, const ewg_params *params, ewo_ctx **out) {
    if (!grid || !fluid || !material || !params || !out) return EWG_INVALID_ARGUMENT;
    if (grid->dim != 2 && grid->dim != 3) return EWG_INVALID_ARGUMENT;
    if (grid->nx < 1 || grid->ny < 1 || grid->nz < 1) return EWG_INVALID_ARGUMENT;
    if (grid->dim == 2 && grid->nz != 1) return EWG_INVALID_ARGUMENT;
    if (material->nregions < 1 || !material->region_params) return EWG_INVALID_ARGUMENT;
    auto *c = new ewo_ctx();
    c->grid = *grid;
    c->fluid = *fluid;
    c->prm = *params;
    c->dim = grid->dim;
    c->nx = grid->nx;
    c->ny = grid->ny;
    c->nz = grid->nz;
    c->N = static_cast<int64_t>(c->nx) * c->ny * c->nz;
    c->h[0] = grid->lx / c->nx;
    c->h[1] = grid->ly / c->ny;
    c->h[2] = (c->dim == 3) ? grid->lz / c->nz : 1.0;
    if (c->dim == 2) {
        c->V = c->h[0] * c->h[1];
        c->area[0] = c->h[1];
        c->area[1] = c->h[0];
        c->area[2] = 0.0;
    } else {
        c->V = c->h[0] * c->h[1] * c->h[2];
        c->area[0] = c->h[1] * c->h[2];
        c->area[1] = c->h[0] * c->h[2];
        c->area[2] = c->h[0] * c->h[1];
    }
    int64_t side[6];
    side[0] = side[1] = static_cast<int64_t>(c->ny) * c->nz;
    side[2] = side[3] = static_cast<int64_t>(c->nx) * c->nz;
    side[4] = side[5] = (c->dim == 3) ? static_cast<int64_t>(c->nx) * c->ny : 0;
    c->side_off[0] = 0;
    for (int s = 0; s < 6; ++s) c->side_off[s + 1] = c->side_off[s] + side[s];
    c->B = c->side_off[6];
    for (int r = 0; r < material->nregions; ++r)
        c->regs.push_back(make_region(material->law, material->region_params + 4 * r));
    c->x0.assign(2 * c->N, 0.0);
    c->x1.assign(2 * c->N, 0.0);
    c->storage_old.assign(2 * c->N, 0.0);
    c->bc_type.assign(c->B, EWG_BC_NO_FLOW);
    c->bc_val.assign(2 * c->B, 0.0);
    build_pattern(*c);
    *out = c;
    return EWG_OK;
}

void ewo_destroy(ewo_ctx *ctx) { delete ctx; }
const char *ewo_last_error(const ewo_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }
int ewo_set_params(ewo_ctx *ctx, const ewg_params *params) {
    if (!ctx || !params) return EWG_INVALID_ARGUMENT;
    ctx->prm = *params;
    return EWG_OK;
}

int ewo_set_static(ewo_ctx *c, const double *K, const double *phi, const int32_t *region,
                   const double *pe) {
    if (!c || !K || !phi || !region) return EWG_INVALID_ARGUMENT;
    c->K.assign(K, K + c->N);
    c->phi.assign(phi, phi + c->N);
    c->region.assign(region, region + c->N);
    for (int64_t i = 0; i < c->N; ++i)
        if (c->region[i] < 0 || c->region[i] >= static_cast<int32_t>(c->regs.size()))
            return fail(c, EWG_INVALID_ARGUMENT, "region index out of range");
    if (pe) c->pe.assign(pe, pe + c->N);
    else c->pe.clear();
    c->static_set = true;
    return EWG_OK;
}

int ewo_set_boundary(ewo_ctx *c, const int32_t *type, const double *values) {
    if (!c || !type || !values) return EWG_INVALID_ARGUMENT;
    c->bc_type.assign(type, type + c->B);
    c->bc_val.assign(values, values + 2 * c->B);
    return EWG_OK;
}

int ewo_set_solution(ewo_ctx *c, int time_idx, const double *x) {
    if (!c || !x || (time_idx != 0 && time_idx != 1)) return EWG_INVALID_ARGUMENT;
    (time_idx == 0 ? c->x0 : c->x1).assign(x, x + 2 * c->N);
    return EWG_OK;
}
int ewo_get_solution(ewo_ctx *c, int time_idx, double *x) {
    if (!c || !x || (time_idx != 0 && time_idx != 1)) return EWG_INVALID_ARGUMENT;
    const auto &v = time_idx == 0 ? c->x0 : c->x1;
    std::copy(v.begin(), v.end(), x);
    return EWG_OK;
}
int ewo_advance_time_level(ewo_ctx *c) {
    if (!c) return EWG_INVALID_ARGUMENT;
    c->x1 = c->x0;
    return EWG_OK;
}
int64_t ewo_num_cells(const ewo_ctx *c) { return c ? c->N : 0; }
int64_t ewo_num_blocks(const ewo_ctx *c) { return c ? static_cast<int64_t>(c->colidx.size()) : 0; }
int64_t ewo_num_boundary_faces(const ewo_ctx *c) { return c ? c->B : 0; }

int ewo_get_pattern(const ewo_ctx *c, int32_t *rowptr, int32_t *colidx) {
    if (!c || !rowptr || !colidx) return EWG_INVALID_ARGUMENT;
    std::copy(c->rowptr.begin(), c->rowptr.end(), rowptr);
    std::copy(c->colidx.begin(), c->colidx.end(), colidx);
    return EWG_OK;
}

int ewo_linearize(ewo_ctx *c, double dt, int newton_iter, double *resid, double *jac) {
    if (!c || !(dt > 0.0)) return EWG_INVALID_ARGUMENT;
    if (!c->static_set) return fail(c, EWG_INVALID_STATE, "static fields not set");
    c->resid.assign(2 * c->N, 0.0);
    c->jac.assign(4 * c->colidx.size(), 0.0);
    c->pref = c->x0[0];
    linearize_all(*c, dt, newton_iter, c->resid.data(), c->jac.data());
    c->have_system = true;
    if (resid) std::copy(c->resid.begin(), c->resid.end(), resid);
    if (jac) std::copy(c->jac.begin(), c->jac.end(), jac);
    return EWG_OK;
}

int ewo_storage(const ewo_ctx *c, const double *x, double *M) {
    if (!c || !x || !M) return EWG_INVALID_ARGUMENT;
    for (int64_t i = 0; i < c->N; ++i) {
        IQ q = intensive(*c, i, x);
        Ev m[2];
        storage_of(*c, i, q, m);
        M[2 * i] = m[0].v;
        M[2 * i + 1] = m[1].v;
    }
    return EWG_OK;
}

int ewo_spmv(const ewo_ctx *c, const double *A, const double *x, double *y) {
    if (!c || !A || !x || !y) return EWG_INVALID_ARGUMENT;
    spmv(*c, A, x, y);
    return EWG_OK;
}

int ewo_ilu0_factor(const ewo_ctx *c, const double *A, double *LU) {
    if (!c || !A || !LU) return EWG_INVALID_ARGUMENT;
    std::copy(A, A + 4 * c->colidx.size(), LU);
    try {
        ilu0_factor(*c, LU);
    } catch (const NumericalIssue &) {
        return EWG_NUMERICAL_ISSUE;
    }
    return EWG_OK;
}

int ewo_ilu0_apply(const ewo_ctx *c, const double *LU, const double *d, double *v) {
    if (!c || !LU || !d || !v) return EWG_INVALID_ARGUMENT;
    ilu0_apply(*c, LU, d, v);
    return EWG_OK;
}

// The FP32 preconditioner's arithmetic: factors and right-hand side rounded to float,
// both sweeps in float, the result widened (the form libewg runs with precond_fp32).
int ewo_ilu0_apply32(const ewo_ctx *c, const float *LU, const double *d, double *v) {
    if (!c || !LU || !d || !v) return EWG_INVALID_ARGUMENT;
    std::vector<float> df(2 * c->N), vf(2 * c->N);
    for (int64_t k = 0; k < 2 * c->N; ++k) df[k] = static_cast<float>(d[k]);
    ilu0_apply(*c, LU, df.data(), vf.data());
    for (int64_t k = 0; k < 2 * c->N; ++k) v[k] = static_cast<double>(vf[k]);
    return EWG_OK;
}

int ewo_bicgstab(ewo_ctx *c, const double *A, const double *b, double *x, int32_t *iterations) {
    if (!c || !A || !b || !x) return EWG_INVALID_ARGUMENT;
    std::vector<double> bv(b, b + 2 * c->N), xv;
    int its = 0;
    bool conv;
    try {
        conv = bicgstab(*c, A, bv, xv, its);
    } catch (const NumericalIssue &e) {
        return fail(c, EWG_NUMERICAL_ISSUE, e.what());
    }
    std::copy(xv.begin(), xv.end(), x);
    if (iterations) *iterations = its;
    return conv ? EWG_OK : EWG_NOT_CONVERGED;
}

// ParallelBaseBackend::prepare/setResidual/setMatrix (parallelbasebackend.hh:231-312)
int ewo_set_system(ewo_ctx *c, const double *resid, const double *jac) {
    if (!c) return EWG_INVALID_ARGUMENT;
    if (resid) c->resid.assign(resid, resid + 2 * c->N);
    if (jac) c->jac.assign(jac, jac + 4 * c->colidx.size());
    if (c->resid.size() != (size_t)(2 * c->N) || c->jac.size() != 4 * c->colidx.size())
        return fail(c, EWG_INVALID_STATE, "incomplete system");
    c->have_system = true;
    return EWG_OK;
}

int ewo_solve(ewo_ctx *c, double *dx, int32_t *iterations) {
    if (!c) return EWG_INVALID_ARGUMENT;
    if (!c->have_system) return fail(c, EWG_INVALID_STATE, "no linearization");
    std::vector<double> xv;
    int its = 0;
    bool conv;
    try {
        conv = bicgstab(*c, c->jac.data(), c->resid, xv, its);
    } catch (const NumericalIssue &e) {
        return fail(c, EWG_NUMERICAL_ISSUE, e.what());
    }
    if (dx) std::copy(xv.begin(), xv.end(), dx);
    if (iterations) *iterations = its;
    return conv ? EWG_OK : EWG_NOT_CONVERGED;
}

int ewo_newton(ewo_ctx *c, double dt, int32_t *iterations, double *final_error) {
    if (!c || !(dt > 0.0)) return EWG_INVALID_ARGUMENT;
    int rc = newton(*c, dt);
    if (iterations) *iterations = c->num_iterations;
    if (final_error) *final_error = c->error;
    c->stats.last_newton_iterations = c->num_iterations;
    c->stats.last_newton_error = c->error;
    return rc;
}

int ewo_run(ewo_ctx *c, double t0, double t_end, double dt0, int32_t *steps_out) {
    if (!c || !(dt0 > 0.0)) return EWG_INVALID_ARGUMENT;
    c->step_its.clear();
    c->step_dts.clear();
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
        for (int fails = 0; fails < c->prm.max_time_step_divisions; ++fails) {  // fvbaseproblem.hh:507-553
            int rc = newton(*c, dt);
            if (rc == EWG_OK) {
                ok = true;
                break;
            }
            c->x0 = c->x1;  // updateFailed (fvbasediscretization.hh:1370-1387)
            c->stats.failed_steps++;
            double nextDt = dt / 2.0;
            if (dt < c->prm.min_time_step_size * (1 + 1e-9))
                return fail(c, EWG_NOT_CONVERGED, "time integration failed at minimum time step");
            if (nextDt < c->prm.min_time_step_size) nextDt = c->prm.min_time_step_size;
            dt = nextDt;
        }
        if (!ok) return fail(c, EWG_NOT_CONVERGED, "Newton did not converge after dt divisions");
        c->x1 = c->x0;  // advanceTimeLevel
        c->step_its.push_back(c->num_iterations);
        c->step_dts.push_back(dt);
        c->stats.time_steps++;
        t += dt;
        ++steps;
        // FvBaseProblem::nextTimeStepSize (fvbaseproblem.hh:586-600) + simulator.hh:809-819
        double dtNext = std::min(c->prm.max_time_step_size, suggest_dt(*c, dt));
        double simMax = remaining(t);
        if (dtNext < simMax && simMax < dtNext * 2) dtNext = simMax / 2 * 1.01;
        dt = std::min(simMax, dtNext);
        if (finished(t, dt)) break;
    }
    if (steps_out) *steps_out = steps;
    return EWG_OK;
}

int ewo_step_log(const ewo_ctx *c, int32_t *newton_its, double *dts, int32_t cap) {
    if (!c) return 0;
    int32_t n = std::min<int32_t>(cap, static_cast<int32_t>(c->step_its.size()));
    for (int32_t i = 0; i < n; ++i) {
        if (newton_its) newton_its[i] = c->step_its[i];
        if (dts) dts[i] = c->step_dts[i];
    }
    return n;
}

int ewo_newton_history(const ewo_ctx *c, double *errs, int32_t cap) {
    if (!c) return 0;
    int32_t n = std::min<int32_t>(cap, static_cast<int32_t>(c->newton_errors.size()));
    for (int32_t i = 0; i < n; ++i) errs[i] = c->newton_errors[i];
    return n;
}

int ewo_get_stats(const ewo_ctx *c, ewg_stats *s) {
    if (!c || !s) return EWG_INVALID_ARGUMENT;
    *s = c->stats;
    s->last_newton_iterations = c->num_iterations;
    s->last_linear_iterations = c->last_linear_iterations;
    s->last_newton_error = c->error;
    return EWG_OK;
}

int ewo_material(const ewo_ctx *c, int region, double pe, double Sw, double *out) {
    if (!c || !out || region < 0 || region >= static_cast<int>(c->regs.size()))
        return EWG_INVALID_ARGUMENT;
    const MatReg &r = c->regs[region];
    Ev s = Ev::var(Sw, 1);
    Ev pc = pcnw(r, pe, s);
    Ev kr[2];
    relperm(r, s, kr);
    out[0] = pc.v;
    out[1] = pc.d[1];
    out[2] = kr[0].v;
    out[3] = kr[0].d[1];
    out[4] = kr[1].v;
    out[5] = kr[1].d[1];
    out[6] = eff_sat(r, s).v;
    out[7] = 0.0;
    out[8] = 0.0;
    return EWG_OK;
}

}  // extern "C"
