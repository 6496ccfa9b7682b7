// ewg_physics.h -- scalar/dual arithmetic, material laws and the per-face flux of the
// ECFV immiscible two-phase model, shared by the host set-up code and the gfx950 kernels.
//
// Everything here is compiled with -ffp-contract=off: each rounding step is the one
// written, so the device results are reproducible and the upwind decisions (strict
// "> 0" on a potential difference) are taken on bit-identical values wherever the
// inputs are identical.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

#define EWG_HD __host__ __device__ __forceinline__

namespace ewg {

// ----------------------------------------------------------------------------------
// Elementary functions from +,-,*,/ and fma only. log: mantissa reduced to
// [sqrt(1/2), sqrt(2)], log m = 2 atanh(s), s = (m-1)/(m+1), series to s^23.
// exp: x = k ln2 + r with k = rint(x/ln2), ln2 split 32+53 bits, Taylor to r^13.
// pow(x, y) = exp(y log x). Accuracy: a few ulp for the arguments used here.
// ----------------------------------------------------------------------------------
constexpr double kLn2Hi = 0x1.62e42fee00000p-1;
constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;
constexpr double kInvLn2 = 0x1.71547652b82fep+0;
constexpr double kSqrt2 = 0x1.6a09e667f3bcdp+0;

EWG_HD double bits_to_double(uint64_t b) { return __builtin_bit_cast(double, b); }
EWG_HD uint64_t double_to_bits(double d) { return __builtin_bit_cast(uint64_t, d); }

EWG_HD double exp2i(int k) { return bits_to_double(static_cast<uint64_t>(k + 1023) << 52); }

EWG_HD double dexp(double x) {
    if (x != x) return x;
    if (x > 709.78) return __builtin_inf();
    if (x < -745.2) return 0.0;
    double kd = rint(x * kInvLn2);
    double r = fma(-kd, kLn2Hi, x);
    r = fma(-kd, kLn2Lo, r);
    double p = 1.0 / 6227020800.0;
    p = fma(p, r, 1.0 / 479001600.0);
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    int k = static_cast<int>(kd);
    if (k > 1023) return (p * exp2i(1023)) * exp2i(k - 1023);
    if (k < -1022) return (p * exp2i(-1022)) * exp2i(k + 1022);
    return p * exp2i(k);
}

EWG_HD double dlog(double x) {
    if (x != x || x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    uint64_t b = double_to_bits(x);
    int e = static_cast<int>((b >> 52) & 0x7ff);
    if (e == 0) {
        b = double_to_bits(x * 0x1p54);
        e = static_cast<int>((b >> 52) & 0x7ff) - 54;
    }
    e -= 1023;
    double m = bits_to_double((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > kSqrt2) {
        m = m * 0.5;
        e += 1;
    }
    double f = m - 1.0;
    double s = f / (m + 1.0);
    double s2 = s * s;
    double p = 2.0 / 23.0;
    p = fma(p, s2, 2.0 / 21.0);
    p = fma(p, s2, 2.0 / 19.0);
    p = fma(p, s2, 2.0 / 17.0);
    p = fma(p, s2, 2.0 / 15.0);
    p = fma(p, s2, 2.0 / 13.0);
    p = fma(p, s2, 2.0 / 11.0);
    p = fma(p, s2, 2.0 / 9.0);
    p = fma(p, s2, 2.0 / 7.0);
    p = fma(p, s2, 2.0 / 5.0);
    p = fma(p, s2, 2.0 / 3.0);
    double logm = fma(s * s2, p, 2.0 * s);
    double ed = static_cast<double>(e);
    return fma(ed, kLn2Hi, fma(ed, kLn2Lo, logm));
}

EWG_HD double dpow(double x, double y) {
    if (x > 0.0) return dexp(y * dlog(x));
    if (x == 0.0) return y > 0.0 ? 0.0 : (y == 0.0 ? 1.0 : __builtin_inf());
    return __builtin_nan("");
}

// ----------------------------------------------------------------------------------
// Forward-mode dual number: value + d/dp_w + d/dS_w of the focus cell. Rules:
// (uv)' = u'v + v'u, (u/v)' = (v u' - u v')/v^2; mixed scalar ops leave the
// derivatives untouched or scale them. Dual<double> (Ev) is the linearization's type,
// Dual<float> (Evf) the one of the FP32 flux form.
// ----------------------------------------------------------------------------------
template <class T>
struct Dual {
    T v, d0, d1;
};
using Ev = Dual<double>;
using Evf = Dual<float>;
EWG_HD Ev ev(double x) { return Ev{x, 0.0, 0.0}; }
template <class T>
EWG_HD Dual<T> operator+(Dual<T> a, Dual<T> b) { return Dual<T>{a.v + b.v, a.d0 + b.d0, a.d1 + b.d1}; }
template <class T>
EWG_HD Dual<T> operator-(Dual<T> a, Dual<T> b) { return Dual<T>{a.v - b.v, a.d0 - b.d0, a.d1 - b.d1}; }
template <class T>
EWG_HD Dual<T> operator-(Dual<T> a) { return Dual<T>{-a.v, -a.d0, -a.d1}; }
template <class T>
EWG_HD Dual<T> operator*(Dual<T> a, Dual<T> b) {
    return Dual<T>{a.v * b.v, a.d0 * b.v + b.d0 * a.v, a.d1 * b.v + b.d1 * a.v};
}
template <class T>
EWG_HD Dual<T> operator/(Dual<T> a, Dual<T> b) {
    T vv = b.v * b.v;
    Dual<T> r;
    r.d0 = (b.v * a.d0 - a.v * b.d0) / vv;
    r.d1 = (b.v * a.d1 - a.v * b.d1) / vv;
    r.v = a.v / b.v;
    return r;
}
template <class T>
EWG_HD Dual<T> operator+(Dual<T> a, T s) { return Dual<T>{a.v + s, a.d0, a.d1}; }
template <class T>
EWG_HD Dual<T> operator+(T s, Dual<T> a) { return Dual<T>{s + a.v, a.d0, a.d1}; }
template <class T>
EWG_HD Dual<T> operator-(Dual<T> a, T s) { return Dual<T>{a.v - s, a.d0, a.d1}; }
template <class T>
EWG_HD Dual<T> operator-(T s, Dual<T> a) { return Dual<T>{s - a.v, -a.d0, -a.d1}; }
template <class T>
EWG_HD Dual<T> operator*(Dual<T> a, T s) { return Dual<T>{a.v * s, a.d0 * s, a.d1 * s}; }
template <class T>
EWG_HD Dual<T> operator*(T s, Dual<T> a) { return Dual<T>{a.v * s, a.d0 * s, a.d1 * s}; }
template <class T>
EWG_HD Dual<T> operator/(Dual<T> a, T s) { return Dual<T>{a.v / s, a.d0 / s, a.d1 / s}; }
// precision changes of a dual
EWG_HD Evf to_f32(const Ev &a) {
    return Evf{static_cast<float>(a.v), static_cast<float>(a.d0), static_cast<float>(a.d1)};
}
EWG_HD Ev to_f64(const Evf &a) {
    return Ev{static_cast<double>(a.v), static_cast<double>(a.d0), static_cast<double>(a.d1)};
}

EWG_HD double vpow(double x, double y) { return dpow(x, y); }
EWG_HD Ev vpow(Ev x, double y) {
    double pw = dpow(x.v, y);
    double df = (x.v == 0.0) ? 0.0 : y * pw / x.v;
    return Ev{pw, x.d0 * df, x.d1 * df};
}
EWG_HD double vsqrt(double x) { return sqrt(x); }
EWG_HD Ev vsqrt(Ev x) {
    double s = sqrt(x.v);
    double ds = (s == 0.0) ? 0.0 : 0.5 / s;
    return Ev{s, x.d0 * ds, x.d1 * ds};
}
EWG_HD double val(double x) { return x; }
EWG_HD double val(Ev x) { return x.v; }
EWG_HD double lift(double, double x) { return x; }
EWG_HD Ev lift(Ev, double x) { return ev(x); }

// ----------------------------------------------------------------------------------
// Material laws (EffToAbsLaw over regularized van Genuchten / Brooks-Corey).
// Regularization: pc linear below Se = 0.01 (slope there); van Genuchten: cubic
// Hermite on [0.99, 1] to (0, -2 pc(0.99)/0.01), linear beyond 1; Brooks-Corey: linear
// with the slope at Se = 1 beyond 1. kr clamped to their limits outside (0, 1).
// ----------------------------------------------------------------------------------
constexpr double kSeLow = 0.01;
constexpr double kSeHigh = 0.99;
constexpr int kLawVG = 0;
constexpr int kLawBC = 1;

struct MatReg {
    int law;
    int pad_;
    double p1, Swr, Snr, denom;
    double alpha, n, m, e_pc1, e_pc2, e_kr1, e_kr2, e_krn1, e_krn2;
    double pcLow, dpcLow, pcHigh, dpcHigh, mEnd;
    double lambda, e_pc, e_krw, e_krn, cLow, dLow;
};

inline double vg_pc_raw(const MatReg &r, double Se) {
    return dpow(dpow(Se, r.e_pc1) - 1.0, r.e_pc2) / r.alpha;
}
inline double vg_dpc_raw(const MatReg &r, double Se) {
    double a = dpow(Se, r.e_pc1);
    double b = a - 1.0;
    double pc = dpow(b, r.e_pc2) / r.alpha;
    return -pc * a / (r.n * r.m * b * Se);
}

// host: derived constants of a region from (p1, p2, Swr, Snr)
inline MatReg make_region(int law, const double *p) {
    MatReg r{};
    r.law = law;
    r.p1 = p[0];
    r.Swr = p[2];
    r.Snr = p[3];
    r.denom = 1.0 - r.Swr - r.Snr;
    if (law == kLawVG) {
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
        r.cLow = dpow(kSeLow, r.e_pc);
        r.dLow = r.e_pc * r.cLow / kSeLow;
    }
    return r;
}

template <class T>
EWG_HD T eff_sat(const MatReg &r, T Sw) { return (Sw - r.Swr) / r.denom; }

template <class T>
EWG_HD T pcnw(const MatReg &r, double pe, T Sw) {
    T Se = eff_sat(r, Sw);
    if (r.law == kLawVG) {
        if (val(Se) < kSeLow) return r.pcLow + r.dpcLow * (Se - kSeLow);
        if (val(Se) <= kSeHigh) return vpow(vpow(Se, r.e_pc1) - 1.0, r.e_pc2) / r.alpha;
        if (val(Se) < 1.0) {
            const double h = 1.0 - kSeHigh;
            T t = (Se - kSeHigh) / h;
            T t2 = t * t;
            T t3 = t2 * t;
            T h00 = (2.0 * t3 - 3.0 * t2) + 1.0;
            T h10 = (t3 - 2.0 * t2) + t;
            T h11 = t3 - t2;
            return (h00 * r.pcHigh + h10 * (h * r.dpcHigh)) + h11 * (h * r.mEnd);
        }
        return r.mEnd * (Se - 1.0);
    }
    double peu = pe > 0.0 ? pe : r.p1;
    if (val(Se) < kSeLow) return peu * (r.cLow + r.dLow * (Se - kSeLow));
    if (val(Se) <= 1.0) return peu * vpow(Se, r.e_pc);
    return peu * (1.0 + r.e_pc * (Se - 1.0));
}

template <class T>
EWG_HD void relperm(const MatReg &r, T Sw, T &krw, T &krn) {
    T Se = eff_sat(r, Sw);
    if (val(Se) <= 0.0) { krw = lift(Se, 0.0); krn = lift(Se, 1.0); return; }
    if (val(Se) >= 1.0) { krw = lift(Se, 1.0); krn = lift(Se, 0.0); return; }
    if (r.law == kLawVG) {
        T a = 1.0 - vpow(1.0 - vpow(Se, r.e_kr1), r.e_kr2);
        krw = vsqrt(Se) * (a * a);
        krn = vpow(1.0 - Se, r.e_krn1) * vpow(1.0 - vpow(Se, r.e_kr1), r.e_krn2);
    } else {
        krw = vpow(Se, r.e_krw);
        T t = 1.0 - Se;
        krn = (t * t) * (1.0 - vpow(Se, r.e_krn));
    }
}

// harmonic mean (entry-wise intersection permeability)
EWG_HD double harmonic_mean(double x, double y) {
    if (x * y <= 0.0) return 0.0;
    return (2.0 * x * y) / (x + y);
}

}  // namespace ewg
