// Device-side building blocks of the RoadRunner model for gfx950 (CDNA4).
//
// Geometry, limb-darkening interpolation, Kepler/Taylor orbit and the epoch
// reduction, written once and instantiated for FP64 and FP32 where the flux
// stage needs both.  Reference citations are to PyTransit 2.8.1.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rr_hip.h"

namespace rr {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kHalfPi = 0.5 * kPi;
constexpr float kPiF = 3.14159265358979323846f;

// ---- overloads so templates read the same for float and double -----------------------
__device__ __forceinline__ double d_sqrt(double x) { return sqrt(x); }
// FP32: the hardware square root (1 ulp) without sqrtf's denormal rescaling; the
// arguments here are squared distances and kite products, never denormal.
__device__ __forceinline__ float d_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// atan2(y, x) for y >= 0 in FP32: octant reduction with one hardware reciprocal
// and atan(t) = t P(t^2) on [0, 1], P of degree 7 fitted by Lawson's minimax
// iteration (pytransit_amd/csrc/atan_poly.py).  Max error 1.4e-7 rad.
__device__ __forceinline__ float atan2_ypos_f(float y, float x) {
    const float ax = fabsf(x);
    const float mx = fmaxf(ax, y), mn = fminf(ax, y);
    const float t = mx > 0.0f ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;
    const float s = t * t;
    float p = -0.00405456312f;
    p = fmaf(p, s, 0.0218629446f);
    p = fmaf(p, s, -0.0559123158f);
    p = fmaf(p, s, 0.0964219719f);
    p = fmaf(p, s, -0.139086291f);
    p = fmaf(p, s, 0.199465662f);
    p = fmaf(p, s, -0.333298594f);
    p = fmaf(p, s, 0.999999344f);
    float r = t * p;
    r = (y > ax) ? 1.57079633f - r : r;
    return (x < 0.0f) ? 3.14159265f - r : r;
}

// atan2 with y >= 0 (the kite angles): libm in FP64, the polynomial in FP32.
__device__ __forceinline__ double d_atan2_ypos(double y, double x) { return atan2(y, x); }
__device__ __forceinline__ float d_atan2_ypos(float y, float x) { return atan2_ypos_f(y, x); }
__device__ __forceinline__ double d_abs(double x) { return fabs(x); }
__device__ __forceinline__ float d_abs(float x) { return fabsf(x); }
__device__ __forceinline__ double d_max(double a, double b) { return fmax(a, b); }
// FP32 min/max as selects: no NaN canonicalisation (the FP32 geometry never sees NaN).
__device__ __forceinline__ float d_max(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ double d_min(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float d_min(float a, float b) { return a < b ? a : b; }
template <typename T> __device__ __forceinline__ T pi_of();
template <> __device__ __forceinline__ double pi_of<double>() { return kPi; }
template <> __device__ __forceinline__ float pi_of<float>() { return kPiF; }

// numpy.linspace element i of n: i*step + start with the last element exactly b.
__device__ __forceinline__ double linspace_at(double a, double b, int n, int i) {
#pragma clang fp contract(off)
    if (i == n - 1) return b;
    double step = (b - a) / (double)(n - 1);
    return (double)i * step + a;
}

// Area of the intersection of two circles; common.py:36-49 (FP64, init-time tables).
__device__ __forceinline__ double cci_area(double r1, double r2, double b) {
    if (r1 < b - r2) return 0.0;
    if (r1 >= b + r2) return kPi * (r2 * r2);
    if (b - r2 <= -r1) return kPi * (r1 * r1);
    return r2 * r2 * acos((b * b + r2 * r2 - r1 * r1) / (2.0 * b * r2)) +
           r1 * r1 * acos((b * b + r1 * r1 - r2 * r2) / (2.0 * b * r1)) -
           0.5 * sqrt((-b + r2 + r1) * (b + r2 - r1) * (b - r2 + r1) * (b + r2 + r1));
}

// Lens area and kappa_0 of circles r1, r2 at separation b by the kite formula
// (Agol et al. 2020); common.py:52-73 with tsort common.py:5-33.  akite is the
// area of the kite spanned by the centres and the intersection points (0 outside
// the lens case).
template <typename T>
__device__ __forceinline__ void ccia_kite3(T r1, T r2, T b, T& area, T& kap0, T& akite_out) {
    akite_out = T(0);
    if (r1 + r2 <= b) { area = T(0); kap0 = T(0); return; }
    if (d_abs(r1 - r2) < b && b <= r1 + r2) {
        // descending sort of (r1, r2, b): x >= y >= z
        const T hi12 = d_max(r1, r2), lo12 = d_min(r1, r2);
        const T x = d_max(hi12, b);
        const T z = d_min(lo12, b);
        const T y = d_max(lo12, d_min(hi12, b));
        const T akite = T(0.5) * d_sqrt((x + (y + z)) * (z - (x - y)) * (z + (x - y)) * (x + (y - z)));
        const T k0 = d_atan2_ypos(T(2) * akite, (r2 - r1) * (r2 + r1) + b * b);
        const T k1 = d_atan2_ypos(T(2) * akite, (r1 - r2) * (r1 + r2) + b * b);
        area = r1 * r1 * k1 + r2 * r2 * k0 - akite;
        kap0 = k0;
        akite_out = akite;
        return;
    }
    if (b <= r1 - r2) { area = pi_of<T>() * (r2 * r2); kap0 = pi_of<T>(); return; }
    if (b <= r2 - r1) { area = pi_of<T>() * (r1 * r1); kap0 = T(0); return; }
    area = T(NAN); kap0 = T(NAN);
}

template <typename T>
__device__ __forceinline__ void ccia_kite(T r1, T r2, T b, T& area, T& kap0) {
    T akite;
    ccia_kite3<T>(r1, r2, b, area, kap0, akite);
}

// Mean limb darkening for grazing parameter g; common.py:225-233.  The interval
// index is clamped to ng-2 (the reference reads lda[ng] for g in (1-1e-7, 1]).
__device__ __forceinline__ double interp_ld(double g, double dg, const double* __restrict__ row, int ng) {
    if (g < 0.0) return NAN;
    if (g > 1.0) return 0.0;
    int i = (int)floor(g / dg);
    i = i > ng - 2 ? ng - 2 : (i < 0 ? 0 : i);
    const double a = (g - i * dg) / dg;
    return (1.0 - a) * row[i] + a * row[i + 1];
}

// FP32 form: g / dg by a multiply with 1/dg.  The lerp is continuous across
// knots, so a knot assigned to the neighbouring interval changes the value by
// rounding only.
__device__ __forceinline__ float interp_ld_f(float g, float inv_dg, const float* __restrict__ row, int ng) {
    if (g > 1.0f) return 0.0f;
    const float x = g * inv_dg;
    int i = (int)x;  // g >= 0
    i = i > ng - 2 ? ng - 2 : (i < 0 ? 0 : i);
    const float a = x - (float)i;
    const float l0 = row[i], l1 = row[i + 1];
    return fmaf(a, l1 - l0, l0);
}

// ---- Orbit: Kepler solve (orbits_py.py) and Taylor coefficients (taylor_z.py) ---------

// numpy.mod: the result has the sign of the divisor.
__device__ __forceinline__ double py_mod(double x, double y) {
    double r = fmod(x, y);
    if (r != 0.0) {
        if ((r < 0.0) != (y < 0.0)) r += y;
    } else {
        r = copysign(0.0, y);
    }
    return r;
}

// orbits_py.py:82-86
__device__ __forceinline__ double mean_anomaly_offset(double e, double w) {
    double m = atan2(sqrt(1.0 - e * e) * sin(kHalfPi - w), e + cos(kHalfPi - w));
    return m - e * sin(m);
}

// True anomaly at time t for t0 = 0: mean anomaly (orbits_py.py:115-119), Newton
// eccentric anomaly with tol 1e-8 and at most 1000 steps (:144-154), true anomaly
// (:191-196).  The Newton loop is bounded, so every lane exits.
__device__ __forceinline__ double ta_newton(double t, double p, double e, double w) {
    const double off = mean_anomaly_offset(e, w);
    const double ma = py_mod(kTwoPi * (t - (0.0 - off * p / kTwoPi)) / p, kTwoPi);
    double ea = ma, err = 0.05;
    int k = 0;
    while (fabs(err) > 1e-8 && k < 1000) {
        err = ea - e * sin(ea) - ma;
        ea = ea - err / (1.0 - e * cos(ea));
        ++k;
    }
    const double ce = cos(ea), se = sin(ea);
    const double sta = sqrt(1.0 - e * e) * se / (1.0 - e * ce);
    const double cta = (ce - e) / (1.0 - e * ce);
    return atan2(sta, cta);
}

// Sky-plane position at time t (t0 = 0) for the finite-difference stencil; taylor_z.py:54-76.
__device__ __forceinline__ void kepler_xy(double t, double p, double a, double ci, double e, double w,
                                          double& x, double& y) {
    const double f = ta_newton(t, p, e, w);
    const double r = a * (1.0 - e * e) / (1.0 + e * cos(f));
    x = -r * cos(w + f);
    y = -r * sin(w + f) * ci;
}

// Taylor coefficients from the seven stencil positions x[0..6], y[0..6] at
// t = (j-3)*dt, dt = 0.02 d; taylor_z.py:78-102.  c[0..4] is x, c[5..9] is y, each
// divided by n!; the x constant term is 0 (taylor_z.py:224).
__device__ __forceinline__ void taylor_from_stencil(const double* x, const double* y, double* c) {
    const double dt = 2e-2;
    const double a1 = 1.0 / 60, b1 = 9.0 / 60, c1 = 45.0 / 60;
    const double a2 = 1.0 / 90, b2 = 3.0 / 20, c2 = 3.0 / 2, d2 = 49.0 / 18;
    const double a3 = 1.0 / 8, b3 = 1.0, c3 = 13.0 / 8;
    const double a4 = 1.0 / 6, b4 = 2.0, c4 = 13.0 / 2, d4 = 28.0 / 3;
    const double dt2 = dt * dt, dt3 = dt2 * dt, dt4 = dt3 * dt;
    const double vx = (a1 * (x[6] - x[0]) + b1 * (x[1] - x[5]) + c1 * (x[4] - x[2])) / dt;
    const double vy = (a1 * (y[6] - y[0]) + b1 * (y[1] - y[5]) + c1 * (y[4] - y[2])) / dt;
    const double ax = (a2 * (x[0] + x[6]) - b2 * (x[1] + x[5]) + c2 * (x[2] + x[4]) - d2 * x[3]) / dt2;
    const double ay = (a2 * (y[0] + y[6]) - b2 * (y[1] + y[5]) + c2 * (y[2] + y[4]) - d2 * y[3]) / dt2;
    const double jx = (a3 * (x[0] - x[6]) + b3 * (x[5] - x[1]) + c3 * (x[2] - x[4])) / dt3;
    const double jy = (a3 * (y[0] - y[6]) + b3 * (y[5] - y[1]) + c3 * (y[2] - y[4])) / dt3;
    const double sx = (-a4 * (x[0] + x[6]) + b4 * (x[1] + x[5]) - c4 * (x[2] + x[4]) + d4 * x[3]) / dt4;
    const double sy = (-a4 * (y[0] + y[6]) + b4 * (y[1] + y[5]) - c4 * (y[2] + y[4]) + d4 * y[3]) / dt4;
    c[0] = 0.0; c[1] = vx; c[2] = 0.5 * ax; c[3] = jx / 6.0; c[4] = sx / 24.0;
    c[5] = y[3]; c[6] = vy; c[7] = 0.5 * ay; c[8] = jy / 6.0; c[9] = sy / 24.0;
}

// Separation z(t) from the Taylor coefficients by Horner's rule (sep_c).
template <typename T>
__device__ __forceinline__ T taylor_z(T t, const T* c) {
    const T px = (((c[4] * t + c[3]) * t + c[2]) * t + c[1]) * t + c[0];
    const T py = (((c[9] * t + c[8]) * t + c[7]) * t + c[6]) * t + c[5];
    return d_sqrt(px * px + py * py);
}

// z_taylor_s(t, t0=0, p=1): Taylor z with a unit-period phase wrap; taylor_z.py:198-226.
__device__ __forceinline__ double taylor_z_wrapped(double t, const double* c) {
    const double epoch = floor((t - 0.0 + 0.5 * 1.0) / 1.0);
    const double tt = t - (0.0 + epoch * 1.0);
    return taylor_z(tt, c);
}

// Contact time by bisection, tol 1e-6 d and at most 100 steps; taylor_z.py:298-328.
// point 1 or 4 (first/fourth contact, z = 1+k).  The first bracket is [0, s 2/v]
// with s = -1 before mid-transit and +1 after; the reference takes v = vx, the
// sky-plane x velocity (positive at a transit).  The eclipse path passes |vx|:
// behind the star the planet moves in -x, and v = vx would put first contact after
// mid-eclipse.
__device__ __forceinline__ double contact_bisect(double k, int point, const double* c, double v) {
    const double s = (point == 1 || point == 2 || point == 12) ? -1.0 : 1.0;
    const double zt = (point == 1 || point == 4) ? 1.0 + k : ((point == 2 || point == 3) ? 1.0 - k : 1.0);
    double t0 = 0.0;
    double t2 = s * 2.0 / v;
    double t1 = 0.5 * t2;
    double z0 = taylor_z_wrapped(t0, c) - zt;
    double z1 = taylor_z_wrapped(t1, c) - zt;
    int i = 0;
    while (fabs(t2 - t0) > 1e-6 && i < 100) {
        if (z0 * z1 < 0.0) {
            t2 = t1;
            t1 = 0.5 * (t0 + t1);
            z1 = taylor_z_wrapped(t1, c) - zt;
        } else {
            t0 = t1;
            t1 = 0.5 * (t1 + t2);
            z0 = z1;
            z1 = taylor_z_wrapped(t1, c) - zt;
        }
        ++i;
    }
    return t1;
}

__device__ __forceinline__ double find_contact_point(double k, int point, const double* c) {
    return contact_bisect(k, point, c, c[1]);
}

// ---- Epoch reduction (model_full.py:88-89) -------------------------------------------

// Exact form: floor((t - t0 + p/2) / p) with an IEEE division, then
// tc = t - (t0 + epoch*p), with no contraction so each step rounds as in numpy.
__device__ __forceinline__ double phase_exact(double t, double t0, double p) {
#pragma clang fp contract(off)
    const double epoch = floor((t - t0 + 0.5 * p) / p);
    return t - (t0 + epoch * p);
}

// Division-free form for the mixed path: epoch from a multiply by 1/p, then one
// correction step so that r = (t - t0 + p/2) - epoch*p lies in [0, p).  It can
// differ from phase_exact only where the quotient is within rounding of an
// integer, i.e. at tc = +-p/2, far outside any transit window.
__device__ __forceinline__ double phase_fast(double t, double t0, double p, double inv_p) {
#pragma clang fp contract(off)
    const double u = t - t0 + 0.5 * p;
    double epoch = floor(u * inv_p);
    const double r = u - epoch * p;
    epoch += (r < 0.0) ? -1.0 : ((r >= p) ? 1.0 : 0.0);
    return t - (t0 + epoch * p);
}

// ---- Limb-darkening laws; models/numba/ldmodels.py:22-139 -----------------------------

__device__ __forceinline__ double ld_law(int law, double mu, const double* c, int nc) {
    switch (law) {
        case RR_LD_UNIFORM: return 1.0;
        case RR_LD_LINEAR: return 1.0 - c[0] * (1.0 - mu);
        case RR_LD_QUADRATIC: return 1.0 - c[0] * (1.0 - mu) - c[1] * ((1.0 - mu) * (1.0 - mu));
        case RR_LD_QUADRATIC_TRI: {
            const double a = sqrt(c[0]), b = 2.0 * c[1];
            const double u = a * b, v = a * (1.0 - b);
            return 1.0 - u * (1.0 - mu) - v * ((1.0 - mu) * (1.0 - mu));
        }
        case RR_LD_NONLINEAR: {
            // mu^1.5 as mu * sqrt(mu): within 2 ulp of pow
            const double sm = sqrt(mu);
            return 1.0 - c[0] * (1.0 - sm) - c[1] * (1.0 - mu) - c[2] * (1.0 - mu * sm) - c[3] * (1.0 - mu * mu);
        }
        case RR_LD_GENERAL: {
            double s = 0.0, mp = mu;  // mp = mu^(i+1) by repeated products
            for (int i = 0; i < nc; ++i, mp *= mu) s += c[i] * (1.0 - mp);
            return s;
        }
        case RR_LD_SQUARE_ROOT: return 1.0 - c[0] * (1.0 - mu) - c[1] * (1.0 - sqrt(mu));
        case RR_LD_LOGARITHMIC: return 1.0 - c[0] * (1.0 - mu) - c[1] * mu * log(mu);
        case RR_LD_EXPONENTIAL: return 1.0 - c[0] * (1.0 - mu) - c[1] / (1.0 - exp(mu));
        case RR_LD_POWER_2: return 1.0 - c[0] * (1.0 - pow(mu, c[1]));
        case RR_LD_POWER_2_PM: {
            const double cc = 1.0 - c[0] + c[1];
            const double a = log2(cc / c[1]);
            return 1.0 - cc * (1.0 - pow(mu, a));
        }
        default: return NAN;
    }
}

// ---- Laws that are linear in their coefficients -------------------------------------
//
// I(mu) = sum_i coef_i(c) B_i(mu) for every named law but power-2, power-2-pm and
// tabulated profiles: the constant and the coefficient terms of ldmodels.py:22-113.
// Then every linear functional of the profile (its limb-darkening means, its disk
// integral) of any passband is the same combination of the functional of the
// basis, which a pv computes once for all its passbands.
constexpr int kMaxBasis = 6;

// Number of basis functions of the law, 0 when it is not linear in its coefficients.
__host__ __device__ inline int ld_basis_count(int law, int nc) {
    switch (law) {
        case RR_LD_UNIFORM: return 1;
        case RR_LD_LINEAR: return 2;
        case RR_LD_QUADRATIC:
        case RR_LD_QUADRATIC_TRI:
        case RR_LD_SQUARE_ROOT:
        case RR_LD_LOGARITHMIC:
        case RR_LD_EXPONENTIAL: return 3;
        case RR_LD_NONLINEAR: return 5;
        case RR_LD_GENERAL: return nc >= 1 && nc <= kMaxBasis ? nc : 0;
        default: return 0;
    }
}

// B_i(mu), in the form of the matching term of ld_law.
__device__ __forceinline__ double ld_basis(int law, int i, double mu) {
    if (law == RR_LD_GENERAL) {
        double mp = mu;  // mu^(i+1) by repeated products
        for (int j = 0; j < i; ++j) mp *= mu;
        return 1.0 - mp;
    }
    if (i == 0) return 1.0;
    switch (law) {
        case RR_LD_LINEAR: return 1.0 - mu;
        case RR_LD_QUADRATIC:
        case RR_LD_QUADRATIC_TRI: return i == 1 ? 1.0 - mu : (1.0 - mu) * (1.0 - mu);
        case RR_LD_NONLINEAR: {
            const double sm = sqrt(mu);
            return i == 1 ? 1.0 - sm : (i == 2 ? 1.0 - mu : (i == 3 ? 1.0 - mu * sm : 1.0 - mu * mu));
        }
        case RR_LD_SQUARE_ROOT: return i == 1 ? 1.0 - mu : 1.0 - sqrt(mu);
        case RR_LD_LOGARITHMIC: return i == 1 ? 1.0 - mu : mu * log(mu);
        case RR_LD_EXPONENTIAL: return i == 1 ? 1.0 - mu : 1.0 / (1.0 - exp(mu));
        default: return NAN;
    }
}

// coef_i of the coefficients c[0..nc).
__device__ __forceinline__ double ld_basis_coef(int law, int i, const double* c) {
    switch (law) {
        case RR_LD_GENERAL: return c[i];
        case RR_LD_QUADRATIC_TRI: {
            if (i == 0) return 1.0;
            const double a = sqrt(c[0]), b = 2.0 * c[1];
            return i == 1 ? -(a * b) : -(a * (1.0 - b));
        }
        default: return i == 0 ? 1.0 : -c[i - 1];
    }
}

// Analytic disk integral where the reference has one (ldmodels.py:27-123); false
// for the laws that use the 200-node trapezoid (rrmodel.py:220-227, Q8).
__device__ __forceinline__ bool ld_integral(int law, const double* c, double& out) {
    switch (law) {
        case RR_LD_UNIFORM: out = kPi; return true;
        case RR_LD_LINEAR: out = 2.0 * kPi * 1.0 / 6.0 * (3.0 - 2.0 * c[0]); return true;  // Q7, as the reference
        case RR_LD_QUADRATIC: out = 2.0 * kPi * 1.0 / 12.0 * (-2.0 * c[0] - c[1] + 6.0); return true;
        case RR_LD_QUADRATIC_TRI: {
            const double a = sqrt(c[0]), b = 2.0 * c[1];
            const double u = a * b, v = a * (1.0 - b);
            out = 2.0 * kPi * 1.0 / 12.0 * (-2.0 * u - v + 6.0);
            return true;
        }
        case RR_LD_POWER_2: out = 2.0 * kPi * (-c[0] * c[1] + c[1] + 2.0) / (2.0 * c[1] + 4.0); return true;
        default: return false;
    }
}

// One trapezoid segment m (1..199) of 2*pi*trapz(z*I(mu), z) on mu = linspace(1, 0, 200).
__device__ __forceinline__ double ld_trapz_segment(int law, const double* c, int nc, int m) {
    const double mu1 = linspace_at(1.0, 0.0, 200, m), mu0 = linspace_at(1.0, 0.0, 200, m - 1);
    const double z1 = sqrt(1.0 - mu1 * mu1), z0 = sqrt(1.0 - mu0 * mu0);
    return (z1 - z0) * (z1 * ld_law(law, mu1, c, nc) + z0 * ld_law(law, mu0, c, nc)) / 2.0;
}

// ---- Orbit for 8 parameter vectors per 64-lane block: lane j (0..6) of a group of 8
// solves Kepler at t = (j-3)*dt, lane 0 forms the stencils, lanes 0 and 1 find
// the first and fourth contacts.  Writes coefficients[10] and the window.
constexpr int kOrbitGroup = 8;

constexpr int kOrbitLds = kOrbitGroup * (8 + 8 + 10 + 2);  // doubles of LDS per block

// `sm` is block-shared storage of kOrbitLds doubles.
__device__ __forceinline__ void orbit_group(int pv, bool valid, double p, double a, double inc, double e, double w,
                                            double kw, double* out_c, double* out_win, double* sm) {
    const int sub = threadIdx.x / kOrbitGroup, j = threadIdx.x % kOrbitGroup;
    double* s_x = sm + sub * 8;
    double* s_y = sm + kOrbitGroup * 8 + sub * 8;
    double* s_c = sm + kOrbitGroup * 16 + sub * 10;
    double* s_w = sm + kOrbitGroup * 26 + sub * 2;
    if (valid && j < 7) {
        double x, y;
        kepler_xy((j - 3) * 2e-2, p, a, cos(inc), e, w, x, y);
        s_x[j] = x;
        s_y[j] = y;
    }
    __syncthreads();
    if (valid && j == 0) taylor_from_stencil(s_x, s_y, s_c);
    __syncthreads();
    if (valid && j < 2) s_w[j] = find_contact_point(kw, j == 0 ? 1 : 4, s_c);
    __syncthreads();
    if (valid) {
        out_c[j] = s_c[j];
        if (j < 2) out_c[8 + j] = s_c[8 + j];
        if (j < 2) out_win[j] = s_w[j];
    }
}

// Workgroups are dealt to the 8 XCDs of the chip round-robin (XCD = blockIdx % 8).
// Block b is given work item xcd_remap(b, n) instead, so that each XCD works through
// one contiguous eighth of the n items: a pattern in the item order (a transit in the
// first tile of every pv, say) is spread over all XCDs rather than landing on the few
// that the round-robin assigns it to, and an XCD's L2 keeps the pvs it works on.
// A bijection of [0, n).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_remap(int b, int n) {
    const int x = b % kXcds, j = b / kXcds;
    const int per = n / kXcds, r = n % kXcds;
    return x * per + min(x, r) + j;
}

// Sum over a wavefront (64 lanes) by butterfly shuffles; every lane gets the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// 2*pi*trapz(z*I(mu), z) on mu = linspace(1, 0, 200) (rrmodel.py:151-152,220-227) by a
// whole wavefront, for the profile I = f(mu).  Each node's value z*I(mu) is computed
// once into `s_y` (kTrapzNodes doubles of LDS), then segment m pairs node m with node
// m-1, each lane summing its segments in node order before the wave sum.  Staging
// the values in LDS keeps the law's evaluation (an FP64 pow for the power laws) out
// of the summation loop, whose registers would otherwise spill.  Every lane of the
// block must call it (it holds two barriers).
constexpr int kTrapzNodes = 200;
template <class F>
__device__ __forceinline__ double trapz_disk_wave(F f, double* s_y) {
    const int lane = threadIdx.x & 63;
    for (int m = lane; m < kTrapzNodes; m += 64) {
        const double mu = linspace_at(1.0, 0.0, kTrapzNodes, m);
        s_y[m] = sqrt(1.0 - mu * mu) * f(mu);
    }
    __syncthreads();
    double part = 0.0;
    for (int m = lane; m < kTrapzNodes; m += 64) {
        if (m >= 1) {
            const double mu = linspace_at(1.0, 0.0, kTrapzNodes, m);
            const double ml = linspace_at(1.0, 0.0, kTrapzNodes, m - 1);
            const double z = sqrt(1.0 - mu * mu), zl = sqrt(1.0 - ml * ml);
            part += (z - zl) * (s_y[m] + s_y[m - 1]) / 2.0;
        }
    }
    __syncthreads();
    return 2.0 * kPi * wave_sum(part);
}

// The disk integral of a law's profile, and of its basis function i.
__device__ __forceinline__ double ld_trapz_wave(int law, const double* c, int nc, double* s_y) {
    return trapz_disk_wave([&](double mu) { return ld_law(law, mu, c, nc); }, s_y);
}
__device__ __forceinline__ double ld_basis_trapz_wave(int law, int i, double* s_y) {
    return trapz_disk_wave([&](double mu) { return ld_basis(law, i, mu); }, s_y);
}

}  // namespace rr
