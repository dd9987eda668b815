// Gaussian log-likelihood of the RoadRunner model and its gradient (hand-written
// adjoint) for one light curve, fused so the flux is never written.
//
//   lnL(pv) = sum_t [ -ln sigma - 1/2 ln 2pi - 1/2 ((o_t - m_t)/sigma)^2 ]
// (the form of lnlike_normal, pytransit/lpf/loglikelihood/wnloglikelihood.py:22-35)
// and d lnL / d(k, t0, p, a, i, c0, c1) with c the two limb-darkening coefficients
// of the quadratic or power-2 law.  The model m_t is rr_full's (model_full.py:
// 75-99); its derivatives follow the chain
//   m <- (I(g), A(k, z)) <- z(Taylor series of the orbit) <- coefficients(p, a, i)
//   and phase tc(t0, p),  I <- table lerp in k and g, ldp(c), istar(c).
// Coefficient derivatives are exact derivatives of the finite-difference stencil
// (taylor_z.py:23-102): d/da = c/a, d/di scales the y row by -tan(i), and d/dp
// propagates dM/dp = -2 pi t / p^2 through Kepler's equation at the converged
// solution, df/dM = (1 + e cos f)^2 / (1 - e^2)^(3/2).  Lens area: dA/dz = -2 a_kite / z (minus the
// chord), dA/dk = 2 k kappa0.  Points outside the window have m = 1 and zero
// gradient.  FP64 epoch reduction and accumulation; per-sample arithmetic in
// Real (float for RR_MIXED, double for RR_FP64).
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "rr_device.hpp"
#include "rr_internal.hpp"

namespace rr {

struct LnlArgs {
    const double* times; const double* obs; const double* sigma;
    const int32_t* nsamples; const double* exptimes; const int32_t* epids;
    const double* k; const double* t0; const double* p; const double* a; const double* inc;
    const double* e; const double* w;
    int law, nldc; const double* ldc;
    const double* W; const double* mu;
    double* lnl; double* grad;
    int npv, npt, kcols, t0cols, nk, ng, nz;
    double kmin, kmax, dk, dg;
};

constexpr int kLnlThreads = 256;
constexpr int kLnlMaxNg = 512;
constexpr int kLnlMaxNz = 1024;

// Position at stencil time t and its derivative with respect to the period:
// dM/dp = -2 pi t / p^2 (t0 = 0) and df/dM = (1 + e cos f)^2 / (1 - e^2)^(3/2).
__device__ void kepler_xy_dp(double t, double p, double a, double ci, double e, double w, double& x, double& y,
                             double& dxdp, double& dydp) {
    const double f = ta_newton(t, p, e, w);
    const double ecf = 1.0 + e * cos(f);
    const double om = 1.0 - e * e;
    const double df_dp = (ecf * ecf / (om * sqrt(om))) * (-kTwoPi * t / (p * p));
    const double r = a * om / ecf;
    const double dr_df = r * e * sin(f) / ecf;
    const double cw = cos(w + f), sw = sin(w + f);
    x = -r * cw;
    y = -r * sw * ci;
    dxdp = (-dr_df * cw + r * sw) * df_dp;
    dydp = (-dr_df * sw - r * cw) * ci * df_dp;
}

__device__ __forceinline__ double dld_dc(int law, double mu, const double* c, int j) {
    if (law == RR_LD_QUADRATIC) return j == 0 ? -(1.0 - mu) : -((1.0 - mu) * (1.0 - mu));
    if (law == RR_LD_POWER_2) {
        const double ma = pow(mu, c[1]);
        return j == 0 ? -(1.0 - ma) : c[0] * ma * log(mu);
    }
    return 0.0;
}

__device__ __forceinline__ double distar_dc(int law, const double* c, int j) {
    if (law == RR_LD_QUADRATIC) return j == 0 ? -kPi / 3.0 : -kPi / 6.0;
    if (law == RR_LD_POWER_2) {
        const double d = 2.0 * c[1] + 4.0;
        return j == 0 ? 2.0 * kPi * (-c[1]) / d : -8.0 * kPi * c[0] / (d * d);
    }
    return 0.0;
}

template <typename Real>
__device__ __forceinline__ Real horner5(const Real* c, Real t) {
    return (((c[4] * t + c[3]) * t + c[2]) * t + c[1]) * t + c[0];
}
template <typename Real>
__device__ __forceinline__ Real dhorner5(const Real* c, Real t) {
    return ((Real(4) * c[4] * t + Real(3) * c[3]) * t + Real(2) * c[2]) * t + c[1];
}

// Lens area of the unit disk and a disk of radius k < 1 at separation z, in the
// lens case |z - 1| < k, from u = z - 1 given to full FP32 precision.  With the
// kite area written as akite = z sin kappa1 (Heron), the kite form
// A = kappa1 + k^2 kappa0 - akite becomes
//   A = k^2 kappa0 + S(kappa1) - u akite / z,   S(x) = x - sin x,
// so a small planet's nearly equal kappa1 and akite are not subtracted (in FP32
// that cancellation leaves ~1e-7 of the area's scale, 0.1 k); S is its Taylor
// series up to x^11 for kappa1 <= 0.6 (truncation < 1e-13), the kite form above
// it.  The kite's square root is ((1 + z)^2 - k^2)(k^2 - u^2), the sorted-sides
// product of common.py:5-33,52-73 regrouped, and z^2 - 1 = u (2 + u).  Returns A,
// kappa0 and akite.
__device__ __forceinline__ void lens_seg_f32(float k, float z, float u, float& area, float& kap0, float& akite) {
    const float zp1 = 1.0f + z;
    const float sq = __builtin_amdgcn_sqrtf((zp1 * zp1 - k * k) * ((k - u) * (k + u)));  // 2 akite
    const float zm = u * (2.0f + u);                                                     // z^2 - 1
    kap0 = atan2_ypos_f(sq, zm + k * k);
    const float kap1 = atan2_ypos_f(sq, 2.0f + zm - k * k);
    akite = 0.5f * sq;
    if (kap1 <= 0.6f) {
        const float x2 = kap1 * kap1;
        float sr = 1.0f / 39916800.0f;
        sr = fmaf(sr, x2, -1.0f / 362880.0f);
        sr = fmaf(sr, x2, 1.0f / 5040.0f);
        sr = fmaf(sr, x2, -1.0f / 120.0f);
        sr = fmaf(sr, x2, 1.0f / 6.0f);
        area = k * k * kap0 + kap1 * x2 * sr - u * akite / z;
    } else {
        area = kap1 + k * k * kap0 - akite;
    }
}

template <typename Real>
__global__ void __launch_bounds__(kLnlThreads) rr_lnlike_kernel(LnlArgs s) {
    __shared__ double s_x[7], s_y[7], s_dx[7], s_dy[7], s_c[10], s_dc[10], s_win[2];
    __shared__ double s_l[kLnlMaxNz], s_d0[kLnlMaxNz], s_d1[kLnlMaxNz];
    __shared__ Real s_L[kLnlMaxNg], s_Lk[kLnlMaxNg], s_Lc0[kLnlMaxNg], s_Lc1[kLnlMaxNg], s_dL[kLnlMaxNg];
    __shared__ double s_L64[kLnlMaxNg];
    __shared__ double s_red[kLnlThreads / 64][8];
    __shared__ double s_ist, s_dist0, s_dist1;
    __shared__ int s_ok;

    constexpr bool kMixed = std::is_same<Real, float>::value;
    const int pv = blockIdx.x;
    const int tid = threadIdx.x;
    const double a = s.a[pv], e = s.e[pv], p = s.p[pv], inc = s.inc[pv], w = s.w[pv];
    const double kk = s.k[(size_t)pv * s.kcols];
    const double* c = s.ldc + (size_t)pv * s.nldc;
    const bool in_table = s.kmin <= kk && kk <= s.kmax;
    // ---- prologue: orbit, window, limb-darkening rows ----------------------------
    if (tid < 7) {
        double x, y, dx, dy;
        kepler_xy_dp((tid - 3) * 2e-2, p, a, cos(inc), e, w, x, y, dx, dy);
        s_x[tid] = x; s_y[tid] = y; s_dx[tid] = dx; s_dy[tid] = dy;
    }
    for (int iz = tid; iz < s.nz; iz += kLnlThreads) {
        const double mu = s.mu[iz];
        s_l[iz] = ld_law(s.law, mu, c, s.nldc);
        s_d0[iz] = dld_dc(s.law, mu, c, 0);
        s_d1[iz] = dld_dc(s.law, mu, c, 1);
    }
    __syncthreads();
    if (tid == 0) {
        taylor_from_stencil(s_x, s_y, s_c);
        taylor_from_stencil(s_dx, s_dy, s_dc);
        double ist;
        if (!ld_integral(s.law, c, ist)) {
            double sum = 0.0;
            for (int m = 1; m < 200; ++m) sum += ld_trapz_segment(s.law, c, s.nldc, m);
            ist = 2.0 * kPi * sum;
        }
        s_ist = ist;
        s_dist0 = distar_dc(s.law, c, 0);
        s_dist1 = distar_dc(s.law, c, 1);
        s_ok = !(isnan(a) || a <= 1.0 || e < 0.0 || isnan(s_l[0])) && in_table;
    }
    __syncthreads();
    if (tid < 2) s_win[tid] = find_contact_point(kk, tid == 0 ? 1 : 4, s_c);
    int ik = 0;
    double ak = 0.0;
    if (in_table) {
        ik = (int)floor((kk - s.kmin) / s.dk);
        ik = ik > s.nk - 2 ? s.nk - 2 : ik;
        ak = (kk - s.kmin - ik * s.dk) / s.dk;
    }
    for (int ig = tid; ig < s.ng; ig += kLnlThreads) {
        double a0 = 0, a1 = 0, b0 = 0, b1 = 0, c0 = 0, c1 = 0;
        const double* w0 = s.W + (size_t)ik * s.nz * s.ng + ig;
        const double* w1 = w0 + (size_t)s.nz * s.ng;
        for (int iz = 0; iz < s.nz; ++iz) {
            const double u0 = w0[(size_t)iz * s.ng], u1 = w1[(size_t)iz * s.ng];
            a0 += u0 * s_l[iz]; a1 += u1 * s_l[iz];
            b0 += u0 * s_d0[iz]; b1 += u1 * s_d0[iz];
            c0 += u0 * s_d1[iz]; c1 += u1 * s_d1[iz];
        }
        const double L = (1.0 - ak) * a0 + ak * a1;
        s_L64[ig] = L;
        s_L[ig] = (Real)L;
        s_Lk[ig] = (Real)((a1 - a0) / s.dk);
        s_Lc0[ig] = (Real)((1.0 - ak) * b0 + ak * b1);
        s_Lc1[ig] = (Real)((1.0 - ak) * c0 + ak * c1);
    }
    __syncthreads();
    // the g slope of each interval from the FP64 means: the difference of two
    // rounded FP32 values would lose ~1e-5 of the slope, which the t0, p, a and i
    // gradients carry through the chain rule
    for (int ig = tid; ig < s.ng - 1; ig += kLnlThreads) s_dL[ig] = (Real)((s_L64[ig + 1] - s_L64[ig]) / s.dg);
    __syncthreads();
    if (!s_ok) {
        if (tid == 0) {
            s.lnl[pv] = NAN;
            if (s.grad)
                for (int j = 0; j < 7; ++j) s.grad[(size_t)pv * 7 + j] = NAN;
        }
        return;
    }

    // ---- main loop over times ----------------------------------------------------
    const int ns = s.nsamples[0];
    const double ext = s.exptimes[0];
    // the light curve's epoch column, clamped as the evaluate path does
    const int ep = min(max(s.epids[0], 0), s.t0cols - 1);
    const double t0v = s.t0[(size_t)pv * s.t0cols + ep];
    const double lo = s_win[0] - (0.003 + ext), hi = s_win[1] + (0.003 + ext);
    const double inv_p = 1.0 / p;
    Real cx[5], cy[5], dcx[5], dcy[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        cx[i] = (Real)s_c[i]; cy[i] = (Real)s_c[5 + i];
        dcx[i] = (Real)s_dc[i]; dcy[i] = (Real)s_dc[5 + i];
    }
    const Real kr = (Real)kk, zlim = Real(1) + kr, inv1pk = Real(1) / zlim;
    const Real dg = (Real)s.dg, inv_dg = (Real)(1.0 / s.dg);
    const double g_scale = 1.0 / ((1.0 + kk) * s.dg);   // g / dg = z g_scale
    const Real tan_i = (Real)tan(inc), inv_a = (Real)(1.0 / a);
    const Real pik2 = (Real)(kPi * kk * kk), dpik2 = (Real)(2.0 * kPi * kk);
    const double scale = 1.0 / (ns * s_ist);
    const Real rscale = (Real)scale;
    // per-sample and per-point divisors as reciprocals (off the samples' chains)
    const double inv_ns = 1.0 / ns, scale_ist = scale / s_ist;
    double chi = 0.0, g[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int it = tid; it < s.npt; it += kLnlThreads) {
        const double t = s.times[it];
        const double om1 = s.obs[it] - 1.0;
        double epoch;
        double tc;
        {
#pragma clang fp contract(off)
            const double u = t - t0v + 0.5 * p;
            double ep = floor(u * inv_p);
            const double rr = u - ep * p;
            ep += (rr < 0.0) ? -1.0 : ((rr >= p) ? 1.0 : 0.0);
            epoch = ep;
            tc = t - (t0v + ep * p);
        }
        if (!(lo <= tc && tc <= hi)) {
            chi += om1 * om1;
            continue;
        }
        Real D = 0, dDk = 0, dDt0 = 0, dDp = 0, dDa = 0, dDi = 0, dDc0 = 0, dDc1 = 0;
        const Real ep = (Real)epoch;
        for (int is = 1; is <= ns; ++is) {
            const double taud = tc + ext * ((is - 0.5) * inv_ns - 0.5);
            const Real tau = (Real)taud;
            Real x, y, z, u;
            if constexpr (kMixed) {
                // the constant and linear terms of the position in FP64: the phase
                // resolves ~1e-9 d that an FP32 time would drop, ~1e-7 of the
                // position at sky speeds of ~40 stellar radii per day.  z - 1 is
                // formed from the FP64 squared distance, so the lens area below
                // does not see the rounding of z.
                const Real t2 = tau * tau;
                const double xd = s_c[0] + s_c[1] * taud + (double)(((cx[4] * tau + cx[3]) * tau + cx[2]) * t2);
                const double yd = s_c[5] + s_c[6] * taud + (double)(((cy[4] * tau + cy[3]) * tau + cy[2]) * t2);
                const double s2 = xd * xd + yd * yd;
                x = (Real)xd;
                y = (Real)yd;
                z = d_sqrt((Real)s2);
                u = (Real)(s2 - 1.0) / (Real(1) + z);
            } else {
                x = horner5(cx, tau);
                y = horner5(cy, tau);
                z = sqrt(x * x + y * y);
                u = z - Real(1);
            }
            if (!(z < zlim)) continue;
            const Real vx = dhorner5(cx, tau), vy = dhorner5(cy, tau);
            const Real px = horner5(dcx, tau), py = horner5(dcy, tau);
            // limb darkening and its slopes at g = z/(1+k)
            // the grazing parameter in units of the g step: in the mixed form from
            // 1 + u in FP64, so the lerp position keeps the precision of u
            Real gg, gx;
            if constexpr (kMixed) {
                const double gxd = (1.0 + (double)u) * g_scale;
                gx = (Real)gxd;
                gg = (Real)(gxd * s.dg);
            } else {
                gg = z * inv1pk;
                gx = gg * inv_dg;
            }
            Real Ii = 0, dIdg = 0, Ik = 0, Ic0 = 0, Ic1 = 0;
            if (gg <= Real(1)) {
                int i = (int)floor(gx);
                i = i > s.ng - 2 ? s.ng - 2 : (i < 0 ? 0 : i);
                const Real al = gx - (Real)i;
                Ii = s_L[i] + al * (s_L[i + 1] - s_L[i]);
                dIdg = s_dL[i];
                Ik = s_Lk[i] + al * (s_Lk[i + 1] - s_Lk[i]);
                Ic0 = s_Lc0[i] + al * (s_Lc0[i + 1] - s_Lc0[i]);
                Ic1 = s_Lc1[i] + al * (s_Lc1[i + 1] - s_Lc1[i]);
            }
            const Real inv_z = Real(1) / z;
            // area and its slopes
            Real A, dAdz, dAdk;
            if (z <= Real(1) - kr) {
                A = pik2; dAdz = 0; dAdk = dpik2;
            } else {
                Real kap, akite;
                if constexpr (kMixed) {
                    if (kk < 1.0) lens_seg_f32(kr, z, u, A, kap, akite);
                    else ccia_kite3<Real>(Real(1), kr, z, A, kap, akite);
                } else {
                    ccia_kite3<Real>(Real(1), kr, z, A, kap, akite);
                }
                dAdz = Real(-2) * akite * inv_z;   // minus the chord of the lens
                dAdk = Real(2) * kr * kap;
            }
            // z derivatives
            const Real dz_t0 = -(x * vx + y * vy) * inv_z;
            const Real dz_p = (x * (px - ep * vx) + y * (py - ep * vy)) * inv_z;
            const Real dz_a = z * inv_a;
            const Real dz_i = -(y * y) * tan_i * inv_z;
            const Real G = dIdg * inv1pk * A + Ii * dAdz;
            D += Ii * A;
            dDk += (Ik - dIdg * z * inv1pk * inv1pk) * A + Ii * dAdk;
            dDt0 += G * dz_t0;
            dDp += G * dz_p;
            dDa += G * dz_a;
            dDi += G * dz_i;
            dDc0 += Ic0 * A;
            dDc1 += Ic1 * A;
        }
        const double r = om1 + (double)(D * rscale);     // o - m, with m = 1 - D/(ns istar)
        chi += r * r;
        const double rs = -r * scale;                     // r * dm/dD
        g[0] += rs * (double)dDk;
        g[1] += rs * (double)dDt0;
        g[2] += rs * (double)dDp;
        g[3] += rs * (double)dDa;
        g[4] += rs * (double)dDi;
        g[5] += rs * (double)dDc0 + r * (double)D * scale_ist * s_dist0;
        g[6] += rs * (double)dDc1 + r * (double)D * scale_ist * s_dist1;
    }
    // ---- block reduction ----------------------------------------------------------
    double v[8] = {chi, g[0], g[1], g[2], g[3], g[4], g[5], g[6]};
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = wave_sum(v[j]);
    if ((tid & 63) == 0)
        for (int j = 0; j < 8; ++j) s_red[tid >> 6][j] = v[j];
    __syncthreads();
    if (tid == 0) {
        double tot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int wv = 0; wv < kLnlThreads / 64; ++wv)
            for (int j = 0; j < 8; ++j) tot[j] += s_red[wv][j];
        const double sig = s.sigma[pv];
        const double is2 = 1.0 / (sig * sig);
        s.lnl[pv] = s.npt * (-log(sig) - 0.5 * log(2.0 * kPi)) - 0.5 * tot[0] * is2;
        if (s.grad)
            for (int j = 0; j < 7; ++j) s.grad[(size_t)pv * 7 + j] = tot[1 + j] * is2;
    }
}

}  // namespace rr

using namespace rr;

extern "C" {

int rr_lnlike(int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv, const rr_ld* ld,
              const double* obs, const double* sigma, double* lnl, double* grad, void* stream) {
    if (!tab || !lc || !pv || !ld) return set_error(RR_EINVAL, "rr_lnlike: null descriptor");
    if (precision != RR_FP64 && precision != RR_MIXED) return set_error(RR_EINVAL, "rr_lnlike: bad precision");
    if (lc->nlc != 1) return set_error(RR_EINVAL, "rr_lnlike: one light curve only (nlc = %d)", lc->nlc);
    if (pv->kcols != 1) return set_error(RR_EINVAL, "rr_lnlike: k must be [npv, 1]");
    if (ld->law != RR_LD_QUADRATIC && ld->law != RR_LD_POWER_2)
        return set_error(RR_EINVAL, "rr_lnlike: the quadratic and power-2 laws are supported");
    if (ld->nldc < 2 || !ld->ldc) return set_error(RR_EINVAL, "rr_lnlike: two limb-darkening coefficients needed");
    if (tab->ng < 2 || tab->ng > kLnlMaxNg || tab->nz < 1 || tab->nz > kLnlMaxNz || !tab->weights || !tab->mu)
        return set_error(RR_EINVAL, "rr_lnlike: bad weight table (ng <= %d, nz <= %d)", kLnlMaxNg, kLnlMaxNz);
    const int npv = pv->npv, npt = lc->npt;
    if (npv < 0 || npt < 0 || pv->t0cols < 1) return set_error(RR_EINVAL, "rr_lnlike: bad sizes");
    if (npv == 0) return RR_OK;
    if (!lc->times || !lc->nsamples || !lc->exptimes || !lc->epids || !obs || !sigma || !lnl)
        return set_error(RR_EINVAL, "rr_lnlike: null array");
    if (!pv->k || !pv->t0 || !pv->p || !pv->a || !pv->inc || !pv->e || !pv->w)
        return set_error(RR_EINVAL, "rr_lnlike: null parameter array");
    LnlArgs s;
    s.times = lc->times; s.obs = obs; s.sigma = sigma; s.nsamples = lc->nsamples; s.exptimes = lc->exptimes;
    s.epids = lc->epids;
    s.k = pv->k; s.t0 = pv->t0; s.p = pv->p; s.a = pv->a; s.inc = pv->inc; s.e = pv->e; s.w = pv->w;
    s.law = ld->law; s.nldc = ld->nldc; s.ldc = ld->ldc;
    s.W = tab->weights; s.mu = tab->mu; s.lnl = lnl; s.grad = grad;
    s.npv = npv; s.npt = npt; s.kcols = pv->kcols; s.t0cols = pv->t0cols; s.nk = tab->nk; s.ng = tab->ng;
    s.nz = tab->nz; s.kmin = tab->kmin; s.kmax = tab->kmax; s.dk = tab->dk; s.dg = tab->dg;
    hipStream_t st = (hipStream_t)stream;
    if (precision == RR_MIXED) hipLaunchKernelGGL(rr_lnlike_kernel<float>, dim3(npv), dim3(kLnlThreads), 0, st, s);
    else hipLaunchKernelGGL(rr_lnlike_kernel<double>, dim3(npv), dim3(kLnlThreads), 0, st, s);
    return check_launch("rr_lnlike_kernel");
}

}  // extern "C"
