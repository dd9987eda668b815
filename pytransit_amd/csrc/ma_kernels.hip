// Mandel & Agol (2002) quadratic limb-darkening transit model on MI355X (gfx950).
//
// Replaces quadratic_model_v (pytransit/models/numba/ma_quadratic_nb.py:532-595) as
// called by QuadraticModel.evaluate (pytransit/models/ma_quadratic.py:74-140), with
// the analytic flux eval_quad_z_s (ma_quadratic_nb.py:277-389) and its complete
// elliptic integrals (ma_quadratic_nb.py:55-126).  FP64 throughout.
//
// Kernels:
//   ma_orbit_kernel   8 pv per 64-lane block: the Taylor orbit and contact window of
//                     the RoadRunner path (taylor_z.py:23-102,298-328)
//   ma_flux_kernel    one lane per (pv, time): epoch, window (contact times
//                     +- 0.025 d), supersampling, analytic flux; flux [npv, npt]
#include <hip/hip_runtime.h>

#include <cmath>

#include "rr_device.hpp"
#include "rr_internal.hpp"

namespace rr {

// Complete elliptic integral of the third kind by Bulirsch's algorithm (Bulirsch
// 1965, Numer. Math. 7, 78-90, as used at ma_quadratic_nb.py:55-91): arguments
// n and the modulus k; at most 1000 iterations, relative tolerance 1e-8.
__device__ double ma_ellpicb(double n, double k) {
    double kc = sqrt(1.0 - k * k);
    double e = kc;
    double p = sqrt(n + 1.0);
    double m0 = 1.0, c = 1.0, d = 1.0 / p;
    for (int it = 0; it < 1000; ++it) {
        const double f = c;
        c = d / p + c;
        double g = e / p;
        d = 2.0 * (f * g + d);
        p = g + p;
        g = m0;
        m0 = kc + m0;
        if (fabs(1.0 - kc / g) > 1e-8) {
            kc = 2.0 * sqrt(e);
            e = kc * m0;
        } else {
            return kHalfPi * (c * m0 + d) / (m0 * (m0 + p));
        }
    }
    return 0.0;
}

// Complete elliptic integrals E(k) and K(k) by Hastings' polynomial approximations
// (Hastings 1955, Approximations for Digital Computers; Abramowitz & Stegun
// 17.3.34 and 17.3.36), the forms of ma_quadratic_nb.py:93-126.
__device__ __forceinline__ double ma_ellec(double k) {
    const double m1 = 1.0 - k * k;
    const double e1 = 1.0 + m1 * (0.44325141463 + m1 * (0.0626060122 + m1 * (0.04757383546 + m1 * 0.01736506451)));
    const double e2 = m1 * (0.2499836831 + m1 * (0.09200180037 + m1 * (0.04069697526 + m1 * 0.00526449639))) *
                      log(1.0 / m1);
    return e1 + e2;
}

__device__ __forceinline__ double ma_ellk(double k) {
    const double m1 = 1.0 - k * k;
    const double k1 = 1.38629436112 + m1 * (0.09666344259 + m1 * (0.03590092383 + m1 * (0.03742563713 + m1 * 0.01451196212)));
    const double k2 = (0.5 + m1 * (0.12498593597 + m1 * (0.06880248576 + m1 * (0.03328355346 + m1 * 0.00441787012)))) *
                      log(m1);
    return k1 - k2;
}

// Flux of a star with quadratic limb darkening (u1, u2) occulted by a disk of
// radius k at separation z; the case analysis of Mandel & Agol (2002) as in
// ma_quadratic_nb.py:277-389 (uniform part le, linear part ld, quadratic part ed).
__device__ double ma_quad_z(double z, double k, double u1, double u2) {
    const double inv_pi = 1.0 / kPi;
    if (fabs(k - 0.5) < 1.0e-4) k = 0.5;
    const double k2 = k * k;
    const double omega = 1.0 - u1 / 3.0 - u2 / 6.0;
    if (fabs(z - k) < 1e-6) z += 1e-6;
    if (z > 1.0 + k || signbit(z)) return 1.0;
    if (k >= 1.0 && z <= k - 1.0) return 0.0;
    const double z2 = z * z;
    const double x1 = (k - z) * (k - z);
    const double x2 = (k + z) * (k + z);
    const double x3 = k * k - z * z;
    double kap0 = 0.0, kap1 = 0.0, le = 0.0;
    if (z <= 1.0 - k) {
        le = k2;
    } else if (fabs(1.0 - k) <= z && z <= 1.0 + k) {
        kap1 = acos(fmin((1.0 - k2 + z2) / (2.0 * z), 1.0));
        kap0 = acos(fmin((k2 + z2 - 1.0) / (2.0 * k * z), 1.0));
        le = k2 * kap0 + kap1;
        le = (le - 0.5 * sqrt(fmax(4.0 * z2 - (1.0 + z2 - k2) * (1.0 + z2 - k2), 0.0))) * inv_pi;
    }
    const bool edge_at_origin = fabs(z - k) < 1.e-4 * (z + k);
    const bool full_transit = k <= 1.0 && z < (1.0 - k);
    double ld, ed;
    if (edge_at_origin) {
        if (k == 0.5) {
            ld = 1.0 / 3.0 - 4.0 * inv_pi / 9.0;
            ed = 3.0 / 32.0;
        } else if (z > 0.5) {
            const double q = 0.5 / k;
            const double kk = ma_ellk(q), ek = ma_ellec(q);
            ld = 1.0 / 3.0 + 16.0 * k / 9.0 * inv_pi * (2.0 * k2 - 1.0) * ek -
                 (32.0 * k2 * k2 - 20.0 * k2 + 3.0) / 9.0 * inv_pi / k * kk;
            ed = 1.0 / 2.0 * inv_pi *
                 (kap1 + k2 * (k2 + 2.0 * z2) * kap0 - (1.0 + 5.0 * k2 + z2) / 4.0 * sqrt((1.0 - x1) * (x2 - 1.0)));
        } else {
            const double q = 2.0 * k;
            const double kk = ma_ellk(q), ek = ma_ellec(q);
            ld = 1.0 / 3.0 + 2.0 / 9.0 * inv_pi * (4.0 * (2.0 * k2 - 1.0) * ek + (1.0 - 4.0 * k2) * kk);
            ed = k2 / 2.0 * (k2 + 2.0 * z2);
        }
    } else if (full_transit) {
        const double q = sqrt((x2 - x1) / (1.0 - x1));
        const double kk = ma_ellk(q), ek = ma_ellec(q);
        const double n = x2 / x1 - 1.0;
        const double pk = ma_ellpicb(n, q);
        ld = 2.0 / 9.0 * inv_pi / sqrt(1.0 - x1) *
             ((1.0 - 5.0 * z2 + k2 + x3 * x3) * kk + (1.0 - x1) * (z2 + 7.0 * k2 - 4.0) * ek - 3.0 * x3 / x1 * pk);
        if (z < k) ld = ld + 2.0 / 3.0;
        if (fabs(k + z - 1.0) < 1.e-4)
            ld = 2.0 / 3.0 * inv_pi * acos(1.0 - 2.0 * k) -
                 4.0 / 9.0 * inv_pi * sqrt(k * (1.0 - k)) * (3.0 + 2.0 * k - 8.0 * k2);
        ed = k2 / 2.0 * (k2 + 2.0 * z2);
    } else {
        const double q = sqrt((1.0 - (k - z) * (k - z)) / 4.0 / z / k);
        const double kk = ma_ellk(q), ek = ma_ellec(q);
        const double n = 1.0 / x1 - 1.0;
        const double pk = ma_ellpicb(n, q);
        ld = 1.0 / 9.0 * inv_pi / sqrt(k * z) *
             (((1.0 - x2) * (2.0 * x2 + x1 - 3.0) - 3.0 * x3 * (x2 - 2.0)) * kk + 4.0 * k * z * (z2 + 7.0 * k2 - 4.0) * ek -
              3.0 * x3 / x1 * pk);
        if (z < k) ld = ld + 2.0 / 3.0;
        ed = 1.0 / 2.0 * inv_pi *
             (kap1 + k2 * (k2 + 2.0 * z2) * kap0 - (1.0 + 5.0 * k2 + z2) / 4.0 * sqrt((1.0 - x1) * (x2 - 1.0)));
    }
    return 1.0 - ((1.0 - u1 - 2.0 * u2) * le + (u1 + 2.0 * u2) * ld + u2 * ed) / omega;
}

struct MaArgs {
    const double* times; const int32_t* lcids; const int32_t* pbids; const int32_t* epids;
    const int32_t* nsamples; const double* exptimes;
    const double* k; const double* t0; const double* p; const double* a; const double* ldc;
    const PvRec* rec;
    double* flux;
    int npv, npt, nlc, npb, kcols, t0cols, ntiles;
};

constexpr int kMaThreads = 256;

// Orbit and contact window of each pv (the reference computes them for every pv).
__global__ void __launch_bounds__(64) ma_orbit_kernel(int npv, const double* P, const double* A, const double* I,
                                                      const double* E, const double* W, const double* K, int kcols,
                                                      PvRec* rec) {
    const int pv = blockIdx.x * kOrbitGroup + threadIdx.x / kOrbitGroup;
    const int j = threadIdx.x % kOrbitGroup;
    const bool inrange = pv < npv;
    const int q = inrange ? pv : 0;
    const double k0 = K[(size_t)q * kcols];
    PvRec* r = rec + q;
    __shared__ double sm[kOrbitLds];
    orbit_group(pv, inrange, P[q], A[q], I[q], E[q], W[q], k0, r->c, &r->bt1, sm);
    if (inrange && j == 0) {
        r->p = P[q];
        r->inv_p = 1.0 / P[q];
        r->k0 = k0;
        // ma_quadratic_nb.py:549-555: a NaN radius ratio or semi-major axis gives inf
        bool good = !isnan(A[q]);
        for (int c = 0; c < kcols; ++c) good = good && !isnan(K[(size_t)q * kcols + c]);
        r->good = good ? 1 : 0;
        r->ldok = 1;
    }
}

// ma_quadratic_nb.py:557-594
__global__ void __launch_bounds__(kMaThreads) ma_flux_kernel(MaArgs f) {
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int pv = bid / f.ntiles;
    const int it = (bid - pv * f.ntiles) * kMaThreads + threadIdx.x;
    if (it >= f.npt) return;
    const PvRec& r = f.rec[pv];
    double* out = f.flux + (size_t)pv * f.npt + it;
    if (!r.good) {
        *out = INFINITY;
        return;
    }
    const int ilc = f.nlc == 1 ? 0 : min(max(f.lcids[it], 0), f.nlc - 1);
    const int iep = min(max(f.epids[ilc], 0), f.t0cols - 1);
    const double t0v = f.t0[(size_t)pv * f.t0cols + iep];
    const double p = r.p;
    double tc;
    {
#pragma clang fp contract(off)
        const double t = f.times[it];
        const double epoch = floor((t - t0v + 0.5 * p) / p);
        tc = t - (t0v + epoch * p);
    }
    if (!(r.bt1 - 0.025 <= tc && tc <= r.bt4 + 0.025)) {
        *out = 1.0;
        return;
    }
    const int ipb = min(max(f.pbids[ilc], 0), f.npb - 1);
    const double k = f.k[(size_t)pv * f.kcols + (f.kcols == 1 ? 0 : ipb)];
    const double u1 = f.ldc[(size_t)pv * 2 * f.npb + 2 * ipb], u2 = f.ldc[(size_t)pv * 2 * f.npb + 2 * ipb + 1];
    if (isnan(u1) || isnan(u2)) {
        *out = INFINITY;
        return;
    }
    const int ns = f.nsamples[ilc];
    const double ext = f.exptimes[ilc];
    double acc = 0.0;
    for (int is = 1; is <= ns; ++is) {
        const double off = ext * ((is - 0.5) / ns - 0.5);
        const double z = taylor_z<double>(tc + off, r.c);
        acc += (z > 1.0 + k) ? 1.0 : ma_quad_z(z, k, u1, u2);
    }
    *out = acc / ns;
}

}  // namespace rr

using namespace rr;

extern "C" {

size_t ma_workspace_bytes(int npv) { return npv < 0 ? 0 : align_up((size_t)npv * sizeof(PvRec), 256); }

int ma_evaluate(const rr_lcdata* lc, const rr_params* pv, const double* ldc, void* workspace, size_t workspace_bytes,
                double* flux, void* stream) {
    if (!lc || !pv) return set_error(RR_EINVAL, "ma_evaluate: null descriptor");
    const int npv = pv->npv, npt = lc->npt, npb = lc->npb, nlc = lc->nlc;
    if (npv < 0 || npt < 0 || npb < 1 || nlc < 1) return set_error(RR_EINVAL, "ma_evaluate: bad sizes");
    if (pv->kcols != 1 && pv->kcols != npb)
        return set_error(RR_EINVAL, "Radius ratios should be given either as an [npv, 1] or [npv, npb] array.");
    if (pv->t0cols < 1) return set_error(RR_EINVAL, "ma_evaluate: t0 needs at least one column");
    if (npv == 0 || npt == 0) return RR_OK;
    if ((long long)npv * npt > (1LL << 40)) return set_error(RR_ERANGE, "ma_evaluate: problem too large");
    if (!lc->times || !lc->pbids || !lc->epids || !lc->nsamples || !lc->exptimes || (nlc > 1 && !lc->lcids))
        return set_error(RR_EINVAL, "ma_evaluate: null light-curve array");
    if (!pv->k || !pv->t0 || !pv->p || !pv->a || !pv->inc || !pv->e || !pv->w || !ldc || !flux)
        return set_error(RR_EINVAL, "ma_evaluate: null parameter array");
    if (!workspace || workspace_bytes < ma_workspace_bytes(npv))
        return set_error(RR_EINVAL, "ma_evaluate: workspace of %zu bytes needed, got %zu", ma_workspace_bytes(npv),
                         workspace_bytes);
    hipStream_t s = (hipStream_t)stream;
    PvRec* rec = (PvRec*)workspace;
    hipLaunchKernelGGL(ma_orbit_kernel, dim3((npv + kOrbitGroup - 1) / kOrbitGroup), dim3(64), 0, s, npv, pv->p,
                       pv->a, pv->inc, pv->e, pv->w, pv->k, pv->kcols, rec);
    int rc = check_launch("ma_orbit_kernel");
    if (rc) return rc;
    MaArgs f;
    f.times = lc->times; f.lcids = lc->lcids; f.pbids = lc->pbids; f.epids = lc->epids;
    f.nsamples = lc->nsamples; f.exptimes = lc->exptimes;
    f.k = pv->k; f.t0 = pv->t0; f.p = pv->p; f.a = pv->a; f.ldc = ldc; f.rec = rec; f.flux = flux;
    f.npv = npv; f.npt = npt; f.nlc = nlc; f.npb = npb; f.kcols = pv->kcols; f.t0cols = pv->t0cols;
    f.ntiles = (npt + kMaThreads - 1) / kMaThreads;
    const long long nb = (long long)f.ntiles * npv;
    if (nb > 0x7fffffffLL) return set_error(RR_ERANGE, "ma_evaluate: grid too large");
    hipLaunchKernelGGL(ma_flux_kernel, dim3((unsigned)nb), dim3(kMaThreads), 0, s, f);
    return check_launch("ma_flux_kernel");
}

}  // extern "C"
