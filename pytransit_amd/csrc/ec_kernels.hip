// Secondary-eclipse, eclipse-spectroscopy and oblate-planet variants of the
// RoadRunner model on MI355X (gfx950).  FP64 throughout, as the reference.
//
// Replaces
//   eclipse_model   pytransit/models/roadrunner/model_eclipse.py:11-81
//                   (EclipseModel.evaluate, models/new_eclipse_model.py:34-70)
//   esmodel         models/roadrunner/model_ecspec.py:13-65
//                   (EclipseSpectroscopyModel.evaluate, models/roadrunner/esmodel.py:46-90)
//   op_full_serial  models/roadrunner/opmodel_full.py:40-134 with the ellipse
//                   scanlines of models/roadrunner/ecintersection.py:93-207
//                   (OblatePlanetModel.evaluate, models/roadrunner/opmodel.py:50-111)
//
// Kernels:
//   ec_orbit_kernel   8 pv per 64-lane block: the mid-eclipse time from the
//                     conjunction phase (orbits/orbits_py.py:544-555), the 7 Kepler
//                     solves of the Taylor stencil about it (orbits/taylor_z.py:
//                     106-185), the light travel time and the contact window
//   ec_flux_kernel    one lane per (pv, time): the planet disk area left visible by
//                     the star, written as the eclipse model's area or fanned out
//                     over the passbands of the spectroscopy model
//   op_flux_kernel    one block per (pv, 256 times): the pv's ellipse scanlines in
//                     LDS, then per lane the supersampled flux with the scanline
//                     ellipse-disk area; the per-pv state comes from rr_setup
#include <hip/hip_runtime.h>

#include <cmath>

#include "rr_device.hpp"
#include "rr_internal.hpp"

namespace rr {

// Per-pv state of the eclipse models.
struct alignas(16) EcRec {
    double c[10];      // Taylor coefficients about mid-eclipse (x then y, divided by n!)
    double bt1, bt4;   // first and fourth contact about mid-eclipse, before the margin
    double p;
    double shift;      // mid-eclipse time after mid-transit, light travel time included
    double k;
    int32_t good, pad_;
};
static_assert(sizeof(EcRec) == 128, "EcRec must be 128 bytes");

constexpr double kRsunM = 6.957e8;          // IAU 2015 nominal solar radius [m]
constexpr double kCLight = 299792458.0;     // [m/s]
constexpr double kDaySec = 86400.0;

// Time from mid-transit to mid-eclipse: the mean-anomaly difference between the
// two conjunctions (true anomalies pi/2 - w and 3pi/2 - w); orbits_py.py:544-555.
__device__ __forceinline__ double eclipse_phase(double p, double e, double w) {
    const double se = sqrt(1.0 - e * e);
    const double etr = atan2(se * sin(kHalfPi - w), e + cos(kHalfPi - w));
    const double eec = atan2(se * sin(kHalfPi + kPi - w), e + cos(kHalfPi + kPi - w));
    const double mtr = etr - e * sin(etr);
    const double mec = eec - e * sin(eec);
    const double ph = (mec - mtr) * p / kTwoPi;
    return ph > 0.0 ? ph : p + ph;
}

// Extra light path of the eclipse over the transit, (r_tr + r_ec) sin i, in days
// for a star of rstar solar radii (r: star-planet distance at each conjunction).
__device__ __forceinline__ double light_travel_time(double a, double inc, double e, double w, double rstar) {
    const double om = a * (1.0 - e * e);
    const double r_tr = om / (1.0 + e * sin(w));
    const double r_ec = om / (1.0 - e * sin(w));
    return (r_tr + r_ec) * sin(inc) * rstar * kRsunM / kCLight / kDaySec;
}

__global__ void __launch_bounds__(64) ec_orbit_kernel(int npv, const double* P, const double* A, const double* I,
                                                      const double* E, const double* W, const double* K,
                                                      const double* rstar, EcRec* rec) {
    const int sub = threadIdx.x / kOrbitGroup, j = threadIdx.x % kOrbitGroup;
    const int pv = blockIdx.x * kOrbitGroup + sub;
    const bool inrange = pv < npv;
    const int q = inrange ? pv : 0;
    __shared__ double s_x[kOrbitGroup][8], s_y[kOrbitGroup][8], s_c[kOrbitGroup][10], s_w[kOrbitGroup][2];
    const double p = P[q], a = A[q], inc = I[q], e = E[q], w = W[q], k = K[q];
    const bool valid = inrange && !(isnan(a) || a <= 1.0 || e < 0.0);
    const double te = valid ? eclipse_phase(p, e, w) : 0.0;
    if (valid && j < 7) {
        double x, y;
        kepler_xy(te + (j - 3) * 2e-2, p, a, cos(inc), e, w, x, y);
        s_x[sub][j] = x;
        s_y[sub][j] = y;
    }
    __syncthreads();
    if (valid && j == 0) taylor_from_stencil(s_x[sub], s_y[sub], s_c[sub]);
    __syncthreads();
    if (valid && j < 2) s_w[sub][j] = contact_bisect(k, j == 0 ? 1 : 4, s_c[sub], fabs(s_c[sub][1]));
    __syncthreads();
    if (!inrange) return;
    EcRec* r = rec + pv;
    if (valid) {
        r->c[j] = s_c[sub][j];
        if (j < 2) r->c[8 + j] = s_c[sub][8 + j];
    }
    if (j == 0) {
        r->bt1 = valid ? s_w[sub][0] : 0.0;
        r->bt4 = valid ? s_w[sub][1] : 0.0;
        r->p = p;
        r->shift = valid ? te + light_travel_time(a, inc, e, w, rstar[q]) : 0.0;
        r->k = k;
        r->good = valid ? 1 : 0;
    }
}

struct EcArgs {
    const double* times; const int32_t* lcids; const int32_t* epids;
    const int32_t* nsamples; const double* exptimes;
    const double* t0; const EcRec* rec; const double* fratio;
    double* flux;
    int npv, npt, nlc, npb, t0cols, ntiles;
    int spectro;       // 0: eclipse model (visible planet area), 1: spectroscopy model
};

constexpr int kEcThreads = 256;

// model_eclipse.py:56-80 (spectro = 0) and model_ecspec.py:51-64 (spectro = 1).
__global__ void __launch_bounds__(kEcThreads) ec_flux_kernel(EcArgs f) {
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int pv = bid / f.ntiles;
    const int it = (bid - pv * f.ntiles) * kEcThreads + threadIdx.x;
    if (it >= f.npt) return;
    const EcRec& r = f.rec[pv];
    const size_t row = (size_t)f.npt;
    double* out = f.flux + (size_t)pv * (f.spectro ? (size_t)f.npb : 1) * row + it;
    if (!r.good) {
        if (f.spectro)
            for (int b = 0; b < f.npb; ++b) out[b * row] = NAN;
        else
            *out = NAN;
        return;
    }
    // the spectroscopy model reads the first light curve's sampling and a 1-D t0
    const int ilc = (f.spectro || f.nlc == 1) ? 0 : min(max(f.lcids[it], 0), f.nlc - 1);
    const int iep = f.spectro ? 0 : min(max(f.epids[ilc], 0), f.t0cols - 1);
    const double te = f.t0[(size_t)pv * f.t0cols + iep] + r.shift;
    const double p = r.p;
    double tc;
    {
#pragma clang fp contract(off)
        const double t = f.times[it];
        const double epoch = floor((t - te + 0.5 * p) / p);
        tc = t - (te + epoch * p);
    }
    const int ns = f.nsamples[ilc];
    const double ext = f.exptimes[ilc];
    const double margin = (f.spectro ? 0.0015 : 0.003) + ext;
    const double k = r.k, k2 = kPi * (k * k);
    const bool inw = r.bt1 - margin <= tc && tc <= r.bt4 + margin;
    if (!f.spectro) {
        double acc = 0.0;
        if (inw) {
            for (int is = 1; is <= ns; ++is) {
                const double z = taylor_z<double>(tc + ext * ((is - 0.5) / ns - 0.5), r.c);
                double area, kap;
                ccia_kite<double>(1.0, k, z, area, kap);
                acc += k2 - area;
            }
        }
        *out = inw ? acc / ns : k2;
        return;
    }
    // spectroscopy: the mean area over the exposure, then per passband
    // 1 - (f A / pi) / (1 + f k^2) (the per-sample terms of the reference, averaged)
    double amean = 0.0;
    if (inw) {
        double s = 0.0;
        for (int is = 1; is <= ns; ++is) {
            const double z = taylor_z<double>(tc + ext * ((is - 0.5) / ns - 0.5), r.c);
            double area, kap;
            ccia_kite<double>(1.0, k, z, area, kap);
            s += area;
        }
        amean = s / ns;
    }
    const double* fr = f.fratio + (size_t)pv * f.npb;
    for (int b = 0; b < f.npb; ++b) {
        const double fb = fr[b];
        out[b * row] = inw ? 1.0 - (fb * amean / kPi) / (1.0 + fb * (k * k)) : 1.0;
    }
}

// ---- Oblate planet ---------------------------------------------------------------

constexpr int kOpThreads = 256;
constexpr int kOpMaxLines = 2048;

struct OpArgs {
    const double* times; const int32_t* lcids; const int32_t* pbids; const int32_t* epids;
    const int32_t* nsamples; const double* exptimes;
    const double* k; const double* t0; const double* fl; const double* alpha;
    const PvRec* rec; const double* istar; const double* ldm;
    double* flux;
    int npv, npt, nlc, npb, kcols, t0cols, ng, nlines, ntiles;
    double dg;
};

// Area of the ellipse with rows (ys, xl, xr) about its centre (cx, cy) inside the
// unit disk; ecintersection.py:137-207, the three branches by the centre's x.  The
// rows are ordered and a row with no chord has xl = NaN.
__device__ double op_ellipse_area(double cx, double cy, double z, double k, double fl, const double* ys,
                                  const double* xl, const double* xr, int ny) {
    if (z <= 1.0 - k) return kPi * k * (1.0 - fl) * k;
    if (z >= 1.0 + k) return 0.0;
    const double dy = ys[1] - ys[0];
    double acc = 0.0;
    if (cx < -k) {
        for (int j = 0; j < ny; ++j) {
            const double x0 = xl[j], x1 = xr[j];
            if (isfinite(x0)) {
                const double v = ys[j] + cy;
                const double xs = -sqrt(1.0 - v * v) - cx;   // NaN off the disk: no branch taken
                if (xs <= x0) acc += x1 - x0;
                else if (xs <= x1) acc += x1 - xs;
            }
        }
    } else if (cx > k) {
        for (int j = 0; j < ny; ++j) {
            const double x0 = xl[j], x1 = xr[j];
            if (isfinite(x0)) {
                const double v = ys[j] + cy;
                const double xs = sqrt(1.0 - v * v) - cx;
                if (xs >= x1) acc += x1 - x0;
                else if (xs >= x0) acc += xs - x0;
            }
        }
    } else {
        for (int j = 0; j < ny; ++j) {
            const double x0 = xl[j], x1 = xr[j];
            if (isfinite(x0)) {
                const double v = ys[j] + cy;
                if (fabs(v) <= 1.0) {
                    const double xs = sqrt(1.0 - v * v);
                    const double a1 = -xs - cx, a2 = xs - cx;
                    const double hi = fmin(a2, x1);
                    if (a1 <= x0) acc += hi - x0;
                    else if (a1 > x0) acc += hi - a1;
                }
            }
        }
    }
    return acc * dy;
}

__global__ void __launch_bounds__(kOpThreads) op_flux_kernel(OpArgs f) {
    __shared__ double s_y[kOpMaxLines], s_xl[kOpMaxLines], s_xr[kOpMaxLines];
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int pv = bid / f.ntiles;
    const int it = (bid - pv * f.ntiles) * kOpThreads + threadIdx.x;
    const PvRec& r = f.rec[pv];
    const bool ok = r.good && r.ldok;
    // the pv's ellipse scanlines (create_ellipse, ecintersection.py:93-134) at k[:, 0]
    if (ok) {
        const double k0 = f.k[(size_t)pv * f.kcols];
        const double fl = f.fl[pv], al = f.alpha[pv];
        const double ca = cos(al), sa = sin(al);
        const double b = 1.0 - fl, b2 = b * b, ca2 = ca * ca, sa2 = sa * sa;
        const double dy = 2.0 / (f.nlines + 1);
        for (int j = threadIdx.x; j < f.nlines; j += kOpThreads) {
            const double y = (double)(j + 1) * dy - 1.0;
            const double y2 = y * y;
            double d = b2 * ca2 + ca2 * ca2 * (-y2) - 2.0 * ca2 * sa2 * y2 - sa2 * sa2 * y2 + sa2;
            double xl = NAN, xr = NAN;
            if (d >= 0.0) {
                d = sqrt(d) / b;
                const double u = (ca * sa * y) / b2 - ca * sa * y;
                const double v = sa2 / b2 + ca2;
                xl = (-d + u) / v;
                xr = (d + u) / v;
            }
            s_y[j] = k0 * y;
            s_xl[j] = k0 * xl;
            s_xr[j] = k0 * xr;
        }
    }
    __syncthreads();
    if (it >= f.npt) return;
    double* out = f.flux + (size_t)pv * f.npt + it;
    if (!ok) {
        *out = NAN;
        return;
    }
    const int ilc = f.nlc == 1 ? 0 : min(max(f.lcids[it], 0), f.nlc - 1);
    const int ipb = min(max(f.pbids[ilc], 0), f.npb - 1);
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
    const double ext = f.exptimes[ilc];
    const double margin = 0.0015 + ext;
    if (!(r.bt1 - margin <= tc && tc <= r.bt4 + margin)) {
        *out = 1.0;
        return;
    }
    const int ns = f.nsamples[ilc];
    const double k = f.k[(size_t)pv * f.kcols + (f.kcols == 1 ? 0 : ipb)];
    const double ist = f.istar[(size_t)pv * f.npb + ipb];
    const double* lrow = f.ldm + ((size_t)pv * f.npb + ipb) * f.ng;
    const double fl = f.fl[pv];
    double acc = 0.0;
    for (int is = 1; is <= ns; ++is) {
        const double tt = tc + ext * ((is - 0.5) / ns - 0.5);
        const double cx = (((r.c[4] * tt + r.c[3]) * tt + r.c[2]) * tt + r.c[1]) * tt + r.c[0];
        const double cy = (((r.c[9] * tt + r.c[8]) * tt + r.c[7]) * tt + r.c[6]) * tt + r.c[5];
        const double z = sqrt(cx * cx + cy * cy);
        const double ip = interp_ld(z / (1.0 + k), f.dg, lrow, f.ng);
        const double ap = op_ellipse_area(cx, cy, z, k, fl, s_y, s_xl, s_xr, f.nlines);
        acc += (ist - ip * ap) / ist;
    }
    *out = acc / ns;
}

}  // namespace rr

using namespace rr;

extern "C" {

size_t ec_workspace_bytes(int npv) { return npv < 0 ? 0 : align_up((size_t)npv * sizeof(EcRec), 256); }

static int ec_common(const char* who, const rr_lcdata* lc, const rr_params* pv, const double* rstar, void* workspace,
                     size_t workspace_bytes, double* flux, int spectro, const double* fratio, int npb,
                     void* stream) {
    if (!lc || !pv) return set_error(RR_EINVAL, "%s: null descriptor", who);
    const int npv = pv->npv, npt = lc->npt, nlc = lc->nlc;
    if (npv < 0 || npt < 0 || nlc < 1 || pv->t0cols < 1 || (spectro && npb < 1))
        return set_error(RR_EINVAL, "%s: bad sizes (npv=%d npt=%d nlc=%d npb=%d)", who, npv, npt, nlc, npb);
    if (pv->kcols != 1) return set_error(RR_EINVAL, "%s: the radius ratio must be one value per pv", who);
    if (npv == 0 || npt == 0) return RR_OK;
    if ((long long)npv * npt * (spectro ? npb : 1) > (1LL << 40))
        return set_error(RR_ERANGE, "%s: problem too large", who);
    if (!lc->times || !lc->nsamples || !lc->exptimes || (!spectro && (!lc->epids || (nlc > 1 && !lc->lcids))))
        return set_error(RR_EINVAL, "%s: null light-curve array", who);
    if (!pv->k || !pv->t0 || !pv->p || !pv->a || !pv->inc || !pv->e || !pv->w || !rstar || !flux ||
        (spectro && !fratio))
        return set_error(RR_EINVAL, "%s: null parameter array", who);
    if (!workspace || workspace_bytes < ec_workspace_bytes(npv))
        return set_error(RR_EINVAL, "%s: workspace of %zu bytes needed, got %zu", who, ec_workspace_bytes(npv),
                         workspace_bytes);
    hipStream_t s = (hipStream_t)stream;
    EcRec* rec = (EcRec*)workspace;
    hipLaunchKernelGGL(ec_orbit_kernel, dim3((npv + kOrbitGroup - 1) / kOrbitGroup), dim3(64), 0, s, npv, pv->p,
                       pv->a, pv->inc, pv->e, pv->w, pv->k, rstar, rec);
    int rc = check_launch("ec_orbit_kernel");
    if (rc) return rc;
    EcArgs f;
    f.times = lc->times; f.lcids = lc->lcids; f.epids = lc->epids; f.nsamples = lc->nsamples;
    f.exptimes = lc->exptimes; f.t0 = pv->t0; f.rec = rec; f.fratio = fratio; f.flux = flux;
    f.npv = npv; f.npt = npt; f.nlc = nlc; f.npb = spectro ? npb : 1; f.t0cols = pv->t0cols;
    f.ntiles = (npt + kEcThreads - 1) / kEcThreads;
    f.spectro = spectro;
    const long long nb = (long long)f.ntiles * npv;
    if (nb > 0x7fffffffLL) return set_error(RR_ERANGE, "%s: grid too large", who);
    hipLaunchKernelGGL(ec_flux_kernel, dim3((unsigned)nb), dim3(kEcThreads), 0, s, f);
    return check_launch("ec_flux_kernel");
}

int eclipse_evaluate(const rr_lcdata* lc, const rr_params* pv, const double* rstar, void* workspace,
                     size_t workspace_bytes, double* flux, void* stream) {
    return ec_common("eclipse_evaluate", lc, pv, rstar, workspace, workspace_bytes, flux, 0, nullptr, 0, stream);
}

int es_evaluate(const rr_lcdata* lc, const rr_params* pv, const double* rstar, const double* fratio, int npb,
                void* workspace, size_t workspace_bytes, double* flux, void* stream) {
    if (pv && pv->t0cols != 1) return set_error(RR_EINVAL, "es_evaluate: t0 must be one value per pv");
    return ec_common("es_evaluate", lc, pv, rstar, workspace, workspace_bytes, flux, 1, fratio, npb, stream);
}

int op_evaluate(const rr_table* tab, const rr_lcdata* lc, const rr_params* pv, const rr_ld* ld, const double* fl,
                const double* alpha, int nlines, void* workspace, size_t workspace_bytes, double* flux,
                void* stream) {
    if (nlines < 2 || nlines > kOpMaxLines)
        return set_error(RR_ERANGE, "op_evaluate: nlines must be in [2, %d], got %d", kOpMaxLines, nlines);
    if (!pv || !lc || !tab) return set_error(RR_EINVAL, "op_evaluate: null descriptor");
    if (pv->npv > 0 && lc->npt > 0 && (!fl || !alpha)) return set_error(RR_EINVAL, "op_evaluate: null f or alpha");
    // the per-pv state (orbit, window, limb-darkening means, disk integrals) is rr_full's
    int rc = rr_setup(RR_FP64, tab, lc, pv, ld, workspace, workspace_bytes, stream);
    if (rc) return rc;
    const int npv = pv->npv, npt = lc->npt, npb = lc->npb;
    if (npv == 0 || npt == 0) return RR_OK;
    if (!flux) return set_error(RR_EINVAL, "op_evaluate: null flux");
    // the workspace layout of rr_workspace_bytes: PvRec[npv] | istar[npv, npb] | ldm64[npv, npb, ng]
    char* ws = (char*)workspace;
    OpArgs f;
    f.rec = (const PvRec*)ws;
    ws += align_up((size_t)npv * sizeof(PvRec), 256);
    f.istar = (const double*)ws;
    ws += align_up((size_t)npv * npb * sizeof(double), 256);
    f.ldm = (const double*)ws;
    f.times = lc->times; f.lcids = lc->lcids; f.pbids = lc->pbids; f.epids = lc->epids;
    f.nsamples = lc->nsamples; f.exptimes = lc->exptimes;
    f.k = pv->k; f.t0 = pv->t0; f.fl = fl; f.alpha = alpha; f.flux = flux;
    f.npv = npv; f.npt = npt; f.nlc = lc->nlc; f.npb = npb; f.kcols = pv->kcols; f.t0cols = pv->t0cols;
    f.ng = tab->ng; f.nlines = nlines; f.dg = tab->dg;
    f.ntiles = (npt + kOpThreads - 1) / kOpThreads;
    const long long nb = (long long)f.ntiles * npv;
    if (nb > 0x7fffffffLL) return set_error(RR_ERANGE, "op_evaluate: grid too large");
    if (!lc->times || !lc->nsamples || !lc->exptimes || (lc->nlc > 1 && !lc->lcids))
        return set_error(RR_EINVAL, "op_evaluate: null light-curve array");
    hipLaunchKernelGGL(op_flux_kernel, dim3((unsigned)nb), dim3(kOpThreads), 0, (hipStream_t)stream, f);
    return check_launch("op_flux_kernel");
}

}  // extern "C"
