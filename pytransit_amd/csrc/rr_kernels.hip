// RoadRunner transit model on MI355X (gfx950): weight table, limb-darkening
// profiles, per-pv setup and the flux kernels, plus the C-ABI of include/rr_hip.h.
//
// Data layout in HBM (all row-major):
//   times[npt] f64, lcids[npt] i32, per-lc arrays [nlc]
//   k[npv, kcols], t0[npv, nep], p/a/inc/e/w[npv] f64
//   ldp[npv, npb, nz] f64, istar[npv, npb] f64
//   workspace: PvRec[npv] (128 B each) | ldm64[npv, npb, ng] | ldm32[npv, npb, ng]
//   flux[npv, npt] f64 (RR_FP64) or f32 (RR_MIXED)
//
// The flux kernel maps one workgroup of 256 lanes to one parameter vector and a
// tile of 1024 consecutive times, so a wavefront reads 64 consecutive times
// (coalesced) and writes 64 consecutive flux values; consecutive times share the
// in/out-of-window branch, so wavefronts rarely diverge.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "rr_device.hpp"
#include "rr_internal.hpp"
#include "../../include/rr_hip.h"

namespace rr {

thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int check_launch(const char* what) {
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return set_error(RR_EHIP, "%s: %s", what, hipGetErrorString(err));
    return RR_OK;
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// ------------------------------------------------------------------------------------
// Weight table; common.py:188-223.  One lane per (k, g) row, areas accumulated
// left to right as in the reference so the normalising sum rounds the same way.
// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) weights_3d_kernel(int nk, double kmin, double kmax, const double* __restrict__ ze,
                                                         int nz, int ng, double* __restrict__ W,
                                                         float* __restrict__ W32) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nk * ng) return;
    const int ik = row / ng, ig = row - ik * ng;
    const double k = linspace_at(kmin, kmax, nk, ik);
    const double g = linspace_at(0.0, 1.0 - 1e-7, ng, ig);
    const double b = g * (1.0 + k);
    double* w = W + (size_t)row * nz;
    double a0 = cci_area(ze[0], k, b);
    w[0] = a0;
    double s = a0;
    for (int i = 1; i < nz; ++i) {
        const double a1 = cci_area(ze[i], k, b);
        w[i] = a1 - a0;
        a0 = a1;
        s += w[i];
    }
    for (int i = 0; i < nz; ++i) {
        const double v = w[i] / s;
        w[i] = v;
        if (W32) W32[(size_t)row * nz + i] = (float)v;
    }
}

// ------------------------------------------------------------------------------------
// Limb-darkening laws; models/numba/ldmodels.py:22-139.
// ------------------------------------------------------------------------------------
__device__ double ld_law(int law, double mu, const double* c, int nc) {
    switch (law) {
        case RR_LD_UNIFORM: return 1.0;
        case RR_LD_LINEAR: return 1.0 - c[0] * (1.0 - mu);
        case RR_LD_QUADRATIC: return 1.0 - c[0] * (1.0 - mu) - c[1] * ((1.0 - mu) * (1.0 - mu));
        case RR_LD_QUADRATIC_TRI: {
            const double a = sqrt(c[0]), b = 2.0 * c[1];
            const double u = a * b, v = a * (1.0 - b);
            return 1.0 - u * (1.0 - mu) - v * ((1.0 - mu) * (1.0 - mu));
        }
        case RR_LD_NONLINEAR:
            return 1.0 - c[0] * (1.0 - sqrt(mu)) - c[1] * (1.0 - mu) - c[2] * (1.0 - pow(mu, 1.5)) -
                   c[3] * (1.0 - mu * mu);
        case RR_LD_GENERAL: {
            double s = 0.0;
            for (int i = 0; i < nc; ++i) s += c[i] * (1.0 - pow(mu, (double)(i + 1)));
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

// Analytic disk integral where the reference has one (ldmodels.py:27-123); returns
// false for the laws that fall back to the trapezoid (rrmodel.py:220-227, Q8).
__device__ bool ld_integral(int law, const double* c, double& out) {
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

__global__ void __launch_bounds__(256) ld_profiles_kernel(int law, const double* __restrict__ ldc, int npv, int npb,
                                                          int nldc, const double* __restrict__ mu, int nmu,
                                                          double* __restrict__ ldp, double* __restrict__ istar) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= npv * npb) return;
    const double* c = ldc + (size_t)j * nldc;
    double* out = ldp + (size_t)j * nmu;
    for (int i = 0; i < nmu; ++i) out[i] = ld_law(law, mu[i], c, nldc);
    double is;
    if (!ld_integral(law, c, is)) {
        // 2*pi*trapezoid(z*I(mu), z) on mu = linspace(1, 0, 200), z = sqrt(1-mu^2)
        const int n = 200;
        double zp = 0.0, yp = 0.0, s = 0.0;
        for (int m = 0; m < n; ++m) {
            const double muv = linspace_at(1.0, 0.0, n, m);
            const double z = sqrt(1.0 - muv * muv);
            const double y = z * ld_law(law, muv, c, nldc);
            if (m > 0) s += (z - zp) * (y + yp) / 2.0;
            zp = z;
            yp = y;
        }
        is = 2.0 * kPi * s;
    }
    istar[j] = is;
}

// istar = 2*pi*sum dz*(z_i I_i + z_{i-1} I_{i-1})/2; ldtkldm.py:77-91.
__global__ void __launch_bounds__(256) integrate_profiles_kernel(const double* __restrict__ mu, int nmu,
                                                                 const double* __restrict__ ldp, int n,
                                                                 double* __restrict__ istar) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double* l = ldp + (size_t)j * nmu;
    double s = 0.0;
    double zp = sqrt(1.0 - mu[0] * mu[0]);
    for (int i = 1; i < nmu; ++i) {
        const double z = sqrt(1.0 - mu[i] * mu[i]);
        s += (z - zp) * 0.5 * (z * l[i] + zp * l[i - 1]);
        zp = z;
    }
    istar[j] = 2.0 * kPi * s;
}

// ------------------------------------------------------------------------------------
// Orbit: Taylor coefficients (7 Kepler solves spread over lanes 0-6) and the two
// contact points (lanes 0-1).  Used by the RR and TS setup kernels.
// ------------------------------------------------------------------------------------
__device__ void orbit_setup_block(double p, double a, double inc, double e, double w, double kw, double* s_x,
                                  double* s_y, double* s_c, double* s_win) {
    const int lane = threadIdx.x;
    if (lane < 7) {
        double x, y;
        kepler_xy((lane - 3) * 2e-2, p, a, cos(inc), e, w, x, y);
        s_x[lane] = x;
        s_y[lane] = y;
    }
    __syncthreads();
    if (lane == 0) taylor_from_stencil(s_x, s_y, s_c);
    __syncthreads();
    if (lane < 2) s_win[lane] = find_contact_point(kw, lane == 0 ? 1 : 4, s_c);
    __syncthreads();
}

__global__ void __launch_bounds__(64) orbit_kernel(int npv, const double* __restrict__ P, const double* __restrict__ A,
                                                   const double* __restrict__ I, const double* __restrict__ E,
                                                   const double* __restrict__ Wp, const double* __restrict__ K0,
                                                   double* __restrict__ coeffs, double* __restrict__ win) {
    __shared__ double s_x[7], s_y[7], s_c[10], s_win[2];
    const int pv = blockIdx.x;
    orbit_setup_block(P[pv], A[pv], I[pv], E[pv], Wp[pv], K0[pv], s_x, s_y, s_c, s_win);
    if (threadIdx.x < 10) coeffs[pv * 10 + threadIdx.x] = s_c[threadIdx.x];
    if (threadIdx.x < 2) win[pv * 2 + threadIdx.x] = s_win[threadIdx.x];
}

// ------------------------------------------------------------------------------------
// Per-pv setup of rr_full; model_full.py:39-70.  One 64-lane workgroup per pv.
// ------------------------------------------------------------------------------------
struct SetupArgs {
    const double* k; const double* p; const double* a; const double* inc; const double* e; const double* w;
    const double* ldp; const double* W; const double* ze;
    PvRec* rec; double* ldm64; float* ldm32;
    int npb, kcols, nk, ng, nz;
    double kmin, kmax, dk;
};

__global__ void __launch_bounds__(64) rr_setup_kernel(SetupArgs s) {
    __shared__ double s_x[7], s_y[7], s_c[10], s_win[2];
    const int pv = blockIdx.x;
    const int lane = threadIdx.x;
    const double a = s.a[pv], e = s.e[pv];
    const double* ldp_pv = s.ldp + (size_t)pv * s.npb * s.nz;
    const bool good = !(isnan(a) || a <= 1.0 || e < 0.0 || isnan(ldp_pv[0]));
    PvRec* rec = s.rec + pv;
    if (!good) {
        if (lane == 0) { rec->good = 0; }
        return;
    }
    const double k0 = s.k[(size_t)pv * s.kcols];

    // Limb-darkening means: lerp in k of two table rows, or exact 2-D weights
    // when k0 is outside [kmin, kmax]; model_full.py:47-55.
    const size_t ldm_base = (size_t)pv * s.npb * s.ng;
    if (s.kmin <= k0 && k0 <= s.kmax) {
        int ik = (int)floor((k0 - s.kmin) / s.dk);
        ik = ik > s.nk - 2 ? s.nk - 2 : ik;
        const double ak = (k0 - s.kmin - ik * s.dk) / s.dk;
        const double* W0 = s.W + (size_t)ik * s.ng * s.nz;
        const double* W1 = W0 + (size_t)s.ng * s.nz;
        for (int ipb = 0; ipb < s.npb; ++ipb) {
            const double* l = ldp_pv + (size_t)ipb * s.nz;
            for (int ig = lane; ig < s.ng; ig += 64) {
                double d0 = 0.0, d1 = 0.0;
                for (int iz = 0; iz < s.nz; ++iz) {
                    d0 += W0[ig * s.nz + iz] * l[iz];
                    d1 += W1[ig * s.nz + iz] * l[iz];
                }
                const double v = (1.0 - ak) * d0 + ak * d1;
                s.ldm64[ldm_base + (size_t)ipb * s.ng + ig] = v;
                if (s.ldm32) s.ldm32[ldm_base + (size_t)ipb * s.ng + ig] = (float)v;
            }
        }
    } else {
        for (int ig = lane; ig < s.ng; ig += 64) {
            const double g = linspace_at(0.0, 1.0 - 1e-7, s.ng, ig);
            const double b = g * (1.0 + k0);
            // pass 1: normalising sum of the annulus weights (common.py:152-185)
            double a0 = cci_area(s.ze[0], k0, b), sum = a0;
            for (int iz = 1; iz < s.nz; ++iz) {
                const double a1 = cci_area(s.ze[iz], k0, b);
                sum += a1 - a0;
                a0 = a1;
            }
            for (int ipb = 0; ipb < s.npb; ++ipb) {
                const double* l = ldp_pv + (size_t)ipb * s.nz;
                double b0 = cci_area(s.ze[0], k0, b);
                double d = (b0 / sum) * l[0];
                for (int iz = 1; iz < s.nz; ++iz) {
                    const double b1 = cci_area(s.ze[iz], k0, b);
                    d += ((b1 - b0) / sum) * l[iz];
                    b0 = b1;
                }
                s.ldm64[ldm_base + (size_t)ipb * s.ng + ig] = d;
                if (s.ldm32) s.ldm32[ldm_base + (size_t)ipb * s.ng + ig] = (float)d;
            }
        }
    }

    // Taylor orbit and contact window; model_full.py:60-67.
    orbit_setup_block(s.p[pv], a, s.inc[pv], e, s.w[pv], k0, s_x, s_y, s_c, s_win);
    if (lane < 10) rec->c[lane] = s_c[lane];
    if (lane == 0) {
        rec->bt1 = s_win[0];
        rec->bt4 = s_win[1];
        rec->p = s.p[pv];
        rec->inv_p = 1.0 / s.p[pv];
        rec->k0 = k0;
        rec->good = 1;
    }
}

// ------------------------------------------------------------------------------------
// Flux; model_full.py:75-99.
// ------------------------------------------------------------------------------------
struct FluxArgs {
    const double* times; const int32_t* lcids; const int32_t* pbids; const int32_t* epids;
    const int32_t* nsamples; const double* exptimes;
    const double* k; const double* t0; const double* istar;
    const PvRec* rec; const double* ldm64; const float* ldm32;
    void* flux;
    int npt, nlc, npb, nep, kcols, t0cols, ng, ntiles;
    double dg;
};

constexpr int kFluxThreads = 256;
constexpr int kFluxItems = 4;
constexpr int kFluxTile = kFluxThreads * kFluxItems;

template <bool kMixed, bool kOneLC>
__global__ void __launch_bounds__(kFluxThreads) rr_flux_kernel(FluxArgs f) {
    const int pv = blockIdx.x / f.ntiles;
    const int tile = blockIdx.x - pv * f.ntiles;
    const int base = tile * kFluxTile + threadIdx.x;
    const PvRec& r = f.rec[pv];
    const size_t out0 = (size_t)pv * f.npt;

    if (!r.good) {
        for (int j = 0; j < kFluxItems; ++j) {
            const int it = base + j * kFluxThreads;
            if (it >= f.npt) break;
            if (kMixed) ((float*)f.flux)[out0 + it] = NAN;
            else ((double*)f.flux)[out0 + it] = NAN;
        }
        return;
    }

    const double p = r.p, inv_p = r.inv_p;
    float cf[10];
    if (kMixed) {
#pragma unroll
        for (int i = 0; i < 10; ++i) cf[i] = (float)r.c[i];
    }
    const float inv_dg = (float)(1.0 / f.dg);

    for (int j = 0; j < kFluxItems; ++j) {
        const int it = base + j * kFluxThreads;
        if (it >= f.npt) break;
        const int ilc = kOneLC ? 0 : min(max(f.lcids[it], 0), f.nlc - 1);
        const int ipb = min(max(f.pbids[ilc], 0), f.npb - 1);
        const int iep = min(max(f.epids[ilc], 0), f.t0cols - 1);
        const double ext = f.exptimes[ilc];
        const double t0v = f.t0[(size_t)pv * f.t0cols + iep];
        const double t = f.times[it];
        const double tc = kMixed ? phase_fast(t, t0v, p, inv_p) : phase_exact(t, t0v, p);
        const double margin = 0.003 + ext;
        const bool inwin = (r.bt1 - margin <= tc) && (tc <= r.bt4 + margin);
        if (!inwin) {
            if (kMixed) ((float*)f.flux)[out0 + it] = 1.0f;
            else ((double*)f.flux)[out0 + it] = 1.0;
            continue;
        }
        const int ns = f.nsamples[ilc];
        const double kk = f.k[(size_t)pv * f.kcols + (f.kcols == 1 ? 0 : ipb)];
        const double ist = f.istar[(size_t)pv * f.npb + ipb];
        const size_t row = ((size_t)pv * f.npb + ipb) * f.ng;
        if (kMixed) {
            const float kf = (float)kk;
            const float zlim = 1.0f + kf;
            const float inv1pk = 1.0f / zlim;
            const float tcf = (float)tc, extf = (float)ext;
            const float inv_ns = 1.0f / (float)ns;
            const float* lrow = f.ldm32 + row;
            float depth = 0.0f;
            for (int is = 0; is < ns; ++is) {
                const float off = extf * (((float)is + 0.5f) * inv_ns - 0.5f);
                const float z = taylor_z<float>(tcf + off, cf);
                if (z < zlim) {
                    const float ip = interp_ld_f(z * inv1pk, inv_dg, lrow, f.ng);
                    float ap, kap;
                    ccia_kite<float>(1.0f, kf, z, ap, kap);
                    depth = fmaf(ip, ap, depth);
                }
            }
            const float scale = (float)(1.0 / ist) * inv_ns;
            ((float*)f.flux)[out0 + it] = 1.0f - depth * scale;
        } else {
            const double* lrow = f.ldm64 + row;
            double acc = 0.0;
            for (int is = 1; is <= ns; ++is) {
                const double off = ext * ((is - 0.5) / ns - 0.5);
                const double z = taylor_z<double>(tc + off, r.c);
                if (z < 1.0 + kk) {
                    const double ip = interp_ld(z / (1.0 + kk), f.dg, lrow, f.ng);
                    double ap, kap;
                    ccia_kite<double>(1.0, kk, z, ap, kap);
                    acc += (ist - ip * ap) / ist;
                } else {
                    acc += 1.0;  // ccia gives 0 and interp 0 for z >= 1+k: (ist - 0)/ist
                }
            }
            ((double*)f.flux)[out0 + it] = acc / ns;
        }
    }
}

}  // namespace rr

using namespace rr;

// ======================================================================================
// C-ABI
// ======================================================================================
extern "C" {

const char* rr_last_error(void) { return g_last_error.c_str(); }

int rr_abi_version(void) { return 1; }

int rr_z_grid(double zcut, int nin, int nedge, double* ze, double* zm, double* mu) {
    if (nin < 1 || nedge < 1 || !ze || !zm) return set_error(RR_EINVAL, "rr_z_grid: bad arguments");
    if (!(zcut > 0.0 && zcut < 1.0)) return set_error(RR_EINVAL, "rr_z_grid: zcut must be in (0, 1)");
    const int n = nin + nedge;
    const double mucut = std::sqrt(1.0 - zcut * zcut);
    const double dz = zcut / nin;
    const double dmu = mucut / nedge;
    for (int i = 0; i < n; ++i) { ze[i] = 0.0; zm[i] = 0.0; }
    for (int i = 0; i < nin - 1; ++i) ze[i] = (i + 1) * dz;
    for (int i = 0; i <= nedge; ++i) ze[n - 1 - i] = std::sqrt(1.0 - (i * dmu) * (i * dmu));
    for (int i = 0; i < n - 1; ++i) zm[i + 1] = 0.5 * (ze[i] + ze[i + 1]);
    if (mu)
        for (int i = 0; i < n; ++i) mu[i] = std::sqrt(1.0 - zm[i] * zm[i]);
    return RR_OK;
}

int rr_weights_3d(int nk, double kmin, double kmax, const double* ze, int nz, int ng, double* W, float* W32,
                  double* dk_out, double* dg_out, void* stream) {
    if (nk < 2 || ng < 2 || nz < 1 || !ze || !W) return set_error(RR_EINVAL, "rr_weights_3d: bad arguments");
    if (!(kmax > kmin) || !(kmin >= 0.0)) return set_error(RR_EINVAL, "rr_weights_3d: need 0 <= kmin < kmax");
    if ((long long)nk * ng * nz > (1LL << 31)) return set_error(RR_ERANGE, "rr_weights_3d: table too large");
    if (dk_out) *dk_out = (kmax - kmin) / nk;
    if (dg_out) {
        const double stop = 1.0 - 1e-7;
        const double g1 = (ng == 2) ? stop : 1.0 * (stop / (double)(ng - 1)) + 0.0;
        *dg_out = g1 - 0.0;
    }
    const int rows = nk * ng;
    hipLaunchKernelGGL(weights_3d_kernel, dim3((rows + 255) / 256), dim3(256), 0, (hipStream_t)stream, nk, kmin,
                       kmax, ze, nz, ng, W, W32);
    return check_launch("weights_3d_kernel");
}

int rr_ld_profiles(int law, const double* ldc, int npv, int npb, int nldc, const double* mu, int nmu, double* ldp,
                   double* istar, void* stream) {
    if (law < RR_LD_UNIFORM || law > RR_LD_POWER_2_PM) return set_error(RR_EINVAL, "rr_ld_profiles: unknown law %d", law);
    if (npv < 0 || npb < 1 || nmu < 1 || nldc < 0) return set_error(RR_EINVAL, "rr_ld_profiles: bad sizes");
    static const int min_coeffs[] = {0, 1, 2, 2, 4, 0, 2, 2, 2, 2, 2};
    if (nldc < min_coeffs[law])
        return set_error(RR_EINVAL, "rr_ld_profiles: law %d needs %d coefficients, got %d", law, min_coeffs[law], nldc);
    if (npv == 0) return RR_OK;
    if (!mu || !ldp || !istar || (nldc > 0 && !ldc)) return set_error(RR_EINVAL, "rr_ld_profiles: null pointer");
    const int n = npv * npb;
    hipLaunchKernelGGL(ld_profiles_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, law, ldc, npv,
                       npb, nldc, mu, nmu, ldp, istar);
    return check_launch("ld_profiles_kernel");
}

int rr_integrate_profiles(const double* mu, int nmu, const double* ldp, int npv, int npb, double* istar,
                          void* stream) {
    if (nmu < 2 || npv < 0 || npb < 1) return set_error(RR_EINVAL, "rr_integrate_profiles: bad sizes");
    if (npv == 0) return RR_OK;
    if (!mu || !ldp || !istar) return set_error(RR_EINVAL, "rr_integrate_profiles: null pointer");
    const int n = npv * npb;
    hipLaunchKernelGGL(integrate_profiles_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, mu, nmu,
                       ldp, n, istar);
    return check_launch("integrate_profiles_kernel");
}

int rr_orbit(int npv, const double* p, const double* a, const double* inc, const double* e, const double* w,
             const double* k0, double* coeffs, double* win, void* stream) {
    if (npv < 0) return set_error(RR_EINVAL, "rr_orbit: npv < 0");
    if (npv == 0) return RR_OK;
    if (!p || !a || !inc || !e || !w || !k0 || !coeffs || !win) return set_error(RR_EINVAL, "rr_orbit: null pointer");
    hipLaunchKernelGGL(orbit_kernel, dim3(npv), dim3(64), 0, (hipStream_t)stream, npv, p, a, inc, e, w, k0, coeffs,
                       win);
    return check_launch("orbit_kernel");
}

size_t rr_workspace_bytes(int npv, int npb, int ng, int nlc) {
    (void)nlc;
    if (npv < 0 || npb < 0 || ng < 0) return 0;
    size_t b = align_up((size_t)npv * sizeof(PvRec), 256);
    b += align_up((size_t)npv * npb * ng * sizeof(double), 256);
    b += align_up((size_t)npv * npb * ng * sizeof(float), 256);
    return b;
}

static int validate(const char* who, int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv,
                    size_t workspace_bytes, const void* workspace) {
    if (!tab || !lc || !pv) return set_error(RR_EINVAL, "%s: null descriptor", who);
    if (precision != RR_FP64 && precision != RR_MIXED) return set_error(RR_EINVAL, "%s: bad precision", who);
    const int npv = pv->npv, npt = lc->npt, npb = lc->npb, nlc = lc->nlc;
    if (npv < 0 || npt < 0 || nlc < 1 || npb < 1 || lc->nep < 1)
        return set_error(RR_EINVAL, "%s: bad sizes (npv=%d npt=%d nlc=%d npb=%d nep=%d)", who, npv, npt, nlc, npb,
                         lc->nep);
    if (pv->kcols != 1 && pv->kcols != npb)
        return set_error(RR_EINVAL, "Radius ratios should be given either as an [npv, 1] or [npv, npb] array.");
    if (pv->t0cols < 1) return set_error(RR_EINVAL, "%s: t0 needs at least one column", who);
    if (tab->nk < 2 || tab->ng < 2 || tab->nz < 1 || !tab->weights || !tab->ze)
        return set_error(RR_EINVAL, "%s: bad weight table", who);
    if ((long long)npv * npt > (1LL << 40) || (long long)npv * npb * tab->ng > (1LL << 31))
        return set_error(RR_ERANGE, "%s: problem too large", who);
    if (npv == 0 || npt == 0) return RR_OK;
    const size_t need = rr_workspace_bytes(npv, npb, tab->ng, nlc);
    if (!workspace || workspace_bytes < need)
        return set_error(RR_EINVAL, "%s: workspace of %zu bytes needed, got %zu", who, need, workspace_bytes);
    return RR_OK;
}

struct WsView {
    PvRec* rec;
    double* ldm64;
    float* ldm32;
};

static WsView carve(void* workspace, int npv, int npb, int ng) {
    char* ws = (char*)workspace;
    WsView v;
    v.rec = (PvRec*)ws;
    ws += align_up((size_t)npv * sizeof(PvRec), 256);
    v.ldm64 = (double*)ws;
    ws += align_up((size_t)npv * npb * ng * sizeof(double), 256);
    v.ldm32 = (float*)ws;
    return v;
}

int rr_setup(int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv, const double* ldp,
             void* workspace, size_t workspace_bytes, void* stream) {
    int rc = validate("rr_setup", precision, tab, lc, pv, workspace_bytes, workspace);
    if (rc) return rc;
    const int npv = pv->npv, npb = lc->npb;
    if (npv == 0 || lc->npt == 0) return RR_OK;
    if (!pv->k || !pv->p || !pv->a || !pv->inc || !pv->e || !pv->w || !ldp)
        return set_error(RR_EINVAL, "rr_setup: null parameter array");
    WsView v = carve(workspace, npv, npb, tab->ng);
    SetupArgs sa;
    sa.k = pv->k; sa.p = pv->p; sa.a = pv->a; sa.inc = pv->inc; sa.e = pv->e; sa.w = pv->w;
    sa.ldp = ldp; sa.W = tab->weights; sa.ze = tab->ze;
    sa.rec = v.rec; sa.ldm64 = v.ldm64; sa.ldm32 = precision == RR_MIXED ? v.ldm32 : nullptr;
    sa.npb = npb; sa.kcols = pv->kcols; sa.nk = tab->nk; sa.ng = tab->ng; sa.nz = tab->nz;
    sa.kmin = tab->kmin; sa.kmax = tab->kmax; sa.dk = tab->dk;
    hipLaunchKernelGGL(rr_setup_kernel, dim3(npv), dim3(64), 0, (hipStream_t)stream, sa);
    return check_launch("rr_setup_kernel");
}

int rr_flux(int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv, const double* istar,
            void* workspace, size_t workspace_bytes, void* flux, void* stream) {
    int rc = validate("rr_flux", precision, tab, lc, pv, workspace_bytes, workspace);
    if (rc) return rc;
    const int npv = pv->npv, npt = lc->npt, npb = lc->npb, nlc = lc->nlc;
    if (npv == 0 || npt == 0) return RR_OK;
    if (!lc->times || !lc->pbids || !lc->epids || !lc->nsamples || !lc->exptimes || (nlc > 1 && !lc->lcids))
        return set_error(RR_EINVAL, "rr_flux: null light-curve array");
    if (!pv->k || !pv->t0 || !istar || !flux) return set_error(RR_EINVAL, "rr_flux: null parameter array");
    WsView v = carve(workspace, npv, npb, tab->ng);
    FluxArgs fa;
    fa.times = lc->times; fa.lcids = lc->lcids; fa.pbids = lc->pbids; fa.epids = lc->epids;
    fa.nsamples = lc->nsamples; fa.exptimes = lc->exptimes;
    fa.k = pv->k; fa.t0 = pv->t0; fa.istar = istar;
    fa.rec = v.rec; fa.ldm64 = v.ldm64; fa.ldm32 = v.ldm32; fa.flux = flux;
    fa.npt = npt; fa.nlc = nlc; fa.npb = npb; fa.nep = lc->nep; fa.kcols = pv->kcols; fa.t0cols = pv->t0cols;
    fa.ng = tab->ng; fa.dg = tab->dg;
    fa.ntiles = (npt + kFluxTile - 1) / kFluxTile;
    const long long nblocks = (long long)fa.ntiles * npv;
    if (nblocks > 0x7fffffffLL) return set_error(RR_ERANGE, "rr_flux: grid too large");
    hipStream_t s = (hipStream_t)stream;
    const bool one = (nlc == 1);
    if (precision == RR_MIXED) {
        if (one) hipLaunchKernelGGL((rr_flux_kernel<true, true>), dim3((unsigned)nblocks), dim3(kFluxThreads), 0, s, fa);
        else hipLaunchKernelGGL((rr_flux_kernel<true, false>), dim3((unsigned)nblocks), dim3(kFluxThreads), 0, s, fa);
    } else {
        if (one) hipLaunchKernelGGL((rr_flux_kernel<false, true>), dim3((unsigned)nblocks), dim3(kFluxThreads), 0, s, fa);
        else hipLaunchKernelGGL((rr_flux_kernel<false, false>), dim3((unsigned)nblocks), dim3(kFluxThreads), 0, s, fa);
    }
    return check_launch("rr_flux_kernel");
}

int rr_evaluate(int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv, const double* ldp,
                const double* istar, void* workspace, size_t workspace_bytes, void* flux, void* stream) {
    int rc = rr_setup(precision, tab, lc, pv, ldp, workspace, workspace_bytes, stream);
    if (rc) return rc;
    return rr_flux(precision, tab, lc, pv, istar, workspace, workspace_bytes, flux, stream);
}

}  // extern "C"
