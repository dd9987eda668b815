// RoadRunner transit model on MI355X (gfx950): weight table, limb-darkening
// profiles, per-pv setup and the flux kernels, plus the C-ABI of include/rr_hip.h.
//
// Data layout in HBM (all row-major):
//   times[npt] f64, lcids[npt] i32, per-lc arrays [nlc]
//   k[npv, kcols], t0[npv, nep], p/a/inc/e/w[npv] f64, ldc[npv, npb, nldc] f64
//   weight table W[nk, nz, ng] f64 (+ f32 copy), g fastest
//   workspace: PvRec[npv] (128 B) | istar[npv, npb] f64 | ldm64[npv, npb, ng] | ldm32[npv, npb, ng]
//   flux[npv, npt] f64 (RR_FP64) or f32 (RR_MIXED)
//
// Setup is one launch with two block roles: orbit blocks (8 pv per 64-lane
// block, the 7 Kepler solves of a pv on 7 lanes) and limb-darkening blocks (one
// (pv, group of up to 8 passbands) per block: profiles at the mu nodes in LDS, disk
// integrals, and the table-lerp contraction with lanes over g reading the table
// coalesced, each table value serving the whole group).
//
// The flux kernel runs one wavefront per workgroup over a tile of 1024 times of
// one pv; each lane owns two runs of 8 consecutive times and issues all its time
// loads up front.  With contiguous light curves sorted in time, a lane classifies
// each run from the phases of its first and last time: runs entirely outside the
// transit window store 1.0 as vectors.  In-window points are compacted into a
// per-wave LDS queue (ballot + popcount, count kept in a register), and all 64
// lanes then share the supersampled geometry of the queued points, reading the
// limb-darkening row from LDS.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "rr_device.hpp"
#include "rr_internal.hpp"

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
// Weight table; common.py:188-223.  One lane per (k, g) row, the areas of a row
// accumulated left to right as in the reference; written as W[ik][iz][ig].
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
    const size_t base = (size_t)ik * nz * ng + ig;
    double a0 = cci_area(ze[0], k, b);
    double s = a0;
    W[base] = a0;
    for (int i = 1; i < nz; ++i) {
        const double a1 = cci_area(ze[i], k, b);
        const double wi = a1 - a0;
        W[base + (size_t)i * ng] = wi;
        a0 = a1;
        s += wi;
    }
    for (int i = 0; i < nz; ++i) {
        const double v = W[base + (size_t)i * ng] / s;
        W[base + (size_t)i * ng] = v;
        if (W32) W32[base + (size_t)i * ng] = (float)v;
    }
}

// ------------------------------------------------------------------------------------
// Limb-darkening profiles for the rr_ld_profiles entry point (one lane per value;
// lane iz == 0 also computes the disk integral).
// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) ld_profiles_kernel(int law, const double* __restrict__ ldc, int npv, int npb,
                                                          int nldc, const double* __restrict__ mu, int nmu,
                                                          double* __restrict__ ldp, double* __restrict__ istar) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)npv * npb * nmu) return;
    const long long j = idx / nmu;
    const int iz = (int)(idx - j * nmu);
    const double* c = ldc + j * nldc;
    ldp[idx] = ld_law(law, mu[iz], c, nldc);
    if (iz == 0) {
        double is;
        if (!ld_integral(law, c, is)) {
            double s = 0.0;
            for (int m = 1; m < 200; ++m) s += ld_trapz_segment(law, c, nldc, m);
            is = 2.0 * kPi * s;
        }
        istar[j] = is;
    }
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
// Trilinear interpolation of a profile grid over (teff, logg, metallicity) for the
// rr_ldtk_interpolate entry point; ldtkldm.py:5-57.  One lane per output value
// ldp[pv][pb][mu]: the lane finds its pv's cell and fractions (a few operations)
// and reads the cell's eight corners, which the lanes of one pv share through L2.
// ------------------------------------------------------------------------------------

// Cell index and fraction along one evenly spaced axis of n >= 2 nodes
// (ldtkldm.py:40-51).  The value is clamped to the axis, and the cell index to
// n - 2, so the last node is reached with a fraction of 1.  (The reference clamps
// the value to x0 + n dx and then indexes cell n - 1, which reads past its grid at
// and above the last node; inside the grid the two are the same.)
__device__ __forceinline__ void grid_cell(double x, double x0, double dx, int n, int& i, double& a) {
    x = fmin(fmax(x, x0), x0 + (n - 1) * dx);
    i = (int)floor((x - x0) / dx);
    i = min(max(i, 0), n - 2);
    a = (x - x0 - i * dx) / dx;
}

__global__ void __launch_bounds__(256) ldtk_interpolate_kernel(const double* __restrict__ grid, int nx, int ny,
                                                               int nz, int nq, double x0, double dx, double y0,
                                                               double dy, double z0, double dz,
                                                               const double* __restrict__ par, int par_stride,
                                                               int npv, double* __restrict__ ldp) {
#pragma clang fp contract(off)
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)npv * nq) return;
    const int pv = (int)(idx / nq), q = (int)(idx - (long long)pv * nq);
    const double x = par[(size_t)pv * par_stride], y = par[(size_t)pv * par_stride + 1],
                 z = par[(size_t)pv * par_stride + 2];
    if (isnan(x) || isnan(y) || isnan(z)) {
        ldp[idx] = NAN;
        return;
    }
    int ix, iy, iz;
    double ax, ay, az;
    grid_cell(x, x0, dx, nx, ix, ax);
    grid_cell(y, y0, dy, ny, iy, ay);
    grid_cell(z, z0, dz, nz, iz, az);
    // the corner weights and the sum in the order of ldtkldm.py:14-35
    const double rx = 1.0 - ax, ry = 1.0 - ay, rz = 1.0 - az;
    const double a1 = rx * ry * rz, a2 = ax * ry * rz, b1 = rx * ay * rz, b2 = rx * ry * az;
    const double c1 = ax * ry * az, c2 = rx * ay * az, d1 = ax * ay * rz, d2 = ax * ay * az;
    const size_t sz = (size_t)nq, sy = sz * nz, sx = sy * ny;
    const double* g = grid + (size_t)ix * sx + (size_t)iy * sy + (size_t)iz * sz + q;
    ldp[idx] = g[0] * a1 + g[sx] * a2 + g[sy] * b1 + g[sz] * b2 + g[sx + sz] * c1 + g[sy + sz] * c2 +
               g[sx + sy] * d1 + g[sx + sy + sz] * d2;
}

__global__ void __launch_bounds__(64) orbit_kernel(int npv, const double* __restrict__ P, const double* __restrict__ A,
                                                   const double* __restrict__ I, const double* __restrict__ E,
                                                   const double* __restrict__ Wp, const double* __restrict__ K0,
                                                   double* __restrict__ coeffs, double* __restrict__ win) {
    const int pv = blockIdx.x * kOrbitGroup + threadIdx.x / kOrbitGroup;
    const bool valid = pv < npv;
    const int q = valid ? pv : 0;
    __shared__ double sm[kOrbitLds];
    orbit_group(pv, valid, P[q], A[q], I[q], E[q], Wp[q], K0[q], coeffs + (size_t)q * 10, win + (size_t)q * 2, sm);
}

// ------------------------------------------------------------------------------------
// Per-pv setup of rr_full; model_full.py:39-70 and rrmodel.py:215-227.
// ------------------------------------------------------------------------------------
constexpr int kMaxNz = 1024;

struct SetupArgs {
    const double* k; const double* p; const double* a; const double* inc; const double* e; const double* w;
    int law, nldc;
    const double* ldc; const double* ldp_in; const double* istar_in;
    const double* W; const float* W32; const double* ze; const double* mu;
    PvRec* rec; double* istar; double* ldm64; float* ldm32;
    const double* t0; const int32_t* pbids; const int32_t* epids;   // light curve 0's context
    int npv, npb, kcols, t0cols, nk, ng, nz, n_orbit_blocks;
    int ld_lds;                       // doubles of the LD role's profile area (<= kLdLds)
    double kmin, kmax, dk;
};

// Passband and epoch index of light curve 0, clamped as the flux kernel clamps.
__device__ __forceinline__ int lc0_pb(const int32_t* pbids, int npb) { return min(max(pbids[0], 0), npb - 1); }
__device__ __forceinline__ int lc0_ep(const int32_t* epids, int nep) { return min(max(epids[0], 0), nep - 1); }

__device__ __forceinline__ void setup_orbit_role(const SetupArgs& s, int blk, double* sm) {
    const int pv = blk * kOrbitGroup + threadIdx.x / kOrbitGroup;
    const int j = threadIdx.x % kOrbitGroup;
    const bool inrange = pv < s.npv;
    const int q = inrange ? pv : 0;
    const double a = s.a[q], e = s.e[q];
    const bool valid = inrange && !(isnan(a) || a <= 1.0 || e < 0.0);
    const double k0 = s.k[(size_t)q * s.kcols];
    PvRec* rec = s.rec + q;
    orbit_group(pv, valid, s.p[q], a, s.inc[q], e, s.w[q], k0, rec->c, &rec->bt1, sm);
    if (inrange && j == 0) {
        rec->p = s.p[q];
        rec->inv_p = 1.0 / s.p[q];
        rec->k0 = k0;
        rec->good = valid ? 1 : 0;
        const int ipb = lc0_pb(s.pbids, s.npb);
        rec->t0_lc0 = s.t0[(size_t)q * s.t0cols + lc0_ep(s.epids, s.t0cols)];
        rec->k_lc0 = s.k[(size_t)q * s.kcols + (s.kcols == 1 ? 0 : ipb)];
    }
}

// One block per (pv, group of up to kLdGroup passbands): the profiles at the mu
// nodes, their disk integrals and the limb-darkening means.  The passbands of a
// group share each table read (every passband of a pv uses the rows of k0,
// model_full.py:47-51).
constexpr int kLdGroup = 8;
constexpr int kLdLds = 1024;  // doubles of profile storage per block (>= kMaxNz)

// A law linear in its coefficients with more passbands than basis functions is
// evaluated through its basis (ld_means_basis): one block takes up to
// kLdGroupBasis passbands, whose basis profiles [nb][nz], coefficients [gs][nb]
// and integrals [gs] share the kLdLds doubles.
constexpr int kLdGroupBasis = 64;

__host__ __device__ inline int ld_basis_group(int nz, int law, int nldc, int npb) {
    const int nb = ld_basis_count(law, nldc);
    if (nb == 0 || npb <= nb) return 0;
    const int room = kLdLds - nb * nz;
    const int gs = room > 0 ? room / (nb + 1) : 0;
    return gs < kLdGroup ? 0 : (gs < kLdGroupBasis ? gs : kLdGroupBasis);
}

__host__ __device__ inline int ld_group_size(int nz, int law, int nldc, int npb) {
    const int gb = ld_basis_group(nz, law, nldc, npb);
    if (gb > 0) return gb;
    const int g = kLdLds / (nz > 0 ? nz : 1);
    return g < kLdGroup ? (g < 1 ? 1 : g) : kLdGroup;
}

// Limb-darkening means of gn (<= G) passbands whose profiles are in s_l, lanes
// over g; each table value read serves all G passbands.
template <int G>
__device__ __forceinline__ void ld_means(const SetupArgs& s, int pv, size_t j0, int gn, bool in_table, int ik,
                                         double ak, double k0, const double* s_l) {
    const int lane = threadIdx.x;
    const int nz = s.nz, ng = s.ng;
    const bool mixed = s.ldm32 != nullptr;
    for (int ig = lane; ig < ng; ig += 64) {
        double v[G];
        if (in_table) {
            double d0[G], d1[G];
#pragma unroll
            for (int q = 0; q < G; ++q) d0[q] = d1[q] = 0.0;
            const size_t o0 = (size_t)ik * nz * ng + ig, o1 = o0 + (size_t)nz * ng;
            // table values in batches of 8 rows, all loads issued before use
            constexpr int kB = 8;
            for (int iz0 = 0; iz0 < nz; iz0 += kB) {
                double w0[kB], w1[kB];
#pragma unroll
                for (int u = 0; u < kB; ++u) {
                    const size_t iz = (size_t)min(iz0 + u, nz - 1);
                    if (mixed && s.W32) {
                        w0[u] = (double)s.W32[o0 + iz * ng];
                        w1[u] = (double)s.W32[o1 + iz * ng];
                    } else {
                        w0[u] = s.W[o0 + iz * ng];
                        w1[u] = s.W[o1 + iz * ng];
                    }
                }
#pragma unroll
                for (int u = 0; u < kB; ++u) {
                    if (iz0 + u < nz) {
                        if constexpr (G == 1) {
                            d0[0] += w0[u] * s_l[iz0 + u];
                            d1[0] += w1[u] * s_l[iz0 + u];
                        } else {
                            // several passbands: lerp the table value once, then one
                            // product per passband (the same sum within rounding)
                            const double wk = (1.0 - ak) * w0[u] + ak * w1[u];
#pragma unroll
                            for (int q = 0; q < G; ++q)
                                if (q < gn) d0[q] += wk * s_l[q * nz + iz0 + u];
                        }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < G; ++q) v[q] = G == 1 ? (1.0 - ak) * d0[q] + ak * d1[q] : d0[q];
        } else {
            // common.py:152-185: normalised annulus areas, recomputed in a second pass
            const double g = linspace_at(0.0, 1.0 - 1e-7, ng, ig);
            const double b = g * (1.0 + k0);
            double a0 = cci_area(s.ze[0], k0, b), sum = a0;
            for (int iz = 1; iz < nz; ++iz) {
                const double a1 = cci_area(s.ze[iz], k0, b);
                sum += a1 - a0;
                a0 = a1;
            }
            double b0 = cci_area(s.ze[0], k0, b);
#pragma unroll
            for (int q = 0; q < G; ++q) v[q] = (b0 / sum) * s_l[(q < gn ? q : 0) * nz];
            for (int iz = 1; iz < nz; ++iz) {
                const double b1 = cci_area(s.ze[iz], k0, b);
                const double wv = (b1 - b0) / sum;
#pragma unroll
                for (int q = 0; q < G; ++q)
                    if (q < gn) v[q] += wv * s_l[q * nz + iz];
                b0 = b1;
            }
        }
#pragma unroll
        for (int q = 0; q < G; ++q) {
            if (q < gn) {
                if (mixed) s.ldm32[(j0 + q) * ng + ig] = (float)v[q];
                else s.ldm64[(j0 + q) * ng + ig] = v[q];
            }
        }
    }
}

// Limb-darkening means of gn (<= G) passbands inside the table with 8-byte (FP32
// table) or 16-byte (FP64 table) loads: a lane owns two neighbouring g, so one
// round of 50 lanes covers ng = 100 and each table row batch is half the load
// instructions of ld_means.  Per value the same lerp and the same sum over the mu
// nodes in order as ld_means.  Needs an even ng and a table aligned to its load.
template <typename T, int G>
__device__ __forceinline__ void ld_means_pair(const SetupArgs& s, size_t j0, int gn, int ik, double ak,
                                              const double* s_l) {
    typedef T vec2 __attribute__((ext_vector_type(2)));
    constexpr int kB = sizeof(T) == 4 ? 8 : 4;
    const int lane = threadIdx.x;
    const int nz = s.nz, ng = s.ng, nv = ng / 2;
    const T* W = std::is_same<T, float>::value ? (const T*)s.W32 : (const T*)s.W;
    const size_t o0 = (size_t)ik * nz * ng, o1 = o0 + (size_t)nz * ng;
    const bool mixed = s.ldm32 != nullptr;
    for (int v = lane; v < nv; v += 64) {
        double d[G][2];
#pragma unroll
        for (int q = 0; q < G; ++q) d[q][0] = d[q][1] = 0.0;
        for (int iz0 = 0; iz0 < nz; iz0 += kB) {
            vec2 w0[kB], w1[kB];
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                const size_t r = (size_t)min(iz0 + u, nz - 1) * ng + 2 * (size_t)v;
                w0[u] = *(const vec2*)(W + o0 + r);
                w1[u] = *(const vec2*)(W + o1 + r);
            }
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                if (iz0 + u < nz) {
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const double wk = (1.0 - ak) * (double)w0[u][e] + ak * (double)w1[u][e];
#pragma unroll
                        for (int q = 0; q < G; ++q)
                            if (q < gn) d[q][e] += wk * s_l[q * nz + iz0 + u];
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < G; ++q) {
            if (q < gn) {
                const size_t o = (j0 + q) * ng + 2 * (size_t)v;
                if (mixed) {
                    s.ldm32[o] = (float)d[q][0];
                    s.ldm32[o + 1] = (float)d[q][1];
                } else {
                    s.ldm64[o] = d[q][0];
                    s.ldm64[o + 1] = d[q][1];
                }
            }
        }
    }
}

// One passband inside the table: each lane takes two g values per round, so twice
// the table loads are in flight per batch (the per-block latency bounds setup).
__device__ __forceinline__ void ld_means_single(const SetupArgs& s, size_t j, int ik, double ak, const double* s_l) {
    const int lane = threadIdx.x;
    const int nz = s.nz, ng = s.ng;
    const bool mixed = s.ldm32 != nullptr;
    constexpr int kB = 8;
    for (int ig0 = lane; ig0 < ng; ig0 += 128) {
        const int ig1 = min(ig0 + 64, ng - 1);
        const size_t a0 = (size_t)ik * nz * ng, a1 = a0 + (size_t)nz * ng;
        double d00 = 0.0, d01 = 0.0, d10 = 0.0, d11 = 0.0;  // [row][g]
        for (int iz0 = 0; iz0 < nz; iz0 += kB) {
            double w00[kB], w01[kB], w10[kB], w11[kB];
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                const size_t r = (size_t)min(iz0 + u, nz - 1) * ng;
                if (mixed && s.W32) {
                    w00[u] = (double)s.W32[a0 + r + ig0];
                    w01[u] = (double)s.W32[a0 + r + ig1];
                    w10[u] = (double)s.W32[a1 + r + ig0];
                    w11[u] = (double)s.W32[a1 + r + ig1];
                } else {
                    w00[u] = s.W[a0 + r + ig0];
                    w01[u] = s.W[a0 + r + ig1];
                    w10[u] = s.W[a1 + r + ig0];
                    w11[u] = s.W[a1 + r + ig1];
                }
            }
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                if (iz0 + u < nz) {
                    const double l = s_l[iz0 + u];
                    d00 += w00[u] * l;
                    d01 += w01[u] * l;
                    d10 += w10[u] * l;
                    d11 += w11[u] * l;
                }
            }
        }
        const double v0 = (1.0 - ak) * d00 + ak * d10, v1 = (1.0 - ak) * d01 + ak * d11;
        if (mixed) {
            s.ldm32[j * ng + ig0] = (float)v0;
            if (ig0 + 64 < ng) s.ldm32[j * ng + ig1] = (float)v1;
        } else {
            s.ldm64[j * ng + ig0] = v0;
            if (ig0 + 64 < ng) s.ldm64[j * ng + ig1] = v1;
        }
    }
}

// One passband inside the table with 16-byte table loads: lane task t covers kV
// consecutive g of one of the two rows (t < nv: row ik, else ik + 1) over all mu
// nodes, so each load instruction moves up to 1 KB; the per-row sums meet in LDS.
// Same sums in the same order as ld_means_single.  Needs ng % kV == 0, a 16-byte
// aligned table and nz + 2 ng doubles of `sm`.
template <typename T>
__device__ __forceinline__ void ld_means_vec(const SetupArgs& s, size_t j, int ik, double ak, const double* s_l,
                                             double* s_part) {
    typedef T vec __attribute__((ext_vector_type(16 / sizeof(T))));
    constexpr int kV = 16 / sizeof(T);
    constexpr int kB = 10;
    const int lane = threadIdx.x;
    const int nz = s.nz, ng = s.ng;
    const int nv = ng / kV;
    const T* W = std::is_same<T, float>::value ? (const T*)s.W32 : (const T*)s.W;
    const size_t base = (size_t)ik * nz * ng;
    for (int t = lane; t < 2 * nv; t += 64) {
        const int row = t >= nv ? 1 : 0;
        const int v = t - row * nv;
        const T* src = W + base + (size_t)row * nz * ng + (size_t)v * kV;
        double acc[kV];
#pragma unroll
        for (int e = 0; e < kV; ++e) acc[e] = 0.0;
        for (int iz0 = 0; iz0 < nz; iz0 += kB) {
            vec w[kB];
#pragma unroll
            for (int u = 0; u < kB; ++u) w[u] = *(const vec*)(src + (size_t)min(iz0 + u, nz - 1) * ng);
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                if (iz0 + u < nz) {
                    const double l = s_l[iz0 + u];
#pragma unroll
                    for (int e = 0; e < kV; ++e) acc[e] += (double)w[u][e] * l;
                }
            }
        }
#pragma unroll
        for (int e = 0; e < kV; ++e) s_part[row * ng + v * kV + e] = acc[e];
    }
    __syncthreads();
    for (int ig = lane; ig < ng; ig += 64) {
        const double v = (1.0 - ak) * s_part[ig] + ak * s_part[ng + ig];
        if (s.ldm32) s.ldm32[j * ng + ig] = (float)v;
        else s.ldm64[j * ng + ig] = v;
    }
}

// The basis means of ld_basis_role inside the table with a lane owning two
// neighbouring g (8-byte loads of the FP32 table, 16-byte of the FP64 one), then
// every passband's means from them: the same lerp and sums per value as the
// per-g loop, in half the dependent load rounds.  Needs an even ng and a table
// aligned to its load.
template <typename T>
__device__ __forceinline__ void basis_means_pair(const SetupArgs& s, int nb, int gn, size_t j0, int ik, double ak,
                                                 const double* s_B, const double* s_coef) {
    typedef T vec2 __attribute__((ext_vector_type(2)));
    constexpr int kB = sizeof(T) == 4 ? 8 : 4;
    const int lane = threadIdx.x;
    const int nz = s.nz, ng = s.ng, nv = ng / 2;
    const T* W = std::is_same<T, float>::value ? (const T*)s.W32 : (const T*)s.W;
    const size_t o0 = (size_t)ik * nz * ng, o1 = o0 + (size_t)nz * ng;
    const bool mixed = s.ldm32 != nullptr;
    for (int v = lane; v < nv; v += 64) {
        double m[kMaxBasis][2];
#pragma unroll
        for (int b = 0; b < kMaxBasis; ++b) m[b][0] = m[b][1] = 0.0;
        for (int iz0 = 0; iz0 < nz; iz0 += kB) {
            vec2 w0[kB], w1[kB];
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                const size_t r = (size_t)min(iz0 + u, nz - 1) * ng + 2 * (size_t)v;
                w0[u] = *(const vec2*)(W + o0 + r);
                w1[u] = *(const vec2*)(W + o1 + r);
            }
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                if (iz0 + u < nz) {
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const double wk = (1.0 - ak) * (double)w0[u][e] + ak * (double)w1[u][e];
#pragma unroll
                        for (int b = 0; b < kMaxBasis; ++b)
                            if (b < nb) m[b][e] += wk * s_B[b * nz + iz0 + u];
                    }
                }
            }
        }
        for (int q = 0; q < gn; ++q) {
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                double val = 0.0;
#pragma unroll
                for (int b = 0; b < kMaxBasis; ++b)
                    if (b < nb) val += s_coef[q * nb + b] * m[b][e];
                const size_t o = (j0 + q) * ng + 2 * (size_t)v + e;
                if (mixed) s.ldm32[o] = (float)val;
                else s.ldm64[o] = val;
            }
        }
    }
}

// The limb-darkening role through the law's basis (ld_basis_count > 0): basis
// profiles at the mu nodes, the basis means M_b(g) from the table rows (lerped in k
// once per value) or the exact 2-D weights, then for each of the gn passbands
// ldm(g) = sum_b coef_b M_b(g) and I* = its analytic integral or
// sum_b coef_b trapz(B_b).  The same sums as the per-passband form up to rounding.
__device__ __forceinline__ void ld_basis_role(const SetupArgs& s, int pv, int pb0, int gn, int gs, size_t j0, bool in_table, int ik,
                              double ak, double k0, double* sm) {
    const int lane = threadIdx.x;
    const int nz = s.nz, ng = s.ng, law = s.law;
    const int nb = ld_basis_count(law, s.nldc);
    double* s_B = sm;                    // [nb][nz]
    double* s_coef = sm + nb * nz;       // [gs][nb]
    double* s_istq = s_coef + gs * nb;   // [gs]
    for (int i = lane; i < nb * nz; i += 64) {
        const int b = i / nz;
        s_B[i] = ld_basis(law, b, s.mu[i - b * nz]);
    }
    for (int i = lane; i < gn * nb; i += 64) {
        const int q = i / nb;
        s_coef[i] = ld_basis_coef(law, i - q * nb, s.ldc + (j0 + q) * s.nldc);
    }
    // disk integrals: analytic per passband, or the basis trapezoids (wave-wide)
    const bool analytic = law == RR_LD_UNIFORM || law == RR_LD_LINEAR || law == RR_LD_QUADRATIC ||
                          law == RR_LD_QUADRATIC_TRI;
    double* s_y = sm + s.ld_lds + kLdGroup;  // trapezoid nodes
    double tb[kMaxBasis];
#pragma unroll
    for (int b = 0; b < kMaxBasis; ++b) tb[b] = (!analytic && b < nb) ? ld_basis_trapz_wave(law, b, s_y) : 0.0;
    __syncthreads();
    for (int q = lane; q < gn; q += 64) {
        double ist;
        if (analytic) {
            ld_integral(law, s.ldc + (j0 + q) * s.nldc, ist);
        } else {
            ist = 0.0;
#pragma unroll
            for (int b = 0; b < kMaxBasis; ++b)
                if (b < nb) ist += s_coef[q * nb + b] * tb[b];
        }
        s_istq[q] = ist;
        s.istar[j0 + q] = ist;
    }
    if (lane == 0 && pb0 == 0)
        s.rec[pv].ldok = isnan(ld_law(law, s.mu[0], s.ldc + j0 * s.nldc, s.nldc)) ? 0 : 1;
    __syncthreads();
    {
        const int q0 = lc0_pb(s.pbids, s.npb) - pb0;
        if (lane == 0 && q0 >= 0 && q0 < gn) s.rec[pv].ist_lc0 = s_istq[q0];
    }
    // basis means, lanes over g
    const bool mixed = s.ldm32 != nullptr;
    const void* wt = mixed ? (const void*)s.W32 : (const void*)s.W;
    if (in_table && wt != nullptr && ng % 2 == 0 && ((uintptr_t)wt & (mixed ? 7 : 15)) == 0) {
        if (mixed) basis_means_pair<float>(s, nb, gn, j0, ik, ak, s_B, s_coef);
        else basis_means_pair<double>(s, nb, gn, j0, ik, ak, s_B, s_coef);
        return;
    }
    for (int ig = lane; ig < ng; ig += 64) {
        double m[kMaxBasis];
#pragma unroll
        for (int b = 0; b < kMaxBasis; ++b) m[b] = 0.0;
        if (in_table) {
            const size_t o0 = (size_t)ik * nz * ng + ig, o1 = o0 + (size_t)nz * ng;
            constexpr int kB = 8;
            for (int iz0 = 0; iz0 < nz; iz0 += kB) {
                double w0[kB], w1[kB];
#pragma unroll
                for (int u = 0; u < kB; ++u) {
                    const size_t r = (size_t)min(iz0 + u, nz - 1) * ng;
                    if (mixed && s.W32) {
                        w0[u] = (double)s.W32[o0 + r];
                        w1[u] = (double)s.W32[o1 + r];
                    } else {
                        w0[u] = s.W[o0 + r];
                        w1[u] = s.W[o1 + r];
                    }
                }
#pragma unroll
                for (int u = 0; u < kB; ++u) {
                    if (iz0 + u < nz) {
                        const double wk = (1.0 - ak) * w0[u] + ak * w1[u];
#pragma unroll
                        for (int b = 0; b < kMaxBasis; ++b)
                            if (b < nb) m[b] += wk * s_B[b * nz + iz0 + u];
                    }
                }
            }
        } else {
            // common.py:152-185: normalised annulus areas
            const double g = linspace_at(0.0, 1.0 - 1e-7, ng, ig);
            const double bb = g * (1.0 + k0);
            double a0 = cci_area(s.ze[0], k0, bb), sum = a0;
            for (int iz = 1; iz < nz; ++iz) {
                const double a1 = cci_area(s.ze[iz], k0, bb);
                sum += a1 - a0;
                a0 = a1;
            }
            double b0 = cci_area(s.ze[0], k0, bb);
#pragma unroll
            for (int b = 0; b < kMaxBasis; ++b) m[b] = b < nb ? (b0 / sum) * s_B[b * nz] : 0.0;
            for (int iz = 1; iz < nz; ++iz) {
                const double b1 = cci_area(s.ze[iz], k0, bb);
                const double wv = (b1 - b0) / sum;
#pragma unroll
                for (int b = 0; b < kMaxBasis; ++b)
                    if (b < nb) m[b] += wv * s_B[b * nz + iz];
                b0 = b1;
            }
        }
        // every passband of the group from the basis means
        for (int q = 0; q < gn; ++q) {
            double v = 0.0;
#pragma unroll
            for (int b = 0; b < kMaxBasis; ++b)
                if (b < nb) v += s_coef[q * nb + b] * m[b];
            const size_t o = (j0 + q) * ng + ig;
            if (mixed) s.ldm32[o] = (float)v;
            else s.ldm64[o] = v;
        }
    }
}

// kMulti = false serves launches whose requests all have one passband (every group
// holds one); it leaves out the group and basis forms, whose registers would make
// the one-passband blocks spill to scratch.
template <bool kMulti>
__device__ __forceinline__ void setup_ld_role(const SetupArgs& s, int pv, int grp, double* sm) {
    double* s_l = sm;                // s.ld_lds
    double* s_ist = sm + s.ld_lds;   // kLdGroup
    const int lane = threadIdx.x;
    const int nz = s.nz, ng = s.ng;
    // one passband: a single group of one, with no group arithmetic on the path
    const int gs = kMulti ? ld_group_size(nz, s.law, s.nldc, s.npb) : 1;
    const int pb0 = kMulti ? grp * gs : 0;
    const int gn = kMulti ? min(gs, s.npb - pb0) : 1;
    // k0 picks the table rows; its load is issued here and first waited on after
    // the profiles and their integrals, which do not need it, so a cold start's
    // memory latency overlaps that work.
    const double k0 = s.k[(size_t)pv * s.kcols];
    int ik = 0;
    double ak = 0.0;
    auto table_pos = [&]() {
        const bool in = s.W != nullptr && s.kmin <= k0 && k0 <= s.kmax;
        if (in) {
            ik = (int)floor((k0 - s.kmin) / s.dk);
            ik = ik > s.nk - 2 ? s.nk - 2 : ik;
            ak = (k0 - s.kmin - ik * s.dk) / s.dk;
        }
        return in;
    };
    const size_t j0 = (size_t)pv * s.npb + pb0;
    if (kMulti && ld_basis_group(nz, s.law, s.nldc, s.npb) > 0) {
        const bool in_table = table_pos();
        ld_basis_role(s, pv, pb0, gn, gs, j0, in_table, ik, ak, k0, sm);
        return;
    }
    // profiles at the mu nodes, lanes over (passband, node), and their disk integrals
    const bool tab = s.law == RR_LD_TABULATED;
    for (int i = lane; i < gn * nz; i += 64) {
        const int q = kMulti ? i / nz : 0, iz = i - q * nz;
        const size_t j = j0 + q;
        s_l[i] = tab ? s.ldp_in[j * nz + iz] : ld_law(s.law, s.mu[iz], s.ldc + j * s.nldc, s.nldc);
    }
    for (int q = 0; q < gn; ++q) {
        const size_t j = j0 + q;
        double ist;
        if (tab) {
            ist = s.istar_in[j];
        } else {
            const double* c = s.ldc + j * s.nldc;
            if (!ld_integral(s.law, c, ist)) ist = ld_trapz_wave(s.law, c, s.nldc, s_ist + kLdGroup);
        }
        if (lane == 0) s_ist[q] = ist;
    }
    __syncthreads();
    if (lane < gn) s.istar[j0 + lane] = s_ist[lane];
    if (lane == 0 && pb0 == 0) s.rec[pv].ldok = isnan(s_l[0]) ? 0 : 1;
    {
        const int q0 = lc0_pb(s.pbids, s.npb) - pb0;
        if (lane == 0 && q0 >= 0 && q0 < gn) s.rec[pv].ist_lc0 = s_ist[q0];
    }
    // limb-darkening means: lerp in k of two table rows, or exact 2-D weights
    // when k0 is outside [kmin, kmax]; model_full.py:47-55
    const bool in_table = table_pos();
    const bool mixed = s.ldm32 != nullptr;
    const void* wt = mixed ? (const void*)s.W32 : (const void*)s.W;
    const int kv = mixed ? 4 : 2;
    const bool vec_ok = wt != nullptr && ng % kv == 0 && ((uintptr_t)wt & 15) == 0 && nz + 2 * ng <= kLdLds;
    if (gn == 1 && in_table && vec_ok) {
        if (mixed) ld_means_vec<float>(s, j0, ik, ak, s_l, s_l + nz);
        else ld_means_vec<double>(s, j0, ik, ak, s_l, s_l + nz);
    } else if (gn == 1 && in_table) ld_means_single(s, j0, ik, ak, s_l);
    else if (gn == 1) ld_means<1>(s, pv, j0, 1, in_table, ik, ak, k0, s_l);
    else if (kMulti && in_table && wt != nullptr && ng % 2 == 0 && ((uintptr_t)wt & (mixed ? 7 : 15)) == 0) {
        if (mixed) ld_means_pair<float, kLdGroup>(s, j0, gn, ik, ak, s_l);
        else ld_means_pair<double, kLdGroup>(s, j0, gn, ik, ak, s_l);
    } else if (kMulti) ld_means<kLdGroup>(s, pv, j0, gn, in_table, ik, ak, k0, s_l);
}

// Doubles of the LD role's profile area for one launch: the group's profiles
// (gs x nz), the 16-byte path's two row sums (nz + 2 ng, taken by any group of one
// passband), or the basis form's profiles, coefficients and integrals.
__host__ __device__ inline int ld_lds_doubles(int nz, int ng, int law, int nldc, int npb) {
    const int gs = ld_group_size(nz, law, nldc, npb);
    int n = gs * nz;
    if (nz + 2 * ng <= kLdLds && nz + 2 * ng > n) n = nz + 2 * ng;
    if (ld_basis_group(nz, law, nldc, npb) > 0) {
        const int nb = ld_basis_count(law, nldc);
        const int m = nb * nz + gs * nb + gs;
        n = m > n ? m : n;
    }
    return n < kLdLds ? n : kLdLds;
}

// Both roles share one dynamic LDS buffer of setup_lds_bytes: a block holds only
// the larger of the two, sized to the launch (3.6 KB for one passband and the
// default grids), so LDS does not cap the number of resident setup blocks.  The
// LD role's layout: profile area [ld_lds], disk integrals [kLdGroup], trapezoid
// nodes [kTrapzNodes].
static size_t setup_lds_bytes(const SetupArgs& s) {
    const int ld = s.ld_lds + kLdGroup + kTrapzNodes;
    const int n = ld > kOrbitLds ? ld : kOrbitLds;
    return sizeof(double) * (size_t)n;
}

template <bool kMulti>
__device__ __forceinline__ void setup_body(const SetupArgs& s, int blk) {
    extern __shared__ double sm[];
    if (blk < s.n_orbit_blocks) {
        setup_orbit_role(s, blk, sm);
    } else {
        const int b = blk - s.n_orbit_blocks;
        if (kMulti) {
            const int gsz = ld_group_size(s.nz, s.law, s.nldc, s.npb);
            const int ngrp = (s.npb + gsz - 1) / gsz;
            const int pv = b / ngrp;
            setup_ld_role<kMulti>(s, pv, b - pv * ngrp, sm);
        } else {
            setup_ld_role<kMulti>(s, b, 0, sm);  // one group per pv
        }
    }
}

// Five waves per SIMD (96 VGPRs): 20 blocks fit a CU when LDS does not bind, so the
// 4096 + 512 blocks of a 4096-pv setup are resident at once on 256 CUs.
template <bool kMulti>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) rr_setup_kernel(SetupArgs s) {
    setup_body<kMulti>(s, blockIdx.x);
}

// Several requests in one launch: item i owns blocks [start[i], start[i+1]).  The
// item index is wave-uniform, so its arguments are scalar loads.
constexpr int kMaxBatch = 16;

struct SetupBatch {
    int n;
    int start[kMaxBatch + 1];
    SetupArgs a[kMaxBatch];
};

__device__ __forceinline__ int batch_item(const int* start, int n, int bid) {
    int i = 0;
    while (i + 1 < n && bid >= start[i + 1]) ++i;
    return i;
}

template <bool kMulti>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) rr_setup_batch_kernel(SetupBatch b) {
    const int i = batch_item(b.start, b.n, blockIdx.x);
    setup_body<kMulti>(b.a[i], blockIdx.x - b.start[i]);
}

// The same with the batch in device memory, for launches that are captured once
// and replayed with new contents: the grid may exceed the batch's blocks
// (start[n]); the blocks past them exit at once.
template <bool kMulti>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5)))
rr_setup_batch_dev_kernel(const SetupBatch* __restrict__ b) {
    const int n = b->n;
    if (n == 0 || (int)blockIdx.x >= b->start[n]) return;
    const int i = batch_item(b->start, n, blockIdx.x);
    if (!kMulti && b->a[i].npb > 1) return;   // a launch shape that does not fit its batch (caller error)
    setup_body<kMulti>(b->a[i], blockIdx.x - b->start[i]);
}

// ------------------------------------------------------------------------------------
// Flux; model_full.py:75-99.
//
// One wavefront per workgroup covers a tile of 1024 times of one pv.  Lane l owns
// two runs of 8 consecutive times (run step m of the wave covers 512 contiguous
// times) and issues its time loads up front.
//
// Phase A: the FP64 epoch reduction classifies points against the transit window.
// With contiguous light curves sorted in time a lane classifies a whole run from the
// phases of its first and last time; when the whole wave's run step is outside
// the window (the common case) the wave stores 1.0 cooperatively, each store
// instruction covering 1 KB of contiguous memory.  In-window points go to an LDS
// queue (ballot + popcount; the count is wave-uniform and kept in a register).
//
// Phase B: with one light curve and one nsamples, work items are (point, sample)
// pairs in sample order, so a round of 64 lanes covers ~64/nss neighbouring
// points and the branch mix (outside / inside / limb) stays coherent; per-point
// sums are formed in sample order from LDS, so results do not depend on the
// schedule.  Otherwise one lane takes one point.  Waves in phase A (HBM-bound)
// and waves in phase B (VALU-bound) share each CU.
// ------------------------------------------------------------------------------------
struct FluxArgs {
    // the tile test's arrays first: one 64-byte kernel-argument load holds them
    const double* times; const PvRec* rec; void* flux; const double* exptimes;
    const int32_t* pbids; const int32_t* nsamples; const int32_t* lcids; const int32_t* epids;
    const double* k; const double* t0; const double* istar;
    const double* ldm64; const float* ldm32;
    int npt, nlc, npb, kcols, t0cols, ng, ntiles;
    int uniform_ns;                   // the common nsamples of all light curves, or 0
    int chunk_pts;                    // points per phase-B chunk (uniform_ns > 0)
    int row_lds;                      // stage the limb-darkening row in LDS (one light curve)
    int one_window;                   // several light curves, times sorted across them, one window
    int npv;                          // pvs of the launch (or of the batch request)
    int k_grp, k_last;                // the shifts of the two division multipliers below
    double dg;
    unsigned long long m_grp;         // n / (kTileGroup * ntiles) = (n * m_grp) >> k_grp
    unsigned long long m_last;        // n / (size of the last pv group) = (n * m_last) >> k_last
};

constexpr int kThreads = 64;
constexpr int kRun = 8;
constexpr int kRuns = 2;
constexpr int kTile = kThreads * kRun * kRuns;       // 1024 times per workgroup
constexpr int kMaxItems = 1024;                       // (point, sample) items per LDS chunk
constexpr int kMaxChunk = 64;
constexpr int kPartQ = 2 * kThreads;                  // queued partial-overlap samples

// Bytes of dynamic LDS the flux kernel uses for these parameters.
constexpr int kCtxWords = 5;  // per-point context of a multi-light-curve chunk (FP32)


// Epoch and phase of time t (the division-free form of model_full.py:88-89).
__device__ __forceinline__ double phase_of(double t, double t0, double p, double inv_p, double& epoch) {
#pragma clang fp contract(off)
    const double u = t - t0 + 0.5 * p;
    double ep = floor(u * inv_p);
    const double r = u - ep * p;
    ep += (r < 0.0) ? -1.0 : ((r >= p) ? 1.0 : 0.0);
    epoch = ep;
    return t - (t0 + ep * p);
}

// t0 + epoch * p, the subtrahend of model_full.py:89, rounded as the reference does.
__device__ __forceinline__ double epoch_base(double t0, double epoch, double p) {
#pragma clang fp contract(off)
    return t0 + epoch * p;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// out[i0 .. i0+n) = v by the 64 lanes of a wave: 16-byte non-temporal vectors,
// consecutive lanes on consecutive vectors, so each store instruction covers 1 KB
// of contiguous memory.
template <typename T>
__device__ __forceinline__ void wave_fill(T* out, int i0, int n, T v, int wl) {
    constexpr int kVec = 16 / sizeof(T);
    if (n <= 0) return;
    T* base = out + i0;
    int head = (int)(((16 - ((uintptr_t)base & 15)) & 15) / sizeof(T));
    if (((uintptr_t)base % sizeof(T)) != 0 || head > n) head = n;
    for (int i = wl; i < head; i += 64) base[i] = v;
    T* vb = base + head;
    const int nv = (n - head) / kVec;
    if constexpr (std::is_same<T, float>::value) {
        const f32x4 vv = {v, v, v, v};
        for (int i = wl; i < nv; i += 64) __builtin_nontemporal_store(vv, (f32x4*)vb + i);
    } else {
        const f64x2 vv = {v, v};
        for (int i = wl; i < nv; i += 64) __builtin_nontemporal_store(vv, (f64x2*)vb + i);
    }
    for (int i = head + nv * kVec + wl; i < n; i += 64) base[i] = v;
}

// The n (<= kRun) times of a run starting at i0 (a multiple of kRun) into
// registers: four 16-byte loads for a full run of a 16-byte aligned array.
__device__ __forceinline__ void load_run(const double* times, int i0, int n, double* tv) {
    if (n == kRun && (((uintptr_t)times) & 15) == 0) {
        const f64x2* v = (const f64x2*)(times + i0);
#pragma unroll
        for (int q = 0; q < kRun / 2; ++q) {
            const f64x2 x = v[q];
            tv[2 * q] = x[0];
            tv[2 * q + 1] = x[1];
        }
    } else {
#pragma unroll
        for (int j = 0; j < kRun; ++j) tv[j] = j < n ? times[i0 + j] : 0.0;
    }
}

// Per-light-curve quantities of one point (model_full.py:84-86, 93-98).
struct PointCtx {
    int ipb, ns;
    double ext, kk, ist, t0v, lo, hi;
};

// A tile of sorted light curves that spans several light curves (at most
// kTileLcs) stages their contexts in LDS: one dependent-load chain for all.  The
// table lives in the chunk-phase area, which only the one-light-curve loops use.
constexpr int kTileLcs = 8;
static_assert(kTileLcs * sizeof(PointCtx) <= kMaxChunk * sizeof(double), "light-curve table");

__device__ __forceinline__ PointCtx point_ctx(const FluxArgs& f, const PvRec& r, int pv, int ilc) {
    PointCtx c;
    c.ipb = min(max(f.pbids[ilc], 0), f.npb - 1);
    c.ns = f.nsamples[ilc];
    c.ext = f.exptimes[ilc];
    c.kk = f.k[(size_t)pv * f.kcols + (f.kcols == 1 ? 0 : c.ipb)];
    c.ist = f.istar[(size_t)pv * f.npb + c.ipb];
    c.t0v = f.t0[(size_t)pv * f.t0cols + min(max(f.epids[ilc], 0), f.t0cols - 1)];
    const double margin = 0.003 + c.ext;
    c.lo = r.bt1 - margin;
    c.hi = r.bt4 + margin;
    return c;
}

// Bytes of dynamic LDS the flux kernel uses for these parameters.
static inline size_t flux_lds_bytes(bool mixed, int ng, int chunk_pts, int uniform_ns, bool row_lds, bool multi_lc) {
    const size_t real = mixed ? 4 : 8;
    size_t b = (size_t)kTile * sizeof(uint16_t);                         // queue
    b += (size_t)kMaxChunk * sizeof(double);                             // phases of a chunk
    if (uniform_ns > 0) b += (size_t)chunk_pts * uniform_ns * real;      // item results
    if (row_lds) b += (size_t)ng * real;                                 // limb-darkening row
    if (mixed && uniform_ns > 0 && multi_lc) b += (size_t)kCtxWords * kMaxChunk * 4;  // per-point context
    if (uniform_ns > 0) b += (size_t)kPartQ * (real + sizeof(uint16_t));                 // partial queue
    return (b + 15) / 16 * 16;
}

// Light curve 0's context from the per-pv record (one light curve: no dependent loads).
__device__ __forceinline__ PointCtx point_ctx0(const FluxArgs& f, const PvRec& r) {
    PointCtx c;
    c.ipb = min(max(f.pbids[0], 0), f.npb - 1);
    c.ns = f.nsamples[0];
    c.ext = f.exptimes[0];
    c.kk = r.k_lc0;
    c.ist = r.ist_lc0;
    c.t0v = r.t0_lc0;
    const double margin = 0.003 + c.ext;
    c.lo = r.bt1 - margin;
    c.hi = r.bt4 + margin;
    return c;
}

// Contribution of sample `is` (0-based) of a point at phase tc: FP64 gives the
// reference's (istar - I*A)/istar, FP32 gives the depth I*A (the flux is
// 1 - mean(depth)/istar).
__device__ __forceinline__ double sample_f64(const double* cc, double dg, int ng, const PointCtx& c,
                                             const double* lrow, double tc, int is) {
    const double off = c.ext * ((is + 1 - 0.5) / c.ns - 0.5);
    const double z = taylor_z<double>(tc + off, cc);
    if (!(z < 1.0 + c.kk)) return 1.0;  // ccia gives 0 and interp 0 for z >= 1+k: (ist - 0)/ist
    const double ip = interp_ld(z / (1.0 + c.kk), dg, lrow, ng);
    double ap, kap;
    ccia_kite<double>(1.0, c.kk, z, ap, kap);
    return (c.ist - ip * ap) / c.ist;
}

__device__ __forceinline__ float sample_f32(const float* cf, float kf, float inv_dg, float tcf, float extf,
                                            float inv_ns, const float* lrow, int ng, int is) {
    const float zlim = 1.0f + kf;
    const float off = extf * (((float)is + 0.5f) * inv_ns - 0.5f);
    const float z = taylor_z<float>(tcf + off, cf);
    if (!(z < zlim)) return 0.0f;
    const float ip = interp_ld_f(z * (1.0f / zlim), inv_dg, lrow, ng);
    float ap, kap;
    ccia_kite<float>(1.0f, kf, z, ap, kap);
    return ip * ap;
}

// sample_f64 in two passes for the item loop with a queue of partial overlaps:
// returns the sample's 1 - I*A/istar unless the planet straddles the limb,
// |1 - k| < z < 1 + k, where it returns I and sets `part`; sample_f64_finish then
// completes it.  The divisions of sample_f64 (by ns, 1 + k, dg and istar) are
// products with the light curve's reciprocals (RecipF64): each changes a result
// by an ulp or so, and they were the long-latency links of the sample's chain.
// Reciprocals of one light curve's per-sample divisors, for the FP64 item loop: the
// divisions become products (a change of the last bit at most).
struct RecipF64 {
    double inv_1pk, inv_ist, inv_dg, inv_ns;
};

// interp_ld with g / dg as g * (1 / dg)
__device__ __forceinline__ double interp_ld_r(double g, double dg, double inv_dg, const double* __restrict__ row, int ng) {
    if (g < 0.0) return NAN;
    if (g > 1.0) return 0.0;
    int i = (int)floor(g * inv_dg);
    i = i > ng - 2 ? ng - 2 : (i < 0 ? 0 : i);
    const double a = (g - i * dg) * inv_dg;
    return (1.0 - a) * row[i] + a * row[i + 1];
}

__device__ __forceinline__ double sample_f64_pass1(const double* cc, double dg, int ng, const PointCtx& c,
                                                   const RecipF64& rc, const double* lrow, double tc, int is,
                                                   double& z, bool& part) {
    const double off = c.ext * ((is + 0.5) * rc.inv_ns - 0.5);
    z = taylor_z<double>(tc + off, cc);
    part = false;
    if (!(z < 1.0 + c.kk)) return 1.0;
    const double ip = interp_ld_r(z * rc.inv_1pk, dg, rc.inv_dg, lrow, ng);
    if (d_abs(1.0 - c.kk) < z) {
        part = true;
        return ip;
    }
    double ap, kap;
    ccia_kite<double>(1.0, c.kk, z, ap, kap);
    return 1.0 - (ip * ap) * rc.inv_ist;
}

__device__ __forceinline__ double sample_f64_finish(double ip, const PointCtx& c, const RecipF64& rc, double z) {
    double ap, kap;
    ccia_kite<double>(1.0, c.kk, z, ap, kap);
    return 1.0 - (ip * ap) * rc.inv_ist;
}

// Lens (kite) area of the unit disk and a planet of radius kf < 1 at separation z,
// 1 - kf < z < 1 + kf: the sides (1, k, z) sorted descending are three selects
// (common.py:5-33,52-73).
__device__ __forceinline__ float kite_small_f32(float kf, float z) {
    const float xs = d_max(1.0f, z), ys = d_max(kf, d_min(1.0f, z)), zs = d_min(kf, z);
    const float sq = d_sqrt((xs + (ys + zs)) * (zs - (xs - ys)) * (zs + (xs - ys)) * (xs + (ys - zs)));
    const float k0 = atan2_ypos_f(sq, (kf - 1.0f) * (kf + 1.0f) + z * z);
    const float k1 = atan2_ypos_f(sq, (1.0f - kf) * (1.0f + kf) + z * z);
    return k1 + kf * kf * k0 - 0.5f * sq;
}

// One sample of the FP32 item loops for a planet smaller than the star (kf < 1),
// where z < 1 + k gives g < 1 (no g > 1 branch).  Returns the depth I * A of a
// sample outside the stellar disk (0) or fully inside it (I pi k^2); for a partial
// overlap, 1 - k < z < 1 + k, it returns the limb-darkening mean I and sets `part`,
// and the caller multiplies by kite_small_f32(kf, z) later.
__device__ __forceinline__ float sample_small_f32_pass1(const float* cf, float kf, float inv_dg, float tz,
                                                        const float* lrow, int ng, float& z, bool& part) {
    const float zlim = 1.0f + kf;
    z = taylor_z<float>(tz, cf);
    part = false;
    if (!(z < zlim)) return 0.0f;
    // limb-darkening mean at g = z / (1 + k) in [0, 1): common.py:225-233
    const float x = z * (1.0f / zlim) * inv_dg;
    const int i = min((int)x, ng - 2);
    const float a = x - (float)i;
    const float l0 = lrow[i], l1 = lrow[i + 1];
    const float ip = fmaf(a, l1 - l0, l0);
    if (z <= 1.0f - kf) return ip * (kPiF * (kf * kf));  // planet fully inside the disk
    part = true;
    return ip;
}

// Partial-overlap samples of an item loop, queued in LDS (item index and z) and
// finished 64 at a time, so that the lens area runs on full waves instead of on
// the few lanes of each round whose sample straddles the limb.
template <typename T>
__device__ __forceinline__ int queue_partial(bool part, int i, T z, int pq, uint16_t* s_pi, T* s_pz, int wl) {
    const unsigned long long mk = __ballot(part);
    if (part) {
        const int slot = pq + __popcll(mk & ((1ull << wl) - 1ull));
        s_pi[slot] = (uint16_t)i;
        s_pz[slot] = z;
    }
    return pq + __popcll(mk);
}

// A tile whose first and last points (phases tca <= tcb of epochs ea <= eb) lie in
// one sorted light curve holds no point of the window [lo, hi] when both lie before
// or after the window of one epoch, or after one window and before the next.
__device__ __forceinline__ bool tile_outside(double tca, double tcb, double ea, double eb, const PointCtx& c) {
    return (ea == eb) ? (tcb < c.lo || tca > c.hi) : ((eb == ea + 1.0) ? (tca > c.hi && tcb < c.lo) : false);
}

// In-window index ranges of a tile that lies in one light curve sorted in time.
// For the window of epoch e, tc = t - (t0 + e p) is monotone in the point index, so
// the window is the index range [first tc >= lo, first tc > hi).  The two edges of
// each of the (at most two) epochs the tile spans are found by a 64-way search:
// one time per lane at 16-point spacing brackets every edge, then 16 lanes per edge
// test the 15 points of its bracket.  Every point gets the classification
// model_full.py:89-91 gives it: near a window, a point's own epoch is the window's
// (windows narrower than half a period).  Returns false, leaving the run
// classification to decide, when the tile spans more than two epochs or the window
// is that wide.
__device__ __forceinline__ bool window_edges(const double* times, int tbase, int cnt, const PointCtx& c, double p,
                                             double ea, double eb, int wl, int (&edge)[4]) {
    if (eb - ea > 1.0 || !(c.hi - c.lo < 0.5 * p)) return false;
    const int nedge = ea == eb ? 2 : 4;
    const double base0 = epoch_base(c.t0v, ea, p), base1 = epoch_base(c.t0v, eb, p);
    // edge k: k even, first tc >= lo; k odd, first tc > hi; epoch ea for k < 2
    auto pred = [&](int k, double t) -> bool {
        const double tc = t - (k < 2 ? base0 : base1);
        return (k & 1) ? (tc > c.hi) : (tc >= c.lo);
    };
    const int s = 16 * wl;
    const double ts = s < cnt ? times[tbase + s] : 0.0;
    int lo_i[4], hi_i[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        edge[k] = cnt;
        lo_i[k] = -1;
        hi_i[k] = cnt;
        if (k < nedge) {
            const unsigned long long mk = __ballot(s >= cnt || pred(k, ts));
            const int fl = mk ? __ffsll((long long)mk) - 1 : 64;  // first sample with the edge passed
            if (fl == 0) {
                edge[k] = 0;
            } else {
                lo_i[k] = 16 * (fl - 1);   // edge in (lo_i, hi_i]
                hi_i[k] = min(16 * fl, cnt);
            }
        }
    }
    // 16 lanes per edge: candidates lo_i + 1 .. lo_i + 15
    const int k2 = wl >> 4, j = wl & 15;
    int lo = lo_i[0], hi = hi_i[0];
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        lo = k2 == k ? lo_i[k] : lo;
        hi = k2 == k ? hi_i[k] : hi;
    }
    const int idx = lo + j;
    const bool active = lo >= 0 && j >= 1 && idx < hi;
    const double tt = active ? times[tbase + idx] : 0.0;
    const unsigned long long m2 = __ballot(active && pred(k2, tt));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (lo_i[k] >= 0) {
            const unsigned int g = (unsigned int)((m2 >> (16 * k)) & 0xffffull);
            edge[k] = g ? lo_i[k] + __ffs(g) - 1 : hi_i[k];
        }
    }
    return true;
}

template <bool kMixed, bool kFast>
__device__ __forceinline__ void flux_body(const FluxArgs& f, int pv, int tile) {
    using Real = typename std::conditional<kMixed, float, double>::type;
    extern __shared__ __align__(16) unsigned char lds[];
    uint16_t* s_q = (uint16_t*)lds;
    double* s_tc = (double*)(lds + kTile * sizeof(uint16_t));
    Real* s_dep = (Real*)(s_tc + kMaxChunk);
    Real* s_row = s_dep + (f.uniform_ns > 0 ? f.chunk_pts * f.uniform_ns : 0);

    const int wl = threadIdx.x;
    const int tbase = tile * kTile;
    int its[kRuns], nrun[kRuns];
#pragma unroll
    for (int m = 0; m < kRuns; ++m) {
        its[m] = tbase + (m * kThreads + wl) * kRun;
        nrun[m] = min(kRun, f.npt - its[m]);  // <= 0 past the end
    }
    const PvRec& r = f.rec[pv];
    Real* out = (Real*)f.flux + (size_t)pv * f.npt;
    // The tile test of a tile in light curve 0's window (one light curve, or one
    // window shared by all), computed ahead of any branch so that its loads (the
    // record, light curve 0's context, the tile's first and last times) are all in
    // flight together; a pv without a valid record gets NaN.
    const int tlast = min(tbase + kTile, f.npt) - 1;
    const double t_first = f.times[tbase], t_last = f.times[tlast];
    const bool rec_ok = (r.good != 0) & (r.ldok != 0);
    const PointCtx ctx0 = point_ctx0(f, r);
    const double p = r.p, inv_p = r.inv_p;
    double ea0, eb0;
    const double tca0 = phase_of(t_first, ctx0.t0v, p, inv_p, ea0);
    const double tcb0 = phase_of(t_last, ctx0.t0v, p, inv_p, eb0);
    const bool one_ctx = kFast && (f.nlc == 1 || f.one_window);
    const bool skip = (!rec_ok) | (one_ctx & tile_outside(tca0, tcb0, ea0, eb0, ctx0));  // no short circuit
    if (skip) {
        wave_fill<Real>(out, tbase, tlast + 1 - tbase, rec_ok ? Real(1) : Real(NAN), wl);
        return;
    }
    // sorted light curves: the tile's points lie in light curves la0..lb0
    // (lcids are non-decreasing); up to kTileLcs of them are staged in LDS
    // (staged once the tile is known to hold a window)
    PointCtx* s_lc = (PointCtx*)s_tc;
    int la0 = 0;
    bool lc_lds = false;
    auto stage_lcs = [&]() {
        if (!(kFast && f.nlc > 1)) return;
        la0 = min(max(f.lcids[tbase], 0), f.nlc - 1);
        const int lb0 = min(max(f.lcids[min(tbase + kTile, f.npt) - 1], 0), f.nlc - 1);
        lc_lds = lb0 != la0 && lb0 - la0 < kTileLcs;
        if (lc_lds) {
            if (wl <= lb0 - la0) s_lc[wl] = point_ctx(f, r, pv, la0 + wl);
            __syncthreads();
        }
    };
    // (the clamp keeps a caller that flags unsorted data as sorted inside the table)
    auto ctx_of = [&](int ilc) -> PointCtx {
        return lc_lds ? s_lc[min(max(ilc - la0, 0), kTileLcs - 1)] : point_ctx(f, r, pv, ilc);
    };
    int count = 0;
    bool ranged = false;  // the queue came from window_edges
    // the ones around a searched window are written after the window's points:
    // phase B's loads would otherwise wait for these stores (one memory counter
    // covers loads and stores)
    int fr0 = 0, fr1 = 0, fr2 = 0, fr3 = 0, fcnt = 0;
    auto fill_deferred = [&]() {
        if (!ranged) return;
        wave_fill<Real>(out, tbase, fr0, Real(1), wl);
        wave_fill<Real>(out, tbase + fr1, fr2 - fr1, Real(1), wl);
        wave_fill<Real>(out, tbase + fr3, fcnt - fr3, Real(1), wl);
    };
    auto push = [&](bool pred, int it) {
        const unsigned long long mk = __ballot(pred);
        if (pred) s_q[count + __popcll(mk & ((1ull << wl) - 1ull))] = (uint16_t)(it - tbase);
        count += __popcll(mk);
    };

    // ---- phase A -----------------------------------------------------------------
    if (kFast) {
        // light curves are contiguous and each is sorted in time: a run whose first
        // and last points share a light curve lies in that light curve, in order.
        // The tile as a whole first: most tiles of a long light curve hold no point
        // of a transit window and are filled with ones at once.
        {
            const int last = tlast;
            const int la = one_ctx ? 0 : min(max(f.lcids[tbase], 0), f.nlc - 1);
            const int lb = one_ctx ? 0 : min(max(f.lcids[last], 0), f.nlc - 1);
            if (la == lb) {
                // the window of the tile's light curve, or the one window all
                // light curves share (times then sorted across light curves too;
                // tested above)
                PointCtx c = ctx0;
                double ea = ea0, eb = eb0;
                if (!one_ctx) {
                    c = point_ctx(f, r, pv, la);
                    const double tca = phase_of(t_first, c.t0v, p, inv_p, ea);
                    const double tcb = phase_of(t_last, c.t0v, p, inv_p, eb);
                    if (tile_outside(tca, tcb, ea, eb, c)) {
                        wave_fill<Real>(out, tbase, last + 1 - tbase, Real(1), wl);
                        return;
                    }
                }
                // the window's index ranges by search: ones around them, the
                // ranges into the queue, no run classification
                const int cnt = last + 1 - tbase;
                int edge[4];
                if (window_edges(f.times, tbase, cnt, c, p, ea, eb, wl, edge)) {
                    const int r0 = edge[0], r1 = edge[1];
                    const int r2 = ea == eb ? r1 : edge[2], r3 = ea == eb ? r1 : edge[3];
                    fr0 = r0;
                    fr1 = r1;
                    fr2 = r2;
                    fr3 = r3;
                    fcnt = cnt;
                    const int n0 = r1 - r0;
                    count = n0 + (r3 - r2);
                    for (int q = wl; q < count; q += kThreads) s_q[q] = (uint16_t)(q < n0 ? r0 + q : r2 + (q - n0));
                    ranged = true;
                }
            }
        }
        stage_lcs();
        const PointCtx c1 = f.nlc == 1 ? ctx0 : ctx_of(la0);
        // the runs' first and last times (only tiles classified run by run need
        // them, and a wave keeps its slot until its loads return)
        double tf[kRuns], tl[kRuns];
#pragma unroll
        for (int m = 0; m < kRuns; ++m) {
            tf[m] = !ranged && nrun[m] > 0 ? f.times[its[m]] : 0.0;
            tl[m] = !ranged && nrun[m] > 0 ? f.times[its[m] + nrun[m] - 1] : 0.0;
        }
#pragma unroll
        for (int m = 0; m < (ranged ? 0 : kRuns); ++m) {
            bool all_out = true, one_lc = true, one_ep = false;
            double base = 0.0;
            PointCtx c = c1;
            if (nrun[m] > 0 && f.nlc > 1) {
                const int la = min(max(f.lcids[its[m]], 0), f.nlc - 1);
                const int lb = min(max(f.lcids[its[m] + nrun[m] - 1], 0), f.nlc - 1);
                one_lc = la == lb;
                c = ctx_of(la);
            }
            if (nrun[m] > 0) {
                all_out = false;
                if (one_lc) {
                    double ef, el;
                    const double tcf = phase_of(tf[m], c.t0v, p, inv_p, ef);
                    const double tcl = phase_of(tl[m], c.t0v, p, inv_p, el);
                    all_out = (ef == el) ? (tcl < c.lo || tcf > c.hi)
                                         : ((el == ef + 1.0) ? (tcf > c.hi && tcl < c.lo) : false);
                    // the epoch is monotone in t: equal at both ends, equal throughout
                    one_ep = ef == el;
                    base = epoch_base(c.t0v, ef, p);
                }
            }
            const int step0 = tbase + m * kThreads * kRun;
            if (__ballot(!all_out) == 0ull) {  // the wave's whole run step is outside the window
                wave_fill<Real>(out, step0, min(kThreads * kRun, f.npt - step0), Real(1), wl);
                continue;
            }
            if (all_out && nrun[m] > 0)
                for (int j = 0; j < nrun[m]; ++j) out[its[m] + j] = Real(1);
            // the run's times, all loads issued before the stores below (which
            // the compiler may not move loads across)
            double tv[kRun];
            if (!all_out) load_run(f.times, its[m], nrun[m], tv);
#pragma unroll
            for (int j = 0; j < kRun; ++j) {
                bool inw = false;
                const int it = its[m] + j;
                if (!all_out && j < nrun[m]) {
                    double tc;
                    PointCtx cc = c;
                    if (one_ep) {
                        tc = tv[j] - base;  // model_full.py:89 with the run's epoch
                    } else {
                        if (!one_lc) cc = ctx_of(min(max(f.lcids[it], 0), f.nlc - 1));
                        double ep;
                        tc = phase_of(tv[j], cc.t0v, p, inv_p, ep);
                    }
                    inw = (cc.lo <= tc) && (tc <= cc.hi);
                    if (!inw) out[it] = Real(1);
                }
                push(inw, it);
            }
        }
    } else {
#pragma unroll
        for (int m = 0; m < kRuns; ++m) {
            double tv[kRun];
            load_run(f.times, its[m], nrun[m], tv);
#pragma unroll
            for (int j = 0; j < kRun; ++j) {
                bool inw = false;
                const int it = its[m] + j;
                if (j < nrun[m]) {
                    const int ilc = f.nlc == 1 ? 0 : min(max(f.lcids[it], 0), f.nlc - 1);
                    const PointCtx c = point_ctx(f, r, pv, ilc);
                    double ep;
                    const double tc = phase_of(tv[j], c.t0v, p, inv_p, ep);
                    inw = (c.lo <= tc) && (tc <= c.hi);
                    if (!inw) out[it] = Real(1);
                }
                push(inw, it);
            }
        }
    }
    const int n = count;
    if (n == 0) {
        fill_deferred();
        return;
    }
    // one light curve with one nsamples: the first chunk's times are loaded here, in
    // flight together with the limb-darkening row staged below, and each chunk loads
    // the next chunk's times before its items (no exposed load latency per chunk)
    const bool uniform_lc = f.uniform_ns > 0 && f.nlc == 1 && f.row_lds;
    double tnext = 0.0;
    if (uniform_lc) {
        __syncthreads();  // the queue entries of other lanes
        if (wl < min(f.chunk_pts, n)) tnext = f.times[tbase + s_q[wl]];
    }

    // ---- phase B -----------------------------------------------------------------
    if (f.row_lds) {
        const int ipb = min(max(f.pbids[0], 0), f.npb - 1);
        const size_t row = ((size_t)pv * f.npb + ipb) * f.ng;
        for (int i = wl; i < f.ng; i += kThreads) s_row[i] = kMixed ? (Real)f.ldm32[row + i] : (Real)f.ldm64[row + i];
    }
    __syncthreads();  // queue entries and the staged row are read by other lanes
    const float inv_dg = (float)(1.0 / f.dg);
    // partial-overlap queue of the FP32 item loops, after the staged row and the
    // multi-light-curve context
    Real* s_pz = (Real*)(s_row + (f.row_lds ? f.ng : 0)) + (kMixed && f.nlc > 1 ? kCtxWords * kMaxChunk : 0);
    uint16_t* s_pi = (uint16_t*)(s_pz + kPartQ);
    float* s_pzf = (float*)s_pz;  // the FP32 loops' view
    if (uniform_lc) {
        // one light curve: the per-point context is uniform
        const int ns = f.uniform_ns;
        const PointCtx c = point_ctx0(f, r);
        const float inv_ns = 1.0f / (float)ns;
        const float kf = (float)c.kk, extf = (float)c.ext;
        float cf[10];
#pragma unroll
        for (int k = 0; k < 10; ++k) cf[k] = (float)r.c[k];
        float* s_tcf = (float*)s_tc;
        // exposure offsets of the samples, exptime ((is + 1/2) / ns - 1/2) (model_full.py:94)
        float* s_off = s_tcf + kMaxChunk;
        if (kMixed && wl < ns) s_off[wl] = extf * (((float)wl + 0.5f) * inv_ns - 0.5f);
        const bool small = kMixed && kf < 1.0f && ns <= kThreads;
        const int cp = f.chunk_pts;
        RecipF64 rc64;
        if (!kMixed) {
            rc64.inv_1pk = 1.0 / (1.0 + c.kk);
            rc64.inv_ist = 1.0 / c.ist;
            rc64.inv_dg = 1.0 / f.dg;
            rc64.inv_ns = 1.0 / c.ns;
        }
        for (int c0 = 0; c0 < n; c0 += cp) {
            const int cn = min(cp, n - c0);
            const double tcur = tnext;
            if (wl < cp && wl < n - (c0 + cp)) tnext = f.times[tbase + s_q[c0 + cp + wl]];
            if (wl < cn) {
                double ep;
                const double tc = phase_of(tcur, c.t0v, p, inv_p, ep);
                if (kMixed) s_tcf[wl] = (float)tc;
                else s_tc[wl] = tc;
            }
            __syncthreads();
            // item i = (point qq, sample is), i = qq * ns + is, advanced by 64 per round
            int qq = (int)(((float)wl + 0.5f) * inv_ns);  // exact i / ns for i, ns <= 1024
            int is = wl - qq * ns;
            const int dq = kThreads / ns, ds = kThreads - dq * ns;
            if (small) {
                // rounds of 64 items over the whole wave; partial overlaps queued
                const int tot = cn * ns;
                int pq = 0;
#pragma unroll 1
                for (int b = 0; b < tot; b += kThreads) {
                    const int i = b + wl;
                    bool part = false;
                    float z = 0.0f;
                    if (i < tot)
                        s_dep[i] = sample_small_f32_pass1(cf, kf, inv_dg, s_tcf[qq] + s_off[is], (const float*)s_row,
                                                          f.ng, z, part);
                    pq = queue_partial<float>(part, i, z, pq, s_pi, s_pzf, wl);
                    if (pq >= kThreads) {
                        __syncthreads();
                        pq -= kThreads;
                        s_dep[s_pi[pq + wl]] *= kite_small_f32(kf, s_pzf[pq + wl]);
                        __syncthreads();
                    }
                    qq += dq;
                    is += ds;
                    if (is >= ns) {
                        is -= ns;
                        ++qq;
                    }
                }
                __syncthreads();
                if (wl < pq) s_dep[s_pi[wl]] *= kite_small_f32(kf, s_pzf[wl]);
            } else if (!kMixed) {
                const int tot = cn * ns;
                int pq = 0;
#pragma unroll 1
                for (int b = 0; b < tot; b += kThreads) {
                    const int i = b + wl;
                    bool part = false;
                    double z = 0.0;
                    if (i < tot)
                        s_dep[i] = sample_f64_pass1(r.c, f.dg, f.ng, c, rc64, (const double*)s_row, s_tc[qq], is, z, part);
                    pq = queue_partial<Real>(part, i, z, pq, s_pi, s_pz, wl);
                    if (pq >= kThreads) {
                        __syncthreads();
                        pq -= kThreads;
                        const int j = s_pi[pq + wl];
                        s_dep[j] = sample_f64_finish(s_dep[j], c, rc64, s_pz[pq + wl]);
                        __syncthreads();
                    }
                    qq += dq;
                    is += ds;
                    if (is >= ns) {
                        is -= ns;
                        ++qq;
                    }
                }
                __syncthreads();
                if (wl < pq) {
                    const int j = s_pi[wl];
                    s_dep[j] = sample_f64_finish(s_dep[j], c, rc64, s_pz[wl]);
                }
            } else {
#pragma unroll 1
                for (int i = wl; i < cn * ns; i += kThreads) {
                    s_dep[i] = sample_f32(cf, kf, inv_dg, s_tcf[qq], extf, inv_ns, (const float*)s_row, f.ng, is);
                    qq += dq;
                    is += ds;
                    if (is >= ns) {
                        is -= ns;
                        ++qq;
                    }
                }
            }
            __syncthreads();
            if (wl < cn) {
                Real acc = Real(0);
                for (int is = 0; is < ns; ++is) acc += s_dep[wl * ns + is];
                const int it = tbase + s_q[c0 + wl];
                if (kMixed) out[it] = Real(1.0f - (float)acc * ((float)(1.0 / c.ist) * inv_ns));
                else out[it] = Real(acc / ns);
            }
            __syncthreads();
        }
        fill_deferred();
        return;
    }
    if (kMixed && f.uniform_ns > 0 && f.nlc > 1) {
        // several light curves with one nsamples: (point, sample) items as above, with
        // each point's light-curve context (phase, k, exposure, 1/istar, LD row) in LDS
        const int ns = f.uniform_ns;
        const float inv_ns = 1.0f / (float)ns;
        float cf[10];
#pragma unroll
        for (int k = 0; k < 10; ++k) cf[k] = (float)r.c[k];
        float* s_ptc = (float*)(s_row + (f.row_lds ? f.ng : 0));
        float* s_pk = s_ptc + kMaxChunk;
        float* s_pext = s_pk + kMaxChunk;
        float* s_pist = s_pext + kMaxChunk;
        int* s_prow = (int*)(s_pist + kMaxChunk);  // passband of the point
        const int cp = f.chunk_pts;
        for (int c0 = 0; c0 < n; c0 += cp) {
            const int cn = min(cp, n - c0);
            if (wl < cn) {
                const int it = tbase + s_q[c0 + wl];
                const int ilc = min(max(f.lcids[it], 0), f.nlc - 1);
                const PointCtx c = ctx_of(ilc);
                double ep;
                s_ptc[wl] = (float)phase_of(f.times[it], c.t0v, p, inv_p, ep);
                s_pk[wl] = (float)c.kk;
                s_pext[wl] = (float)c.ext;
                s_pist[wl] = (float)(1.0 / c.ist);
                s_prow[wl] = c.ipb;
            }
            __syncthreads();
            int qq = (int)(((float)wl + 0.5f) * inv_ns);  // exact i / ns for i, ns <= 1024
            int is = wl - qq * ns;
            const int dq = kThreads / ns, ds = kThreads - dq * ns;
            const int tot = cn * ns;
            int pq = 0;
#pragma unroll 1
            for (int b = 0; b < tot; b += kThreads) {
                const int i = b + wl;
                bool part = false;
                float z = 0.0f;
                if (i < tot) {
                    const float kf = s_pk[qq];
                    const float* lrow = f.ldm32 + ((size_t)pv * f.npb + s_prow[qq]) * f.ng;
                    if (kf < 1.0f) {
                        const float off = s_pext[qq] * (((float)is + 0.5f) * inv_ns - 0.5f);
                        s_dep[i] = sample_small_f32_pass1(cf, kf, inv_dg, s_ptc[qq] + off, lrow, f.ng, z, part);
                    } else {
                        s_dep[i] = sample_f32(cf, kf, inv_dg, s_ptc[qq], s_pext[qq], inv_ns, lrow, f.ng, is);
                    }
                }
                pq = queue_partial<float>(part, i, z, pq, s_pi, s_pzf, wl);
                if (pq >= kThreads) {
                    __syncthreads();
                    pq -= kThreads;
                    const int j = s_pi[pq + wl];
                    s_dep[j] *= kite_small_f32(s_pk[(int)(((float)j + 0.5f) * inv_ns)], s_pzf[pq + wl]);
                    __syncthreads();
                }
                qq += dq;
                is += ds;
                if (is >= ns) {
                    is -= ns;
                    ++qq;
                }
            }
            __syncthreads();
            if (wl < pq) {
                const int j = s_pi[wl];
                s_dep[j] *= kite_small_f32(s_pk[(int)(((float)j + 0.5f) * inv_ns)], s_pzf[wl]);
            }
            __syncthreads();
            if (wl < cn) {
                float acc = 0.0f;
                for (int k = 0; k < ns; ++k) acc += s_dep[wl * ns + k];
                out[tbase + s_q[c0 + wl]] = Real(1.0f - acc * (s_pist[wl] * inv_ns));
            }
            __syncthreads();
        }
        fill_deferred();
        return;
    }
    // general case: one queued point per lane, per-point light-curve context
#pragma unroll 1
    for (int q = wl; q < n; q += kThreads) {
        const int it = tbase + s_q[q];
        const int ilc = f.nlc == 1 ? 0 : min(max(f.lcids[it], 0), f.nlc - 1);
        const PointCtx c = ctx_of(ilc);
        double ep;
        const double tc = phase_of(f.times[it], c.t0v, p, inv_p, ep);
        const size_t row = ((size_t)pv * f.npb + c.ipb) * f.ng;
        if (kMixed) {
            float cf[10];
#pragma unroll
            for (int k = 0; k < 10; ++k) cf[k] = (float)r.c[k];
            const float inv_ns = 1.0f / (float)c.ns;
            const float* lrow = f.row_lds ? (const float*)s_row : f.ldm32 + row;
            float depth = 0.0f;
            for (int is = 0; is < c.ns; ++is)
                depth += sample_f32(cf, (float)c.kk, inv_dg, (float)tc, (float)c.ext, inv_ns, lrow, f.ng, is);
            out[it] = Real(1.0f - depth * ((float)(1.0 / c.ist) * inv_ns));
        } else {
            const double* lrow = f.row_lds ? (const double*)s_row : f.ldm64 + row;
            double acc = 0.0;
            for (int is = 0; is < c.ns; ++is) acc += sample_f64(r.c, f.dg, f.ng, c, lrow, tc, is);
            out[it] = Real(acc / c.ns);
        }
    }
    fill_deferred();
}

// Work items of the XCD-aware order go through the pvs in groups of kTileGroup:
// within a group, the same time tile of every pv before the next tile.  Measured at
// 4096 pv x 20 000 t against the pv-major order (each pv's tiles in a row): FP32
// 102-103 -> 96-98 us, FP64 258 -> 237 us, the same HBM reads.  A tile-major order
// over all 4096 pvs gained less (FP64 248 us) and read 4.5x more (each XCD then
// touches every pv's record and limb-darkening row).
constexpr int kTileGroup = 256;

// n / d for 0 <= n < 2^31 by the host's multiplier m = ceil(2^k / d), k = 31 + ceil(log2 d):
// n m / 2^k exceeds n / d by less than n / 2^k <= 1 / d, so the floor is exact.
__device__ __forceinline__ int udiv31(int n, unsigned long long m, int k) {
    return (int)(((unsigned long long)(unsigned)n * m) >> k);
}

static inline void udiv31_make(int d, unsigned long long& m, int& k) {
    int l = 0;
    while ((1ll << l) < (long long)d) ++l;
    k = 31 + l;
    m = (unsigned long long)(((1ull << k) + (unsigned long long)d - 1ull) / (unsigned long long)d);
}

// (pv, tile) of work item `item` of npv x ntiles, in groups of kTileGroup pvs.
__device__ __forceinline__ void grouped_tile(const FluxArgs& f, int item, int& pv, int& tile) {
    const int per = kTileGroup * f.ntiles;
    const int grp = udiv31(item, f.m_grp, f.k_grp), rem = item - grp * per;
    const int gsz = min(kTileGroup, f.npv - grp * kTileGroup);   // the last group may be smaller
    // a full group divides by kTileGroup = 2^8: multiplier 2^31, shift 39 (no branch)
    const bool full = gsz == kTileGroup;
    tile = udiv31(rem, full ? (1ull << 31) : f.m_last, full ? 39 : f.k_last);
    pv = grp * kTileGroup + (rem - tile * gsz);
}

template <bool kMixed, bool kFast>
__global__ void __launch_bounds__(kThreads) rr_flux_kernel(FluxArgs f) {
    int pv, tile;
    grouped_tile(f, xcd_remap(blockIdx.x, gridDim.x), pv, tile);
    flux_body<kMixed, kFast>(f, pv, tile);
}

struct FluxBatch {
    int n;
    int start[kMaxBatch + 1];
    FluxArgs a[kMaxBatch];
};

template <bool kMixed, bool kFast>
__global__ void __launch_bounds__(kThreads) rr_flux_batch_kernel(FluxBatch b) {
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int i = batch_item(b.start, b.n, bid);
    int pv, tile;
    grouped_tile(b.a[i], bid - b.start[i], pv, tile);
    flux_body<kMixed, kFast>(b.a[i], pv, tile);
}

// Device-memory batch (see rr_setup_batch_dev_kernel).  The XCD remap is over the
// batch's own blocks, so a grid padded past them still spreads the work over all
// XCDs; block b < start[n] is on XCD b % 8 as in the unpadded launch.
template <bool kMixed, bool kFast>
__global__ void __launch_bounds__(kThreads) rr_flux_batch_dev_kernel(const FluxBatch* __restrict__ b) {
    const int n = b->n;
    if (n == 0) return;
    const int total = b->start[n];
    if ((int)blockIdx.x >= total) return;
    const int bid = xcd_remap(blockIdx.x, total);
    const int i = batch_item(b->start, n, bid);
    int pv, tile;
    grouped_tile(b->a[i], bid - b->start[i], pv, tile);
    flux_body<kMixed, kFast>(b->a[i], pv, tile);
}

// A batch as the replayable launch reads it from device memory.
struct BatchDesc {
    SetupBatch setup;
    FluxBatch flux[2];   // [0] general, [1] contiguous sorted light curves
};

static_assert(sizeof(SetupBatch) <= 4096 && sizeof(FluxBatch) <= 4096, "kernel arguments above 4 KB");

}  // namespace rr

using namespace rr;

// ======================================================================================
// C-ABI
// ======================================================================================
extern "C" {

const char* rr_last_error(void) { return g_last_error.c_str(); }

int rr_abi_version(void) { return 2; }

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

static const int kMinCoeffs[] = {0, 1, 2, 2, 4, 0, 2, 2, 2, 2, 2};

int rr_ld_profiles(int law, const double* ldc, int npv, int npb, int nldc, const double* mu, int nmu, double* ldp,
                   double* istar, void* stream) {
    if (law < RR_LD_UNIFORM || law > RR_LD_POWER_2_PM) return set_error(RR_EINVAL, "rr_ld_profiles: unknown law %d", law);
    if (npv < 0 || npb < 1 || nmu < 1 || nldc < 0) return set_error(RR_EINVAL, "rr_ld_profiles: bad sizes");
    if (nldc < kMinCoeffs[law])
        return set_error(RR_EINVAL, "rr_ld_profiles: law %d needs %d coefficients, got %d", law, kMinCoeffs[law], nldc);
    if (npv == 0) return RR_OK;
    if (!mu || !ldp || !istar || (nldc > 0 && !ldc)) return set_error(RR_EINVAL, "rr_ld_profiles: null pointer");
    const long long n = (long long)npv * npb * nmu;
    hipLaunchKernelGGL(ld_profiles_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, law,
                       ldc, npv, npb, nldc, mu, nmu, ldp, istar);
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

int rr_ldtk_interpolate(const double* grid, int nx, int ny, int nz, int npb, int nmu, double x0, double dx,
                        double y0, double dy, double z0, double dz, const double* par, int par_stride, int npv,
                        double* ldp, void* stream) {
    if (nx < 2 || ny < 2 || nz < 2 || npb < 1 || nmu < 1 || npv < 0 || par_stride < 3)
        return set_error(RR_EINVAL, "rr_ldtk_interpolate: bad sizes (each axis needs >= 2 nodes, par_stride >= 3)");
    if (!(dx > 0.0 && dy > 0.0 && dz > 0.0) || !std::isfinite(x0) || !std::isfinite(y0) || !std::isfinite(z0) ||
        !std::isfinite(dx) || !std::isfinite(dy) || !std::isfinite(dz))
        return set_error(RR_EINVAL, "rr_ldtk_interpolate: axis origins must be finite and steps positive");
    if ((long long)npb * nmu > 0x7fffffffLL || (long long)npv * npb * nmu > 0x7fffffffLL * 256LL)
        return set_error(RR_ERANGE, "rr_ldtk_interpolate: too many values");
    if (npv == 0) return RR_OK;
    if (!grid || !par || !ldp) return set_error(RR_EINVAL, "rr_ldtk_interpolate: null pointer");
    const long long n = (long long)npv * npb * nmu;
    hipLaunchKernelGGL(ldtk_interpolate_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, grid, nx, ny, nz, npb * nmu, x0, dx, y0, dy, z0, dz, par, par_stride,
                       npv, ldp);
    return check_launch("ldtk_interpolate_kernel");
}

int rr_orbit(int npv, const double* p, const double* a, const double* inc, const double* e, const double* w,
             const double* k0, double* coeffs, double* win, void* stream) {
    if (npv < 0) return set_error(RR_EINVAL, "rr_orbit: npv < 0");
    if (npv == 0) return RR_OK;
    if (!p || !a || !inc || !e || !w || !k0 || !coeffs || !win) return set_error(RR_EINVAL, "rr_orbit: null pointer");
    hipLaunchKernelGGL(orbit_kernel, dim3((npv + kOrbitGroup - 1) / kOrbitGroup), dim3(64), 0, (hipStream_t)stream,
                       npv, p, a, inc, e, w, k0, coeffs, win);
    return check_launch("orbit_kernel");
}

size_t rr_workspace_bytes(int npv, int npt, int npb, int ng, int nlc) {
    (void)nlc;
    (void)npt;
    if (npv < 0 || npt < 0 || npb < 0 || ng < 0) return 0;
    size_t b = align_up((size_t)npv * sizeof(PvRec), 256);
    b += align_up((size_t)npv * npb * sizeof(double), 256);
    b += align_up((size_t)npv * npb * ng * sizeof(double), 256);
    b += align_up((size_t)npv * npb * ng * sizeof(float), 256);
    return b;
}

struct WsView {
    PvRec* rec;
    double* istar;
    double* ldm64;
    float* ldm32;
};

static WsView carve(void* workspace, int npv, int npt, int npb, int ng) {
    (void)npt;
    char* ws = (char*)workspace;
    WsView v;
    v.rec = (PvRec*)ws;
    ws += align_up((size_t)npv * sizeof(PvRec), 256);
    v.istar = (double*)ws;
    ws += align_up((size_t)npv * npb * sizeof(double), 256);
    v.ldm64 = (double*)ws;
    ws += align_up((size_t)npv * npb * ng * sizeof(double), 256);
    v.ldm32 = (float*)ws;
    return v;
}

static int device_cus() {
    static int cache[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cache[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cache[dev] = n;
    }
    return cache[dev];
}

static int validate(const char* who, int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv,
                    size_t workspace_bytes, const void* workspace) {
    if (!tab || !lc || !pv) return set_error(RR_EINVAL, "%s: null descriptor", who);
    if (precision != RR_FP64 && precision != RR_MIXED) return set_error(RR_EINVAL, "%s: bad precision", who);
    const int npv = pv->npv, npt = lc->npt, npb = lc->npb, nlc = lc->nlc;
    if (npv < 0 || npt < 0 || nlc < 1 || npb < 1 || lc->nep < 1)
        return set_error(RR_EINVAL, "%s: bad sizes (npv=%d npt=%d nlc=%d npb=%d nep=%d)", who, npv, npt, nlc, npb,
                         lc->nep);
    if (nlc > 32767) return set_error(RR_ERANGE, "%s: at most 32767 light curves", who);
    if (pv->kcols != 1 && pv->kcols != npb)
        return set_error(RR_EINVAL, "Radius ratios should be given either as an [npv, 1] or [npv, npb] array.");
    if (pv->t0cols < 1) return set_error(RR_EINVAL, "%s: t0 needs at least one column", who);
    if (tab->nk < 2 || tab->ng < 2 || tab->nz < 1 || tab->nz > kMaxNz || !tab->weights || !tab->ze || !tab->mu)
        return set_error(RR_EINVAL, "%s: bad weight table (nz must be <= %d)", who, kMaxNz);
    if ((long long)npv * npt > (1LL << 40) || (long long)npv * npb * tab->ng > (1LL << 31))
        return set_error(RR_ERANGE, "%s: problem too large", who);
    if (npv == 0 || npt == 0) return RR_OK;
    const size_t need = rr_workspace_bytes(npv, npt, npb, tab->ng, nlc);
    if (tab->ng > 8192) return set_error(RR_ERANGE, "%s: ng must be <= 8192", who);
    if (!workspace || workspace_bytes < need)
        return set_error(RR_EINVAL, "%s: workspace of %zu bytes needed, got %zu", who, need, workspace_bytes);
    return RR_OK;
}

// Validates one request and fills the setup kernel's arguments; *nblocks = 0 when
// there is nothing to do.
static int prepare_setup(const char* who, int precision, const rr_table* tab, const rr_lcdata* lc,
                         const rr_params* pv, const rr_ld* ld, void* workspace, size_t workspace_bytes,
                         SetupArgs& sa, int& nblocks) {
    nblocks = 0;
    int rc = validate(who, precision, tab, lc, pv, workspace_bytes, workspace);
    if (rc) return rc;
    const int npv = pv->npv, npb = lc->npb;
    if (npv == 0 || lc->npt == 0) return RR_OK;
    if (!ld) return set_error(RR_EINVAL, "%s: null limb-darkening descriptor", who);
    if (ld->law == RR_LD_TABULATED) {
        if (!ld->ldp || !ld->istar) return set_error(RR_EINVAL, "%s: tabulated LD needs ldp and istar", who);
    } else {
        if (ld->law < RR_LD_UNIFORM || ld->law > RR_LD_POWER_2_PM)
            return set_error(RR_EINVAL, "%s: unknown limb-darkening law %d", who, ld->law);
        if (ld->nldc < kMinCoeffs[ld->law])
            return set_error(RR_EINVAL, "%s: law %d needs %d coefficients, got %d", who, ld->law,
                             kMinCoeffs[ld->law], ld->nldc);
        if (ld->nldc > 0 && !ld->ldc) return set_error(RR_EINVAL, "%s: null ldc", who);
    }
    if (!pv->k || !pv->p || !pv->a || !pv->inc || !pv->e || !pv->w || !pv->t0)
        return set_error(RR_EINVAL, "%s: null parameter array", who);
    if (!lc->pbids || !lc->epids) return set_error(RR_EINVAL, "%s: null light-curve array", who);
    WsView v = carve(workspace, npv, lc->npt, npb, tab->ng);
    sa.t0 = pv->t0; sa.pbids = lc->pbids; sa.epids = lc->epids; sa.t0cols = pv->t0cols;
    sa.k = pv->k; sa.p = pv->p; sa.a = pv->a; sa.inc = pv->inc; sa.e = pv->e; sa.w = pv->w;
    sa.law = ld->law; sa.nldc = ld->nldc; sa.ldc = ld->ldc; sa.ldp_in = ld->ldp; sa.istar_in = ld->istar;
    sa.W = tab->weights; sa.W32 = tab->weights32; sa.ze = tab->ze; sa.mu = tab->mu;
    sa.rec = v.rec; sa.istar = v.istar;
    sa.ldm64 = precision == RR_FP64 ? v.ldm64 : nullptr;
    sa.ldm32 = precision == RR_MIXED ? v.ldm32 : nullptr;
    sa.npv = npv; sa.npb = npb; sa.kcols = pv->kcols; sa.nk = tab->nk; sa.ng = tab->ng; sa.nz = tab->nz;
    sa.kmin = tab->kmin; sa.kmax = tab->kmax; sa.dk = tab->dk;
    sa.n_orbit_blocks = (npv + kOrbitGroup - 1) / kOrbitGroup;
    sa.ld_lds = ld_lds_doubles(tab->nz, tab->ng, ld->law, ld->nldc, npb);
    const int gs = ld_group_size(tab->nz, ld->law, ld->nldc, npb);
    const long long nb = (long long)sa.n_orbit_blocks + (long long)npv * ((npb + gs - 1) / gs);
    if (nb > 0x7fffffffLL) return set_error(RR_ERANGE, "%s: setup grid too large", who);
    nblocks = (int)nb;
    return RR_OK;
}

// Validates one request and fills the flux kernel's arguments, its dynamic LDS
// bytes and its variant (light curves contiguous and sorted, or general).
static int prepare_flux(const char* who, int precision, const rr_table* tab, const rr_lcdata* lc,
                        const rr_params* pv, void* workspace, size_t workspace_bytes, void* flux, FluxArgs& fa,
                        long long& nblocks, size_t& lds, bool& fast) {
    nblocks = 0;
    int rc = validate(who, precision, tab, lc, pv, workspace_bytes, workspace);
    if (rc) return rc;
    const int npv = pv->npv, npt = lc->npt, npb = lc->npb, nlc = lc->nlc;
    if (npv == 0 || npt == 0) return RR_OK;
    if (!lc->times || !lc->pbids || !lc->epids || !lc->nsamples || !lc->exptimes || (nlc > 1 && !lc->lcids))
        return set_error(RR_EINVAL, "%s: null light-curve array", who);
    if (!pv->k || !pv->t0 || !flux) return set_error(RR_EINVAL, "%s: null parameter array", who);
    WsView v = carve(workspace, npv, npt, npb, tab->ng);
    fa.times = lc->times; fa.lcids = lc->lcids; fa.pbids = lc->pbids; fa.epids = lc->epids;
    fa.nsamples = lc->nsamples; fa.exptimes = lc->exptimes;
    fa.k = pv->k; fa.t0 = pv->t0; fa.istar = v.istar;
    fa.rec = v.rec; fa.ldm64 = v.ldm64; fa.ldm32 = v.ldm32; fa.flux = flux;
    fa.npt = npt; fa.nlc = nlc; fa.npb = npb; fa.kcols = pv->kcols; fa.t0cols = pv->t0cols;
    fa.ng = tab->ng; fa.dg = tab->dg;
    const bool mixed = precision == RR_MIXED;
    const int uns = lc->uniform_nsamples;
    fa.uniform_ns = (uns > 0 && uns <= kMaxItems) ? uns : 0;
    fa.chunk_pts = fa.uniform_ns > 0 ? (kMaxItems / fa.uniform_ns < kMaxChunk ? kMaxItems / fa.uniform_ns : kMaxChunk) : 0;
    fa.row_lds = (nlc == 1 && tab->ng <= 2048) ? 1 : 0;
    fa.ntiles = (npt + kTile - 1) / kTile;
    nblocks = (long long)fa.ntiles * npv;
    if (nblocks > 0x7fffffffLL) return set_error(RR_ERANGE, "%s: grid too large", who);
    fa.npv = npv;
    udiv31_make(kTileGroup * fa.ntiles, fa.m_grp, fa.k_grp);
    udiv31_make(npv - (npv - 1) / kTileGroup * kTileGroup, fa.m_last, fa.k_last);
    fast = (lc->sorted & RR_SORTED_RUNS) != 0;
    fa.one_window = nlc > 1 && (lc->sorted & (RR_SORTED_RUNS | RR_SORTED_TIMES | RR_ONE_WINDOW)) ==
                                   (RR_SORTED_RUNS | RR_SORTED_TIMES | RR_ONE_WINDOW);
    lds = flux_lds_bytes(mixed, tab->ng, fa.chunk_pts, fa.uniform_ns, fa.row_lds != 0, nlc > 1);
    return RR_OK;
}

int rr_setup(int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv, const rr_ld* ld,
             void* workspace, size_t workspace_bytes, void* stream) {
    SetupArgs sa;
    int nblocks = 0;
    int rc = prepare_setup("rr_setup", precision, tab, lc, pv, ld, workspace, workspace_bytes, sa, nblocks);
    if (rc || nblocks == 0) return rc;
    if (sa.npb > 1)
        hipLaunchKernelGGL(rr_setup_kernel<true>, dim3(nblocks), dim3(64), setup_lds_bytes(sa), (hipStream_t)stream, sa);
    else
        hipLaunchKernelGGL(rr_setup_kernel<false>, dim3(nblocks), dim3(64), setup_lds_bytes(sa), (hipStream_t)stream, sa);
    return check_launch("rr_setup_kernel");
}

int rr_flux(int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv, void* workspace,
            size_t workspace_bytes, void* flux, void* stream) {
    FluxArgs fa;
    long long nblocks = 0;
    size_t lds = 0;
    bool fast = false;
    int rc = prepare_flux("rr_flux", precision, tab, lc, pv, workspace, workspace_bytes, flux, fa, nblocks, lds, fast);
    if (rc || nblocks == 0) return rc;
    hipStream_t s = (hipStream_t)stream;
    const dim3 grid((unsigned)nblocks), block(kThreads);
    if (precision == RR_MIXED) {
        if (fast) hipLaunchKernelGGL((rr_flux_kernel<true, true>), grid, block, lds, s, fa);
        else hipLaunchKernelGGL((rr_flux_kernel<true, false>), grid, block, lds, s, fa);
    } else {
        if (fast) hipLaunchKernelGGL((rr_flux_kernel<false, true>), grid, block, lds, s, fa);
        else hipLaunchKernelGGL((rr_flux_kernel<false, false>), grid, block, lds, s, fa);
    }
    return check_launch("rr_flux_kernel");
}

int rr_evaluate(int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv, const rr_ld* ld,
                void* workspace, size_t workspace_bytes, void* flux, void* stream) {
    int rc = rr_setup(precision, tab, lc, pv, ld, workspace, workspace_bytes, stream);
    if (rc) return rc;
    return rr_flux(precision, tab, lc, pv, workspace, workspace_bytes, flux, stream);
}

int rr_max_batch(void) { return kMaxBatch; }

// Validates every request of a batch and fills its setup and flux batches, grid
// sizes and dynamic LDS bytes.
static int plan_batch(const char* who0, int precision, const rr_table* tab, int nreq, const rr_request* reqs,
                      BatchDesc& bd, rr_batch_shape& sh) {
    if (nreq < 0 || nreq > kMaxBatch)
        return set_error(RR_EINVAL, "%s: %d requests, at most %d per call", who0, nreq, kMaxBatch);
    if (nreq > 0 && !reqs) return set_error(RR_EINVAL, "%s: null request array", who0);
    SetupBatch& sb = bd.setup;
    FluxBatch* fb = bd.flux;
    size_t lds[2] = {0, 0}, slds = 0;
    bool multi = false;
    long long fblocks[2] = {0, 0};
    long long sblocks = 0;
    sb.n = 0;
    fb[0].n = fb[1].n = 0;
    sb.start[0] = fb[0].start[0] = fb[1].start[0] = 0;
    char who[48];
    for (int r = 0; r < nreq; ++r) {
        const rr_request& q = reqs[r];
        snprintf(who, sizeof(who), "%s[%d]", who0, r);
        SetupArgs sa;
        int nb = 0;
        int rc = prepare_setup(who, precision, tab, &q.lc, &q.pv, &q.ld, q.workspace, q.workspace_bytes, sa, nb);
        if (rc) return rc;
        FluxArgs fa;
        long long nf = 0;
        size_t l = 0;
        bool fast = false;
        rc = prepare_flux(who, precision, tab, &q.lc, &q.pv, q.workspace, q.workspace_bytes, q.flux, fa, nf, l, fast);
        if (rc) return rc;
        if (nb == 0) continue;
        sb.start[sb.n] = (int)sblocks;
        sb.a[sb.n++] = sa;
        sblocks += nb;
        slds = setup_lds_bytes(sa) > slds ? setup_lds_bytes(sa) : slds;
        multi = multi || sa.npb > 1;
        FluxBatch& b = fb[fast ? 1 : 0];
        b.start[b.n] = (int)fblocks[fast ? 1 : 0];
        b.a[b.n++] = fa;
        fblocks[fast ? 1 : 0] += nf;
        lds[fast ? 1 : 0] = l > lds[fast ? 1 : 0] ? l : lds[fast ? 1 : 0];
    }
    if (sblocks > 0x7fffffffLL || fblocks[0] > 0x7fffffffLL || fblocks[1] > 0x7fffffffLL)
        return set_error(RR_ERANGE, "%s: grid too large", who0);
    sb.start[sb.n] = (int)sblocks;
    fb[0].start[fb[0].n] = (int)fblocks[0];
    fb[1].start[fb[1].n] = (int)fblocks[1];
    sh.setup_blocks = (int32_t)sblocks;
    sh.setup_multi = multi ? 1 : 0;
    sh.setup_lds = (int32_t)slds;
    for (int v = 0; v < 2; ++v) {
        sh.flux_blocks[v] = (int32_t)fblocks[v];
        sh.flux_lds[v] = (int32_t)lds[v];
    }
    return RR_OK;
}

int rr_evaluate_batch(int precision, const rr_table* tab, int nreq, const rr_request* reqs, void* stream) {
    BatchDesc bd;
    rr_batch_shape sh;
    int rc = plan_batch("rr_evaluate_batch", precision, tab, nreq, reqs, bd, sh);
    if (rc || bd.setup.n == 0) return rc;
    hipStream_t s = (hipStream_t)stream;
    const SetupBatch& sb = bd.setup;
    if (sh.setup_multi)
        hipLaunchKernelGGL(rr_setup_batch_kernel<true>, dim3((unsigned)sh.setup_blocks), dim3(64), sh.setup_lds, s, sb);
    else
        hipLaunchKernelGGL(rr_setup_batch_kernel<false>, dim3((unsigned)sh.setup_blocks), dim3(64), sh.setup_lds, s, sb);
    rc = check_launch("rr_setup_batch_kernel");
    if (rc) return rc;
    const bool mixed = precision == RR_MIXED;
    for (int v = 0; v < 2; ++v) {
        if (bd.flux[v].n == 0) continue;
        const dim3 grid((unsigned)sh.flux_blocks[v]), block(kThreads);
        const size_t l = sh.flux_lds[v];
        if (mixed) {
            if (v) hipLaunchKernelGGL((rr_flux_batch_kernel<true, true>), grid, block, l, s, bd.flux[v]);
            else hipLaunchKernelGGL((rr_flux_batch_kernel<true, false>), grid, block, l, s, bd.flux[v]);
        } else {
            if (v) hipLaunchKernelGGL((rr_flux_batch_kernel<false, true>), grid, block, l, s, bd.flux[v]);
            else hipLaunchKernelGGL((rr_flux_batch_kernel<false, false>), grid, block, l, s, bd.flux[v]);
        }
        rc = check_launch("rr_flux_batch_kernel");
        if (rc) return rc;
    }
    return RR_OK;
}

size_t rr_batch_desc_bytes(void) { return sizeof(BatchDesc); }

int rr_batch_plan(int precision, const rr_table* tab, int nreq, const rr_request* reqs, void* desc,
                  rr_batch_shape* shape) {
    if (!desc || !shape) return set_error(RR_EINVAL, "rr_batch_plan: null descriptor or shape");
    BatchDesc& bd = *(BatchDesc*)desc;
    int rc = plan_batch("rr_batch_plan", precision, tab, nreq, reqs, bd, *shape);
    if (rc) {
        bd.setup.n = bd.flux[0].n = bd.flux[1].n = 0;
        bd.setup.start[0] = bd.flux[0].start[0] = bd.flux[1].start[0] = 0;
    }
    return rc;
}

int rr_batch_launch(int precision, const void* desc_dev, const rr_batch_shape* launch, void* stream) {
    if (!desc_dev || !launch) return set_error(RR_EINVAL, "rr_batch_launch: null descriptor or shape");
    if (precision != RR_FP64 && precision != RR_MIXED) return set_error(RR_EINVAL, "rr_batch_launch: bad precision");
    const rr_batch_shape& sh = *launch;
    if (sh.setup_blocks < 0 || sh.flux_blocks[0] < 0 || sh.flux_blocks[1] < 0 || sh.setup_lds < 0 ||
        sh.flux_lds[0] < 0 || sh.flux_lds[1] < 0 || sh.setup_lds > 65536 || sh.flux_lds[0] > 65536 ||
        sh.flux_lds[1] > 65536)
        return set_error(RR_EINVAL, "rr_batch_launch: bad launch shape");
    const BatchDesc* bd = (const BatchDesc*)desc_dev;
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if (sh.setup_blocks > 0) {
        if (sh.setup_multi)
            hipLaunchKernelGGL(rr_setup_batch_dev_kernel<true>, dim3((unsigned)sh.setup_blocks), dim3(64),
                               sh.setup_lds, s, &bd->setup);
        else
            hipLaunchKernelGGL(rr_setup_batch_dev_kernel<false>, dim3((unsigned)sh.setup_blocks), dim3(64),
                               sh.setup_lds, s, &bd->setup);
        rc = check_launch("rr_setup_batch_dev_kernel");
        if (rc) return rc;
    }
    const bool mixed = precision == RR_MIXED;
    for (int v = 0; v < 2; ++v) {
        if (sh.flux_blocks[v] == 0) continue;
        const dim3 grid((unsigned)sh.flux_blocks[v]), block(kThreads);
        const size_t l = sh.flux_lds[v];
        const FluxBatch* fb = &bd->flux[v];
        if (mixed) {
            if (v) hipLaunchKernelGGL((rr_flux_batch_dev_kernel<true, true>), grid, block, l, s, fb);
            else hipLaunchKernelGGL((rr_flux_batch_dev_kernel<true, false>), grid, block, l, s, fb);
        } else {
            if (v) hipLaunchKernelGGL((rr_flux_batch_dev_kernel<false, true>), grid, block, l, s, fb);
            else hipLaunchKernelGGL((rr_flux_batch_dev_kernel<false, false>), grid, block, l, s, fb);
        }
        rc = check_launch("rr_flux_batch_dev_kernel");
        if (rc) return rc;
    }
    return RR_OK;
}

// ------------------------------------------------------------------------------------
// Evaluation without allocation.  The per-pv workspace (records, disk integrals,
// limb-darkening means) lives in a file-scope device array of RR_NOALLOC_SCRATCH_BYTES.
// Every GPU holds its own copy at its own address, found once per device with
// hipGetSymbolAddress; host code uses only that address, never the array's name
// (which on the host is a stand-in, not a device address).  The setup kernel
// writes every workspace byte the flux kernel reads, so nothing needs clearing
// between calls; calls that overlap on one GPU share the array and race.
static __device__ __align__(256) unsigned char scratch[RR_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_address(void** out) {
    int dev = -1;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return set_error(RR_EHIP, "rr_evaluate_noalloc: hipGetDevice: %s", hipGetErrorString(err));
    if (dev < 0 || dev >= kMaxDevices)
        return set_error(RR_ERANGE, "rr_evaluate_noalloc: device %d beyond the %d supported", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !p)
            return set_error(RR_EHIP, "rr_evaluate_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                             hipGetErrorString(err));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return RR_OK;
}
}  // namespace

int rr_evaluate_noalloc(int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv,
                        const rr_ld* ld, void* flux, void* stream) {
    if (!tab || !lc || !pv) return set_error(RR_EINVAL, "rr_evaluate_noalloc: null descriptor");
    if (pv->npv < 0 || lc->npt < 0 || lc->npb < 1 || tab->ng < 0)
        return set_error(RR_EINVAL, "rr_evaluate_noalloc: bad sizes (npv=%d npt=%d npb=%d ng=%d)", pv->npv, lc->npt,
                         lc->npb, tab->ng);
    const size_t need = rr_workspace_bytes(pv->npv, lc->npt, lc->npb, tab->ng, lc->nlc);
    if (need > (size_t)RR_NOALLOC_SCRATCH_BYTES)
        return set_error(RR_ERANGE,
                         "rr_evaluate_noalloc: npv=%d x npb=%d x ng=%d needs %zu bytes of workspace, above the "
                         "%zu-byte scratch array; use rr_evaluate with a caller-allocated workspace",
                         pv->npv, lc->npb, tab->ng, need, (size_t)RR_NOALLOC_SCRATCH_BYTES);
    void* ws = nullptr;
    int rc = scratch_address(&ws);
    if (rc) return rc;
    return rr_evaluate(precision, tab, lc, pv, ld, ws, (size_t)RR_NOALLOC_SCRATCH_BYTES, flux, stream);
}

}  // extern "C"
