// Transmission-spectroscopy RoadRunner model (TSModel) on MI355X (gfx950).
//
// Replaces tsmodel_serial (pytransit/models/roadrunner/model_trspec.py:11-93):
//   per pv: k_mean = mean(k), k_max = max(k), Taylor orbit, contact window at
//   k_mean widened by 0.0015 + exptimes[0]; limb-darkening means
//   ldm[pb, g] = W(k_mean)[g, :] . ldp[pb, :] with exact 2-D weights at k_mean
//   (precompute_weights=False) or a k-lerp of the 3-D table; per (t, sample) one
//   lens geometry at k_mean, fanned out over passbands with the area scaled by
//   k^2/k_mean^2 (z <= 1 - k_max) or extrapolated linearly in k otherwise.
//
// Kernels:
//   ts_setup_kernel    one 64-lane workgroup per pv (orbit on lanes 0-6, disk integrals)
//   ts_weights_kernel  up to 16 (pv, g) rows per workgroup: the 2-D weight rows at
//                      k_mean, the lens areas that need arccos evaluated from a
//                      list on full waves, written transposed W2T[pv][z][g] (FP64
//                      and FP32, g padded to 16)
//   ts_ldm_tab_mfma_kernel  FP32 MFMA contraction ldm = ldp . W2T for tabulated FP64
//                      profiles: one 8-wave workgroup per pv, W2T in LDS, each row
//                      read as 128-byte pieces (16-byte loads, a permuted z order)
//   ts_ldm_mfma_kernel the same for named laws (profiles evaluated per element) and
//                      the shapes the first does not take
//   ts_ldm_mfma64_kernel the FP64 contraction on v_mfma_f64_16x16x4f64 (W2T in LDS
//                      when it fits in 128 KB), else ts_ldm_f64_kernel on the VALU
//   ts_flux_kernel     one lane per time, a workgroup per (pv, 256 times): the FP64
//                      geometry once per (time, sample), then the passband fan-out
//                      over chunks of 64 (FP32) or 32 (FP64) LD rows staged in LDS
//
// Flux layout [npv, npb, npt]; for one passband row 256 lanes write 1 KB (FP32)
// contiguously.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <type_traits>

#include "rr_device.hpp"
#include "rr_internal.hpp"

namespace rr {

struct TsRec {
    double c[10];
    double bt1, bt4;     // window incl. the 0.0015 + exptimes[0] margin
    double p, inv_p;
    double kmean, kmax;
    double ak;           // table mode: k-lerp weight
    int32_t ik;          // table mode: row index, -1 for exact 2-D weights
    int32_t good;
    int32_t pad[2];
};

struct TsArgs {
    const double* times; const double* k; const double* t0;
    const double* exptimes; const int32_t* nsamples;   // element 0 applies to every point (Q16)
    const double* p; const double* a; const double* inc; const double* e; const double* w;
    int law, nldc; const double* ldc; const double* ldp_in; const double* istar_in;
    const double* W; const double* ze; const double* mu;
    TsRec* rec; double* istar; double* w2_64; float* w2_32; double* ldm64; float* ldm32;
    int npv, npb, npt, nk, ng, ngp, nz, nzp, use_table;
    double kmin, kmax_tab, dk, dg;
};

constexpr int kTsGroup = 8;

// ---- per-pv setup; model_trspec.py:36-68 ----------------------------------------------
constexpr int kTsKChunk = 512;   // radius ratios of a pv staged in LDS at a time

__global__ void __launch_bounds__(64) ts_setup_kernel(TsArgs s) {
    __shared__ double s_x[8][8], s_y[8][8], s_c[8][10], s_w[8][2], s_km[8];
    __shared__ double s_kb[kTsGroup][kTsKChunk];
    const int sub = threadIdx.x / kTsGroup, j = threadIdx.x % kTsGroup;
    const int pv = blockIdx.x * kTsGroup + sub;
    const bool inrange = pv < s.npv;
    const int q = inrange ? pv : 0;
    const double a = s.a[q], e = s.e[q];
    const bool valid = inrange && !(isnan(a) || a <= 1.0 || e < 0.0);
    // mean as numba's left-to-right sum, and the max, of the pv's row of radius
    // ratios: the row is loaded into LDS by the pv's 8 lanes (all loads in flight
    // at once), then summed in order by one lane
    double sum = 0.0, kmx = -INFINITY;
    const double* kr = s.k + (size_t)q * s.npb;
    for (int c0 = 0; c0 < s.npb; c0 += kTsKChunk) {
        const int cn = min(kTsKChunk, s.npb - c0);
        if (inrange)
            for (int i = j; i < cn; i += kTsGroup) s_kb[sub][i] = kr[c0 + i];
        __syncthreads();
        if (inrange && j == 0)
#pragma unroll 8
            for (int i = 0; i < cn; ++i) {
                sum += s_kb[sub][i];
                kmx = fmax(kmx, s_kb[sub][i]);
            }
        __syncthreads();
    }
    if (inrange && j == 0) {
        s_km[sub] = sum / s.npb;
        TsRec* r = s.rec + q;
        r->kmean = sum / s.npb;
        r->kmax = kmx;
        r->p = s.p[q];
        r->inv_p = 1.0 / s.p[q];
        r->good = valid ? 1 : 0;
        // The reference tests kmin <= kmean <= max(k) (its kmax is shadowed by max(k)),
        // which passes for every kmean >= kmin and reads past the table above kmax.
        // Defined here: the table serves kmean in [kmin, kmax] (the table's bounds);
        // elsewhere the exact 2-D weights at kmean are used, as the RR setup does.
        const double km = sum / s.npb;
        if (s.use_table && s.kmin <= km && km <= s.kmax_tab) {
            int ik = (int)floor((km - s.kmin) / s.dk);
            ik = ik > s.nk - 2 ? s.nk - 2 : ik;
            r->ik = ik;
            r->ak = (km - s.kmin - ik * s.dk) / s.dk;
        } else {
            r->ik = -1;
            r->ak = 0.0;
        }
    }
    __syncthreads();
    if (valid && j < 7) {
        double x, y;
        kepler_xy((j - 3) * 2e-2, s.p[q], a, cos(s.inc[q]), e, s.w[q], x, y);
        s_x[sub][j] = x;
        s_y[sub][j] = y;
    }
    __syncthreads();
    if (valid && j == 0) taylor_from_stencil(s_x[sub], s_y[sub], s_c[sub]);
    __syncthreads();
    if (valid && j < 2) s_w[sub][j] = find_contact_point(s_km[sub], j == 0 ? 1 : 4, s_c[sub]);
    __syncthreads();
    if (valid) {
        TsRec* r = s.rec + q;
        r->c[j] = s_c[sub][j];
        if (j < 2) r->c[8 + j] = s_c[sub][8 + j];
        if (j == 0) r->bt1 = s_w[sub][0] - (0.0015 + s.exptimes[0]);
        if (j == 1) r->bt4 = s_w[sub][1] + (0.0015 + s.exptimes[0]);
    }
}

// Disk integrals per (pv, pb): given, analytic, or the 200-node trapezoid.
__global__ void __launch_bounds__(256) ts_istar_kernel(TsArgs s) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= s.npv * s.npb) return;
    double is;
    if (s.law == RR_LD_TABULATED) {
        is = s.istar_in[j];
    } else {
        const double* c = s.ldc + (size_t)j * s.nldc;
        if (!ld_integral(s.law, c, is)) {
            double sum = 0.0;
            for (int m = 1; m < 200; ++m) sum += ld_trapz_segment(s.law, c, s.nldc, m);
            is = 2.0 * kPi * sum;
        }
    }
    s.istar[j] = is;
}

// ---- weights at k_mean, transposed: W2T[pv][z][g] --------------------------------------
// A workgroup takes R rows (g values) of one pv.  The exact row
// (calculate_weights_2d, common.py:152-185) is the differences of the lens areas
// at consecutive z edges over their sum.  The areas go to LDS: the cheap cases of
// the lens area (no overlap, one disk inside the other) at once, the (r, z) pairs
// that need the arccos form to a list that all lanes then work through, so no lane
// idles in a wave that evaluates it.  The sum runs over the differences in z order
// as in the reference; the rows' values for one z are written next to each other.
constexpr int kTwMaxRows = 16;
__host__ __device__ inline int ts_weight_rows(int nz) {
    const int r = 2048 / (nz > 0 ? nz : 1);
    return r < 1 ? 1 : (r > kTwMaxRows ? kTwMaxRows : r);
}
// LDS of the exact-weight kernel: areas [R][nz], the list [R][nz] (16-bit, padded
// to 8 bytes), differences [R][nz]
__host__ __device__ inline size_t ts_list_bytes(int nz) {
    return ((size_t)ts_weight_rows(nz) * nz * sizeof(uint16_t) + 7) / 8 * 8;
}
__host__ __device__ inline size_t ts_weights_lds(int nz) {
    return (size_t)ts_weight_rows(nz) * nz * 2 * sizeof(double) + ts_list_bytes(nz);
}

__global__ void __launch_bounds__(256) ts_weights_kernel(TsArgs s) {
    extern __shared__ __align__(16) double s_a[];   // [R][nz] lens areas, the list, the differences
    __shared__ int s_n;
    __shared__ double s_sum[kTwMaxRows];
    const int R = ts_weight_rows(s.nz), nz = s.nz;
    const int nchunk = (s.ngp + R - 1) / R;
    const int pv = blockIdx.x / nchunk;
    const int g0 = (blockIdx.x - pv * nchunk) * R;
    const int nr = min(R, s.ngp - g0);
    const int tid = threadIdx.x;
    const size_t base = (size_t)pv * s.nzp * s.ngp + g0;
    auto put = [&](int iz, int row, double v) {
        if (s.w2_64) s.w2_64[base + (size_t)iz * s.ngp + row] = v;
        if (s.w2_32) s.w2_32[base + (size_t)iz * s.ngp + row] = (float)v;
    };
    const TsRec& r = s.rec[pv];
    if (r.ik >= 0) {
        // (1-ak) W[ik] + ak W[ik+1] (device table layout [nk][nz][ng])
        const double* w0 = s.W + (size_t)r.ik * nz * s.ng;
        const double* w1 = w0 + (size_t)nz * s.ng;
        for (int i = tid; i < s.nzp * nr; i += blockDim.x) {
            const int iz = i / nr, row = i - iz * nr, ig = g0 + row;
            double v = 0.0;
            if (iz < nz && ig < s.ng) v = (1.0 - r.ak) * w0[(size_t)iz * s.ng + ig] + r.ak * w1[(size_t)iz * s.ng + ig];
            put(iz, row, v);
        }
        return;
    }
    uint16_t* s_list = (uint16_t*)(s_a + (size_t)R * nz);
    __shared__ double s_b[kTwMaxRows];
    const double km = r.kmean;
    if (tid == 0) s_n = 0;
    if (tid < nr) s_b[tid] = linspace_at(0.0, 1.0 - 1e-7, s.ng, g0 + tid) * (1.0 + km);
    __syncthreads();
    auto b_of = [&](int ig) { return s_b[ig - g0]; };
    const int lane = tid % 64;
    for (int i0 = tid - lane; i0 < nr * nz; i0 += blockDim.x) {   // wave-uniform trip count
        const int i = i0 + lane;
        const int row = i / nz, iz = i - row * nz;
        bool listed = false;
        if (i < nr * nz && g0 + row < s.ng) {
            const double b = b_of(g0 + row), r1 = s.ze[iz];
            if (r1 < b - km || r1 >= b + km || b - km <= -r1) s_a[i] = cci_area(r1, km, b);   // cci_area's cheap cases
            else listed = true;
        }
        // one LDS atomic per wave for its listed pairs
        const unsigned long long m = __ballot(listed);
        if (m) {
            const int first = __ffsll((long long)m) - 1;
            int at = 0;
            if (lane == first) at = atomicAdd(&s_n, __popcll(m));
            at = __shfl(at, first, 64);
            if (listed) s_list[at + __popcll(m & ((1ull << lane) - 1ull))] = (uint16_t)i;
        }
    }
    __syncthreads();
    const int nl = s_n;
    for (int j = tid; j < nl; j += blockDim.x) {
        const int i = s_list[j], row = i / nz, iz = i - row * nz;
        s_a[i] = cci_area(s.ze[iz], km, b_of(g0 + row));
    }
    __syncthreads();
    // the differences, all at once, then each row's sum over them in z order
    double* s_d = (double*)((unsigned char*)s_list + ts_list_bytes(nz));
    for (int i = tid; i < nr * nz; i += blockDim.x) {
        const int iz = i % nz;
        s_d[i] = iz == 0 ? s_a[i] : s_a[i] - s_a[i - 1];
    }
    __syncthreads();
    if (tid < nr) {
        double sum = 0.0;
        if (g0 + tid < s.ng) {
            const double* d = s_d + (size_t)tid * nz;
#pragma unroll 8
            for (int iz = 0; iz < nz; ++iz) sum += d[iz];
        }
        s_sum[tid] = sum;
    }
    __syncthreads();
    for (int i = tid; i < s.nzp * nr; i += blockDim.x) {
        const int iz = i / nr, row = i - iz * nr;
        double v = 0.0;
        if (iz < nz && g0 + row < s.ng) v = s_d[(size_t)row * nz + iz] / s_sum[row];
        put(iz, row, v);
    }
}

// Profile value of (pv, pb) at mu node iz: a table read, or the named law (the
// template keeps the law code out of the table-read kernels' registers).
template <bool kTab>
__device__ __forceinline__ double ts_profile(const TsArgs& s, int pv, int pb, int iz) {
    const size_t j = (size_t)pv * s.npb + pb;
    if (kTab) return s.ldp_in[j * s.nz + iz];
    return ld_law(s.law, s.mu[iz], s.ldc + j * s.nldc, s.nldc);
}

// ---- FP32 MFMA contraction: ldm[pv][pb][g] = sum_z ldp[pv][pb][z] * W2T[pv][z][g] ----
// D(16x16) += A(16x4) B(4x16) per v_mfma_f32_16x16x4f32.  Lane l supplies
// A[l%16][l/16] and B[l/16][l%16] and holds D[4*(l/16)+r][l%16], r = 0..3.
// A rows are passbands, B columns are g.  W2T of the pv sits in LDS with a padded
// row stride so the 16 lanes reading one k hit distinct banks.
typedef float f32x4_t __attribute__((ext_vector_type(4)));
constexpr int kMfmaK = 128;   // k per register-resident A fragment

template <bool kTab>
__global__ void __launch_bounds__(256) ts_ldm_mfma_kernel(TsArgs s) {
    extern __shared__ float lds_b[];  // [nzp][ngp + 1]
    const int pv = blockIdx.x;
    const int ldb = s.ngp + 1;
    for (int i = threadIdx.x; i < s.nzp * s.ngp; i += blockDim.x) {
        const int iz = i / s.ngp, ig = i - iz * s.ngp;
        lds_b[iz * ldb + ig] = s.w2_32[(size_t)pv * s.nzp * s.ngp + i];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int nrt = (s.npb + 15) / 16, nct = s.ngp / 16;
    for (int rt = wave; rt < nrt; rt += 4) {
        const int pb = rt * 16 + li;
        const bool row_ok = pb < s.npb;
        for (int ct0 = 0; ct0 < nct; ct0 += 8) {
            f32x4_t acc[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) acc[c] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
            for (int k0 = 0; k0 < s.nzp; k0 += kMfmaK) {
                float afrag[kMfmaK / 4];
#pragma unroll
                for (int st = 0; st < kMfmaK / 4; ++st) {
                    const int iz = k0 + st * 4 + lk;
                    afrag[st] = (row_ok && iz < s.nz) ? (float)ts_profile<kTab>(s, pv, pb, iz) : 0.0f;
                }
#pragma unroll
                for (int st = 0; st < kMfmaK / 4; ++st) {
                    const int iz = k0 + st * 4 + lk;
                    if (k0 + st * 4 >= s.nzp) break;
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        const int ct = ct0 + c;
                        if (ct < nct) {
                            const float b = lds_b[iz * ldb + ct * 16 + li];
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(afrag[st], b, acc[c], 0, 0, 0);
                        }
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int ct = ct0 + c;
                if (ct >= nct) break;
                const int g = ct * 16 + li;
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int pbo = rt * 16 + lk * 4 + rr;
                    if (pbo < s.npb && g < s.ng) s.ldm32[((size_t)pv * s.npb + pbo) * s.ng + g] = acc[c][rr];
                }
            }
        }
    }
}

// ---- FP32 MFMA contraction of tabulated FP64 profiles, row-contiguous loads ---------
// ldm[pv][pb][g] = sum_z ldp[pv][pb][z] W2T[pv][z][g] for profiles given as an
// FP64 table (the TSModel's LDModel input).  One workgroup of 8 waves per pv with
// W2T in LDS; a wave takes one 16-passband row tile at a time.  The sum over z is
// taken in a permuted order: lane (li, lk) holds row li at z = 16 j + 4 lk + q for
// MFMA step 4 j + q, so each lane loads 4 consecutive FP64 values (two 16-byte
// loads) and the four lanes of a row read 128 contiguous bytes; the B fragment
// reads of LDS follow the same permutation.  The LDS row stride (4 mod 8 floats)
// puts the four z rows of an MFMA step in distinct banks.
typedef double f64x2_t __attribute__((ext_vector_type(2)));
constexpr int kTabMfmaThreads = 512;
constexpr int kTabMfmaMaxJ = 8;    // z groups of 16: nz <= 128
constexpr int kTabMfmaMaxCt = 8;   // g tiles of 16: ngp <= 128

__host__ __device__ inline int ts_tab_ldb(int ngp) { return ngp + ((4 - ngp % 8) + 8) % 8; }

template <int NCT>
__global__ void __launch_bounds__(kTabMfmaThreads) ts_ldm_tab_mfma_kernel(TsArgs s) {
    extern __shared__ float lds_t[];  // [nz16][ldb]
    const int pv = blockIdx.x;
    const int ldb = ts_tab_ldb(s.ngp);
    const int nj = (s.nz + 15) / 16, nz16 = nj * 16;
    for (int i = threadIdx.x; i < nz16 * s.ngp; i += blockDim.x) {
        const int iz = i / s.ngp, ig = i - iz * s.ngp;
        lds_t[iz * ldb + ig] = iz < s.nzp ? s.w2_32[(size_t)pv * s.nzp * s.ngp + i] : 0.0f;
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int nrt = (s.npb + 15) / 16;
    constexpr int kWaves = kTabMfmaThreads / 64;
    for (int rt = wave; rt < nrt; rt += kWaves) {
        const int pb = rt * 16 + li;
        const bool row_ok = pb < s.npb;
        const double* arow = s.ldp_in + ((size_t)pv * s.npb + (row_ok ? pb : 0)) * s.nz;
        f32x4_t acc[NCT];
#pragma unroll
        for (int c = 0; c < NCT; ++c) acc[c] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
        // the row in two halves of 64 z values: 16 FP64 values in flight per lane
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            constexpr int kJh = kTabMfmaMaxJ / 2;
            float af[4 * kJh];
#pragma unroll
            for (int jj = 0; jj < kJh; ++jj) {
                const int j = h * kJh + jj, z0 = 16 * j + 4 * lk;
                f64x2_t v0 = {0.0, 0.0}, v1 = {0.0, 0.0};
                if (j < nj && row_ok && z0 < s.nz) {
                    v0 = *(const f64x2_t*)(arow + z0);
                    v1 = *(const f64x2_t*)(arow + z0 + 2);
                }
                af[4 * jj + 0] = (float)v0.x;
                af[4 * jj + 1] = (float)v0.y;
                af[4 * jj + 2] = (float)v1.x;
                af[4 * jj + 3] = (float)v1.y;
            }
#pragma unroll
            for (int jj = 0; jj < kJh; ++jj) {
                const int j = h * kJh + jj;
                if (j < nj) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const float* brow = lds_t + (16 * j + 4 * lk + q) * ldb + li;
#pragma unroll
                        for (int c = 0; c < NCT; ++c)
                            acc[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[4 * jj + q], brow[c * 16], acc[c], 0, 0,
                                                                          0);
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < NCT; ++c) {
            const int g = c * 16 + li;
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                const int pbo = rt * 16 + lk * 4 + rr;
                if (pbo < s.npb && g < s.ng) s.ldm32[((size_t)pv * s.npb + pbo) * s.ng + g] = acc[c][rr];
            }
        }
    }
}

// ---- FP64 MFMA contraction (v_mfma_f64_16x16x4f64), same tiling as the FP32 one ---------
// Lane l supplies A[l%16][l/16] and B[l/16][l%16] and holds D[(l/16)+4r][l%16]:
// unlike the FP32 16x16x4 form, the four result rows of a lane are interleaved.
// One workgroup of 8 waves per pv; W2T (FP64) sits in LDS with a padded row stride.
typedef double f64x4_t __attribute__((ext_vector_type(4)));
constexpr int kMfma64Threads = 512;
constexpr size_t kMfma64MaxLds = 128 * 1024;

template <bool kTab>
__global__ void __launch_bounds__(kMfma64Threads) ts_ldm_mfma64_kernel(TsArgs s) {
    extern __shared__ double lds_b64[];  // [nzp][ngp + 1]
    const int pv = blockIdx.x;
    const int ldb = s.ngp + 1;
    for (int i = threadIdx.x; i < s.nzp * s.ngp; i += blockDim.x) {
        const int iz = i / s.ngp, ig = i - iz * s.ngp;
        lds_b64[iz * ldb + ig] = s.w2_64[(size_t)pv * s.nzp * s.ngp + i];
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int nrt = (s.npb + 15) / 16, nct = s.ngp / 16;
    constexpr int kWaves = kMfma64Threads / 64;
    for (int rt = wave; rt < nrt; rt += kWaves) {
        const int pb = rt * 16 + li;
        const bool row_ok = pb < s.npb;
        for (int ct0 = 0; ct0 < nct; ct0 += 8) {
            f64x4_t acc[8];
#pragma unroll
            for (int c = 0; c < 8; ++c) acc[c] = (f64x4_t){0.0, 0.0, 0.0, 0.0};
            for (int k0 = 0; k0 < s.nzp; k0 += 4) {
                const int iz = k0 + lk;
                const double a = (row_ok && iz < s.nz) ? ts_profile<kTab>(s, pv, pb, iz) : 0.0;
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const int ct = ct0 + c;
                    if (ct < nct) {
                        const double b = lds_b64[iz * ldb + ct * 16 + li];
                        acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int ct = ct0 + c;
                if (ct >= nct) break;
                const int g = ct * 16 + li;
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int pbo = rt * 16 + lk + 4 * rr;
                    if (pbo < s.npb && g < s.ng) s.ldm64[((size_t)pv * s.npb + pbo) * s.ng + g] = acc[c][rr];
                }
            }
        }
    }
}

// ---- FP64 contraction on the vector ALUs: one lane per (pv, pb, g) --------------------
template <bool kTab>
__global__ void __launch_bounds__(256) ts_ldm_f64_kernel(TsArgs s) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)s.npv * s.npb * s.ng) return;
    const int ig = (int)(idx % s.ng);
    const long long j = idx / s.ng;
    const int pv = (int)(j / s.npb), pb = (int)(j - (long long)pv * s.npb);
    const double* w = s.w2_64 + (size_t)pv * s.nzp * s.ngp + ig;
    double d = 0.0;
    for (int iz = 0; iz < s.nz; ++iz) d += w[(size_t)iz * s.ngp] * ts_profile<kTab>(s, pv, pb, iz);
    s.ldm64[idx] = d;
}

// The FP32 weights (W2T of the mixed path) when W2T does not fit the MFMA kernel's
// LDS: one lane per (pv, passband, g), the sum in FP64, the mean stored in FP32.
template <bool kTab>
__global__ void __launch_bounds__(256) ts_ldm_f32_kernel(TsArgs s) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)s.npv * s.npb * s.ng) return;
    const int ig = (int)(idx % s.ng);
    const long long j = idx / s.ng;
    const int pv = (int)(j / s.npb), pb = (int)(j - (long long)pv * s.npb);
    const float* w = s.w2_32 + (size_t)pv * s.nzp * s.ngp + ig;
    double d = 0.0;
    for (int iz = 0; iz < s.nz; ++iz) d += w[(size_t)iz * s.ngp] * ts_profile<kTab>(s, pv, pb, iz);
    s.ldm32[idx] = (float)d;
}

// ---- flux --------------------------------------------------------------------------
constexpr int kTsThreads = 256;

// Passbands per LDS chunk: 64 (FP32) / 32 (FP64) with one accumulator per lane,
// half that when a lane keeps one accumulator per passband of the chunk.
constexpr int ts_pb_chunk(bool mixed, bool cached) { return cached ? (mixed ? 64 : 32) : (mixed ? 32 : 16); }

// Geometry of one (time, sample): the lens at k_mean in FP64 (model_trspec.py:79-84)
// and the lerp position in g; shared by every passband.
template <typename Real>
struct TsGeo {
    Real ra, rb;      // lerp weights of l[ig + 1] and l[ig]
    Real ap0, dadk;   // area at k_mean and its k-derivative 2 k_mean kappa_0
    int ig;
    bool inside;      // z <= 1 - k_max: area scales as k^2 / k_mean^2
    bool gz;          // g > 1: the limb-darkening mean is 0
};

template <typename Real>
__device__ __forceinline__ TsGeo<Real> ts_geometry(const TsRec& r, double tc, double ext, int ns, int is,
                                                   double dg, int ng) {
    TsGeo<Real> o;
    const double km = r.kmean;
    const double off = ext * ((is + 1 - 0.5) / ns - 0.5);
    const double z = taylor_z<double>(tc + off, r.c);
    double ap0, kap;
    ccia_kite<double>(1.0, km, z, ap0, kap);
    o.inside = z <= 1.0 - r.kmax;
    const double g = z / (1.0 + km);
    // lerp position (common.py:225-233, index clamped to ng-2)
    o.gz = g > 1.0;
    o.ig = 0;
    double al = 0.0;
    if (!o.gz) {
        int ig = (int)floor(g / dg);
        o.ig = ig > ng - 2 ? ng - 2 : (ig < 0 ? 0 : ig);
        al = (g - o.ig * dg) / dg;
    }
    o.ra = (Real)al;
    o.rb = (Real)(1.0 - al);
    o.ap0 = (Real)ap0;
    o.dadk = (Real)(2.0 * km * kap);
    return o;
}

// With nsamples <= kTsGeoRegs the geometry of a lane's samples stays in registers,
// the passbands are the outer loop and a lane owns 16 bytes of consecutive times
// (4 FP32 or 2 FP64), so a wave stores 1 KB of one passband row per instruction.  Otherwise a lane
// owns one time, the samples are the outer loop over a chunk of PC accumulators
// and the geometry is recomputed per chunk.
constexpr int kTsGeoRegs = 2;
// times per lane of the cached form: one 16-byte store per passband row
template <typename Real> constexpr int ts_tpl() { return 16 / (int)sizeof(Real); }

template <typename Real>
__device__ __forceinline__ void ts_store(Real* o, const Real* v, int n, bool vec) {
    if (vec && n == ts_tpl<Real>()) {
        if constexpr (std::is_same<Real, float>::value) {
            const f32x4_t w = {v[0], v[1], v[2], v[3]};
            __builtin_nontemporal_store(w, (f32x4_t*)o);
        } else {
            typedef double f64x2_t __attribute__((ext_vector_type(2)));
            const f64x2_t w = {v[0], v[1]};
            __builtin_nontemporal_store(w, (f64x2_t*)o);
        }
    } else {
        for (int j = 0; j < n; ++j) o[j] = v[j];
    }
}

// ---- FP32 flux, one workgroup per pv, rows written in order ----------------------------
//
// The tile form above has each workgroup write a 4 KB piece of every passband row
// (rows npt apart), and HBM takes such strided pieces at ~4.3 TB/s where it takes
// the same bytes written in address order at ~5.3 TB/s.  With one sample and
// npt <= 8192 (the usual transmission-spectroscopy shape: a few thousand times,
// hundreds of passbands), one 1024-thread workgroup per pv writes the pv's
// [npb][npt] block front to back: thread i owns the 16-byte time vectors i and
// i + 1024 of every row, and each step writes one whole row.
//
// A row costs few operations per value: the geometry of each time is reduced once
// to the LDS offset of its limb-darkening interval and five coefficients, with
//   I = rb l[ig] + ra l[ig+1],  area = c1 afac + (c2 + c3 (k - k_mean)),
//   F = 1 - (I area) / I*
// (c1 = A0 inside z <= 1 - k_max, else c2 = A0 and c3 = 2 k_mean kappa0, the
// serial model's two branches); outside the window, and for g > 1 where both the
// mean and the area vanish, F = 1 is selected, so a NaN row of limb-darkening
// means reaches only the in-window values, as in the reference.  The LD rows of a
// chunk of passbands sit in LDS transposed, passband fastest, so the rows of a
// 16-passband group are read with immediate offsets from one address; the row
// stride is 65 values, so lanes reading different g hit different banks (a stride
// of 64 would put every lane of a row read in one bank).  A wave none of whose
// times is in the window stores ones without arithmetic.
constexpr int kTsRowThreads = 1024;
constexpr int kTsRowPc = 64;      // passbands per staged chunk
constexpr int kTsRowLd = kTsRowPc + 1;   // LDS stride of one g
constexpr int kTsRowQ = 16;       // passbands per unrolled group
// time vectors (16 bytes: 4 FP32 or 2 FP64 times) per thread; npt <= 8192 (FP32), 6144 (FP64)
template <typename Real> constexpr int ts_row_vec() { return sizeof(Real) == 4 ? 2 : 3; }
template <typename Real> constexpr int ts_row_max_npt() { return kTsRowThreads * ts_row_vec<Real>() * ts_tpl<Real>(); }

__host__ __device__ inline size_t ts_rows_lds(int ng, size_t real) { return ((size_t)kTsRowLd * ng + 3 * kTsRowPc) * real; }

template <typename Real>
__global__ void __launch_bounds__(kTsRowThreads) ts_flux_rows_kernel(TsArgs s, Real* flux) {
    constexpr int TPL = ts_tpl<Real>();
    constexpr int NV = ts_row_vec<Real>();
    extern __shared__ __align__(16) unsigned char lds[];
    Real* s_ldm = (Real*)lds;                        // [ng][kTsRowLd]
    Real* s_afac = s_ldm + (size_t)kTsRowLd * s.ng;
    Real* s_kd = s_afac + kTsRowPc;
    Real* s_iist = s_kd + kTsRowPc;

    const int pv = xcd_remap(blockIdx.x, gridDim.x);
    const int npt = s.npt, nv = npt / TPL;
    const TsRec& r = s.rec[pv];
    Real* out = flux + (size_t)pv * s.npb * npt;
    Real vnan[TPL], vone[TPL];
#pragma unroll
    for (int j = 0; j < TPL; ++j) {
        vnan[j] = Real(NAN);
        vone[j] = Real(1);
    }
    if (!r.good) {
        for (size_t i = threadIdx.x; i < (size_t)s.npb * nv; i += kTsRowThreads) ts_store(out + i * TPL, vnan, TPL, true);
        return;
    }
    // the geometry of the thread's times, once, as row coefficients
    const double ext = s.exptimes[0];
    int off[NV][TPL];
    Real rb[NV][TPL], ra[NV][TPL], c1[NV][TPL], c2[NV][TPL], c3[NV][TPL];
    bool use[NV][TPL];   // in the window with g <= 1
    bool act[NV];        // the wave has a time of vector u in the window
#pragma unroll
    for (int u = 0; u < NV; ++u) {
        const int v = threadIdx.x + u * kTsRowThreads;
        bool any = false;
#pragma unroll
        for (int j = 0; j < TPL; ++j) {
            off[u][j] = 0;
            rb[u][j] = ra[u][j] = c1[u][j] = c2[u][j] = c3[u][j] = Real(0);
            use[u][j] = false;
            if (v < nv) {
                double tc;
                {
#pragma clang fp contract(off)
                    const double t = s.times[TPL * v + j];
                    const double t0v = s.t0[pv];
                    const double epoch = floor((t - t0v + 0.5 * r.p) * r.inv_p);
                    const double uu = t - t0v + 0.5 * r.p;
                    const double rem = uu - epoch * r.p;
                    const double ep = epoch + ((rem < 0.0) ? -1.0 : ((rem >= r.p) ? 1.0 : 0.0));
                    tc = t - (t0v + ep * r.p);
                }
                if ((r.bt1 <= tc) && (tc <= r.bt4)) {
                    const TsGeo<Real> g = ts_geometry<Real>(r, tc, ext, 1, 0, s.dg, s.ng);
                    any = true;
                    if (!g.gz) {
                        use[u][j] = true;
                        off[u][j] = g.ig * kTsRowLd;
                        rb[u][j] = g.rb;
                        ra[u][j] = g.ra;
                        c1[u][j] = g.inside ? g.ap0 : Real(0);
                        c2[u][j] = g.inside ? Real(0) : g.ap0;
                        c3[u][j] = g.inside ? Real(0) : g.dadk;
                    }
                }
            }
        }
        act[u] = __ballot(any) != 0ull;
    }
    for (int pb0 = 0; pb0 < s.npb; pb0 += kTsRowPc) {
        const int npc = min(kTsRowPc, s.npb - pb0);
        __syncthreads();  // the previous chunk's rows are no longer read
        for (int i = threadIdx.x; i < kTsRowPc * s.ng; i += kTsRowThreads) {
            const int q = i / s.ng, g = i - q * s.ng;   // coalesced global reads along g
            const size_t src = ((size_t)pv * s.npb + pb0 + q) * s.ng + g;
            s_ldm[g * kTsRowLd + q] = q < npc ? (sizeof(Real) == 4 ? (Real)s.ldm32[src] : (Real)s.ldm64[src]) : Real(0);
        }
        for (int q = threadIdx.x; q < npc; q += kTsRowThreads) {
            const double kq = s.k[(size_t)pv * s.npb + pb0 + q];
            s_afac[q] = (Real)(kq * kq / (r.kmean * r.kmean));
            s_kd[q] = (Real)(kq - r.kmean);
            s_iist[q] = (Real)(1.0 / s.istar[(size_t)pv * s.npb + pb0 + q]);
        }
        __syncthreads();
#pragma unroll 1
        for (int q0 = 0; q0 < npc; q0 += kTsRowQ) {
            const Real* lq = s_ldm + q0;
#pragma unroll
            for (int qq = 0; qq < kTsRowQ; ++qq) {
                const int q = q0 + qq;
                if (q < npc) {
                    const Real f = s_iist[q], afac = s_afac[q], kd = s_kd[q];
                    Real* o = out + (size_t)(pb0 + q) * npt;
#pragma unroll
                    for (int u = 0; u < NV; ++u) {
                        const int v = threadIdx.x + u * kTsRowThreads;
                        if (v < nv) {
                            if (act[u]) {
                                Real w[TPL];
#pragma unroll
                                for (int j = 0; j < TPL; ++j) {
                                    const Real* l = lq + off[u][j] + qq;
                                    const Real ip = rb[u][j] * l[0] + ra[u][j] * l[kTsRowLd];
                                    const Real area = c1[u][j] * afac + (c2[u][j] + c3[u][j] * kd);
                                    w[j] = use[u][j] ? Real(1) - (ip * area) * f : Real(1);
                                }
                                ts_store(o + (size_t)v * TPL, w, TPL, true);
                            } else {
                                ts_store(o + (size_t)v * TPL, vone, TPL, true);
                            }
                        }
                    }
                }
            }
        }
    }
}

// ---- FP32 rows in whole cache lines ----------------------------------------------------
//
// A row of npt FP32 values starts 4 (R npt mod 32) bytes past a 128-byte boundary
// (R the row's index in the flux array).  In the row form above, thread i owns the
// same time vectors in every row, so in a row that starts off a boundary every
// wave's 1 KB store straddles cache lines, and HBM takes those pieces at ~4.0
// TB/s against ~4.9 TB/s for whole lines (the same kernel at npt = 5000 and 4992).
// Here the geometry of every time is reduced once to its row coefficients in an
// LDS table, one array per coefficient, and in each row slot s takes time vector
// (s + head) mod nv, head being the number of 16-byte vectors before the row's
// first 128-byte boundary: the 64 slots of a wave then cover one aligned
// kilobyte.  Rows P = 32 / gcd(npt, 32) apart have the same head and share one
// load of the coefficients from the table into registers; the per-value
// arithmetic is the row form's.
constexpr int kTsLine = 32;   // FP32 values per 128-byte cache line
constexpr int kTsStageRegs = 8;   // LD values per thread of a staged chunk: kTsRowPc x ng <= 8 x 1024
constexpr size_t kTsRowsAlignedMaxLds = 160 * 1024;

__host__ __device__ inline size_t ts_rows_table_offset(int ng) { return (ts_rows_lds(ng, 4) + 15) / 16 * 16; }
__host__ __device__ inline size_t ts_rows_aligned_lds(int ng, int npt) { return ts_rows_table_offset(ng) + (size_t)npt * 5 * 4; }

__global__ void __launch_bounds__(kTsRowThreads) ts_flux_rows_aligned_kernel(TsArgs s, float* flux) {
    using Real = float;
    constexpr int TPL = ts_tpl<float>();
    constexpr int NV = ts_row_vec<float>();
    extern __shared__ __align__(16) unsigned char lds[];
    Real* s_ldm = (Real*)lds;                        // [ng][kTsRowLd]
    Real* s_afac = s_ldm + (size_t)kTsRowLd * s.ng;
    Real* s_kd = s_afac + kTsRowPc;
    Real* s_iist = s_kd + kTsRowPc;
    const int npt = s.npt, nv = npt / TPL;
    // the row coefficients of every time: LDS offset << 2 | inside << 1 | use, and
    // the lerp weights, the area at k_mean and its k-derivative
    int* t_pk = (int*)(lds + ts_rows_table_offset(s.ng));
    Real* t_ra = (Real*)(t_pk + npt);
    Real* t_rb = t_ra + npt;
    Real* t_a0 = t_rb + npt;
    Real* t_dk = t_a0 + npt;

    const int pv = xcd_remap(blockIdx.x, gridDim.x);
    const TsRec& r = s.rec[pv];
    float* out = flux + (size_t)pv * s.npb * npt;
    if (!r.good) {
        const f32x4_t vnan = {NAN, NAN, NAN, NAN};
        for (size_t i = threadIdx.x; i < (size_t)s.npb * nv; i += kTsRowThreads)
            __builtin_nontemporal_store(vnan, (f32x4_t*)(out + i * TPL));
        return;
    }
    // a chunk's LD rows and per-passband k and disk integral, loaded into registers
    // one chunk ahead (every load of a chunk in flight at once, the wait for them
    // after the previous chunk's rows) and stored to LDS after the chunk barrier
    float st[kTsStageRegs];
    double kq = 0.0, isq = 1.0;
    auto stage_load = [&](int pb0) {
        const int npc = min(kTsRowPc, s.npb - pb0);
#pragma unroll
        for (int m = 0; m < kTsStageRegs; ++m) {
            const int i = threadIdx.x + m * kTsRowThreads;
            const int q = i / s.ng, g = i - q * s.ng;
            st[m] = (i < kTsRowPc * s.ng && q < npc) ? s.ldm32[((size_t)pv * s.npb + pb0 + q) * s.ng + g] : 0.0f;
        }
        if ((int)threadIdx.x < npc) {
            kq = s.k[(size_t)pv * s.npb + pb0 + threadIdx.x];
            isq = s.istar[(size_t)pv * s.npb + pb0 + threadIdx.x];
        }
    };
    auto stage_store = [&](int pb0) {
        const int npc = min(kTsRowPc, s.npb - pb0);
#pragma unroll
        for (int m = 0; m < kTsStageRegs; ++m) {
            const int i = threadIdx.x + m * kTsRowThreads;
            const int q = i / s.ng, g = i - q * s.ng;
            if (i < kTsRowPc * s.ng) s_ldm[g * kTsRowLd + q] = st[m];
        }
        if ((int)threadIdx.x < npc) {
            s_afac[threadIdx.x] = (Real)(kq * kq / (r.kmean * r.kmean));
            s_kd[threadIdx.x] = (Real)(kq - r.kmean);
            s_iist[threadIdx.x] = (Real)(1.0 / isq);
        }
    };
    const double ext = s.exptimes[0];
    const double t0v = s.t0[pv];
    for (int t = threadIdx.x; t < npt; t += kTsRowThreads) {
        double tc;
        {
#pragma clang fp contract(off)
            const double tt = s.times[t];
            const double epoch = floor((tt - t0v + 0.5 * r.p) * r.inv_p);
            const double uu = tt - t0v + 0.5 * r.p;
            const double rem = uu - epoch * r.p;
            const double ep = epoch + ((rem < 0.0) ? -1.0 : ((rem >= r.p) ? 1.0 : 0.0));
            tc = tt - (t0v + ep * r.p);
        }
        int pk = 0;
        Real ra = Real(0), rb = Real(0), a0 = Real(0), dk = Real(0);
        if ((r.bt1 <= tc) && (tc <= r.bt4)) {
            const TsGeo<Real> g = ts_geometry<Real>(r, tc, ext, 1, 0, s.dg, s.ng);
            if (!g.gz) {
                pk = ((g.ig * kTsRowLd) << 2) | (g.inside ? 2 : 0) | 1;
                ra = g.ra;
                rb = g.rb;
                a0 = g.ap0;
                dk = g.dadk;
            }
        }
        t_pk[t] = pk;
        t_ra[t] = ra;
        t_rb[t] = rb;
        t_a0[t] = a0;
        t_dk[t] = dk;
    }
    stage_load(0);
    // rows P apart start equally far past a 128-byte boundary
    const int gl = npt & -npt;   // lowest set bit of npt
    const int P = kTsLine / (gl < kTsLine ? gl : kTsLine);
    for (int pb0 = 0; pb0 < s.npb; pb0 += kTsRowPc) {
        const int npc = min(kTsRowPc, s.npb - pb0);
        __syncthreads();  // the previous chunk's rows are no longer read (and the table is written)
        stage_store(pb0);
        __syncthreads();
        if (pb0 + kTsRowPc < s.npb) stage_load(pb0 + kTsRowPc);
        for (int q0 = pb0; q0 < pb0 + min(P, npc); ++q0) {
            // this class's head, and each slot's coefficients from the table
            const int mis = (int)(((uintptr_t)(out + (size_t)q0 * npt) >> 2) & (kTsLine - 1));
            const int head = ((kTsLine - mis) & (kTsLine - 1)) / TPL;
            // per value: LDS offset, lerp weights, the area at k_mean (negated when
            // the planet is inside the disk, where it scales by afac) and its
            // k-derivative (0 inside); bit u * TPL + j of `use` marks a value in the
            // window with g <= 1
            int vi[NV], off[NV][TPL];
            Real rb[NV][TPL], ra[NV][TPL], as[NV][TPL], dk[NV][TPL];
            unsigned use = 0u;
            bool act[NV];
#pragma unroll
            for (int u = 0; u < NV; ++u) {
                const int slot = threadIdx.x + u * kTsRowThreads;
                vi[u] = slot + head < nv ? slot + head : slot + head - nv;
                bool any = false;
#pragma unroll
                for (int j = 0; j < TPL; ++j) {
                    off[u][j] = 0;
                    ra[u][j] = rb[u][j] = as[u][j] = dk[u][j] = Real(0);
                }
                if (slot < nv) {
                    const int4 pk4 = *(const int4*)(t_pk + TPL * vi[u]);
                    const f32x4_t ra4 = *(const f32x4_t*)(t_ra + TPL * vi[u]);
                    const f32x4_t rb4 = *(const f32x4_t*)(t_rb + TPL * vi[u]);
                    const f32x4_t a04 = *(const f32x4_t*)(t_a0 + TPL * vi[u]);
                    const f32x4_t dk4 = *(const f32x4_t*)(t_dk + TPL * vi[u]);
                    const int pk[TPL] = {pk4.x, pk4.y, pk4.z, pk4.w};
#pragma unroll
                    for (int j = 0; j < TPL; ++j) {
                        const bool inside = (pk[j] & 2) != 0;
                        if (pk[j] & 1) use |= 1u << (u * TPL + j);
                        any = any || (pk[j] & 1) != 0;
                        off[u][j] = pk[j] >> 2;
                        ra[u][j] = ra4[j];
                        rb[u][j] = rb4[j];
                        as[u][j] = inside ? -a04[j] : a04[j];
                        dk[u][j] = inside ? Real(0) : dk4[j];
                    }
                }
                act[u] = __ballot(any) != 0ull;
            }
#pragma unroll 1
            for (int q = q0; q < pb0 + npc; q += P) {
                const int ql = q - pb0;
                const Real f = s_iist[ql], afac = s_afac[ql], kd = s_kd[ql];
                const Real* lq = s_ldm + ql;
                float* o = out + (size_t)q * npt;
#pragma unroll
                for (int u = 0; u < NV; ++u) {
                    const int slot = threadIdx.x + u * kTsRowThreads;
                    if (slot < nv) {
                        f32x4_t w = {1.0f, 1.0f, 1.0f, 1.0f};
                        if (act[u]) {
#pragma unroll
                            for (int j = 0; j < TPL; ++j) {
                                const Real* l = lq + off[u][j];
                                const Real ip = rb[u][j] * l[0] + ra[u][j] * l[kTsRowLd];
                                // the row form's c1 afac + (c2 + c3 (k - k_mean)), with c1 = A0
                                // inside and c2 = A0, c3 = 2 k_mean kappa_0 otherwise
                                const Real area = as[u][j] < Real(0) ? (-as[u][j]) * afac : dk[u][j] * kd + as[u][j];
                                w[j] = ((use >> (u * TPL + j)) & 1u) ? Real(1) - (ip * area) * f : Real(1);
                            }
                        }
                        __builtin_nontemporal_store(w, (f32x4_t*)(o + (size_t)vi[u] * TPL));
                    }
                }
            }
        }
    }
}

template <bool kMixed, bool kCached>
__global__ void __launch_bounds__(kTsThreads) ts_flux_kernel(TsArgs s, void* flux, int ntiles) {
    using Real = typename std::conditional<kMixed, float, double>::type;
    constexpr int PC = ts_pb_chunk(kMixed, kCached);
    constexpr int TPL = kCached ? ts_tpl<Real>() : 1;
    extern __shared__ __align__(16) unsigned char lds[];
    Real* s_ldm = (Real*)lds;                       // [PC][ng]
    Real* s_afac = s_ldm + (size_t)PC * s.ng;       // k^2 / k_mean^2
    Real* s_kd = s_afac + PC;                       // k - k_mean
    Real* s_iist = s_kd + PC;                       // 1 / istar
    __shared__ int s_any;

    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int tile = bid % ntiles;
    const int pv = bid / ntiles;
    const int t0l = (tile * kTsThreads + (int)threadIdx.x) * TPL;  // first time of the lane
    const int nt = max(0, min(TPL, s.npt - t0l));                   // its valid times
    const TsRec& r = s.rec[pv];
    Real* out = (Real*)flux + (size_t)pv * s.npb * s.npt;
    // whole-vector stores need every row start 16-B aligned
    const bool vec = TPL > 1 && (s.npt % TPL) == 0 && (((uintptr_t)out) % 16) == 0;
    Real ones[TPL], v[TPL];
#pragma unroll
    for (int j = 0; j < TPL; ++j) ones[j] = Real(1);
    if (!r.good) {
#pragma unroll
        for (int j = 0; j < TPL; ++j) v[j] = Real(NAN);
        if (nt > 0)
            for (int q = 0; q < s.npb; ++q) ts_store(out + (size_t)q * s.npt + t0l, v, nt, vec);
        return;
    }
    double tc[TPL];
    bool inw[TPL];
    bool any = false;
#pragma unroll
    for (int j = 0; j < TPL; ++j) {
        tc[j] = 0.0;
        inw[j] = false;
        if (j < nt) {
#pragma clang fp contract(off)
            const double t = s.times[t0l + j];
            const double t0v = s.t0[pv];
            const double epoch = floor((t - t0v + 0.5 * r.p) * r.inv_p);
            const double u = t - t0v + 0.5 * r.p;
            const double rem = u - epoch * r.p;
            const double ep = epoch + ((rem < 0.0) ? -1.0 : ((rem >= r.p) ? 1.0 : 0.0));
            tc[j] = t - (t0v + ep * r.p);
            inw[j] = (r.bt1 <= tc[j]) && (tc[j] <= r.bt4);
            any = any || inw[j];
        }
    }
    if (threadIdx.x == 0) s_any = 0;
    __syncthreads();
    if (any) s_any = 1;
    __syncthreads();
    if (!s_any) {
        if (nt > 0)
            for (int q = 0; q < s.npb; ++q) ts_store(out + (size_t)q * s.npt + t0l, ones, nt, vec);
        return;
    }
    const int ns = s.nsamples[0];
    const double ext = s.exptimes[0];
    const Real inv_ns = Real(1) / Real(ns);
    const double km = r.kmean;
    TsGeo<Real> geo[TPL][kCached ? kTsGeoRegs : 1];
    if (kCached) {
#pragma unroll
        for (int j = 0; j < TPL; ++j)
#pragma unroll
            for (int is = 0; is < kTsGeoRegs; ++is)
                if (inw[j] && is < ns) geo[j][is] = ts_geometry<Real>(r, tc[j], ext, ns, is, s.dg, s.ng);
    }

    for (int pb0 = 0; pb0 < s.npb; pb0 += PC) {
        const int npc = min(PC, s.npb - pb0);
        __syncthreads();  // the previous chunk's rows are no longer read
        // stage the chunk's limb-darkening means and per-passband factors
        for (int i = threadIdx.x; i < npc * s.ng; i += kTsThreads) {
            const size_t src = ((size_t)pv * s.npb + pb0) * s.ng + i;
            s_ldm[i] = kMixed ? (Real)s.ldm32[src] : (Real)s.ldm64[src];
        }
        for (int q = threadIdx.x; q < npc; q += kTsThreads) {
            const double kq = s.k[(size_t)pv * s.npb + pb0 + q];
            s_afac[q] = (Real)(kq * kq / (km * km));
            s_kd[q] = (Real)(kq - km);
            s_iist[q] = (Real)(1.0 / s.istar[(size_t)pv * s.npb + pb0 + q]);
        }
        __syncthreads();
        if (nt == 0) continue;
        Real* o = out + (size_t)pb0 * s.npt + t0l;
        // depth of passband q for one sample; the flux is 1 - mean(depth) / istar
        auto depth = [&](const TsGeo<Real>& g, int q) {
            const Real* l = s_ldm + (size_t)q * s.ng;
            const Real ip = g.gz ? Real(0) : (g.rb * l[g.ig] + g.ra * l[g.ig + 1]);
            const Real area = g.inside ? g.ap0 * s_afac[q] : g.ap0 + s_kd[q] * g.dadk;
            return ip * area;
        };
        if (kCached) {
#pragma unroll 2
            for (int q = 0; q < npc; ++q) {
                const Real f = s_iist[q] * inv_ns;
#pragma unroll
                for (int j = 0; j < TPL; ++j) {
                    Real acc = Real(0);
                    if (inw[j]) {
                        acc = depth(geo[j][0], q);
                        if (ns > 1) acc += depth(geo[j][kTsGeoRegs - 1], q);
                    }
                    v[j] = inw[j] ? Real(1) - acc * f : Real(1);
                }
                ts_store(o + (size_t)q * s.npt, v, nt, vec);
            }
        } else {
            Real acc[PC];
#pragma unroll
            for (int q = 0; q < PC; ++q) acc[q] = Real(0);
            if (inw[0]) {
                for (int is = 0; is < ns; ++is) {
                    const TsGeo<Real> g = ts_geometry<Real>(r, tc[0], ext, ns, is, s.dg, s.ng);
#pragma unroll
                    for (int q = 0; q < PC; ++q)
                        if (q < npc) acc[q] += depth(g, q);
                }
            }
#pragma unroll
            for (int q = 0; q < PC; ++q)
                if (q < npc) o[(size_t)q * s.npt] = inw[0] ? Real(1) - acc[q] * (s_iist[q] * inv_ns) : Real(1);
        }
    }
}

}  // namespace rr

using namespace rr;

extern "C" {

static size_t ts_sizes(int npv, int npb, int ng, int nz, int precision, size_t* off) {
    const int ngp = (ng + 15) / 16 * 16, nzp = (nz + 3) / 4 * 4;
    size_t b = 0;
    off[0] = b; b += align_up((size_t)npv * sizeof(TsRec), 256);
    off[1] = b; b += align_up((size_t)npv * npb * sizeof(double), 256);
    const size_t w2e = (size_t)npv * nzp * ngp;
    off[2] = b; b += align_up(w2e * (precision == RR_MIXED ? sizeof(float) : sizeof(double)), 256);
    off[3] = b; b += align_up((size_t)npv * npb * ng * (precision == RR_MIXED ? sizeof(float) : sizeof(double)), 256);
    return b;
}

size_t ts_workspace_bytes(int npv, int npb, int ng, int nz, int precision) {
    if (npv < 0 || npb < 0 || ng < 0 || nz < 0) return 0;
    size_t off[4];
    return ts_sizes(npv, npb, ng, nz, precision, off);
}

int ts_evaluate(int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv, const rr_ld* ld,
                void* workspace, size_t workspace_bytes, void* flux, void* stream) {
    if (!tab || !lc || !pv || !ld) return set_error(RR_EINVAL, "ts_evaluate: null descriptor");
    if (precision != RR_FP64 && precision != RR_MIXED) return set_error(RR_EINVAL, "ts_evaluate: bad precision");
    const int npv = pv->npv, npt = lc->npt, npb = lc->npb;
    if (npv < 0 || npt < 0 || npb < 1) return set_error(RR_EINVAL, "ts_evaluate: bad sizes");
    if (pv->kcols != npb)
        return set_error(RR_EINVAL, "The transmission spectrum transit model requires that the number or radius "
                                    "ratios and the number of passbands match.");
    if (tab->ng < 2 || tab->nz < 1 || tab->nz > 1024 || tab->ng > 2048 || !tab->ze || !tab->mu)
        return set_error(RR_EINVAL, "ts_evaluate: bad z grid / ng");
    if (npv == 0 || npt == 0) return RR_OK;
    if (!lc->times || !lc->nsamples || !lc->exptimes) return set_error(RR_EINVAL, "ts_evaluate: null time arrays");
    if (!pv->k || !pv->t0 || !pv->p || !pv->a || !pv->inc || !pv->e || !pv->w || !flux)
        return set_error(RR_EINVAL, "ts_evaluate: null parameter array");
    if (ld->law == RR_LD_TABULATED) {
        if (!ld->ldp || !ld->istar) return set_error(RR_EINVAL, "ts_evaluate: tabulated LD needs ldp and istar");
    } else if (ld->law < RR_LD_UNIFORM || ld->law > RR_LD_POWER_2_PM || (ld->nldc > 0 && !ld->ldc)) {
        return set_error(RR_EINVAL, "ts_evaluate: bad limb-darkening descriptor");
    }
    size_t off[4];
    const size_t need = ts_sizes(npv, npb, tab->ng, tab->nz, precision, off);
    if (!workspace || workspace_bytes < need)
        return set_error(RR_EINVAL, "ts_evaluate: workspace of %zu bytes needed, got %zu", need, workspace_bytes);
    if ((long long)npv * npb * npt > (1LL << 40)) return set_error(RR_ERANGE, "ts_evaluate: problem too large");

    char* ws = (char*)workspace;
    const bool mixed = precision == RR_MIXED;
    TsArgs a;
    a.times = lc->times; a.k = pv->k; a.t0 = pv->t0;
    a.exptimes = lc->exptimes; a.nsamples = lc->nsamples;
    a.p = pv->p; a.a = pv->a; a.inc = pv->inc; a.e = pv->e; a.w = pv->w;
    a.law = ld->law; a.nldc = ld->nldc; a.ldc = ld->ldc; a.ldp_in = ld->ldp; a.istar_in = ld->istar;
    a.W = tab->weights; a.ze = tab->ze; a.mu = tab->mu;
    a.rec = (TsRec*)(ws + off[0]);
    a.istar = (double*)(ws + off[1]);
    a.w2_64 = mixed ? nullptr : (double*)(ws + off[2]);
    a.w2_32 = mixed ? (float*)(ws + off[2]) : nullptr;
    a.ldm64 = mixed ? nullptr : (double*)(ws + off[3]);
    a.ldm32 = mixed ? (float*)(ws + off[3]) : nullptr;
    a.npv = npv; a.npb = npb; a.npt = npt; a.nk = tab->nk; a.ng = tab->ng; a.nz = tab->nz;
    a.ngp = (tab->ng + 15) / 16 * 16; a.nzp = (tab->nz + 3) / 4 * 4;
    a.use_table = tab->weights != nullptr && tab->nk >= 2;
    a.kmin = tab->kmin; a.kmax_tab = tab->kmax; a.dk = tab->dk; a.dg = tab->dg;
    hipStream_t st = (hipStream_t)stream;
    const bool tab_ld = ld->law == RR_LD_TABULATED;
    // the weight kernel lists (g, z) pairs of a workgroup as 16-bit indices
    if ((size_t)ts_weight_rows(a.nz) * a.nz > 65535u) return set_error(RR_ERANGE, "ts_evaluate: nz must be <= 65535");

    hipLaunchKernelGGL(ts_setup_kernel, dim3((npv + kTsGroup - 1) / kTsGroup), dim3(64), 0, st, a);
    int rc = check_launch("ts_setup_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(ts_istar_kernel, dim3((npv * npb + 255) / 256), dim3(256), 0, st, a);
    if ((rc = check_launch("ts_istar_kernel"))) return rc;
    const int wrows = ts_weight_rows(a.nz);
    const long long nw = (long long)npv * ((a.ngp + wrows - 1) / wrows);
    hipLaunchKernelGGL(ts_weights_kernel, dim3((unsigned)nw), dim3(256), ts_weights_lds(a.nz), st, a);
    if ((rc = check_launch("ts_weights_kernel"))) return rc;
    // tabulated FP64 profiles with nz a multiple of 4 (16-byte aligned rows), nz <= 128
    // and ngp <= 128: the row-contiguous kernel
    const int nct = a.ngp / 16;
    const bool tab_rows = mixed && tab_ld && a.nz % 4 == 0 && a.nz <= 16 * kTabMfmaMaxJ && nct <= kTabMfmaMaxCt &&
                          ((uintptr_t)ld->ldp & 15) == 0;
    if (tab_rows) {
        const size_t lds = (size_t)((a.nz + 15) / 16 * 16) * ts_tab_ldb(a.ngp) * sizeof(float);
        const dim3 g(npv), b(kTabMfmaThreads);
        switch (nct) {
            case 1: hipLaunchKernelGGL(ts_ldm_tab_mfma_kernel<1>, g, b, lds, st, a); break;
            case 2: hipLaunchKernelGGL(ts_ldm_tab_mfma_kernel<2>, g, b, lds, st, a); break;
            case 3: hipLaunchKernelGGL(ts_ldm_tab_mfma_kernel<3>, g, b, lds, st, a); break;
            case 4: hipLaunchKernelGGL(ts_ldm_tab_mfma_kernel<4>, g, b, lds, st, a); break;
            case 5: hipLaunchKernelGGL(ts_ldm_tab_mfma_kernel<5>, g, b, lds, st, a); break;
            case 6: hipLaunchKernelGGL(ts_ldm_tab_mfma_kernel<6>, g, b, lds, st, a); break;
            case 7: hipLaunchKernelGGL(ts_ldm_tab_mfma_kernel<7>, g, b, lds, st, a); break;
            default: hipLaunchKernelGGL(ts_ldm_tab_mfma_kernel<8>, g, b, lds, st, a); break;
        }
        if ((rc = check_launch("ts_ldm_tab_mfma_kernel"))) return rc;
    } else if (mixed) {
        const size_t lds = (size_t)a.nzp * (a.ngp + 1) * sizeof(float);
        if (lds <= 160 * 1024) {
            if (tab_ld) hipLaunchKernelGGL(ts_ldm_mfma_kernel<true>, dim3(npv), dim3(256), lds, st, a);
            else hipLaunchKernelGGL(ts_ldm_mfma_kernel<false>, dim3(npv), dim3(256), lds, st, a);
            if ((rc = check_launch("ts_ldm_mfma_kernel"))) return rc;
        } else {
            const long long n = (long long)npv * npb * a.ng;
            const dim3 g((unsigned)((n + 255) / 256));
            if (tab_ld) hipLaunchKernelGGL(ts_ldm_f32_kernel<true>, g, dim3(256), 0, st, a);
            else hipLaunchKernelGGL(ts_ldm_f32_kernel<false>, g, dim3(256), 0, st, a);
            if ((rc = check_launch("ts_ldm_f32_kernel"))) return rc;
        }
    } else {
        const size_t lds = (size_t)a.nzp * (a.ngp + 1) * sizeof(double);
        if (lds <= kMfma64MaxLds) {
            if (tab_ld) hipLaunchKernelGGL(ts_ldm_mfma64_kernel<true>, dim3(npv), dim3(kMfma64Threads), lds, st, a);
            else hipLaunchKernelGGL(ts_ldm_mfma64_kernel<false>, dim3(npv), dim3(kMfma64Threads), lds, st, a);
            if ((rc = check_launch("ts_ldm_mfma64_kernel"))) return rc;
        } else {
            const long long n = (long long)npv * npb * a.ng;
            const dim3 g((unsigned)((n + 255) / 256));
            if (tab_ld) hipLaunchKernelGGL(ts_ldm_f64_kernel<true>, g, dim3(256), 0, st, a);
            else hipLaunchKernelGGL(ts_ldm_f64_kernel<false>, g, dim3(256), 0, st, a);
            if ((rc = check_launch("ts_ldm_f64_kernel"))) return rc;
        }
    }
    // nsamples[0] as given by the host (0: unknown) selects the register-cached form
    const bool cached = lc->uniform_nsamples >= 1 && lc->uniform_nsamples <= kTsGeoRegs;
    const int per_tile = kTsThreads * (cached ? (mixed ? ts_tpl<float>() : ts_tpl<double>()) : 1);
    const int ntiles = (npt + per_tile - 1) / per_tile;
    const long long nb = (long long)npv * ntiles;
    if (nb > 0x7fffffffLL) return set_error(RR_ERANGE, "ts_evaluate: grid too large");
    const dim3 grid((unsigned)nb), block(kTsThreads);
    const int uns = lc->uniform_nsamples;
    // one sample, rows of whole 16-byte vectors: the row form
    if (uns == 1 && ((uintptr_t)flux % 16) == 0) {
        if (mixed && npt % 4 == 0 && npt <= ts_row_max_npt<float>() && kTsRowPc * a.ng <= kTsStageRegs * kTsRowThreads &&
            ts_rows_aligned_lds(a.ng, npt) <= kTsRowsAlignedMaxLds) {
            hipLaunchKernelGGL(ts_flux_rows_aligned_kernel, dim3((unsigned)npv), dim3(kTsRowThreads),
                               ts_rows_aligned_lds(a.ng, npt), st, a, (float*)flux);
            return check_launch("ts_flux_rows_aligned_kernel");
        }
        if (mixed && npt % 4 == 0 && npt <= ts_row_max_npt<float>() && ts_rows_lds(a.ng, 4) <= 64 * 1024) {
            hipLaunchKernelGGL(ts_flux_rows_kernel<float>, dim3((unsigned)npv), dim3(kTsRowThreads),
                               ts_rows_lds(a.ng, 4), st, a, (float*)flux);
            return check_launch("ts_flux_rows_kernel");
        }
        if (!mixed && npt % 2 == 0 && npt <= ts_row_max_npt<double>() && ts_rows_lds(a.ng, 8) <= 64 * 1024) {
            hipLaunchKernelGGL(ts_flux_rows_kernel<double>, dim3((unsigned)npv), dim3(kTsRowThreads),
                               ts_rows_lds(a.ng, 8), st, a, (double*)flux);
            return check_launch("ts_flux_rows_kernel");
        }
    }
    if (mixed) {
        const int PC = ts_pb_chunk(true, cached);
        const size_t lds = ((size_t)PC * a.ng + 3 * PC) * sizeof(float);
        if (cached) hipLaunchKernelGGL((ts_flux_kernel<true, true>), grid, block, lds, st, a, flux, ntiles);
        else hipLaunchKernelGGL((ts_flux_kernel<true, false>), grid, block, lds, st, a, flux, ntiles);
    } else {
        const int PC = ts_pb_chunk(false, cached);
        const size_t lds = ((size_t)PC * a.ng + 3 * PC) * sizeof(double);
        if (cached) hipLaunchKernelGGL((ts_flux_kernel<false, true>), grid, block, lds, st, a, flux, ntiles);
        else hipLaunchKernelGGL((ts_flux_kernel<false, false>), grid, block, lds, st, a, flux, ntiles);
    }
    return check_launch("ts_flux_kernel");
}

}  // extern "C"
