/*
 * rr_hip.h -- C-ABI of the MI355X RoadRunner transit-model library (librr_hip.so).
 *
 * Drop-in boundary for PyTransit's RoadRunner hot path.  Every entry point takes
 * plain pointers and sizes, caller-owned DEVICE buffers (row-major, C order) and a
 * hipStream_t passed as `void*` (NULL = the default stream).  No exceptions cross
 * the ABI: each call returns RR_OK (0) or a negative error code, and
 * rr_last_error() returns the text of the calling thread's last error.  The
 * library keeps no state between calls apart from what the caller passes in (and
 * the per-GPU scratch array of rr_evaluate_noalloc), so calls on different
 * streams are independent (rr_evaluate_noalloc's excepted).
 *
 * Reference citations are to PyTransit 2.8.1 (paths relative to its checkout).
 */
#ifndef RR_HIP_H
#define RR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RR_OK 0
#define RR_EINVAL (-1)   /* bad argument (shape, size, null pointer)          */
#define RR_EHIP (-2)     /* a HIP runtime call or kernel launch failed         */
#define RR_ERANGE (-3)   /* a size exceeds a compiled-in limit                 */

/* Arithmetic of the flux stage.  RR_FP64: every step in FP64, flux is double*.
 * RR_MIXED: FP64 epoch/phase reduction and per-pv setup, FP32 z, area,
 * limb-darkening interpolation and accumulation; flux is float*. */
#define RR_FP64 0
#define RR_MIXED 1

/* Limb-darkening laws; the string keys of RoadRunnerModel.ldmodels
 * (pytransit/models/roadrunner/rrmodel.py:48-58, laws in models/numba/ldmodels.py:22-139). */
#define RR_LD_UNIFORM 0
#define RR_LD_LINEAR 1
#define RR_LD_QUADRATIC 2
#define RR_LD_QUADRATIC_TRI 3
#define RR_LD_NONLINEAR 4
#define RR_LD_GENERAL 5
#define RR_LD_SQUARE_ROOT 6
#define RR_LD_LOGARITHMIC 7
#define RR_LD_EXPONENTIAL 8
#define RR_LD_POWER_2 9
#define RR_LD_POWER_2_PM 10

/* Limb-darkening weight table (common.py:188-223), device-resident.  The device
 * layout is [nk, nz, ng] (g fastest, so lanes indexed by g read it coalesced);
 * the reference's array is weights[ik, ig, iz]. */
typedef struct rr_table {
    int32_t nk, ng, nz;
    int32_t pad_;
    double kmin, kmax;
    double dk;               /* (kmax-kmin)/nk, the reference's indexing step (Q1)   */
    double dg;               /* gs[1]-gs[0] of linspace(0, 1-1e-7, ng)               */
    const double* weights;   /* [nk, nz, ng] FP64; NULL selects exact 2-D weights    */
    const float* weights32;  /* [nk, nz, ng] FP32 copy used by RR_MIXED; nullable    */
    const double* ze;        /* [nz] annulus edges                                   */
    const double* mu;        /* [nz] mu at the annulus means (rrmodel.py:172)        */
} rr_table;

/* Limb darkening of one evaluate() call.  law >= 0: a named law (RR_LD_*) with
 * coefficients ldc[npv, npb, nldc], evaluated on the device at tab->mu.
 * law == RR_LD_TABULATED: profiles ldp[npv, npb, nz] and integrals istar[npv, npb]
 * given by the caller (the LDModel protocol, ldmodel.py:21-39). */
#define RR_LD_TABULATED (-1)
typedef struct rr_ld {
    int32_t law;
    int32_t nldc;
    const double* ldc;
    const double* ldp;
    const double* istar;
} rr_ld;

/* Layout bits of rr_lcdata.sorted.  RR_SORTED_RUNS: lcids are non-decreasing and the
 * times of each light curve are non-decreasing (enables run skipping).  With it,
 * RR_SORTED_TIMES: the times are non-decreasing over the whole array, and
 * RR_ONE_WINDOW: every light curve has the same epoch index and exposure time, so a
 * pv has one transit window for all of them.  0 is always valid. */
#define RR_SORTED_RUNS 1
#define RR_SORTED_TIMES 2
#define RR_ONE_WINDOW 4

/* Light-curve data set by set_data() (transitmodel.py:56-125), device-resident. */
typedef struct rr_lcdata {
    int32_t npt, nlc, npb, nep;
    int32_t sorted;           /* layout bits, RR_SORTED_*: see below                  */
    int32_t uniform_nsamples; /* the common nsamples of all light curves, or 0         */
    const double* times;      /* [npt]                                               */
    const int32_t* lcids;     /* [npt], values in [0, nlc)                           */
    const int32_t* pbids;     /* [nlc], values in [0, npb)                           */
    const int32_t* epids;     /* [nlc], values in [0, nep)                           */
    const int32_t* nsamples;  /* [nlc], >= 1                                         */
    const double* exptimes;   /* [nlc]                                               */
} rr_lcdata;

/* Model parameters of evaluate(k, ldc, t0, p, a, i, e, w), device-resident. */
typedef struct rr_params {
    int32_t npv;
    int32_t kcols;            /* columns of k: 1 or npb                              */
    int32_t t0cols;           /* columns of t0: nep                                  */
    int32_t pad_;
    const double* k;          /* [npv, kcols]                                        */
    const double* t0;         /* [npv, t0cols]                                       */
    const double* p;          /* [npv]                                               */
    const double* a;          /* [npv]                                               */
    const double* inc;        /* [npv]                                               */
    const double* e;          /* [npv]                                               */
    const double* w;          /* [npv]                                               */
} rr_params;

/* One model evaluation of a batch: the arguments of rr_evaluate for one request. */
typedef struct rr_request {
    rr_lcdata lc;
    rr_params pv;
    rr_ld ld;
    void* workspace;          /* rr_workspace_bytes(...) bytes, device               */
    size_t workspace_bytes;
    void* flux;               /* [npv, npt], double (RR_FP64) or float (RR_MIXED)    */
} rr_request;

const char* rr_last_error(void);
int rr_abi_version(void);

/* Annulus edges ze[nin+nedge], means zm[nin+nedge] and mu = sqrt(1-zm^2), computed
 * on the host.  Replaces create_z_grid (common.py:131-149) and rrmodel.py:171-172. */
int rr_z_grid(double zcut, int nin, int nedge, double* ze, double* zm, double* mu);

/* 3-D weight table (device) on linspace(kmin, kmax, nk) x linspace(0, 1-1e-7, ng),
 * written in the device layout W[nk, nz, ng] (FP64) with an optional FP32 copy
 * W32 (NULL to skip).  Writes the
 * reference's dk = (kmax-kmin)/nk and dg to *dk_out / *dg_out (host pointers).
 * Replaces calculate_weights_3d (common.py:188-223), called at rrmodel.py:173. */
int rr_weights_3d(int nk, double kmin, double kmax, const double* ze, int nz, int ng,
                  double* W, float* W32, double* dk_out, double* dg_out, void* stream);

/* Limb-darkening profiles ldp[npv, npb, nmu] at mu[nmu] and disk integrals
 * istar[npv, npb] for a named law; ldc is [npv, npb, nldc].  Laws without an
 * analytic integral use the 200-node trapezoid of rrmodel.py:220-227.
 * Replaces evaluate_ld / evaluate_ldi (models/numba/ldmodels.py:142-175) as used
 * at rrmodel.py:215-227 and tsmodel.py:103-115. */
int rr_ld_profiles(int law, const double* ldc, int npv, int npb, int nldc,
                   const double* mu, int nmu, double* ldp, double* istar, void* stream);

/* Disk integrals of tabulated profiles, istar[npv, npb] = 2 pi trapz_z(z * ldp).
 * Replaces integrate_profiles_set (models/numba/ldtkldm.py:77-91). */
int rr_integrate_profiles(const double* mu, int nmu, const double* ldp, int npv, int npb,
                          double* istar, void* stream);

/* Profiles ldp[npv, npb, nmu] interpolated trilinearly in a grid of stellar
 * intensity profiles grid[nx, ny, nz, npb, nmu] (teff, logg, metallicity axes,
 * evenly spaced: axis node i at x0 + i dx) at each pv's parameters
 * (par[pv * par_stride + 0..2] = teff, logg, metallicity).  Parameters are
 * clamped to the grid; a NaN parameter gives a NaN profile row.  Each axis needs
 * at least 2 nodes.  Replaces trilinear_interpolation_set
 * (models/numba/ldtkldm.py:5-57) as called by LDTkLDModel.__call__
 * (models/ldtkldm.py:74-89); the disk integrals then come from
 * rr_integrate_profiles. */
int rr_ldtk_interpolate(const double* grid, int nx, int ny, int nz, int npb, int nmu, double x0, double dx,
                        double y0, double dy, double z0, double dz, const double* par, int par_stride, int npv,
                        double* ldp, void* stream);

/* Bytes of device workspace rr_evaluate / rr_setup / rr_flux need for these sizes:
 * the per-pv state (orbit coefficients, contact window, validity), the stellar
 * disk integrals and the limb-darkening means. */
size_t rr_workspace_bytes(int npv, int npt, int npb, int ng, int nlc);

/* The full RoadRunner model: per-pv setup (validity, limb-darkening profiles and
 * means, Taylor orbit, contact-time window) then the flux over all (pv, time)
 * points.  flux is [npv, npt]: double* for RR_FP64, float* for RR_MIXED.
 * Replaces rr_full (pytransit/models/roadrunner/model_full.py:9-100) and its
 * npv=1 specialisation rr_simple (model_simple.py:11-138), together with the
 * limb-darkening evaluation of RoadRunnerModel.evaluate (rrmodel.py:215-227). */
int rr_evaluate(int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv,
                const rr_ld* ld, void* workspace, size_t workspace_bytes, void* flux, void* stream);

/* The two stages of rr_evaluate, for callers that time or reuse them separately.
 * rr_setup writes the per-pv state (model_full.py:39-70 plus the profiles) into
 * the workspace; rr_flux reads it and writes the flux (the hot loop,
 * model_full.py:75-99).  rr_flux must follow an rr_setup with the same
 * precision, sizes and workspace. */
int rr_setup(int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv,
             const rr_ld* ld, void* workspace, size_t workspace_bytes, void* stream);
int rr_flux(int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv,
            void* workspace, size_t workspace_bytes, void* flux, void* stream);

/* Up to rr_max_batch() independent requests (each its own light curves,
 * population, limb-darkening law and workspace) evaluated with one setup launch
 * and at most two flux launches on `stream`; the results equal rr_evaluate's for
 * each request alone.  Every request is validated before anything is launched.
 * No counterpart in the reference, which evaluates one model at a time
 * (rrmodel.py:175-238); this is the launch primitive of the request server. */
int rr_max_batch(void);
int rr_evaluate_batch(int precision, const rr_table* tab, int nreq, const rr_request* reqs,
                      void* stream);

/* The batch of rr_evaluate_batch split into a plan on the host and a launch that
 * reads the plan from device memory, so the launch can be captured in a HIP graph
 * once and replayed for every later batch whose shape it covers.  rr_batch_plan
 * validates the requests and writes the batch's descriptor (rr_batch_desc_bytes()
 * bytes, plain data) into `desc` (host memory) and its launch shape into *shape.
 * The caller copies the descriptor to device memory and calls rr_batch_launch with
 * that device copy and a shape covering the plan's: the same setup_multi, and
 * every block count and LDS size at least the plan's (blocks beyond the plan's
 * exit at once).  The results equal rr_evaluate_batch's.  No counterpart in the
 * reference; these are the replayable launch of the request server. */
typedef struct rr_batch_shape {
    int32_t setup_blocks, setup_multi, setup_lds;
    int32_t flux_blocks[2];   /* [0] general light-curve layout, [1] contiguous sorted */
    int32_t flux_lds[2];
} rr_batch_shape;
size_t rr_batch_desc_bytes(void);
int rr_batch_plan(int precision, const rr_table* tab, int nreq, const rr_request* reqs, void* desc,
                  rr_batch_shape* shape);
int rr_batch_launch(int precision, const void* desc_dev, const rr_batch_shape* launch, void* stream);

/* Orbit boundary alone: Taylor coefficients coeffs[npv, 2, 5] (row 0 x, row 1 y,
 * coefficient n = d^n/dt^n / n!) and contact window win[npv, 2] = (t1, t4) for
 * radius ratio k0[npv].  Replaces meepmeep's solve2d + bounding_box as called at
 * model_full.py:60,65 (method: orbits/taylor_z.py:23-102,298-328,391-394). */
int rr_orbit(int npv, const double* p, const double* a, const double* inc, const double* e,
             const double* w, const double* k0, double* coeffs, double* win, void* stream);

/* Transmission-spectroscopy model (TSModel): k is [npv, npb] (kcols = npb), t0 is
 * [npv] (t0cols = 1), nsamples[0] and exptimes[0] apply to every point and
 * lcids/pbids are ignored (Q16).  tab->weights == NULL selects exact per-pv 2-D
 * weights at mean(k) (the default precompute_weights=False path); otherwise the
 * 3-D table is lerped at mean(k).  flux is [npv, npb, npt].  With RR_MIXED the
 * weights x profile contraction runs on FP32 MFMA (v_mfma_f32_16x16x4f32).
 * Replaces tsmodel_serial (pytransit/models/roadrunner/model_trspec.py:11-93) as
 * called by TransmissionSpectroscopyModel.evaluate (tsmodel.py:46-130). */
size_t ts_workspace_bytes(int npv, int npb, int ng, int nz, int precision);
int ts_evaluate(int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv,
                const rr_ld* ld, void* workspace, size_t workspace_bytes, void* flux, void* stream);

/* Gaussian log-likelihood of the model for one light curve and its gradient,
 * fused with the model so the flux is never written:
 *   lnl[pv] = sum_t -ln sigma[pv] - 1/2 ln 2 pi - 1/2 ((obs_t - m_t) / sigma[pv])^2
 *   grad[pv, :] = d lnl / d(k, t0, p, a, i, c0, c1)   (nullable)
 * c0, c1 are the coefficients of the quadratic or power-2 law; k must lie in the
 * weight table's [kmin, kmax] (else NaN); nlc = 1, kcols = 1.  The likelihood is
 * lnlike_normal (pytransit/lpf/loglikelihood/wnloglikelihood.py:22-35) applied to
 * rr_full's flux (model_full.py:9-100); the gradient has no counterpart in the
 * reference (its dfdk/dfdb in common.py:104-128 are approximate and unused). */
int rr_lnlike(int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv,
              const rr_ld* ld, const double* obs, const double* sigma, double* lnl, double* grad,
              void* stream);

/* Mandel & Agol (2002) quadratic limb-darkening model: flux [npv, npt] (FP64) for
 * ldc [npv, 2*npb] = (u1, u2) per passband, k [npv, 1|npb], t0 [npv, nep], with the
 * Taylor orbit and contact window of the RoadRunner path widened by 0.025 d and
 * per-light-curve supersampling.  A NaN radius ratio or semi-major axis gives an
 * inf row, NaN coefficients inf points.  Replaces quadratic_model_v
 * (pytransit/models/numba/ma_quadratic_nb.py:532-595) and eval_quad_z_s (:277-389)
 * as called by QuadraticModel.evaluate (pytransit/models/ma_quadratic.py:74-140). */
size_t ma_workspace_bytes(int npv);
int ma_evaluate(const rr_lcdata* lc, const rr_params* pv, const double* ldc, void* workspace,
                size_t workspace_bytes, double* flux, void* stream);

/* Secondary eclipse of a uniform planet disk: flux[npv, npt] (FP64) is the planet
 * disk area left visible by the star, pi k^2 outside the eclipse window.  k is
 * [npv] (kcols = 1), t0 [npv, t0cols] indexed by lc->epids (EclipseModel passes
 * the light curves' passband indices there, new_eclipse_model.py:70), rstar[npv]
 * the stellar radius in solar radii (light travel time).  The orbit is the
 * in-repo Taylor series about mid-eclipse (orbits/taylor_z.py:106-185) at the
 * conjunction phase (orbits/orbits_py.py:544-555); the contact window gets
 * 0.003 d + the exposure time.  Invalid pv (a <= 1, a NaN, e < 0) give NaN rows.
 * Replaces eclipse_model (pytransit/models/roadrunner/model_eclipse.py:11-81). */
size_t ec_workspace_bytes(int npv);
int eclipse_evaluate(const rr_lcdata* lc, const rr_params* pv, const double* rstar, void* workspace,
                     size_t workspace_bytes, double* flux, void* stream);

/* Eclipse spectroscopy: flux[npv, npb, npt] (FP64) = 1 - (f A / pi) / (1 + f k^2)
 * averaged over the exposure for flux ratios fratio[npv, npb], 1 outside the
 * window (0.0015 d + the exposure time).  t0 is [npv] (t0cols = 1); nsamples[0]
 * and exptimes[0] apply to every time.  Replaces esmodel
 * (pytransit/models/roadrunner/model_ecspec.py:13-65) as called by
 * EclipseSpectroscopyModel.evaluate (esmodel.py:46-88). */
int es_evaluate(const rr_lcdata* lc, const rr_params* pv, const double* rstar, const double* fratio, int npb,
                void* workspace, size_t workspace_bytes, double* flux, void* stream);

/* Oblate planet: the RoadRunner model (FP64) with the planet's sky projection an
 * ellipse of semi-axes k and k (1 - f[pv]) rotated by alpha[pv], its area inside
 * the stellar disk summed over nlines scanlines (2 <= nlines <= 2048).  The
 * workspace is rr_evaluate's (rr_workspace_bytes); the per-pv setup is rr_setup's.
 * Replaces op_full_serial (pytransit/models/roadrunner/opmodel_full.py:40-134) and
 * the scanline area of ecintersection.py:93-207, as called by
 * OblatePlanetModel.evaluate (opmodel.py:50-111). */
int op_evaluate(const rr_table* tab, const rr_lcdata* lc, const rr_params* pv, const rr_ld* ld, const double* f,
                const double* alpha, int nlines, void* workspace, size_t workspace_bytes, double* flux,
                void* stream);

/* rr_evaluate without allocation: the workspace is a device array inside the
 * library, RR_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; the
 * caller passes inputs and the flux buffer only, and no memory is allocated or
 * freed.  Requests whose workspace (rr_workspace_bytes) exceeds the array fail
 * with RR_ERANGE.  Calls that overlap on one GPU share the array and race:
 * serialise them (one stream, or events).  No counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). */
#define RR_NOALLOC_SCRATCH_BYTES (32u << 20)
int rr_evaluate_noalloc(int precision, const rr_table* tab, const rr_lcdata* lc, const rr_params* pv,
                        const rr_ld* ld, void* flux, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RR_HIP_H */
