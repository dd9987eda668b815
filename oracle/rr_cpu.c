/*
 * rr_cpu.c -- C11/OpenMP FP64 restatement of PyTransit's RoadRunner rr_full.
 *
 * TEST INFRASTRUCTURE ONLY: the CPU baseline of bench.py and a cross-check of the
 * numpy oracle.  Never linked into the product library.
 *
 * Loop structure follows pytransit/models/roadrunner/model_full.py:9-100: a serial
 * per-parameter-vector setup (validity, limb-darkening means, Taylor orbit,
 * contact window; :39-70) and one `omp parallel for schedule(static)` over the
 * flattened npv*npt index, like `prange(npv * npt)` at :76.  Helpers restate
 * common.py (geometry, weights, interpolation), numba/ldmodels.py (laws) and
 * orbits/taylor_z.py + orbits/orbits_py.py (orbit), with the same defined
 * behaviour as the numpy oracle where the reference is undefined (Q2, Q3).
 * Compiled without -ffast-math and with -ffp-contract=off, as the reference's
 * kernels run with fastmath=False.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double PI_ = 3.14159265358979323846;

static double lin_at(double a, double b, int n, int i) {
    if (i == n - 1) return b;
    double step = (b - a) / (double)(n - 1);
    return (double)i * step + a;
}

/* common.py:36-49 */
static double cci_area(double r1, double r2, double b) {
    if (r1 < b - r2) return 0.0;
    if (r1 >= b + r2) return PI_ * (r2 * r2);
    if (b - r2 <= -r1) return PI_ * (r1 * r1);
    return r2 * r2 * acos((b * b + r2 * r2 - r1 * r1) / (2 * b * r2)) +
           r1 * r1 * acos((b * b + r1 * r1 - r2 * r2) / (2 * b * r1)) -
           0.5 * sqrt((-b + r2 + r1) * (b + r2 - r1) * (b - r2 + r1) * (b + r2 + r1));
}

/* common.py:52-73 with tsort :5-33 */
static double ccia_kite(double r1, double r2, double b) {
    if (r1 + r2 <= b) return 0.0;
    if (fabs(r1 - r2) < b && b <= r1 + r2) {
        double hi = r1 > r2 ? r1 : r2, lo = r1 > r2 ? r2 : r1;
        double x = hi > b ? hi : b;
        double z = lo < b ? lo : b;
        double m = hi < b ? hi : b;
        double y = lo > m ? lo : m;
        double ak = 0.5 * sqrt((x + (y + z)) * (z - (x - y)) * (z + (x - y)) * (x + (y - z)));
        double k0 = atan2(2.0 * ak, (r2 - r1) * (r2 + r1) + b * b);
        double k1 = atan2(2.0 * ak, (r1 - r2) * (r1 + r2) + b * b);
        return r1 * r1 * k1 + r2 * r2 * k0 - ak;
    }
    if (b <= r1 - r2) return PI_ * (r2 * r2);
    if (b <= r2 - r1) return PI_ * (r1 * r1);
    return NAN;
}

/* common.py:225-233 (index clamped to ng-2) */
static double interp_ld(double g, double dg, const double* l, int ng) {
    if (g < 0.0) return NAN;
    if (g > 1.0) return 0.0;
    int i = (int)floor(g / dg);
    if (i > ng - 2) i = ng - 2;
    if (i < 0) i = 0;
    double a = (g - i * dg) / dg;
    return (1.0 - a) * l[i] + a * l[i + 1];
}

/* common.py:188-223; row sums accumulated left to right */
void rrcpu_weights_3d(int nk, double kmin, double kmax, const double* ze, int nz, int ng, double* W) {
#pragma omp parallel for schedule(static)
    for (int row = 0; row < nk * ng; ++row) {
        int ik = row / ng, ig = row % ng;
        double k = lin_at(kmin, kmax, nk, ik), g = lin_at(0.0, 1.0 - 1e-7, ng, ig);
        double b = g * (1.0 + k);
        double* w = W + (size_t)row * nz;
        double a0 = cci_area(ze[0], k, b), s = a0;
        w[0] = a0;
        for (int i = 1; i < nz; ++i) {
            double a1 = cci_area(ze[i], k, b);
            w[i] = a1 - a0;
            a0 = a1;
            s += w[i];
        }
        for (int i = 0; i < nz; ++i) w[i] /= s;
    }
}

/* models/numba/ldmodels.py:22-139; law ids as RR_LD_* in include/rr_hip.h */
static double ld_law(int law, double mu, const double* c, int nc) {
    switch (law) {
        case 0: return 1.0;
        case 1: return 1.0 - c[0] * (1.0 - mu);
        case 2: return 1.0 - c[0] * (1.0 - mu) - c[1] * ((1.0 - mu) * (1.0 - mu));
        case 3: {
            double a = sqrt(c[0]), b = 2 * c[1], u = a * b, v = a * (1.0 - b);
            return 1.0 - u * (1.0 - mu) - v * ((1.0 - mu) * (1.0 - mu));
        }
        case 4: return 1.0 - c[0] * (1.0 - sqrt(mu)) - c[1] * (1.0 - mu) - c[2] * (1.0 - pow(mu, 1.5)) - c[3] * (1.0 - mu * mu);
        case 5: {
            double s = 0.0;
            for (int i = 0; i < nc; ++i) s += c[i] * (1.0 - pow(mu, (double)(i + 1)));
            return s;
        }
        case 6: return 1.0 - c[0] * (1.0 - mu) - c[1] * (1.0 - sqrt(mu));
        case 7: return 1.0 - c[0] * (1.0 - mu) - c[1] * mu * log(mu);
        case 8: return 1.0 - c[0] * (1.0 - mu) - c[1] / (1.0 - exp(mu));
        case 9: return 1.0 - c[0] * (1.0 - pow(mu, c[1]));
        case 10: {
            double cc = 1.0 - c[0] + c[1], a = log2(cc / c[1]);
            return 1.0 - cc * (1.0 - pow(mu, a));
        }
    }
    return NAN;
}

static double ld_istar(int law, const double* c, int nc) {
    switch (law) {
        case 0: return PI_;
        case 1: return 2 * PI_ * 1 / 6.0 * (3 - 2 * c[0]);
        case 2: return 2 * PI_ * 1 / 12.0 * (-2 * c[0] - c[1] + 6);
        case 3: {
            double a = sqrt(c[0]), b = 2 * c[1], u = a * b, v = a * (1.0 - b);
            return 2 * PI_ * 1 / 12.0 * (-2 * u - v + 6);
        }
        case 9: return 2 * PI_ * (-c[0] * c[1] + c[1] + 2) / (2 * c[1] + 4);
    }
    /* rrmodel.py:220-227: 2*pi*trapz(z*I, z) on mu = linspace(1, 0, 200) */
    double s = 0.0, zp = 0.0, yp = 0.0;
    for (int m = 0; m < 200; ++m) {
        double mu = lin_at(1.0, 0.0, 200, m), z = sqrt(1.0 - mu * mu), y = z * ld_law(law, mu, c, nc);
        if (m > 0) s += (z - zp) * (y + yp) / 2.0;
        zp = z;
        yp = y;
    }
    return 2 * PI_ * s;
}

/* orbits_py.py:82-86, 115-119, 144-154, 191-200 */
static double pymod(double x, double y) {
    double r = fmod(x, y);
    if (r != 0.0) {
        if ((r < 0.0) != (y < 0.0)) r += y;
    } else {
        r = copysign(0.0, y);
    }
    return r;
}

static double ta_newton(double t, double p, double e, double w) {
    double off = atan2(sqrt(1.0 - e * e) * sin(0.5 * PI_ - w), e + cos(0.5 * PI_ - w));
    off -= e * sin(off);
    double ma = pymod(2 * PI_ * (t - (0.0 - off * p / (2 * PI_))) / p, 2 * PI_);
    double ea = ma, err = 0.05;
    int k = 0;
    while (fabs(err) > 1e-8 && k < 1000) {
        err = ea - e * sin(ea) - ma;
        ea = ea - err / (1.0 - e * cos(ea));
        ++k;
    }
    double sta = sqrt(1.0 - e * e) * sin(ea) / (1.0 - e * cos(ea));
    double cta = (cos(ea) - e) / (1.0 - e * cos(ea));
    return atan2(sta, cta);
}

/* taylor_z.py:23-102; c[0..4] x, c[5..9] y, divided by n! */
static void taylor_coeffs(double p, double a, double inc, double e, double w, double* c) {
    const double dt = 2e-2;
    double ae = a * (1.0 - e * e), ci = cos(inc), x[7], y[7];
    for (int j = 0; j < 7; ++j) {
        double f = ta_newton((j - 3) * dt, p, e, w);
        double r = ae / (1.0 + e * cos(f));
        x[j] = -r * cos(w + f);
        y[j] = -r * sin(w + f) * ci;
    }
    double vx = (1.0 / 60 * (x[6] - x[0]) + 9.0 / 60 * (x[1] - x[5]) + 45.0 / 60 * (x[4] - x[2])) / dt;
    double vy = (1.0 / 60 * (y[6] - y[0]) + 9.0 / 60 * (y[1] - y[5]) + 45.0 / 60 * (y[4] - y[2])) / dt;
    double ax = (1.0 / 90 * (x[0] + x[6]) - 3.0 / 20 * (x[1] + x[5]) + 1.5 * (x[2] + x[4]) - 49.0 / 18 * x[3]) / (dt * dt);
    double ay = (1.0 / 90 * (y[0] + y[6]) - 3.0 / 20 * (y[1] + y[5]) + 1.5 * (y[2] + y[4]) - 49.0 / 18 * y[3]) / (dt * dt);
    double jx = (1.0 / 8 * (x[0] - x[6]) + (x[5] - x[1]) + 13.0 / 8 * (x[2] - x[4])) / (dt * dt * dt);
    double jy = (1.0 / 8 * (y[0] - y[6]) + (y[5] - y[1]) + 13.0 / 8 * (y[2] - y[4])) / (dt * dt * dt);
    double sx = (-1.0 / 6 * (x[0] + x[6]) + 2 * (x[1] + x[5]) - 6.5 * (x[2] + x[4]) + 28.0 / 3 * x[3]) / (dt * dt * dt * dt);
    double sy = (-1.0 / 6 * (y[0] + y[6]) + 2 * (y[1] + y[5]) - 6.5 * (y[2] + y[4]) + 28.0 / 3 * y[3]) / (dt * dt * dt * dt);
    c[0] = 0.0; c[1] = vx; c[2] = 0.5 * ax; c[3] = jx / 6.0; c[4] = sx / 24.0;
    c[5] = y[3]; c[6] = vy; c[7] = 0.5 * ay; c[8] = jy / 6.0; c[9] = sy / 24.0;
}

static double sep(double t, const double* c) {
    double px = (((c[4] * t + c[3]) * t + c[2]) * t + c[1]) * t + c[0];
    double py = (((c[9] * t + c[8]) * t + c[7]) * t + c[6]) * t + c[5];
    return sqrt(px * px + py * py);
}

static double z_wrapped(double t, const double* c) {
    double ep = floor((t - 0.0 + 0.5 * 1.0) / 1.0);
    return sep(t - (0.0 + ep * 1.0), c);
}

/* taylor_z.py:298-328 */
static double contact(double k, int point, const double* c) {
    double s = (point == 1) ? -1.0 : 1.0, zt = 1.0 + k;
    double t0 = 0.0, t2 = s * 2.0 / c[1], t1 = 0.5 * t2;
    double z0 = z_wrapped(t0, c) - zt, z1 = z_wrapped(t1, c) - zt;
    for (int i = 0; fabs(t2 - t0) > 1e-6 && i < 100; ++i) {
        if (z0 * z1 < 0.0) {
            t2 = t1;
            t1 = 0.5 * (t0 + t1);
            z1 = z_wrapped(t1, c) - zt;
        } else {
            t0 = t1;
            t1 = 0.5 * (t1 + t2);
            z0 = z1;
            z1 = z_wrapped(t1, c) - zt;
        }
    }
    return t1;
}

/*
 * RoadRunnerModel.evaluate for a named law (rrmodel.py:212-235) + rr_full.
 * Layouts as include/rr_hip.h; ldc is [npv, npb, nldc]; flux [npv, npt].
 * nthreads <= 0 keeps the OpenMP default.  Returns 0.
 */
int rrcpu_evaluate(int npt, const double* times, int nlc, const int32_t* lcids, int npb, const int32_t* pbids,
                   const int32_t* epids, const int32_t* nsamples, const double* exptimes, int npv, int kcols,
                   const double* k, int t0cols, const double* t0, const double* p, const double* a,
                   const double* inc, const double* e, const double* w, int law, int nldc, const double* ldc,
                   int nz, const double* mu, int nk, int ng, double kmin, double kmax, double dk, double dg,
                   const double* W, const double* ze, double* flux, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    double* ldp = (double*)malloc(sizeof(double) * (size_t)npv * npb * nz);
    double* istar = (double*)malloc(sizeof(double) * (size_t)npv * npb);
    double* ldm = (double*)calloc((size_t)npv * npb * ng, sizeof(double));
    double* xyc = (double*)calloc((size_t)npv * 10, sizeof(double));
    double* bbs = (double*)calloc((size_t)npv * nlc * 2, sizeof(double));
    char* good = (char*)calloc((size_t)npv, 1);
    double* wg = (double*)malloc(sizeof(double) * (size_t)ng * nz);

    /* limb darkening (evaluate_ld / evaluate_ldi) */
    for (int j = 0; j < npv * npb; ++j) {
        const double* c = ldc + (size_t)j * nldc;
        for (int i = 0; i < nz; ++i) ldp[(size_t)j * nz + i] = ld_law(law, mu[i], c, nldc);
        istar[j] = ld_istar(law, c, nldc);
    }

    /* serial per-pv setup; model_full.py:39-70 */
    for (int ipv = 0; ipv < npv; ++ipv) {
        const double* lp = ldp + (size_t)ipv * npb * nz;
        if (isnan(a[ipv]) || a[ipv] <= 1.0 || e[ipv] < 0.0 || isnan(lp[0])) continue;
        good[ipv] = 1;
        double k0 = k[(size_t)ipv * kcols];
        if (kmin <= k0 && k0 <= kmax) {
            int ik = (int)floor((k0 - kmin) / dk);
            if (ik > nk - 2) ik = nk - 2;
            double ak = (k0 - kmin - ik * dk) / dk;
            const double* W0 = W + (size_t)ik * ng * nz;
            const double* W1 = W0 + (size_t)ng * nz;
            for (int ipb = 0; ipb < npb; ++ipb)
                for (int ig = 0; ig < ng; ++ig) {
                    double d0 = 0.0, d1 = 0.0;
                    for (int iz = 0; iz < nz; ++iz) {
                        d0 += W0[ig * nz + iz] * lp[ipb * nz + iz];
                        d1 += W1[ig * nz + iz] * lp[ipb * nz + iz];
                    }
                    ldm[((size_t)ipv * npb + ipb) * ng + ig] = (1.0 - ak) * d0 + ak * d1;
                }
        } else {
            for (int ig = 0; ig < ng; ++ig) {
                double b = lin_at(0.0, 1.0 - 1e-7, ng, ig) * (1.0 + k0);
                double a0 = cci_area(ze[0], k0, b), s = a0;
                wg[ig * nz] = a0;
                for (int iz = 1; iz < nz; ++iz) {
                    double a1 = cci_area(ze[iz], k0, b);
                    wg[ig * nz + iz] = a1 - a0;
                    a0 = a1;
                    s += wg[ig * nz + iz];
                }
                for (int iz = 0; iz < nz; ++iz) wg[ig * nz + iz] /= s;
            }
            for (int ipb = 0; ipb < npb; ++ipb)
                for (int ig = 0; ig < ng; ++ig) {
                    double d = 0.0;
                    for (int iz = 0; iz < nz; ++iz) d += wg[ig * nz + iz] * lp[ipb * nz + iz];
                    ldm[((size_t)ipv * npb + ipb) * ng + ig] = d;
                }
        }
        double* c = xyc + (size_t)ipv * 10;
        taylor_coeffs(p[ipv], a[ipv], inc[ipv], e[ipv], w[ipv], c);
        double bt1 = contact(k0, 1, c), bt4 = contact(k0, 4, c);
        for (int ilc = 0; ilc < nlc; ++ilc) {
            bbs[((size_t)ipv * nlc + ilc) * 2 + 0] = bt1 - (0.003 + exptimes[ilc]);
            bbs[((size_t)ipv * nlc + ilc) * 2 + 1] = bt4 + (0.003 + exptimes[ilc]);
        }
    }

    /* hot loop; model_full.py:75-99 */
    const long long n = (long long)npv * npt;
#pragma omp parallel for schedule(static)
    for (long long j = 0; j < n; ++j) {
        int ipv = (int)(j / npt), ipt = (int)(j % npt);
        if (!good[ipv]) {
            flux[j] = NAN;
            continue;
        }
        int ilc = nlc > 1 ? lcids[ipt] : 0, ipb = pbids[ilc], iep = epids[ilc];
        double tz = t0[(size_t)ipv * t0cols + iep], pp = p[ipv];
        double epoch = floor((times[ipt] - tz + 0.5 * pp) / pp);
        double tc = times[ipt] - (tz + epoch * pp);
        const double* bb = bbs + ((size_t)ipv * nlc + ilc) * 2;
        if (!(bb[0] <= tc && tc <= bb[1])) {
            flux[j] = 1.0;
            continue;
        }
        double kk = k[(size_t)ipv * kcols + (kcols == 1 ? 0 : ipb)], ist = istar[(size_t)ipv * npb + ipb];
        const double* l = ldm + ((size_t)ipv * npb + ipb) * ng;
        const double* c = xyc + (size_t)ipv * 10;
        int ns = nsamples[ilc];
        double f = 0.0;
        for (int is = 1; is <= ns; ++is) {
            double off = exptimes[ilc] * ((is - 0.5) / ns - 0.5);
            double z = sep(tc + off, c);
            double ip = interp_ld(z / (1.0 + kk), dg, l, ng);
            double ap = ccia_kite(1.0, kk, z);
            f += (ist - ip * ap) / ist;
        }
        flux[j] = f / ns;
    }
    free(ldp);
    free(istar);
    free(ldm);
    free(xyc);
    free(bbs);
    free(good);
    free(wg);
    return 0;
}
