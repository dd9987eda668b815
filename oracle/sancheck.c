/* Drives the C restatement (rr_cpu.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer (oracle/Makefile: make sanitize).  TEST
 * INFRASTRUCTURE: a small workload with two light curves in two passbands,
 * supersampling, a radius ratio outside the table (exact weights), an invalid
 * parameter vector and a short period (several epochs), checked for a NaN row
 * where the pv is invalid and fluxes in (0, 1] elsewhere. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

void rrcpu_weights_3d(int nk, double kmin, double kmax, const double* ze, int nz, int ng, double* W);
int rrcpu_evaluate(int npt, const double* times, int nlc, const int32_t* lcids, int npb, const int32_t* pbids,
                   const int32_t* epids, const int32_t* nsamples, const double* exptimes, int npv, int kcols,
                   const double* k, int t0cols, const double* t0, const double* p, const double* a,
                   const double* inc, const double* e, const double* w, int law, int nldc, const double* ldc,
                   int nz, const double* mu, int nk, int ng, double kmin, double kmax, double dk, double dg,
                   const double* W, const double* ze, double* flux, int nthreads);

int main(void) {
    enum { NZ = 40, NG = 100, NK = 64, NPT = 600, NPV = 5, NPB = 2 };
    double ze[NZ], mu[NZ];
    for (int i = 0; i < NZ; ++i) {
        ze[i] = (i + 1) / (double)NZ;
        const double zm = i == 0 ? 0.0 : 0.5 * (ze[i - 1] + ze[i]);
        mu[i] = sqrt(1.0 - zm * zm);
    }
    const double kmin = 0.005, kmax = 0.5;
    double* W = malloc(sizeof(double) * NK * NG * NZ);
    rrcpu_weights_3d(NK, kmin, kmax, ze, NZ, NG, W);
    double times[NPT];
    int32_t lcids[NPT];
    for (int i = 0; i < NPT; ++i) {
        times[i] = -0.3 + 0.6 * (i % (NPT / 2)) / (NPT / 2) + (i >= NPT / 2 ? 1.7 : 0.0);
        lcids[i] = i >= NPT / 2;
    }
    const int32_t pbids[NPB] = {0, 1}, epids[NPB] = {0, 1}, nsamples[NPB] = {3, 1};
    const double exptimes[NPB] = {0.01, 0.0};
    double k[NPV * NPB], t0[NPV * 2], p[NPV], a[NPV], inc[NPV], e[NPV], w[NPV], ldc[NPV * NPB * 2];
    for (int j = 0; j < NPV; ++j) {
        k[j * NPB] = 0.08 + 0.01 * j;
        k[j * NPB + 1] = 0.09 + 0.01 * j;
        t0[j * 2] = 0.001 * j;
        t0[j * 2 + 1] = 1.7 + 0.001 * j;
        p[j] = j == 3 ? 0.4 : 2.0 + j;
        a[j] = 6.0 + j;
        inc[j] = acos(0.3 / a[j]);
        e[j] = j == 4 ? 0.2 : 0.0;
        w[j] = 0.5;
        for (int b = 0; b < NPB; ++b) {
            ldc[(j * NPB + b) * 2] = 0.4;
            ldc[(j * NPB + b) * 2 + 1] = 0.2;
        }
    }
    k[2 * NPB] = 0.7;   /* beyond kmax: exact 2-D weights */
    a[1] = 0.8;         /* invalid: a NaN row */
    double* flux = malloc(sizeof(double) * NPV * NPT);
    rrcpu_evaluate(NPT, times, 2, lcids, NPB, pbids, epids, nsamples, exptimes, NPV, NPB, k, 2, t0, p, a, inc, e, w,
                   2, 2, ldc, NZ, mu, NK, NG, kmin, kmax, (kmax - kmin) / NK, (1.0 - 1e-7) / (NG - 1), W, ze, flux,
                   1);
    int bad = 0, intransit = 0;
    for (int j = 0; j < NPV; ++j)
        for (int i = 0; i < NPT; ++i) {
            const double f = flux[j * NPT + i];
            if (j == 1) bad += !isnan(f);
            else bad += !(f > 0.0 && f <= 1.0);
            intransit += j != 1 && f < 1.0;
        }
    free(W);
    free(flux);
    if (bad || intransit < 100) {
        fprintf(stderr, "sancheck: %d unexpected values, %d in-transit points\n", bad, intransit);
        return 1;
    }
    printf("sancheck passed (%d in-transit points)\n", intransit);
    return 0;
}
