/* CF -> SF transform (test infrastructure).
 *
 * connes_apodization  src/hawaii.c:5401-5410   factor (1 - (t/tmax)^2)^2, tmax = dt*n
 * dct                 src/hawaii.c:5412-5473   Makhoul DCT-II via a 2n-point real FFT
 * idct                src/hawaii.c:5478-5553   Makhoul: pack 2n complex values
 *                        v_k e^{i pi k/2n}/(2n), 0, -v_{2n-k} e^{i pi k/2n}/(2n), backward FFT,
 *                        real part of the first n
 * idct_cf_to_sf       src/hawaii.c:5569-5601
 * GSL's radix-2 FFTs are replaced by a textbook iterative radix-2 transform.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define ATU 2.4188843265857e-17
#define LightSpeed_cm 2.99792458e10
#define ADIPMOMU 8.4783536255e-30
#define EPSILON0 8.8541878128e-12

/* sign = -1: forward e^{-2 pi i jk/n}; sign = +1: backward (unscaled) */
void orc_fft_radix2(double *re, double *im, size_t n, int sign)
{
    size_t i, j, len, k;
    for (i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (len = 2; len <= n; len <<= 1) {
        size_t half = len >> 1;
        for (k = 0; k < half; ++k) {
            double ang = sign * 2.0 * M_PI * (double)k / (double)len;
            double wr = cos(ang), wi = sin(ang);
            for (i = k; i < n; i += len) {
                size_t m = i + half;
                double xr = re[m] * wr - im[m] * wi;
                double xi = re[m] * wi + im[m] * wr;
                re[m] = re[i] - xr;
                im[m] = im[i] - xi;
                re[i] += xr;
                im[i] += xi;
            }
        }
    }
}

void orc_idct(const double *v, size_t len, double *out)
{
    size_t n2 = 2 * len, k;
    double *re = calloc(n2, sizeof(double)), *im = calloc(n2, sizeof(double));
    for (k = 0; k < len; ++k) {
        double complex cx = v[k] * cexp(I * M_PI * (double)k / (2.0 * len)) / (2.0 * len);
        re[k] = creal(cx);
        im[k] = cimag(cx);
    }
    for (k = len + 1; k < n2; ++k) {
        size_t i = n2 - k;
        double complex cx = -v[i] * cexp(I * M_PI * (double)k / (2.0 * len)) / (2.0 * len);
        re[k] = creal(cx);
        im[k] = cimag(cx);
    }
    orc_fft_radix2(re, im, n2, +1);
    for (k = 0; k < len; ++k) out[k] = re[k];
    free(re);
    free(im);
}

void orc_dct(const double *v, size_t len, double *out)
{
    size_t n2 = 2 * len, i;
    double *re = calloc(n2, sizeof(double)), *im = calloc(n2, sizeof(double));
    for (i = 0; i < len; ++i) {
        re[i] = v[i];
        re[len + i] = v[len - 1 - i];
    }
    orc_fft_radix2(re, im, n2, -1);
    out[0] = re[0];
    for (i = 1; i < len; ++i) {
        double complex cx = (re[i] + I * im[i]) * cexp(-I * M_PI * (double)i / (2.0 * len));
        out[i] = creal(cx);
    }
    free(re);
    free(im);
}

void orc_connes(double *a, size_t n, double dt)
{
    double tmax = dt * n;
    size_t i;
    for (i = 0; i < n; ++i) {
        double tcurr = dt * i;
        double factor = (1.0 - (tcurr / tmax) * (tcurr / tmax)) * (1.0 - (tcurr / tmax) * (tcurr / tmax));
        a[i] *= factor;
    }
}

void orc_cf_to_sf(const double *t, double *cf, size_t n, double *nu, double *sf)
{
    double Xscale = 1.0 / LightSpeed_cm / ATU / 2.0 / M_PI;
    double Yscale = ATU * ADIPMOMU * ADIPMOMU / (4.0 * M_PI * EPSILON0);
    double dt = t[1] - t[0], tmax;
    size_t i;
    orc_connes(cf, n, dt);
    orc_idct(cf, n, sf);
    tmax = n * dt;
    for (i = 0; i < n; ++i) {
        nu[i] = i * M_PI / tmax * Xscale;
        sf[i] = sf[i] * tmax / M_PI * Yscale;
    }
}

/* One PR/mu trajectory's spectrum (src/hawaii.c:3806-3822): each component's series of L
 * samples (x[c * L + i], zero after the trajectory) Connes-apodised, transformed (a real radix-2
 * FFT of length L), squared as gsl_fft_square does (|X_k|^2), summed over the components:
 * out[k], k < nfreq. */
void orc_prmu_power(const double *x, size_t L, double dt, size_t nfreq, double *out)
{
    double *re = malloc(L * sizeof(double)), *im = malloc(L * sizeof(double));
    size_t c, k;
    for (k = 0; k < nfreq; ++k) out[k] = 0.0;
    for (c = 0; c < 3; ++c) {
        memcpy(re, x + c * L, L * sizeof(double));
        memset(im, 0, L * sizeof(double));
        orc_connes(re, L, dt);
        orc_fft_radix2(re, im, L, -1);
        for (k = 0; k < nfreq; ++k) out[k] += re[k] * re[k] + im[k] * im[k];
    }
    free(re);
    free(im);
}
