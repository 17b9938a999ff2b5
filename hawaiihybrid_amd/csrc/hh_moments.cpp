// Spectral moments from a correlation function and the analytic partition function.
//
//  * hh_pf_analytic   analytic_full_partition_function_by_V (src/hawaii.c:5071-5090): the
//                     classical kinetic partition function per unit volume,
//                     atom-atom (2 pi mu T/HkT)^(3/2); linear molecule + atom
//                     4 pi (2 pi T/HkT)^(5/2) mu^(3/2) I; rotor + linear molecule
//                     32 pi^3 (2 pi T/HkT)^4 mu^(3/2) sqrt(I1 I2 I3) I_lin.
//  * hh_cf_moment     compute_Mn_from_cf_using_classical_detailed_balance
//                     (src/hawaii.c:5945-5979): M0 = CF(0) ZeroCoeff / ALU^3 and
//                     M2 = -SecondCoeff CF''(0) / ALU^3, CF''(0) from the central difference
//                     over 21 points (11 samples, CF even in t) or, for 5..10 samples, over 9.
//                     The stencil weights come from Fornberg's recursion (B. Fornberg, Math.
//                     Comp. 51 (1988) 699) on the integer nodes -N..N, in long double.
#include <cmath>

#include "../../include/hawaii_gpu.h"
#include "hh_internal.h"

namespace {

constexpr double kHkT = 4.3597447222071e-18 / 1.380649e-23;   // Hartree / Boltzmann (src/constants.h:24)
constexpr double kALU = 5.29177210903e-11;                     // src/constants.h:8
constexpr double kZeroCoeff = 0.00361479637 / (4.0 * M_PI);    // src/constants.h:29
constexpr double kSecondCoeff = 13856114.29344114;             // src/constants.h:30

// Weights w[0..N] of f''(0) ~ sum_{j=-N..N} w_|j| f(j h) / h^2 on the nodes -N..N.
void second_derivative_weights(int N, long double* w)
{
    const int n = 2 * N + 1;
    long double x[21], c[21][3] = {};
    for (int i = 0; i < n; ++i) x[i] = (long double)(i - N);
    const int m = 2;
    long double c1 = 1.0L, c4 = x[0];
    c[0][0] = 1.0L;
    for (int i = 1; i < n; ++i) {
        const int mn = i < m ? i : m;
        long double c2 = 1.0L;
        const long double c5 = c4;
        c4 = x[i];
        for (int j = 0; j < i; ++j) {
            const long double c3 = x[i] - x[j];
            c2 *= c3;
            if (j == i - 1) {
                for (int k = mn; k >= 1; --k) c[i][k] = c1 * (k * c[i - 1][k - 1] - c5 * c[i - 1][k]) / c2;
                c[i][0] = -c1 * c5 * c[i - 1][0] / c2;
            }
            for (int k = mn; k >= 1; --k) c[j][k] = (c4 * c[j][k] - k * c[j][k - 1]) / c3;
            c[j][0] = c4 * c[j][0] / c3;
        }
        c1 = c2;
    }
    for (int j = 0; j <= N; ++j) w[j] = c[N + j][2];
}

}  // namespace

double hh_pf_analytic(const hh_system* sys, double T)
{
    if (!sys || !(T > 0.0)) return hh::fail(HH_E_INVALID, "hh_pf_analytic: system and T > 0 required"), -1.0;
    const double a = 2.0 * M_PI * T / kHkT;
    switch (sys->system) {
    case HH_SYS_HEAR:
        return std::pow(a * sys->mu, 1.5);
    case HH_SYS_CO2AR:
    case HH_SYS_N2AR:
        return 4.0 * M_PI * std::pow(a, 2.5) * std::pow(sys->mu, 1.5) * sys->I;
    case HH_SYS_CH4CO2:
        return 32.0 * M_PI * M_PI * M_PI * std::pow(a, 4.0) * std::pow(sys->mu, 1.5) *
               std::sqrt(sys->I_rotor[0] * sys->I_rotor[1] * sys->I_rotor[2]) * sys->I;
    default:
        return hh::fail(HH_E_UNSUPPORTED, "hh_pf_analytic: unknown system %d", sys->system), -1.0;
    }
}

int hh_cf_moment(const double* cf, long len, double dt, double ntraj, int n, double* out)
{
    if (!cf || !out || len < 1) return hh::fail(HH_E_INVALID, "hh_cf_moment: cf, out and len >= 1 required");
    const double scale = (ntraj > 0.0 ? 1.0 / ntraj : 1.0) / (kALU * kALU * kALU);
    if (n == 0) {
        *out = cf[0] * kZeroCoeff * scale;
        return HH_OK;
    }
    if (n != 2) return hh::fail(HH_E_UNSUPPORTED, "hh_cf_moment: moment of order %d is not provided", n);
    if (!(dt > 0.0)) return hh::fail(HH_E_INVALID, "hh_cf_moment: dt > 0 required");
    int N;
    if (len >= 11) N = 10;
    else if (len >= 5) N = 4;
    else return hh::fail(HH_E_INVALID, "hh_cf_moment: %ld samples are too few to estimate M2", len);
    long double w[11];
    second_derivative_weights(N, w);
    long double d2 = w[0] * cf[0];
    for (int j = 1; j <= N; ++j) d2 += 2.0L * w[j] * cf[j];
    *out = -kSecondCoeff * (double)(d2 / ((long double)dt * dt)) * scale;
    return HH_OK;
}

// compute_Mn_from_sf_using_classical_detailed_balance (src/hawaii.c:5981-6013):
// Mn = 2 Moment_SF_Coeff * integral of sf(nu) nu^n d nu (cm^-1) by the composite Simpson 3/8
// weights of integrate_composite_simpson (src/hawaii.c:5926-5943): 1, 3, 3, 2, 3, 3, 2, ..., 1.
int hh_sf_moment(const double* nu, const double* sf, long len, double ntraj, int n, double* out)
{
    if (!nu || !sf || !out || len < 2) return hh::fail(HH_E_INVALID, "hh_sf_moment: nu, sf, out and len >= 2 required");
    if (n < 0) return hh::fail(HH_E_INVALID, "hh_sf_moment: order must be non-negative");
    if (n % 2 == 1) {
        *out = 0.0;
        return HH_OK;
    }
    const double NL = 2.686780111e25;          // src/constants.h:22
    const double HBar = 1.054571817e-34;       // src/constants.h:21
    const double coeff = (8.0 * M_PI * M_PI * M_PI) * NL * NL / 3.0 / HBar;   // Moment_SF_Coeff
    const double h = nu[1] - nu[0];
    double sum = sf[0] * std::pow(nu[0], n) + sf[len - 1] * std::pow(nu[len - 1], n);
    for (long j = 1; j < len - 1; ++j) sum += (j % 3 == 0 ? 2.0 : 3.0) * (sf[j] * std::pow(nu[j], n));
    *out = 2.0 * coeff * (sum * 3.0 * h / 8.0);
    if (ntraj > 0) *out /= ntraj;
    return HH_OK;
}
