/*
 * Spectral moments from a correlation function (test infrastructure).
 *
 * compute_Mn_from_cf_using_classical_detailed_balance (src/hawaii.c:5945-5979):
 *   M0 = CF[0] ZeroCoeff / ALU^3,  M2 = -SecondCoeff CF''(0) / ALU^3  (/ ntraj if unnormalised),
 * CF''(0) by the maximal-order central difference on 2N+1 points (N = 10 when at least 11
 * samples exist, N = 4 when 5..10 do), CF even in t.  The weights here come from the closed
 * form of the central second-difference weights,
 *   f''(0) ~ sum_{j=1..N} c_j (f(j h) - 2 f(0) + f(-j h)) / h^2,
 *   c_j = 2 (-1)^(j+1) (N!)^2 / (j^2 (N-j)! (N+j)!),
 * an expression independent of the recursion the engine uses.
 */
#include <math.h>

#include "oracle.h"

#define ALU 5.29177210903e-11                     /* src/constants.h:8 */
#define ZERO_COEFF (0.00361479637 / (4.0 * M_PI)) /* src/constants.h:29 */
#define SECOND_COEFF 13856114.29344114            /* src/constants.h:30 */

static long double fact(int n)
{
    long double f = 1.0L;
    int i;
    for (i = 2; i <= n; ++i) f *= i;
    return f;
}

int orc_cf_moment(const double *cf, size_t len, double dt, double ntraj, int n, double *out)
{
    const double scale = (ntraj > 0.0 ? 1.0 / ntraj : 1.0) / (ALU * ALU * ALU);
    long double d2 = 0.0L;
    int N, j;
    if (n == 0) {
        *out = cf[0] * ZERO_COEFF * scale;
        return 0;
    }
    if (n != 2) return -1;
    if (len >= 11) N = 10;
    else if (len >= 5) N = 4;
    else return -1;
    for (j = 1; j <= N; ++j) {
        const long double c = 2.0L * ((j & 1) ? 1.0L : -1.0L) * fact(N) * fact(N) /
                              ((long double)j * j * fact(N - j) * fact(N + j));
        d2 += c * (2.0L * cf[j] - 2.0L * cf[0]);
    }
    *out = -SECOND_COEFF * (double)(d2 / ((long double)dt * dt)) * scale;
    return 0;
}

/*
 * CORRELATION_ARRAY weights (calculate_correlation_array_and_save, src/hawaii.c:2639-2674).
 * For trajectory t of energy E[t] (H at its initial point) and satellite temperature Ts:
 *   WTT = exp(-pesmin HkT / Tb / Ts (Tb - Ts)) for bound pairs with Ts < Tb, else 1,
 *   weight = exp(-E HkT / Tb / Ts (Tb - Ts)) / WTT,
 * free pairs (pair_state 1) take no weight when Ts > Tb.  w[t * M + st] = weight * ratio[st]
 * (ratios NULL: 1) and nstar[st] = sum of the weights in trajectory order.  A trajectory with
 * counted[t] == 0 takes no weight at all: the array loop skips a trajectory whose last
 * integrator call ended on the step limit (`if (status == -1) continue;`, src/hawaii.c:2637),
 * unlike calculate_correlation_and_save, which still counts it (counted NULL: all counted).
 */
#define HKT (4.3597447222071e-18 / 1.380649e-23) /* src/constants.h:24 */

void orc_cf_array_weights(const double *E, long n, double Tb, int pair_state, double pesmin, int M,
                          const double *temps, const double *ratios, const int *counted, double *w, double *nstar)
{
    long t;
    int st;
    for (st = 0; st < M; ++st) nstar[st] = 0.0;
    for (t = 0; t < n; ++t) {
        for (st = 0; st < M; ++st) {
            const double Ts = temps[st];
            double WTT = 1.0, weight;
            w[t * M + st] = 0.0;
            if (counted && !counted[t]) continue;
            if (pair_state == 2 && Ts < Tb) WTT = exp(-pesmin * HKT / Tb / Ts * (Tb - Ts));
            weight = exp(-E[t] * HKT / Tb / Ts * (Tb - Ts)) / WTT;
            if (pair_state == 1 && Ts > Tb) continue;
            w[t * M + st] = weight * (ratios ? ratios[st] : 1.0);
            nstar[st] += weight;
        }
    }
}

/* compute_Mn_from_sf_using_classical_detailed_balance (src/hawaii.c:5981-6013) with
 * integrate_composite_simpson (src/hawaii.c:5926-5943) */
double orc_sf_moment(const double *nu, const double *sf, size_t len, double ntraj, int n)
{
    const double coeff = (8.0 * M_PI * M_PI * M_PI) * 2.686780111e25 * 2.686780111e25 / 3.0 / 1.054571817e-34;
    double h = nu[1] - nu[0], sum;
    size_t j;
    if (n % 2 == 1) return 0.0;
    sum = sf[0] * pow(nu[0], n) + sf[len - 1] * pow(nu[len - 1], n);
    for (j = 1; j < len - 1; ++j) sum += (j % 3 == 0 ? 2.0 : 3.0) * (sf[j] * pow(nu[j], n));
    return 2.0 * coeff * (sum * 3.0 * h / 8.0) / (ntraj > 0 ? ntraj : 1.0);
}
