// PROCESSING numerics on the host: the baseline fit of FIT_BASELINE and the LOESS smoothing of
// SMOOTH.  Both work on one CF or SF of at most a few 10^5 points after the device has produced
// it, so they run where the reference runs them, on the host.
//
//  * hh_fit_baseline  fit_baseline (src/hawaii.c:5313-5400) with gsl_nonlinear_opt
//                     (src/hawaii.c:5215-5311): the wing model c(t) = C + A / (1 + B^2 t^2)
//                     (src/hawaii.c:5147-5149) fitted by least squares to the CF points k >= begin,
//                     times scaled by the time of the secondary maximum (the largest value after the
//                     CF first turns negative) and values by that maximum.  The reference allocates
//                     len - begin + 1 data points and fills len - begin of them; glibc serves
//                     allocations of this size from fresh, zeroed pages, so its last point is (0, 0)
//                     and its starting C is max(0, 1) = 1.  That point is part of the fit here too.
//                     The minimiser is Levenberg-Marquardt with More's column scaling and Nielsen's
//                     damping update (K. Madsen, H. B. Nielsen, O. Tingleff, "Methods for non-linear
//                     least squares problems", 2004, algorithm 3.16), stopped when every step
//                     component |dx_j| <= xtol (|x_j| + xtol) or the scaled gradient is below gtol.
//                     The reference runs GSL's trust-region LM at xtol = gtol = 1e-8 for at most 100
//                     steps.  The wing fits lie in a flat valley (parameter sets 1e-4 apart have
//                     sums of squares equal to 1e-15) where a step test of 1e-8 stops this
//                     minimiser short of the minimum the reference's outputs show it reached, so
//                     xtol = 1e-12 and gtol = 1e-15 here, up to 1000 steps, ending early when 40
//                     damping increases in a row find no lower sum (the minimum to rounding).  B is
//                     returned as |B|; A and C are multiplied back by the maximum, B divided by its
//                     time.
//  * hh_loess         loess_init / loess_estimate / loess_apply_smoothing (src/loess.cpp:98-451),
//                     as SMOOTH calls them (src/driver.c:2849-2951): x and y mapped to [0, 1] by
//                     their ranges (the range starts from FLT_MAX / FLT_MIN, as the reference's
//                     statics do, so an SF whose values all lie below FLT_MIN is scaled by
//                     FLT_MIN), grid point i at GRID_XMIN + i (grid_max - GRID_XMIN) / (n - 1) with
//                     GRID_XMIN = 0, a window of ws_min points (ws_min + ws_step i, truncated, for
//                     i > ws_delay) grown from the nearest data point as make_window does (ties go
//                     to the upper side; the window is the window_size indices from its lower end),
//                     tricube weights of the distance over the window's largest, and a weighted
//                     cubic fit evaluated at the grid point.  The reference fits in powers of the
//                     normalised x and solves the normal equations through Eigen's pseudo-inverse;
//                     for windows of a few thousandths of the range (beyond ~200 cm^-1 of an SF
//                     sampled every 0.05 cm^-1) that normal matrix has a condition number above
//                     1e16 and its results carry rounding noise of up to ~1 % (visible as an
//                     oscillation in its output).  The same cubic fit is solved here in the basis
//                     ((x - x_grid) / window radius)^j, well conditioned, by Gaussian elimination
//                     with partial pivoting in long double.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/hawaii_gpu.h"
#include "hh_internal.h"

using hh::fail;

namespace {

struct WingFit {
    const double* x;
    const double* y;
    long n;
    // residuals r, Jacobian columns and 1/2 |r|^2 at p
    double eval(const double* p, double* r, double (*J)[3]) const
    {
        const double A = p[0], B = p[1], C = p[2];
        double F = 0.0;
        for (long k = 0; k < n; ++k) {
            const double x2 = x[k] * x[k], den = 1.0 + B * B * x2;
            const double v = C + A / den - y[k];
            if (r) r[k] = v;
            if (J) {
                J[k][0] = 1.0 / den;
                J[k][1] = -2.0 * A * B * x2 / den / den;
                J[k][2] = 1.0;
            }
            F += 0.5 * v * v;
        }
        return F;
    }
};

// x <- solution of M x = b (n <= 4), Gaussian elimination with partial pivoting; false if singular
template <int N>
bool solve_small(long double (&M)[N][N], long double (&b)[N], int n)
{
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (fabsl(M[r][c]) > fabsl(M[piv][c])) piv = r;
        if (M[piv][c] == 0.0L) return false;
        if (piv != c) {
            for (int k = 0; k < n; ++k) {
                const long double t = M[c][k];
                M[c][k] = M[piv][k];
                M[piv][k] = t;
            }
            const long double t = b[c];
            b[c] = b[piv];
            b[piv] = t;
        }
        for (int r = c + 1; r < n; ++r) {
            const long double f = M[r][c] / M[c][c];
            for (int k = c; k < n; ++k) M[r][k] -= f * M[c][k];
            b[r] -= f * b[c];
        }
    }
    for (int c = n - 1; c >= 0; --c) {
        long double s = b[c];
        for (int k = c + 1; k < n; ++k) s -= M[c][k] * b[k];
        b[c] = s / M[c][c];
    }
    return true;
}

// Levenberg-Marquardt on the three wing parameters; p holds the start and receives the result
int lm_fit(const WingFit& fit, double* p)
{
    const double xtol = 1e-12, gtol = 1e-15;
    const int max_steps = 1000, max_rejects = 40;
    int rejects = 0;
    std::vector<double> r((size_t)fit.n), rn((size_t)fit.n), Jv(3 * (size_t)fit.n);
    double (*J)[3] = reinterpret_cast<double (*)[3]>(Jv.data());
    double F = fit.eval(p, r.data(), J);
    double D[3] = {0.0, 0.0, 0.0}, mu = -1.0, nu = 2.0;
    int steps = 0;
    while (steps < max_steps && rejects < max_rejects) {
        long double JtJ[3][3] = {}, g[3] = {};
        for (long k = 0; k < fit.n; ++k)
            for (int i = 0; i < 3; ++i) {
                g[i] += (long double)J[k][i] * r[k];
                for (int j = 0; j < 3; ++j) JtJ[i][j] += (long double)J[k][i] * J[k][j];
            }
        double gmax = 0.0;
        for (int i = 0; i < 3; ++i) {
            D[i] = std::fmax(D[i], std::sqrt((double)JtJ[i][i]));
            gmax = std::fmax(gmax, std::fabs((double)g[i]) * std::fmax(std::fabs(p[i]), 1.0));
        }
        if (gmax <= gtol * std::fmax(2.0 * F, 1.0)) return HH_OK;
        if (mu < 0) {
            double m = 0.0;
            for (int i = 0; i < 3; ++i) m = std::fmax(m, (double)JtJ[i][i] / (D[i] * D[i] > 0 ? D[i] * D[i] : 1.0));
            mu = 1e-3 * (m > 0 ? m : 1.0);
        }
        long double M[3][3], dx[3];
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) M[i][j] = JtJ[i][j];
            M[i][i] += mu * D[i] * D[i];
            dx[i] = -g[i];
        }
        if (!solve_small<3>(M, dx, 3)) return fail(HH_E_INVALID, "FIT_BASELINE: singular normal equations");
        double pn[3];
        for (int i = 0; i < 3; ++i) pn[i] = p[i] + (double)dx[i];
        const double Fn = fit.eval(pn, rn.data(), nullptr);
        long double pred = 0.0L;   // predicted decrease 1/2 dx^T (mu D^2 dx - g)
        for (int i = 0; i < 3; ++i) pred += 0.5L * dx[i] * (mu * D[i] * D[i] * dx[i] - g[i]);
        const double rho = pred > 0 ? (F - Fn) / (double)pred : -1.0;
        if (rho > 0 && std::isfinite(Fn)) {
            bool small = true;
            for (int i = 0; i < 3; ++i) small = small && std::fabs((double)dx[i]) <= xtol * (std::fabs(p[i]) + xtol);
            for (int i = 0; i < 3; ++i) p[i] = pn[i];
            F = fit.eval(p, r.data(), J);
            const double c = 2.0 * rho - 1.0;
            mu *= std::fmax(1.0 / 3.0, 1.0 - c * c * c);
            nu = 2.0;
            ++steps;
            rejects = 0;
            if (small) return HH_OK;
        } else {
            // rounding, not the model, decides once no step of any length lowers the sum
            mu *= nu;
            nu *= 2.0;
            ++rejects;
        }
    }
    return HH_OK;   // the reference reports the status and uses the parameters it has
}

}  // namespace

extern "C" {

int hh_fit_baseline(const double* t, const double* cf, long len, long begin, double* abc)
{
    if (!t || !cf || !abc || len < 2) return fail(HH_E_INVALID, "FIT_BASELINE: bad argument");
    if (begin < 0 || len <= begin)
        return fail(HH_E_INVALID, "FIT_BASELINE: cf_extrapolation_begin_index %ld must be below the CF length %ld",
                    begin, len);
    bool wasnegative = false;
    double max2time = 0.0, CFmax2 = FLT_MIN;
    for (long k = 0; k < len; ++k) {
        if (cf[k] < 0.0) wasnegative = true;
        if (wasnegative && cf[k] > CFmax2) {
            CFmax2 = cf[k];
            max2time = t[k];
        }
    }
    if (!(max2time > 0.0))
        return fail(HH_E_INVALID, "FIT_BASELINE: the correlation function has no secondary maximum after turning negative");
    const long n = len - begin + 1;
    std::vector<double> x((size_t)n, 0.0), y((size_t)n, 0.0);
    for (long k = begin, i = 0; k < len; ++k, ++i) {
        x[(size_t)i] = t[k] / max2time;
        y[(size_t)i] = cf[k] / CFmax2;
    }
    double p[3] = {0.1, 0.1, std::fmax(y[(size_t)n - 1], 1.0)};
    const int rc = lm_fit(WingFit{x.data(), y.data(), n}, p);
    if (rc < 0) return rc;
    abc[0] = p[0] * CFmax2;
    abc[1] = std::fabs(p[1]) / max2time;
    abc[2] = p[2] * CFmax2;
    return HH_OK;
}

int hh_loess(const double* x, const double* y, long len, double grid_max, long npoints, long ws_min, double ws_step,
             long ws_delay, double* nu_out, double* y_out)
{
    if (!x || !y || !nu_out || !y_out) return fail(HH_E_INVALID, "SMOOTH: bad argument");
    if (len < 2) return fail(HH_E_INVALID, "SMOOTH: expected length of arrays to be at least 2");
    if (npoints <= 0) return fail(HH_E_INVALID, "SMOOTH: number of grid points (grid_npoints) must be provided");
    if (!(grid_max > 0.0)) return fail(HH_E_INVALID, "SMOOTH: the ending point of the grid (grid_max) must be provided");
    double minx = FLT_MAX, maxx = FLT_MIN, miny = FLT_MAX, maxy = FLT_MIN;
    for (long i = 0; i < len; ++i) {
        if (x[i] > maxx) maxx = x[i];
        if (x[i] < minx) minx = x[i];
        if (y[i] > maxy) maxy = y[i];
        if (y[i] < miny) miny = y[i];
    }
    const double gmin = 0.0;
    if (gmin < minx) return fail(HH_E_INVALID, "SMOOTH: the grid starts (0) below the smallest frequency (%.5e)", minx);
    if (grid_max > maxx)
        return fail(HH_E_INVALID, "SMOOTH: the ending point of the grid (%.5e) is greater than the maximum frequency (%.5e)",
                    grid_max, maxx);
    std::vector<double> xp((size_t)len), yp((size_t)len), dist((size_t)len);
    for (long i = 0; i < len; ++i) {
        xp[(size_t)i] = (x[i] - minx) / (maxx - minx);
        yp[(size_t)i] = (y[i] - miny) / (maxy - miny);
    }
    const double step = npoints > 1 ? (grid_max - gmin) / (double)(npoints - 1) : 0.0;
    for (long i = 0; i < npoints; ++i) {
        const double xg = gmin + step * (double)i;
        nu_out[i] = xg;
        long ws = ws_min;
        if (i > ws_delay) ws = (long)((double)ws_min + ws_step * (double)i);
        if (ws < 4 || ws > len)
            return fail(HH_E_INVALID, "SMOOTH: window of %ld points at grid point %ld (need 4..%ld)", ws, i, len);
        const double nx = (xg - minx) / (maxx - minx);
        long imin = 0;
        for (long k = 0; k < len; ++k) {
            dist[(size_t)k] = std::fabs(xp[(size_t)k] - nx);
            if (dist[(size_t)k] < dist[(size_t)imin]) imin = k;
        }
        long lo;
        if (imin == 0) {
            lo = 0;
        } else if (imin == len - 1) {
            lo = len - ws;
        } else {
            lo = imin;
            long up = imin + 1;
            for (long c = 0; c < ws; ++c) {
                if (lo == 0)
                    ++up;
                else if (up == len - 1)
                    --lo;
                else if (dist[(size_t)lo - 1] < dist[(size_t)up + 1])
                    --lo;
                else
                    ++up;
            }
            if (lo + ws > len) return fail(HH_E_INVALID, "SMOOTH: window past the data at grid point %ld", i);
        }
        double dmax = FLT_MIN;
        for (long k = lo; k < lo + ws; ++k) dmax = std::fmax(dmax, dist[(size_t)k]);
        // the weighted cubic least-squares fit, in the basis ((x - nx) / dmax)^j: the same cubic
        // as the reference's fit in powers of x, whose normal matrix is too ill-conditioned in a
        // window a few thousandths of the data range wide to be solved to more than a few digits;
        // its value at the grid point is the constant coefficient
        long double M[4][4] = {}, b[4] = {};
        for (long k = lo; k < lo + ws; ++k) {
            const double u = dist[(size_t)k] / dmax, w1 = 1.0 - u * u * u, w = w1 * w1 * w1;
            const long double v = ((long double)xp[(size_t)k] - nx) / dmax;
            const long double pw[4] = {1.0L, v, v * v, v * v * v};
            for (int a = 0; a < 4; ++a) {
                b[a] += w * pw[a] * yp[(size_t)k];
                for (int c = 0; c < 4; ++c) M[a][c] += w * pw[a] * pw[c];
            }
        }
        if (!solve_small<4>(M, b, 4)) return fail(HH_E_INVALID, "SMOOTH: singular local fit at grid point %ld", i);
        y_out[i] = (double)b[0] * (maxy - miny) + miny;
    }
    return HH_OK;
}

}  // extern "C"
