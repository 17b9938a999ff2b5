/* Initial-condition sampler (test infrastructure).
 *
 * Draw order and arithmetic follow the reference exactly:
 *   generate_normal  src/hawaii.c:1206-1230  (Box-Muller, cosine branch, retry on non-finite log U)
 *   q_generator      src/hawaii.c:1232-1290  (R^3 uniform, Phi, Theta, then monomer angles)
 *   p_generator      src/hawaii.c:1330-1380  (x0, x1, x2 for the pair, then the rotor's x0, x1)
 *   p_generator_linear_molecule src/hawaii.c:1292-1303
 *   reject           src/hawaii.c:1382-1421  (u drawn first; weight = desired/proposal/M < u)
 *   candidate loop   src/hawaii.c:2969-2993  (Hamiltonian before reject; FREE keeps E >= 0)
 *   R0 injection     src/hawaii.c:3486-3512  (q_generator, R := R0, p_generator, pR := -|pR|)
 *   calculate_M0     src/hawaii.c:1423-1491, calculate_M2 src/hawaii.c:1524-1617 (same loop)
 */
#define _GNU_SOURCE   /* sincos */
#include <math.h>
#include "oracle.h"

#define HkT (4.3597447222071e-18 / 1.380649e-23)   /* src/constants.h:24 */

typedef struct {
    orc_mt mt;
    long draws;
} rng;

static double drand(rng *r)
{
    r->draws++;
    return orc_mt_drand(&r->mt);
}

static double generate_normal(rng *r, double sigma)
{
    for (;;) {
        double U = drand(r);
        double V = drand(r);
        double logU = log(U);
        if (isnan(logU) || isinf(logU)) continue;
        return sigma * sqrt(-2 * logU) * cos(2.0 * M_PI * V);
    }
}

static void q_generator(const orc_system *s, const orc_sampler *sp, rng *r, double *y)
{
    double RMIN = sp->Rmin, RMAX = sp->Rmax;
    double R3 = drand(r) * (RMAX * RMAX * RMAX - RMIN * RMIN * RMIN) + RMIN * RMIN * RMIN;
    y[4] = pow(R3, 1.0 / 3.0);
    y[0] = drand(r) * 2.0 * M_PI;
    y[2] = acos(2.0 * drand(r) - 1.0);
    if (s->system == ORC_SYS_CO2AR) {
        y[6] = drand(r) * 2.0 * M_PI;
        y[8] = acos(2.0 * drand(r) - 1.0);
    }
    if (s->system == ORC_SYS_CH4CO2) {       /* monomer 1 ROTOR, monomer 2 LINEAR_MOLECULE */
        y[6] = drand(r) * 2.0 * M_PI;
        y[8] = acos(2.0 * drand(r) - 1.0);
        y[10] = drand(r) * 2.0 * M_PI;
        y[12] = drand(r) * 2.0 * M_PI;
        y[14] = acos(2.0 * drand(r) - 1.0);
    }
}

/* p_generator_linear_molecule (src/hawaii.c:1292-1303): monomer angles m[0] = phi, m[2] = theta */
static void p_linear(double II0, double T, rng *r, double *m)
{
    double sqrt_IIKT = sqrt(II0 / HkT * T);
    double sin_theta = sin(m[2]);
    double x0 = generate_normal(r, 1.0);
    double x1 = generate_normal(r, 1.0);
    m[1] = sqrt_IIKT * sin_theta * x1;
    m[3] = sqrt_IIKT * x0;
}

/* p_generator_rotor (src/hawaii.c:1305-1327): m = (phi, pphi, theta, ptheta, psi, ppsi) */
static void p_rotor(const double *II, double T, rng *r, double *m)
{
    double theta = m[2], psi = m[4];
    double sqrt_IIx = sqrt(II[0] / HkT * T), sqrt_IIy = sqrt(II[1] / HkT * T), sqrt_IIz = sqrt(II[2] / HkT * T);
    double sin_theta, cos_theta, sin_psi, cos_psi;
    double x0, x1, x2;
    /* the reference calls sincos, whose cosine can differ from cos in the last bit */
    sincos(theta, &sin_theta, &cos_theta);
    sincos(psi, &sin_psi, &cos_psi);
    x0 = generate_normal(r, 1.0);
    x1 = generate_normal(r, 1.0);
    x2 = generate_normal(r, 1.0);
    m[1] = sqrt_IIx * sin_theta * sin_psi * x0 + sqrt_IIy * sin_theta * cos_psi * x1 + sqrt_IIz * cos_theta * x2;
    m[3] = sqrt_IIx * cos_psi * x0 - sqrt_IIy * sin_psi * x1;
    m[5] = sqrt_IIz * x2;
}

static void p_generator(const orc_system *s, double T, rng *r, double *y)
{
    double sqrt_MUKT = sqrt(s->mu / HkT * T);
    double sinTheta = sin(y[2]);
    double x[3];
    int i;
    for (i = 0; i < 3; ++i) x[i] = generate_normal(r, 1.0);
    y[1] = sqrt_MUKT * y[4] * sinTheta * x[0];
    y[3] = sqrt_MUKT * y[4] * x[1];
    y[5] = sqrt_MUKT * x[2];
    if (s->system == ORC_SYS_CO2AR) p_linear(s->II[0], T, r, y + 6);
    if (s->system == ORC_SYS_CH4CO2) {
        p_rotor(s->IR, T, r, y + 6);
        p_linear(s->II[0], T, r, y + 12);
    }
}

/* true = rejected */
static int reject(const orc_system *s, const double *y, double T, double pesmin, rng *r)
{
    double u = drand(r);
    double Mval = exp(-pesmin * HkT / T);
    double M = 1.5 * Mval;
    double kin_part = orc_kinetic(s, y);
    double q[8];
    int i, nq = orc_qp_size(s) / 2;
    double pot_part, proposal, desired, weight;
    for (i = 0; i < nq; ++i) q[i] = y[2 * i];
    pot_part = orc_pes(s, q);
    proposal = exp(-kin_part * HkT / T);
    desired = exp(-(kin_part + pot_part) * HkT / T);
    weight = desired / proposal / M;
    return weight < u;
}

long orc_sample_ics(const orc_system *s, const orc_sampler *sp, uint32_t seed,
                    long ntraj, double *ic, long *n_draws)
{
    rng r;
    long accepted = 0, candidates = 0;
    int qp = orc_qp_size(s), i;
    double y[16];
    r.draws = 0;
    orc_mt_seed(&r.mt, seed);

    while (accepted < ntraj) {
        for (i = 0; i < 16; ++i) y[i] = 0.0;
        q_generator(s, sp, &r, y);
        if (sp->mode == ORC_IC_R0_INJECT) {
            y[4] = sp->R0;
            p_generator(s, sp->T, &r, y);
            if (y[5] > 0.0) y[5] = -y[5];
            ++candidates;
        } else if (sp->mode == ORC_IC_PHASE_SPACE) {   /* q_generator + p_generator only */
            p_generator(s, sp->T, &r, y);
            ++candidates;
        } else {
            double energy;
            p_generator(s, sp->T, &r, y);
            energy = orc_hamiltonian(s, y);
            ++candidates;
            if (reject(s, y, sp->T, sp->pesmin, &r)) continue;
            if (sp->pair_state == ORC_PAIR_FREE && energy < 0.0) continue;
            if (sp->pair_state == ORC_PAIR_BOUND && energy > 0.0) continue;
        }
        for (i = 0; i < qp; ++i) ic[accepted * qp + i] = y[i];
        ++accepted;
    }
    if (n_draws) *n_draws = r.draws;
    return candidates;
}

double orc_dipole_rate_sq(const orc_system *s, const double *y)
{
    const double h = 1.0e-3;
    int i, k, nq = orc_qp_size(s) / 2;
    double ydot[16], qv[8], D[8][3], md[3];
    orc_rhs(s, y, ydot);
    for (i = 0; i < nq; ++i) qv[i] = y[2 * i];
    for (i = 0; i < nq; ++i) {
        double dp[3], dm[3], tmp = qv[i];
        qv[i] = tmp + h;
        orc_dipole(s, qv, dp);
        qv[i] = tmp - h;
        orc_dipole(s, qv, dm);
        for (k = 0; k < 3; ++k) D[i][k] = (dp[k] - dm[k]) / (2.0 * h);
        qv[i] = tmp;
    }
    /* dip_lab_dot = dHdp (1 x nq) . D (nq x 3), summed over the coordinates in order */
    for (k = 0; k < 3; ++k) {
        md[k] = 0.0;
        for (i = 0; i < nq; ++i) md[k] += ydot[2 * i] * D[i][k];
    }
    return md[0] * md[0] + md[1] * md[1] + md[2] * md[2];
}

/* the candidate loop shared by calculate_M0 and calculate_M2 (order 0 or 2) */
static void moment_loop(const orc_system *s, const orc_sampler *sp, uint32_t seed, long npoints, double ppf,
                        int order, double *m, double *q)
{
    const double ZeroCoeff = 0.00361479637 / (4.0 * M_PI);   /* src/constants.h:29 */
    const double SecondCoeff = 13856114.29344114;            /* src/constants.h:30 */
    const double coeff = order == 2 ? SecondCoeff : ZeroCoeff;
    rng r;
    long n = 0;
    int i, nq = orc_qp_size(s) / 2;
    double y[16], qv[8], d[3];
    double mm = 0.0, qq = 0.0;
    r.draws = 0;
    orc_mt_seed(&r.mt, seed);
    while (n < npoints) {
        double energy, fval, diff;
        for (i = 0; i < 16; ++i) y[i] = 0.0;
        q_generator(s, sp, &r, y);
        p_generator(s, sp->T, &r, y);
        energy = orc_hamiltonian(s, y);
        if (reject(s, y, sp->T, sp->pesmin, &r)) continue;
        if (sp->pair_state == ORC_PAIR_FREE && energy < 0.0) continue;
        if (sp->pair_state == ORC_PAIR_BOUND && energy > 0.0) continue;
        if (order == 2) {
            fval = orc_dipole_rate_sq(s, y);
        } else {
            for (i = 0; i < nq; ++i) qv[i] = y[2 * i];
            orc_dipole(s, qv, d);
            fval = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        }
        diff = fval - mm;
        mm += diff / (n + 1.0);
        qq += diff * diff * (n / (n + 1.0));
        n++;
    }
    *m = mm * coeff * ppf;
    *q = sqrt(qq / n / (n - 1)) * coeff * ppf;
}

void orc_m0(const orc_system *s, const orc_sampler *sp, uint32_t seed, long npoints, double ppf,
            double *m, double *q)
{
    moment_loop(s, sp, seed, npoints, ppf, 0, m, q);
}

void orc_m2(const orc_system *s, const orc_sampler *sp, uint32_t seed, long npoints, double ppf,
            double *m, double *q)
{
    moment_loop(s, sp, seed, npoints, ppf, 2, m, q);
}
