/* He-Ar and CO2-Ar surfaces and Hamilton's equations (test infrastructure).
 *
 * Equations of motion: src/hawaii.c:474-615 (rhsMonomer, rhs), kinetic energy
 * src/hawaii.c:1041-1137, Hamiltonian src/hawaii.c:1190-1204.  The arithmetic
 * is written in the same order as the reference so that, built without FMA
 * contraction, it rounds the same way.
 *
 * State layout (src/hawaii.h:64-72): [Phi, pPhi, Theta, pTheta, R, pR | phi1, pphi1,
 * theta1, ptheta1]; coordinate vector q = [Phi, Theta, R | phi1, theta1].
 */
#include <math.h>
#include "oracle.h"
#include "co2ar_ref_table.h"

#define HTOCM 2.1947463136320e5   /* src/constants.h:11 */

int orc_qp_size(const orc_system *s)
{
    return s->system == ORC_SYS_CH4CO2 ? 16 : s->system == ORC_SYS_CO2AR ? 10 : 6;
}

/* ------------------------------------------------------------------ He-Ar */
/* PES-IDS/HeAr.h:61-94.  The surface is a Tang-Toennies form
 *   V = A exp(-a1 R - a2 R^2) - sum_{n=3..8} C_2n f_2n(R) / R^2n,
 *   f_2n = 1 - exp(-b R) sum_{k=0}^{2n} (b R)^k / k!
 * restated here with the damping sums built from b (the reference prints the
 * b^k/k! terms as 14-digit decimals).  dV/dR is the analytic derivative:
 *   d/dR[exp(-bR) S_m(bR)] = -b exp(-bR) (bR)^m / m!. */
static const double HE_A = 0.2303058634e2, HE_A1 = 0.163051044e1, HE_A2 = 0.4680016e-1;
static const double HE_B = 0.16274792500000e1;
static const double HE_C[6] = {0.940835513e1, 0.165523018e3, 0.379715796e4,
                               0.1165179999e6, 0.466258000e7, 0.236861000e9};

double orc_hear_V(double R)
{
    double x = HE_B * R, e = exp(-x);
    double term = 1.0, sum = 1.0, V = HE_A * exp(-HE_A1 * R - HE_A2 * R * R);
    int k, n;
    double R2n = R * R * R * R; /* becomes R^6 in the first pass */
    for (n = 3, k = 0; n <= 8; ++n) {
        R2n *= R * R;
        while (k < 2 * n) { ++k; term *= x / k; sum += term; }
        V -= HE_C[n - 3] * (1.0 - e * sum) / R2n;
    }
    return V;
}

double orc_hear_dV(double R)
{
    double x = HE_B * R, e = exp(-x);
    double term = 1.0, sum = 1.0;
    double dV = HE_A * exp(-HE_A1 * R - HE_A2 * R * R) * (-HE_A1 - 2.0 * HE_A2 * R);
    int k, n;
    double R2n = R * R * R * R;
    for (n = 3, k = 0; n <= 8; ++n) {
        R2n *= R * R;
        while (k < 2 * n) { ++k; term *= x / k; sum += term; }
        /* d/dR [C f / R^2n] = C [ f' / R^2n - 2n f / R^(2n+1) ],  f' = b e term */
        dV -= HE_C[n - 3] * (HE_B * e * term / R2n - 2.0 * n * (1.0 - e * sum) / (R2n * R));
    }
    return dV;
}

double orc_hear_dip(double R)
{
    /* PES-IDS/HeAr.h:44-47 */
    return exp(0.47293181 - 0.40100613 * R - 0.10292726 * R * R) - 148.55391 * pow(R, -7);
}

/* ------------------------------------------------------------------ CO2-Ar */
/* Legendre P_l(x), l = 0..10 (the values gsl_sf_legendre_Pl_array returns). */
static void legendre_P(double x, double *P, int lmax)
{
    int l;
    P[0] = 1.0;
    if (lmax >= 1) P[1] = x;
    for (l = 1; l < lmax; ++l) P[l + 1] = ((2 * l + 1) * x * P[l] - l * P[l - 1]) / (l + 1);
}

/* dP_l/dx by P'_{l+1} = P'_{l-1} + (2l+1) P_l (no division by x^2 - 1). */
static void legendre_dP(const double *P, double *dP, int lmax)
{
    int l;
    dP[0] = 0.0;
    if (lmax >= 1) dP[1] = 1.0;
    for (l = 1; l < lmax; ++l) dP[l + 1] = dP[l - 1] + (2 * l + 1) * P[l];
}

/* Associated P_l^1(x) with the Condon-Shortley phase, GSL_SF_LEGENDRE_NONE
 * normalisation (what ai_ids_co2ar.cpp:12-19 reads at index l(l+1)/2 + 1). */
static void legendre_P1(double x, double *P1, int lmax)
{
    int l;
    double s = sqrt((1.0 - x) * (1.0 + x));
    P1[0] = 0.0;
    P1[1] = -s;
    for (l = 1; l < lmax; ++l) P1[l + 1] = ((2 * l + 1) * x * P1[l] - (l + 1) * P1[l - 1]) / l;
}

static int co2ar_branch(double R)
{
    int b;
    for (b = 0; b < CO2AR_NBRANCH - 1; ++b)
        if (R < CO2AR_BOUNDS[b]) return b;
    return CO2AR_NBRANCH - 1;
}

/* radial coefficient: sum_k a_k R^k (k = -7..3) + A exp(alpha R) */
static double radial(const double *poly, const double *ex, double R)
{
    double v = poly[7] + poly[8] * R + poly[9] * (R * R) + poly[10] * (R * R * R);
    double inv = 1.0 / R, ip = 1.0;
    int k;
    for (k = 1; k <= 7; ++k) {
        ip *= inv;
        if (poly[7 - k] != 0.0) v += poly[7 - k] * ip;
    }
    if (ex[0] != 0.0) v += ex[0] * exp(ex[1] * R);
    return v;
}

static double dradial(const double *poly, const double *ex, double R)
{
    double v = poly[8] + 2.0 * poly[9] * R + 3.0 * poly[10] * (R * R);
    double inv = 1.0 / R, ip = inv;
    int k;
    for (k = 1; k <= 7; ++k) {
        ip *= inv;
        if (poly[7 - k] != 0.0) v -= k * poly[7 - k] * ip;
    }
    if (ex[0] != 0.0) v += ex[0] * ex[1] * exp(ex[1] * R);
    return v;
}

/* pes_co2ar (PES-IDS/ai_pes_co2ar.c:27-459): sum_{l even} c_l(R) P_l(cos theta), cm^-1 */
double orc_co2ar_V_cm(double R, double theta)
{
    double P[11];
    int b = co2ar_branch(R), s;
    double V = 0.0;
    legendre_P(cos(theta), P, 10);
    for (s = 0; s < 6; ++s) V += radial(CO2AR_PES_POLY[b][s], CO2AR_PES_EXP[b][s], R) * P[2 * s];
    return V;
}

/* dpes_dR / dpes_dTheta (PES-IDS/ai_pes_co2ar.c:461-1028), cm^-1 per bohr / rad */
void orc_co2ar_dV_cm(double R, double theta, double *dR, double *dTheta)
{
    double P[11], dP[11];
    double x = cos(theta), sn = sin(theta);
    int b = co2ar_branch(R), s;
    legendre_P(x, P, 10);
    legendre_dP(P, dP, 10);
    *dR = 0.0;
    *dTheta = 0.0;
    for (s = 0; s < 6; ++s) {
        *dR += dradial(CO2AR_PES_POLY[b][s], CO2AR_PES_EXP[b][s], R) * P[2 * s];
        *dTheta += radial(CO2AR_PES_POLY[b][s], CO2AR_PES_EXP[b][s], R) * (-sn * dP[2 * s]);
    }
}

/* AI_IDS_co2_ar::dipole_vector (PES-IDS/ai_ids_co2ar.cpp:12-19): (dipx, 0, dipz) */
void orc_co2ar_dip_mol(double R, double theta, double d[3])
{
    double x = cos(theta), P[11], P1[11];
    int b = co2ar_branch(R), s;
    legendre_P(x, P, 10);
    legendre_P1(x, P1, 10);
    d[0] = 0.0;
    d[1] = 0.0;
    d[2] = 0.0;
    for (s = 0; s < 5; ++s) d[0] += radial(CO2AR_DIPX_POLY[b][s], CO2AR_DIPX_EXP[b][s], R) * P1[2 * s + 2];
    for (s = 0; s < 6; ++s) d[2] += radial(CO2AR_DIPZ_POLY[b][s], CO2AR_DIPZ_EXP[b][s], R) * P[2 * s];
}

/* 3x3 helpers with the reference's rotation conventions (src/angles_handler.cpp:45-100) */
typedef double m3[3][3];
static void Sz(m3 m, double s, double c) { m[0][0] = c; m[0][1] = s; m[0][2] = 0; m[1][0] = -s; m[1][1] = c; m[1][2] = 0; m[2][0] = 0; m[2][1] = 0; m[2][2] = 1; }
static void Sx(m3 m, double s, double c) { m[0][0] = 1; m[0][1] = 0; m[0][2] = 0; m[1][0] = 0; m[1][1] = c; m[1][2] = s; m[2][0] = 0; m[2][1] = -s; m[2][2] = c; }
static void Sy(m3 m, double s, double c) { m[0][0] = c; m[0][1] = 0; m[0][2] = -s; m[1][0] = 0; m[1][1] = 1; m[1][2] = 0; m[2][0] = s; m[2][1] = 0; m[2][2] = c; }
static void Sz_dot(m3 m, double s, double c) { m[0][0] = -s; m[0][1] = c; m[0][2] = 0; m[1][0] = -c; m[1][1] = -s; m[1][2] = 0; m[2][0] = 0; m[2][1] = 0; m[2][2] = 0; }
static void Sx_dot(m3 m, double s, double c) { m[0][0] = 0; m[0][1] = 0; m[0][2] = 0; m[1][0] = 0; m[1][1] = -s; m[1][2] = c; m[2][0] = 0; m[2][1] = -c; m[2][2] = -s; }
static void Sy_dot(m3 m, double s, double c) { m[0][0] = -s; m[0][1] = 0; m[0][2] = -c; m[1][0] = 0; m[1][1] = 0; m[1][2] = 0; m[2][0] = c; m[2][1] = 0; m[2][2] = -s; }
static void mv(m3 m, const double *v, double *o) { int i; for (i = 0; i < 3; ++i) o[i] = m[i][0] * v[0] + m[i][1] * v[1] + m[i][2] * v[2]; }
static void mtv(m3 m, const double *v, double *o) { int i; for (i = 0; i < 3; ++i) o[i] = m[0][i] * v[0] + m[1][i] * v[1] + m[2][i] * v[2]; }

/* dd = A * B * C^T * D^T * z */
static void chain(m3 A, m3 B, m3 C, m3 D, double *dd)
{
    double z[3] = {0, 0, 1}, t1[3], t2[3], t3[3];
    mtv(D, z, t1);
    mtv(C, t1, t2);
    mv(B, t2, t3);
    mv(A, t3, dd);
}

/* linear_molecule_atom_lab_to_mol (src/angles_handler.cpp:106-138)
 * qlab: Phi Theta R phi1 theta1 -> qmol: R phiem thetaem psiem thetam */
void orc_co2ar_lab_to_mol(const double *qlab, double *qmol)
{
    m3 Sphiem, Sthetaem, Sphi1, Stheta1;
    double dd[3], s;
    qmol[0] = qlab[2];
    qmol[1] = qlab[0] + M_PI / 2.0;
    qmol[2] = qlab[1];
    Sz(Sphiem, sin(qmol[1]), cos(qmol[1]));
    Sx(Sthetaem, sin(qmol[2]), cos(qmol[2]));
    Sz(Sphi1, sin(qlab[3]), cos(qlab[3]));
    Sy(Stheta1, sin(qlab[4]), cos(qlab[4]));
    chain(Sthetaem, Sphiem, Sphi1, Stheta1, dd);
    qmol[4] = acos(dd[2]);
    s = sqrt(1 - dd[2] * dd[2]);
    qmol[3] = atan2(dd[1] / s, dd[0] / s);
}

/* linear_molecule_atom_Jacobi_mol_by_lab (src/angles_handler.cpp:140-193):
 * jac[i][j] = d qmol_j / d qlab_i for the entries the gradient needs. */
void orc_co2ar_jacobian(const double *qlab, const double *qmol, double jac[5][5])
{
    m3 Sphiem, Sthetaem, Sphi1, Stheta1, Dphiem, Dthetaem, Dphi1, Dtheta1;
    double dd[3], sp = sin(qmol[3]), cp = cos(qmol[3]), st = sin(qmol[4]);
    int i, j;
    for (i = 0; i < 5; ++i) for (j = 0; j < 5; ++j) jac[i][j] = 0.0;
    jac[2][0] = 1.0;
    jac[0][1] = 1.0;
    jac[1][2] = 1.0;
    Sz(Sphiem, sin(qmol[1]), cos(qmol[1]));
    Sx(Sthetaem, sin(qmol[2]), cos(qmol[2]));
    Sz(Sphi1, sin(qlab[3]), cos(qlab[3]));
    Sy(Stheta1, sin(qlab[4]), cos(qlab[4]));
    Sx_dot(Dthetaem, sin(qmol[2]), cos(qmol[2]));
    Sz_dot(Dphiem, sin(qmol[1]), cos(qmol[1]));
    Sz_dot(Dphi1, sin(qlab[3]), cos(qlab[3]));
    Sy_dot(Dtheta1, sin(qlab[4]), cos(qlab[4]));

    chain(Sthetaem, Sphiem, Dphi1, Stheta1, dd);
    jac[3][4] = -dd[2] / st;
    jac[3][3] = (dd[1] * cp - dd[0] * sp) / st;
    chain(Sthetaem, Sphiem, Sphi1, Dtheta1, dd);
    jac[4][4] = -dd[2] / st;
    jac[4][3] = (dd[1] * cp - dd[0] * sp) / st;
    chain(Sthetaem, Dphiem, Sphi1, Stheta1, dd);
    jac[0][4] = -dd[2] / st;
    jac[0][3] = (dd[1] * cp - dd[0] * sp) / st;
    chain(Dthetaem, Sphiem, Sphi1, Stheta1, dd);
    jac[1][4] = -dd[2] / st;
    jac[1][3] = (dd[1] * cp - dd[0] * sp) / st;
}

/* ------------------------------------------------------------------ plugin-level API */
double orc_pes(const orc_system *s, const double *q)
{
    if (s->system == ORC_SYS_CH4CO2) return orc_ch4co2_pes(q);
    if (s->system == ORC_SYS_HEAR) return orc_hear_V(q[2]);
    {
        double qmol[5];
        orc_co2ar_lab_to_mol(q, qmol);                     /* ai_pes_co2ar_lib.cpp:6-10 */
        return orc_co2ar_V_cm(qmol[0], qmol[4]) / HTOCM;
    }
}

void orc_dpes(const orc_system *s, const double *q, double *dvdq)
{
    if (s->system == ORC_SYS_CH4CO2) {
        orc_ch4co2_dpes(q, dvdq);
        return;
    }
    if (s->system == ORC_SYS_HEAR) {                      /* HeAr.h:17-21 */
        dvdq[0] = 0.0;
        dvdq[1] = 0.0;
        dvdq[2] = orc_hear_dV(q[2]);
        return;
    }
    {                                                      /* ai_pes_co2ar_lib.cpp:12-29 */
        double qmol[5], jac[5][5], dR, dT, dm[5] = {0, 0, 0, 0, 0};
        int i, j;
        orc_co2ar_lab_to_mol(q, qmol);
        orc_co2ar_jacobian(q, qmol, jac);
        orc_co2ar_dV_cm(qmol[0], qmol[4], &dR, &dT);
        dm[0] = dR / HTOCM;
        dm[4] = dT / HTOCM;
        for (i = 0; i < 5; ++i) {
            double acc = 0.0;
            for (j = 0; j < 5; ++j) acc += jac[i][j] * dm[j];
            dvdq[i] = acc;
        }
    }
}

void orc_dipole(const orc_system *s, const double *q, double mu[3])
{
    if (s->system == ORC_SYS_CH4CO2) {
        orc_ch4co2_dipole(q, mu);
        return;
    }
    if (s->system == ORC_SYS_HEAR) {                      /* HeAr.h:23-31 */
        double d = orc_hear_dip(q[2]);
        mu[0] = d * sin(q[1]) * cos(q[0]);
        mu[1] = d * sin(q[1]) * sin(q[0]);
        mu[2] = d * cos(q[1]);
        return;
    }
    {                                                      /* ai_ids_co2ar_lib.cpp:12-37 */
        double qmol[5], dm[3];
        orc_co2ar_lab_to_mol(q, qmol);
        orc_co2ar_dip_mol(qmol[0], qmol[4], dm);
        orc_lm_mol_to_lab(qmol, dm, mu);
    }
}

/* molecular-frame vector -> lab: Sphiem^T Sthetaem^T Spsiem^T d (ai_ids_co2ar_lib.cpp:25-33) */
void orc_lm_mol_to_lab(const double *qmol, const double dm[3], double mu[3])
{
    double t1[3], t2[3];
    m3 Sphiem, Sthetaem, Spsiem;
    Sz(Sphiem, sin(qmol[1]), cos(qmol[1]));
    Sx(Sthetaem, sin(qmol[2]), cos(qmol[2]));
    Sz(Spsiem, sin(qmol[3]), cos(qmol[3]));
    mtv(Spsiem, dm, t1);
    mtv(Sthetaem, t1, t2);
    mtv(Sphiem, t2, mu);
}

static void extract_q(const orc_system *s, const double *y, double *q)
{
    int i, n = orc_qp_size(s) / 2;
    for (i = 0; i < n; ++i) q[i] = y[2 * i];
}

void orc_rhs(const orc_system *s, const double *y, double *ydot)
{
    double q[5], dV[5];
    if (s->system == ORC_SYS_CH4CO2) {
        orc_ch4co2_rhs(s, y, ydot);
        return;
    }
    double pPhi = y[1], Theta = y[2], pTheta = y[3], R = y[4], pR = y[5];
    double mu = s->mu;
    double R2, R3, sinTheta, cosTheta, sinTheta2, sinTheta3;
    extract_q(s, y, q);
    orc_dpes(s, q, dV);

    R2 = R * R;
    R3 = R2 * R;
    sinTheta = sin(Theta);
    cosTheta = cos(Theta);
    sinTheta2 = sinTheta * sinTheta;
    sinTheta3 = sinTheta2 * sinTheta;

    /* src/hawaii.c:586-591 */
    ydot[4] = pR / mu;
    ydot[5] = pTheta * pTheta / (mu * R3) + pPhi * pPhi / (mu * R3 * sinTheta2) - dV[2];
    ydot[0] = pPhi / (mu * R2 * sinTheta2);
    ydot[1] = -dV[0];
    ydot[2] = pTheta / (mu * R2);
    ydot[3] = pPhi * pPhi * cosTheta / (mu * R2 * sinTheta3) - dV[1];

    if (s->system == ORC_SYS_CO2AR) {                     /* src/hawaii.c:494-497 */
        double I = s->II[0];
        double pphi = y[7], theta = y[8], ptheta = y[9];
        double st = sin(theta), ct = cos(theta);
        ydot[6] = pphi / I / st / st;
        ydot[7] = -dV[3];
        ydot[8] = ptheta / I;
        ydot[9] = pphi * pphi * ct / I / st / st / st - dV[4];
    }
}

double orc_kinetic(const orc_system *s, const double *y)
{
    if (s->system == ORC_SYS_CH4CO2) return orc_ch4co2_kinetic(s, y);
    double pPhi = y[1], Theta = y[2], pTheta = y[3], R = y[4], pR = y[5], mu = s->mu;
    double sinTheta = sin(Theta);
    double K = pR * pR / (2.0 * mu) + pTheta * pTheta / (2.0 * mu * R * R) +
               pPhi * pPhi / (2.0 * mu * R * R * sinTheta * sinTheta);
    if (s->system == ORC_SYS_CO2AR) {                     /* src/hawaii.c:1076-1086 */
        double st = sin(y[8]);
        K += y[9] * y[9] / (2.0 * s->II[0]) + y[7] * y[7] / (2.0 * s->II[1] * st * st);
    }
    return K;
}

double orc_hamiltonian(const orc_system *s, const double *y)
{
    double q[8];
    double V, K;
    extract_q(s, y, q);
    V = orc_pes(s, q);
    K = orc_kinetic(s, y);
    return K + V;
}

double orc_pf_analytic(const orc_system *s, double T)
{
    const double HkT = 4.3597447222071e-18 / 1.380649e-23;
    if (s->system == ORC_SYS_HEAR) return pow(2.0 * M_PI * s->mu * T / HkT, 1.5);
    return 4.0 * M_PI * pow(2.0 * M_PI * T / HkT, 2.5) * pow(s->mu, 1.5) * s->II[0];
}
