/* CH4-CO2 (ROTOR + LINEAR_MOLECULE) surfaces and equations of motion (test infrastructure).
 *
 * Lab coordinates q = [Phi, Theta, R, phi1, theta1, psi1, phi2, theta2] (monomer 1 = CH4, a
 * rotor with Euler angles phi1, theta1, psi1; monomer 2 = CO2, a linear molecule); state
 * [Phi, pPhi, Theta, pTheta, R, pR | phi1, pphi1, theta1, ptheta1, psi1, ppsi1 | phi2, pphi2,
 * theta2, ptheta2].
 *
 * Followed:
 *   frame of CH4 ("kal")     CH4_linear_molecule_lab_to_kal, src/angles_handler.cpp:357-419:
 *                            S1 = Spsi1 Stheta1 Sphi1, dd = S1 SPhi^T STheta^T z (R in the CH4
 *                            frame), dd2 = S1 Sphi2^T Stheta2^T z (CO2 axis), angles by acos/atan2
 *   Jacobian (8 x 5)         CH4_linear_molecule_Jacobi_kal_by_lab, src/angles_handler.cpp:421-555
 *   plugin wrappers          PES-IDS/ai_pes_ch4_co2_lib.cpp:8-37 (gradient = J dV_kal),
 *                            PES-IDS/ai_ids_ch4_co2_lib.cpp:8-34 (mu_lab = S1^T mu_kal)
 *   dipole in the CH4 frame  dipole_vector, PES-IDS/ai_ids_ch4_co2.cpp:672-1066, from the tables
 *                            that hawaiihybrid_amd/data/gen_ch4co2_tables.py expands out of the
 *                            reference's generated basis and radial functions; associated Legendre
 *                            functions by the textbook recurrence with the Condon-Shortley phase
 *                            (GSL_SF_LEGENDRE_NONE, csphase = -1), cos(m phi) and sin(m phi) as
 *                            the reference computes them
 *   rotor equations          rhsMonomer ROTOR, src/hawaii.c:505-531
 *   kinetic energy           kinetic_energy, src/hawaii.c:1041-1137 (ROTOR m1, LINEAR m2)
 *
 * The potential is a stand-in: the reference's CH4-CO2 PES sources are not in its repository
 * (.MISSING_LARGE_BLOBS lists PES-IDS/ai_pes_ch4_co2*.c).  It keeps their signature
 * (pes_ch4co2(R, phi1, theta1, phi2, theta2) and the five derivatives, PES-IDS/ai_pes_ch4_co2.h:
 * 19-29, angles of the kal frame) and is a sum of Lennard-Jones site-site terms between the
 * CH4 sites (C at the origin, H at r_CH (+-1, +-1, +-1)/sqrt 3 with an even number of minus
 * signs) and the CO2 sites (C at R a, O at R a +- r_CO b; a, b the unit vectors of (theta1,
 * phi1), (theta2, phi2)), with the well depth scaled to -320.0 cm^-1 (above the reference's
 * sampling bound pesmin = -342.934 cm^-1, example_configs/CF-F-CH4-CO2.conf:27).
 */
#include <math.h>
#include <string.h>

#include "oracle.h"
#include "ch4co2_dipole_table.h"

/* ------------------------------------------------------------------ rotations */
typedef double m3[3][3];

static void Sz(m3 m, double s, double c)
{
    m[0][0] = c; m[0][1] = s; m[0][2] = 0;
    m[1][0] = -s; m[1][1] = c; m[1][2] = 0;
    m[2][0] = 0; m[2][1] = 0; m[2][2] = 1;
}
static void Sz_dot(m3 m, double s, double c)
{
    m[0][0] = -s; m[0][1] = c; m[0][2] = 0;
    m[1][0] = -c; m[1][1] = -s; m[1][2] = 0;
    m[2][0] = 0; m[2][1] = 0; m[2][2] = 0;
}
static void Sx(m3 m, double s, double c)
{
    m[0][0] = 1; m[0][1] = 0; m[0][2] = 0;
    m[1][0] = 0; m[1][1] = c; m[1][2] = s;
    m[2][0] = 0; m[2][1] = -s; m[2][2] = c;
}
static void Sx_dot(m3 m, double s, double c)
{
    m[0][0] = 0; m[0][1] = 0; m[0][2] = 0;
    m[1][0] = 0; m[1][1] = -s; m[1][2] = c;
    m[2][0] = 0; m[2][1] = -c; m[2][2] = -s;
}
/* Sy_filler_non_standard (src/angles_handler.cpp:88-95) and its derivative */
static void Sy_ns(m3 m, double s, double c)
{
    m[0][0] = c; m[0][1] = 0; m[0][2] = s;
    m[1][0] = 0; m[1][1] = 1; m[1][2] = 0;
    m[2][0] = -s; m[2][1] = 0; m[2][2] = c;
}
static void Sy_ns_dot(m3 m, double s, double c)
{
    m[0][0] = -s; m[0][1] = 0; m[0][2] = c;
    m[1][0] = 0; m[1][1] = 0; m[1][2] = 0;
    m[2][0] = -c; m[2][1] = 0; m[2][2] = -s;
}
static void mm(m3 a, m3 b, m3 o)   /* o = a b */
{
    int i, j;
    m3 t;
    for (i = 0; i < 3; ++i)
        for (j = 0; j < 3; ++j) t[i][j] = a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j];
    memcpy(o, t, sizeof(m3));
}
static void mmt(m3 a, m3 b, m3 o)  /* o = a b^T */
{
    int i, j;
    m3 t;
    for (i = 0; i < 3; ++i)
        for (j = 0; j < 3; ++j) t[i][j] = a[i][0] * b[j][0] + a[i][1] * b[j][1] + a[i][2] * b[j][2];
    memcpy(o, t, sizeof(m3));
}

/* A B^T C^T z evaluated left to right, as the reference's Eigen expression (the z column) */
static void chain_z(m3 A, m3 B, m3 C, double *dd)
{
    m3 t;
    int i;
    mmt(A, B, t);
    mmt(t, C, t);
    for (i = 0; i < 3; ++i) dd[i] = t[i][2];
}

typedef struct {
    m3 SPhi, STheta, Sphi1, Stheta1, Spsi1, Sphi2, Stheta2;
    m3 SPhi_d, STheta_d, Sphi1_d, Stheta1_d, Spsi1_d, Sphi2_d, Stheta2_d;
} frames;

static void fill_frames(const double *q, frames *f)
{
    Sz(f->SPhi, sin(q[0]), cos(q[0]));
    Sy_ns(f->STheta, sin(q[1]), cos(q[1]));
    Sz(f->Sphi1, sin(q[3]), cos(q[3]));
    Sx(f->Stheta1, sin(q[4]), cos(q[4]));
    Sz(f->Spsi1, sin(q[5]), cos(q[5]));
    Sz(f->Sphi2, sin(q[6]), cos(q[6]));
    Sy_ns(f->Stheta2, sin(q[7]), cos(q[7]));
    Sz_dot(f->SPhi_d, sin(q[0]), cos(q[0]));
    Sy_ns_dot(f->STheta_d, sin(q[1]), cos(q[1]));
    Sz_dot(f->Sphi1_d, sin(q[3]), cos(q[3]));
    Sx_dot(f->Stheta1_d, sin(q[4]), cos(q[4]));
    Sz_dot(f->Spsi1_d, sin(q[5]), cos(q[5]));
    Sz_dot(f->Sphi2_d, sin(q[6]), cos(q[6]));
    Sy_ns_dot(f->Stheta2_d, sin(q[7]), cos(q[7]));
}

/* S = Spsi Stheta Sphi */
static void s1_of(m3 Spsi, m3 Stheta, m3 Sphi, m3 S)
{
    mm(Spsi, Stheta, S);
    mm(S, Sphi, S);
}

void orc_ch4co2_lab_to_kal(const double *qlab, double *qkal)
{
    frames f;
    m3 S1;
    double dd[3], c, s;
    fill_frames(qlab, &f);
    s1_of(f.Spsi1, f.Stheta1, f.Sphi1, S1);
    qkal[0] = qlab[2];
    chain_z(S1, f.SPhi, f.STheta, dd);
    c = dd[2];
    s = sqrt(1 - c * c);
    qkal[2] = acos(c);
    qkal[1] = atan2(dd[1] / s, dd[0] / s);
    chain_z(S1, f.Sphi2, f.Stheta2, dd);
    c = dd[2];
    s = sqrt(1.0 - dd[2] * dd[2]);
    qkal[4] = acos(c);
    qkal[3] = atan2(dd[1] / s, dd[0] / s);
}

void orc_ch4co2_jacobian(const double *qlab, const double *qkal, double jac[8][5])
{
    frames f;
    m3 S;
    double dd[3], dd2[3];
    double sp1 = sin(qkal[1]), cp1 = cos(qkal[1]), st1 = sin(qkal[2]);
    double sp2 = sin(qkal[3]), cp2 = cos(qkal[3]), st2 = sin(qkal[4]);
    int k;
    memset(jac, 0, sizeof(double) * 40);
    fill_frames(qlab, &f);
    jac[2][0] = 1.0;
    /* row k: the lab coordinate; (dd, dd2) with the differentiated rotation in its place */
    for (k = 0; k < 8; ++k) {
        m3 *A = &f.Spsi1, *B = &f.Stheta1, *C = &f.Sphi1, *D = &f.SPhi, *E = &f.STheta, *F = &f.Sphi2,
           *G = &f.Stheta2;
        if (k == 2) continue;
        switch (k) {
        case 0: D = &f.SPhi_d; break;
        case 1: E = &f.STheta_d; break;
        case 3: C = &f.Sphi1_d; break;
        case 4: B = &f.Stheta1_d; break;
        case 5: A = &f.Spsi1_d; break;
        case 6: F = &f.Sphi2_d; break;
        case 7: G = &f.Stheta2_d; break;
        }
        s1_of(*A, *B, *C, S);
        chain_z(S, *D, *E, dd);
        chain_z(S, *F, *G, dd2);
        jac[k][1] = (-dd[0] * sp1 + dd[1] * cp1) / st1;
        jac[k][2] = -dd[2] / st1;
        jac[k][3] = (-dd2[0] * sp2 + dd2[1] * cp2) / st2;
        jac[k][4] = -dd2[2] / st2;
    }
    /* src/angles_handler.cpp:548-554: theta2k does not depend on Phi, Theta; theta1k not on
     * phi2t, theta2t */
    jac[0][4] = 0.0;
    jac[1][4] = 0.0;
    jac[6][2] = 0.0;
    jac[7][2] = 0.0;
}

/* ------------------------------------------------------------------ stand-in PES */
#define ANG 1.8897261254578281            /* bohr per angstrom */
#define KELVIN_HARTREE 3.166811563e-6     /* Hartree per kelvin */
#define R_CH (1.087 * ANG)
#define R_CO (1.160 * ANG)
/* site parameters (sigma in angstrom, epsilon in kelvin): CH4 C, CH4 H, CO2 C, CO2 O */
static const double SIG[4] = {3.40, 2.65, 2.80, 3.05};
static const double EPS[4] = {51.0, 8.6, 27.0, 79.0};

static void ch4_sites(double h[5][3])
{
    static const double sg[4][3] = {{1, 1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
    int i, j;
    for (j = 0; j < 3; ++j) h[0][j] = 0.0;
    for (i = 0; i < 4; ++i)
        for (j = 0; j < 3; ++j) h[i + 1][j] = R_CH * sg[i][j] / sqrt(3.0);
}

/* V (unscaled, Hartree) and the gradients with respect to the three CO2 site positions */
static double lj_sites(const double co2[3][3], double g[3][3])
{
    double h[5][3], V = 0.0;
    int i, j, k;
    ch4_sites(h);
    memset(g, 0, sizeof(double) * 9);
    for (i = 0; i < 5; ++i) {
        int ti = i == 0 ? 0 : 1;
        for (j = 0; j < 3; ++j) {
            int tj = j == 0 ? 2 : 3;
            double sig = 0.5 * (SIG[ti] + SIG[tj]) * ANG, eps = sqrt(EPS[ti] * EPS[tj]) * KELVIN_HARTREE;
            double d[3], r2 = 0.0, sr2, sr6, dVdr_over_r;
            for (k = 0; k < 3; ++k) {
                d[k] = co2[j][k] - h[i][k];
                r2 += d[k] * d[k];
            }
            sr2 = sig * sig / r2;
            sr6 = sr2 * sr2 * sr2;
            V += 4.0 * eps * (sr6 * sr6 - sr6);
            /* dV/dr / r = 4 eps (-12 sr12 + 6 sr6) / r^2 */
            dVdr_over_r = 4.0 * eps * (-12.0 * sr6 * sr6 + 6.0 * sr6) / r2;
            for (k = 0; k < 3; ++k) g[j][k] += dVdr_over_r * d[k];
        }
    }
    return V;
}

static void unit(double phi, double theta, double u[3], double du_dphi[3], double du_dtheta[3])
{
    double sp = sin(phi), cp = cos(phi), st = sin(theta), ct = cos(theta);
    u[0] = st * cp;
    u[1] = st * sp;
    u[2] = ct;
    if (du_dphi) {
        du_dphi[0] = -st * sp;
        du_dphi[1] = st * cp;
        du_dphi[2] = 0.0;
    }
    if (du_dtheta) {
        du_dtheta[0] = ct * cp;
        du_dtheta[1] = ct * sp;
        du_dtheta[2] = -st;
    }
}

/* the stand-in's well depth is scaled to this value at its minimum (cm^-1) */
#define CH4CO2_WELL_CM (-320.0)
double orc_ch4co2_well_scale(void);

static double raw_V(double R, double phi1, double theta1, double phi2, double theta2, double g[3][3])
{
    double a[3], b[3], co2[3][3];
    int k;
    unit(phi1, theta1, a, 0, 0);
    unit(phi2, theta2, b, 0, 0);
    for (k = 0; k < 3; ++k) {
        co2[0][k] = R * a[k];
        co2[1][k] = R * a[k] + R_CO * b[k];
        co2[2][k] = R * a[k] - R_CO * b[k];
    }
    return lj_sites(co2, g);
}

double orc_ch4co2_pes_kal(double R, double phi1, double theta1, double phi2, double theta2)
{
    double g[3][3];
    return orc_ch4co2_well_scale() * raw_V(R, phi1, theta1, phi2, theta2, g);
}

void orc_ch4co2_dpes_kal(double R, double phi1, double theta1, double phi2, double theta2, double d[5])
{
    double g[3][3], a[3], b[3], da_dp[3], da_dt[3], db_dp[3], db_dt[3], gs[3], gd[3];
    double sc = orc_ch4co2_well_scale();
    int k;
    raw_V(R, phi1, theta1, phi2, theta2, g);
    unit(phi1, theta1, a, da_dp, da_dt);
    unit(phi2, theta2, b, db_dp, db_dt);
    for (k = 0; k < 3; ++k) {
        gs[k] = g[0][k] + g[1][k] + g[2][k];   /* dV / d(R a) */
        gd[k] = g[1][k] - g[2][k];             /* dV / d(r_CO b) */
    }
    d[0] = sc * (gs[0] * a[0] + gs[1] * a[1] + gs[2] * a[2]);
    d[1] = sc * R * (gs[0] * da_dp[0] + gs[1] * da_dp[1] + gs[2] * da_dp[2]);
    d[2] = sc * R * (gs[0] * da_dt[0] + gs[1] * da_dt[1] + gs[2] * da_dt[2]);
    d[3] = sc * R_CO * (gd[0] * db_dp[0] + gd[1] * db_dp[1] + gd[2] * db_dp[2]);
    d[4] = sc * R_CO * (gd[0] * db_dt[0] + gd[1] * db_dt[1] + gd[2] * db_dt[2]);
}

/* Scale of the raw site sum that puts the minimum over the orientations at CH4CO2_WELL_CM:
 * the minimum of the unscaled sum is found once by a coarse grid over (R, theta1, phi1,
 * theta2, phi2) refined by coordinate descent (the site sum is smooth; the search is
 * deterministic). */
double orc_ch4co2_well_scale(void)
{
    static double scale = 0.0;
    if (scale == 0.0) {
        double g[3][3], best = 0.0, x[5] = {7.0, 0.0, 1.0, 0.0, 1.0}, step[5] = {0.5, 0.5, 0.5, 0.5, 0.5};
        int i1, i2, i3, i4, it, k;
        for (i1 = 0; i1 < 12; ++i1)
            for (i2 = 0; i2 < 12; ++i2)
                for (i3 = 0; i3 < 12; ++i3)
                    for (i4 = 0; i4 < 12; ++i4) {
                        double p1 = 2 * M_PI * i1 / 12, t1 = M_PI * (i2 + 0.5) / 12;
                        double p2 = 2 * M_PI * i3 / 12, t2 = M_PI * (i4 + 0.5) / 12, R;
                        for (R = 5.0; R <= 10.0; R += 0.25) {
                            double v = raw_V(R, p1, t1, p2, t2, g);
                            if (v < best) {
                                best = v;
                                x[0] = R; x[1] = p1; x[2] = t1; x[3] = p2; x[4] = t2;
                            }
                        }
                    }
        for (it = 0; it < 200; ++it)
            for (k = 0; k < 5; ++k) {
                int dir;
                for (dir = -1; dir <= 1; dir += 2) {
                    double y[5];
                    memcpy(y, x, sizeof y);
                    y[k] += dir * step[k];
                    {
                        double v = raw_V(y[0], y[1], y[2], y[3], y[4], g);
                        if (v < best) {
                            best = v;
                            memcpy(x, y, sizeof y);
                        }
                    }
                }
                step[k] *= 0.9;
            }
        scale = (CH4CO2_WELL_CM / 2.1947463136320e5) / best;
    }
    return scale;
}

/* ------------------------------------------------------------------ dipole */
/* P_l^m(x) with the Condon-Shortley phase for l <= lmax, m <= l, at index l(l+1)/2 + m */
static void legendre_array(int lmax, double x, double *P)
{
    int l, m;
    double s = sqrt((1.0 - x) * (1.0 + x)), pmm = 1.0;
    for (m = 0; m <= lmax; ++m) {
        double p0, p1;
        if (m > 0) pmm *= -(2.0 * m - 1.0) * s;
        p0 = pmm;
        P[m * (m + 1) / 2 + m] = p0;
        if (m == lmax) break;
        p1 = x * (2.0 * m + 1.0) * p0;
        P[(m + 1) * (m + 2) / 2 + m] = p1;
        for (l = m + 2; l <= lmax; ++l) {
            double p2 = (x * (2.0 * l - 1.0) * p1 - (l + m - 1.0) * p0) / (l - m);
            P[l * (l + 1) / 2 + m] = p2;
            p0 = p1;
            p1 = p2;
        }
    }
}

static double radial_fn(int b, double R)
{
    int k = 0;
    if (R >= HH_CH4CO2_BOUNDS[HH_CH4CO2_NINT - 1]) {
        double r2 = R * R, r4 = r2 * r2;
        return (HH_CH4CO2_TAIL[b][0] + (HH_CH4CO2_TAIL[b][1] + HH_CH4CO2_TAIL[b][2] / R) / R) / r4;
    }
    while (k < HH_CH4CO2_NINT - 1 && !(R < HH_CH4CO2_BOUNDS[k])) ++k;
    {
        const double *c = HH_CH4CO2_RADIAL[k][b];
        double t = R - HH_CH4CO2_CENTRE[k];
        return ((c[3] * t + c[2]) * t + c[1]) * t + c[0];
    }
}

void orc_ch4co2_dipole_kal(const double *qkal, double d[3])
{
    double P1[16 * 17 / 2 + 17], P2[13 * 14 / 2 + 13], X[HH_CH4CO2_NX], Y[HH_CH4CO2_NY];
    double R = qkal[0], phi1 = qkal[1], theta1 = qkal[2], phi2 = qkal[3], theta2 = qkal[4];
    int l, m, b, c;
    legendre_array(13, cos(theta1), P1);
    legendre_array(12, cos(theta2), P2);
    for (l = 0; l <= 13; ++l) {
        X[l * l] = P1[l * (l + 1) / 2];
        for (m = 1; m <= l; ++m) {
            X[l * l + 2 * m - 1] = P1[l * (l + 1) / 2 + m] * cos(m * phi1);
            X[l * l + 2 * m] = P1[l * (l + 1) / 2 + m] * sin(m * phi1);
        }
    }
    for (l = 0; l <= 12; l += 2) {
        int o = l * (l - 1) / 2;
        Y[o] = P2[l * (l + 1) / 2];
        for (m = 1; m <= l; ++m) {
            Y[o + 2 * m - 1] = P2[l * (l + 1) / 2 + m] * cos(m * phi2);
            Y[o + 2 * m] = P2[l * (l + 1) / 2 + m] * sin(m * phi2);
        }
    }
    d[0] = d[1] = d[2] = 0.0;
    for (b = 0; b < HH_CH4CO2_NB; ++b) {
        double r = radial_fn(b, R);
        for (c = 0; c < 3; ++c) {
            double s = 0.0;
            int t;
            for (t = HH_CH4CO2_START[b][c]; t < HH_CH4CO2_START[b][c + 1]; ++t) {
                unsigned xy = HH_CH4CO2_XY[t];
                s += HH_CH4CO2_COEF[t] * X[xy & 0xff] * Y[xy >> 8];
            }
            d[c] += s * r;
        }
    }
}

/* ------------------------------------------------------------------ lab-frame plugin wrappers */
double orc_ch4co2_pes(const double *q)
{
    double k[5];
    orc_ch4co2_lab_to_kal(q, k);
    return orc_ch4co2_pes_kal(k[0], k[1], k[2], k[3], k[4]);
}

void orc_ch4co2_dpes(const double *q, double *dvdq)
{
    double k[5], dk[5], jac[8][5];
    int i, j;
    orc_ch4co2_lab_to_kal(q, k);
    orc_ch4co2_jacobian(q, k, jac);
    orc_ch4co2_dpes_kal(k[0], k[1], k[2], k[3], k[4], dk);
    for (i = 0; i < 8; ++i) {
        double acc = 0.0;
        for (j = 0; j < 5; ++j) acc += jac[i][j] * dk[j];
        dvdq[i] = acc;
    }
}

void orc_ch4co2_dipole(const double *q, double mu[3])
{
    double k[5], dk[3], t1[3], t2[3];
    m3 Sphi, Stheta, Spsi;
    int i;
    orc_ch4co2_lab_to_kal(q, k);
    orc_ch4co2_dipole_kal(k, dk);
    Sz(Sphi, sin(q[3]), cos(q[3]));
    Sx(Stheta, sin(q[4]), cos(q[4]));
    Sz(Spsi, sin(q[5]), cos(q[5]));
    /* Sphi1^T Stheta1^T Spsi1^T d */
    for (i = 0; i < 3; ++i) t1[i] = Spsi[0][i] * dk[0] + Spsi[1][i] * dk[1] + Spsi[2][i] * dk[2];
    for (i = 0; i < 3; ++i) t2[i] = Stheta[0][i] * t1[0] + Stheta[1][i] * t1[1] + Stheta[2][i] * t1[2];
    for (i = 0; i < 3; ++i) mu[i] = Sphi[0][i] * t2[0] + Sphi[1][i] * t2[1] + Sphi[2][i] * t2[2];
}

/* ------------------------------------------------------------------ equations of motion */
/* rhsMonomer ROTOR (src/hawaii.c:505-531) for monomer state m[6] = (phi, pphi, theta, ptheta,
 * psi, ppsi) with the potential derivatives dV[3] = dV/d(phi, theta, psi) */
static void rotor_rhs(const double *II, const double *m, const double *dV, double *deriv)
{
    double pPhi = m[1], theta = m[2], pTheta = m[3], psi = m[4], pPsi = m[5];
    double sin_theta = sin(theta), cos_theta = cos(theta), sin_psi = sin(psi), cos_psi = cos(psi);
    double t1 = (pPhi - pPsi * cos_theta) * sin_psi + pTheta * sin_theta * cos_psi;
    double t2 = (pPhi - pPsi * cos_theta) * cos_psi - pTheta * sin_theta * sin_psi;
    double t3 = pPhi * cos_theta - pPsi;
    double t4 = pPsi * cos_theta - pPhi;
    deriv[0] = t1 * sin_psi / (II[0] * sin_theta * sin_theta) + t2 * cos_psi / (II[1] * sin_theta * sin_theta);
    deriv[1] = -dV[0];
    deriv[2] = t1 * cos_psi / (II[0] * sin_theta) - t2 * sin_psi / (II[1] * sin_theta);
    deriv[3] = -t3 * (t4 * (II[0] - II[1]) * cos_psi * cos_psi + pTheta * (II[0] - II[1]) * sin_theta * sin_psi * cos_psi +
                      t4 * II[1]) / (II[0] * II[1] * sin_theta * sin_theta * sin_theta) - dV[1];
    deriv[4] = -t1 * cos_theta * sin_psi / (II[0] * sin_theta * sin_theta) -
               t2 * cos_theta * cos_psi / (II[1] * sin_theta * sin_theta) + pPsi / II[2];
    deriv[5] = -t1 * t2 / (II[0] * sin_theta * sin_theta) + t1 * t2 / (II[1] * sin_theta * sin_theta) - dV[2];
}

void orc_ch4co2_rhs(const orc_system *s, const double *y, double *ydot)
{
    double q[8], dV[8];
    double pPhi = y[1], Theta = y[2], pTheta = y[3], R = y[4], pR = y[5], mu = s->mu;
    double R2 = R * R, R3 = R2 * R, sinTheta = sin(Theta), cosTheta = cos(Theta);
    double sinTheta2 = sinTheta * sinTheta, sinTheta3 = sinTheta2 * sinTheta;
    int i;
    for (i = 0; i < 8; ++i) q[i] = y[2 * i];
    orc_ch4co2_dpes(q, dV);
    /* src/hawaii.c:586-591 */
    ydot[4] = pR / mu;
    ydot[5] = pTheta * pTheta / (mu * R3) + pPhi * pPhi / (mu * R3 * sinTheta2) - dV[2];
    ydot[0] = pPhi / (mu * R2 * sinTheta2);
    ydot[1] = -dV[0];
    ydot[2] = pTheta / (mu * R2);
    ydot[3] = pPhi * pPhi * cosTheta / (mu * R2 * sinTheta3) - dV[1];
    rotor_rhs(s->IR, y + 6, dV + 3, ydot + 6);
    {                                                      /* src/hawaii.c:494-497, monomer 2 */
        double I = s->II[0];
        double pphi = y[13], theta = y[14], ptheta = y[15];
        double st = sin(theta), ct = cos(theta);
        ydot[12] = pphi / I / st / st;
        ydot[13] = -dV[6];
        ydot[14] = ptheta / I;
        ydot[15] = pphi * pphi * ct / I / st / st / st - dV[7];
    }
}

/* kinetic_energy (src/hawaii.c:1041-1137): KIN1 + KIN2 (rotor) + KIN3 (linear) */
double orc_ch4co2_kinetic(const orc_system *s, const double *y)
{
    double pPhi = y[1], Theta = y[2], pTheta = y[3], R = y[4], pR = y[5], mu = s->mu;
    double sinTheta = sin(Theta);
    double KIN1 = pR * pR / (2.0 * mu) + pTheta * pTheta / (2.0 * mu * R * R) +
                  pPhi * pPhi / (2.0 * mu * R * R * sinTheta * sinTheta);
    double pphi1t = y[7], theta = y[8], ptheta1t = y[9], psi = y[10], ppsi1t = y[11];
    double sin_theta = sin(theta), cos_theta = cos(theta), sin_psi = sin(psi), cos_psi = cos(psi);
    double KIN2 = ((pphi1t - ppsi1t * cos_theta) * sin_psi + ptheta1t * sin_theta * cos_psi) *
                  ((pphi1t - ppsi1t * cos_theta) * sin_psi + ptheta1t * sin_theta * cos_psi) /
                  (2.0 * s->IR[0] * sin_theta * sin_theta) +
                  ((pphi1t - ppsi1t * cos_theta) * cos_psi - ptheta1t * sin_theta * sin_psi) *
                  ((pphi1t - ppsi1t * cos_theta) * cos_psi - ptheta1t * sin_theta * sin_psi) /
                  (2.0 * s->IR[1] * sin_theta * sin_theta) +
                  ppsi1t * ppsi1t / (2.0 * s->IR[2]);
    double sin_theta2t = sin(y[14]);
    double KIN3 = y[15] * y[15] / (2.0 * s->II[0]) + y[13] * y[13] / (2.0 * s->II[1] * sin_theta2t * sin_theta2t);
    return KIN1 + KIN2 + KIN3;
}
