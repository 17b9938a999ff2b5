// Device-side molecular systems: Hamilton's equations, PES gradients and induced dipoles
// for He-Ar (atom-atom) and CO2-Ar (linear rotor + atom).
//
// Equations of motion follow src/hawaii.c:553-615 (rhs) and src/hawaii.c:474-538
// (rhsMonomer, linear molecule); kinetic energy src/hawaii.c:1041-1137.  They stay in FP64.
// The surface evaluation (PES gradient, dipole) is templated on its arithmetic type P
// (double, or float in the mixed-precision engine).
//
// CO2-Ar surfaces (PES-IDS/ai_pes_co2ar_lib.cpp:5-30, PES-IDS/ai_ids_co2ar_lib.cpp:12-38):
// the reference builds the molecular-frame angles through rotation matrices, acos and
// atan2 (src/angles_handler.cpp:106-193) and maps the gradient back with a 5x5 Jacobian.
// Here the same quantities come from two unit vectors:
//     Rhat = (sin T cos P, sin T sin P, cos T),  u = (sin t cos p, sin t sin p, cos t),
//     cos(theta_m) = c = Rhat . u,
//     dV/dq_i = dV/dR delta_{i,R} + (sum_l c_l(R) P_l'(c)) dc/dq_i,
//     mu_lab = dipz Rhat - (sum_l dx_l(R) P_l'(c)) (u - c Rhat)
// (P_l^1 = -(1-c^2)^{1/2} P_l' with the Condon-Shortley phase, and the molecular x axis is
// (u - c Rhat)/sin(theta_m)), which is the reference's chain without the 1/sin(theta_m)
// singularity and without inverse trigonometric functions.
#pragma once
#include <hip/hip_runtime.h>

#include "ch4co2_standin.h"

namespace hh {

constexpr double kALU = 5.29177210903e-11;           // src/constants.h:8
constexpr double kHTOCM = 2.1947463136320e5;         // src/constants.h:11
constexpr double kTWO_PI = 6.283185307179586476925;

enum SystemId : int { SYS_HEAR = 0, SYS_CO2AR = 1, SYS_CH4CO2 = 3 };

// Surface tables resident in device memory (filled by the host from co2ar_tables.h).
template <typename P>
struct Co2ArTables {
    P pes[32][6][8];     // centred cubic (4) + R^-4..R^-7 (4), Hartree
    P pes_exp[6][2];     // R < 4.5: A exp(alpha R), Hartree
    P dipx[32][5][8];    // P_l^1 slots l = 2..10
    P dipz[32][6][8];    // P_l slots l = 0..10
    double center[32];
};

struct SysConst {
    double mu;           // reduced mass
    double I;            // linear-rotor moment of inertia (II[0] == II[1]); CH4-CO2: of CO2
    const void* tables;  // Co2ArTables<P>* for CO2-Ar
    double IR[3];        // CH4-CO2: principal moments of the rotor (CH4)
};

// ------------------------------------------------------------------ He-Ar (PES-IDS/HeAr.h)
// Tang-Toennies form V = A exp(-a1 R - a2 R^2) - sum_{n=3}^{8} C_2n f_2n(R)/R^2n with
// f_2n = 1 - exp(-bR) sum_{k<=2n} (bR)^k/k!; derivative analytic.  Kept in FP64 in every
// engine mode: f_2n cancels to 1e-5 below R ~ 5 bohr, beyond what FP32 can carry.
__device__ __forceinline__ double hear_dV(double R)
{
    const double A = 0.2303058634e2, a1 = 0.163051044e1, a2 = 0.4680016e-1, b = 0.16274792500000e1;
    const double C[6] = {0.940835513e1, 0.165523018e3, 0.379715796e4, 0.1165179999e6, 0.466258000e7,
                         0.236861000e9};
    const double x = b * R, e = exp(-x), inv = 1.0 / R, inv2 = inv * inv;
    double term = 1.0, sum = 1.0;
    double dV = A * exp(-a1 * R - a2 * R * R) * (-a1 - 2.0 * a2 * R);
    double ip = inv2 * inv2;  // R^-4
    int k = 0;
#pragma unroll
    for (int n = 3; n <= 8; ++n) {
        ip *= inv2;  // R^-2n
#pragma unroll
        for (int j = 0; j < 2; ++j) { ++k; term *= x * (1.0 / k); sum += term; }
        if (n == 3) {
#pragma unroll
            for (int j = 0; j < 4; ++j) { ++k; term *= x * (1.0 / k); sum += term; }
        }
        dV -= C[n - 3] * (b * e * term * ip - 2.0 * n * (1.0 - e * sum) * ip * inv);
    }
    return dV;
}

__device__ __forceinline__ double hear_V(double R)
{
    const double A = 0.2303058634e2, a1 = 0.163051044e1, a2 = 0.4680016e-1, b = 0.16274792500000e1;
    const double C[6] = {0.940835513e1, 0.165523018e3, 0.379715796e4, 0.1165179999e6, 0.466258000e7,
                         0.236861000e9};
    const double x = b * R, e = exp(-x), inv = 1.0 / R, inv2 = inv * inv;
    double term = 1.0, sum = 1.0, V = A * exp(-a1 * R - a2 * R * R), ip = inv2 * inv2;
    int k = 0;
#pragma unroll
    for (int n = 3; n <= 8; ++n) {
        ip *= inv2;
#pragma unroll
        for (int j = 0; j < 2; ++j) { ++k; term *= x * (1.0 / k); sum += term; }
        if (n == 3) {
#pragma unroll
            for (int j = 0; j < 4; ++j) { ++k; term *= x * (1.0 / k); sum += term; }
        }
        V -= C[n - 3] * (1.0 - e * sum) * ip;
    }
    return V;
}

template <typename P>
__device__ __forceinline__ P hear_dip(P R)
{
    // PES-IDS/HeAr.h:44-47
    const P r2 = R * R, inv = P(1) / R, inv2 = inv * inv, inv7 = inv2 * inv2 * inv2 * inv;
    return exp(P(0.47293181) - P(0.40100613) * R - P(0.10292726) * r2) - P(148.55391) * inv7;
}

// 1/x within an ulp: the hardware reciprocal estimate refined by two Newton steps (each at
// least doubles the correct bits).  The engines with FP32 surfaces (P = float) use it in place
// of the correctly rounded FP64 division, which costs twice the instructions and a longer
// dependent chain; the FP64 engines divide.
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

template <typename P>
__device__ __forceinline__ double recip(double x)
{
    if constexpr (sizeof(P) == 4)
        return rcp_nr(x);
    else
        return 1.0 / x;
}

// ------------------------------------------------------------------ CO2-Ar
__device__ __forceinline__ int co2ar_branch(double R)
{
    // breakpoints 4.5, 4.75, ..., 10 (step 1/4), then 11, 12, 14, 16, 18, 20, 25, 30
    if (R < 4.5) return 0;
    if (R < 10.0) return 1 + (int)((R - 4.5) * 4.0);   // exact: R - 4.5 and *4 are exact here
    int b = 23;
    b += (R >= 11.0) + (R >= 12.0) + (R >= 14.0) + (R >= 16.0) + (R >= 18.0) + (R >= 20.0) + (R >= 25.0) +
         (R >= 30.0);
    return b;
}

// sin and cos for the bounded angles of a trajectory (|x| < 1.6e6 rad).  Cody-Waite
// reduction by pi/2 in FP64 (pi/2 split into three doubles whose leading two carry 33
// significant bits, so k*C1 and k*C2 are exact for |k| < 2^20), then the Taylor series of
// sin and cos on |r| <= pi/4, truncated where the next term is below 2^-60 (FP64) or
// 2^-30 (FP32) relative, and quadrant selection.  No table, no large-argument path.
struct Reduced {
    double r;
    int q;
};

__device__ __forceinline__ Reduced reduce_pio2(double x)
{
    const double k = rint(x * 0.6366197723675814);
    double r = fma(-k, 1.5707963267341256, x);
    r = fma(-k, 6.077100506303966e-11, r);
    r = fma(-k, 2.0222662487959506e-21, r);
    return {r, (int)k & 3};
}

template <typename P>
__device__ __forceinline__ void sincos_p(double x, P* s, P* c)
{
    const Reduced red = reduce_pio2(x);
    P sr, cr;
    if constexpr (sizeof(P) == 8) {
        const double r = red.r, r2 = r * r;
        sr = r + r * r2 * (-1.0 / 6 + r2 * (1.0 / 120 + r2 * (-1.0 / 5040 + r2 * (1.0 / 362880 + r2 * (-1.0 / 39916800 +
             r2 * (1.0 / 6227020800.0 + r2 * (-1.0 / 1307674368000.0 + r2 * (1.0 / 355687428096000.0))))))));
        cr = 1.0 + r2 * (-0.5 + r2 * (1.0 / 24 + r2 * (-1.0 / 720 + r2 * (1.0 / 40320 + r2 * (-1.0 / 3628800 +
             r2 * (1.0 / 479001600.0 + r2 * (-1.0 / 87178291200.0 + r2 * (1.0 / 20922789888000.0))))))));
    } else {
        const float r = (float)red.r, r2 = r * r;
        sr = r + r * r2 * (-1.0f / 6 + r2 * (1.0f / 120 + r2 * (-1.0f / 5040 + r2 * (1.0f / 362880))));
        cr = 1.0f + r2 * (-0.5f + r2 * (1.0f / 24 + r2 * (-1.0f / 720 + r2 * (1.0f / 40320 + r2 * (-1.0f / 3628800)))));
    }
    const bool swap = red.q & 1;
    const P s0 = swap ? cr : sr, c0 = swap ? sr : cr;
    *s = (red.q == 2 || red.q == 3) ? -s0 : s0;
    *c = (red.q == 1 || red.q == 2) ? -c0 : c0;
}

// Legendre P_l(c), P_l'(c), l = 0..10 (three-term recurrences, P'_{l+1} = P'_{l-1} + (2l+1) P_l)
template <typename P>
__device__ __forceinline__ void legendre10(P c, P* Pl, P* dPl)
{
    Pl[0] = P(1);
    Pl[1] = c;
    dPl[0] = P(0);
    dPl[1] = P(1);
#pragma unroll
    for (int l = 1; l < 10; ++l) {
        Pl[l + 1] = (P(2 * l + 1) * c * Pl[l] - P(l) * Pl[l - 1]) * P(1.0 / (l + 1));
        dPl[l + 1] = dPl[l - 1] + P(2 * l + 1) * Pl[l];
    }
}

// radial slot value and derivative: centred cubic in x plus R^-4..R^-7 in s = 1/R
template <typename P>
__device__ __forceinline__ void slot_eval(const P* c, P x, P s, P& v, P& dv)
{
    const P s2 = s * s, s4 = s2 * s2;
    v = ((c[3] * x + c[2]) * x + c[1]) * x + c[0] + s4 * (((c[7] * s + c[6]) * s + c[5]) * s + c[4]);
    dv = (P(3) * c[3] * x + P(2) * c[2]) * x + c[1] -
         s4 * s * (((P(7) * c[7] * s + P(6) * c[6]) * s + P(5) * c[5]) * s + P(4) * c[4]);
}

template <typename P>
__device__ __forceinline__ P slot_value(const P* c, P x, P s)
{
    const P s2 = s * s, s4 = s2 * s2;
    return ((c[3] * x + c[2]) * x + c[1]) * x + c[0] + s4 * (((c[7] * s + c[6]) * s + c[5]) * s + c[4]);
}

struct Trig {
    double sT, cT, st, ct;   // FP64 sin/cos of Theta and theta1 (EOM)
};

// Rounding-noise scale of an FP32 gradient: 8 units of FP32 roundoff (2^-24) times the sum of
// the magnitudes of the Legendre terms that are added up (their cancellation is what makes the
// FP32 gradient noisy, 1e-6 relative and more where the terms nearly cancel).
constexpr double kF32NoiseUnits = 8.0 / 16777216.0;

// CO2-Ar lab-frame gradient dV/dq, q = [Phi, Theta, R, phi1, theta1] (Hartree / bohr, rad).
// noise: for P = float an absolute bound on the rounding noise of every component (0 for FP64).
template <typename P>
__device__ __forceinline__ void co2ar_dpes(const Co2ArTables<P>* __restrict__ tb, const double* y,
                                           const Trig& tr, double* dV, double& noise)
{
    const double R = y[4];
    const int b = co2ar_branch(R);
    // FP32: the radial coefficients of the point's branch are read from LDS first, so their
    // latency overlaps the angular part (one wait instead of one per slot)
    P cf[6][8];
    double cen = 0.0;
    if constexpr (sizeof(P) == 4) {
        const int bb = b > 0 ? b : 1;   // the exponential branch (b == 0) uses other constants
#pragma unroll
        for (int q = 0; q < 6; ++q)
#pragma unroll
            for (int j = 0; j < 8; ++j) cf[q][j] = tb->pes[bb][q][j];
        cen = tb->center[bb];
        __builtin_amdgcn_sched_barrier(0);
    }
    P sP, cP, sp, cp;
    sincos_p<P>(y[0], &sP, &cP);
    sincos_p<P>(y[6], &sp, &cp);
    const P sT = (P)tr.sT, cT = (P)tr.cT, st = (P)tr.st, ct = (P)tr.ct;
    const P Rx = sT * cP, Ry = sT * sP, Rz = cT;
    const P ux = st * cp, uy = st * sp, uz = ct;
    P c = Rx * ux + Ry * uy + Rz * uz;
    c = c > P(1) ? P(1) : (c < P(-1) ? P(-1) : c);
    const P dc_dPhi = -Ry * ux + Rx * uy;                       // dRhat/dPhi . u
    const P dc_dTheta = cT * cP * ux + cT * sP * uy - sT * uz;  // dRhat/dTheta . u
    const P dc_dphi = -Rx * uy + Ry * ux;                       // Rhat . du/dphi
    const P dc_dtheta = ct * cp * Rx + ct * sp * Ry - st * Rz;  // Rhat . du/dtheta

    P Pl[11], dPl[11];
    legendre10<P>(c, Pl, dPl);

    P dVdR = P(0), dVdc = P(0);
    P mag = P(0);   // sum of |terms| (FP32 only)
    if (b == 0) {
        const P r = (P)R;
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const P e = tb->pes_exp[s][0] * exp(tb->pes_exp[s][1] * r);
            const P tR = tb->pes_exp[s][1] * e * Pl[2 * s], tc = e * dPl[2 * s];
            dVdR += tR;
            dVdc += tc;
            if constexpr (sizeof(P) == 4) mag += fabs(tR) + fabs(tc);
        }
    } else {
        const P x = (P)(R - (sizeof(P) == 4 ? cen : tb->center[b]));
        const P sinv = (P)recip<P>(R);
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            P v, dv;
            slot_eval<P>(sizeof(P) == 4 ? cf[s] : tb->pes[b][s], x, sinv, v, dv);
            const P tR = dv * Pl[2 * s], tc = v * dPl[2 * s];
            dVdR += tR;
            dVdc += tc;
            if constexpr (sizeof(P) == 4) mag += fabs(tR) + fabs(tc);
        }
    }
    noise = sizeof(P) == 4 ? kF32NoiseUnits * (double)mag : 0.0;
    dV[0] = (double)(dVdc * dc_dPhi);
    dV[1] = (double)(dVdc * dc_dTheta);
    dV[2] = (double)dVdR;
    dV[3] = (double)(dVdc * dc_dphi);
    dV[4] = (double)(dVdc * dc_dtheta);
}

template <typename P>
__device__ __forceinline__ P co2ar_pes(const Co2ArTables<P>* __restrict__ tb, const double* y)
{
    double sTd, cTd, std_, ctd;
    sincos_p<double>(y[2], &sTd, &cTd);
    sincos_p<double>(y[8], &std_, &ctd);
    P sP, cP, sp, cp;
    sincos_p<P>(y[0], &sP, &cP);
    sincos_p<P>(y[6], &sp, &cp);
    const P sT = (P)sTd, cT = (P)cTd, st = (P)std_, ct = (P)ctd;
    P c = sT * cP * st * cp + sT * sP * st * sp + cT * ct;
    c = c > P(1) ? P(1) : (c < P(-1) ? P(-1) : c);
    P Pl[11], dPl[11];
    legendre10<P>(c, Pl, dPl);
    const double R = y[4];
    const int b = co2ar_branch(R);
    P V = P(0);
    if (b == 0) {
#pragma unroll
        for (int s = 0; s < 6; ++s) V += tb->pes_exp[s][0] * exp(tb->pes_exp[s][1] * (P)R) * Pl[2 * s];
    } else {
        const P x = (P)(R - tb->center[b]);
        const P sinv = (P)(1.0 / R);
#pragma unroll
        for (int s = 0; s < 6; ++s) V += slot_value<P>(tb->pes[b][s], x, sinv) * Pl[2 * s];
    }
    return V;
}

// CO2-Ar lab-frame induced dipole at state y
template <typename P>
__device__ __forceinline__ void co2ar_dipole(const Co2ArTables<P>* __restrict__ tb, const double* y, P* mu)
{
    const double R = y[4];
    const int b = co2ar_branch(R);
    // FP32: the branch's radial coefficients are read from LDS before the angular part (as in
    // co2ar_dpes)
    P cx[5][8], cz[6][8];
    double cen = 0.0;
    if constexpr (sizeof(P) == 4) {
#pragma unroll
        for (int q = 0; q < 5; ++q)
#pragma unroll
            for (int j = 0; j < 8; ++j) cx[q][j] = tb->dipx[b][q][j];
#pragma unroll
        for (int q = 0; q < 6; ++q)
#pragma unroll
            for (int j = 0; j < 8; ++j) cz[q][j] = tb->dipz[b][q][j];
        cen = tb->center[b];
        __builtin_amdgcn_sched_barrier(0);
    }
    P sT, cT, sP, cP, st, ct, sp, cp;
    sincos_p<P>(y[2], &sT, &cT);
    sincos_p<P>(y[0], &sP, &cP);
    sincos_p<P>(y[8], &st, &ct);
    sincos_p<P>(y[6], &sp, &cp);
    const P Rx = sT * cP, Ry = sT * sP, Rz = cT;
    const P ux = st * cp, uy = st * sp, uz = ct;
    P c = Rx * ux + Ry * uy + Rz * uz;
    c = c > P(1) ? P(1) : (c < P(-1) ? P(-1) : c);
    P Pl[11], dPl[11];
    legendre10<P>(c, Pl, dPl);
    const P x = (P)(R - (sizeof(P) == 4 ? cen : tb->center[b]));
    const P sinv = (P)recip<P>(R);
    P gx = P(0), dz = P(0);
#pragma unroll
    for (int s = 0; s < 5; ++s) gx += slot_value<P>(sizeof(P) == 4 ? cx[s] : tb->dipx[b][s], x, sinv) * dPl[2 * s + 2];
#pragma unroll
    for (int s = 0; s < 6; ++s) dz += slot_value<P>(sizeof(P) == 4 ? cz[s] : tb->dipz[b][s], x, sinv) * Pl[2 * s];
    // mu = dz Rhat - gx (u - c Rhat)
    mu[0] = dz * Rx - gx * (ux - c * Rx);
    mu[1] = dz * Ry - gx * (uy - c * Ry);
    mu[2] = dz * Rz - gx * (uz - c * Rz);
}

// ------------------------------------------------------------------ CH4-CO2 (ROTOR + LINEAR_MOLECULE)
// Lab coordinates q = [Phi, Theta, R, phi1, theta1, psi1, phi2, theta2] at state indices 0, 2, ...,
// 14; monomer 1 = CH4 (Euler angles phi1, theta1, psi1), monomer 2 = CO2 (phi2, theta2).  The frame
// of CH4 is S1 = Sz(psi1) Sx(theta1) Sz(phi1) (src/angles_handler.cpp:357-419); the intermolecular
// direction and the CO2 axis in the lab are Rhat = (-sin T cos P, -sin T sin P, cos T) and
// u = (-sin t2 cos p2, -sin t2 sin p2, cos t2) (the reference's SPhi^T STheta^T z with its
// non-standard Sy), a = S1 Rhat and b = S1 u in the CH4 frame.  The stand-in surface
// (ch4co2_standin.h) is a function of R, a and b; its lab gradient follows from the generators of
// the rotations: with ga = dV/da, gb = dV/db (CH4 frame) and Ga = S1^T ga, Gb = S1^T gb (lab),
//   dV/dPhi = Ga . dRhat/dPhi          dV/dTheta = Ga . dRhat/dTheta
//   dV/dphi2 = Gb . du/dphi2           dV/dtheta2 = Gb . du/dtheta2
//   dV/dphi1 = Ga . Jz Rhat + Gb . Jz u              (dS1/dphi1 = S1 Jz)
//   dV/dtheta1 = A^T ga . Jx A^T a + A^T gb . Jx A^T b   (A = Sz(psi1); dS1/dtheta1 = A Jx A^T S1)
//   dV/dpsi1 = ga . Jz a + gb . Jz b                 (dS1/dpsi1 = Jz S1)
// with Jz v = (v1, -v0, 0), Jx v = (0, v2, -v1): the Jacobian of the reference
// (src/angles_handler.cpp:421-555) without its 1/sin(theta_K) factors.
struct Ch4Co2Trig {
    double sT, cT, s1, c1, sp, cp, s2, c2;   // FP64 sin/cos of Theta, theta1, psi1, theta2
};

template <typename P>
struct Ch4Co2Frame {
    P Rh[3], u[3], a[3], b[3];
    P sP, cP, sf, cf, sP2, cP2;   // sin/cos of Phi, phi1, phi2
    P s1, c1, sp, cp, sT, cT, s2, c2;
};

template <typename P>
__device__ __forceinline__ void ch4co2_frame(const double* y, const Ch4Co2Trig& tr, Ch4Co2Frame<P>& f)
{
    sincos_p<P>(y[0], &f.sP, &f.cP);
    sincos_p<P>(y[6], &f.sf, &f.cf);
    sincos_p<P>(y[12], &f.sP2, &f.cP2);
    f.sT = (P)tr.sT, f.cT = (P)tr.cT, f.s1 = (P)tr.s1, f.c1 = (P)tr.c1;
    f.sp = (P)tr.sp, f.cp = (P)tr.cp, f.s2 = (P)tr.s2, f.c2 = (P)tr.c2;
    f.Rh[0] = -f.sT * f.cP;
    f.Rh[1] = -f.sT * f.sP;
    f.Rh[2] = f.cT;
    f.u[0] = -f.s2 * f.cP2;
    f.u[1] = -f.s2 * f.sP2;
    f.u[2] = f.c2;
    // a = Sz(psi) Sx(theta) Sz(phi) Rhat, b likewise
    const P* v[2] = {f.Rh, f.u};
    P* o[2] = {f.a, f.b};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const P w0 = f.cf * v[k][0] + f.sf * v[k][1], w1 = -f.sf * v[k][0] + f.cf * v[k][1], w2 = v[k][2];
        const P x1 = f.c1 * w1 + f.s1 * w2, x2 = -f.s1 * w1 + f.c1 * w2;
        o[k][0] = f.cp * w0 + f.sp * x1;
        o[k][1] = -f.sp * w0 + f.cp * x1;
        o[k][2] = x2;
    }
}

// Stand-in site sum at R, a, b: the energy (Hartree, scaled) and gs = dV/d(R a) summed over the
// CO2 sites, gd = dV/d(r_CO b); mag: a bound on the sum of the pair forces (FP32 noise bound)
template <typename P>
__device__ __forceinline__ P ch4co2_sites(P R, const P* a, const P* b, P* gs, P* gd, P& mag)
{
    using namespace ch4co2;
    P V = P(0);
    mag = P(0);
    gs[0] = gs[1] = gs[2] = gd[0] = gd[1] = gd[2] = P(0);
    const P hs = P(kRCH / 1.7320508075688772);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const P sgn = j == 0 ? P(0) : (j == 1 ? P(1) : P(-1));
        const P X[3] = {R * a[0] + sgn * P(kRCO) * b[0], R * a[1] + sgn * P(kRCO) * b[1], R * a[2] + sgn * P(kRCO) * b[2]};
        const int tj = j == 0 ? 2 : 3;
        P g[3] = {P(0), P(0), P(0)};
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const int ti = i == 0 ? 0 : 1;
            const double sig = 0.5 * (kSigma[ti] + kSigma[tj]) * kAng;
            const P s2 = P(sig * sig);
            const P e4 = P(4.0 * kWellScale * kKelvinHartree) * (P)__builtin_sqrt(kEpsilon[ti] * kEpsilon[tj]);
            const P h0 = i == 0 ? P(0) : hs * P(kH[i - 1][0]), h1 = i == 0 ? P(0) : hs * P(kH[i - 1][1]);
            const P h2 = i == 0 ? P(0) : hs * P(kH[i - 1][2]);
            const P d0 = X[0] - h0, d1 = X[1] - h1, d2 = X[2] - h2;
            const P r2 = d0 * d0 + d1 * d1 + d2 * d2;
            const P ir2 = P(1) / r2;
            const P sr2 = s2 * ir2, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
            V += e4 * (sr12 - sr6);
            const P w = e4 * (P(-12) * sr12 + P(6) * sr6) * ir2;   // (dV/dr) / r
            g[0] += w * d0;
            g[1] += w * d1;
            g[2] += w * d2;
            mag += fabs(w) * (fabs(d0) + fabs(d1) + fabs(d2));   // >= |dV/dr| of the pair
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            gs[k] += g[k];
            gd[k] += j == 0 ? P(0) : sgn * g[k];
        }
    }
    return V;
}

// lab gradient dV/dq[8] of the stand-in; noise: FP32 rounding bound of the components (0 in FP64)
template <typename P>
__device__ __forceinline__ void ch4co2_dpes(const double* y, const Ch4Co2Trig& tr, double* dV, double& noise)
{
    Ch4Co2Frame<P> f;
    ch4co2_frame<P>(y, tr, f);
    const P R = (P)y[4];
    P gs[3], gd[3], mag;
    (void)ch4co2_sites<P>(R, f.a, f.b, gs, gd, mag);
    P ga[3], gb[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ga[k] = R * gs[k];
        gb[k] = P(ch4co2::kRCO) * gd[k];
    }
    const P dR = gs[0] * f.a[0] + gs[1] * f.a[1] + gs[2] * f.a[2];
    // lab-frame gradients: S1^T g = Sz(phi)^T Sx(theta)^T Sz(psi)^T g
    P Ga[3], Gb[3], Aa[3], Ab[3], Ag[3], Bg[3];
    const P* gg[2] = {ga, gb};
    P* GG[2] = {Ga, Gb};
    P* AG[2] = {Ag, Bg};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const P* g = gg[k];
        const P x0 = f.cp * g[0] - f.sp * g[1], x1 = f.sp * g[0] + f.cp * g[1], x2 = g[2];   // A^T g
        AG[k][0] = x0;
        AG[k][1] = x1;
        AG[k][2] = x2;
        const P w1 = f.c1 * x1 - f.s1 * x2, w2 = f.s1 * x1 + f.c1 * x2;                    // Sx^T
        GG[k][0] = f.cf * x0 - f.sf * w1;
        GG[k][1] = f.sf * x0 + f.cf * w1;
        GG[k][2] = w2;
    }
    // A^T a, A^T b
    Aa[0] = f.cp * f.a[0] - f.sp * f.a[1];
    Aa[1] = f.sp * f.a[0] + f.cp * f.a[1];
    Aa[2] = f.a[2];
    Ab[0] = f.cp * f.b[0] - f.sp * f.b[1];
    Ab[1] = f.sp * f.b[0] + f.cp * f.b[1];
    Ab[2] = f.b[2];
    // dRhat/dPhi = (sT sP, -sT cP, 0), dRhat/dTheta = (-cT cP, -cT sP, -sT); u likewise
    const P dPhi = Ga[0] * (f.sT * f.sP) - Ga[1] * (f.sT * f.cP);
    const P dTheta = -Ga[0] * (f.cT * f.cP) - Ga[1] * (f.cT * f.sP) - Ga[2] * f.sT;
    const P dphi2 = Gb[0] * (f.s2 * f.sP2) - Gb[1] * (f.s2 * f.cP2);
    const P dtheta2 = -Gb[0] * (f.c2 * f.cP2) - Gb[1] * (f.c2 * f.sP2) - Gb[2] * f.s2;
    const P dphi1 = Ga[0] * f.Rh[1] - Ga[1] * f.Rh[0] + Gb[0] * f.u[1] - Gb[1] * f.u[0];
    const P dtheta1 = Ag[1] * Aa[2] - Ag[2] * Aa[1] + Bg[1] * Ab[2] - Bg[2] * Ab[1];
    const P dpsi1 = ga[0] * f.a[1] - ga[1] * f.a[0] + gb[0] * f.b[1] - gb[1] * f.b[0];
    dV[0] = (double)dPhi;
    dV[1] = (double)dTheta;
    dV[2] = (double)dR;
    dV[3] = (double)dphi1;
    dV[4] = (double)dtheta1;
    dV[5] = (double)dpsi1;
    dV[6] = (double)dphi2;
    dV[7] = (double)dtheta2;
    // each component is a sum of pair forces times a lever arm of at most R + r_CO
    noise = sizeof(P) == 4 ? kF32NoiseUnits * (double)mag * ((double)R + ch4co2::kRCO) : 0.0;
}

template <typename P>
__device__ __forceinline__ P ch4co2_pes(const double* y)
{
    Ch4Co2Trig tr;
    sincos_p<double>(y[2], &tr.sT, &tr.cT);
    sincos_p<double>(y[8], &tr.s1, &tr.c1);
    sincos_p<double>(y[10], &tr.sp, &tr.cp);
    sincos_p<double>(y[14], &tr.s2, &tr.c2);
    Ch4Co2Frame<P> f;
    ch4co2_frame<P>(y, tr, f);
    P gs[3], gd[3], mag;
    return ch4co2_sites<P>((P)y[4], f.a, f.b, gs, gd, mag);
}

// ------------------------------------------------------------------ systems as types
template <int SYS>
struct SysTraits;

template <>
struct SysTraits<SYS_HEAR> {
    static constexpr int QP = 6;
};
template <>
struct SysTraits<SYS_CO2AR> {
    static constexpr int QP = 10;
};
template <>
struct SysTraits<SYS_CH4CO2> {
    static constexpr int QP = 16;
};

// rhsMonomer ROTOR (src/hawaii.c:505-531): monomer state m = (phi, pphi, theta, ptheta, psi, ppsi),
// dV = dV/d(phi, theta, psi), sin/cos of theta and psi given
__device__ __forceinline__ void rotor_rhs(const double* II, const double* m, const double* dV, double st, double ct,
                                          double sp, double cp, double* d)
{
    const double pPhi = m[1], pTheta = m[3], pPsi = m[5];
    const double t1 = (pPhi - pPsi * ct) * sp + pTheta * st * cp;
    const double t2 = (pPhi - pPsi * ct) * cp - pTheta * st * sp;
    const double t3 = pPhi * ct - pPsi, t4 = pPsi * ct - pPhi;
    const double is = 1.0 / st, is2 = is * is, i0 = 1.0 / II[0], i1 = 1.0 / II[1];
    d[0] = (t1 * sp * i0 + t2 * cp * i1) * is2;
    d[1] = -dV[0];
    d[2] = (t1 * cp * i0 - t2 * sp * i1) * is;
    d[3] = -t3 * (t4 * (II[0] - II[1]) * cp * cp + pTheta * (II[0] - II[1]) * st * sp * cp + t4 * II[1]) * i0 * i1 *
               is2 * is - dV[1];
    d[4] = -(t1 * sp * i0 + t2 * cp * i1) * ct * is2 + pPsi / II[2];
    d[5] = t1 * t2 * (i1 - i0) * is2 - dV[2];
}

// Right-hand side of Hamilton's equations.  Intermolecular part src/hawaii.c:586-591, linear
// rotor src/hawaii.c:494-497.  noise: bound on the FP32 rounding noise of the momentum
// derivatives (0 when the surfaces are evaluated in FP64).
template <int SYS, typename P>
__device__ __forceinline__ void rhs(const SysConst& sc, const double* __restrict__ y, double* __restrict__ yd,
                                    double& noise)
{
    noise = 0.0;
    const double pPhi = y[1], Theta = y[2], pTheta = y[3], R = y[4], pR = y[5], mu = sc.mu;
    if constexpr (SYS == SYS_CH4CO2) {
        Ch4Co2Trig ct;
        sincos_p<double>(Theta, &ct.sT, &ct.cT);
        sincos_p<double>(y[8], &ct.s1, &ct.c1);
        sincos_p<double>(y[10], &ct.sp, &ct.cp);
        sincos_p<double>(y[14], &ct.s2, &ct.c2);
        double dV[8];
        ch4co2_dpes<P>(y, ct, dV, noise);
        const double iR = recip<P>(R), iR2 = iR * iR, imu = 1.0 / mu;
        const double is = recip<P>(ct.sT), is2 = is * is;
        const double a = imu * iR2;
        yd[4] = pR * imu;
        yd[5] = (pTheta * pTheta + pPhi * pPhi * is2) * a * iR - dV[2];
        yd[0] = pPhi * a * is2;
        yd[1] = -dV[0];
        yd[2] = pTheta * a;
        yd[3] = pPhi * pPhi * ct.cT * a * is2 * is - dV[1];
        rotor_rhs(sc.IR, y + 6, dV + 3, ct.s1, ct.c1, ct.sp, ct.cp, yd + 6);
        const double iI = 1.0 / sc.I, ist = recip<P>(ct.s2), ist2 = ist * ist;
        yd[12] = y[13] * iI * ist2;
        yd[13] = -dV[6];
        yd[14] = y[15] * iI;
        yd[15] = y[13] * y[13] * ct.c2 * iI * ist2 * ist - dV[7];
        return;
    }
    Trig tr;
    sincos_p<double>(Theta, &tr.sT, &tr.cT);
    double dV[5];
    if constexpr (SYS == SYS_HEAR) {
        dV[0] = 0.0;
        dV[1] = 0.0;
        dV[2] = hear_dV(R);
    } else {
        sincos_p<double>(y[8], &tr.st, &tr.ct);
        co2ar_dpes<P>((const Co2ArTables<P>*)sc.tables, y, tr, dV, noise);
    }
    const double iR = recip<P>(R), iR2 = iR * iR, imu = 1.0 / mu;
    const double is = recip<P>(tr.sT), is2 = is * is;
    const double a = imu * iR2;               // 1/(mu R^2)
    yd[4] = pR * imu;
    yd[5] = (pTheta * pTheta + pPhi * pPhi * is2) * a * iR - dV[2];
    yd[0] = pPhi * a * is2;
    yd[1] = -dV[0];
    yd[2] = pTheta * a;
    yd[3] = pPhi * pPhi * tr.cT * a * is2 * is - dV[1];
    if constexpr (SYS == SYS_CO2AR) {
        const double iI = 1.0 / sc.I, ist = recip<P>(tr.st), ist2 = ist * ist;
        const double pphi = y[7], ptheta = y[9];
        yd[6] = pphi * iI * ist2;
        yd[7] = -dV[3];
        yd[8] = ptheta * iI;
        yd[9] = pphi * pphi * tr.ct * iI * ist2 * ist - dV[4];
    }
}

template <int SYS, typename P>
__device__ __forceinline__ void rhs(const SysConst& sc, const double* __restrict__ y, double* __restrict__ yd)
{
    double noise;
    rhs<SYS, P>(sc, y, yd, noise);
}

template <int SYS, typename P>
__device__ __forceinline__ void dipole(const SysConst& sc, const double* y, P* mu)
{
    if constexpr (SYS == SYS_CH4CO2) {
        // The CH4-CO2 dipole (320 basis functions) is evaluated by its own batched kernel over
        // recorded sample coordinates (hh_ch4co2.hip); integrator code records coordinates
        // instead of calling this.  A call here is a bug: it yields NaN, which every dipole
        // check reports.
        mu[0] = mu[1] = mu[2] = (P)__builtin_nan("");
        return;
    } else if constexpr (SYS == SYS_HEAR) {
        P sT, cT, sP, cP;
        sincos_p<P>(y[2], &sT, &cT);
        sincos_p<P>(y[0], &sP, &cP);
        const P d = hear_dip<P>((P)y[4]);
        mu[0] = d * sT * cP;
        mu[1] = d * sT * sP;
        mu[2] = d * cT;
    } else {
        co2ar_dipole<P>((const Co2ArTables<P>*)sc.tables, y, mu);
    }
}

// kinetic_energy (src/hawaii.c:1041-1137), operation order of the reference
template <int SYS>
__device__ __forceinline__ double kinetic(const SysConst& sc, const double* y)
{
    const double pPhi = y[1], Theta = y[2], pTheta = y[3], R = y[4], pR = y[5], mu = sc.mu;
    double s, c_unused;
    sincos_p<double>(Theta, &s, &c_unused);
    double K = pR * pR / (2.0 * mu) + pTheta * pTheta / (2.0 * mu * R * R) + pPhi * pPhi / (2.0 * mu * R * R * s * s);
    if constexpr (SYS == SYS_CO2AR) {
        double st, ct_unused;
        sincos_p<double>(y[8], &st, &ct_unused);
        K = K + (y[9] * y[9] / (2.0 * sc.I) + y[7] * y[7] / (2.0 * sc.I * st * st));
    }
    if constexpr (SYS == SYS_CH4CO2) {
        // rotor (src/hawaii.c:1089-1110) and linear molecule (src/hawaii.c:1117-1128)
        double st, ct, sp, cp, s2, c2_unused;
        sincos_p<double>(y[8], &st, &ct);
        sincos_p<double>(y[10], &sp, &cp);
        sincos_p<double>(y[14], &s2, &c2_unused);
        const double A = (y[7] - y[11] * ct) * sp + y[9] * st * cp, B = (y[7] - y[11] * ct) * cp - y[9] * st * sp;
        const double K2 = A * A / (2.0 * sc.IR[0] * st * st) + B * B / (2.0 * sc.IR[1] * st * st) +
                          y[11] * y[11] / (2.0 * sc.IR[2]);
        const double K3 = y[15] * y[15] / (2.0 * sc.I) + y[13] * y[13] / (2.0 * sc.I * s2 * s2);
        K = K + K2 + K3;
    }
    return K;
}

template <int SYS, typename P>
__device__ __forceinline__ double pes_value(const SysConst& sc, const double* y)
{
    if constexpr (SYS == SYS_HEAR)
        return hear_V(y[4]);
    else if constexpr (SYS == SYS_CH4CO2)
        return (double)ch4co2_pes<P>(y);
    else
        return (double)co2ar_pes<P>((const Co2ArTables<P>*)sc.tables, y);
}

template <int SYS, typename P>
__device__ __forceinline__ double hamiltonian(const SysConst& sc, const double* y)
{
    if constexpr (SYS == SYS_CH4CO2) return kinetic<SYS_CH4CO2>(sc, y) + (double)ch4co2_pes<P>(y);
    const double pPhi = y[1], Theta = y[2], pTheta = y[3], R = y[4], pR = y[5], mu = sc.mu;
    double s, c_unused;
    sincos_p<double>(Theta, &s, &c_unused);
    double K = pR * pR / (2.0 * mu) + pTheta * pTheta / (2.0 * mu * R * R) + pPhi * pPhi / (2.0 * mu * R * R * s * s);
    double V;
    if constexpr (SYS == SYS_HEAR) {
        V = hear_V(R);
    } else {
        double st, ct_unused;
        sincos_p<double>(y[8], &st, &ct_unused);
        K += y[9] * y[9] / (2.0 * sc.I) + y[7] * y[7] / (2.0 * sc.I * st * st);
        V = (double)co2ar_pes<P>((const Co2ArTables<P>*)sc.tables, y);
    }
    return K + V;
}

// |d mu/dt|^2 at phase point y, the M2 integrand (calculate_M2, src/hawaii.c:1564-1594;
// integrand_M2, src/hep_hawaii.cpp:296-325): the lab dipole's derivative in every coordinate by
// central differences with h = 1e-3, contracted with dH/dp (the coordinate velocities, the even
// components of Hamilton's equations, compute_dHdp src/hawaii.c:1493-1522), summed over the
// coordinates in state order.  One dipole call site in a rolled loop over the 2 x (QP/2)
// displaced states keeps the CO2-Ar surface inlined once (no register spill); the displaced
// state is built with selects so it stays in registers.
template <int SYS>
__device__ __forceinline__ double dipole_rate_sq(const SysConst& sc, const double* y)
{
    constexpr int QP = SysTraits<SYS>::QP;
    constexpr double h = 1.0e-3;
    double yd[QP];
    rhs<SYS, double>(sc, y, yd);
    double md0 = 0.0, md1 = 0.0, md2 = 0.0;
    double dp0 = 0.0, dp1 = 0.0, dp2 = 0.0;
#pragma unroll 1
    for (int j = 0; j < QP; ++j) {
        const int c = j & ~1;   // the displaced coordinate: + h at even j, - h at odd j
        const double step = (j & 1) ? -h : h;
        double yy[QP];
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < QP; ++k) {
            yy[k] = k == c ? y[k] + step : y[k];
            v = k == c ? yd[k] : v;
        }
        double d[3];
        dipole<SYS, double>(sc, yy, d);
        if (j & 1) {
            md0 += v * ((dp0 - d[0]) / (2.0 * h));
            md1 += v * ((dp1 - d[1]) / (2.0 * h));
            md2 += v * ((dp2 - d[2]) / (2.0 * h));
        } else {
            dp0 = d[0];
            dp1 = d[1];
            dp2 = d[2];
        }
    }
    return md0 * md0 + md1 * md1 + md2 * md2;
}

}  // namespace hh
