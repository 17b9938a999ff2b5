/*
 * CPU oracle for the trajectory-ensemble path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker
 * (or, for cpu_baseline, as the CPU baseline being timed).  The engine under
 * hawaiihybrid_amd/ never links or calls it.
 *
 * It restates, in plain C, the reference CPU algorithm of the path:
 *   Mersenne Twister (thirdparty/mtwist.c), initial-condition sampling
 *   (src/hawaii.c:1206-1421, 2969-2993, 3486-3512), Hamilton's equations
 *   (src/hawaii.c:474-615, 1041-1204), the He-Ar and CO2-Ar surfaces
 *   (PES-IDS/HeAr.h, PES-IDS/ai_pes_co2ar*, PES-IDS/ai_ids_co2ar*,
 *   src/angles_handler.cpp:45-193), the per-trajectory correlation function
 *   (src/hawaii.c:2116-2347), ensemble accumulation (src/hawaii.c:2969-3114),
 *   CF -> SF (src/hawaii.c:5401-5601) and the rejection-sampler M0
 *   (src/hawaii.c:1423-1491).
 *
 * The one deliberate replacement: SUNDIALS CVODE (BDF, not vendored) is replaced by
 * the explicit embedded Runge-Kutta 8(5,3) pair of Dormand & Prince with its 7th-order
 * dense output (Hairer, Norsett & Wanner, "Solving ODEs I", II.5 and II.6: DOP853),
 * sampled at t_k = k*dt like CVode's CV_NORMAL interpolation (src/trajectory.c:155).
 */
#ifndef HH_ORACLE_H
#define HH_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- MT19937 */
typedef struct {
    uint32_t mt[624];
    int idx;
} orc_mt;

void     orc_mt_seed(orc_mt *s, uint32_t seed);
uint32_t orc_mt_u32(orc_mt *s);
double   orc_mt_drand(orc_mt *s);
void     orc_mt_fill(uint32_t seed, uint32_t *out, size_t n);

/* ---------------------------------------------------------------- systems */
enum { ORC_SYS_HEAR = 0, ORC_SYS_CO2AR = 1, ORC_SYS_CH4CO2 = 3 };

typedef struct {
    int system;      /* ORC_SYS_* */
    double mu;       /* reduced mass, a.u. */
    double II[2];    /* linear-rotor inertia (II[0] == II[1]); unused for He-Ar */
    double IR[3];    /* CH4-CO2: principal moments of the rotor (monomer 1) */
} orc_system;

int    orc_qp_size(const orc_system *s);
double orc_pes(const orc_system *s, const double *q);
void   orc_dpes(const orc_system *s, const double *q, double *dvdq);
void   orc_dipole(const orc_system *s, const double *q, double mu[3]);
void   orc_rhs(const orc_system *s, const double *y, double *ydot);
double orc_kinetic(const orc_system *s, const double *y);
double orc_hamiltonian(const orc_system *s, const double *y);

/* He-Ar radial functions (PES-IDS/HeAr.h) */
double orc_hear_V(double R);
double orc_hear_dV(double R);
double orc_hear_dip(double R);

/* CO2-Ar molecular-frame functions (R, theta_m) */
double orc_co2ar_V_cm(double R, double theta);            /* cm^-1 */
void   orc_co2ar_dV_cm(double R, double theta, double *dR, double *dTheta);
void   orc_co2ar_dip_mol(double R, double theta, double d[3]);
void   orc_co2ar_lab_to_mol(const double *qlab, double *qmol);
void   orc_co2ar_jacobian(const double *qlab, const double *qmol, double jac[5][5]);
/* molecular-frame vector of a linear molecule + atom -> lab frame */
void   orc_lm_mol_to_lab(const double *qmol, const double dm[3], double mu[3]);

/* CH4-CO2 (ch4co2.c): lab q[8] = [Phi, Theta, R, phi1, theta1, psi1, phi2, theta2]; the frame
 * of CH4 ("kal": R, phi1K, theta1K, phi2K, theta2K) and its 8 x 5 Jacobian as the reference;
 * the stand-in PES in kal coordinates (Hartree) and its five derivatives; the dipole in the
 * CH4 frame and in the lab */
void   orc_ch4co2_lab_to_kal(const double *qlab, double *qkal);
void   orc_ch4co2_jacobian(const double *qlab, const double *qkal, double jac[8][5]);
double orc_ch4co2_pes_kal(double R, double phi1, double theta1, double phi2, double theta2);
void   orc_ch4co2_dpes_kal(double R, double phi1, double theta1, double phi2, double theta2, double d[5]);
double orc_ch4co2_well_scale(void);
void   orc_ch4co2_dipole_kal(const double *qkal, double d[3]);
double orc_ch4co2_pes(const double *q);
void   orc_ch4co2_dpes(const double *q, double *dvdq);
void   orc_ch4co2_dipole(const double *q, double mu[3]);
void   orc_ch4co2_rhs(const orc_system *s, const double *y, double *ydot);
double orc_ch4co2_kinetic(const orc_system *s, const double *y);

/* ---------------------------------------------------------------- sampler */
enum { ORC_IC_REJECTION = 0, ORC_IC_R0_INJECT = 1, ORC_IC_PHASE_SPACE = 2 };
enum { ORC_PAIR_FREE = 1, ORC_PAIR_BOUND = 2, ORC_PAIR_ALL = 3 };

typedef struct {
    int mode;          /* ORC_IC_* */
    int pair_state;    /* ORC_PAIR_*; rejection mode only */
    double T;          /* K */
    double Rmin, Rmax; /* bohr */
    double pesmin;     /* Hartree */
    double R0;         /* bohr; r0_inject mode only */
} orc_sampler;

/* Draw ntraj accepted initial conditions from one MT stream seeded with `seed`.
 * ic is row-major [ntraj][QP].  Returns the number of candidates drawn. */
long orc_sample_ics(const orc_system *s, const orc_sampler *sp, uint32_t seed,
                    long ntraj, double *ic, long *n_draws);

/* ---------------------------------------------------------------- integrator / CF */
enum { ORC_HALVES_BOTH = 0, ORC_HALVES_FORWARD = 1 };
enum { ORC_OK = 0, ORC_FAIL_DIPOLE_NAN = 1, ORC_FAIL_CF_HUGE = 2, ORC_FAIL_CF_JUMP = 3,
       ORC_FAIL_STEPS = 4, ORC_FAIL_STEPSIZE = 5 };

typedef struct {
    double rtol, atol;  /* CVODE_TOLERANCE (src/trajectory.c:214-230): rtol = atol */
    double dt;          /* sampling_time */
    int L;              /* MaxTrajectoryLength */
    double Rcut;
    int halves;         /* ORC_HALVES_* */
    long max_steps;     /* per output interval (CVodeSetMaxNumSteps, src/trajectory.c:62) */
    int zimmermann;     /* use_zimmermann_trick (both halves; src/hawaii.c:1888-2114) */
} orc_cf_params;

typedef struct {
    long nrhs;          /* right-hand-side evaluations */
    long ndip;          /* dipole evaluations */
    long naccept, nreject;
    double e_drift;     /* max |H(end) - H(0)| over the halves */
} orc_traj_stats;

/* One half: propagate y0 and fill c[0..L-1] with mu(0).mu(t_k).  Returns ORC_* and
 * writes the number of filled samples to *len (c[len..L-1] is zero). */
int orc_half_cf(const orc_system *s, const orc_cf_params *p, const double *y0,
                double *c, int *len, orc_traj_stats *st);

/* One trajectory as correlation_eval: crln[k] = 1/2 (c_f + c_b) ALU^3 (both halves) or
 * c_f ALU^3 (forward only).  A dipole-check failure zeroes crln (src/hawaii.c:2183-2237), and
 * so does a last half that ended on the step limit (CV_TOO_MUCH_WORK, src/hawaii.c:3001). */
int orc_traj_cf(const orc_system *s, const orc_cf_params *p, const double *y0,
                double *crln, int *len_f, int *len_b, orc_traj_stats *st);

/* The PR/mu trajectory: the dipole series vec[3k + i] at t = k dt from y0 (forward) up to and
 * including the first sample with R > p->Rcut, at most p->L samples; dipole check only. */
int orc_prmu_series(const orc_system *s, const orc_cf_params *p, const double *y0, double *vec, int *len,
                    orc_traj_stats *st);

/* The Zimmermann sum over time origins of one dipole series d[3 * (2L - 1)] (d[3c + i] is
 * component i at series index c, mu(0) at c = L - 1):
 * crln[s] = ALU^3 / L sum_{c=0}^{L-1} d(c) . d(c + s)  (src/hawaii.c:2089-2094). */
void orc_zimmermann_series_cf(const double *d, int L, double *crln);

/* Per-trajectory rows, threads over trajectories.  crln: [ntraj][L]. */
void orc_cf_rows(const orc_system *s, const orc_cf_params *p, const double *ic, long ntraj,
                 int nthreads, double *crln, int *status, int *len_f, int *len_b,
                 long *nrhs, long *ndip, double *e_drift);

/* Ensemble CF with per-thread partial sums (for timing): cf[k] = sum ppf*crln / ntraj. */
void orc_cf_ensemble(const orc_system *s, const orc_cf_params *p, const double *ic, long ntraj,
                     int nthreads, double ppf, double *cf, long *nrhs_total);

/* ---------------------------------------------------------------- CF -> SF */
void orc_fft_radix2(double *re, double *im, size_t n, int sign); /* in place, unscaled */
void orc_idct(const double *v, size_t n, double *out);
void orc_dct(const double *v, size_t n, double *out);
void orc_connes(double *a, size_t n, double dt);
/* t[n], cf[n] (apodized in place as the reference does) -> nu[n] (cm^-1), sf[n] */
void orc_cf_to_sf(const double *t, double *cf, size_t n, double *nu, double *sf);
/* one PR/mu spectrum: x[3][L] series -> sum_c |FFT(Connes(x_c))_k|^2, k < nfreq */
void orc_prmu_power(const double *x, size_t L, double dt, size_t nfreq, double *out);

/* ---------------------------------------------------------------- M0 */
/* Rejection-sampler M0 (calculate_M0, src/hawaii.c:1423-1491): m = mean(mu.mu)*ZeroCoeff*ppf
 * q = sqrt(var/n)*ZeroCoeff*ppf, from `seed`. */
void orc_m0(const orc_system *s, const orc_sampler *sp, uint32_t seed, long npoints, double ppf,
            double *m, double *q);

/* ---------------------------------------------------------------- M2 */
/* |d mu/dt|^2 at the full state y (calculate_M2's per-point value, src/hawaii.c:1564-1594):
 * central differences (h = 1e-3) of the dipole in every coordinate, contracted with dH/dp (the
 * even components of orc_rhs, compute_dHdp src/hawaii.c:1493-1522). */
double orc_dipole_rate_sq(const orc_system *s, const double *y);
/* Rejection-sampler M2 (calculate_M2, src/hawaii.c:1524-1617): the orc_m0 loop with
 * orc_dipole_rate_sq as the value, scaled by SecondCoeff * ppf. */
void orc_m2(const orc_system *s, const orc_sampler *sp, uint32_t seed, long npoints, double ppf,
            double *m, double *q);

/* analytic_full_partition_function_by_V (src/hawaii.c:5071-5090) */
double orc_pf_analytic(const orc_system *s, double T);

/* compute_Mn_from_cf_using_classical_detailed_balance (src/hawaii.c:5945-5979), n = 0 or 2;
 * ntraj > 0 divides (an unnormalised CF).  Returns 0, or -1 when the CF is too short. */
int orc_cf_moment(const double *cf, size_t len, double dt, double ntraj, int n, double *out);
/* compute_Mn_from_sf_using_classical_detailed_balance (src/hawaii.c:5981-6013); ntraj > 0 divides */
double orc_sf_moment(const double *nu, const double *sf, size_t len, double ntraj, int n);
/* CORRELATION_ARRAY weights per trajectory and satellite temperature (moments.c) */
void orc_cf_array_weights(const double *E, long n, double Tb, int pair_state, double pesmin, int M,
                          const double *temps, const double *ratios, const int *counted, double *w, double *nstar);

/* ---------------------------------------------------------------- N2-Ar PIP-NN surfaces */
/* A multilayer perceptron as the reference's MLPES / qm::MLP (PES-IDS/mlp.hpp:178-209,
 * PES-IDS/QModel.hpp:206-230): layer k maps dims[k] -> dims[k+1] with W[k] row-major
 * [dims[k]][dims[k+1]] (x W + b); the activation follows every layer but the last. */
enum { ORC_ACT_SILU = 0, ORC_ACT_TANH = 1 };
typedef struct {
    int n_layers;
    int dims[5];
    int act;
    const double *W[4];
    const double *b[4];
} orc_mlp;

/* The N2-Ar surfaces (PES-IDS/ai_pes_n2_ar_pip_nn.cpp, PES-IDS/ai_ids_n2_ar_pip_nn.cpp):
 * PES = network on 17 standardised PIPs, inverse-scaled, minus en_inf, in cm^-1;
 * dipole = two charge models (short a0 = 2, long a0 = 6) of three subnets each on 308
 * standardised PIPs, blended over R in [9.5, 12.5]. */
typedef struct {
    orc_mlp pes;
    const double *pes_xmean, *pes_xscale;
    double pes_ymean, pes_yscale, pes_en_inf;
    orc_mlp dip[2][3];
    const double *dip_xmean[2], *dip_xscale[2];
} orc_n2ar;

double orc_mlp_forward(const orc_mlp *m, const double *x, double *out);   /* returns out[0] */
void   orc_pip_n2ar_pes(const double y[3], double p[17]);
void   orc_pip_n2ar_dip(const double y[10], double p[308]);
double orc_n2ar_pes_xyz(const orc_n2ar *m, const double xyz[9]);             /* cm^-1 */
void   orc_n2ar_dpes_xyz(const orc_n2ar *m, const double xyz[9], double g[9]); /* cm^-1 / bohr */
double orc_n2ar_pes(const orc_n2ar *m, const double q[5]);                    /* Hartree, lab q */
void   orc_n2ar_dpes(const orc_n2ar *m, const double q[5], double dvdq[5]);   /* Hartree / (bohr, rad) */
void   orc_n2ar_dipole_bf(const orc_n2ar *m, double R, double theta, double d[3]);
/* one charge model alone: 0 short-range (a0 = 2), 1 long-range (a0 = 6) */
void   orc_n2ar_dipole_model_bf(const orc_n2ar *m, int model, double R, double theta, double d[3]);
void   orc_n2ar_dipole(const orc_n2ar *m, const double q[5], double mu[3]);   /* lab, a.u. */
/* plain network forward of the synthetic shape on the PES features (no scalers) */
double orc_n2ar_mlp_on_pes_features(const orc_mlp *net, const double *xmean, const double *xscale,
                                    const double q[5]);

#ifdef __cplusplus
}
#endif
#endif
