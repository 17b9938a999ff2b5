/*
 * hawaii_gpu.h -- C ABI of the MI355X trajectory-ensemble engine (libhawaii_gpu.so).
 *
 * Plain C: pointers, sizes and POD structs; no HIP or torch types.  Device buffers are
 * passed as plain `double*` device addresses (hipMalloc'd or torch-allocated on the same
 * device).  Every entry point returns 0 on success or a negative HH_E* code; the message
 * of the last error on the calling thread is available from hh_last_error().  Nothing in
 * this library calls exit() or abort().
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   hh_sample_ics        q_generator / p_generator / reject / candidate loop
 *                        (src/hawaii.c:1206-1421, src/hawaii.c:2969-2993) and the R0
 *                        injection of the PR/mu path (src/hawaii.c:3486-3512), seeded as
 *                        init_ms(..., seed) -> mt_seed32 (src/hawaii.c:157-163)
 *   hh_mt_fill           mt_seed32 / mt_lrand (thirdparty/mtwist.c:372-406, 640-796;
 *                        thirdparty/mtwist.h:560-571)
 *   hh_cf_run            the trajectory loop of calculate_correlation_and_save
 *                        (src/hawaii.c:2969-3114) with correlation_eval
 *                        (src/hawaii.c:2116-2347) and the CVODE propagation it drives
 *                        (src/trajectory.c:146-200)
 *   hh_cf_array          the trajectory loop of calculate_correlation_array_and_save
 *                        (src/hawaii.c:2361-2793): one ensemble, CFs at satellite
 *                        temperatures by Boltzmann reweighting (src/hawaii.c:2639-2674)
 *   hh_prmu              the trajectory loop and spectrum accumulation of
 *                        calculate_spectral_function_using_prmu_representation_and_save
 *                        (src/hawaii.c:3148-3977), without requantised rotors
 *   hh_cf_to_sf          idct_cf_to_sf (src/hawaii.c:5569-5601)
 *   hh_m0                calculate_M0 / mpi_calculate_M0 (src/hawaii.c:1423-1491,
 *                        src/hawaii.c:1620-1738)
 *   hh_m2                calculate_M2 / mpi_calculate_M2 (src/hawaii.c:1524-1617,
 *                        src/hawaii.c:1740-1858)
 *   hh_vegas             c_mpi_perform_integration with integrand_M0 / integrand_M2 /
 *                        integrand_pf (src/hep_hawaii.cpp:116-380)
 *   hh_pf_analytic       analytic_full_partition_function_by_V (src/hawaii.c:5071-5090)
 *   hh_cf_moment         compute_Mn_from_cf_using_classical_detailed_balance
 *                        (src/hawaii.c:5945-5979)
 *   hh_sf_moment         compute_Mn_from_sf_using_classical_detailed_balance
 *                        (src/hawaii.c:5981-6013)
 *   hh_fit_baseline      FIT_BASELINE's wing-model fit (fit_baseline, src/hawaii.c:5313-5400)
 *   hh_loess             SMOOTH's LOESS smoothing (src/loess.cpp:98-451)
 *   hh_pipnn_*           the N2-Ar PIP-NN plugin surfaces (PES-IDS/ai_pes_n2_ar_pip_nn.cpp,
 *                        PES-IDS/ai_ids_n2_ar_pip_nn.cpp) as batched device evaluations
 * The .conf driver front end is declared in hawaii_conf.h.
 */
#ifndef HAWAII_GPU_H
#define HAWAII_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ errors */
enum {
    HH_OK = 0,
    HH_E_INVALID = -1,     /* bad argument */
    HH_E_DEVICE = -2,      /* HIP runtime error or no usable gfx950 device */
    HH_E_NOMEM = -3,       /* device or host allocation failed */
    HH_E_UNSUPPORTED = -4, /* system / mode not provided on the device */
    HH_E_BUSY = -5,        /* serving: no capacity */
    HH_E_NOTFOUND = -6,
};
const char* hh_last_error(void);

/* ------------------------------------------------------------------ systems */
/* He-Ar (ATOM + ATOM), CO2-Ar and N2-Ar (LINEAR_MOLECULE + ATOM), CH4-CO2 (ROTOR +
 * LINEAR_MOLECULE).  The trajectory engine integrates He-Ar and CO2-Ar; N2-Ar surfaces are
 * evaluated by hh_pipnn_*; the partition function covers all four. */
enum { HH_SYS_HEAR = 0, HH_SYS_CO2AR = 1, HH_SYS_N2AR = 2, HH_SYS_CH4CO2 = 3 };

typedef struct {
    int system;          /* HH_SYS_* */
    double mu;           /* reduced mass, a.u. (pair_reduced_mass) */
    double I;            /* linear-rotor moment of inertia II[0] (== II[1]); ignored for He-Ar */
    double I_rotor[3];   /* CH4-CO2: principal moments of the rotor (monomer 1); else ignored */
} hh_system;

/* Plugin surface ABI of the reference (src/hawaii.h:480-499): q = [Phi, Theta, R, monomer angles] */
typedef double (*hh_pes_fn)(double* q);
typedef void (*hh_dpes_fn)(double* q, double* dVdq);
typedef void (*hh_dipole_fn)(double* q, double mu[3]);

int hh_qp_size(int system);

/* ------------------------------------------------------------------ Mersenne Twister */
void hh_mt_fill(uint32_t seed, uint32_t* out, size_t n);

/* ------------------------------------------------------------------ initial conditions */
/* REJECTION: calculate_correlation_and_save's candidate loop; R0_INJECT: the PR/mu path's R0
 * injection; PHASE_SPACE: q_generator + p_generator only (every candidate kept) */
enum { HH_IC_REJECTION = 0, HH_IC_R0_INJECT = 1, HH_IC_PHASE_SPACE = 2 };
enum { HH_PAIR_FREE = 1, HH_PAIR_BOUND = 2, HH_PAIR_ALL = 3 }; /* PairState values of src/hawaii.h:291-297 */

typedef struct {
    int mode;          /* HH_IC_* */
    int pair_state;    /* rejection mode: FREE keeps E >= 0, BOUND keeps E <= 0 */
    double T;          /* K */
    double Rmin, Rmax; /* sampler_Rmin / sampler_Rmax, bohr */
    double pesmin;     /* Hartree (rejection mode) */
    double R0;         /* bohr (R0-injection mode) */
} hh_sampler;

/* Sample `ntraj` initial conditions from the seeded stream, row-major [ntraj][QP], on the
 * host, in the reference's draw order and arithmetic (bit-exact given the same pes).  The
 * rejection step calls `pes` (a plugin's pes_lab); NULL selects the engine's built-in
 * host surface.  Returns the number of candidates drawn (>= ntraj) or a negative code. */
long hh_sample_ics(const hh_system* sys, const hh_sampler* smp, uint32_t seed, long ntraj, hh_pes_fn pes,
                   double* ic_rowmajor, long* n_draws);

/* The same stream continued: the first skip_draws draws of the seeded stream are passed over
 * (the draws of an earlier call, its *n_draws), so consecutive calls sample as one loop would,
 * the preliminary moments' points first and then the trajectories' (src/hawaii.c:3443-3490).
 * *n_draws counts from the start of the stream, skipped draws included. */
long hh_sample_ics_from(const hh_system* sys, const hh_sampler* smp, uint32_t seed, long skip_draws, long ntraj,
                        hh_pes_fn pes, double* ic_rowmajor, long* n_draws);

/* The engine's built-in host potential of `system` in the plugin ABI (a pes_lab: q = [Phi,
 * Theta, R, monomer angles], Hartree), the one the sampler's rejection step uses when its pes is
 * NULL; NULL for a system without one.  No counterpart in the reference: it stands in for a
 * plugin's pes_lab in programs that have none. */
hh_pes_fn hh_host_pes(int system);

/* ------------------------------------------------------------------ engine */
typedef struct hh_engine hh_engine;

int hh_engine_create(int device, hh_engine** out);
void hh_engine_destroy(hh_engine* e);
int hh_engine_info(hh_engine* e, char* buf, size_t len);

enum { HH_ENGINE_PLAIN = 0, HH_ENGINE_BATCHED = 1 };
enum { HH_PREC_FP64 = 0, HH_PREC_MIXED = 1 };
enum { HH_HALVES_BOTH = 0, HH_HALVES_FORWARD = 1 };

typedef struct {
    double rtol, atol;   /* CVODE_TOLERANCE: rtol = atol in the reference (src/trajectory.c:214-230) */
    double dt;           /* sampling_time, a.t.u. */
    double Rcut;         /* bohr */
    double ppf;          /* partial_partition_function_ratio weight */
    int L;               /* MaxTrajectoryLength */
    int halves;          /* HH_HALVES_* */
    int max_steps;       /* max integrator steps between two samples (<= 0: 50000) */
    int engine;          /* HH_ENGINE_* */
    int precision;       /* HH_PREC_*; the plain engine is FP64 only */
    int block;           /* batched: 0 or 256, the one launch shape (256-lane workgroups, one
                            wavefront per SIMD) */
    int grid;            /* reserved (0) */
    int zimmermann;      /* use_zimmermann_trick: CF averaged over every time origin of the
                            trajectory's dipole series (correlation_eval_zimmerman_trick,
                            src/hawaii.c:1888-2114); both halves, L <= 4096 */
    double orbit_rtol;   /* mixed precision, systems with a rotor: an item whose pair reaches an
                            outer radial turning point (pR from > 0 to <= 0: an orbiting complex,
                            whose trajectory is chaotic) is integrated again from its initial
                            condition in FP64 at rtol = atol = orbit_rtol (0: 1e-12), concurrently
                            with the rest of the ensemble; < 0 keeps every item in mixed precision */
} hh_cf_params;

typedef struct {
    double integrate_ms;    /* device time of the integration launch (HIP events) */
    double reduce_ms;       /* device time of the CF reductions */
    long long n_items;      /* work items (trajectory halves) */
    long long nrhs;         /* right-hand-side evaluations */
    long long nsamples;     /* recorded dipole samples */
    long long n_failed;     /* items that ended on a dipole check (their trajectory counts as zero) */
    long long n_stepfail;   /* items that ended on a step-count / step-size limit: the partial CF is
                               kept, except that a trajectory whose last half hit the step-count
                               limit (CVode's CV_TOO_MUCH_WORK) counts as zero (src/hawaii.c:3001) */
    double max_e_drift;     /* max |H(end) - H(0)|, Hartree */
    int block, grid;        /* launch shape used */
    int launches;           /* integration kernel launches (time slices of the batched engine) */
    long long n_orbit;      /* items integrated again in FP64 (orbit_rtol); their rows, statuses and
                               lengths are the FP64 run's, nrhs counts both runs */
    long long nrhs_orbit;   /* right-hand sides of the FP64 runs of those items (part of nrhs) */
    long long nsamples_orbit; /* samples recorded by the FP64 runs (part of nsamples) */
    int pool_launches;      /* batched engine: launches of the run's own pool (its system and
                               precision; the FP64 orbit pool's launches are not counted) */
    double pool_kernel_ms;  /* their summed durations, each from its slice's start to the end of
                               that pool's launch on its stream (HIP events) */
} hh_cf_stats;

/* Run `njobs` correlation-function jobs of one system on the device.  ic_dev[j] holds
 * job j's initial conditions, SoA [QP][ntraj[j]] doubles; cf_dev[j] receives its ensemble
 * CF[L] = sum_traj ppf crln / ntraj (crln as correlation_eval).  The batched engine streams
 * all jobs through one persistent launch.  Synchronous. */
int hh_cf_run(hh_engine* e, const hh_system* sys, const hh_cf_params* p, int njobs, const double* const* ic_dev,
              const long* ntraj, double* const* cf_dev, hh_cf_stats* stats);

/* Per-item results of job `job` of the last hh_cf_run (host buffers of n_items entries;
 * any pointer may be NULL).  rows: [n_items][L] samples mu(0).mu(t_k), or with zimmermann set
 * [n_items][3][L] dipole components; entries past len are zero. */
int hh_cf_items(hh_engine* e, int job, int* len, int* status, int* nrhs, float* e_drift, double* rows);

/* ------------------------------------------------------------------ batched surfaces */
/* Evaluate the engine's device surfaces at n phase points y_dev (SoA [QP][n] doubles, state
 * layout [Phi, pPhi, Theta, pTheta, R, pR, (phi1, pphi1, theta1, ptheta1)]), FP64:
 *   HH_EVAL_PES        out[n]       V (Hartree)                    pes_lab (src/hawaii.h:480-499)
 *   HH_EVAL_DIPOLE     out[3][n]    lab-frame dipole (a.u.)        dipole_lab
 *   HH_EVAL_HAMILTONIAN out[n]      H = K + V                      Hamiltonian (src/hawaii.c:1190-1204)
 *   HH_EVAL_RHS        out[QP][n]   Hamilton's equations           rhs (src/hawaii.c:553-615)
 *   HH_EVAL_DIPOLE_RATE out[n]      |d mu/dt|^2 (a.u.), central    calculate_M2's per-point value
 *                                   differences h = 1e-3 . dH/dp   (src/hawaii.c:1564-1594)
 * The batched device form of a plugin's entry points. */
enum { HH_EVAL_PES = 0, HH_EVAL_DIPOLE = 1, HH_EVAL_HAMILTONIAN = 2, HH_EVAL_RHS = 3, HH_EVAL_DIPOLE_RATE = 4 };
int hh_eval(hh_engine* e, const hh_system* sys, int what, const double* y_dev, long n, double* out_dev);

/* CORRELATION_ARRAY: calculate_correlation_array_and_save's accumulation (src/hawaii.c:2361-2793,
 * weights src/hawaii.c:2639-2674).  The ensemble ic_dev was sampled at T_base (hh_sample_ics with
 * the same pair state); hh_cf_run integrates it once (p->ppf is ignored) and every trajectory's
 * CF is accumulated at each satellite temperature T_s with weight
 *     w = exp(-E HkT / T_base / T_s (T_base - T_s)) / WTT,     E = H at the initial point,
 * WTT = exp(-pesmin HkT / T_base / T_s (T_base - T_s)) for bound pairs with T_s < T_base, else 1;
 * free pairs take no weight above T_base.  Outputs, per temperature: cf_dev[st * L + k] =
 * sum(w ratio_st crln) / nstar_st and nstar[st] = sum(w) (the "trajectories" count the reference
 * writes).  A dipole failure adds its weight and a zero CF; a trajectory whose last half ended on
 * the step limit adds nothing (the array loop's `continue`, src/hawaii.c:2637, where hh_cf_run
 * counts it in ntraj as calculate_correlation_and_save does).  Zimmermann jobs are not accepted. */
typedef struct {
    double T_base;        /* K: the ensemble's sampling temperature */
    int pair_state;       /* HH_PAIR_FREE or HH_PAIR_BOUND */
    int ntemp;            /* 1..64 satellite temperatures */
    double pesmin;        /* Hartree: the bound-pair weight cap */
    const double* temps;  /* [ntemp] K, host */
    const double* ratios; /* [ntemp] partial partition function ratios, host (NULL: all 1) */
} hh_cf_array_params;

int hh_cf_array(hh_engine* e, const hh_system* sys, const hh_cf_params* p, const hh_cf_array_params* a,
                const double* ic_dev, long ntraj, double* cf_dev, double* nstar, hh_cf_stats* stats);

/* ------------------------------------------------------------------ PR/mu spectral function */
/* calculate_spectral_function_using_prmu_representation_and_save (src/hawaii.c:3148-3977) for
 * pairs without requantised rotors (ATOM + ATOM, LINEAR_MOLECULE + ATOM), without the Poisson
 * free-path cut-off and spin weights (both need requantised rotors in the reference).  ic_dev
 * (SoA [QP][ntraj]) holds R0-injected initial conditions (hh_sample_ics, HH_IC_R0_INJECT).
 * Each trajectory is integrated forward with its lab dipole sampled at t = dt, 2 dt, ... up to
 * and including the first sample beyond R0; a trajectory still inside R0 after L samples, or
 * one whose dipole failed, is not counted (counted[t] = 0: the reference draws another one in
 * its place).  For every counted trajectory
 *     sf_dev[k] += SF_COEFF * (-pR / mu) * sum_c |FFT_L(Connes(mu_c))_k|^2,  k < nfreq,
 * SF_COEFF = Hartree ALU^6 ATU R0^2 dt^2 (src/hawaii.c:3395, 3806-3822); the caller divides
 * by the number of counted trajectories.  L is the FFT length (a power of two). */
typedef struct {
    double rtol, atol;      /* CVODE_TOLERANCE */
    double dt;              /* sampling_time */
    double R0;              /* injection distance; a trajectory ends at the first sample beyond it */
    int L;                  /* MaxTrajectoryLength: samples per trajectory and FFT length */
    int nfreq;              /* spectral points kept (frequency_array_length), <= L/2 + 1 */
    int precision;          /* HH_PREC_FP64 or HH_PREC_MIXED */
    int max_steps;          /* max integrator steps between two samples (<= 0: 50000) */
    int allow_truncating;   /* count trajectories still inside R0 after L samples
                               (allow_truncating_trajectories_at_length_limit) */
    double orbit_rtol;      /* mixed precision: orbiting pairs integrated again in FP64 at this
                               tolerance (as hh_cf_params.orbit_rtol; 0: 1e-12, < 0: off) */
} hh_prmu_params;

int hh_prmu(hh_engine* e, const hh_system* sys, const hh_prmu_params* p, const double* ic_dev, long ntraj,
            double* sf_dev, int* counted, long* ncounted, hh_cf_stats* stats);

/* ------------------------------------------------------------------ CF -> SF on the device */
/* Connes apodization in place on cf_dev, Makhoul IDCT through a 2n-point complex FFT,
 * unit scaling: nu_dev[n] in cm^-1, sf_dev[n] (idct_cf_to_sf).  n must be a power of two. */
int hh_cf_to_sf(hh_engine* e, double* cf_dev, long n, double dt, double* nu_dev, double* sf_dev, double* ms);

/* hh_cf_to_sf without any device allocation: no counterpart in the reference; for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The transform's work sequences live in a static device array of the library
 * (4 MiB per device, held until the process ends), so n <= 65536 (HH_E_INVALID above).  Same
 * kernels and results as hh_cf_to_sf.  Every call on a device shares that array: calls that
 * overlap on one device race. */
int hh_cf_to_sf_noalloc(hh_engine* e, double* cf_dev, long n, double dt, double* nu_dev, double* sf_dev);

/* ------------------------------------------------------------------ M0 phase-space moment */
/* Rejection-sampler M0 over `npoints` accepted points (calculate_M0 semantics) with the
 * candidate stream split into `nstreams` independent Mersenne Twister streams seeded
 * from `seed` (stream s: seed + s); m, q as in the reference (times ppf). */
int hh_m0(hh_engine* e, const hh_system* sys, const hh_sampler* smp, uint32_t seed, long npoints, int nstreams,
          double ppf, double* m, double* q, double* ms);

/* Rejection-sampler M2 (calculate_M2 / mpi_calculate_M2, src/hawaii.c:1524-1617, 1740-1858):
 * the same streams and candidate loop as hh_m0, averaging |d mu/dt|^2 (HH_EVAL_DIPOLE_RATE) over
 * the accepted points; m, q scaled by SecondCoeff (src/constants.h:30) times ppf. */
int hh_m2(hh_engine* e, const hh_system* sys, const hh_sampler* smp, uint32_t seed, long npoints, int nstreams,
          double ppf, double* m, double* q, double* ms);

/* ------------------------------------------------------------------ VEGAS phase-space integrals */
/* c_mpi_perform_integration / mpi_perform_integration (src/hep_hawaii.cpp:332-380) with the
 * integrands integrand_M0, integrand_M2 and integrand_pf (src/hep_hawaii.cpp:116-330) over the
 * phase space of transform_variables (src/hep_hawaii.cpp:8-114): adaptive Monte Carlo (Lepage's
 * VEGAS) on the device, niterations of npoints, the first iteration discarded.  M0 = value *
 * ZeroCoeff / hh_pf_analytic; M2 = value * SecondCoeff / hh_pf_analytic (src/driver.c:3377-3381);
 * the partition-function ratio = value(PF) / hh_pf_analytic.  The estimate
 * agrees with the reference's hep-mc run statistically (same integrand, own sampler). */
enum { HH_INTEGRAND_M0 = 0, HH_INTEGRAND_M2 = 1, HH_INTEGRAND_PF = 2 };   /* IntegrandType order */

typedef struct {
    int integrand;          /* HH_INTEGRAND_M0, HH_INTEGRAND_M2 or HH_INTEGRAND_PF */
    int pair_state;         /* HH_PAIR_FREE (E >= 0), HH_PAIR_BOUND (E <= 0), HH_PAIR_ALL */
    double T;               /* K */
    double Rmin, Rmax;      /* sampler_Rmin, sampler_Rmax */
    int niterations;        /* the reference's default: 12 */
    long npoints;           /* per iteration; the reference's default: 1000000 */
    uint32_t seed;
} hh_vegas_params;

typedef struct {
    double value, error;    /* the integral and its standard error */
    double chi2_dof;        /* consistency of the accumulated iterations */
    int iterations;
    long long points;
    long long nonfinite;    /* points whose integrand was not finite (counted as 0) */
    double ms;              /* device time */
    double iter_value[32], iter_error[32];   /* per iteration (first 32) */
} hh_vegas_result;

int hh_vegas(hh_engine* e, const hh_system* sys, const hh_vegas_params* p, hh_vegas_result* out);

/* ------------------------------------------------------------------ moments, partition function */
/* analytic_full_partition_function_by_V (src/hawaii.c:5071-5090): classical kinetic partition
 * function per unit volume at T (K).  Returns a negative value (and sets the error) on bad input. */
double hh_pf_analytic(const hh_system* sys, double T);

/* compute_Mn_from_cf_using_classical_detailed_balance (src/hawaii.c:5945-5979) on a host CF:
 * n = 0: M0 = cf[0] ZeroCoeff / ALU^3;  n = 2: M2 = -SecondCoeff cf''(0) / ALU^3 from the central
 * difference over 21 points (len >= 11) or 9 points (len >= 5), cf even in t, sample step dt.
 * ntraj > 0 marks an unnormalised CF (divided by ntraj); ntraj <= 0 a normalised one. */
int hh_cf_moment(const double* cf, long len, double dt, double ntraj, int n, double* out);

/* compute_Mn_from_sf_using_classical_detailed_balance (src/hawaii.c:5981-6013) on a host SF:
 * Mn = 2 Moment_SF_Coeff * int sf nu^n d nu over the uniform grid nu[len] (cm^-1) by the composite
 * Simpson 3/8 weights of integrate_composite_simpson (src/hawaii.c:5926-5943); odd n give 0;
 * ntraj > 0 divides (an unnormalised SF, as the PR/mu loop's estimates). */
int hh_sf_moment(const double* nu, const double* sf, long len, double ntraj, int n, double* out);

/* FIT_BASELINE (execute_fit_baseline, src/driver.c:2072-2142; fit_baseline and
 * gsl_nonlinear_opt, src/hawaii.c:5215-5400): the wing model C + A / (1 + B^2 t^2) fitted by
 * Levenberg-Marquardt to cf[begin..len) (times t in a.u., the CF's own units), with times scaled
 * by that of the CF's secondary maximum and values by the maximum; abc = {A, |B|, C} in the CF's
 * units.  HH_E_INVALID for begin >= len or a CF without a secondary maximum (it never turns
 * negative and rises again).  Host computation; no device is used. */
int hh_fit_baseline(const double* t, const double* cf, long len, long begin, double* abc);

/* SMOOTH (execute_smooth, src/driver.c:2849-2951; loess_init, loess_create_grid,
 * loess_apply_smoothing and loess_estimate, src/loess.cpp:98-451): LOESS with local cubics and
 * tricube weights of the series (x, y)[len] evaluated on npoints equidistant points from 0 to
 * grid_max (nu_out, y_out [npoints]); the window holds ws_min points, ws_min + ws_step i
 * (truncated) beyond grid point ws_delay.  HH_E_INVALID for a grid outside the data, npoints <= 0,
 * grid_max <= 0 or a window of fewer than 4 or more than len points.  Host computation. */
int hh_loess(const double* x, const double* y, long len, double grid_max, long npoints, long ws_min, double ws_step,
             long ws_delay, double* nu_out, double* y_out);

/* ------------------------------------------------------------------ PIP-NN surfaces (N2-Ar) */
/* The N2-Ar permutation-invariant-polynomial neural-network surfaces
 * (PES-IDS/ai_pes_n2_ar_pip_nn.cpp, PES-IDS/ai_ids_n2_ar_pip_nn.cpp, PES-IDS/mlp.hpp,
 * PES-IDS/QModel.hpp; the plugin's pes_lab / dipole_lab of examples/correlation_n2_ar.cpp:11-75)
 * evaluated in batches on the device: the polynomial features in FP32, every network layer as
 * BF16 matrix-core products with FP32 accumulation across the batch of phase points. */
enum { HH_ACT_SILU = 0, HH_ACT_TANH = 1 };
enum { HH_MLP_BF16 = 0,      /* one BF16 product per layer */
       HH_MLP_BF16X3 = 1 };  /* operands split into BF16 high + low parts, three products
                                (hi*hi + hi*lo + lo*hi): FP32-level accuracy on the same units */

typedef struct {
    int n_layers;            /* linear layers (hidden + output), 2..4; 0 = absent */
    int dims[5];             /* dims[0] inputs ... dims[n_layers] outputs */
    int activation;          /* HH_ACT_*, after every layer but the last */
    const double* W[4];      /* layer k: row-major [dims[k]][dims[k+1]] (y = x W + b) */
    const double* b[4];      /* [dims[k+1]] */
} hh_mlp_desc;

typedef struct {
    /* PES: 17 polynomials -> network -> E = (yscale * out + ymean - en_inf) / HTOCM (Hartree).
     * Shapes: 17-32-1 (the reference model) or 17-H-H-H-1 with H = 64 or 128. */
    hh_mlp_desc pes;
    const double* pes_xmean;     /* [17] StandardScaler of the inputs */
    const double* pes_xscale;    /* [17] */
    double pes_ymean, pes_yscale, pes_en_inf;
    /* dipole: [short (a0 = 2), long (a0 = 6)] x 3 subnets 119-16-2, 119-16-2, 70-8-1 (SiLU) on
     * 308 standardised polynomials; dip[0][0].n_layers = 0 leaves the dipole out */
    hh_mlp_desc dip[2][3];
    const double* dip_xmean[2];  /* [308] */
    const double* dip_xscale[2];
} hh_pipnn_desc;

typedef struct hh_pipnn hh_pipnn;
int hh_pipnn_create(hh_engine* e, const hh_pipnn_desc* desc, int precision, hh_pipnn** out);
/* q_dev: lab coordinates SoA [5][n] (Phi, Theta, R, phi1, theta1).  energy_dev[n] (Hartree,
 * pes_lab) and dipole_dev[3][n] (a.u., dipole_lab) may each be NULL.  ms (optional): device
 * time of the launches.  Synchronous. */
int hh_pipnn_eval(hh_pipnn* m, const double* q_dev, long n, double* energy_dev, double* dipole_dev, double* ms);
void hh_pipnn_destroy(hh_pipnn* m);

/* ------------------------------------------------------------------ serving session */
/* A session keeps lane pools resident on the device and runs submitted requests with
 * continuous batching: a background thread runs the pools in time slices; between slices it
 * admits waiting requests while their device buffers fit the memory budget (FIFO), reduces
 * every CF job whose trajectories have all finished and advances M0 requests by one slice.
 * Requests of different systems, temperatures and sizes overlap on the device, and each
 * result is the same bits as the same request run alone.  Caller buffers (ic_dev, cf_dev)
 * must stay valid until the request has finished.  Thread-safe.  New interface (the
 * reference has no serving path); it drives the same work as hh_cf_run and hh_m0. */
typedef struct hh_session hh_session;

typedef struct {
    double slice_ms;          /* pool kernel time slice (<= 0: 10 ms) */
    double max_device_gib;    /* admission budget for request buffers (<= 0: 80% of the free HBM) */
    int block;                /* 0 or 256: the lane pools' one launch shape */
} hh_session_opts;

enum { HH_REQ_WAITING = 0, HH_REQ_RUNNING = 1, HH_REQ_DONE = 2, HH_REQ_FAILED = 3 };
enum { HH_REQ_CF = 0, HH_REQ_M0 = 1 };

typedef struct {
    int kind;                 /* HH_REQ_CF / HH_REQ_M0 */
    int state;                /* HH_REQ_* */
    double submit_ms;         /* session clock (ms since hh_session_create) */
    double admit_ms;
    double done_ms;
    double latency_ms;        /* done - submit */
    long long trajectories;   /* CF: ntraj; M0: accepted points */
    hh_cf_stats cf;           /* CF requests (integrate_ms / reduce_ms are the session's
                                 device time while the request was running) */
    double m0, m0_err;        /* M0 requests: m, q as hh_m0 */
    char error[160];          /* HH_REQ_FAILED */
} hh_request_info;

int hh_session_create(hh_engine* e, const hh_session_opts* opts, hh_session** out);
int hh_submit_cf(hh_session* s, const hh_system* sys, const hh_cf_params* p, const double* ic_dev, long ntraj,
                 double* cf_dev, uint64_t* id);
int hh_submit_m0(hh_session* s, const hh_system* sys, const hh_sampler* smp, uint32_t seed, long npoints,
                 int nstreams, double ppf, uint64_t* id);
/* Wait up to timeout_ms (< 0: forever) for request `id`; returns 1 when it has finished
 * (done or failed; `info` filled, the request forgotten), 0 on time-out. */
int hh_poll(hh_session* s, uint64_t id, double timeout_ms, hh_request_info* info);
/* Finish every submitted request, stop the worker and free the session. */
int hh_session_destroy(hh_session* s);

#ifdef __cplusplus
}
#endif
#endif
