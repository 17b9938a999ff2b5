/*
 * hawaii_conf.h -- the reference driver's .conf front end on the MI355X engine.
 *
 * Reads the configuration format of the reference driver (src/driver.c:54-1522: &INPUT,
 * &MONOMER and &PROCESSING blocks, `key = value` lines, `!` comments, case-insensitive
 * keywords) and runs the part of it that is this engine's path:
 *   CALCULATION_TYPE = CORRELATION_SINGLE   calculate_correlation_and_save
 *                                           (src/hawaii.c:2795-3117): preliminary M0 and M2
 *                                           (INITIALM0_NPOINTS, INITIALM2_NPOINTS) from the
 *                                           rejection sampler, NITERATIONS iterations of
 *                                           TOTAL_TRAJECTORIES/NITERATIONS trajectories from the
 *                                           seeded stream, the CF file rewritten after each
 *                                           iteration (write_correlation_function_ext,
 *                                           src/hawaii.c:4150-4212)
 *   CALCULATION_TYPE = CORRELATION_ARRAY    calculate_correlation_array_and_save
 *                                           (src/hawaii.c:2361-2793): one ensemble per
 *                                           iteration reweighted to SATELLITE_TEMPERATURES
 *   CALCULATION_TYPE = CALCULATE_PHASE_SPACE_M0   the VEGAS M0 integral at TEMPERATURE or each
 *                                           of TEMPERATURES (src/driver.c:3389-3445)
 *   CALCULATION_TYPE = CALCULATE_PHASE_SPACE_M2   the VEGAS M2 integral at TEMPERATURE
 *                                           (src/driver.c:3344-3388)
 *   CALCULATION_TYPE = PR_MU                calculate_spectral_function_using_prmu_...
 *                                           (src/hawaii.c:3148-3977) for ATOM + ATOM and
 *                                           LINEAR_MOLECULE + ATOM: preliminary M0 and M2, then
 *                                           R0-injected trajectories, SF_FILENAME rewritten after
 *                                           each iteration
 *   CALCULATION_TYPE = PROCESSING           the stack machine of run_processing
 *                                           (src/driver.c:2956-3081) restricted to READ_CF,
 *                                           READ_SF, READ_SPECTRUM, CF_TO_SF, CMP, DUP, DUP2,
 *                                           DROP, DROP2, WRITE_CF, WRITE_SF, WRITE_SPECTRUM,
 *                                           ALPHA, AVERAGE_CFS, ADD_SPECTRA, COMPUTE_Mn_CLASSICAL_/
 *                                           QUANTUM_DETAILED_BALANCE and the desymmetrizations D1, D2,
 *                                           D3, D4, D4a
 * Systems: two ATOM monomers (He-Ar surfaces), LINEAR_MOLECULE + ATOM (CO2-Ar surfaces) or
 * ROTOR + LINEAR_MOLECULE (CH4-CO2 surfaces, CORRELATION_SINGLE).
 * Without PARTIAL_PARTITION_FUNCTION_RATIO the ratio is integrated by VEGAS as the reference does
 * (src/hawaii.c:2826-2880), and the VEGAS M0 and M2 follow.  Plugins named by so_potential and
 * so_dipole are dlopen'ed; pes_lab drives the host rejection sampler, so the sampled phase points
 * follow the reference's draws, and the device integrates its built-in surfaces, so a plugin is
 * accepted only when it is the built-in surface of the configured monomers (it declares
 * `int hh_system_id(void)`, or its functions agree with the built-in surface at 48 points).
 */
#ifndef HAWAII_CONF_H
#define HAWAII_CONF_H

#include "hawaii_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { HH_CALC_NONE = 0, HH_CALC_CORRELATION_SINGLE = 1, HH_CALC_PROCESSING = 2, HH_CALC_PHASE_SPACE_M0 = 3,
       HH_CALC_CORRELATION_ARRAY = 4, HH_CALC_PHASE_SPACE_M2 = 5, HH_CALC_PR_MU = 6 };
enum { HH_MONOMER_NONE = 0, HH_MONOMER_ATOM = 1, HH_MONOMER_LINEAR_MOLECULE = 2, HH_MONOMER_OTHER = 3,
       HH_MONOMER_ROTOR = 4 };
enum { HH_PROC_READ_CF = 0, HH_PROC_READ_SF, HH_PROC_CF_TO_SF, HH_PROC_CMP, HH_PROC_DUP, HH_PROC_DUP2,
       HH_PROC_DROP, HH_PROC_DROP2, HH_PROC_WRITE_CF, HH_PROC_WRITE_SF, HH_PROC_READ_SPECTRUM, HH_PROC_WRITE_SPECTRUM,
       HH_PROC_ALPHA, HH_PROC_D1, HH_PROC_D2, HH_PROC_D3, HH_PROC_D4, HH_PROC_D4A, HH_PROC_AVERAGE_CFS,
       HH_PROC_ADD_SPECTRA, HH_PROC_MN_CLASSICAL, HH_PROC_MN_QUANTUM, HH_PROC_FIT_BASELINE, HH_PROC_SMOOTH,
       HH_PROC_INT3 };

#define HH_CONF_PATH 512
#define HH_CONF_MAX_PROC 128

typedef struct {
    int op;                       /* HH_PROC_* */
    char arg[HH_CONF_PATH];       /* file name for READ_* / WRITE_*; the arguments of the others */
} hh_conf_proc;

typedef struct {
    int calculation_type;         /* HH_CALC_* */
    int pair_state;               /* HH_PAIR_* (FREE_AND_METASTABLE -> HH_PAIR_FREE) */
    double temperature;
    double reduced_mass;
    char so_potential[HH_CONF_PATH];
    char so_dipole[HH_CONF_PATH];
    long niterations;
    long total_trajectories;
    double cvode_tolerance;
    double sampling_time;
    long max_trajectory_length;
    double sampler_Rmin, sampler_Rmax;
    double pesmin;
    long initialM0_npoints;
    char cf_filename[HH_CONF_PATH];
    char sf_filename[HH_CONF_PATH];
    double R0, Rcut;
    double ppf;                   /* partial_partition_function_ratio (0: not given) */
    int use_zimmermann_trick;
    int n_monomers;
    int monomer_type[2];          /* HH_MONOMER_* */
    double II[2][3];
    int nproc;
    hh_conf_proc proc[HH_CONF_MAX_PROC];
    long hep_ppf_niterations, hep_ppf_npoints;   /* VEGAS settings (0: the reference's 12 x 10^6) */
    long hep_m0_niterations, hep_m0_npoints;
    int ntemperatures;                           /* TEMPERATURES = {...} (CALCULATE_PHASE_SPACE_M0) */
    double temperatures[64];
    /* CORRELATION_ARRAY (src/driver.c:1170-1247) */
    int nsatellite;                              /* SATELLITE_TEMPERATURES = {...} */
    double satellite_temperatures[64];
    int nppf_ratios;                             /* PARTIAL_PARTITION_FUNCTION_RATIOS = {...} (0: VEGAS) */
    double ppf_ratios[64];
    int ncf_filenames;                           /* CF_FILENAMES = {"...", ...} */
    char cf_filenames[64][HH_CONF_PATH];
    long initialM2_npoints;                      /* rejection-sampler M2 before the trajectories */
    long hep_m2_niterations, hep_m2_npoints;     /* VEGAS M2 (0: the reference's 12 x 10^6) */
    /* PR_MU (src/driver.c:1180-1190) */
    double approx_frequency_max;                 /* APPROXIMATEFREQUENCYMAX, cm^-1 */
    double average_time_between_collisions;      /* a.t.u.; > 0 (Poisson cut-off) is not provided */
    double spectrum_frequency_max;               /* PROCESSING: WRITE_SPECTRUM truncation, cm^-1 (0: none) */
} hh_conf;

/* Parse a configuration file (no device use).  Unknown keywords and unsupported PROCESSING
 * functions are errors naming the line; keywords of the reference that only steer parts
 * outside this engine (the PR/mu path's histograms and spin weights) are accepted and ignored. */
int hh_conf_parse(const char* path, hh_conf* out);

typedef struct {
    uint32_t seed;                /* 0: from /dev/urandom, as the reference driver */
    int precision;                /* HH_PREC_FP64 (default) or HH_PREC_MIXED */
    double rtol_mixed;            /* rtol = atol of the mixed engine (<= 0: 1e-10) */
    int device;
    int quiet;
} hh_conf_run_opts;

/* Run a parsed configuration on the engine (creates one on opts->device). */
int hh_conf_run(const hh_conf* conf, const hh_conf_run_opts* opts);

#ifdef __cplusplus
}
#endif
#endif
