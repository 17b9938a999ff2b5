/*
 * hawaii_ref.h -- the reference's engine-side C API (src/hawaii.h, src/hep_hawaii.h) on the
 * MI355X engine, in libhawaii.so.
 *
 * A program written against the reference's library (examples/correlation_co2_ar.cpp:21-51:
 * init_ms, the pes / dipole_1 function pointers, a CalcParams, calculate_correlation_and_save,
 * free_cfnc) relinks against this one.  The types keep the reference's names, member order and
 * C types, so a caller's structs have the reference's layout; the histogram and file members
 * that only the reference's requantised-rotor path uses are opaque pointers here.
 *
 * Each entry point runs on the process's engine (device 0, or $HAWAII_DEVICE), created at the
 * first call.  The seeded Mersenne Twister stream is global, as the reference's: init_ms seeds
 * it (seed 0: a fresh seed) and every later sampling call continues where the previous one
 * stopped.  The device integrates its built-in surfaces for the monomers given to init_ms
 * (ATOM + ATOM: He-Ar, LINEAR_MOLECULE + ATOM: CO2-Ar, ROTOR + LINEAR_MOLECULE: CH4-CO2); the
 * functions in `pes`, `dpes`, `dipole_1` must be those surfaces (checked at 48 points on first
 * use, or declared by an `int hh_system_id(void)` of their plugin) and `pes` drives the
 * host rejection sampler as in the reference.  NULL selects the built-in surface.
 * Failures print the reason and exit(1), as the reference's asserts and exits do.
 */
#ifndef HAWAII_REF_H
#define HAWAII_REF_H

#include <stdbool.h>
#include <stddef.h>
#include <stdio.h>
#include <time.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/hawaii.h:183-190: phase-point size of a monomer = type % 100 */
typedef enum {
    ATOM = 0,
    LINEAR_MOLECULE = 4,
    LINEAR_MOLECULE_REQ_INTEGER = 104,
    LINEAR_MOLECULE_REQ_HALFINTEGER = 204,
    ROTOR = 6,
    ROTOR_REQUANTIZED_ROTATION = 106,
} MonomerType;

typedef struct hh_ref_histogram hh_ref_histogram;   /* the reference's gsl_histogram, unused here */

/* src/hawaii.h:210-245 */
typedef struct {
    MonomerType t;
    double II[3];
    double DJ;
    double* qp;
    double* dVdq;
    bool apply_requantization;
    size_t req_switch_counter;
    size_t torque_cache_len;
    double torque_limit;
    double* torque_cache;
    size_t nswitch_histogram_bins;
    double nswitch_histogram_max;
    const char* nswitch_histogram_filename;
    hh_ref_histogram* nswitch_histogram;
    FILE* fp_nswitch_histogram;
    double initial_j;
    size_t jini_histogram_bins;
    double jini_histogram_max;
    const char* jini_histogram_filename;
    hh_ref_histogram* jini_histogram;
    FILE* fp_jini_histogram;
    size_t jfin_histogram_bins;
    double jfin_histogram_max;
    const char* jfin_histogram_filename;
    hh_ref_histogram* jfin_histogram;
    FILE* fp_jfin_histogram;
} Monomer;

/* src/hawaii.h:267-284 */
typedef struct {
    double intermolecular_qp[6];
    Monomer m1;
    Monomer m2;
    double mu;
    size_t Q_SIZE;
    size_t QP_SIZE;
    double* intermediate_q;
    double* dVdq;
    size_t seed;
    time_t init_rawtime;
    time_t temp_rawtime;
} MoleculeSystem;

/* src/hawaii.h:291-297 */
typedef enum {
    PAIR_STATE_NONE,
    PAIR_STATE_FREE_AND_METASTABLE,
    PAIR_STATE_BOUND,
    PAIR_STATE_ALL,
    PAIR_STATE_COUNT,
} PairState;

/* src/hawaii.h:301-310 */
typedef enum {
    CALCULATION_NONE,
    CALCULATION_PR_MU,
    CALCULATION_CORRELATION_SINGLE,
    CALCULATION_CORRELATION_ARRAY,
    CALCULATION_PROCESSING,
    CALCULATION_PHASE_SPACE_M0,
    CALCULATION_PHASE_SPACE_M2,
    CALCULATION_TYPES_COUNT,
} CalculationType;

/* src/hawaii.h:314-365 */
typedef struct {
    PairState ps;
    CalculationType calculation_type;
    double sampler_Rmin;
    double sampler_Rmax;
    double pesmin;
    size_t initialM0_npoints;
    size_t initialM2_npoints;
    double partial_partition_function_ratio;
    size_t hep_m0_niterations;
    size_t hep_m0_npoints;
    size_t hep_m2_niterations;
    size_t hep_m2_npoints;
    size_t hep_ppf_niterations;
    size_t hep_ppf_npoints;
    double odd_j_spin_weight;
    double even_j_spin_weight;
    double sampling_time;
    size_t MaxTrajectoryLength;
    bool allow_truncating_trajectories_at_length_limit;
    double cvode_tolerance;
    size_t niterations;
    size_t total_trajectories;
    const char* cf_filename;
    double Rcut;
    bool use_zimmermann_trick;
    const char* sf_filename;
    double ApproximateFrequencyMax;
    double R0;
    double average_time_between_collisions;
    double* partial_partition_function_ratios;
    double* satellite_temperatures;
    size_t num_satellite_temperatures;
    const char** cf_filenames;
} CalcParams;

/* src/hawaii.h:375-383 */
typedef struct {
    double* t;
    double* data;
    size_t len;
    size_t capacity;
    double ntraj;
    double Temperature;
    bool normalized;
} CFnc;

/* src/hawaii.h:397-404 */
typedef struct {
    double* t;
    double** data;
    size_t ntemp;
    size_t len;
    double* nstar;
    size_t ntraj;
} CFncArray;

/* src/hawaii.h:409-417 */
typedef struct {
    double* nu;
    double* data;
    size_t len;
    size_t capacity;
    double ntraj;
    double Temperature;
    bool normalized;
} SFnc;

/* src/hawaii.h:594-598 */
typedef struct {
    char* items;
    size_t count;
    size_t capacity;
} String_Builder;

/* src/hep_hawaii.h:10-14 */
typedef enum {
    INTEGRAND_M0,
    INTEGRAND_M2,
    INTEGRAND_PF,
} IntegrandType;

/* the plugin function pointers (src/hawaii.h:471-499) */
typedef void (*dipolePtr)(double*, double[3]);
typedef void (*dipoleFree)(void);
typedef double (*pesPtr)(double*);
typedef void (*dpesPtr)(double*, double*);
extern dipolePtr dipole_1;
extern dipolePtr dipole_2;
extern dipoleFree free_dipole_1;
extern dipoleFree free_dipole_2;
extern pesPtr pes;
extern dpesPtr dpes;

/* src/hawaii.h:502-504 (init_ms, src/hawaii.c:173-320; init_ms_from_monomers, src/hawaii.c:74-171) */
MoleculeSystem init_ms(double mu, MonomerType t1, MonomerType t2, double* I1, double* I2, size_t seed);
MoleculeSystem init_ms_from_monomers(double mu, Monomer* m1, Monomer* m2, size_t seed);
void free_ms(MoleculeSystem* ms);

/* src/hawaii.h:507-509 (src/hawaii.c:2795-3117, 3148-3977, 2361-2793).  The CF / SF files are
 * written after every iteration in the reference's format; the returned CF / SF is normalised
 * by the trajectory count. */
CFnc calculate_correlation_and_save(MoleculeSystem* ms, CalcParams* params, double Temperature);
SFnc calculate_spectral_function_using_prmu_representation_and_save(MoleculeSystem* ms, CalcParams* params,
                                                                    double Temperature);
CFncArray calculate_correlation_array_and_save(MoleculeSystem* ms, CalcParams* params, double base_temperature);

/* src/hawaii.h:552 (src/hawaii.c:5071-5090) */
double analytic_full_partition_function_by_V(MoleculeSystem* ms, double Temperature);

/* src/hawaii.h:558-567: compute_Mn_from_cf_using_classical_detailed_balance (src/hawaii.c:5945-5979),
 * compute_Mn_from_sf_using_classical_detailed_balance (src/hawaii.c:5981-6013), the rejection
 * sampler's M0 and M2 over params->initialM0_npoints / initialM2_npoints points of the stream
 * (calculate_M0, src/hawaii.c:1423-1491; calculate_M2, src/hawaii.c:1493-1617; the mpi_ forms are
 * the same on one process). */
bool compute_Mn_from_cf_using_classical_detailed_balance(CFnc cf, size_t n, double* result);
double compute_Mn_from_sf_using_classical_detailed_balance(SFnc sf, size_t n);
void calculate_M0(MoleculeSystem* ms, CalcParams* params, double Temperature, double* m, double* q);
void calculate_M2(MoleculeSystem* ms, CalcParams* params, double Temperature, double* m, double* q);
void mpi_calculate_M0(MoleculeSystem* ms, CalcParams* params, double Temperature, double* m, double* q);
void mpi_calculate_M2(MoleculeSystem* ms, CalcParams* params, double Temperature, double* m, double* q);

/* src/hep_hawaii.h:16 (src/hep_hawaii.cpp:8-380): the VEGAS integral and its error, unscaled */
void c_mpi_perform_integration(MoleculeSystem* ms, IntegrandType integrand_type, CalcParams* params,
                               double Temperature, size_t niterations, size_t npoints, double* m, double* q);

/* src/hawaii.h:617-648: files and transforms */
bool write_correlation_function(const char* filename, CFnc cf);
bool write_spectral_function(const char* filename, SFnc sf);
bool read_correlation_function(const char* filename, String_Builder* sb, CFnc* cf);
bool read_spectral_function(const char* filename, String_Builder* sb, SFnc* sf);
SFnc idct_cf_to_sf(CFnc cf);
void free_cfnc(CFnc cf);
void free_sfnc(SFnc sf);
void free_cfnc_array(CFncArray ca);


#ifdef __cplusplus
}
#endif
#endif
