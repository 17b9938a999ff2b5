/* A program written the way the reference's examples/correlation_co2_ar.cpp:21-51 is, linked
 * against libhawaii.so instead of the reference's library: init_ms, the pes / dipole function
 * pointers, CalcParams, calculate_correlation_and_save, idct_cf_to_sf, the CF moments and a VEGAS
 * integral through c_mpi_perform_integration.
 *
 *   correlation_co2_ar MU I SEED L TOTAL NITER M0POINTS CF_FILE SF_FILE
 */
#include <stdio.h>
#include <stdlib.h>

#include "hawaii_gpu.h"
#include "hawaii_ref.h"

int main(int argc, char* argv[])
{
    if (argc != 10) {
        fprintf(stderr, "usage: %s MU I SEED L TOTAL NITER M0POINTS CF_FILE SF_FILE\n", argv[0]);
        return 2;
    }
    const double MU = atof(argv[1]);
    double I1[2] = {atof(argv[2]), atof(argv[2])};
    MoleculeSystem ms = init_ms(MU, LINEAR_MOLECULE, ATOM, I1, NULL, (size_t)atol(argv[3]));

    pes = hh_host_pes(HH_SYS_CO2AR);   /* the plugin's pes_lab in the reference's example */
    dipole_1 = NULL;                   /* NULL: the built-in CO2-Ar dipole */
    dipole_2 = NULL;

    CalcParams params = {0};
    params.ps = PAIR_STATE_FREE_AND_METASTABLE;
    params.sampler_Rmin = 4.0;
    params.sampler_Rmax = 40.0;
    params.niterations = (size_t)atol(argv[6]);
    params.total_trajectories = (size_t)atol(argv[5]);
    params.cvode_tolerance = 1e-12;
    params.sampling_time = 200.0;
    params.MaxTrajectoryLength = (size_t)atol(argv[4]);
    params.Rcut = 40.0;
    params.partial_partition_function_ratio = 2.68854e+05;
    params.initialM0_npoints = (size_t)atol(argv[7]);
    params.initialM2_npoints = 0;
    params.pesmin = -195.6337098547 / 2.1947463136320e5;
    params.cf_filename = argv[8];

    const double Temperature = 300.0;
    CFnc cf = calculate_correlation_and_save(&ms, &params, Temperature);
    printf("CFNC len %zu ntraj %.1f normalized %d T %.2f\n", cf.len, cf.ntraj, (int)cf.normalized, cf.Temperature);

    double m0 = 0.0, m2 = 0.0;
    if (!compute_Mn_from_cf_using_classical_detailed_balance(cf, 0, &m0) ||
        !compute_Mn_from_cf_using_classical_detailed_balance(cf, 2, &m2))
        return 1;
    printf("CF M0 %.12e M2 %.12e\n", m0, m2);

    SFnc sf = idct_cf_to_sf(cf);
    if (!write_spectral_function(argv[9], sf)) return 1;

    double pf = 0.0, pf_err = 0.0;
    c_mpi_perform_integration(&ms, INTEGRAND_PF, &params, Temperature, 4, 100000, &pf, &pf_err);
    printf("PF %.12e +- %.12e analytic %.12e\n", pf, pf_err, analytic_full_partition_function_by_V(&ms, Temperature));

    free_sfnc(sf);
    free_cfnc(cf);
    free_ms(&ms);
    return 0;
}
