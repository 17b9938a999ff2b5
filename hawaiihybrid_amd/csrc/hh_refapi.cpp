// The reference's engine-side C API (include/hawaii_ref.h) over the engine, built as
// libhawaii.so next to libhawaii_gpu.so.  Each call maps the reference's MoleculeSystem and
// CalcParams onto the calculations the .conf driver runs (hh_conf.cpp), so a program and a .conf
// file with the same settings and seed give the same files.  The process holds one engine and
// one position in the seeded Mersenne Twister stream, as the reference holds one global stream.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/hawaii_gpu.h"
#include "../../include/hawaii_ref.h"
#include "hh_conf_impl.h"

extern "C" {
dipolePtr dipole_1 = nullptr;
dipolePtr dipole_2 = nullptr;
dipoleFree free_dipole_1 = nullptr;
dipoleFree free_dipole_2 = nullptr;
pesPtr pes = nullptr;
dpesPtr dpes = nullptr;
}

namespace {

constexpr double kZeroCoeff = 0.00361479637 / (4.0 * M_PI);   // src/constants.h:29
constexpr double kSecondCoeff = 13856114.29344114;             // src/constants.h:30

struct Process {
    std::recursive_mutex mu;
    hh_engine* e = nullptr;
    uint32_t seed = 0;   // the stream init_ms seeded
    long draws = 0;      // draws consumed since then
};
Process& proc()
{
    static Process p;
    return p;
}

[[noreturn]] void die(const char* where)
{
    std::fprintf(stderr, "ERROR: %s: %s\n", where, hh_last_error());
    std::exit(1);
}

void destroy_engine()
{
    if (proc().e) hh_engine_destroy(proc().e);
    proc().e = nullptr;
}

hh_engine* engine()
{
    Process& p = proc();
    if (!p.e) {
        const char* d = std::getenv("HAWAII_DEVICE");
        if (hh_engine_create(d ? std::atoi(d) : 0, &p.e) < 0) die("engine");
        std::atexit(destroy_engine);
    }
    return p.e;
}

uint32_t fresh_seed()
{
    uint32_t s = 0;
    if (FILE* f = std::fopen("/dev/urandom", "rb")) {
        if (std::fread(&s, sizeof s, 1, f) != 1) s = 0;
        std::fclose(f);
    }
    return s ? s : (uint32_t)time(nullptr);
}

int monomer_kind(MonomerType t)
{
    switch (t) {
    case ATOM: return HH_MONOMER_ATOM;
    case LINEAR_MOLECULE: return HH_MONOMER_LINEAR_MOLECULE;
    case ROTOR: return HH_MONOMER_ROTOR;
    default: return HH_MONOMER_OTHER;   // the requantised rotors are not part of this engine
    }
}

void copy_path(char* dst, const char* src)
{
    dst[0] = 0;
    if (src) std::snprintf(dst, HH_CONF_PATH, "%s", src);
}

// the .conf settings a MoleculeSystem and CalcParams stand for (src/driver.c:54-1522 maps the
// same keywords onto CalcParams)
hh_conf conf_of(const MoleculeSystem* ms, const CalcParams* p, double T)
{
    hh_conf c;
    std::memset(&c, 0, sizeof c);
    c.pair_state = (int)p->ps;   // PairState values are the HH_PAIR_* values
    c.temperature = T;
    c.reduced_mass = ms->mu;
    c.niterations = (long)p->niterations;
    c.total_trajectories = (long)p->total_trajectories;
    c.cvode_tolerance = p->cvode_tolerance;
    c.sampling_time = p->sampling_time;
    c.max_trajectory_length = (long)p->MaxTrajectoryLength;
    c.sampler_Rmin = p->sampler_Rmin;
    c.sampler_Rmax = p->sampler_Rmax;
    c.pesmin = p->pesmin;
    c.initialM0_npoints = (long)p->initialM0_npoints;
    c.initialM2_npoints = (long)p->initialM2_npoints;
    copy_path(c.cf_filename, p->cf_filename);
    copy_path(c.sf_filename, p->sf_filename);
    c.R0 = p->R0;
    c.Rcut = p->Rcut;
    c.ppf = p->partial_partition_function_ratio;
    c.use_zimmermann_trick = p->use_zimmermann_trick ? 1 : 0;
    c.n_monomers = 2;
    c.monomer_type[0] = monomer_kind(ms->m1.t);
    c.monomer_type[1] = monomer_kind(ms->m2.t);
    for (int k = 0; k < 3; ++k) {
        c.II[0][k] = ms->m1.II[k];
        c.II[1][k] = ms->m2.II[k];
    }
    c.hep_ppf_niterations = (long)p->hep_ppf_niterations;
    c.hep_ppf_npoints = (long)p->hep_ppf_npoints;
    c.hep_m0_niterations = (long)p->hep_m0_niterations;
    c.hep_m0_npoints = (long)p->hep_m0_npoints;
    c.hep_m2_niterations = (long)p->hep_m2_niterations;
    c.hep_m2_npoints = (long)p->hep_m2_npoints;
    c.approx_frequency_max = p->ApproximateFrequencyMax;
    c.average_time_between_collisions = p->average_time_between_collisions;
    const size_t nsat = p->num_satellite_temperatures < 64 ? p->num_satellite_temperatures : 64;
    c.nsatellite = (int)nsat;
    for (size_t k = 0; k < nsat; ++k) {
        if (p->satellite_temperatures) c.satellite_temperatures[k] = p->satellite_temperatures[k];
        if (p->cf_filenames) copy_path(c.cf_filenames[k], p->cf_filenames[k]);
    }
    c.ncf_filenames = p->cf_filenames ? (int)nsat : 0;
    c.nppf_ratios = p->partial_partition_function_ratios ? (int)nsat : 0;
    for (int k = 0; k < c.nppf_ratios; ++k) c.ppf_ratios[k] = p->partial_partition_function_ratios[k];
    return c;
}

// the program's surface functions (null: the built-in ones); one dipole correlated with itself
void surfaces(hh::conf::HostSurfaces& hs)
{
    if (dipole_1 != dipole_2 && dipole_2 != nullptr) {
        std::fprintf(stderr, "ERROR: dipole_1 and dipole_2 differ: the engine correlates one dipole with itself\n");
        std::exit(1);
    }
    hs.pes = pes;
    hs.dpes = dpes;
    hs.dipole = dipole_1;
    hs.pes_name = "pes";
    hs.dipole_name = "dipole_1";
}

hh_conf_run_opts run_opts()
{
    hh_conf_run_opts o{};
    o.seed = proc().seed;
    const char* prec = std::getenv("HAWAII_PRECISION");
    o.precision = (prec && std::strcmp(prec, "mixed") == 0) ? HH_PREC_MIXED : HH_PREC_FP64;
    o.quiet = 0;
    return o;
}

hh_system system_of_ms(const MoleculeSystem* ms)
{
    hh_conf c;
    std::memset(&c, 0, sizeof c);
    c.reduced_mass = ms->mu;
    c.n_monomers = 2;
    c.monomer_type[0] = monomer_kind(ms->m1.t);
    c.monomer_type[1] = monomer_kind(ms->m2.t);
    for (int k = 0; k < 3; ++k) {
        c.II[0][k] = ms->m1.II[k];
        c.II[1][k] = ms->m2.II[k];
    }
    hh_system sys{};
    if (hh::conf::system_of(&c, &sys) < 0) die("init_ms");
    return sys;
}

void init_monomer(Monomer& m, MonomerType t, const double* II)
{
    std::memset(&m, 0, sizeof m);
    m.t = t;
    const int n = t == ATOM ? 0 : (t % 100 == 4 ? 2 : 3);
    for (int k = 0; k < n; ++k) m.II[k] = II[k];
    const int qp = t % 100;
    if (qp > 0) {
        m.qp = static_cast<double*>(std::calloc(qp, sizeof(double)));
        m.dVdq = static_cast<double*>(std::calloc(qp / 2, sizeof(double)));
    }
}

MoleculeSystem make_ms(double mu, const Monomer& m1, const Monomer& m2, size_t seed)
{
    MoleculeSystem ms;
    std::memset(&ms, 0, sizeof ms);
    ms.mu = mu;
    ms.m1 = m1;
    ms.m2 = m2;
    // src/hawaii.c:173-320: Q_SIZE = 3 + the monomers' angles, QP_SIZE = 2 Q_SIZE
    ms.QP_SIZE = 6 + (size_t)(m1.t % 100) + (size_t)(m2.t % 100);
    ms.Q_SIZE = ms.QP_SIZE / 2;
    ms.intermediate_q = static_cast<double*>(std::calloc(ms.Q_SIZE, sizeof(double)));
    ms.dVdq = static_cast<double*>(std::calloc(ms.Q_SIZE, sizeof(double)));
    std::lock_guard<std::recursive_mutex> lock(proc().mu);
    proc().seed = seed > 0 ? (uint32_t)seed : fresh_seed();
    proc().draws = 0;
    ms.seed = proc().seed;
    std::printf("Generator seed is set to %zu\n", ms.seed);
    time(&ms.init_rawtime);
    ms.temp_rawtime = ms.init_rawtime;
    (void)system_of_ms(&ms);   // refuses monomers without an engine here
    return ms;
}

// accepted phase points of the rejection sampler from the stream's current position
std::vector<double> sample_points(const MoleculeSystem* ms, const CalcParams* p, double T, long n, hh_system& sys)
{
    sys = system_of_ms(ms);
    hh::conf::HostSurfaces hs;
    surfaces(hs);
    if (hh::conf::check_surfaces(engine(), sys, p->sampler_Rmin, p->sampler_Rmax, hs) < 0) die("surfaces");
    const hh_sampler smp{HH_IC_REJECTION, (int)p->ps, T, p->sampler_Rmin, p->sampler_Rmax, p->pesmin, 0.0};
    std::vector<double> ic((size_t)n * hh_qp_size(sys.system));
    long pos = 0;
    if (hh_sample_ics_from(&sys, &smp, proc().seed, proc().draws, n, pes, ic.data(), &pos) < 0) die("sampler");
    proc().draws = pos;
    return ic;
}

// Welford mean and standard error of one value per point (calculate_M0's running sums)
void welford(const std::vector<double>& f, double coeff, double* m, double* q)
{
    double mm = 0.0, qq = 0.0;
    const size_t n = f.size();
    for (size_t k = 0; k < n; ++k) {
        const double diff = f[k] - mm;
        mm += diff / (k + 1.0);
        qq += diff * diff * (k / (k + 1.0));
    }
    *m = mm * coeff;
    *q = std::sqrt(qq / n / (n - 1.0)) * coeff;
}

// calculate_M0 / calculate_M2 on the device: the dipoles (|mu|^2) or |dmu/dt|^2 of the accepted
// points, the running mean on the host
void moment(MoleculeSystem* ms, CalcParams* p, double T, int what, double* m, double* q)
{
    std::lock_guard<std::recursive_mutex> lock(proc().mu);
    const long n = (long)(what == HH_EVAL_DIPOLE ? p->initialM0_npoints : p->initialM2_npoints);
    if (n < 2 || std::fabs(p->pesmin) <= 1e-15) {
        std::fprintf(stderr, "ERROR: the moment needs at least 2 points and pesmin\n");
        std::exit(1);
    }
    hh_system sys{};
    const std::vector<double> ic = sample_points(ms, p, T, n, sys);
    const int qp = hh_qp_size(sys.system);
    std::vector<double> soa((size_t)n * qp);
    for (long t = 0; t < n; ++t)
        for (int i = 0; i < qp; ++i) soa[(size_t)i * n + t] = ic[(size_t)t * qp + i];
    const int rows = what == HH_EVAL_DIPOLE ? 3 : 1;
    double *y = nullptr, *out = nullptr;
    if (hipMalloc(&y, soa.size() * sizeof(double)) != hipSuccess ||
        hipMalloc(&out, (size_t)rows * n * sizeof(double)) != hipSuccess)
        die("device allocation");
    std::vector<double> v((size_t)rows * n), f((size_t)n);
    if (hipMemcpy(y, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) die("upload");
    if (hh_eval(engine(), &sys, what, y, n, out) < 0) die("hh_eval");
    if (hipMemcpy(v.data(), out, v.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) die("download");
    (void)hipFree(y);
    (void)hipFree(out);
    for (long k = 0; k < n; ++k)
        f[k] = rows == 3 ? v[k] * v[k] + v[n + k] * v[n + k] + v[2 * n + k] * v[2 * n + k] : v[k];
    welford(f, (rows == 3 ? kZeroCoeff : kSecondCoeff) * p->partial_partition_function_ratio, m, q);
}

double* copy_of(const std::vector<double>& v)
{
    double* p = static_cast<double*>(std::malloc(v.size() * sizeof(double)));
    if (!p) {
        std::fprintf(stderr, "ERROR: out of host memory\n");
        std::exit(1);
    }
    std::memcpy(p, v.data(), v.size() * sizeof(double));
    return p;
}

}  // namespace

extern "C" {

MoleculeSystem init_ms(double mu, MonomerType t1, MonomerType t2, double* I1, double* I2, size_t seed)
{
    if (mu <= 0) {
        std::fprintf(stderr, "ERROR: invalid reduced mass (%.5e) -- must be positive\n", mu);
        std::exit(1);
    }
    static const double none[3] = {0.0, 0.0, 0.0};
    Monomer m1, m2;
    init_monomer(m1, t1, I1 ? I1 : none);
    init_monomer(m2, t2, I2 ? I2 : none);
    return make_ms(mu, m1, m2, seed);
}

MoleculeSystem init_ms_from_monomers(double mu, Monomer* m1, Monomer* m2, size_t seed)
{
    Monomer a, b;
    init_monomer(a, m1->t, m1->II);
    init_monomer(b, m2->t, m2->II);
    return make_ms(mu, a, b, seed);
}

void free_ms(MoleculeSystem* ms)
{
    if (!ms) return;
    std::free(ms->intermediate_q);
    std::free(ms->dVdq);
    std::free(ms->m1.qp);
    std::free(ms->m1.dVdq);
    std::free(ms->m2.qp);
    std::free(ms->m2.dVdq);
    ms->intermediate_q = ms->dVdq = ms->m1.qp = ms->m1.dVdq = ms->m2.qp = ms->m2.dVdq = nullptr;
}

CFnc calculate_correlation_and_save(MoleculeSystem* ms, CalcParams* params, double Temperature)
{
    std::lock_guard<std::recursive_mutex> lock(proc().mu);
    const hh_conf c = conf_of(ms, params, Temperature);
    hh::conf::HostSurfaces hs;
    surfaces(hs);
    hh::conf::Series cf;
    hh::conf::RunIO io;
    io.skip_draws = proc().draws;
    io.cf = &cf;
    if (hh::conf::run_correlation_single(engine(), &c, run_opts(), hs, &io) < 0) die("calculate_correlation_and_save");
    proc().draws = io.draws_used;
    CFnc out{};
    out.t = copy_of(cf.x);
    out.data = copy_of(cf.y);
    out.len = out.capacity = cf.x.size();
    out.ntraj = cf.ntraj;
    out.Temperature = Temperature;
    out.normalized = true;
    return out;
}

CFncArray calculate_correlation_array_and_save(MoleculeSystem* ms, CalcParams* params, double base_temperature)
{
    std::lock_guard<std::recursive_mutex> lock(proc().mu);
    const hh_conf c = conf_of(ms, params, base_temperature);
    hh::conf::HostSurfaces hs;
    surfaces(hs);
    std::vector<hh::conf::Series> cfs;
    hh::conf::RunIO io;
    io.skip_draws = proc().draws;
    io.cfs = &cfs;
    if (hh::conf::run_correlation_array(engine(), &c, run_opts(), hs, &io) < 0)
        die("calculate_correlation_array_and_save");
    proc().draws = io.draws_used;
    CFncArray out{};
    out.ntemp = cfs.size();
    out.len = cfs.empty() ? 0 : cfs[0].x.size();
    out.t = cfs.empty() ? nullptr : copy_of(cfs[0].x);
    out.data = static_cast<double**>(std::malloc(out.ntemp * sizeof(double*)));
    out.nstar = static_cast<double*>(std::malloc(out.ntemp * sizeof(double)));
    for (size_t k = 0; k < out.ntemp; ++k) {
        out.data[k] = copy_of(cfs[k].y);
        out.nstar[k] = cfs[k].ntraj;
    }
    out.ntraj = (size_t)(params->total_trajectories / params->niterations * params->niterations);
    return out;
}

SFnc calculate_spectral_function_using_prmu_representation_and_save(MoleculeSystem* ms, CalcParams* params,
                                                                    double Temperature)
{
    std::lock_guard<std::recursive_mutex> lock(proc().mu);
    const hh_conf c = conf_of(ms, params, Temperature);
    hh::conf::HostSurfaces hs;
    surfaces(hs);
    hh::conf::Series sf;
    hh::conf::RunIO io;
    io.skip_draws = proc().draws;
    io.sf = &sf;
    if (hh::conf::run_prmu(engine(), &c, run_opts(), hs, &io) < 0)
        die("calculate_spectral_function_using_prmu_representation_and_save");
    proc().draws = io.draws_used;
    SFnc out{};
    out.nu = copy_of(sf.x);
    out.data = copy_of(sf.y);
    out.len = out.capacity = sf.x.size();
    out.ntraj = sf.ntraj;
    out.Temperature = Temperature;
    out.normalized = true;
    return out;
}

double analytic_full_partition_function_by_V(MoleculeSystem* ms, double Temperature)
{
    const hh_system sys = system_of_ms(ms);
    const double v = hh_pf_analytic(&sys, Temperature);
    if (v < 0) die("analytic_full_partition_function_by_V");
    return v;
}

bool compute_Mn_from_cf_using_classical_detailed_balance(CFnc cf, size_t n, double* result)
{
    if (n != 0 && n != 2) {
        std::printf("ERROR: estimating spectral moment of order %zu is not implemented\n", n);
        std::exit(1);
    }
    const double dt = cf.len > 1 ? cf.t[1] - cf.t[0] : 0.0;
    if (n == 2 && cf.len < 5) {
        std::printf("ERROR: length of correlation function (%zu) is too short to estimate M2\n", cf.len);
        return false;
    }
    return hh_cf_moment(cf.data, (long)cf.len, dt, cf.normalized ? 0.0 : cf.ntraj, (int)n, result) == HH_OK;
}

double compute_Mn_from_sf_using_classical_detailed_balance(SFnc sf, size_t n)
{
    double r = 0.0;
    if (hh_sf_moment(sf.nu, sf.data, (long)sf.len, sf.normalized ? 0.0 : sf.ntraj, (int)n, &r) < 0)
        die("compute_Mn_from_sf_using_classical_detailed_balance");
    return r;
}

void calculate_M0(MoleculeSystem* ms, CalcParams* params, double Temperature, double* m, double* q)
{
    moment(ms, params, Temperature, HH_EVAL_DIPOLE, m, q);
}

void calculate_M2(MoleculeSystem* ms, CalcParams* params, double Temperature, double* m, double* q)
{
    moment(ms, params, Temperature, HH_EVAL_DIPOLE_RATE, m, q);
}

void mpi_calculate_M0(MoleculeSystem* ms, CalcParams* params, double Temperature, double* m, double* q)
{
    calculate_M0(ms, params, Temperature, m, q);
}

void mpi_calculate_M2(MoleculeSystem* ms, CalcParams* params, double Temperature, double* m, double* q)
{
    calculate_M2(ms, params, Temperature, m, q);
}

void c_mpi_perform_integration(MoleculeSystem* ms, IntegrandType integrand_type, CalcParams* params,
                               double Temperature, size_t niterations, size_t npoints, double* m, double* q)
{
    std::lock_guard<std::recursive_mutex> lock(proc().mu);
    const hh_system sys = system_of_ms(ms);
    hh_vegas_params vp{(int)integrand_type, (int)params->ps, Temperature, params->sampler_Rmin,
                       params->sampler_Rmax, (int)niterations, (long)npoints, (uint32_t)ms->seed};
    hh_vegas_result r{};
    if (hh_vegas(engine(), &sys, &vp, &r) < 0) die("c_mpi_perform_integration");
    *m = r.value;
    *q = r.error;
}

bool write_correlation_function(const char* filename, CFnc cf)
{
    hh::conf::Series s;
    s.x.assign(cf.t, cf.t + cf.len);
    s.y.assign(cf.data, cf.data + cf.len);
    if (!cf.normalized)
        for (double& v : s.y) v /= cf.ntraj;
    s.T = cf.Temperature;
    s.ntraj = cf.ntraj;
    return hh::conf::write_series(filename, s) == HH_OK;
}

bool write_spectral_function(const char* filename, SFnc sf)
{
    hh::conf::Series s;
    s.sf = true;
    s.x.assign(sf.nu, sf.nu + sf.len);
    s.y.assign(sf.data, sf.data + sf.len);
    if (!sf.normalized)
        for (double& v : s.y) v /= sf.ntraj;
    s.T = sf.Temperature;
    s.ntraj = sf.ntraj;
    return hh::conf::write_series(filename, s) == HH_OK;
}

bool read_correlation_function(const char* filename, String_Builder* sb, CFnc* cf)
{
    (void)sb;
    hh::conf::Series s;
    if (hh::conf::read_series(filename, false, s) < 0) return false;
    cf->t = copy_of(s.x);
    cf->data = copy_of(s.y);
    cf->len = cf->capacity = s.x.size();
    cf->ntraj = s.ntraj;
    cf->Temperature = s.T;
    cf->normalized = true;
    return true;
}

bool read_spectral_function(const char* filename, String_Builder* sb, SFnc* sf)
{
    (void)sb;
    hh::conf::Series s;
    if (hh::conf::read_series(filename, true, s) < 0) return false;
    sf->nu = copy_of(s.x);
    sf->data = copy_of(s.y);
    sf->len = sf->capacity = s.x.size();
    sf->ntraj = s.ntraj;
    sf->Temperature = s.T;
    sf->normalized = true;
    return true;
}

SFnc idct_cf_to_sf(CFnc cf)
{
    std::lock_guard<std::recursive_mutex> lock(proc().mu);
    const long n = (long)cf.len;
    const double dt = n > 1 ? cf.t[1] - cf.t[0] : 0.0;
    double *d = nullptr, *nu = nullptr, *s = nullptr;
    if (hipMalloc(&d, n * sizeof(double)) != hipSuccess || hipMalloc(&nu, n * sizeof(double)) != hipSuccess ||
        hipMalloc(&s, n * sizeof(double)) != hipSuccess)
        die("idct_cf_to_sf: device allocation");
    if (hipMemcpy(d, cf.data, n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) die("idct_cf_to_sf: upload");
    if (hh_cf_to_sf(engine(), d, n, dt, nu, s, nullptr) < 0) die("idct_cf_to_sf");
    SFnc sf{};
    sf.nu = static_cast<double*>(std::malloc(n * sizeof(double)));
    sf.data = static_cast<double*>(std::malloc(n * sizeof(double)));
    // the CF is Connes-apodised in place, as the reference's transform leaves it
    if (hipMemcpy(cf.data, d, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(sf.nu, nu, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(sf.data, s, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        die("idct_cf_to_sf: download");
    (void)hipFree(d);
    (void)hipFree(nu);
    (void)hipFree(s);
    sf.len = sf.capacity = cf.len;
    sf.Temperature = cf.Temperature;
    sf.ntraj = cf.ntraj;
    sf.normalized = cf.normalized;
    return sf;
}

void free_cfnc(CFnc cf)
{
    std::free(cf.t);
    std::free(cf.data);
}

void free_sfnc(SFnc sf)
{
    std::free(sf.nu);
    std::free(sf.data);
}

void free_cfnc_array(CFncArray ca)
{
    std::free(ca.t);
    for (size_t k = 0; k < ca.ntemp; ++k) std::free(ca.data[k]);
    std::free(ca.data);
    std::free(ca.nstar);
}

}  // extern "C"
