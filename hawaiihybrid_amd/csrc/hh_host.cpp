// Host side of the MI355X trajectory engine: device set-up, surface-table upload, work-item
// bookkeeping, launches, reductions and the C ABI of include/hawaii_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "co2ar_tables.h"
#include "hh_internal.h"

extern "C" __attribute__((visibility("hidden"))) long hh_sample_ics_impl(const hh_system* sys, const hh_sampler* smp,
                                                                        uint32_t seed, long skip, long ntraj,
                                                                        hh_pes_fn pes, double* ic, long* n_draws);

namespace hh {
namespace {
thread_local std::string g_err;
}

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

const char* last_error() { return g_err.c_str(); }
}  // namespace hh

using hh::DevBuf;
using hh::fail;

namespace {

template <typename P>
void fill_tables(hh::Co2ArTables<P>& t)
{
    for (int b = 0; b < 32; ++b) {
        t.center[b] = HH_CO2AR_CENTER[b];
        for (int s = 0; s < 6; ++s)
            for (int c = 0; c < 8; ++c) t.pes[b][s][c] = (P)HH_CO2AR_PES[b][s][c];
        for (int s = 0; s < 5; ++s)
            for (int c = 0; c < 8; ++c) t.dipx[b][s][c] = (P)HH_CO2AR_DIPX[b][s][c];
        for (int s = 0; s < 6; ++s)
            for (int c = 0; c < 8; ++c) t.dipz[b][s][c] = (P)HH_CO2AR_DIPZ[b][s][c];
    }
    for (int s = 0; s < 6; ++s) {
        t.pes_exp[s][0] = (P)HH_CO2AR_PES_EXP[s][0];
        t.pes_exp[s][1] = (P)HH_CO2AR_PES_EXP[s][1];
    }
}

}  // namespace

namespace hh {

// per-item summary of jobs into stats
int collect_stats(const std::vector<hh::JobDesc>& jobs, hh_cf_stats* stats)
{
    for (const hh::JobDesc& jb : jobs) {
        const size_t n = (size_t)jb.n_items;
        std::vector<int> ln(n), st(n), nr(n);
        std::vector<float> dr(n);
        HH_CHECK(hipMemcpy(ln.data(), jb.len, n * sizeof(int), hipMemcpyDeviceToHost));
        HH_CHECK(hipMemcpy(st.data(), jb.status, n * sizeof(int), hipMemcpyDeviceToHost));
        HH_CHECK(hipMemcpy(nr.data(), jb.nrhs, n * sizeof(int), hipMemcpyDeviceToHost));
        HH_CHECK(hipMemcpy(dr.data(), jb.e_drift, n * sizeof(float), hipMemcpyDeviceToHost));
        stats->n_items += (long long)n;
        if (jb.orbit_count) {
            int no = 0;
            unsigned long long nr64 = 0;
            HH_CHECK(hipMemcpy(&no, jb.orbit_count, sizeof no, hipMemcpyDeviceToHost));
            HH_CHECK(hipMemcpy(&nr64, jb.orbit_nrhs, sizeof nr64, hipMemcpyDeviceToHost));
            stats->n_orbit += no;
            stats->nrhs_orbit += (long long)nr64;
            if (no > 0) {
                std::vector<int> items((size_t)no);
                HH_CHECK(hipMemcpy(items.data(), jb.orbit_items, items.size() * sizeof(int), hipMemcpyDeviceToHost));
                for (int it : items) stats->nsamples_orbit += ln[(size_t)it];
            }
        }
        for (size_t i = 0; i < n; ++i) {
            stats->nrhs += nr[i];
            stats->nsamples += ln[i];
            if (st[i] >= 1 && st[i] <= 3) ++stats->n_failed;
            if (st[i] == hh::ITEM_FAIL_STEPS || st[i] == hh::ITEM_FAIL_STEPSIZE) ++stats->n_stepfail;
            stats->max_e_drift = std::max(stats->max_e_drift, (double)dr[i]);
        }
    }
    return HH_OK;
}

int check_cf_args(const hh_system* sys, const hh_cf_params* p)
{
    if (sys->system != HH_SYS_HEAR && sys->system != HH_SYS_CO2AR && sys->system != HH_SYS_CH4CO2)
        return fail(HH_E_UNSUPPORTED, "system %d has no trajectory engine", sys->system);
    if (sys->system == HH_SYS_CH4CO2) {
        if (p->engine != HH_ENGINE_BATCHED) return fail(HH_E_UNSUPPORTED, "CH4-CO2 runs on the batched engine");
        if (p->zimmermann) return fail(HH_E_UNSUPPORTED, "CH4-CO2: Zimmermann's trick is not provided");
        if (!(sys->I > 0) || !(sys->I_rotor[0] > 0) || !(sys->I_rotor[1] > 0) || !(sys->I_rotor[2] > 0))
            return fail(HH_E_INVALID, "CH4-CO2 needs the CO2 moment of inertia I and the three of CH4 (I_rotor)");
    }
    if (sys->mu <= 0) return fail(HH_E_INVALID, "invalid reduced mass (%.5e) -- must be positive", sys->mu);
    if (p->L <= 0 || p->dt <= 0 || p->Rcut <= 0 || p->rtol <= 0 || p->atol <= 0)
        return fail(HH_E_INVALID, "MaxTrajectoryLength, sampling_time, Rcut and tolerances must be positive");
    if (p->engine != HH_ENGINE_PLAIN && p->engine != HH_ENGINE_BATCHED)
        return fail(HH_E_INVALID, "unknown engine %d", p->engine);
    if (p->precision != HH_PREC_FP64 && p->precision != HH_PREC_MIXED)
        return fail(HH_E_INVALID, "unknown precision %d", p->precision);
    if (p->engine == HH_ENGINE_PLAIN && p->precision != HH_PREC_FP64)
        return fail(HH_E_INVALID, "the plain engine is FP64 only");
    if (p->halves != HH_HALVES_BOTH && p->halves != HH_HALVES_FORWARD)
        return fail(HH_E_INVALID, "unknown halves %d", p->halves);
    if (p->zimmermann && p->halves != HH_HALVES_BOTH)
        return fail(HH_E_INVALID, "Zimmermann's trick integrates both halves of every trajectory");
    if (!(p->orbit_rtol <= 0.0 || p->orbit_rtol < 1.0))
        return fail(HH_E_INVALID, "orbit_rtol must be below 1 (0: 1e-12, negative: off)");
    if (p->zimmermann && p->L > zimm_max_length())
        return fail(HH_E_UNSUPPORTED, "Zimmermann's trick: MaxTrajectoryLength %d above %d", p->L, zimm_max_length());
    return HH_OK;
}

// The ensemble CF of a finished job: the per-trajectory rows summed in a fixed order, or for the
// Zimmermann series the FFT cross-correlation over every time origin.
int reduce_job(hh_engine* e, const JobDesc& jd, double ppf, double* cf_out)
{
    if (jd.record == 2) return HH_OK;   // PR/mu series: the spectrum is formed by its caller
    if (jd.record == 3) {
        // CH4-CO2: dipoles of the recorded samples and the CF rows, then the usual reduction
        HH_CHECK(run_ch4co2_cf_rows(jd, jd.f32_dipole ? e->ch4tab32.p : e->ch4tab64.p, jd.f32_dipole != 0, e->stream));
        JobDesc jr = jd;
        jr.rows = jd.cf_rows;
        jr.record = 0;
        return reduce_job(e, jr, ppf, cf_out);
    }
    if (jd.record) {
        const size_t ws = zimm_workspace_doubles(jd.L);
        HH_CHECK(e->psum.ensure(ws * sizeof(double)));
        HH_CHECK(e->pcomp.ensure(ws * sizeof(double)));
        const int logP = zimm_log_points(jd.L);
        if (e->zimm_logP != logP) {
            const long P = 1L << logP;
            std::vector<double> tw((size_t)P);   // e^{+2 pi i j/P}, j < P/2, as (re, im) pairs
            for (long j = 0; j < P / 2; ++j) {
                const long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)j / (long double)P;
                tw[2 * j] = (double)cosl(a);
                tw[2 * j + 1] = (double)sinl(a);
            }
            HH_CHECK(e->zimm_tw.ensure(tw.size() * sizeof(double)));
            HH_CHECK(hipMemcpyAsync(e->zimm_tw.p, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice,
                                    e->stream));
            HH_CHECK(hipStreamSynchronize(e->stream));
            e->zimm_logP = logP;
        }
        HH_CHECK(run_zimm_reduce(jd, ppf, e->zimm_tw.as<double2>(), e->psum.as<double>(), e->pcomp.as<double>(),
                                 cf_out, e->stream));
        return HH_OK;
    }
    const long long chunk = 256;
    const int nchunks = (int)((jd.ntraj + chunk - 1) / chunk);
    HH_CHECK(e->psum.ensure((size_t)nchunks * jd.L * sizeof(double)));
    HH_CHECK(e->pcomp.ensure((size_t)nchunks * jd.L * sizeof(double)));
    HH_CHECK(run_reduce(jd, ppf, e->psum.as<double>(), e->pcomp.as<double>(), chunk, nchunks, cf_out, e->stream));
    return HH_OK;
}

}  // namespace hh

namespace hh {

const int kZeroIntStorage = 0;

int m0_check(const hh_system* sys, const hh_sampler* smp, long npoints, int nstreams)
{
    if (sys->system != HH_SYS_HEAR && sys->system != HH_SYS_CO2AR)
        return fail(HH_E_UNSUPPORTED, "unknown system %d", sys->system);
    if (npoints < 2 || nstreams <= 0 || npoints < nstreams * 2L)
        return fail(HH_E_INVALID, "M0: need npoints >= 2 * nstreams > 0");
    if (smp->mode != HH_IC_REJECTION) return fail(HH_E_INVALID, "M0 uses the rejection sampler");
    if (!(smp->pesmin < 0 || smp->pesmin > 0)) return fail(HH_E_INVALID, "M0: pesmin must be non-zero");
    if (!(smp->Rmin > 0 && smp->Rmax > smp->Rmin)) return fail(HH_E_INVALID, "M0: need 0 < Rmin < Rmax");
    if (smp->T <= 0) return fail(HH_E_INVALID, "M0: temperature must be positive");
    return HH_OK;
}

// [624 x ns] u32 MT states, then m, q (double), n, cand (long), pos (int) per stream, finished
size_t m0_bytes(long ns)
{
    return (size_t)ns * (624 * sizeof(uint32_t) + 2 * sizeof(double) + 2 * sizeof(long) + sizeof(int)) + sizeof(int);
}

M0Acc m0_view(void* base, long ns)
{
    M0Acc a;
    char* p = static_cast<char*>(base) + (size_t)ns * 624 * sizeof(uint32_t);
    a.m = reinterpret_cast<double*>(p);
    a.q = a.m + ns;
    a.n = reinterpret_cast<long*>(a.q + ns);
    a.cand = a.n + ns;
    a.pos = reinterpret_cast<int*>(a.cand + ns);
    a.finished = a.pos + ns;
    return a;
}

int m0_combine(const M0Acc& acc, long npoints, long nstreams, double ppf, double* m, double* q, int order)
{
    std::vector<double> hm(nstreams), hq(nstreams);
    std::vector<long> hn(nstreams);
    HH_CHECK(hipMemcpy(hm.data(), acc.m, nstreams * sizeof(double), hipMemcpyDeviceToHost));
    HH_CHECK(hipMemcpy(hq.data(), acc.q, nstreams * sizeof(double), hipMemcpyDeviceToHost));
    HH_CHECK(hipMemcpy(hn.data(), acc.n, nstreams * sizeof(long), hipMemcpyDeviceToHost));
    for (long s = 0; s < nstreams; ++s) {
        const long quota = npoints / nstreams + (s < npoints % nstreams ? 1 : 0);
        if (hn[s] != quota)
            return fail(HH_E_INVALID, "M%d: stream %ld accepted %ld of %ld points within the candidate limit "
                                      "(check pesmin, T and the pair state)", order, s, hn[s], quota);
    }
    // combine as mpi_calculate_M0 / mpi_calculate_M2 do across ranks (src/hawaii.c:1722-1733,
    // 1850-1854): mean of the per-stream means and of the per-stream q, q normalised by the
    // per-stream count
    double sm = 0.0, sq = 0.0;
    for (long s = 0; s < nstreams; ++s) {
        sm += hm[s];
        sq += hq[s];
    }
    const double ZeroCoeff = 0.00361479637 / (4.0 * M_PI);   // src/constants.h:29
    const double SecondCoeff = 13856114.29344114;            // src/constants.h:30
    const double coeff = order == 2 ? SecondCoeff : ZeroCoeff;
    const double n0 = (double)hn[0];
    *m = (sm / nstreams) * coeff * ppf;
    *q = std::sqrt((sq / nstreams) / n0 / (n0 - 1)) * coeff * ppf;
    return HH_OK;
}

}  // namespace hh

namespace {
const int& kZeroInt = hh::kZeroIntStorage;
int moment_run(hh_engine* e, const hh_system* sys, const hh_sampler* smp, uint32_t seed, long npoints, int nstreams,
               double ppf, int order, double* m, double* q, double* ms);
}

extern "C" {

const char* hh_last_error(void) { return hh::last_error(); }

int hh_qp_size(int system)
{
    switch (system) {
    case HH_SYS_HEAR: return 6;
    case HH_SYS_CO2AR:
    case HH_SYS_N2AR: return 10;
    case HH_SYS_CH4CO2: return 16;
    default: return -1;
    }
}

long hh_sample_ics(const hh_system* sys, const hh_sampler* smp, uint32_t seed, long ntraj, hh_pes_fn pes,
                   double* ic, long* n_draws)
{
    return hh_sample_ics_from(sys, smp, seed, 0, ntraj, pes, ic, n_draws);
}

long hh_sample_ics_from(const hh_system* sys, const hh_sampler* smp, uint32_t seed, long skip_draws, long ntraj,
                        hh_pes_fn pes, double* ic, long* n_draws)
{
    if (!sys || !smp || !ic || ntraj < 0) return fail(HH_E_INVALID, "hh_sample_ics: null argument or ntraj < 0");
    if (skip_draws < 0) return fail(HH_E_INVALID, "hh_sample_ics_from: skip_draws < 0");
    if (sys->system != HH_SYS_HEAR && sys->system != HH_SYS_CO2AR && sys->system != HH_SYS_N2AR &&
        sys->system != HH_SYS_CH4CO2)
        return fail(HH_E_UNSUPPORTED, "hh_sample_ics: system %d has no sampler", sys->system);
    if (sys->system == HH_SYS_CH4CO2 &&
        !(sys->I > 0 && sys->I_rotor[0] > 0 && sys->I_rotor[1] > 0 && sys->I_rotor[2] > 0))
        return fail(HH_E_INVALID, "CH4-CO2 needs the CO2 moment of inertia I and the three of CH4 (I_rotor)");
    if (smp->mode != HH_IC_REJECTION && smp->mode != HH_IC_R0_INJECT && smp->mode != HH_IC_PHASE_SPACE)
        return fail(HH_E_INVALID, "hh_sample_ics: unknown mode %d", smp->mode);
    if (sys->system == HH_SYS_N2AR && smp->mode == HH_IC_REJECTION && !pes)
        return fail(HH_E_UNSUPPORTED, "hh_sample_ics: N2-Ar rejection sampling needs the plugin's pes_lab");
    if (sys->mu <= 0) return fail(HH_E_INVALID, "invalid reduced mass (%.5e) -- must be positive", sys->mu);
    if (smp->T <= 0) return fail(HH_E_INVALID, "temperature must be positive");
    if (smp->mode != HH_IC_R0_INJECT && !(smp->Rmin > 0 && smp->Rmax > smp->Rmin))
        return fail(HH_E_INVALID, "sampler_Rmin/sampler_Rmax must satisfy 0 < Rmin < Rmax");
    if (smp->mode == HH_IC_REJECTION && !(smp->pesmin < 0 || smp->pesmin > 0))
        return fail(HH_E_INVALID, "pesmin must be non-zero in rejection mode");
    return hh_sample_ics_impl(sys, smp, seed, skip_draws, ntraj, pes, ic, n_draws);
}

int hh_engine_create(int device, hh_engine** out)
{
    if (!out) return fail(HH_E_INVALID, "hh_engine_create: out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(HH_E_DEVICE, "no HIP device is visible");
    if (device < 0 || device >= n) return fail(HH_E_INVALID, "device %d out of range (%d visible)", device, n);
    HH_CHECK(hipSetDevice(device));
    hh_engine* e = new hh_engine();
    e->device = device;
    if (hipGetDeviceProperties(&e->prop, device) != hipSuccess) {
        delete e;
        return fail(HH_E_DEVICE, "hipGetDeviceProperties failed");
    }
    if (std::strncmp(e->prop.gcnArchName, "gfx950", 6) != 0) {
        std::string arch = e->prop.gcnArchName;
        delete e;
        return fail(HH_E_DEVICE, "device %d is %s; this build targets gfx950 (MI355X)", device, arch.c_str());
    }
    // graph-captured slices are opt-in (HH_GRAPHS=1)
    if (const char* g = std::getenv("HH_GRAPHS")) e->use_graphs = g[0] && g[0] != '0';
    // a blocking stream: the engine's work is ordered after the caller's copies and kernels on
    // the legacy default stream (hipMemcpy of pageable memory can return before its DMA ends)
    if (hipStreamCreateWithFlags(&e->stream, hipStreamDefault) != hipSuccess) {
        delete e;
        return fail(HH_E_DEVICE, "hipStreamCreate failed");
    }
    for (auto& ev : e->ev) (void)hipEventCreate(&ev);
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
            e->wall_khz = khz;
    }
    {
        static hh::Co2ArTables<double> t64;
        static hh::Co2ArTables<float> t32;
        fill_tables(t64);
        fill_tables(t32);
        if (e->tab64.ensure(sizeof t64) != hipSuccess || e->tab32.ensure(sizeof t32) != hipSuccess) {
            hh_engine_destroy(e);
            return fail(HH_E_NOMEM, "table allocation failed");
        }
        if (hipMemcpy(e->tab64.p, &t64, sizeof t64, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(e->tab32.p, &t32, sizeof t32, hipMemcpyHostToDevice) != hipSuccess) {
            hh_engine_destroy(e);
            return fail(HH_E_DEVICE, "table upload failed");
        }
    }
    {
        // CH4-CO2 dipole tables
        for (int f32 = 0; f32 < 2; ++f32) {
            const size_t nb = hh::ch4co2_table_bytes(f32);
            std::vector<char> host(nb);
            hh::ch4co2_tables_host(f32, host.data());
            hh::DevBuf& d = f32 ? e->ch4tab32 : e->ch4tab64;
            if (d.ensure(nb) != hipSuccess || hipMemcpy(d.p, host.data(), nb, hipMemcpyHostToDevice) != hipSuccess) {
                hh_engine_destroy(e);
                return fail(HH_E_NOMEM, "CH4-CO2 table upload failed");
            }
        }
    }
    // the device address of hh_cf_to_sf_noalloc's static work array on this device
    if (hh::noalloc_scratch_address(&e->noalloc_scratch) != hipSuccess || !e->noalloc_scratch) {
        hh_engine_destroy(e);
        return fail(HH_E_DEVICE, "hipGetSymbolAddress of the CF -> SF work array failed");
    }
    *out = e;
    return HH_OK;
}

void hh_engine_destroy(hh_engine* e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    e->last_pool_jobs.clear();
    e->pools.clear();
    for (DevBuf* b : {&e->tab64, &e->tab32, &e->ch4tab64, &e->ch4tab32, &e->rows, &e->len, &e->status, &e->nrhs, &e->drift, &e->jobs, &e->queue,
                      &e->psum, &e->pcomp, &e->fft_a, &e->fft_b, &e->m0_state, &e->arena})
        b->release();
    for (auto& ev : e->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    if (e->capture) (void)hipStreamDestroy(e->capture);
    delete e;
}

int hh_engine_info(hh_engine* e, char* buf, size_t len)
{
    if (!e || !buf) return fail(HH_E_INVALID, "hh_engine_info: null argument");
    snprintf(buf, len, "{\"device\": %d, \"arch\": \"%s\", \"cus\": %d, \"clock_khz\": %d, \"hbm_bytes\": %zu, "
                       "\"lds_per_block\": %zu}",
             e->device, e->prop.gcnArchName, e->prop.multiProcessorCount, e->prop.clockRate,
             (size_t)e->prop.totalGlobalMem, (size_t)e->prop.sharedMemPerBlock);
    return HH_OK;
}

int hh_cf_run(hh_engine* e, const hh_system* sys, const hh_cf_params* p, int njobs, const double* const* ic_dev,
              const long* ntraj, double* const* cf_dev, hh_cf_stats* stats)
{
    return hh::cf_run(e, sys, p, njobs, ic_dev, ntraj, cf_dev, stats, -1);
}

}  // extern "C"

int hh::cf_run(hh_engine* e, const hh_system* sys, const hh_cf_params* p, int njobs, const double* const* ic_dev,
               const long* ntraj, double* const* cf_dev, hh_cf_stats* stats, int record)
{
    if (record < 0) record = p && p->zimmermann ? 1 : 0;
    if (sys && sys->system == HH_SYS_CH4CO2) {
        if (record != 0) return fail(HH_E_UNSUPPORTED, "CH4-CO2: only correlation functions are provided");
        record = 3;   // coordinates per sample; the dipole pass follows the integration
    }
    if (!e || !sys || !p || njobs <= 0 || !ic_dev || !ntraj || !cf_dev)
        return fail(HH_E_INVALID, "hh_cf_run: null argument or njobs <= 0");
    int rc = hh::check_cf_args(sys, p);
    if (rc < 0) return rc;
    for (int j = 0; j < njobs; ++j)
        if (ntraj[j] <= 0 || !ic_dev[j] || !cf_dev[j]) return fail(HH_E_INVALID, "job %d: empty or null buffers", j);
    const int halves = p->halves == HH_HALVES_BOTH ? 2 : 1;
    std::lock_guard<std::recursive_mutex> lock(e->mu);
    HH_CHECK(hipSetDevice(e->device));
    e->last_jobs.clear();
    e->last_pool_jobs.clear();
    if (stats) std::memset(stats, 0, sizeof *stats);

    if (p->engine == HH_ENGINE_BATCHED) {
        // resident lane pool: all jobs are admitted at once and streamed through time slices
        if (p->block != 0 && p->block != hh::kPoolBlock)
            return fail(HH_E_INVALID, "block must be 0 or 256 (the batched engine's one launch shape)");
        hh::LanePool* pl = nullptr;
        rc = hh::pool_get(e, e->pools, *sys, p->precision == HH_PREC_MIXED, &pl);
        if (rc < 0) return rc;
        if (!pl->inflight.empty()) return fail(HH_E_BUSY, "lane pool of this class is in use");
        if (njobs > pl->ring_cap) return fail(HH_E_INVALID, "at most %d jobs per call", pl->ring_cap);
        std::vector<hh::PoolJob*> mine;
        for (int j = 0; j < njobs; ++j) {
            hh::PoolJob* pj = nullptr;
            rc = hh::pool_admit(e, *pl, *p, ic_dev[j], ntraj[j], cf_dev[j], 0, &pj, record);
            if (rc < 0) {
                while (!pl->inflight.empty()) pl->inflight.pop_back();
                return rc;
            }
            mine.push_back(pj);
        }
        rc = hh::pool_order_group(e, *pl);
        if (rc < 0) {
            while (!pl->inflight.empty()) pl->inflight.pop_back();
            return rc;
        }
        double kms = 0.0, rms = 0.0;
        int launches = 0;
        const double pool_ms0 = pl->launch_ms;
        const long long pool_n0 = pl->launch_count;
        for (;;) {
            bool all = true;
            for (hh::PoolJob* pj : mine) all &= pj->state == 2;
            if (all) break;
            bool progress = false;
            rc = hh::pool_step(e, e->pools, 20.0, &kms, &rms, &launches, &progress);
            if (rc == HH_OK && !progress)
                rc = fail(HH_E_DEVICE, "hh_cf_run: jobs left with no work in flight (internal bookkeeping error)");
            if (rc < 0) {
                // drop the pools with their state (the orbit pool holds jobs on this call's
                // buffers); the next call builds fresh ones
                (void)hipStreamSynchronize(e->stream);
                e->pools.clear();
                return rc;
            }
        }
        for (hh::PoolJob* pj : mine) {
            e->last_jobs.push_back(pj->jd);
            e->last_pool_jobs.push_back(hh::pool_take(*pl, pj));
        }
        if (stats) {
            stats->integrate_ms = kms;
            stats->reduce_ms = rms;
            stats->block = hh::kPoolBlock;
            stats->grid = (int)(pl->S / hh::kPoolBlock);
            stats->launches = launches;
            stats->pool_launches = (int)(pl->launch_count - pool_n0);
            stats->pool_kernel_ms = pl->launch_ms - pool_ms0;
            return hh::collect_stats(e->last_jobs, stats);
        }
        return HH_OK;
    }

    // plain: one thread per item, every job in one launch, each item runs to its end
    std::vector<hh::JobDesc> jobs(njobs);
    long long total = 0;
    for (int j = 0; j < njobs; ++j) {
        jobs[j].item_begin = total;
        jobs[j].n_items = (long long)ntraj[j] * halves;
        jobs[j].ntraj = ntraj[j];
        total += jobs[j].n_items;
    }
    const int rec = record ? 3 : 1;
    const size_t rows_bytes = (size_t)total * (size_t)p->L * rec * sizeof(double);
    if (e->rows.ensure(rows_bytes) != hipSuccess)
        return fail(HH_E_NOMEM, "cannot allocate %.2f GiB of per-trajectory CF rows", rows_bytes / 1073741824.0);
    HH_CHECK(e->len.ensure(total * sizeof(int)));
    HH_CHECK(e->status.ensure(total * sizeof(int)));
    HH_CHECK(e->nrhs.ensure(total * sizeof(int)));
    HH_CHECK(e->drift.ensure(total * sizeof(float)));
    HH_CHECK(e->jobs.ensure(njobs * sizeof(hh::JobDesc)));
    for (int j = 0; j < njobs; ++j) {
        hh::JobDesc& jb = jobs[j];
        const long long off = jb.item_begin;
        jb.ic = ic_dev[j];
        jb.rows = e->rows.as<double>() + off * p->L * rec;
        jb.record = record;
        jb.len = e->len.as<int>() + off;
        jb.status = e->status.as<int>() + off;
        jb.nrhs = e->nrhs.as<int>() + off;
        jb.e_drift = e->drift.as<float>() + off;
        jb.done = nullptr;
        jb.halves = halves;
        jb.L = p->L;
        jb.max_steps = p->max_steps > 0 ? p->max_steps : 50000;
        jb.dt = p->dt;
        jb.Rcut = p->Rcut;
        jb.rtol = p->rtol;
        jb.atol = p->atol;
    }
    HH_CHECK(hipMemcpyAsync(e->jobs.p, jobs.data(), njobs * sizeof(hh::JobDesc), hipMemcpyHostToDevice, e->stream));
    const hh::SysConst sc = hh::sysconst(e, *sys, false);
    const int grid = (int)((total + 255) / 256);
    // per-wavefront stage arena (kSlots * QP * 64 doubles per wavefront)
    const size_t per_lane = (size_t)hh::arena_doubles_per_lane(sys->system) * sizeof(double);
    if (e->arena.ensure(per_lane * (size_t)grid * 256) != hipSuccess)
        return fail(HH_E_NOMEM, "cannot allocate the integrator arena");
    HH_CHECK(hipEventRecord(e->ev[0], e->stream));
    HH_CHECK(hh::run_plain(sys->system, sc, e->jobs.as<hh::JobDesc>(), njobs, total, e->arena.as<double>(), e->stream));
    HH_CHECK(hipEventRecord(e->ev[1], e->stream));
    for (int j = 0; j < njobs; ++j) {
        const int rc = hh::reduce_job(e, jobs[j], p->ppf, cf_dev[j]);
        if (rc != HH_OK) return rc;
    }
    HH_CHECK(hipEventRecord(e->ev[2], e->stream));
    HH_CHECK(hipStreamSynchronize(e->stream));
    e->last_jobs = jobs;
    if (stats) {
        float a = 0, b = 0;
        (void)hipEventElapsedTime(&a, e->ev[0], e->ev[1]);
        (void)hipEventElapsedTime(&b, e->ev[1], e->ev[2]);
        stats->integrate_ms = a;
        stats->reduce_ms = b;
        stats->block = 256;
        stats->grid = grid;
        stats->launches = 1;
        return hh::collect_stats(jobs, stats);
    }
    return HH_OK;
}

extern "C" {

int hh_cf_items(hh_engine* e, int job, int* len, int* status, int* nrhs, float* e_drift, double* rows)
{
    if (!e) return fail(HH_E_INVALID, "hh_cf_items: engine is NULL");
    std::lock_guard<std::recursive_mutex> lock(e->mu);
    if (job < 0 || job >= (int)e->last_jobs.size()) return fail(HH_E_NOTFOUND, "no job %d in the last run", job);
    HH_CHECK(hipSetDevice(e->device));
    const hh::JobDesc& jb = e->last_jobs[job];
    const size_t n = (size_t)jb.n_items;
    if (len) HH_CHECK(hipMemcpy(len, jb.len, n * sizeof(int), hipMemcpyDeviceToHost));
    if (status) HH_CHECK(hipMemcpy(status, jb.status, n * sizeof(int), hipMemcpyDeviceToHost));
    if (nrhs) HH_CHECK(hipMemcpy(nrhs, jb.nrhs, n * sizeof(int), hipMemcpyDeviceToHost));
    if (e_drift) HH_CHECK(hipMemcpy(e_drift, jb.e_drift, n * sizeof(float), hipMemcpyDeviceToHost));
    if (rows) {
        // one copy of the whole block, then zero what lies past each item's length (never written
        // on the device); a CH4-CO2 job's rows are its CF rows (the coordinates are internal)
        std::vector<int> ln(n);
        HH_CHECK(hipMemcpy(ln.data(), jb.len, n * sizeof(int), hipMemcpyDeviceToHost));
        const int rec = jb.record == 3 ? 1 : (jb.record ? 3 : 1);
        const double* src = jb.record == 3 ? jb.cf_rows : jb.rows;
        HH_CHECK(hipMemcpy(rows, src, n * (size_t)jb.L * rec * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n * rec; ++i) {
            const size_t from = (size_t)std::max(0, std::min(ln[i / rec], jb.L));
            std::memset(rows + i * jb.L + from, 0, (jb.L - from) * sizeof(double));
        }
    }
    return HH_OK;
}

int hh_cf_array(hh_engine* e, const hh_system* sys, const hh_cf_params* p, const hh_cf_array_params* a,
                const double* ic_dev, long ntraj, double* cf_dev, double* nstar, hh_cf_stats* stats)
{
    if (!e || !sys || !p || !a || !ic_dev || ntraj <= 0 || !cf_dev || !nstar)
        return fail(HH_E_INVALID, "hh_cf_array: null argument or ntraj <= 0");
    if (a->ntemp < 1 || a->ntemp > 64 || !a->temps) return fail(HH_E_INVALID, "hh_cf_array: 1..64 temperatures");
    if (!(a->T_base > 0)) return fail(HH_E_INVALID, "hh_cf_array: T_base must be positive");
    for (int st = 0; st < a->ntemp; ++st)
        if (!(a->temps[st] > 0)) return fail(HH_E_INVALID, "hh_cf_array: temperature %d is not positive", st);
    // src/hawaii.c:2657-2659: only free (and metastable) and bound pairs have weights
    if (a->pair_state != HH_PAIR_FREE && a->pair_state != HH_PAIR_BOUND)
        return fail(HH_E_INVALID, "hh_cf_array: pair state must be free or bound");
    if (p->zimmermann) return fail(HH_E_INVALID, "hh_cf_array: Zimmermann jobs are not supported");
    // one CF job over the ensemble (its per-item rows stay with the engine)
    hh_cf_params q = *p;
    q.ppf = 1.0;
    hh::DevBuf cf_tmp;
    HH_CHECK(cf_tmp.ensure((size_t)p->L * sizeof(double)));
    const double* icp[1] = {ic_dev};
    const long nt[1] = {ntraj};
    double* cfp[1] = {cf_tmp.as<double>()};
    // held across the run and the read of its rows, so no other call on the engine replaces them
    std::lock_guard<std::recursive_mutex> lock(e->mu);
    int rc = hh_cf_run(e, sys, &q, 1, icp, nt, cfp, stats);
    if (rc < 0) return rc;
    HH_CHECK(hipSetDevice(e->device));
    if (e->last_jobs.size() != 1) return fail(HH_E_INVALID, "hh_cf_array: the engine lost the job's items");
    const hh::JobDesc jd = e->last_jobs[0];
    // energies at the initial points (Hamiltonian(ms) right after p_generator, src/hawaii.c:2626)
    const int M = a->ntemp;
    hh::DevBuf en, wr_dev, ns_dev;
    HH_CHECK(en.ensure((size_t)ntraj * sizeof(double)));
    const hh::SysConst sc = hh::sysconst(e, *sys, false);
    HH_CHECK(hh::run_eval(sys->system, sc, HH_EVAL_HAMILTONIAN, ic_dev, ntraj, en.as<double>(), e->stream));
    std::vector<double> E((size_t)ntraj);
    std::vector<int> status((size_t)jd.n_items);
    HH_CHECK(hipMemcpyAsync(E.data(), en.p, E.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HH_CHECK(hipMemcpyAsync(status.data(), jd.status, status.size() * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HH_CHECK(hipStreamSynchronize(e->stream));
    // The array loop skips a trajectory whose last integrator call ended on the step limit
    // (`if (status == -1) continue;`, src/hawaii.c:2637): no weight, not in nstar.  A dipole
    // failure ends correlation_eval first (status 1) and counts with a zero CF.
    auto counted = [&](long t) {
        if (jd.halves == 2) {
            const int f = status[2 * t], b = status[2 * t + 1];
            if (hh::dipole_failure(f) || hh::dipole_failure(b)) return true;
            return b != hh::ITEM_FAIL_STEPS;
        }
        return hh::dipole_failure(status[t]) || status[t] != hh::ITEM_FAIL_STEPS;
    };
    // weights in the reference's operation order, nstar summed in trajectory order
    constexpr double kHkT = 4.3597447222071e-18 / 1.380649e-23;   // Hartree / Boltzmann (src/constants.h:24)
    const double Tb = a->T_base;
    std::vector<double> wr((size_t)ntraj * M, 0.0), ns(M, 0.0);
    for (long t = 0; t < ntraj; ++t) {
        if (!counted(t)) continue;
        for (int st = 0; st < M; ++st) {
            const double Ts = a->temps[st];
            double WTT = 1.0;
            if (a->pair_state == HH_PAIR_BOUND && Ts < Tb) WTT = std::exp(-a->pesmin * kHkT / Tb / Ts * (Tb - Ts));
            const double weight = std::exp(-E[t] * kHkT / Tb / Ts * (Tb - Ts)) / WTT;
            if (a->pair_state == HH_PAIR_FREE && Ts > Tb) continue;
            const double ratio = a->ratios ? a->ratios[st] : 1.0;
            wr[(size_t)t * M + st] = weight * ratio;
            ns[st] += weight;
        }
    }
    HH_CHECK(wr_dev.ensure(wr.size() * sizeof(double)));
    HH_CHECK(ns_dev.ensure(ns.size() * sizeof(double)));
    HH_CHECK(hipMemcpyAsync(wr_dev.p, wr.data(), wr.size() * sizeof(double), hipMemcpyHostToDevice, e->stream));
    HH_CHECK(hipMemcpyAsync(ns_dev.p, ns.data(), ns.size() * sizeof(double), hipMemcpyHostToDevice, e->stream));
    // at most 256 chunks of trajectories per temperature
    const long long chunk = std::max(256LL, (jd.ntraj + 255) / 256);
    const int nchunks = (int)((jd.ntraj + chunk - 1) / chunk);
    const size_t pw = (size_t)M * nchunks * jd.L * sizeof(double);
    HH_CHECK(e->psum.ensure(pw));
    HH_CHECK(e->pcomp.ensure(pw));
    HH_CHECK(hh::run_reduce_weighted(jd, M, wr_dev.as<double>(), ns_dev.as<double>(), e->psum.as<double>(),
                                     e->pcomp.as<double>(), chunk, nchunks, cf_dev, e->stream));
    HH_CHECK(hipStreamSynchronize(e->stream));
    for (int st = 0; st < M; ++st) nstar[st] = ns[st];
    return HH_OK;
}

int hh_eval(hh_engine* e, const hh_system* sys, int what, const double* y_dev, long n, double* out_dev)
{
    if (!e || !sys || !y_dev || !out_dev || n <= 0) return fail(HH_E_INVALID, "hh_eval: bad argument");
    if (sys->system != HH_SYS_HEAR && sys->system != HH_SYS_CO2AR && sys->system != HH_SYS_CH4CO2)
        return fail(HH_E_UNSUPPORTED, "unknown system %d", sys->system);
    if (what < HH_EVAL_PES || what > HH_EVAL_DIPOLE_RATE)
        return fail(HH_E_INVALID, "hh_eval: unknown quantity %d", what);
    if (sys->system == HH_SYS_CH4CO2 && what == HH_EVAL_DIPOLE_RATE)
        return fail(HH_E_UNSUPPORTED, "CH4-CO2: |d mu/dt|^2 is not provided");
    std::lock_guard<std::recursive_mutex> lock(e->mu);
    HH_CHECK(hipSetDevice(e->device));
    if (sys->system == HH_SYS_CH4CO2 && what == HH_EVAL_DIPOLE) {
        hh::DevBuf rec, ln;
        if (rec.ensure((size_t)n * 8 * sizeof(double)) != hipSuccess || ln.ensure(sizeof(int)) != hipSuccess)
            return fail(HH_E_NOMEM, "hh_eval: no memory for the CH4-CO2 dipole records");
        HH_CHECK(hh::run_ch4co2_eval_dipole(y_dev, n, rec.as<double>(), ln.as<int>(), e->ch4tab64.p, out_dev, e->stream));
        HH_CHECK(hipStreamSynchronize(e->stream));
        return HH_OK;
    }
    const hh::SysConst sc = hh::sysconst(e, *sys, false);
    HH_CHECK(hh::run_eval(sys->system, sc, what, y_dev, n, out_dev, e->stream));
    HH_CHECK(hipStreamSynchronize(e->stream));
    return HH_OK;
}

int hh_cf_to_sf(hh_engine* e, double* cf_dev, long n, double dt, double* nu_dev, double* sf_dev, double* ms)
{
    if (!e || !cf_dev || !nu_dev || !sf_dev) return fail(HH_E_INVALID, "hh_cf_to_sf: null argument");
    if (n < 2 || (n & (n - 1))) return fail(HH_E_INVALID, "Length of the input array should be a power of 2");
    if (dt <= 0) return fail(HH_E_INVALID, "sampling time must be positive");
    std::lock_guard<std::recursive_mutex> lock(e->mu);
    HH_CHECK(hipSetDevice(e->device));
    HH_CHECK(e->fft_a.ensure(2 * n * sizeof(double2)));
    HH_CHECK(e->fft_b.ensure(2 * n * sizeof(double2)));
    const bool lds_ok = (size_t)(2 * n) * sizeof(double2) <= e->prop.sharedMemPerBlock;
    HH_CHECK(hipEventRecord(e->ev[0], e->stream));
    HH_CHECK(hh::run_cf_to_sf(cf_dev, n, dt, e->fft_a.as<double2>(), e->fft_b.as<double2>(), nu_dev, sf_dev, e->stream,
                              lds_ok));
    HH_CHECK(hipEventRecord(e->ev[1], e->stream));
    HH_CHECK(hipStreamSynchronize(e->stream));
    if (ms) {
        float a = 0;
        (void)hipEventElapsedTime(&a, e->ev[0], e->ev[1]);
        *ms = a;
    }
    return HH_OK;
}

int hh_cf_to_sf_noalloc(hh_engine* e, double* cf_dev, long n, double dt, double* nu_dev, double* sf_dev)
{
    if (!e || !cf_dev || !nu_dev || !sf_dev) return fail(HH_E_INVALID, "hh_cf_to_sf_noalloc: null argument");
    if (n < 2 || (n & (n - 1))) return fail(HH_E_INVALID, "Length of the input array should be a power of 2");
    if (n > hh::kCfToSfNoAllocMaxN)
        return fail(HH_E_INVALID, "hh_cf_to_sf_noalloc: n = %ld above its static work array's %ld", n,
                    hh::kCfToSfNoAllocMaxN);
    if (dt <= 0) return fail(HH_E_INVALID, "sampling time must be positive");
    std::lock_guard<std::recursive_mutex> lock(e->mu);
    HH_CHECK(hipSetDevice(e->device));
    double2* a = static_cast<double2*>(e->noalloc_scratch);
    double2* b = a + 2 * hh::kCfToSfNoAllocMaxN;
    const bool lds_ok = (size_t)(2 * n) * sizeof(double2) <= e->prop.sharedMemPerBlock;
    HH_CHECK(hh::run_cf_to_sf(cf_dev, n, dt, a, b, nu_dev, sf_dev, e->stream, lds_ok));
    HH_CHECK(hipStreamSynchronize(e->stream));
    return HH_OK;
}

int hh_m0(hh_engine* e, const hh_system* sys, const hh_sampler* smp, uint32_t seed, long npoints, int nstreams,
          double ppf, double* m, double* q, double* ms)
{
    return moment_run(e, sys, smp, seed, npoints, nstreams, ppf, 0, m, q, ms);
}

int hh_m2(hh_engine* e, const hh_system* sys, const hh_sampler* smp, uint32_t seed, long npoints, int nstreams,
          double ppf, double* m, double* q, double* ms)
{
    return moment_run(e, sys, smp, seed, npoints, nstreams, ppf, 2, m, q, ms);
}

}  // extern "C"

namespace {

// the rejection-sampler moments of hh_m0 (order 0) and hh_m2 (order 2)
int moment_run(hh_engine* e, const hh_system* sys, const hh_sampler* smp, uint32_t seed, long npoints, int nstreams,
               double ppf, int order, double* m, double* q, double* ms)
{
    if (!e || !sys || !smp || !m || !q) return fail(HH_E_INVALID, "hh_m%d: null argument", order);
    int rc = hh::m0_check(sys, smp, npoints, nstreams);
    if (rc < 0) return rc;
    std::lock_guard<std::recursive_mutex> lock(e->mu);
    HH_CHECK(hipSetDevice(e->device));
    HH_CHECK(e->m0_state.ensure(hh::m0_bytes(nstreams)));
    const hh::M0Acc acc = hh::m0_view(e->m0_state.p, nstreams);
    HH_CHECK(hipMemcpyAsync(acc.finished, &kZeroInt, sizeof(int), hipMemcpyHostToDevice, e->stream));
    const hh::SysConst sc = hh::sysconst(e, *sys, false);
    hh::M0Params mp{smp->T, smp->Rmin, smp->Rmax, smp->pesmin, smp->pair_state, 0};
    HH_CHECK(hipEventRecord(e->ev[0], e->stream));
    HH_CHECK(hh::run_m0(sys->system, sc, mp, seed, nstreams, npoints, reinterpret_cast<uint32_t*>(e->m0_state.p), acc, 1,
                        0, e->stream, order));
    HH_CHECK(hipEventRecord(e->ev[1], e->stream));
    HH_CHECK(hipStreamSynchronize(e->stream));
    rc = hh::m0_combine(acc, npoints, nstreams, ppf, m, q, order);
    if (rc < 0) return rc;
    if (ms) {
        float a = 0;
        (void)hipEventElapsedTime(&a, e->ev[0], e->ev[1]);
        *ms = a;
    }
    return HH_OK;
}

}  // namespace
