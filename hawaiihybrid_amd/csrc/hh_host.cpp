// Host side of the MI355X trajectory engine: device set-up, surface-table upload, work-item
// bookkeeping, launches, reductions and the C ABI of include/hawaii_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hawaii_gpu.h"
#include "co2ar_tables.h"
#include "hh_engine.h"
#include "hh_systems.h"

namespace hh {
int arena_doubles_per_lane(int sys);
hipError_t run_plain(int sys, const SysConst& sc, const JobDesc* jobs_dev, int njobs, long long total, double* ws,
                     hipStream_t st);
hipError_t run_persistent(int sys, bool mixed, int block, const SysConst& sc, const JobDesc* jobs_dev, int njobs,
                          long long total, QueueState* q, int grid, double* ws, hipStream_t st);
hipError_t run_reduce(const JobDesc& jb, double ppf, double* psum, double* pcomp, long long chunk, int nchunks,
                      double* cf_out, hipStream_t st);
hipError_t run_cf_to_sf(double* cf, long n, double dt, double2* work_a, double2* work_b, double* nu, double* sf,
                        hipStream_t st, bool lds_ok);
struct M0Params {
    double T, Rmin, Rmax, pesmin;
    int pair_state;
    int pad;
};
hipError_t run_m0(int sys, const SysConst& sc, const M0Params& p, uint32_t seed, long nstreams, long npoints,
                  uint32_t* state, double* m, double* q, long* n, hipStream_t st);
}  // namespace hh

extern "C" __attribute__((visibility("hidden"))) long hh_sample_ics_impl(const hh_system* sys, const hh_sampler* smp, uint32_t seed, long ntraj,
                                   hh_pes_fn pes, double* ic, long* n_draws);

namespace {

thread_local std::string g_err;

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

#define HH_CHECK(expr)                                                                                 \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail(HH_E_DEVICE, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                         \
    } while (0)

// grow-only device buffer
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n)
    {
        if (n <= bytes) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            bytes = 0;
        }
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

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

struct hh_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipDeviceProp_t prop{};
    DevBuf tab64, tab32;          // Co2ArTables<double>, Co2ArTables<float>
    DevBuf rows, len, status, nrhs, drift, jobs, queue, psum, pcomp, fft_a, fft_b, m0_state, m0_out, arena;
    std::vector<hh::JobDesc> last_jobs;
    std::mutex mu;
};

extern "C" {

const char* hh_last_error(void) { return g_err.c_str(); }

int hh_qp_size(int system) { return system == HH_SYS_CO2AR ? 10 : (system == HH_SYS_HEAR ? 6 : -1); }

long hh_sample_ics(const hh_system* sys, const hh_sampler* smp, uint32_t seed, long ntraj, hh_pes_fn pes,
                   double* ic, long* n_draws)
{
    if (!sys || !smp || !ic || ntraj < 0) return fail(HH_E_INVALID, "hh_sample_ics: null argument or ntraj < 0");
    if (sys->system != HH_SYS_HEAR && sys->system != HH_SYS_CO2AR)
        return fail(HH_E_UNSUPPORTED, "hh_sample_ics: unknown system %d", sys->system);
    if (sys->mu <= 0) return fail(HH_E_INVALID, "invalid reduced mass (%.5e) -- must be positive", sys->mu);
    if (smp->T <= 0) return fail(HH_E_INVALID, "temperature must be positive");
    if (smp->mode == HH_IC_REJECTION && !(smp->Rmin > 0 && smp->Rmax > smp->Rmin))
        return fail(HH_E_INVALID, "sampler_Rmin/sampler_Rmax must satisfy 0 < Rmin < Rmax");
    if (smp->mode == HH_IC_REJECTION && !(smp->pesmin < 0 || smp->pesmin > 0))
        return fail(HH_E_INVALID, "pesmin must be non-zero in rejection mode");
    return hh_sample_ics_impl(sys, smp, seed, ntraj, pes, ic, n_draws);
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
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        delete e;
        return fail(HH_E_DEVICE, "hipStreamCreate failed");
    }
    for (auto& ev : e->ev) (void)hipEventCreate(&ev);
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
    *out = e;
    return HH_OK;
}

void hh_engine_destroy(hh_engine* e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (DevBuf* b : {&e->tab64, &e->tab32, &e->rows, &e->len, &e->status, &e->nrhs, &e->drift, &e->jobs, &e->queue,
                      &e->psum, &e->pcomp, &e->fft_a, &e->fft_b, &e->m0_state, &e->m0_out, &e->arena})
        b->release();
    for (auto& ev : e->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (e->stream) (void)hipStreamDestroy(e->stream);
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
    if (!e || !sys || !p || njobs <= 0 || !ic_dev || !ntraj || !cf_dev)
        return fail(HH_E_INVALID, "hh_cf_run: null argument or njobs <= 0");
    if (sys->system != HH_SYS_HEAR && sys->system != HH_SYS_CO2AR)
        return fail(HH_E_UNSUPPORTED, "hh_cf_run: unknown system %d", sys->system);
    if (p->L <= 0 || p->dt <= 0 || p->Rcut <= 0 || p->rtol <= 0 || p->atol <= 0)
        return fail(HH_E_INVALID, "MaxTrajectoryLength, sampling_time, Rcut and tolerances must be positive");
    if (p->engine == HH_ENGINE_PLAIN && p->precision != HH_PREC_FP64)
        return fail(HH_E_INVALID, "the plain engine is FP64 only");
    const int halves = p->halves == HH_HALVES_BOTH ? 2 : 1;
    std::lock_guard<std::mutex> lock(e->mu);
    HH_CHECK(hipSetDevice(e->device));

    // job table and per-item buffers
    std::vector<hh::JobDesc> jobs(njobs);
    long long total = 0, total_traj = 0;
    for (int j = 0; j < njobs; ++j) {
        if (ntraj[j] <= 0 || !ic_dev[j] || !cf_dev[j]) return fail(HH_E_INVALID, "job %d: empty or null buffers", j);
        jobs[j].item_begin = total;
        jobs[j].n_items = (long long)ntraj[j] * halves;
        jobs[j].ntraj = ntraj[j];
        total += jobs[j].n_items;
        total_traj += ntraj[j];
    }
    const size_t rows_bytes = (size_t)total * (size_t)p->L * sizeof(double);
    if (e->rows.ensure(rows_bytes) != hipSuccess)
        return fail(HH_E_NOMEM, "cannot allocate %.2f GiB of per-trajectory CF rows", rows_bytes / 1073741824.0);
    HH_CHECK(e->len.ensure(total * sizeof(int)));
    HH_CHECK(e->status.ensure(total * sizeof(int)));
    HH_CHECK(e->nrhs.ensure(total * sizeof(int)));
    HH_CHECK(e->drift.ensure(total * sizeof(float)));
    HH_CHECK(e->jobs.ensure(njobs * sizeof(hh::JobDesc)));
    HH_CHECK(e->queue.ensure(sizeof(hh::QueueState)));
    for (int j = 0; j < njobs; ++j) {
        hh::JobDesc& jb = jobs[j];
        const long long off = jb.item_begin;
        jb.ic = ic_dev[j];
        jb.rows = e->rows.as<double>() + off * p->L;
        jb.len = e->len.as<int>() + off;
        jb.status = e->status.as<int>() + off;
        jb.nrhs = e->nrhs.as<int>() + off;
        jb.e_drift = e->drift.as<float>() + off;
        jb.halves = halves;
        jb.L = p->L;
        jb.max_steps = p->max_steps > 0 ? p->max_steps : 50000;
        jb.dt = p->dt;
        jb.Rcut = p->Rcut;
        jb.rtol = p->rtol;
        jb.atol = p->atol;
    }
    HH_CHECK(hipMemcpyAsync(e->jobs.p, jobs.data(), njobs * sizeof(hh::JobDesc), hipMemcpyHostToDevice, e->stream));
    hh::QueueState q0{0, 0};
    HH_CHECK(hipMemcpyAsync(e->queue.p, &q0, sizeof q0, hipMemcpyHostToDevice, e->stream));

    const bool mixed = p->precision == HH_PREC_MIXED;
    hh::SysConst sc{sys->mu, sys->I, mixed ? e->tab32.p : e->tab64.p};
    int block = p->block, grid = p->grid;
    // per-wavefront stage arena (kSlots * QP * 64 doubles per wavefront)
    const size_t per_lane = (size_t)hh::arena_doubles_per_lane(sys->system) * sizeof(double);
    if (p->engine == HH_ENGINE_PLAIN) {
        sc.tables = e->tab64.p;
        block = 256;
        grid = (int)((total + 255) / 256);
        if (e->arena.ensure(per_lane * (size_t)grid * 256) != hipSuccess)
            return fail(HH_E_NOMEM, "cannot allocate the integrator arena");
        HH_CHECK(hipEventRecord(e->ev[0], e->stream));
        HH_CHECK(hh::run_plain(sys->system, sc, e->jobs.as<hh::JobDesc>(), njobs, total, e->arena.as<double>(),
                               e->stream));
    } else {
        if (block != 64 && block != 128 && block != 256) block = 256;
        if (grid <= 0) grid = e->prop.multiProcessorCount * 2 * (256 / block);   // two waves per SIMD
        if (e->arena.ensure(per_lane * (size_t)grid * block) != hipSuccess)
            return fail(HH_E_NOMEM, "cannot allocate the integrator arena");
        HH_CHECK(hipEventRecord(e->ev[0], e->stream));
        HH_CHECK(hh::run_persistent(sys->system, mixed, block, sc, e->jobs.as<hh::JobDesc>(), njobs, total,
                                    e->queue.as<hh::QueueState>(), grid, e->arena.as<double>(), e->stream));
    }
    HH_CHECK(hipEventRecord(e->ev[1], e->stream));

    // reductions: chunks of trajectories per column tile
    const long long chunk = 256;
    for (int j = 0; j < njobs; ++j) {
        const int nchunks = (int)((jobs[j].ntraj + chunk - 1) / chunk);
        HH_CHECK(e->psum.ensure((size_t)nchunks * p->L * sizeof(double)));
        HH_CHECK(e->pcomp.ensure((size_t)nchunks * p->L * sizeof(double)));
        HH_CHECK(hh::run_reduce(jobs[j], p->ppf, e->psum.as<double>(), e->pcomp.as<double>(), chunk, nchunks, cf_dev[j],
                                e->stream));
    }
    HH_CHECK(hipEventRecord(e->ev[2], e->stream));
    HH_CHECK(hipStreamSynchronize(e->stream));
    e->last_jobs = jobs;

    if (stats) {
        float a = 0, b = 0;
        (void)hipEventElapsedTime(&a, e->ev[0], e->ev[1]);
        (void)hipEventElapsedTime(&b, e->ev[1], e->ev[2]);
        std::memset(stats, 0, sizeof *stats);
        stats->integrate_ms = a;
        stats->reduce_ms = b;
        stats->n_items = total;
        stats->block = block;
        stats->grid = grid;
        std::vector<int> ln(total), st(total), nr(total);
        std::vector<float> dr(total);
        HH_CHECK(hipMemcpy(ln.data(), e->len.p, total * sizeof(int), hipMemcpyDeviceToHost));
        HH_CHECK(hipMemcpy(st.data(), e->status.p, total * sizeof(int), hipMemcpyDeviceToHost));
        HH_CHECK(hipMemcpy(nr.data(), e->nrhs.p, total * sizeof(int), hipMemcpyDeviceToHost));
        HH_CHECK(hipMemcpy(dr.data(), e->drift.p, total * sizeof(float), hipMemcpyDeviceToHost));
        for (long long i = 0; i < total; ++i) {
            stats->nrhs += nr[i];
            stats->nsamples += ln[i];
            if (st[i] >= 1 && st[i] <= 3) ++stats->n_failed;
            if (st[i] >= 4) ++stats->n_stepfail;
            stats->max_e_drift = std::max(stats->max_e_drift, (double)dr[i]);
        }
    }
    return HH_OK;
}

int hh_cf_items(hh_engine* e, int job, int* len, int* status, int* nrhs, float* e_drift, double* rows)
{
    if (!e) return fail(HH_E_INVALID, "hh_cf_items: engine is NULL");
    std::lock_guard<std::mutex> lock(e->mu);
    if (job < 0 || job >= (int)e->last_jobs.size()) return fail(HH_E_NOTFOUND, "no job %d in the last run", job);
    HH_CHECK(hipSetDevice(e->device));
    const hh::JobDesc& jb = e->last_jobs[job];
    const size_t n = (size_t)jb.n_items;
    if (len) HH_CHECK(hipMemcpy(len, jb.len, n * sizeof(int), hipMemcpyDeviceToHost));
    if (status) HH_CHECK(hipMemcpy(status, jb.status, n * sizeof(int), hipMemcpyDeviceToHost));
    if (nrhs) HH_CHECK(hipMemcpy(nrhs, jb.nrhs, n * sizeof(int), hipMemcpyDeviceToHost));
    if (e_drift) HH_CHECK(hipMemcpy(e_drift, jb.e_drift, n * sizeof(float), hipMemcpyDeviceToHost));
    if (rows) {
        std::vector<int> ln(n);
        HH_CHECK(hipMemcpy(ln.data(), jb.len, n * sizeof(int), hipMemcpyDeviceToHost));
        std::memset(rows, 0, n * (size_t)jb.L * sizeof(double));
        for (size_t i = 0; i < n; ++i)
            if (ln[i] > 0)
                HH_CHECK(hipMemcpy(rows + i * jb.L, jb.rows + i * (size_t)jb.L, ln[i] * sizeof(double),
                                   hipMemcpyDeviceToHost));
    }
    return HH_OK;
}

int hh_cf_to_sf(hh_engine* e, double* cf_dev, long n, double dt, double* nu_dev, double* sf_dev, double* ms)
{
    if (!e || !cf_dev || !nu_dev || !sf_dev) return fail(HH_E_INVALID, "hh_cf_to_sf: null argument");
    if (n < 2 || (n & (n - 1))) return fail(HH_E_INVALID, "Length of the input array should be a power of 2");
    if (dt <= 0) return fail(HH_E_INVALID, "sampling time must be positive");
    std::lock_guard<std::mutex> lock(e->mu);
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

int hh_m0(hh_engine* e, const hh_system* sys, const hh_sampler* smp, uint32_t seed, long npoints, int nstreams,
          double ppf, double* m, double* q, double* ms)
{
    if (!e || !sys || !smp || !m || !q) return fail(HH_E_INVALID, "hh_m0: null argument");
    if (npoints < 2 || nstreams <= 0 || npoints < nstreams * 2L)
        return fail(HH_E_INVALID, "hh_m0: need npoints >= 2 * nstreams > 0");
    if (smp->mode != HH_IC_REJECTION) return fail(HH_E_INVALID, "hh_m0: M0 uses the rejection sampler");
    if (!(smp->pesmin < 0 || smp->pesmin > 0)) return fail(HH_E_INVALID, "hh_m0: pesmin must be non-zero");
    std::lock_guard<std::mutex> lock(e->mu);
    HH_CHECK(hipSetDevice(e->device));
    HH_CHECK(e->m0_state.ensure((size_t)nstreams * 624 * sizeof(uint32_t)));
    HH_CHECK(e->m0_out.ensure((size_t)nstreams * 3 * sizeof(double)));
    double* om = e->m0_out.as<double>();
    double* oq = om + nstreams;
    long* on = reinterpret_cast<long*>(oq + nstreams);
    hh::SysConst sc{sys->mu, sys->I, e->tab64.p};
    hh::M0Params mp{smp->T, smp->Rmin, smp->Rmax, smp->pesmin, smp->pair_state, 0};
    HH_CHECK(hipEventRecord(e->ev[0], e->stream));
    HH_CHECK(hh::run_m0(sys->system, sc, mp, seed, nstreams, npoints, e->m0_state.as<uint32_t>(), om, oq, on, e->stream));
    HH_CHECK(hipEventRecord(e->ev[1], e->stream));
    HH_CHECK(hipStreamSynchronize(e->stream));
    std::vector<double> hm(nstreams), hq(nstreams);
    std::vector<long> hn(nstreams);
    HH_CHECK(hipMemcpy(hm.data(), om, nstreams * sizeof(double), hipMemcpyDeviceToHost));
    HH_CHECK(hipMemcpy(hq.data(), oq, nstreams * sizeof(double), hipMemcpyDeviceToHost));
    HH_CHECK(hipMemcpy(hn.data(), on, nstreams * sizeof(long), hipMemcpyDeviceToHost));
    // combine as mpi_calculate_M0 does across ranks (src/hawaii.c:1722-1733): mean of the
    // per-stream means and of the per-stream q, q normalised by the per-stream count
    for (int s = 0; s < nstreams; ++s) {
        const long quota = npoints / nstreams + (s < npoints % nstreams ? 1 : 0);
        if (hn[s] != quota)
            return fail(HH_E_INVALID, "hh_m0: stream %d accepted %ld of %ld points within the candidate limit "
                                      "(check pesmin, T and the pair state)", s, hn[s], quota);
    }
    double sm = 0.0, sq = 0.0;
    for (int s = 0; s < nstreams; ++s) {
        sm += hm[s];
        sq += hq[s];
    }
    const double ZeroCoeff = 0.00361479637 / (4.0 * M_PI);
    const double n0 = (double)hn[0];
    *m = (sm / nstreams) * ZeroCoeff * ppf;
    *q = std::sqrt((sq / nstreams) / n0 / (n0 - 1)) * ZeroCoeff * ppf;
    if (ms) {
        float a = 0;
        (void)hipEventElapsedTime(&a, e->ev[0], e->ev[1]);
        *ms = a;
    }
    return HH_OK;
}

}  // extern "C"
