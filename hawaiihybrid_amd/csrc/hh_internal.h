// Host-side internals shared by hh_host.cpp (C ABI, synchronous entry points) and
// hh_pool.cpp (resident lane pool and serving session).
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hawaii_gpu.h"
#include "hh_engine.h"
#include "hh_systems.h"

namespace hh {
// kernel launchers (hh_kernels.hip, hh_aux.hip)
int arena_doubles_per_lane(int sys);
hipError_t run_plain(int sys, const SysConst& sc, const JobDesc* jobs_dev, int njobs, long long total, double* ws,
                     hipStream_t st);
hipError_t run_pool(int sys, bool mixed, int block, int waves, const SysConst& sc, const JobDesc* ring,
                    const PoolCtl& ctl, const LaneSlots& sl, double* ws, hipStream_t st);
hipError_t run_slots_clear(const LaneSlots& sl, hipStream_t st);
hipError_t run_clamp_next(unsigned long long* next, unsigned long long total, hipStream_t st);
// right-hand sides spent on the job in ring slot `ring_slot` (added to *out)
hipError_t run_pool_work(const LaneSlots& sl, int ring_slot, const int* item_nrhs, long long n_items,
                         unsigned long long* out, hipStream_t st);
hipError_t run_repack(int sys, const LaneSlots& src, const LaneSlots& dst, int* scratch, hipStream_t st);
hipError_t run_eval(int sys, const SysConst& sc, int what, const double* y, long long n, double* out, hipStream_t st);
// CH4-CO2 dipole tables (hh_ch4co2.hip): bytes and host image in FP32 or FP64; the dipole pass of a
// record-3 job (dipoles over its coordinate rows, then the CF rows jd.cf_rows with the sample checks)
size_t ch4co2_table_bytes(bool f32);
void ch4co2_tables_host(bool f32, void* out);
hipError_t run_ch4co2_cf_rows(const JobDesc& jd, const void* tables, bool f32, hipStream_t st);
// CH4-CO2 lab dipoles at n phase points (SoA [16][n]) into out[3][n] (FP64 tables); records holds
// 8 n doubles of scratch, len one int
hipError_t run_ch4co2_eval_dipole(const double* y, long long n, double* records, int* len, const void* tables,
                                  double* out, hipStream_t st);
// scheduling keys of a job's items: what = 0 the straight-line transit time, 1 the squared impact
// parameter, 2 the transit time with the items marked 1 in keys moved to a front tier
hipError_t run_order_keys(const JobDesc& jb, double mu, int what, float* keys, int* idx, int idx_offset,
                          hipStream_t st);
// keys[order[r]] = (r < n_front) for r in [0, n)
hipError_t run_order_front(const int* order, long long n, long long n_front, float* keys, hipStream_t st);
size_t slots_bytes(int sys, long long S);
LaneSlots slots_view(int sys, void* base, long long S);
hipError_t run_reduce(const JobDesc& jb, double ppf, double* psum, double* pcomp, long long chunk, int nchunks,
                      double* cf_out, hipStream_t st);
// CORRELATION_ARRAY: per-temperature weighted sums of a job's items, normalised by nstar
// (device [ntemp]); wr is device [ntraj][ntemp], psum/pcomp hold ntemp * nchunks * L doubles
hipError_t run_reduce_weighted(const JobDesc& jb, int ntemp, const double* wr, const double* nstar, double* psum,
                               double* pcomp, long long chunk, int nchunks, double* cf_out, hipStream_t st);
// Zimmermann's trick (hh_zimm.hip)
int zimm_log_points(int L);
size_t zimm_workspace_doubles(int L);
int zimm_max_length();
hipError_t run_zimm_reduce(const JobDesc& jb, double ppf, const double2* tw, double* psum, double* pcomp,
                           double* cf_out, hipStream_t st);
hipError_t run_cf_to_sf(double* cf, long n, double dt, double2* work_a, double2* work_b, double* nu, double* sf,
                        hipStream_t st, bool lds_ok);
// hh_cf_to_sf_noalloc: the device address of its static work array on the current device
hipError_t noalloc_scratch_address(void** ptr);
hipError_t run_m0(int sys, const SysConst& sc, const M0Params& p, uint32_t seed, long nstreams, long npoints,
                  uint32_t* state, const M0Acc& acc, int init, long long ticks, hipStream_t st, int order = 0);
int m0_check(const hh_system* sys, const hh_sampler* smp, long npoints, int nstreams);
size_t m0_bytes(long nstreams);
M0Acc m0_view(void* base, long nstreams);
// M0 (order 0) or M2 (order 2) of one request: combine the per-stream results as
// mpi_calculate_M0 / mpi_calculate_M2 combine ranks
int m0_combine(const M0Acc& acc, long npoints, long nstreams, double ppf, double* m, double* q, int order = 0);

// error reporting: thread-local message, negative code
int fail(int code, const char* fmt, ...);

#define HH_CHECK(expr)                                                                                     \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return ::hh::fail(HH_E_DEVICE, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

// grow-only device buffer
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n)
    {
        if (n <= bytes) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess)
            bytes = n;
        else
            p = nullptr;
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
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
};

// One correlation-function job inside a lane pool.
struct PoolJob {
    long long seq = 0;          // sequence number in its pool (ring slot seq % ring)
    JobDesc jd{};               // device descriptor (points into `buf` and the caller's ic)
    DevBuf buf;                 // rows [n_items][L] + len, status, nrhs, e_drift
    double* cf_out = nullptr;   // caller's CF[L]
    double ppf = 1.0;
    int state = 0;              // 0 integrating, 1 reduced (launched), 2 complete
    uint64_t request = 0;       // serving request id (0: synchronous run)
    size_t bytes = 0;           // device bytes held by `buf`
    std::shared_ptr<DevBuf> group_order;   // order of the admission group this job belongs to
    bool items_done = false;    // every item of the job has finished (seen after a slice)
    // mixed precision: items handed over to FP64 (jd.orbit_items / jd.orbit_count) run as orbit
    // jobs in the FP64 pool `orbit_pool` at rtol = atol = orbit_rtol; the job is reduced once its
    // own items and all its orbit jobs have finished
    double orbit_rtol = 0.0;
    long long orbit_admitted = 0;           // hand-overs already admitted as orbit jobs
    struct LanePool* orbit_pool = nullptr;
    int orbit_pending = 0;                  // orbit jobs admitted and not yet finished
    PoolJob* parent = nullptr;              // an orbit job: the job whose items it integrates
};

// A resident lane pool for one (system, precision, block) class: S = grid * block lane slots
// with their saved integrator state, the per-wavefront stage arena, a ring of in-flight job
// descriptors and the class's work queue.
struct LanePool {
    int sys = 0;
    hh_system system{};         // mu and I of every job in the pool
    bool mixed = false;
    int block = 256;
    int waves = 2;                  // resident wavefronts per SIMD
    long long S = 0;
    // in-flight jobs live in ring slots seq % ring_cap: the sequence numbers of the jobs in flight
    // must span fewer than ring_cap (orbit jobs arrive every slice while a long one still runs)
    int ring_cap = 8192;
    DevBuf slots[2];            // double-buffered for repacking
    int cur = 0;
    DevBuf arena, ring, ring_done, ring_orbit, ring_orbit_nrhs, next, scratch, sort_tmp;
    // every pool's slices run on its own stream (default priority), so the slices of different
    // pools (an FP64 orbit pool beside its mixed pool, classes of a serving mix) overlap
    hipStream_t stream = nullptr;
    hipEvent_t slice_end = nullptr;
    double launch_ms = 0.0;         // summed durations of this pool's slice launches
    long long launch_count = 0;     // and their number (pool_step)
    std::deque<std::unique_ptr<PoolJob>> inflight;   // sequence order
    long long next_seq = 0;
    unsigned long long total = 0;   // items admitted
    bool drained_repacked = false;
    long long unordered_seq = 0;    // first job not yet ordered (pool_order_group)
    LaneSlots view(int which) const { return slots_view(sys, slots[which].p, S); }
    // one slice as a captured graph: [repack (when the queue has drained), clamp of the queue
    // head, the pool kernel, copies of the done counters and the head into pinned memory]
    struct SliceGraph {
        bool repack = false;
        int cur = 0;
        PoolCtl ctl{};
        hipGraphExec_t exec = nullptr;
    };
    std::vector<SliceGraph> graphs;   // a few variants (repack or not, buffer, admission state)
    bool repack_pending = false;
    int* h_done = nullptr;            // pinned copies read by the host after each slice
    int* h_orbit = nullptr;
    unsigned long long* h_next = nullptr;
    LanePool() = default;
    LanePool(const LanePool&) = delete;
    LanePool& operator=(const LanePool&) = delete;
    ~LanePool()
    {
        for (auto& g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (h_done) (void)hipHostFree(h_done);
        if (h_orbit) (void)hipHostFree(h_orbit);
        if (h_next) (void)hipHostFree(h_next);
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        if (slice_end) (void)hipEventDestroy(slice_end);
    }
};

}  // namespace hh

struct hh_session;

struct hh_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t capture = nullptr;      // non-blocking stream used only to capture slice graphs
    bool use_graphs = false;            // HH_GRAPHS=1: slices as captured graphs (opt-in: see hh_pool.cpp)
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hipDeviceProp_t prop{};
    int wall_khz = 100000;        // s_memrealtime frequency
    hh::DevBuf tab64, tab32;      // Co2ArTables<double>, Co2ArTables<float>
    hh::DevBuf ch4tab64, ch4tab32;   // CH4-CO2 dipole tables (FP64, FP32)
    hh::DevBuf rows, len, status, nrhs, drift, jobs, queue, psum, pcomp, fft_a, fft_b, m0_state, arena;
    hh::DevBuf zimm_tw;           // twiddles e^{2 pi i j/P} of the Zimmermann transforms
    void* noalloc_scratch = nullptr;    // device address of hh_cf_to_sf_noalloc's static work array
    hh::DevBuf prmu_a, prmu_b, prmu_w;   // PR/mu: FFT ping-pong series, weights and spectra
    hh::DevBuf prmu_tw;                  // PR/mu: twiddles e^{2 pi i j/L}, j < L, of FFT length prmu_tw_L
    long long prmu_tw_L = 0;
    int zimm_logP = 0;
    std::vector<hh::JobDesc> last_jobs;
    std::vector<std::unique_ptr<hh::PoolJob>> last_pool_jobs;   // kept for hh_cf_items
    std::vector<std::unique_ptr<hh::LanePool>> pools;
    // serialises calls on the engine; recursive so that hh_cf_array and hh_prmu hold it across
    // their inner run and the read of its per-item results
    std::recursive_mutex mu;
};

namespace hh {
// the surface constants of a system for the device kernels (f32: the FP32 tables of CO2-Ar)
inline SysConst sysconst(const hh_engine* e, const hh_system& sys, bool f32)
{
    SysConst sc{sys.mu, sys.I, (f32 && sys.system == HH_SYS_CO2AR) ? e->tab32.p : e->tab64.p, {0.0, 0.0, 0.0}};
    for (int i = 0; i < 3; ++i) sc.IR[i] = sys.I_rotor[i];
    return sc;
}
// the ensemble CF of a finished job into cf_out (on the engine stream)
int reduce_job(hh_engine* e, const JobDesc& jd, double ppf, double* cf_out);
using PoolSet = std::vector<std::unique_ptr<LanePool>>;
// The batched engine's one launch shape: 256-lane workgroups, one wavefront per SIMD.  Builds at
// two wavefronts per SIMD (64- and 128-lane workgroups) gave rows that differed between runs of
// the same input on a few items in 10^4, with extra right-hand sides, in FP64 and in mixed
// precision, while the one-wavefront build and the plain engine reproduce their bits.
constexpr int kPoolBlock = 256, kPoolWaves = 1;
// share of an admission group of a rotating-monomer system handed out first, smallest impact
// parameter first (pool_order_group; the environment variable HH_FRONT_SHARE overrides it)
constexpr double kFrontShare = 0.10;
// the lane pool of class (system, mu, I, precision) in `set`, created on first use
int pool_get(hh_engine* e, PoolSet& set, const hh_system& sys, bool mixed, LanePool** out);
// device bytes a job of this shape holds while it runs
size_t pool_job_bytes(const hh_cf_params& p, long ntraj, int record = -1);
// true when a job of this class hands its orbiting items over to FP64 (mixed precision, a
// system with a rotor, orbit_rtol >= 0)
bool orbit_handover(const hh_system& sys, const hh_cf_params& p);
// admit a correlation-function job: allocates its buffers, uploads its descriptor into the
// pool's ring and appends its items to the pool's queue; *out (optional) sees the job.
// record: the JobDesc row mode (-1: 1 with p.zimmermann, else 0; 2: a PR/mu dipole series)
int pool_admit(hh_engine* e, LanePool& pl, const hh_cf_params& p, const double* ic_dev, long ntraj, double* cf_dev,
               uint64_t request, PoolJob** out, int record = -1);
// hh_cf_run with the row mode of every job (-1: from p->zimmermann; 2: PR/mu series, no reduce)
int cf_run(hh_engine* e, const hh_system* sys, const hh_cf_params* p, int njobs, const double* const* ic_dev,
           const long* ntraj, double* const* cf_dev, hh_cf_stats* stats, int record);
// hand out the items of the jobs admitted since the last call (a group) longest-expected-first
// across the whole group; call after admitting them and before the next pool_step
int pool_order_group(hh_engine* e, LanePool& pl);
// true when the ring has room for one more job
bool pool_has_room(const LanePool& pl);
// one slice of every pool with work (each pool on its own stream, concurrently), then
// admission of the items handed over to FP64 as orbit jobs, reductions of the jobs whose items
// (orbit jobs included) are all done and repacking of drained pools; synchronises the engine
// stream.  Jobs whose CF is final have state 2 afterwards; their orbit jobs are gone from the
// orbit pool.  kernel_ms / reduce_ms accumulate device time.
int pool_step(hh_engine* e, PoolSet& set, double slice_ms, double* kernel_ms, double* reduce_ms,
              int* launches = nullptr, bool* progress = nullptr);
// argument checks shared by hh_cf_run and hh_submit_cf
int check_cf_args(const hh_system* sys, const hh_cf_params* p);
// per-item summaries (items, RHS, samples, failures, drift) of jobs, added into stats
int collect_stats(const std::vector<JobDesc>& jobs, hh_cf_stats* stats);
// remove the complete job from its pool and hand it over
std::unique_ptr<PoolJob> pool_take(LanePool& pl, PoolJob* job);
}  // namespace hh
