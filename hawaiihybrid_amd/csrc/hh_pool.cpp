// Resident lane pools of the batched engine.
//
// A pool owns S = 1 wavefront/SIMD x 64 lanes of integrator slots (65536 on 256 CUs) for one
// class of work (system, mu, I, precision).  Work runs in slices: one launch of pool_cf_kernel
// restores every slot, integrates and refills from the pool's queue until its time is up, and
// saves the slots again.  Between slices the host
//   * admits new jobs (their items are appended to the queue: continuous batching),
//   * launches the ensemble reduction of every job whose items have all finished, so a job's
//     CF does not wait for the slowest trajectory of other jobs,
//   * repacks the slots once the queue is drained, spreading the remaining items over all
//     wavefronts so the long tail of the ensemble runs with little divergence.
// Every item's arithmetic is independent of the slice boundaries and of its neighbours, and
// each job's reduction has a fixed order, so a job's CF is the same bits alone or in a mix.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include <hipcub/hipcub.hpp>

#include "hh_internal.h"

namespace hh {

namespace {
const int kZero = 0;
}  // namespace

int pool_get(hh_engine* e, PoolSet& set, const hh_system& sys, bool mixed, LanePool** out)
{
    for (auto& p : set)
        if (p->sys == sys.system && p->mixed == mixed && p->system.mu == sys.mu && p->system.I == sys.I && p->system.I_rotor[0] == sys.I_rotor[0] &&
            p->system.I_rotor[1] == sys.I_rotor[1] && p->system.I_rotor[2] == sys.I_rotor[2]) {
            *out = p.get();
            return HH_OK;
        }
    auto pl = std::make_unique<LanePool>();
    pl->sys = sys.system;
    pl->system = sys;
    pl->mixed = mixed;
    pl->block = kPoolBlock;
    pl->waves = kPoolWaves;
    // one 256-lane workgroup (a wavefront on each of its four SIMDs) per CU
    const long long grid = (long long)e->prop.multiProcessorCount * kPoolWaves * (256 / kPoolBlock);
    pl->S = grid * kPoolBlock;
    const size_t sb = slots_bytes(pl->sys, pl->S);
    const size_t arena = (size_t)arena_doubles_per_lane(pl->sys) * sizeof(double) * (size_t)pl->S;
    if (pl->slots[0].ensure(sb) != hipSuccess || pl->slots[1].ensure(sb) != hipSuccess ||
        pl->arena.ensure(arena) != hipSuccess || pl->ring.ensure(pl->ring_cap * sizeof(JobDesc)) != hipSuccess ||
        pl->ring_done.ensure(pl->ring_cap * sizeof(int)) != hipSuccess ||
        pl->ring_orbit.ensure(pl->ring_cap * sizeof(int)) != hipSuccess ||
        pl->ring_orbit_nrhs.ensure(pl->ring_cap * sizeof(unsigned long long)) != hipSuccess ||
        pl->next.ensure(sizeof(unsigned long long)) != hipSuccess ||
        pl->scratch.ensure((size_t)(pl->S / 256 + 1) * sizeof(int)) != hipSuccess)
        return fail(HH_E_NOMEM, "cannot allocate a lane pool (%.1f MiB of slots and arena)",
                    (2.0 * sb + arena) / 1048576.0);
    HH_CHECK(hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking));
    HH_CHECK(hipEventCreate(&pl->slice_end));
    HH_CHECK(run_slots_clear(pl->view(0), e->stream));
    const unsigned long long z = 0;
    HH_CHECK(hipMemcpyAsync(pl->next.p, &z, sizeof z, hipMemcpyHostToDevice, e->stream));
    HH_CHECK(hipStreamSynchronize(e->stream));
    *out = pl.get();
    set.push_back(std::move(pl));
    return HH_OK;
}

size_t pool_job_bytes(const hh_cf_params& p, long ntraj, int record)
{
    const long long items = (long long)ntraj * (p.halves == HH_HALVES_BOTH ? 2 : 1);
    if (record < 0) record = p.zimmermann ? 1 : 0;
    // rows ([3][L] per item for a dipole series), len, status, nrhs, e_drift, the job's share
    // of its group's order and its list of items handed over to FP64
    // CH4-CO2 (record 3): eight coordinates per sample and the CF rows the dipole pass writes
    const size_t row = (size_t)p.L * (record == 3 ? 9 : row_width(record));
    return (size_t)items * (row * sizeof(double) + 5 * sizeof(int) + sizeof(float));
}

bool orbit_handover(const hh_system& sys, const hh_cf_params& p)
{
    // a central potential (He-Ar) has no orbiting chaos and its mixed engine keeps FP64 forces
    return p.precision == HH_PREC_MIXED && sys.system != HH_SYS_HEAR && p.orbit_rtol >= 0.0;
}

bool pool_has_room(const LanePool& pl)
{
    return pl.inflight.empty() || pl.next_seq - pl.inflight.front()->seq < pl.ring_cap;
}

int pool_admit(hh_engine* e, LanePool& pl, const hh_cf_params& p, const double* ic_dev, long ntraj, double* cf_dev,
               uint64_t request, PoolJob** out, int record)
{
    if (record < 0) record = p.zimmermann ? 1 : 0;
    if (!pool_has_room(pl)) return fail(HH_E_BUSY, "lane pool ring full (%d jobs in flight)", pl.ring_cap);
    auto job = std::make_unique<PoolJob>();
    const int halves = p.halves == HH_HALVES_BOTH ? 2 : 1;
    const long long n_items = (long long)ntraj * halves;
    const size_t bytes = pool_job_bytes(p, ntraj, record);
    if (job->buf.ensure(bytes) != hipSuccess)
        return fail(HH_E_NOMEM, "cannot allocate %.2f GiB for a job of %ld trajectories", bytes / 1073741824.0, ntraj);
    job->bytes = bytes;
    job->seq = pl.next_seq;
    job->cf_out = cf_dev;
    job->ppf = p.ppf;
    job->request = request;
    const int slot = (int)(job->seq % pl.ring_cap);
    JobDesc& jd = job->jd;
    jd.ic = ic_dev;
    jd.rows = job->buf.as<double>();
    jd.record = record;
    jd.cf_rows = record == 3 ? jd.rows + (size_t)n_items * p.L * 8 : nullptr;
    jd.f32_dipole = pl.mixed ? 1 : 0;
    jd.len = reinterpret_cast<int*>(jd.rows + (size_t)n_items * p.L * (record == 3 ? 9 : row_width(record)));
    jd.status = jd.len + n_items;
    jd.nrhs = jd.status + n_items;
    jd.e_drift = reinterpret_cast<float*>(jd.nrhs + n_items);
    jd.ntraj = ntraj;
    jd.item_begin = (long long)pl.total;
    jd.n_items = n_items;
    jd.done = pl.ring_done.as<int>() + slot;
    jd.halves = halves;
    jd.L = p.L;
    jd.max_steps = p.max_steps > 0 ? p.max_steps : 50000;
    jd.dt = p.dt;
    jd.Rcut = p.Rcut;
    jd.rtol = p.rtol;
    jd.atol = p.atol;
    jd.order = nullptr;
    jd.order_base = 0;
    jd.map = nullptr;
    jd.orbit_items = nullptr;
    jd.orbit_count = nullptr;
    jd.orbit_nrhs = nullptr;
    if (orbit_handover(pl.system, p)) {
        static const unsigned long long zero64 = 0;
        jd.orbit_items = reinterpret_cast<int*>(jd.e_drift + n_items);
        jd.orbit_count = pl.ring_orbit.as<int>() + slot;
        jd.orbit_nrhs = pl.ring_orbit_nrhs.as<unsigned long long>() + slot;
        job->orbit_rtol = p.orbit_rtol > 0.0 ? p.orbit_rtol : 1e-12;
        HH_CHECK(hipMemcpyAsync(jd.orbit_count, &kZero, sizeof(int), hipMemcpyHostToDevice, e->stream));
        HH_CHECK(hipMemcpyAsync(jd.orbit_nrhs, &zero64, sizeof zero64, hipMemcpyHostToDevice, e->stream));
    }
    HH_CHECK(hipMemcpyAsync(pl.ring.as<JobDesc>() + slot, &job->jd, sizeof(JobDesc), hipMemcpyHostToDevice, e->stream));
    HH_CHECK(hipMemcpyAsync(jd.done, &kZero, sizeof(int), hipMemcpyHostToDevice, e->stream));
    // lanes that found the queue empty in the last slice left the head past the end; none of
    // those indices was handed out, so the head goes back to the first item of this job
    HH_CHECK(run_clamp_next(pl.next.as<unsigned long long>(), pl.total, e->stream));
    pl.total += (unsigned long long)n_items;
    pl.next_seq++;
    pl.drained_repacked = false;
    if (out) *out = job.get();
    pl.inflight.push_back(std::move(job));
    return HH_OK;
}

int pool_order_group(hh_engine* e, LanePool& pl)
{
    std::vector<PoolJob*> group;
    for (auto& j : pl.inflight)
        if (j->seq >= pl.unordered_seq) group.push_back(j.get());
    pl.unordered_seq = pl.next_seq;
    if (group.empty()) return HH_OK;
    const long long base = group.front()->jd.item_begin;
    const long long n = group.back()->jd.item_begin + group.back()->jd.n_items - base;
    if (n > 0x7fffffffLL) return fail(HH_E_INVALID, "admission group of %lld items is too large", n);
    auto order = std::make_shared<DevBuf>();
    size_t tmp_bytes = 0, tmp_asc = 0;
    HH_CHECK(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp_bytes, (const float*)nullptr, (float*)nullptr,
                                                          (const int*)nullptr, (int*)nullptr, (int)n, 0, 32,
                                                          e->stream));
    HH_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_asc, (const float*)nullptr, (float*)nullptr,
                                                (const int*)nullptr, (int*)nullptr, (int)n, 0, 32, e->stream));
    tmp_bytes = std::max(tmp_bytes, tmp_asc);
    const size_t kv = (size_t)n * (2 * sizeof(float) + sizeof(int));
    if (order->ensure((size_t)n * sizeof(int)) != hipSuccess || pl.sort_tmp.ensure(kv + tmp_bytes + 256) != hipSuccess)
        return fail(HH_E_NOMEM, "cannot allocate the ordering workspace");
    float* keys_in = pl.sort_tmp.as<float>();
    float* keys_out = keys_in + n;
    int* idx_in = reinterpret_cast<int*>(keys_out + n);
    void* tmp = reinterpret_cast<char*>(pl.sort_tmp.p) + ((kv + 255) / 256) * 256;
    // Around a rotating monomer, a pair that collides nearly head-on can hand enough energy to the
    // rotation to stay bound in an orbiting complex: by far the longest trajectories (and, in the
    // mixed-precision engine, the ones handed over to FP64 when they turn back).  Over the group,
    // the kFrontShare of the items with the smallest impact parameter therefore go first, each
    // tier longest-transit first.  On 65536 CO2-Ar trajectories at 200 K every orbiting one has
    // b^2 within the smallest 7 %.
    static const double share = [] {
        const char* v = std::getenv("HH_FRONT_SHARE");   // a share in [0, 1]; 0 turns the tier off
        const double s = v ? std::atof(v) : kFrontShare;
        return s >= 0.0 && s <= 1.0 ? s : kFrontShare;
    }();
    const long long n_front = pl.sys == HH_SYS_HEAR ? 0 : (long long)(share * (double)n);
    int what = ORDER_TRANSIT_KEY;
    if (n_front > 0) {
        for (PoolJob* j : group) {
            const long long off = j->jd.item_begin - base;
            HH_CHECK(run_order_keys(j->jd, pl.system.mu, ORDER_IMPACT_KEY, keys_in + off, idx_in + off, (int)off,
                                    e->stream));
        }
        HH_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, keys_in, keys_out, idx_in, order->as<int>(),
                                                    (int)n, 0, 32, e->stream));
        HH_CHECK(run_order_front(order->as<int>(), n, n_front, keys_in, e->stream));
        what = ORDER_TIERED_KEY;
    }
    for (PoolJob* j : group) {
        const long long off = j->jd.item_begin - base;
        HH_CHECK(run_order_keys(j->jd, pl.system.mu, what, keys_in + off, idx_in + off, (int)off, e->stream));
    }
    HH_CHECK(hipcub::DeviceRadixSort::SortPairsDescending(tmp, tmp_bytes, keys_in, keys_out, idx_in, order->as<int>(),
                                                          (int)n, 0, 32, e->stream));
    for (PoolJob* j : group) {
        j->group_order = order;
        j->jd.order = order->as<int>();
        j->jd.order_base = base;
        HH_CHECK(hipMemcpyAsync(pl.ring.as<JobDesc>() + j->seq % pl.ring_cap, &j->jd, sizeof(JobDesc),
                                hipMemcpyHostToDevice, e->stream));
    }
    return HH_OK;
}

// The graph-captured slices below are opt-in (HH_GRAPHS=1); the default path launches each
// slice directly.  A benchmark run with captured slices as the default (serving mix of
// lane-pool sessions and M0 requests) ended in an illegal-address fault whose cause was not
// found, so they stay off unless asked for.
namespace {

bool same_ctl(const PoolCtl& a, const PoolCtl& b)
{
    return a.next == b.next && a.total == b.total && a.first_seq == b.first_seq && a.last_seq == b.last_seq &&
           a.ticks == b.ticks && a.ring == b.ring;
}

// The slice of one pool as a graph launched on the engine stream; captured on the engine's
// capture stream the first time a (repack, buffer, admission state) combination is needed.
int launch_slice_graph(hh_engine* e, LanePool& pl, const PoolCtl& ctl, const SysConst& sc)
{
    if (!e->capture) HH_CHECK(hipStreamCreateWithFlags(&e->capture, hipStreamNonBlocking));
    if (!pl.h_done) HH_CHECK(hipHostMalloc((void**)&pl.h_done, pl.ring_cap * sizeof(int), hipHostMallocDefault));
    if (!pl.h_orbit) HH_CHECK(hipHostMalloc((void**)&pl.h_orbit, pl.ring_cap * sizeof(int), hipHostMallocDefault));
    if (!pl.h_next) HH_CHECK(hipHostMalloc((void**)&pl.h_next, sizeof(unsigned long long), hipHostMallocDefault));
    const bool repack = pl.repack_pending;
    LanePool::SliceGraph* g = nullptr;
    for (auto& c : pl.graphs)
        if (c.repack == repack && c.cur == pl.cur && same_ctl(c.ctl, ctl)) g = &c;
    if (!g) {
        if (pl.graphs.size() >= 4) {
            (void)hipGraphExecDestroy(pl.graphs.front().exec);
            pl.graphs.erase(pl.graphs.begin());
        }
        const int src = pl.cur, dst = repack ? 1 - pl.cur : pl.cur;
        hipGraph_t graph = nullptr;
        HH_CHECK(hipStreamBeginCapture(e->capture, hipStreamCaptureModeThreadLocal));
        hipError_t err = hipSuccess;
        if (repack) err = run_repack(pl.sys, pl.view(src), pl.view(dst), pl.scratch.as<int>(), e->capture);
        if (err == hipSuccess) err = run_clamp_next(pl.next.as<unsigned long long>(), pl.total, e->capture);
        if (err == hipSuccess)
            err = run_pool(pl.sys, pl.mixed, pl.block, pl.waves, sc, pl.ring.as<JobDesc>(), ctl, pl.view(dst),
                           pl.arena.as<double>(), e->capture);
        if (err == hipSuccess)
            err = hipMemcpyAsync(pl.h_done, pl.ring_done.p, pl.ring_cap * sizeof(int), hipMemcpyDeviceToHost,
                                 e->capture);
        if (err == hipSuccess)
            err = hipMemcpyAsync(pl.h_orbit, pl.ring_orbit.p, pl.ring_cap * sizeof(int), hipMemcpyDeviceToHost,
                                 e->capture);
        if (err == hipSuccess)
            err = hipMemcpyAsync(pl.h_next, pl.next.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, e->capture);
        const hipError_t end = hipStreamEndCapture(e->capture, &graph);
        if (err == hipSuccess) err = end;
        hipGraphExec_t exec = nullptr;
        if (err == hipSuccess) err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        HH_CHECK(err);
        LanePool::SliceGraph ng;
        ng.repack = repack;
        ng.cur = pl.cur;
        ng.ctl = ctl;
        ng.exec = exec;
        pl.graphs.push_back(ng);
        g = &pl.graphs.back();
    }
    HH_CHECK(hipGraphLaunch(g->exec, pl.stream));
    if (repack) {
        pl.cur = 1 - pl.cur;
        pl.repack_pending = false;
    }
    return HH_OK;
}

}  // namespace

namespace {

SysConst pool_sysconst(const hh_engine* e, const LanePool& pl) { return sysconst(e, pl.system, pl.mixed); }

// The items that `parent` handed over since its last orbit job, [parent.orbit_admitted, count)
// of its list, as one orbit job of the FP64 pool of the same class: the job shares the parent's
// buffers and maps its item k to the parent's item orbit_items[orbit_admitted + k].
int admit_orbit_job(hh_engine* e, PoolSet& set, LanePool& parent_pool, PoolJob& parent, int count)
{
    LanePool* op = parent.orbit_pool;
    if (!op) {
        const int rc = pool_get(e, set, parent_pool.system, false, &op);
        if (rc < 0) return rc;
        parent.orbit_pool = op;
    }
    if (!pool_has_room(*op)) return HH_OK;   // the ring is full: admitted after a later slice
    auto job = std::make_unique<PoolJob>();
    const long long n = count - parent.orbit_admitted;
    job->seq = op->next_seq;
    job->parent = &parent;
    const int slot = (int)(job->seq % op->ring_cap);
    JobDesc& jd = job->jd;
    jd = parent.jd;
    jd.item_begin = (long long)op->total;
    jd.n_items = n;
    jd.done = op->ring_done.as<int>() + slot;
    jd.order = nullptr;
    jd.order_base = 0;
    jd.map = parent.jd.orbit_items + parent.orbit_admitted;
    jd.orbit_items = nullptr;
    jd.orbit_count = nullptr;
    jd.orbit_nrhs = parent.jd.orbit_nrhs;
    jd.rtol = jd.atol = parent.orbit_rtol;
    HH_CHECK(hipMemcpyAsync(op->ring.as<JobDesc>() + slot, &jd, sizeof(JobDesc), hipMemcpyHostToDevice, e->stream));
    HH_CHECK(hipMemcpyAsync(jd.done, &kZero, sizeof(int), hipMemcpyHostToDevice, e->stream));
    HH_CHECK(run_clamp_next(op->next.as<unsigned long long>(), op->total, e->stream));
    op->total += (unsigned long long)n;
    op->next_seq++;
    op->unordered_seq = op->next_seq;   // orbit jobs run in hand-over order
    parent.orbit_admitted = count;
    parent.orbit_pending++;
    op->inflight.push_back(std::move(job));
    return HH_OK;
}

}  // namespace

int pool_step(hh_engine* e, PoolSet& set, double slice_ms, double* kernel_ms, double* reduce_ms, int* launches,
              bool* progress)
{
    if (progress) *progress = false;
    static const bool trace = std::getenv("HH_POOL_TRACE") != nullptr;   // progress lines on stderr
    // 1. one slice of every pool with items in flight, each pool on its own stream so the
    //    slices of different pools share the device
    std::vector<LanePool*> run;
    for (auto& pp : set) {
        bool work = false;
        for (auto& j : pp->inflight) work |= j->state == 0 && !j->items_done;
        if (work) run.push_back(pp.get());
    }
    if (!run.empty()) {
        HH_CHECK(hipEventRecord(e->ev[0], e->stream));
        for (LanePool* pl : run) {
            PoolCtl ctl{};
            ctl.next = pl->next.as<unsigned long long>();
            ctl.total = pl->total;
            ctl.first_seq = pl->inflight.front()->seq;
            ctl.last_seq = pl->inflight.back()->seq;
            ctl.ticks = (long long)(slice_ms * e->wall_khz);
            ctl.ring = pl->ring_cap;
            HH_CHECK(hipStreamWaitEvent(pl->stream, e->ev[0], 0));
            if (e->use_graphs) {
                const int rc = launch_slice_graph(e, *pl, ctl, pool_sysconst(e, *pl));
                if (rc != HH_OK) return rc;
            } else {
                HH_CHECK(run_pool(pl->sys, pl->mixed, pl->block, pl->waves, pool_sysconst(e, *pl),
                                  pl->ring.as<JobDesc>(), ctl, pl->view(pl->cur), pl->arena.as<double>(), pl->stream));
            }
            HH_CHECK(hipEventRecord(pl->slice_end, pl->stream));
            HH_CHECK(hipStreamWaitEvent(e->stream, pl->slice_end, 0));
            if (launches) ++*launches;
        }
        HH_CHECK(hipEventRecord(e->ev[1], e->stream));
        HH_CHECK(hipStreamSynchronize(e->stream));
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e->ev[0], e->ev[1]);
        if (kernel_ms) *kernel_ms += ms;
        for (LanePool* pl : run) {
            float pms = 0.0f;
            (void)hipEventElapsedTime(&pms, e->ev[0], pl->slice_end);
            pl->launch_ms += pms;
            pl->launch_count++;
        }
        if (trace) std::fprintf(stderr, "[hh pool] slice of %zu pools: %.2f ms\n", run.size(), ms);
    }

    // 2. progress of every job of the pools that ran: finished items, hand-overs to FP64
    struct Progress {
        LanePool* pl;
        unsigned long long next;
    };
    std::vector<Progress> prog;
    for (LanePool* pl : run) {
        std::vector<int> done(pl->ring_cap), orbit(pl->ring_cap);
        unsigned long long next = 0;
        if (e->use_graphs) {
            // the graph copied the counters into pinned memory at the end of the slice
            std::memcpy(done.data(), pl->h_done, pl->ring_cap * sizeof(int));
            std::memcpy(orbit.data(), pl->h_orbit, pl->ring_cap * sizeof(int));
            next = *pl->h_next;
        } else {
            HH_CHECK(hipMemcpy(done.data(), pl->ring_done.p, pl->ring_cap * sizeof(int), hipMemcpyDeviceToHost));
            HH_CHECK(hipMemcpy(orbit.data(), pl->ring_orbit.p, pl->ring_cap * sizeof(int), hipMemcpyDeviceToHost));
            HH_CHECK(hipMemcpy(&next, pl->next.p, sizeof next, hipMemcpyDeviceToHost));
        }
        // lanes that found the queue empty advanced the head past the end; none of those
        // indices was handed out (every admission clamps the device head back to the end)
        if (next > pl->total) next = pl->total;
        prog.push_back({pl, next});
        for (auto& jp : pl->inflight) {
            PoolJob& j = *jp;
            if (j.state != 0 || j.items_done) continue;
            const int slot = (int)(j.seq % pl->ring_cap);
            if (j.jd.orbit_count && orbit[slot] > j.orbit_admitted) {
                const int rc = admit_orbit_job(e, set, *pl, j, orbit[slot]);
                if (rc != HH_OK) return rc;
            }
            if (done[slot] >= j.jd.n_items) j.items_done = true;
        }
    }
    // a finished orbit job leaves its pool at once (its slot frees for the next hand-overs)
    for (auto& pp : set) {
        std::vector<PoolJob*> fin;
        for (auto& jp : pp->inflight)
            if (jp->parent && jp->items_done) fin.push_back(jp.get());
        for (PoolJob* o : fin) {
            o->parent->orbit_pending--;
            (void)pool_take(*pp, o);
        }
    }

    // 3. reductions: a job whose items have finished and whose orbit jobs have all finished;
    //    an orbit job is complete (state 2) once its items have finished
    bool launched = false;
    HH_CHECK(hipEventRecord(e->ev[2], e->stream));
    for (auto& pp : set) {
        LanePool& pl = *pp;
        for (auto& jp : pl.inflight) {
            PoolJob& j = *jp;
            if (j.state != 0 || !j.items_done || j.parent) continue;
            if (j.jd.orbit_count) {
                // the hand-over list is final once the items are done: admit what is left of it
                // (an orbit pool whose ring was full takes it after a later slice)
                int count = 0;
                HH_CHECK(hipMemcpy(&count, j.jd.orbit_count, sizeof count, hipMemcpyDeviceToHost));
                if (count > j.orbit_admitted) {
                    const int rc = admit_orbit_job(e, set, pl, j, count);
                    if (rc != HH_OK) return rc;
                }
                if (count != j.orbit_admitted || j.orbit_pending > 0) continue;
            }
            const int rc = reduce_job(e, j.jd, j.ppf, j.cf_out);
            if (rc != HH_OK) return rc;
            j.state = 1;
            launched = true;
        }
    }
    // 4. repacking: once every pool's queue is drained (the bulk of the work has been handed out),
    //    the items left in a pool are spread over its wavefronts
    bool all_drained = true;
    for (const Progress& g : prog) all_drained &= g.next >= g.pl->total;
    for (const Progress& g : prog) {
        LanePool* pl = g.pl;
        bool integrating = false;
        for (auto& j : pl->inflight) integrating |= j->state == 0 && !j->items_done;
        if (integrating && all_drained && e->use_graphs) {
            // the next slice's graph spreads them over the wavefronts
            pl->repack_pending = true;
        } else if (integrating && all_drained) {
            HH_CHECK(run_repack(pl->sys, pl->view(pl->cur), pl->view(1 - pl->cur), pl->scratch.as<int>(), e->stream));
            pl->cur = 1 - pl->cur;
            launched = true;
            if (trace) {
                int active = 0;
                HH_CHECK(hipMemcpy(&active, pl->scratch.as<int>() + pl->S / 256, sizeof(int), hipMemcpyDeviceToHost));
                std::fprintf(stderr, "[hh pool] sys %d %s: queue drained, %d items in flight\n", pl->sys,
                             pl->mixed ? "mixed" : "fp64", active);
            }
        } else if (trace && integrating) {
            std::fprintf(stderr, "[hh pool] sys %d %s: %llu of %llu items handed out\n", pl->sys,
                         pl->mixed ? "mixed" : "fp64", g.next, pl->total);
        }
    }
    HH_CHECK(hipEventRecord(e->ev[3], e->stream));
    if (launched) {
        HH_CHECK(hipStreamSynchronize(e->stream));
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e->ev[2], e->ev[3]);
        if (reduce_ms) *reduce_ms += ms;
    }
    // reduced jobs are final
    for (auto& pp : set)
        for (auto& jp : pp->inflight)
            if (jp->state == 1) {
                jp->state = 2;
                if (progress) *progress = true;
            }
    if (progress && !run.empty()) *progress = true;
    return HH_OK;
}

std::unique_ptr<PoolJob> pool_take(LanePool& pl, PoolJob* job)
{
    for (auto it = pl.inflight.begin(); it != pl.inflight.end(); ++it)
        if (it->get() == job) {
            std::unique_ptr<PoolJob> j = std::move(*it);
            pl.inflight.erase(it);
            return j;
        }
    return nullptr;
}

}  // namespace hh
