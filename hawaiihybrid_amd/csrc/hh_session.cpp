// Serving session: continuous batching of correlation-function and M0 requests on one
// device.  A worker thread repeats
//   admit waiting requests (FIFO) while their buffers fit the device-memory budget
//   -> one time slice of every lane pool with work (pool_step) and of every running M0
//   -> hand over the requests that finished.
// CF requests of one class (system, mu, I, precision) share a lane pool, so a request that
// arrives while others integrate takes lanes as soon as they free up; its reduction runs
// right after its own last trajectory ends.
#include <chrono>
#include <cstring>
#include <map>

#include "hh_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0;
    int kind = HH_REQ_CF;
    size_t bytes = 0;
    bool admitted = false;
    hh_request_info info{};
    // CF
    hh_system sys{};
    hh_cf_params p{};
    const double* ic = nullptr;
    long ntraj = 0;
    double* cf = nullptr;
    hh::PoolJob* job = nullptr;
    hh::LanePool* pool = nullptr;
    double kms0 = 0.0, rms0 = 0.0;
    // M0
    hh_sampler smp{};
    uint32_t seed = 0;
    long npoints = 0;
    int nstreams = 0;
    double ppf = 1.0;
    hh::DevBuf m0buf;
    bool m0_started = false;
};

}  // namespace

struct hh_session {
    hh_engine* e = nullptr;
    hh_session_opts opts{};
    std::mutex mu;
    std::condition_variable cv;
    std::thread worker;
    bool stop = false;
    hh::PoolSet pools;
    size_t budget = 0, in_use = 0;
    uint64_t next_id = 1;
    Clock::time_point t0;
    double kms = 0.0, rms = 0.0;   // device time of all slices / reductions so far
    hipStream_t m0_stream = nullptr;   // M0 slices, beside the pools' streams (default priority)
    std::deque<std::unique_ptr<Request>> waiting;
    std::vector<std::unique_ptr<Request>> running;
    std::map<uint64_t, std::unique_ptr<Request>> finished;

    double now_ms() const { return std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
    void run();
    int admit(Request& r);
    int slice(std::string& err);
    void finish(std::unique_ptr<Request> r, int state, const char* err);
};

namespace {
const int kZero = 0;
}

// engine lock held by the caller
int hh_session::admit(Request& r)
{
    if (r.kind == HH_REQ_CF) {
        int rc = hh::pool_get(e, pools, r.sys, r.p.precision == HH_PREC_MIXED, &r.pool);
        if (rc < 0) return rc;
        if (!hh::pool_has_room(*r.pool)) return HH_E_BUSY;
        r.kms0 = kms;
        r.rms0 = rms;
        return hh::pool_admit(e, *r.pool, r.p, r.ic, r.ntraj, r.cf, r.id, &r.job);
    }
    if (r.m0buf.ensure(hh::m0_bytes(r.nstreams)) != hipSuccess)
        return hh::fail(HH_E_NOMEM, "cannot allocate M0 stream state");
    const hh::M0Acc acc = hh::m0_view(r.m0buf.p, r.nstreams);
    if (!m0_stream) HH_CHECK(hipStreamCreateWithFlags(&m0_stream, hipStreamNonBlocking));
    HH_CHECK(hipMemcpyAsync(acc.finished, &kZero, sizeof(int), hipMemcpyHostToDevice, m0_stream));
    return HH_OK;
}

// engine lock held by the caller: one slice of every running M0 request on the M0 stream and of
// the pools on theirs, so the phase-space requests share the device with the trajectories
int hh_session::slice(std::string& err)
{
    bool m0 = false;
    for (auto& rp : running) {
        Request& r = *rp;
        if (r.kind != HH_REQ_M0) continue;
        const hh::M0Acc acc = hh::m0_view(r.m0buf.p, r.nstreams);
        const hh::SysConst sc = hh::sysconst(e, r.sys, false);
        hh::M0Params mp{r.smp.T, r.smp.Rmin, r.smp.Rmax, r.smp.pesmin, r.smp.pair_state, 0};
        const hipError_t he = hh::run_m0(r.sys.system, sc, mp, r.seed, r.nstreams, r.npoints,
                                         reinterpret_cast<uint32_t*>(r.m0buf.p), acc, r.m0_started ? 0 : 1,
                                         (long long)(opts.slice_ms * e->wall_khz), m0_stream);
        if (he != hipSuccess) {
            err = hipGetErrorString(he);
            return HH_E_DEVICE;
        }
        r.m0_started = true;
        m0 = true;
    }
    int rc = hh::pool_step(e, pools, opts.slice_ms, &kms, &rms);
    if (m0) {
        const hipError_t he = hipStreamSynchronize(m0_stream);
        if (he != hipSuccess && rc >= 0) {
            err = hipGetErrorString(he);
            return HH_E_DEVICE;
        }
    }
    if (rc < 0) {
        err = hh_last_error();
        return rc;
    }
    return HH_OK;
}

void hh_session::finish(std::unique_ptr<Request> r, int state, const char* err)
{
    r->info.state = state;
    r->info.done_ms = now_ms();
    r->info.latency_ms = r->info.done_ms - r->info.submit_ms;
    if (err) snprintf(r->info.error, sizeof r->info.error, "%s", err);
    if (r->admitted) in_use -= std::min(in_use, r->bytes);
    const uint64_t id = r->id;
    finished[id] = std::move(r);
}

void hh_session::run()
{
    (void)hipSetDevice(e->device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        // admission, first come first served, by device memory
        while (!waiting.empty()) {
            Request& r = *waiting.front();
            if (r.bytes > budget) {
                std::unique_ptr<Request> rr = std::move(waiting.front());
                waiting.pop_front();
                finish(std::move(rr), HH_REQ_FAILED, "request needs more device memory than the session budget");
                continue;
            }
            if (in_use + r.bytes > budget) break;
            int rc;
            {
                std::lock_guard<std::recursive_mutex> el(e->mu);
                rc = admit(r);
            }
            if (rc == HH_E_BUSY) break;   // ring full: wait for a job to finish
            std::unique_ptr<Request> rr = std::move(waiting.front());
            waiting.pop_front();
            if (rc < 0) {
                finish(std::move(rr), HH_REQ_FAILED, hh_last_error());
                continue;
            }
            rr->info.state = HH_REQ_RUNNING;
            rr->info.admit_ms = now_ms();
            rr->admitted = true;
            in_use += rr->bytes;
            running.push_back(std::move(rr));
        }
        {
            // the CF jobs admitted in this round form one group: longest-expected-first across it
            std::lock_guard<std::recursive_mutex> el(e->mu);
            for (auto& pl : pools)
                if (hh::pool_order_group(e, *pl) < 0) {
                    // ordering is an optimisation: on failure the items run in queue order
                    pl->unordered_seq = pl->next_seq;
                }
        }
        cv.notify_all();
        if (running.empty()) {
            if (stop && waiting.empty()) break;
            cv.wait(lk);
            continue;
        }
        lk.unlock();
        std::string err;
        int rc;
        {
            std::lock_guard<std::recursive_mutex> el(e->mu);
            rc = slice(err);
        }
        lk.lock();
        if (rc < 0) {
            // device error: every running request fails; the pools start afresh
            for (auto& r : running) finish(std::move(r), HH_REQ_FAILED, err.c_str());
            running.clear();
            std::lock_guard<std::recursive_mutex> el(e->mu);
            (void)hipStreamSynchronize(e->stream);
            pools.clear();
            continue;
        }
        // hand over finished requests
        std::vector<std::unique_ptr<Request>> still;
        for (auto& rp : running) {
            Request& r = *rp;
            bool done = false;
            int state = HH_REQ_DONE;
            std::string msg;
            std::lock_guard<std::recursive_mutex> el(e->mu);
            if (r.kind == HH_REQ_CF) {
                if (r.job->state == 2) {
                    done = true;
                    std::unique_ptr<hh::PoolJob> job = hh::pool_take(*r.pool, r.job);
                    r.job = nullptr;
                    hh_cf_stats& st = r.info.cf;
                    std::memset(&st, 0, sizeof st);
                    st.integrate_ms = kms - r.kms0;
                    st.reduce_ms = rms - r.rms0;
                    st.block = r.pool->block;
                    st.grid = (int)(r.pool->S / r.pool->block);
                    std::vector<hh::JobDesc> one{job->jd};
                    if (hh::collect_stats(one, &st) < 0) {
                        state = HH_REQ_FAILED;
                        msg = hh_last_error();
                    }
                    r.info.trajectories = r.ntraj;
                }
            } else {
                const hh::M0Acc acc = hh::m0_view(r.m0buf.p, r.nstreams);
                int fin = 0;
                if (hipMemcpy(&fin, acc.finished, sizeof fin, hipMemcpyDeviceToHost) != hipSuccess) {
                    done = true;
                    state = HH_REQ_FAILED;
                    msg = "reading M0 progress failed";
                } else if (fin >= r.nstreams) {
                    done = true;
                    if (hh::m0_combine(acc, r.npoints, r.nstreams, r.ppf, &r.info.m0, &r.info.m0_err) < 0) {
                        state = HH_REQ_FAILED;
                        msg = hh_last_error();
                    }
                    r.info.trajectories = r.npoints;
                    r.m0buf.release();
                }
            }
            if (done)
                finish(std::move(rp), state, msg.empty() ? nullptr : msg.c_str());
            else
                still.push_back(std::move(rp));
        }
        running = std::move(still);
        cv.notify_all();
    }
    cv.notify_all();
}

extern "C" {

int hh_session_create(hh_engine* e, const hh_session_opts* opts, hh_session** out)
{
    if (!e || !out) return hh::fail(HH_E_INVALID, "hh_session_create: null argument");
    *out = nullptr;
    auto s = std::make_unique<hh_session>();
    s->e = e;
    if (opts) s->opts = *opts;
    if (s->opts.slice_ms <= 0) s->opts.slice_ms = 10.0;
    if (s->opts.block != 0 && s->opts.block != hh::kPoolBlock)
        return hh::fail(HH_E_INVALID, "block must be 0 or 256 (the lane pools' one launch shape)");
    s->opts.block = hh::kPoolBlock;
    {
        std::lock_guard<std::recursive_mutex> el(e->mu);
        HH_CHECK(hipSetDevice(e->device));
        size_t free_b = 0, total_b = 0;
        HH_CHECK(hipMemGetInfo(&free_b, &total_b));
        s->budget = s->opts.max_device_gib > 0 ? (size_t)(s->opts.max_device_gib * 1073741824.0) : free_b / 10 * 8;
    }
    s->t0 = Clock::now();
    hh_session* raw = s.get();
    s->worker = std::thread([raw] { raw->run(); });
    *out = s.release();
    return HH_OK;
}

int hh_submit_cf(hh_session* s, const hh_system* sys, const hh_cf_params* p, const double* ic_dev, long ntraj,
                 double* cf_dev, uint64_t* id)
{
    if (!s || !sys || !p || !ic_dev || !cf_dev || ntraj <= 0) return hh::fail(HH_E_INVALID, "hh_submit_cf: bad argument");
    int rc = hh::check_cf_args(sys, p);
    if (rc < 0) return rc;
    if (p->engine != HH_ENGINE_BATCHED) return hh::fail(HH_E_INVALID, "serving runs the batched engine");
    auto r = std::make_unique<Request>();
    r->kind = HH_REQ_CF;
    r->sys = *sys;
    r->p = *p;
    r->ic = ic_dev;
    r->ntraj = ntraj;
    r->cf = cf_dev;
    r->bytes = hh::pool_job_bytes(*p, ntraj);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return hh::fail(HH_E_INVALID, "session is closing");
    r->id = s->next_id++;
    r->info.kind = HH_REQ_CF;
    r->info.state = HH_REQ_WAITING;
    r->info.submit_ms = s->now_ms();
    if (id) *id = r->id;
    s->waiting.push_back(std::move(r));
    s->cv.notify_all();
    return HH_OK;
}

int hh_submit_m0(hh_session* s, const hh_system* sys, const hh_sampler* smp, uint32_t seed, long npoints, int nstreams,
                 double ppf, uint64_t* id)
{
    if (!s || !sys || !smp) return hh::fail(HH_E_INVALID, "hh_submit_m0: null argument");
    int rc = hh::m0_check(sys, smp, npoints, nstreams);
    if (rc < 0) return rc;
    auto r = std::make_unique<Request>();
    r->kind = HH_REQ_M0;
    r->sys = *sys;
    r->smp = *smp;
    r->seed = seed;
    r->npoints = npoints;
    r->nstreams = nstreams;
    r->ppf = ppf;
    r->bytes = hh::m0_bytes(nstreams);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return hh::fail(HH_E_INVALID, "session is closing");
    r->id = s->next_id++;
    r->info.kind = HH_REQ_M0;
    r->info.state = HH_REQ_WAITING;
    r->info.submit_ms = s->now_ms();
    if (id) *id = r->id;
    s->waiting.push_back(std::move(r));
    s->cv.notify_all();
    return HH_OK;
}

int hh_poll(hh_session* s, uint64_t id, double timeout_ms, hh_request_info* info)
{
    if (!s) return hh::fail(HH_E_INVALID, "hh_poll: session is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    auto known = [&] {
        if (s->finished.count(id)) return true;
        for (auto& r : s->waiting)
            if (r->id == id) return true;
        for (auto& r : s->running)
            if (r->id == id) return true;
        return false;
    };
    if (!known()) return hh::fail(HH_E_NOTFOUND, "no request %llu", (unsigned long long)id);
    auto ready = [&] { return s->finished.count(id) > 0; };
    if (timeout_ms < 0)
        s->cv.wait(lk, ready);
    else
        s->cv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready);
    auto it = s->finished.find(id);
    if (it == s->finished.end()) return 0;
    if (info) *info = it->second->info;
    s->finished.erase(it);
    return 1;
}

int hh_session_destroy(hh_session* s)
{
    if (!s) return HH_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    {
        std::lock_guard<std::recursive_mutex> el(s->e->mu);
        (void)hipSetDevice(s->e->device);
        (void)hipStreamSynchronize(s->e->stream);
        s->pools.clear();
        if (s->m0_stream) {
            (void)hipStreamSynchronize(s->m0_stream);
            (void)hipStreamDestroy(s->m0_stream);
        }
    }
    delete s;
    return HH_OK;
}

}  // extern "C"
