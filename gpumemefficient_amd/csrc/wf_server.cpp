// wf_server.cpp -- continuous-batching alignment server (no counterpart in
// the reference, which runs one pair per process).
//
// A scheduler thread owns S launch slots (one HIP stream each, with its own
// grow-only device arenas and pinned result buffer).  Requests are queued by
// wf_server_submit; the scheduler
//   1. admits queued requests in arrival order while their device footprint
//      (codes + strip-boundary rows + descriptors) fits the HBM budget,
//   2. whenever a slot is free, launches every admitted request not yet on
//      the device (up to max_batch) as one persistent plan on that slot's
//      stream: upload, reset, engine, result copy, event,
//   3. polls the slots' events, publishes completions and returns the
//      requests' bytes to the budget.
// Requests that arrive while others run are admitted and launched on the
// next free slot, so their work overlaps on the device with the launches
// already running.  Each launch's grid is capped at 1/S of the resident
// capacity so a large request cannot starve the other slots.
//
// wf_server_replay drives the server with a timed request trace from the
// calling thread (the serving benchmark's request loop).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/wfdp.h"
#include "wf_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket;
    int32_t problem, flags;
    std::vector<int32_t> a, b;
    int64_t bytes;      // device footprint charged to the HBM budget
    bool bitpar;        // can run on the bit-parallel engine (LCS/ED, <= 32 symbols)
    double t_submit;
    double t_start = 0; // launch time
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    WfArenas ar;
    WfPlanHost ph;
    std::vector<Request> reqs;
    WfResultDev *pinned = nullptr;
    int64_t pinned_cap = 0;
    bool busy = false;
};

}  // namespace

struct wf_server {
    wf_ctx *ctx = nullptr;
    wf_server_opts opts{};
    std::vector<Slot> slots;
    std::thread sched;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request> pending;           // submitted, not admitted
    std::deque<Request> admitted;          // admitted, not launched
    std::vector<wf_completion> done;       // completed, not yet polled
    int64_t budget = 0, used = 0;
    uint64_t next_ticket = 1;
    int64_t inflight = 0;                  // admitted or running requests
    std::atomic<bool> stop{false};
    Clock::time_point t0;
    int error = WF_OK;
    // statistics
    uint64_t launches = 0, completed = 0, cells = 0;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

static int64_t request_bytes(int64_t n1, int64_t n2, int strip_rows)
{
    const int64_t strips = n2 > 0 ? (n2 + strip_rows - 1) / strip_rows : 0;
    return ((n1 + 15) & ~15LL) + ((n2 + 15) & ~15LL) + std::max<int64_t>(0, strips - 1) * n1 * 8 +
           (int64_t)sizeof(WfPairDev) + strips * (int64_t)sizeof(WfTaskDev) + (int64_t)sizeof(WfResultDev);
}

// At most 256 distinct symbols: every engine encodes a pair's symbols to
// one-byte codes (wf_encode_pair).  Checked at submission so that one bad
// request never reaches a launch shared with others.
static bool alphabet_fits(const wf_job *job)
{
    bool small[256] = {};
    std::vector<int32_t> other;
    int distinct = 0;
    auto see = [&](int32_t v) {
        if ((uint32_t)v < 256u) {
            if (!small[v]) {
                small[v] = true;
                ++distinct;
            }
        } else {
            other.push_back(v);
        }
    };
    for (int64_t i = 0; i < job->n1; ++i) see(job->a[i]);
    for (int64_t i = 0; i < job->n2; ++i) see(job->b[i]);
    if (!other.empty()) {
        std::sort(other.begin(), other.end());
        distinct += (int)(std::unique(other.begin(), other.end()) - other.begin());
    }
    return distinct <= 256;
}

static void finish_slot(wf_server *s, Slot &sl)
{
    std::vector<wf_completion> out;
    out.reserve(sl.reqs.size());
    const double t = s->now();
    for (size_t i = 0; i < sl.reqs.size(); ++i) {
        const Request &r = sl.reqs[i];
        wf_completion c{};
        c.ticket = r.ticket;
        if (sl.ph.host_only[i]) {
            c.last = sl.ph.host_result[i];
        } else {
            c.last = sl.pinned[i].last;
            c.max_score = sl.pinned[i].maxs;
            c.status = sl.pinned[i].status ? WF_ERR_TIMEOUT : WF_OK;
        }
        c.t_submit = r.t_submit;
        c.t_start = r.t_start;
        c.t_done = t;
        out.push_back(c);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    for (size_t i = 0; i < sl.reqs.size(); ++i) {
        s->used -= sl.reqs[i].bytes;
        s->cells += (uint64_t)(sl.reqs[i].a.size() * sl.reqs[i].b.size());
    }
    s->inflight -= (int64_t)sl.reqs.size();
    s->completed += sl.reqs.size();
    s->done.insert(s->done.end(), out.begin(), out.end());
    sl.reqs.clear();
    sl.busy = false;
    s->cv.notify_all();
}

static int launch_slot(wf_server *s, Slot &sl, std::vector<Request> &&batch)
{
    sl.reqs = std::move(batch);
    std::vector<wf_job> jobs(sl.reqs.size());
    for (size_t i = 0; i < sl.reqs.size(); ++i) {
        const Request &r = sl.reqs[i];
        jobs[i] = wf_job{r.problem, r.flags, r.a.data(), (int64_t)r.a.size(), r.b.data(), (int64_t)r.b.size()};
    }
    sl.ph = WfPlanHost();
    int st = wf_build_plan(s->ctx, jobs.data(), (int64_t)jobs.size(), nullptr, &sl.ph);
    if (st != WF_OK) return st;
    const int nslots = (int)s->slots.size();
    auto share = [&](const WfVariant *v) {
        return std::max(1, wf_max_resident_groups(s->ctx, v) / std::max(1, nslots));
    };
    // The asynchronous engine only pays off when all strips of the launch are
    // resident within the slot's share of the chip; otherwise the lock-step
    // engine's higher occupancy wins.
    if (sl.ph.variant->engine == WF_ENGINE_ASYNC && (int64_t)sl.ph.tasks.size() > share(sl.ph.variant)) {
        const wf_tile lock{0, 0, 0, WF_TILE_LOCKSTEP};
        sl.ph = WfPlanHost();
        if ((st = wf_build_plan(s->ctx, jobs.data(), (int64_t)jobs.size(), &lock, &sl.ph)) != WF_OK) return st;
    }
    if ((st = wf_upload_plan_async(s->ctx, &sl.ph, &sl.ar, sl.stream, false)) != WF_OK) return st;
    const int maxg = share(sl.ph.variant);
    if ((st = wf_enqueue_plan(s->ctx, &sl.ph, &sl.ar, WF_RUN_PERSISTENT, maxg, sl.stream)) != WF_OK) return st;
    const int64_t n = sl.ph.npairs;
    if (n > sl.pinned_cap) {
        if (sl.pinned) (void)hipHostFree(sl.pinned);
        sl.pinned = nullptr;
        if (hipHostMalloc((void **)&sl.pinned, n * sizeof(WfResultDev), hipHostMallocDefault) != hipSuccess)
            return WF_ERR_NOMEM;
        sl.pinned_cap = n;
    }
    if (hipMemcpyAsync(sl.pinned, sl.ar.res.ptr, n * sizeof(WfResultDev), hipMemcpyDeviceToHost, sl.stream) !=
        hipSuccess)
        return WF_ERR_DEVICE;
    if (hipEventRecord(sl.done, sl.stream) != hipSuccess) return WF_ERR_DEVICE;
    sl.busy = true;
    s->launches++;
    return WF_OK;
}

static void scheduler(wf_server *s)
{
    (void)hipSetDevice(s->ctx->device);
    const int strip_rows = 512;
    while (true) {
        bool did = false;
        // 1. completions
        for (Slot &sl : s->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipSuccess) {
                finish_slot(s, sl);
                did = true;
            } else if (q != hipErrorNotReady) {
                std::lock_guard<std::mutex> lk(s->mu);
                s->error = WF_ERR_DEVICE;
            }
        }
        // 2. admission by HBM budget, in arrival order
        {
            std::lock_guard<std::mutex> lk(s->mu);
            while (!s->pending.empty()) {
                Request &r = s->pending.front();
                if (r.bytes > s->budget) {
                    // Can never fit: fail it instead of blocking the queue.
                    wf_completion c{};
                    c.ticket = r.ticket;
                    c.status = WF_ERR_NOMEM;
                    c.t_submit = r.t_submit;
                    c.t_done = s->now();
                    s->done.push_back(c);
                    s->inflight--;
                    s->pending.pop_front();
                    did = true;
                    continue;
                }
                if (s->used + r.bytes > s->budget) break;
                s->used += r.bytes;
                s->admitted.push_back(std::move(r));
                s->pending.pop_front();
                did = true;
            }
        }
        // 3. launch admitted requests on free slots
        for (Slot &sl : s->slots) {
            if (sl.busy) continue;
            std::vector<Request> batch;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                const int maxb = s->opts.per_request ? 1 : std::max(1, s->opts.max_batch);
                // One launch runs one engine: take the oldest request and the
                // admitted requests of its class (bit-parallel LCS/ED, or the
                // rest), in arrival order; the other class waits for the
                // next free slot.
                if (!s->admitted.empty()) {
                    const bool cls = s->admitted.front().bitpar;
                    for (auto it = s->admitted.begin(); it != s->admitted.end() && (int)batch.size() < maxb;) {
                        if (it->bitpar == cls) {
                            batch.push_back(std::move(*it));
                            it = s->admitted.erase(it);
                        } else {
                            ++it;
                        }
                    }
                }
            }
            if (batch.empty()) break;
            const double t = s->now();
            for (Request &r : batch) r.t_start = t;
            const int st = launch_slot(s, sl, std::move(batch));
            if (st != WF_OK) {
                // The launch's requests fail with its status; only a HIP
                // runtime error stops the server (requests are validated at
                // submission, so a plan cannot fail for one of them).
                std::lock_guard<std::mutex> lk(s->mu);
                if (st == WF_ERR_DEVICE) s->error = st;
                for (Request &r : sl.reqs) {
                    wf_completion c{};
                    c.ticket = r.ticket;
                    c.status = st;
                    c.t_submit = r.t_submit;
                    c.t_done = s->now();
                    s->done.push_back(c);
                    s->used -= r.bytes;
                    s->inflight--;
                }
                sl.reqs.clear();
                sl.busy = false;
            }
            did = true;
        }
        if (s->stop.load()) {
            bool idle;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                idle = s->pending.empty() && s->admitted.empty();
            }
            bool any_busy = false;
            for (Slot &sl : s->slots) any_busy |= sl.busy;
            if (idle && !any_busy) break;
        }
        if (!did) {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait_for(lk, std::chrono::microseconds(20));
        }
        (void)strip_rows;
    }
}

extern "C" {

int wf_server_create(int device, const wf_server_opts *opts, wf_server **out)
{
    if (!out) return WF_ERR_INVALID;
    *out = nullptr;
    wf_server *s = new wf_server();
    int st = wf_ctx_create(device, &s->ctx);
    if (st != WF_OK) {
        delete s;
        return st;
    }
    if (opts) s->opts = *opts;
    if (s->opts.streams <= 0) s->opts.streams = 4;
    if (s->opts.max_batch <= 0) s->opts.max_batch = 512;
    if (s->opts.queue_capacity <= 0) s->opts.queue_capacity = 1 << 20;
    if (s->opts.per_request) s->opts.streams = 1;
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    s->budget = s->opts.hbm_budget > 0 ? s->opts.hbm_budget : (int64_t)(fr / 2);
    s->slots.resize(s->opts.streams);
    for (Slot &sl : s->slots) {
        if (hipStreamCreate(&sl.stream) != hipSuccess || hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) !=
                                                             hipSuccess) {
            wf_server_destroy(s);
            return WF_ERR_DEVICE;
        }
    }
    s->t0 = Clock::now();
    s->sched = std::thread(scheduler, s);
    *out = s;
    return WF_OK;
}

int wf_server_submit(wf_server *s, const wf_job *job, uint64_t *ticket)
{
    if (!s || !job || job->problem < WF_LCS || job->problem > WF_SW || job->n1 < 0 || job->n2 < 0 ||
        (job->n1 > 0 && !job->a) || (job->n2 > 0 && !job->b) || job->n1 > (1LL << 30) || job->n2 > (1LL << 30))
        return WF_ERR_INVALID;
    if (!alphabet_fits(job)) return WF_ERR_ALPHABET;
    Request r;
    r.problem = job->problem;
    r.flags = job->flags;
    r.a.assign(job->a, job->a + job->n1);
    r.b.assign(job->b, job->b + job->n2);
    r.bytes = request_bytes(job->n1, job->n2, 512);
    r.bitpar = wf_bitpar_eligible(job, 1);
    std::lock_guard<std::mutex> lk(s->mu);
    if ((int64_t)s->pending.size() >= s->opts.queue_capacity) return WF_ERR_AGAIN;
    r.ticket = s->next_ticket++;
    r.t_submit = s->now();
    if (ticket) *ticket = r.ticket;
    s->pending.push_back(std::move(r));
    s->inflight++;
    s->cv.notify_all();
    return WF_OK;
}

int wf_server_poll(wf_server *s, wf_completion *out, int max, int *n)
{
    if (!s || !n || (max > 0 && !out)) return WF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    const int k = (int)std::min<size_t>((size_t)std::max(0, max), s->done.size());
    for (int i = 0; i < k; ++i) out[i] = s->done[i];
    s->done.erase(s->done.begin(), s->done.begin() + k);
    *n = k;
    return s->error;
}

int wf_server_wait_idle(wf_server *s, double timeout_s)
{
    if (!s) return WF_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    const bool ok = s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s),
                                   [&] { return s->inflight == 0 || s->error != WF_OK; });
    if (s->error != WF_OK) return s->error;
    return ok ? WF_OK : WF_ERR_TIMEOUT;
}

int wf_server_stats(wf_server *s, wf_server_stats_t *st)
{
    if (!s || !st) return WF_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    st->launches = s->launches;
    st->completed = s->completed;
    st->cells = s->cells;
    st->hbm_budget = s->budget;
    st->hbm_used = s->used;
    st->pending = (int64_t)s->pending.size();
    return WF_OK;
}

void wf_server_destroy(wf_server *s)
{
    if (!s) return;
    s->stop.store(true);
    s->cv.notify_all();
    if (s->sched.joinable()) s->sched.join();
    if (s->ctx) (void)hipSetDevice(s->ctx->device);
    for (Slot &sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        sl.ar.release();
        if (sl.pinned) (void)hipHostFree(sl.pinned);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    if (s->ctx) wf_ctx_destroy(s->ctx);
    delete s;
}

int wf_server_replay(wf_server *s, const wf_job *jobs, const double *arrival_s, int64_t n, wf_completion *out)
{
    if (!s || (n > 0 && (!jobs || !arrival_s || !out))) return WF_ERR_INVALID;
    std::vector<int64_t> index_of_ticket;
    std::vector<uint64_t> tickets((size_t)n);
    const double base = s->now();
    std::vector<wf_completion> got;
    got.reserve((size_t)n);
    std::vector<wf_completion> buf(1024);
    int64_t next = 0;
    while ((int64_t)got.size() < n) {
        // Submit everything whose arrival time has passed.
        const double t = s->now() - base;
        while (next < n && arrival_s[next] <= t) {
            int st;
            while ((st = wf_server_submit(s, &jobs[next], &tickets[next])) == WF_ERR_AGAIN)
                std::this_thread::yield();
            if (st != WF_OK) return st;
            ++next;
        }
        int k = 0;
        const int e = wf_server_poll(s, buf.data(), (int)buf.size(), &k);
        if (e != WF_OK) return e;
        got.insert(got.end(), buf.begin(), buf.begin() + k);
        if (k == 0) {
            if (next < n) {
                const double wait = arrival_s[next] - (s->now() - base);
                if (wait > 50e-6) std::this_thread::sleep_for(std::chrono::duration<double>(wait - 40e-6));
            } else {
                std::this_thread::sleep_for(std::chrono::microseconds(10));
            }
        }
    }
    // Completions in submission order; times relative to the trace start,
    // with t_submit replaced by the scheduled arrival.
    const uint64_t t0 = n > 0 ? tickets[0] : 0;
    for (const wf_completion &c : got) {
        const int64_t i = (int64_t)(c.ticket - t0);
        if (i < 0 || i >= n) continue;
        out[i] = c;
        out[i].t_submit = arrival_s[i];
        out[i].t_start -= base;
        out[i].t_done -= base;
    }
    return WF_OK;
}

}  // extern "C"
