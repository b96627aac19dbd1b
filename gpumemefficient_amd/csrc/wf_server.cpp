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
// Resident mode (wf_server_opts.resident) replaces the slots with one
// persistent launch of the lock-step engine that lives as long as the
// server (wf_serve_engine, wf_kernels.hip).  The scheduler thread admits
// requests by the same HBM budget (a range of one device arena per request:
// codes and strip boundaries), encodes their codes into pinned staging
// memory and publishes a submission entry; the kernel takes it in, queues
// its strips in a task ring in HBM and runs them as workgroups free up, so a
// request that arrives while others run starts within microseconds, with no
// launch of its own.  Completions come back through a pinned ring.
//
// wf_server_replay drives the server with a timed request trace from the
// calling thread (the serving benchmark's request loop).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <cstdio>
#include <deque>
#include <map>
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

// First-fit allocator of byte ranges in [0, capacity), 256-byte granules.
class RangeAlloc {
  public:
    void reset(int64_t capacity)
    {
        free_.clear();
        if (capacity > 0) free_[0] = capacity;
    }
    int64_t alloc(int64_t bytes)
    {
        bytes = (std::max<int64_t>(bytes, 1) + 255) & ~int64_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const int64_t off = it->first, left = it->second - bytes;
            free_.erase(it);
            if (left > 0) free_[off + bytes] = left;
            return off;
        }
        return -1;
    }
    void release(int64_t off, int64_t bytes)
    {
        bytes = (std::max<int64_t>(bytes, 1) + 255) & ~int64_t(255);
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

  private:
    std::map<int64_t, int64_t> free_;
};

// A request on the resident engine, from submission to completion.
struct ResidentReq {
    Request req;
    int64_t codes_off = -1, codes_bytes = 0;  // codes in the device code arena
    int64_t bnd_off = -1, bnd_bytes = 0;      // strip boundaries in the device boundary arena
    int64_t stage_off = -1, stage_bytes = 0;  // codes in pinned staging memory
    int32_t ngroups = 0;
    bool nomax = false;  // SW without the running maximum: report max_score -1
};

// Resident mode: the persistent launch and everything it reads and writes.
struct Resident {
    static constexpr int64_t kSlots = 8192;        // requests in flight
    static constexpr int64_t kTasks = 1LL << 20;   // strips in flight
    static constexpr int64_t kStage = 64LL << 20;  // pinned staging bytes
    static constexpr int kStripRows = 512;         // lock-step <2, 4, 64>
    hipStream_t stream = nullptr;
    // Codes and strip boundaries live in separate arenas: the boundary arena
    // only ever holds zeros or words tagged by some request, and every
    // request has a tag of its own, so a recycled range never shows a stale
    // word with the current request's tag.
    DevBuf codes, bnd, pairs, res, done, taskq, ctl;
    WfSubmitDev *subq = nullptr;  // pinned, device-mapped (same addresses on this platform)
    WfCompDev *compq = nullptr;
    WfHostCtl *hctl = nullptr;
    uint8_t *staging = nullptr;
    RangeAlloc codes_alloc, bnd_alloc, stage_alloc;
    int64_t codes_cap = 0, bnd_cap = 0;
    std::vector<int32_t> free_slots;
    std::vector<ResidentReq> inflight;            // by slot
    uint64_t next_sub = 0, next_task = 0, next_comp = 0;
    uint32_t next_tag = 1;
    int64_t tasks_in_flight = 0;
    int grid = 0;
    bool launched = false;
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
    Resident *rs = nullptr;  // resident mode

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
            c.max_score = (sl.ph.pairs[i].mode & WFK_MODE_NOMAX) ? -1 : sl.pinned[i].maxs;
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
    const int nslots = (int)s->slots.size();
    auto share = [&](const WfVariant *v) {
        return std::max(1, wf_max_resident_groups(s->ctx, v) / std::max(1, nslots));
    };
    // The variant's share bounds an SW column split (the split's strips must
    // fit in the launch's workgroups, or the launch would go to the lock-step
    // engine below).
    sl.ph = WfPlanHost();
    int st = wf_build_plan(s->ctx, jobs.data(), (int64_t)jobs.size(), nullptr, &sl.ph);
    if (st != WF_OK) return st;
    if (sl.ph.variant->engine == WF_ENGINE_ASYNC && !sl.ph.splits.empty() &&
        (int64_t)sl.ph.tasks.size() > share(sl.ph.variant)) {
        const int64_t cap = share(sl.ph.variant);
        sl.ph = WfPlanHost();
        if ((st = wf_build_plan(s->ctx, jobs.data(), (int64_t)jobs.size(), nullptr, &sl.ph, cap)) != WF_OK)
            return st;
    }
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

// ---------------------------------------------------------------------------
// Resident mode
// ---------------------------------------------------------------------------

static int64_t round16(int64_t v) { return (v + 15) & ~int64_t(15); }

// Pinned host memory the device reads and writes coherently (system-scope
// atomics on both sides); on this platform its device address is its host
// address.
template <class T>
static T *pinned(size_t bytes)
{
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    std::memset(p, 0, bytes);
    return static_cast<T *>(p);
}

template <class T>
static T *device_view(T *host)
{
    void *d = nullptr;
    if (hipHostGetDevicePointer(&d, host, 0) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return static_cast<T *>(d);
}

static void resident_free(Resident *r)
{
    if (r->subq) (void)hipHostFree(r->subq);
    if (r->compq) (void)hipHostFree(r->compq);
    if (r->hctl) (void)hipHostFree(r->hctl);
    if (r->staging) (void)hipHostFree(r->staging);
    r->codes.release();
    r->bnd.release();
    r->pairs.release();
    r->res.release();
    r->done.release();
    r->taskq.release();
    r->ctl.release();
    if (r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
}

static int resident_create(wf_server *s)
{
    Resident *r = new Resident();
    s->rs = r;
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) return WF_ERR_DEVICE;
    // Device arenas within the HBM budget (capped at 16 GiB): 1/16 for codes.
    const int64_t total = std::min<int64_t>(s->budget, 16LL << 30);
    r->codes_cap = std::max<int64_t>(1 << 20, (total / 16) & ~int64_t(255));
    r->bnd_cap = std::max<int64_t>(1 << 20, (total - r->codes_cap) & ~int64_t(255));
    int st;
    if ((st = r->codes.ensure(r->codes_cap)) != WF_OK || (st = r->bnd.ensure(r->bnd_cap, true)) != WF_OK ||
        (st = r->pairs.ensure(Resident::kSlots * sizeof(WfPairDev))) != WF_OK ||
        (st = r->res.ensure(Resident::kSlots * sizeof(WfResultDev))) != WF_OK ||
        (st = r->done.ensure(Resident::kSlots * sizeof(int32_t))) != WF_OK ||
        (st = r->taskq.ensure(Resident::kTasks * sizeof(WfTaskQDev), true)) != WF_OK ||
        (st = r->ctl.ensure(sizeof(WfServeCtl), true)) != WF_OK)
        return st;
    r->subq = pinned<WfSubmitDev>(Resident::kSlots * sizeof(WfSubmitDev));
    r->compq = pinned<WfCompDev>(Resident::kSlots * sizeof(WfCompDev));
    r->hctl = pinned<WfHostCtl>(sizeof(WfHostCtl));
    r->staging = pinned<uint8_t>(Resident::kStage);
    if (!r->subq || !r->compq || !r->hctl || !r->staging) return WF_ERR_NOMEM;
    r->codes_alloc.reset(r->codes_cap);
    r->bnd_alloc.reset(r->bnd_cap);
    r->stage_alloc.reset(Resident::kStage);
    r->inflight.resize(Resident::kSlots);
    for (int64_t i = Resident::kSlots - 1; i >= 0; --i) r->free_slots.push_back((int32_t)i);

    int block = 0;
    const void *fn = wf_serve_kernel(&block);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, block, 0) != hipSuccess || per_cu < 1) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    // Every workgroup of the launch must be resident at once: a strip waits
    // only on strips claimed before it, by running workgroups.
    r->grid = per_cu * s->ctx->prop.multiProcessorCount;
    WfServeArgs sa{};
    sa.ka.pairs = (const WfPairDev *)r->pairs.ptr;
    sa.ka.codes = (const uint8_t *)r->codes.ptr;
    sa.ka.bnd = (uint64_t *)r->bnd.ptr;
    sa.ka.res = (WfResultDev *)r->res.ptr;
    sa.ka.timeout_ticks = (int64_t)(s->ctx->timeout_s * 1.0e8);
    sa.taskq = (WfTaskQDev *)r->taskq.ptr;
    sa.taskq_cap = Resident::kTasks;
    sa.subq = device_view(r->subq);
    sa.subq_cap = Resident::kSlots;
    sa.staging = device_view(r->staging);
    sa.compq = device_view(r->compq);
    sa.compq_cap = Resident::kSlots;
    sa.ctl = (WfServeCtl *)r->ctl.ptr;
    sa.hctl = device_view(r->hctl);
    sa.pair_done = (int32_t *)r->done.ptr;
    if (!sa.subq || !sa.staging || !sa.compq || !sa.hctl) return WF_ERR_DEVICE;
    if (hipDeviceSynchronize() != hipSuccess) return WF_ERR_DEVICE;  // arenas cleared
    if (wf_launch_serve(r->grid, &sa, r->stream) != hipSuccess) return WF_ERR_DEVICE;
    r->launched = true;
    s->launches = 1;
    return WF_OK;
}

// Push a completion (caller holds s->mu).
static void push_done(wf_server *s, const Request &q, int32_t last, int32_t maxs, int32_t status, double t)
{
    wf_completion c{};
    c.ticket = q.ticket;
    c.last = last;
    c.max_score = maxs;
    c.status = status;
    c.t_submit = q.t_submit;
    c.t_start = q.t_start;
    c.t_done = t;
    s->done.push_back(c);
    s->inflight--;
    s->completed++;
    s->cells += (uint64_t)(q.a.size() * q.b.size());
}

static void resident_scheduler(wf_server *s)
{
    Resident *r = s->rs;
    (void)hipSetDevice(s->ctx->device);
    int64_t running = 0;  // published, not completed
    while (true) {
        bool did = false;
        // 1. completions, in the order the device finished them
        for (;;) {
            WfCompDev *c = r->compq + (r->next_comp % Resident::kSlots);
            if (__atomic_load_n(&c->seq, __ATOMIC_ACQUIRE) != r->next_comp + 1) break;
            const int32_t slot = c->slot;
            ResidentReq &q = r->inflight[slot];
            const int32_t st = c->status ? WF_ERR_TIMEOUT : WF_OK;
            r->codes_alloc.release(q.codes_off, q.codes_bytes);
            if (q.bnd_bytes) r->bnd_alloc.release(q.bnd_off, q.bnd_bytes);
            r->stage_alloc.release(q.stage_off, q.stage_bytes);
            r->tasks_in_flight -= q.ngroups;
            r->free_slots.push_back(slot);
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->used -= q.req.bytes;
                const bool nomax = q.req.problem == WF_SW && (q.nomax);
                push_done(s, q.req, c->last, nomax ? -1 : c->maxs, st, s->now());
                s->cv.notify_all();
            }
            q.req = Request();
            r->next_comp++;
            --running;
            did = true;
        }
        // 2. admission by HBM budget, in arrival order, and publication
        for (;;) {
            ResidentReq q;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                if (s->pending.empty()) break;
                Request &f = s->pending.front();
                const int64_t n1 = (int64_t)f.a.size(), n2 = (int64_t)f.b.size();
                if (n1 == 0 || n2 == 0) {
                    // Nothing for the device: the boundary is the answer.
                    f.t_start = s->now();
                    push_done(s, f, wf_empty_score(f.problem, n1, n2, f.flags), 0, WF_OK, s->now());
                    s->pending.pop_front();
                    s->cv.notify_all();
                    did = true;
                    continue;
                }
                const int32_t ngroups = (int32_t)((n2 + Resident::kStripRows - 1) / Resident::kStripRows);
                const int64_t cb = round16(n1) + round16(n2), bb = (int64_t)(ngroups - 1) * n1 * 8;
                if (f.bytes > s->budget || cb > r->codes_cap || bb > r->bnd_cap || cb > Resident::kStage ||
                    ngroups > Resident::kTasks) {
                    f.t_start = s->now();
                    push_done(s, f, 0, 0, WF_ERR_NOMEM, s->now());
                    s->pending.pop_front();
                    s->cv.notify_all();
                    did = true;
                    continue;
                }
                if (s->used + f.bytes > s->budget || r->free_slots.empty() ||
                    r->tasks_in_flight + ngroups > Resident::kTasks)
                    break;
                q.codes_off = r->codes_alloc.alloc(cb);
                q.bnd_off = bb ? r->bnd_alloc.alloc(bb) : 0;
                q.stage_off = r->stage_alloc.alloc(cb);
                if (q.codes_off < 0 || q.bnd_off < 0 || q.stage_off < 0) {
                    if (q.codes_off >= 0) r->codes_alloc.release(q.codes_off, cb);
                    if (bb && q.bnd_off >= 0) r->bnd_alloc.release(q.bnd_off, bb);
                    if (q.stage_off >= 0) r->stage_alloc.release(q.stage_off, cb);
                    break;  // wait for completions to free room
                }
                q.codes_bytes = q.stage_bytes = cb;
                q.bnd_bytes = bb;
                q.ngroups = ngroups;
                s->used += f.bytes;
                q.req = std::move(f);
                s->pending.pop_front();
            }
            // Encode into staging and publish (no lock: only this thread
            // touches the rings and allocators).
            const int64_t n1 = (int64_t)q.req.a.size(), n2 = (int64_t)q.req.b.size();
            const wf_job j{q.req.problem, q.req.flags, q.req.a.data(), n1, q.req.b.data(), n2};
            uint8_t *sa = r->staging + q.stage_off;
            int32_t mode = 0;
            const int est = wf_encode_job(j, 0, wf_perm4_enabled(), sa, sa + round16(n1), &mode);
            // A job's no-maximum flag, as in a plan (the kernel has the
            // instances without the running maximum).
            if (est == WF_OK && j.problem == WF_SW && !(mode & WFK_MODE_CMP) && (j.flags & WF_TILE_NO_MAX))
                mode |= WFK_MODE_NOMAX;
            const int32_t slot = r->free_slots.back();
            if (est != WF_OK) {
                r->codes_alloc.release(q.codes_off, q.codes_bytes);
                if (q.bnd_bytes) r->bnd_alloc.release(q.bnd_off, q.bnd_bytes);
                r->stage_alloc.release(q.stage_off, q.stage_bytes);
                std::lock_guard<std::mutex> lk(s->mu);
                s->used -= q.req.bytes;
                q.req.t_start = s->now();
                push_done(s, q.req, 0, 0, est, s->now());
                s->cv.notify_all();
                did = true;
                continue;
            }
            r->free_slots.pop_back();
            WfPairDev pd{};
            pd.a_off = q.codes_off;
            pd.b_off = q.codes_off + round16(n1);
            pd.bnd_off = q.bnd_off / 8;
            pd.n1 = (int32_t)n1;
            pd.n2 = (int32_t)n2;
            pd.prob = q.req.problem;
            pd.mode = mode;
            q.nomax = (mode & WFK_MODE_NOMAX) != 0;
            // Longer dependency chains (columns) issue first, as in a plan
            // launch, so that a long pair's strips are not starved by the
            // short ones that share their SIMDs.
            pd.prio = n1 >= 16384 ? 3 : n1 >= 4096 ? 2 : n1 >= 1024 ? 1 : 0;
            pd.ngroups = q.ngroups;
            pd.col_off = -1;
            pd.tag = 0x80000000u | (r->next_tag++ & 0x7fffffffu);
            if (r->next_tag > 0x7fffffffu) r->next_tag = 1;
            WfSubmitDev *e = r->subq + (r->next_sub % Resident::kSlots);
            e->pd = pd;
            e->task_base = (int64_t)r->next_task;
            e->stage_off = q.stage_off;
            e->code_bytes = q.codes_bytes;
            e->slot = slot;
            __atomic_store_n(&e->seq, r->next_sub + 1, __ATOMIC_RELEASE);
            r->next_sub++;
            r->next_task += (uint64_t)q.ngroups;
            r->tasks_in_flight += q.ngroups;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                q.req.t_start = s->now();
            }
            r->inflight[slot] = std::move(q);
            __atomic_store_n(&r->hctl->sub_tail, r->next_sub, __ATOMIC_RELEASE);
            ++running;
            did = true;
        }
        if (s->stop.load()) {
            bool idle;
            {
                std::lock_guard<std::mutex> lk(s->mu);
                idle = s->pending.empty();
            }
            if (idle && running == 0) break;
        }
        if (!did) {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait_for(lk, std::chrono::microseconds(20));
        }
    }
}

// End the persistent launch (the server is idle) and free its memory.
static void resident_destroy(wf_server *s)
{
    Resident *r = s->rs;
    if (!r) return;
    s->rs = nullptr;
    if (r->launched) {
        __atomic_store_n(&r->hctl->stop, 1, __ATOMIC_RELEASE);
        // Every waiting workgroup sees the stop flag within microseconds; a
        // strip still running ends within its dependency timeout.
        const auto t0 = Clock::now();
        hipError_t q;
        while ((q = hipStreamQuery(r->stream)) == hipErrorNotReady &&
               std::chrono::duration<double>(Clock::now() - t0).count() < 2.0 * s->ctx->timeout_s + 5.0)
            std::this_thread::sleep_for(std::chrono::microseconds(200));
        if (q == hipErrorNotReady) {
            // Leave the memory to the running launch rather than free it under it.
            std::fprintf(stderr, "wf_server_destroy: the resident launch did not end; its memory is kept\n");
            return;
        }
    }
    resident_free(r);
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
    if (s->opts.per_request && s->opts.resident) {
        wf_server_destroy(s);
        return WF_ERR_INVALID;
    }
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    s->budget = s->opts.hbm_budget > 0 ? s->opts.hbm_budget : (int64_t)(fr / 2);
    if (s->opts.resident) {
        s->t0 = Clock::now();
        if ((st = resident_create(s)) != WF_OK) {
            wf_server_destroy(s);
            return st;
        }
        s->sched = std::thread(resident_scheduler, s);
        *out = s;
        return WF_OK;
    }
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
    resident_destroy(s);
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
