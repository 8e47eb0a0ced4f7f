// sgbm_serve.hip - serving form: requests admitted into a device pool by footprint and run in
// continuous batches over several lanes.
//
// A lane is a stream with its own BatchEngine (persistent per-stage kernels in one hipGraph, or
// the per-pair kernels for a batch of one), so batches on different lanes overlap on the
// device, and requests that arrive while a batch runs start on the next free lane instead of
// waiting for it. Each pair's result is the batched pipeline's, which is bit-identical to the
// per-pair path.
//
// Who does what:
//   - the submitting thread validates the request, copies host images into a pinned block of
//     the server's pinned pool, queues the request for admission, and, when a lane is idle and
//     nothing else is ready or waiting, admits it and starts it on that lane itself (no
//     hand-over to another thread);
//   - the dispatcher thread admits waiting requests whose working set fits the device pool
//     (serve_core.h Admission), retires lanes whose batch has finished (pool ranges released,
//     requests marked done) and starts ready requests on idle lanes, as one batch of requests
//     of similar size (pick_batch), so small ones never wait for a large one;
//   - a completion thread copies finished outputs from their pinned blocks into the callers'
//     buffers and marks the requests done, in the order their batches finish, so a lane is
//     free again as soon as its batch is, and no result waits for another one's waiter.
// A request that found no room in the pinned pool is staged through its lane's own pinned
// buffers instead, by whichever thread starts and retires its batch.
#include <hip/hip_runtime.h>
#include <sys/prctl.h>
#include <time.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch_engine.h"
#include "host_copy.h"
#include "lumen_sgbm.h"
#include "serve_core.h"
#include "sgbm_host.h"

using namespace lumen;

namespace {

thread_local std::string g_serve_error;

int64_t monotonic_ns() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return static_cast<int64_t>(ts.tv_sec) * 1000000000LL + ts.tv_nsec;
}

int hip_code(hipError_t e) { return e == hipErrorOutOfMemory ? LUMEN_ENOMEM : LUMEN_EDEVICE; }

struct HipBacking : PoolBacking {
    void* acquire(size_t bytes) override {
        void* p = nullptr;
        return hipMalloc(&p, bytes) == hipSuccess ? p : nullptr;
    }
    void release(void* p) override { (void)hipFree(p); }
};

struct PinnedBacking : PoolBacking {
    void* acquire(size_t bytes) override {
        void* p = nullptr;
        return hipHostMalloc(&p, bytes, hipHostMallocDefault) == hipSuccess ? p : nullptr;
    }
    void release(void* p) override { (void)hipHostFree(p); }
};

struct Request {
    uint64_t id = 0;
    lumen_pair_desc d{};
    size_t bytes = 0;
    double work = 0;  // algorithmic HBM bytes of the pair: W*H*Dp*(1+5*paths)
    bool large = false;  // above the server's large_us: runs alone on lane 0
    char* region = nullptr;
    int16_t* devOut = nullptr;  // the region's output map (host-memory requests)
    // Pinned staging of a host request, filled at submit: left at 0, right at
    // align_up(W*H, 256) (the gap of the device region's layout), the output at pinOut.
    char* pin = nullptr;
    size_t pinOut = 0;
    size_t inOff = 0, outOff = 0;  // offsets in the lane's staging (requests without `pin`)
    int status = LUMEN_OK;
    std::string err;
    lumen_request_info info{};
    bool done = false;     // device work finished and retired (output staged or written)
    bool waiting = false;  // a thread is in lumen_server_wait for this ticket
};

struct Lane {
    int index = 0;
    BatchEngine engine;
    hipEvent_t fin = nullptr;
    bool busy = false;      // claimed for a batch
    bool launched = false;  // its batch is enqueued and `fin` recorded after it
    std::vector<Request*> reqs;
    std::string err;
    void* pinIn = nullptr;
    size_t pinInCap = 0;
    void* pinOut = nullptr;
    size_t pinOutCap = 0;
    ~Lane() {
        if (engine.stream()) (void)hipStreamSynchronize(engine.stream());
        if (fin) (void)hipEventDestroy(fin);
        if (pinIn) (void)hipHostFree(pinIn);
        if (pinOut) (void)hipHostFree(pinOut);
    }
};

hipError_t grow_pinned(void*& p, size_t& cap, size_t need) {
    if (need <= cap && p) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    const size_t n = need * 5 / 4 + 4096;
    hipError_t e = hipHostMalloc(&p, n, hipHostMallocDefault);
    if (e == hipSuccess) cap = n;
    return e;
}

size_t staged_input_bytes(size_t n) { return align_up(n, 256) + n; }

}  // namespace

struct lumen_server {
    lumen_server_config cfg{};
    int device = 0;
    HipBacking backing;
    DevicePool pool;
    PinnedBacking pinBacking;
    DevicePool pinned;  // host staging blocks of host-memory requests
    Admission admission;
    std::vector<std::unique_ptr<Lane>> lanes;
    std::mutex m;
    std::condition_variable wake;    // dispatcher: new work or shutdown
    std::condition_variable doneCv;  // waiters: a request finished
    std::condition_variable copyCv;  // completion thread: outputs to deliver or shutdown
    std::deque<Request*> finished;   // device work done, output still in its pinned block
    bool copyStop = false;
    std::thread copier;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    std::deque<Request*> ready;      // admitted, not started, in admission order
    std::deque<Request*> readyLarge; // the same for large requests (lane 0, one at a time)
    uint64_t nextId = 1;
    bool stop = false;
    int busyLanes = 0;
    lumen_server_stats stats{};
    std::thread thread;

    bool idle_locked() const {
        return admission.empty() && ready.empty() && readyLarge.empty() && busyLanes == 0 &&
               finished.empty() && copying == 0;
    }
    int copying = 0;  // outputs the completion thread is copying right now
    // Whether lane i may start a batch now, and from which queue.
    std::deque<Request*>* queue_for(size_t i) {
        if (cfg.large_us > 0 && i == 0 && !readyLarge.empty()) return &readyLarge;
        return ready.empty() ? nullptr : &ready;
    }
};

namespace {

// Enqueues one batch on a lane: uploads (host requests), the pipeline, downloads, the finish
// event. Runs without the server lock; a claimed lane and its requests belong to the thread
// that claimed it until it is marked launched.
int start_lane(lumen_server* S, Lane& L) {
    const bool hostIo = S->cfg.device_io == 0;
    hipStream_t s = L.engine.stream();
    BatchPlan pl;
    pl.jobs.reserve(L.reqs.size());
    size_t inBytes = 0, outBytes = 0;
    const int64_t t = monotonic_ns();
    for (Request* r : L.reqs) {
        PairJob j{};
        fill_job(j, r->d.params, r->d.width, r->d.height);
        r->devOut = layout_pair(j, r->region);
        const size_t n = (size_t)j.W * j.H;
        if (hostIo) {
            if (!r->pin) {
                // left and right are adjacent in the region (each rounded up to 256 bytes), so
                // one copy uploads both from a staging block with the same gap.
                r->inOff = inBytes;
                inBytes += 2 * align_up(n, 256);
                r->outOff = outBytes;
                outBytes += align_up(n * 2, 256);
            }
        } else {
            j.left = r->d.left;
            j.right = r->d.right;
            j.out = r->d.out;
        }
        pl.jobs.push_back(j);
        r->info.t_start_ns = t;
        r->info.lane = L.index;
        r->info.batch_size = static_cast<int32_t>(L.reqs.size());
    }
    auto fail = [&](hipError_t e, const char* what) {
        L.err = std::string(what) + ": " + hipGetErrorString(e);
        return hip_code(e);
    };
    if (hostIo) {
        if (inBytes) {
            hipError_t e = grow_pinned(L.pinIn, L.pinInCap, inBytes);
            if (e == hipSuccess) e = grow_pinned(L.pinOut, L.pinOutCap, outBytes);
            if (e != hipSuccess) return fail(e, "pinned staging");
            uint8_t* pin = static_cast<uint8_t*>(L.pinIn);
            std::vector<CopyJob> stage;
            for (Request* r : L.reqs) {
                if (r->pin) continue;
                const size_t n = (size_t)r->d.width * r->d.height;
                stage.push_back({pin + r->inOff, r->d.left, n});
                stage.push_back({pin + r->inOff + align_up(n, 256), r->d.right, n});
            }
            parallel_copy(stage);
        }
        for (Request* r : L.reqs) {
            const size_t n = (size_t)r->d.width * r->d.height;
            const void* src = r->pin ? static_cast<const void*>(r->pin)
                                     : static_cast<uint8_t*>(L.pinIn) + r->inOff;
            hipError_t e = hipMemcpyAsync(r->region, src, staged_input_bytes(n),
                                          hipMemcpyHostToDevice, s);
            if (e != hipSuccess) return fail(e, "upload");
        }
    }
    int rc = L.engine.enqueue(pl, L.err);
    if (rc) return rc;
    if (hostIo) {
        for (Request* r : L.reqs) {
            const size_t n = (size_t)r->d.width * r->d.height;
            void* dst = r->pin ? static_cast<void*>(r->pin + r->pinOut)
                               : static_cast<uint8_t*>(L.pinOut) + r->outOff;
            hipError_t e = hipMemcpyAsync(dst, r->devOut, n * 2, hipMemcpyDeviceToHost, s);
            if (e != hipSuccess) return fail(e, "download");
        }
    }
    hipError_t e = hipEventRecord(L.fin, s);
    if (e != hipSuccess) return fail(e, "event record");
    return LUMEN_OK;
}

// Copies a finished lane's outputs of requests without a pinned block to their callers (the
// others are copied by their waiters). No server lock needed.
void finish_lane(lumen_server* S, Lane& L, hipError_t status) {
    if (status != hipSuccess) {
        for (Request* r : L.reqs) {
            r->status = hip_code(status);
            r->err = std::string("batch: ") + hipGetErrorString(status);
        }
        return;
    }
    if (S->cfg.device_io == 0) {
        const uint8_t* pout = static_cast<const uint8_t*>(L.pinOut);
        std::vector<CopyJob> outs;
        for (Request* r : L.reqs)
            if (!r->pin)
                outs.push_back({r->d.out, pout + r->outOff, (size_t)r->d.width * r->d.height * 2});
        parallel_copy(outs);
    }
}

void complete_locked(lumen_server* S, Request* r, int64_t t) {
    r->info.status = r->status;
    r->info.t_done_ns = t;
    r->done = true;
    if (r->status == LUMEN_OK)
        ++S->stats.completed;
    else
        ++S->stats.failed;
}

// Frees a finished lane and its requests' device ranges. Requests whose output waits in a
// pinned block go to the completion thread; the others are complete.
void retire_locked(lumen_server* S, Lane& L) {
    const int64_t t = monotonic_ns();
    bool toCopy = false;
    for (Request* r : L.reqs) {
        S->pool.release(r->region);
        r->region = nullptr;
        if (r->pin && r->status == LUMEN_OK) {
            S->finished.push_back(r);
            toCopy = true;
        } else {
            complete_locked(S, r, t);
        }
    }
    L.reqs.clear();
    L.busy = false;
    L.launched = false;
    --S->busyLanes;
    if (toCopy) S->copyCv.notify_one();
}

// Completion thread: delivers staged outputs to the callers' buffers, then completes them.
void copier(lumen_server* S) {
    std::unique_lock<std::mutex> lk(S->m);
    for (;;) {
        S->copyCv.wait(lk, [&] { return S->copyStop || !S->finished.empty(); });
        if (S->finished.empty()) return;  // stopping, nothing left
        Request* r = S->finished.front();
        S->finished.pop_front();
        ++S->copying;
        lk.unlock();
        parallel_copy({{r->d.out, r->pin + r->pinOut, (size_t)r->d.width * r->d.height * 2}});
        lk.lock();
        --S->copying;
        S->pinned.release(r->pin);
        r->pin = nullptr;
        complete_locked(S, r, monotonic_ns());
        S->doneCv.notify_all();
    }
}

// Hands out pool ranges to waiting requests that fit; they join a ready queue.
void admit_locked(lumen_server* S) {
    for (const Admission::Admitted& a : S->admission.admit()) {
        Request* r = S->reqs[a.id].get();
        r->region = static_cast<char*>(a.region);
        r->info.t_admit_ns = monotonic_ns();
        (r->large ? S->readyLarge : S->ready).push_back(r);
    }
    S->stats.pool_high_water = S->pool.high_water();
}

// Moves the next batch of queue q onto idle lane L and claims it (a large request runs alone).
void claim_locked(lumen_server* S, Lane& L, std::deque<Request*>& q) {
    const size_t maxBatch = &q == &S->readyLarge ? 1 : static_cast<size_t>(S->cfg.max_batch);
    const double budget = S->cfg.batch_budget_us * 5e6;  // bytes at 5 TB/s
    std::vector<double> work;
    for (const Request* r : q) work.push_back(r->work);
    const std::vector<size_t> pick = pick_batch(work, maxBatch, budget, 4.0);
    for (size_t k = pick.size(); k-- > 0;) {
        L.reqs.push_back(q[pick[k]]);
        q.erase(q.begin() + static_cast<std::ptrdiff_t>(pick[k]));
    }
    std::reverse(L.reqs.begin(), L.reqs.end());
    L.busy = true;
    L.launched = false;
    ++S->busyLanes;
    ++S->stats.batches;
    S->stats.max_batch_seen = std::max<uint64_t>(S->stats.max_batch_seen, L.reqs.size());
    S->stats.max_lanes_busy = std::max<uint64_t>(S->stats.max_lanes_busy, S->busyLanes);
}

// Enqueues a claimed lane's batch; called with the lock held through `lk`, which is released
// while the copies and launches are enqueued.
void launch_claimed(lumen_server* S, Lane& L, std::unique_lock<std::mutex>& lk) {
    lk.unlock();
    const int rc = start_lane(S, L);
    if (rc) {
        // Nothing of a failed start is left in flight once the lane's stream drains.
        (void)hipStreamSynchronize(L.engine.stream());
        for (Request* r : L.reqs) {
            r->status = rc;
            r->err = L.err;
        }
    }
    lk.lock();
    if (rc) {
        retire_locked(S, L);
        S->doneCv.notify_all();
    } else {
        L.launched = true;
    }
}

void dispatcher(lumen_server* S) {
    (void)hipSetDevice(S->device);
    // Timed waits of a few microseconds: the default 50 us timer slack would stretch each poll
    // of the lanes several times over.
    (void)prctl(PR_SET_TIMERSLACK, 1UL, 0UL, 0UL, 0UL);
    std::unique_lock<std::mutex> lk(S->m);
    for (;;) {
        bool progress = false;
        const size_t readyBefore = S->ready.size();
        admit_locked(S);
        progress |= S->ready.size() != readyBefore;
        for (auto& L : S->lanes) {
            if (!L->launched) continue;
            const hipError_t q = hipEventQuery(L->fin);
            if (q == hipErrorNotReady) continue;
            lk.unlock();
            finish_lane(S, *L, q);
            lk.lock();
            retire_locked(S, *L);
            S->doneCv.notify_all();
            progress = true;
        }
        if (progress) admit_locked(S);  // retired ranges may let waiting requests in
        for (size_t i = 0; i < S->lanes.size(); ++i) {
            Lane& L = *S->lanes[i];
            std::deque<Request*>* q = S->queue_for(i);
            if (L.busy || !q) continue;
            claim_locked(S, L, *q);
            launch_claimed(S, L, lk);
            progress = true;
        }
        if (S->stop && S->idle_locked()) break;
        if (!progress) {
            // Poll the lanes often while batches run; otherwise sleep until a submit.
            if (S->busyLanes > 0)
                S->wake.wait_for(lk, std::chrono::microseconds(10));
            else
                S->wake.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
}

}  // namespace

extern "C" {

int64_t lumen_now_ns(void) { return monotonic_ns(); }

int lumen_server_create(const lumen_server_config* cfg, lumen_server** out) {
    if (!out) {
        g_serve_error = "out is NULL";
        return LUMEN_EINVAL;
    }
    *out = nullptr;
    lumen_server_config c{};
    if (cfg) c = *cfg;
    if (c.lanes == 0) c.lanes = 2;
    if (c.max_batch == 0) c.max_batch = 64;
    if (c.max_bypass == 0) c.max_bypass = 8;
    if (c.batch_budget_us <= 0) c.batch_budget_us = 2000;
    if (c.lanes < 1 || c.lanes > 4 || c.max_batch < 1 || c.max_batch > 65535 ||
        c.pinned_mb < 0 || c.large_us < 0) {
        g_serve_error = "lanes must be 1..4, max_batch 1..65535, pinned_mb and large_us >= 0";
        return LUMEN_EINVAL;
    }
    std::unique_ptr<lumen_server> S(new (std::nothrow) lumen_server());
    if (!S) {
        g_serve_error = "out of host memory";
        return LUMEN_ENOMEM;
    }
    hipError_t e = hipGetDevice(&S->device);
    if (e == hipSuccess && c.pool_bytes == 0) {
        size_t freeB = 0, totalB = 0;
        e = hipMemGetInfo(&freeB, &totalB);
        c.pool_bytes = freeB / 2;
    }
    if (e != hipSuccess) {
        g_serve_error = std::string("lumen_server_create: ") + hipGetErrorString(e);
        return hip_code(e);
    }
    S->cfg = c;
    if (!S->pool.open(&S->backing, c.pool_bytes)) {
        g_serve_error = "lumen_server_create: cannot allocate a device pool of " +
                        std::to_string(c.pool_bytes) + " bytes";
        return LUMEN_ENOMEM;
    }
    S->admission = Admission(&S->pool, c.max_bypass < 0 ? 0 : c.max_bypass);
    S->stats.pool_bytes = S->pool.capacity();
    if (c.device_io == 0) {
        // Without a pinned pool every request is staged through its lane's buffers instead.
        const size_t mb = c.pinned_mb > 0 ? static_cast<size_t>(c.pinned_mb) : 1024;
        (void)S->pinned.open(&S->pinBacking, mb << 20);
    }
    for (int i = 0; i < c.lanes; ++i) {
        std::unique_ptr<Lane> L(new Lane());
        L->index = i;
        e = L->engine.init();
        if (e == hipSuccess) e = hipEventCreateWithFlags(&L->fin, hipEventDisableTiming);
        if (e != hipSuccess) {
            g_serve_error = std::string("lumen_server_create: ") + hipGetErrorString(e);
            return hip_code(e);
        }
        S->lanes.push_back(std::move(L));
    }
    lumen_server* raw = S.get();
    try {
        S->copier = std::thread(copier, raw);
        S->thread = std::thread(dispatcher, raw);
    } catch (...) {
        if (S->copier.joinable()) {
            {
                std::lock_guard<std::mutex> g(S->m);
                S->copyStop = true;
            }
            S->copyCv.notify_all();
            S->copier.join();
        }
        g_serve_error = "lumen_server_create: cannot start the server threads";
        return LUMEN_ENOMEM;
    }
    *out = S.release();
    return LUMEN_OK;
}

int lumen_server_destroy(lumen_server* s) {
    if (!s) return LUMEN_OK;
    {
        std::lock_guard<std::mutex> g(s->m);
        s->stop = true;
    }
    s->wake.notify_all();
    if (s->thread.joinable()) s->thread.join();  // returns once nothing is queued or copying
    {
        std::lock_guard<std::mutex> g(s->m);
        s->copyStop = true;
    }
    s->copyCv.notify_all();
    if (s->copier.joinable()) s->copier.join();
    delete s;
    return LUMEN_OK;
}

int lumen_server_submit(lumen_server* s, const lumen_pair_desc* pair, uint64_t* ticket) {
    if (!s || !pair || !ticket) {
        g_serve_error = "NULL argument";
        return LUMEN_EINVAL;
    }
    const lumen_pair_desc& d = *pair;
    const std::string msg = validate_pair(d.params, d.width, d.height, d.left && d.right);
    if (!msg.empty()) {
        g_serve_error = msg;
        return LUMEN_EINVAL;
    }
    if (!d.out) {
        g_serve_error = "Output buffer is NULL";
        return LUMEN_EINVAL;
    }
    const size_t bytes = pair_region_bytes(d.params, d.width, d.height);
    if (!s->pool.fits_ever(bytes)) {
        g_serve_error = "request needs " + std::to_string(bytes) + " bytes, the pool has " +
                        std::to_string(s->pool.capacity());
        return LUMEN_ENOMEM;
    }
    std::unique_ptr<Request> r(new (std::nothrow) Request());
    if (!r) {
        g_serve_error = "out of host memory";
        return LUMEN_ENOMEM;
    }
    r->d = d;
    r->bytes = bytes;
    r->work = static_cast<double>(d.width) * d.height *
              padded_range(d.params.maxDisparity - d.params.minDisparity) *
              (1 + 5 * num_paths(d.params));
    r->large = s->cfg.large_us > 0 && r->work > s->cfg.large_us * 5e6;
    r->info.t_submit_ns = monotonic_ns();
    r->info.lane = -1;
    const bool hostIo = s->cfg.device_io == 0;
    const size_t n = (size_t)d.width * d.height;
    if (hostIo) {
        // Stage the images now, in this thread: the dispatcher then only enqueues copies.
        r->pinOut = align_up(staged_input_bytes(n), 256);
        {
            std::lock_guard<std::mutex> g(s->m);
            r->pin = static_cast<char*>(s->pinned.alloc(r->pinOut + 2 * n));
            if (!r->pin) ++s->stats.unstaged;
        }
        if (r->pin)
            parallel_copy({{r->pin, d.left, n}, {r->pin + align_up(n, 256), d.right, n}});
    }
    Lane* direct = nullptr;
    std::unique_lock<std::mutex> lk(s->m);
    if (s->stop) {
        if (r->pin) s->pinned.release(r->pin);
        g_serve_error = "server is shutting down";
        return LUMEN_EINVAL;
    }
    const uint64_t id = s->nextId++;
    r->id = id;
    *ticket = id;
    s->admission.push(id, bytes);
    s->reqs[id] = std::move(r);
    ++s->stats.submitted;
    s->stats.max_waiting = std::max<uint64_t>(s->stats.max_waiting, s->admission.waiting());
    // Nothing ahead of it and a lane idle: admit and start it from here, without waking the
    // dispatcher first.
    if (s->ready.empty() && s->readyLarge.empty() && s->admission.waiting() == 1) {
        admit_locked(s);
        for (size_t i = 0; i < s->lanes.size(); ++i) {
            std::deque<Request*>* q = s->queue_for(i);
            if (s->lanes[i]->busy || !q) continue;
            direct = s->lanes[i].get();
            claim_locked(s, *direct, *q);
            ++s->stats.direct_starts;
            break;
        }
    }
    if (direct) launch_claimed(s, *direct, lk);
    lk.unlock();
    s->wake.notify_one();
    return LUMEN_OK;
}

int lumen_server_wait(lumen_server* s, uint64_t ticket, int64_t timeout_us,
                      lumen_request_info* info) {
    if (!s) {
        g_serve_error = "NULL server";
        return LUMEN_EINVAL;
    }
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) {
        g_serve_error = "unknown ticket";
        return LUMEN_EINVAL;
    }
    Request* r = it->second.get();
    // One waiter per ticket: the waiter that sees the request done erases it, so a second
    // waiter would be left holding a freed request.
    if (r->waiting) {
        g_serve_error = "ticket already has a waiter";
        return LUMEN_EINVAL;
    }
    r->waiting = true;
    auto isDone = [&] { return r->done; };
    if (timeout_us < 0) {
        s->doneCv.wait(lk, isDone);
    } else if (!s->doneCv.wait_for(lk, std::chrono::microseconds(timeout_us), isDone)) {
        r->waiting = false;
        g_serve_error = "timed out";
        return LUMEN_ETIMEDOUT;
    }
    if (r->pin) {  // a failed request's staging block (delivered ones are released already)
        s->pinned.release(r->pin);
        r->pin = nullptr;
    }
    if (info) *info = r->info;
    const int status = r->status;
    if (status != LUMEN_OK) g_serve_error = r->err;
    s->reqs.erase(ticket);
    return status;
}

int lumen_server_drain(lumen_server* s) {
    if (!s) return LUMEN_EINVAL;
    std::unique_lock<std::mutex> lk(s->m);
    s->doneCv.wait(lk, [&] { return s->idle_locked(); });
    return LUMEN_OK;
}

int lumen_server_get_stats(lumen_server* s, lumen_server_stats* out) {
    if (!s || !out) return LUMEN_EINVAL;
    std::lock_guard<std::mutex> g(s->m);
    *out = s->stats;
    out->pool_high_water = s->pool.high_water();
    return LUMEN_OK;
}

const char* lumen_server_last_error(const lumen_server* s) {
    (void)s;
    return g_serve_error.c_str();
}

}  // extern "C"
