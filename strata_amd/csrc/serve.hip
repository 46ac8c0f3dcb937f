// serve.hip — a request loop over a device-resident raster: window
// requests are admitted in arrival order into batches (the batch kernel:
// one persistent launch per batch) under an HBM budget for results, with
// several batches in flight on their own streams (continuous batching).
//
// Each result is the chain on the request's window as a raster of its own
// (exec.ProcessN semantics, internal/exec/process.go:28, then point ops),
// so it is the same alone or among other requests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "host.hpp"
#include "kernels.hpp"
#include "strata_hip.h"

using namespace strata;

namespace {

using Clock = std::chrono::steady_clock;

double seconds(Clock::time_point t) { return std::chrono::duration<double>(t.time_since_epoch()).count(); }

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// First-fit allocator over [0, cap) with coalescing of freed ranges.
struct Arena {
    std::map<size_t, size_t> free_;  // offset -> size
    size_t cap = 0, used = 0, peak = 0;

    void init(size_t c)
    {
        cap = c;
        free_.clear();
        free_[0] = c;
    }
    bool alloc(size_t n, size_t &off)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            off = it->first;
            const size_t rest = it->second - n;
            free_.erase(it);
            if (rest) free_[off + n] = rest;
            used += n;
            peak = std::max(peak, used);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n)
    {
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
};

enum ReqState { PENDING = 0, RUNNING = 1, DONE = 2, RELEASED = 3 };

struct Request {
    strata_window w;
    double t_arrival = 0, t_launch = 0, t_done = 0;
    int state = PENDING;
    bool drop = false;  // release the result as soon as it completes
    size_t off = 0, bytes = 0, mask_off = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    char *dev = nullptr;   // header | op | tasks | counter
    char *host = nullptr;  // pinned staging of the same layout
    std::vector<int64_t> reqs;
    bool busy = false;
};

} // namespace

struct strata_server {
    strata_serve_config cfg{};
    std::vector<std::vector<float>> taps;
    BatchOp op{};
    Arena arena;
    char *arena_dev = nullptr;
    std::vector<Slot> slots;
    size_t slot_bytes = 0;
    int max_blocks[kBatchFamilies] = {1, 1, 1};
    int device = 0;

    std::mutex mu;
    std::condition_variable cv;
    std::deque<int64_t> pending;
    std::vector<Request> reqs;
    std::thread sched;
    bool stop = false;
    std::atomic<uint64_t> gen{0};  // bumped (under mu) by submit, release and stop
    int64_t in_flight = 0;
    strata_serve_stats stats{};
    double busy_since = -1;
    int sched_error = STRATA_OK;

    void loop();
    bool launch(Slot &s, std::vector<int64_t> &batch, std::unique_lock<std::mutex> &lk);
    void complete(Slot &s, double now);
};

namespace {

// Result memory of a request: w*h floats, then its mask words.
void result_bytes(const strata_window &w, size_t &bytes, size_t &mask_off)
{
    const size_t cells = (size_t)(w.w * w.h);
    mask_off = align256(cells * sizeof(float));
    bytes = mask_off + align256(((cells + 63) / 64 + 1) * sizeof(uint64_t));
}

} // namespace

void strata_server::complete(Slot &s, double now)
{
    for (int64_t id : s.reqs) {
        Request &r = reqs[(size_t)id];
        r.t_done = now;
        r.state = DONE;
        stats.requests += 1;
        stats.queue_ms += (r.t_launch - r.t_arrival) * 1e3;
        stats.service_ms += (r.t_done - r.t_launch) * 1e3;
        stats.cells += r.w.w * r.w.h;
        if (r.drop) {
            arena.release(r.off, r.bytes);
            r.state = RELEASED;
        }
    }
    in_flight -= (int64_t)s.reqs.size();
    s.reqs.clear();
    s.busy = false;
}

// Builds the batch's work list in the slot's pinned staging, then (without
// the lock) uploads it and launches on the slot's stream.
bool strata_server::launch(Slot &s, std::vector<int64_t> &batch, std::unique_lock<std::mutex> &lk)
{
    std::sort(batch.begin(), batch.end(), [&](int64_t a, int64_t b) {
        return reqs[(size_t)a].w.w * reqs[(size_t)a].w.h > reqs[(size_t)b].w.w * reqs[(size_t)b].w.h;
    });
    std::vector<BatchTask> tasks(batch.size());
    for (size_t i = 0; i < batch.size(); ++i) {
        const Request &r = reqs[(size_t)batch[i]];
        strata_task t;
        memset(&t, 0, sizeof t);
        t.k = cfg.k;
        t.npost = cfg.npost;
        for (int j = 0; j < cfg.npost; ++j) t.post[j] = cfg.post[j];
        const int64_t start = r.w.y * cfg.src.stride + r.w.x;
        t.src = cfg.src;
        t.src.data = cfg.src.data + start;
        t.src.width = r.w.w;
        t.src.height = r.w.h;
        if (cfg.src.valid) t.src.valid_offset = cfg.src.valid_offset + start;
        t.dst.data = reinterpret_cast<float *>(arena_dev + r.off);
        t.dst.width = r.w.w;
        t.dst.height = r.w.h;
        t.dst.stride = r.w.w;
        t.dst.valid = reinterpret_cast<uint64_t *>(arena_dev + r.off + r.mask_off);
        t.dst.valid_offset = 0;
        t.dst.mem = STRATA_MEM_DEVICE;
        tasks[i] = make_task(t, 0);
    }
    std::vector<char> masked(batch.size(), 1);
    const ListLayout L = write_lists(&op, 1, tasks.data(), reinterpret_cast<const bool *>(masked.data()),
                                     tasks.size(), cfg.rows, s.host);
    s.reqs = batch;
    s.busy = true;
    in_flight += (int64_t)batch.size();
    stats.max_in_flight = std::max(stats.max_in_flight, in_flight);
    stats.batches += 1;
    // The upload carries zeroed counters: no clearing kernel.
    const BatchLaunch view = launch_view(L, s.dev, max_blocks[BF_HORN], max_blocks[BF_FOCAL], false);

    lk.unlock();
    hipError_t e = hipMemcpyAsync(s.dev, s.host, L.bytes, hipMemcpyHostToDevice, s.stream);
    if (e == hipSuccess) e = launch_batch(view, s.stream);
    if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
    const double launched = seconds(Clock::now());
    lk.lock();
    for (int64_t id : s.reqs) reqs[(size_t)id].t_launch = launched;
    if (e != hipSuccess) {
        sched_error = hip_fail(e, "launching a served batch");
        return false;
    }
    return true;
}

void strata_server::loop()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    std::vector<int64_t> batch;
    while (!stop) {
        bool progressed = false;
        const double now = seconds(Clock::now());
        for (Slot &s : slots) {
            if (s.busy && hipEventQuery(s.done) == hipSuccess) {
                complete(s, seconds(Clock::now()));
                progressed = true;
            }
        }
        if (progressed) cv.notify_all();
        bool any_busy = false, blocked = false;
        for (Slot &s : slots) {
            if (!s.busy && !pending.empty() && sched_error == STRATA_OK) {
                batch.clear();
                while (!pending.empty() && (int)batch.size() < cfg.max_batch) {
                    Request &r = reqs[(size_t)pending.front()];
                    if (!arena.alloc(r.bytes, r.off)) {
                        if (batch.empty()) {
                            stats.admission_waits += 1;
                            blocked = true;
                        }
                        break;
                    }
                    r.state = RUNNING;
                    batch.push_back(pending.front());
                    pending.pop_front();
                }
                if (!batch.empty()) {
                    if (!launch(s, batch, lk)) break;
                    progressed = true;
                }
            }
            any_busy = any_busy || s.busy;
        }
        stats.peak_bytes = (int64_t)arena.peak;
        if (any_busy && busy_since < 0) busy_since = now;
        if (!any_busy && busy_since >= 0) {
            stats.busy_ms += (now - busy_since) * 1e3;
            busy_since = -1;
        }
        if (progressed) continue;
        if (any_busy) {
            // Batches in flight: poll their events.
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        } else {
            // Nothing runs: the queue is empty, or its head waits for result
            // memory. Only a submission, a release or stop changes that. Spin
            // on their generation count for a while (a request arriving soon
            // is launched without a thread wake-up), then sleep.
            (void)blocked;
            const uint64_t g = gen.load();
            lk.unlock();
            const Clock::time_point until = Clock::now() + std::chrono::milliseconds(2);
            bool changed = false;
            while (Clock::now() < until) {
                if (gen.load() != g) {
                    changed = true;
                    break;
                }
                std::this_thread::yield();
            }
            lk.lock();
            if (!changed) cv.wait(lk, [&] { return stop || gen.load() != g; });
        }
    }
}

extern "C" {

int strata_serve_create(const strata_serve_config *cfg, strata_server **out)
{
    if (!cfg || !out) return fail(STRATA_EINVAL, "strata: nil serve config or result pointer");
    *out = nullptr;
    // The chain is checked as a batch task over the whole raster.
    strata_task t;
    memset(&t, 0, sizeof t);
    t.k = cfg->k;
    t.npost = cfg->npost;
    for (int j = 0; j < 2; ++j) t.post[j] = cfg->post[j];
    t.src = cfg->src;
    t.dst = cfg->src;
    int rc = check_task(t, 0);
    if (rc) return rc;
    strata_server *s = new strata_server;
    s->cfg = *cfg;
    if (s->cfg.slots <= 0) s->cfg.slots = 4;
    if (s->cfg.max_batch <= 0) s->cfg.max_batch = 256;
    if (s->cfg.rows < 0) s->cfg.rows = 0;  // 0: rows chosen per request for latency
    if (s->cfg.hbm_budget <= 0) s->cfg.hbm_budget = (int64_t)4 << 30;
    auto keep = [&](strata_kernel &k) {
        if (k.ntaps > 0) {
            s->taps.emplace_back(k.taps, k.taps + k.ntaps);
            k.taps = s->taps.back().data();
        }
    };
    s->taps.reserve(3);
    keep(s->cfg.k);
    for (int j = 0; j < s->cfg.npost; ++j) keep(s->cfg.post[j]);
    t.k = s->cfg.k;
    for (int j = 0; j < s->cfg.npost; ++j) t.post[j] = s->cfg.post[j];
    s->op = make_op(t);

    auto bail = [&](int code) {
        strata_serve_destroy(s);
        return code;
    };
    int cus = 1;
    if (hipGetDevice(&s->device) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess)
        return bail(fail(STRATA_EDEVICE, "strata: cannot query the device"));
    for (int f = 0; f < kBatchFamilies; ++f) s->max_blocks[f] = std::max(1, cus * batch_blocks_per_cu(f));
    if (hipMalloc((void **)&s->arena_dev, (size_t)s->cfg.hbm_budget) != hipSuccess)
        return bail(fail(STRATA_ENOMEM, "strata: cannot allocate %lld bytes of result memory",
                         (long long)s->cfg.hbm_budget));
    s->arena.init((size_t)s->cfg.hbm_budget);
    s->slot_bytes = list_bytes(1, (size_t)s->cfg.max_batch);
    s->slots.resize((size_t)s->cfg.slots);
    for (Slot &sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess ||
            hipMalloc((void **)&sl.dev, s->slot_bytes) != hipSuccess ||
            hipHostMalloc((void **)&sl.host, s->slot_bytes, hipHostMallocDefault) != hipSuccess)
            return bail(fail(STRATA_ENOMEM, "strata: cannot set up a batch slot"));
    }
    s->sched = std::thread([s] { s->loop(); });
    *out = s;
    return STRATA_OK;
}

static int submit_at(strata_server *s, const strata_window *w, double t_arrival, bool drop, int64_t *id)
{
    if (!s || !w) return fail(STRATA_EINVAL, "strata: nil server or window");
    if (w->w <= 0 || w->h <= 0 || w->x < 0 || w->y < 0 || w->x + w->w > s->cfg.src.width ||
        w->y + w->h > s->cfg.src.height)
        return fail(STRATA_EINVAL, "raster: window (%lld,%lld %lld×%lld) is outside the %lld×%lld raster",
                    (long long)w->x, (long long)w->y, (long long)w->w, (long long)w->h,
                    (long long)s->cfg.src.width, (long long)s->cfg.src.height);
    Request r;
    r.w = *w;
    r.t_arrival = t_arrival;
    r.drop = drop || s->cfg.auto_release;
    result_bytes(r.w, r.bytes, r.mask_off);
    if (r.bytes > s->arena.cap)
        return fail(STRATA_EINVAL, "strata: a %lld×%lld result needs %zu bytes, over the budget of %zu",
                    (long long)w->w, (long long)w->h, r.bytes, s->arena.cap);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->sched_error) return s->sched_error;
    const int64_t n = (int64_t)s->reqs.size();
    s->reqs.push_back(r);
    s->pending.push_back(n);
    if (id) *id = n;
    s->gen.fetch_add(1);
    s->cv.notify_all();
    return STRATA_OK;
}

int strata_serve_submit(strata_server *s, const strata_window *w, int64_t *id)
{
    return submit_at(s, w, seconds(Clock::now()), false, id);
}

int strata_serve_wait(strata_server *s, int64_t id, double *latency_ms)
{
    if (!s) return fail(STRATA_EINVAL, "strata: nil server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (id < 0 || id >= (int64_t)s->reqs.size()) return fail(STRATA_EINVAL, "strata: unknown request %lld", (long long)id);
    s->cv.wait(lk, [&] { return s->reqs[(size_t)id].state >= DONE || s->sched_error != STRATA_OK; });
    if (s->reqs[(size_t)id].state < DONE) return s->sched_error;
    if (latency_ms) *latency_ms = (s->reqs[(size_t)id].t_done - s->reqs[(size_t)id].t_arrival) * 1e3;
    return STRATA_OK;
}

int strata_serve_fetch(strata_server *s, int64_t id, float *data, uint64_t *mask)
{
    if (!s) return fail(STRATA_EINVAL, "strata: nil server");
    Request r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (id < 0 || id >= (int64_t)s->reqs.size() || s->reqs[(size_t)id].state != DONE)
            return fail(STRATA_EINVAL, "strata: request %lld is not a completed, unreleased result", (long long)id);
        r = s->reqs[(size_t)id];
    }
    const size_t cells = (size_t)(r.w.w * r.w.h);
    if (data)
        HIP_TRY(hipMemcpy(data, s->arena_dev + r.off, cells * sizeof(float), hipMemcpyDefault), "fetching a result");
    if (mask)
        HIP_TRY(hipMemcpy(mask, s->arena_dev + r.off + r.mask_off, (cells + 63) / 64 * sizeof(uint64_t),
                          hipMemcpyDefault),
                "fetching a result");
    return STRATA_OK;
}

int strata_serve_release(strata_server *s, int64_t id)
{
    if (!s) return fail(STRATA_EINVAL, "strata: nil server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (id < 0 || id >= (int64_t)s->reqs.size() || s->reqs[(size_t)id].state != DONE)
        return fail(STRATA_EINVAL, "strata: request %lld is not a completed, unreleased result", (long long)id);
    Request &r = s->reqs[(size_t)id];
    s->arena.release(r.off, r.bytes);
    r.state = RELEASED;
    s->gen.fetch_add(1);
    s->cv.notify_all();
    return STRATA_OK;
}

int strata_serve_replay(strata_server *s, const strata_window *windows, const double *arrival_s, int n,
                        double *latency_ms)
{
    if (!s || (n > 0 && (!windows || !arrival_s))) return fail(STRATA_EINVAL, "strata: nil replay arguments");
    for (int i = 1; i < n; ++i)
        if (arrival_s[i] < arrival_s[i - 1]) return fail(STRATA_EINVAL, "strata: arrivals must not decrease");
    std::vector<int64_t> ids((size_t)n);
    const Clock::time_point t0 = Clock::now();
    for (int i = 0; i < n; ++i) {
        const Clock::time_point at =
            t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[i]));
        // Sleep while the arrival is far, then spin to it.
        while (true) {
            const auto left = at - Clock::now();
            if (left <= Clock::duration::zero()) break;
            if (left > std::chrono::microseconds(300))
                std::this_thread::sleep_for(left - std::chrono::microseconds(200));
            else
                std::this_thread::yield();
        }
        const int rc = submit_at(s, &windows[i], seconds(at), true, &ids[(size_t)i]);
        if (rc) return rc;
    }
    for (int i = 0; i < n; ++i) {
        double ms = 0;
        const int rc = strata_serve_wait(s, ids[(size_t)i], &ms);
        if (rc) return rc;
        if (latency_ms) latency_ms[i] = ms;
    }
    return STRATA_OK;
}

int strata_serve_stats_get(strata_server *s, strata_serve_stats *out)
{
    if (!s || !out) return fail(STRATA_EINVAL, "strata: nil server or stats");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    return STRATA_OK;
}

void strata_serve_destroy(strata_server *s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->gen.fetch_add(1);
    }
    s->cv.notify_all();
    if (s->sched.joinable()) s->sched.join();
    for (Slot &sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.dev) (void)hipFree(sl.dev);
        if (sl.host) (void)hipHostFree(sl.host);
    }
    if (s->arena_dev) (void)hipFree(s->arena_dev);
    delete s;
}

} // extern "C"
