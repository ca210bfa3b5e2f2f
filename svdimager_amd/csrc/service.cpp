// Streaming request service: snapshot requests that arrive over time are admitted against a
// device-memory budget, grouped by image size into batched calls, and timed from arrival to
// image-ready, by a loop that runs on a thread of its own inside the library.
//
// The reference images one snapshot per process (main, SVDIntegratedImager.cu:1768-1853) and
// sizes its slab batches from free device memory (choose_slab_batch_size,
// SVDIntegratedImager.cu:1470-1496); this is the request loop above the batched call.
//
// * admission: a request costs its visibilities (the bytes the caller states for its snapshot)
//   plus its image, held from admission until its image is ready.  Requests are admitted in
//   arrival order while the context's workspace plus the bytes held plus the next cost fit the
//   budget.  The workspace is meant to be sized in advance (batched calls of the shapes to be
//   served, before the service starts) and only grows;
// * batching: admitted requests wait per (image size, cell).  While fewer than max_inflight
//   batched calls are in flight, the key of the oldest waiting request is taken with up to
//   max_batch requests of that key and enqueued as one asynchronous batched call;
// * completion: an event is recorded behind each batched call on the context stream; the loop
//   polls the oldest call's event, and a request's latency runs from its submission to the loop
//   seeing that event complete.
//
// GPU work cannot enqueue other kernels here, so the loop that launches each snapshot's kernels
// is host code: native, on its own thread, touching only device-resident, pre-sized buffers.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "svdi_internal.hpp"

using Clock = std::chrono::steady_clock;

struct svdi_service {
    struct Request {
        uint64_t id = 0;
        svdi_snapshot snap = {};
        size_t cost = 0;
        float* image = nullptr;
        size_t N = 0;
        float cell = 0.0f;
        Clock::time_point arrival, launched;
    };
    struct Batch {
        std::vector<Request> reqs;
        hipEvent_t ev = nullptr;
        int status = SVDI_OK;
        uint64_t seq = 0;
    };

    svdi_context* ctx = nullptr;
    size_t budget = 0;
    uint32_t max_batch = 1, max_inflight = 1;
    svdi_options opts = {};

    std::mutex mu;
    std::condition_variable cv;          // new requests, new results, stop
    std::thread worker;
    bool stop = false;
    std::deque<Request> incoming;        // submitted, not admitted (arrival order)
    std::map<std::pair<size_t, float>, std::deque<Request>> ready;   // admitted, not launched
    std::deque<Batch> inflight;
    std::vector<hipEvent_t> free_events;
    std::vector<svdi_service_result> results;   // finished, not collected
    size_t held = 0;                     // bytes of admitted, unfinished requests
    size_t workspace = 0;
    size_t outstanding = 0;              // submitted, not finished
    uint64_t calls = 0;                  // batched calls made (and refused requests)

    size_t workspace_bytes() {
        size_t b = 0;
        if (svdi_workspace_bytes(ctx, &b) != SVDI_OK) b = workspace;
        return b;
    }

    void finish(Batch& b, Clock::time_point t) {
        for (const Request& r : b.reqs) {
            svdi_service_result x = {};
            x.id = r.id;
            x.status = b.status;
            x.batch = static_cast<uint32_t>(b.reqs.size());
            x.call = b.seq;
            x.queue_ms = std::chrono::duration<double, std::milli>(r.launched - r.arrival).count();
            x.latency_ms = std::chrono::duration<double, std::milli>(t - r.arrival).count();
            results.push_back(x);
            held -= r.cost;
            --outstanding;
        }
        if (b.ev) free_events.push_back(b.ev);
    }

    // Under mu: retire the finished calls, oldest first.
    void retire() {
        while (!inflight.empty()) {
            Batch& b = inflight.front();
            if (b.ev) {
                const hipError_t q = hipEventQuery(b.ev);
                if (q == hipErrorNotReady) break;
                if (q != hipSuccess && b.status == SVDI_OK) b.status = SVDI_ERR_HIP;
            }
            finish(b, Clock::now());
            inflight.pop_front();
            cv.notify_all();
        }
    }

    // Under mu: admit in arrival order while the budget allows.  A request that does not fit
    // when nothing is held (the workspace has grown past what the budget leaves for it: a shape
    // the warm-up did not cover) fails with SVDI_ERR_OUT_OF_MEMORY instead of waiting forever.
    void admit() {
        while (!incoming.empty()) {
            Request& r = incoming.front();
            if (workspace + held + r.cost > budget) {
                if (held != 0) break;
                Batch b;
                b.seq = calls++;
                b.status = SVDI_ERR_OUT_OF_MEMORY;
                b.reqs.push_back(r);
                held += r.cost;
                b.reqs.back().launched = r.arrival;
                incoming.pop_front();
                finish(b, Clock::now());
                cv.notify_all();
                continue;
            }
            held += r.cost;
            ready[{r.N, r.cell}].push_back(r);
            incoming.pop_front();
        }
    }

    // Under mu (released around the call): launch batched calls while fewer than max_inflight
    // are in flight.
    void launch(std::unique_lock<std::mutex>& lk) {
        while (inflight.size() < max_inflight) {
            auto pick = ready.end();
            for (auto it = ready.begin(); it != ready.end(); ++it)
                if (!it->second.empty() &&
                    (pick == ready.end() || it->second.front().arrival < pick->second.front().arrival))
                    pick = it;
            if (pick == ready.end()) return;
            Batch b;
            b.seq = calls++;
            std::deque<Request>& q = pick->second;
            while (!q.empty() && b.reqs.size() < max_batch) {
                b.reqs.push_back(q.front());
                q.pop_front();
            }
            const Clock::time_point t = Clock::now();
            for (Request& r : b.reqs) r.launched = t;
            if (free_events.empty()) {
                hipEvent_t e = nullptr;
                if (hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess) free_events.push_back(e);
            }
            if (free_events.empty()) {               // the call could not be tracked: not made
                b.status = SVDI_ERR_HIP;
                finish(b, Clock::now());
                cv.notify_all();
                continue;
            }
            b.ev = free_events.back();
            free_events.pop_back();
            lk.unlock();
            std::vector<svdi_snapshot> snaps;
            std::vector<float*> images;
            std::vector<size_t> sizes;
            std::vector<float> cells;
            for (const Request& r : b.reqs) {
                snaps.push_back(r.snap);
                images.push_back(r.image);
                sizes.push_back(r.N);
                cells.push_back(r.cell);
            }
            b.status = svdi::image_batch(ctx, snaps.data(), snaps.size(), images.data(), sizes.data(), cells.data(),
                                         &opts, nullptr, false);
            if (b.status == SVDI_OK) {
                std::lock_guard<std::mutex> cl(ctx->mu);
                (void)hipSetDevice(ctx->device);
                if (hipEventRecord(b.ev, ctx->stream) != hipSuccess) {
                    // Wait for the context's work instead, so that the requests are not reported
                    // done while their images are still being written.
                    b.status = SVDI_ERR_HIP;
                    (void)hipStreamSynchronize(ctx->stream);
                }
            }
            const size_t ws = workspace_bytes();
            lk.lock();
            workspace = ws;
            if (b.status != SVDI_OK) {
                finish(b, Clock::now());
                cv.notify_all();
            } else {
                inflight.push_back(std::move(b));
            }
        }
    }

    void run() {
        (void)hipSetDevice(ctx->device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            retire();
            admit();
            launch(lk);
            if (stop && outstanding == 0) break;
            if (!inflight.empty()) cv.wait_for(lk, std::chrono::microseconds(20));
            else cv.wait(lk, [&] { return stop || !incoming.empty(); });
        }
    }
};

namespace {

int service_error(const char* msg) {
    svdi::report_error(msg);
    return SVDI_ERR_INVALID_ARGUMENT;
}

}  // namespace

extern "C" {

int svdi_service_create(svdi_context* ctx, const svdi_service_config* cfg, svdi_service** out) {
    if (!ctx || !cfg || !out) return service_error("ctx / config / out is null");
    if (cfg->max_batch < 1 || cfg->max_inflight < 1 || cfg->budget_bytes == 0)
        return service_error("max_batch, max_inflight and budget_bytes must be positive");
    svdi_service* s = nullptr;
    try {
        s = new svdi_service();
        s->ctx = ctx;
        s->budget = cfg->budget_bytes;
        s->max_batch = cfg->max_batch;
        s->max_inflight = cfg->max_inflight;
        s->opts = cfg->options;
        s->workspace = s->workspace_bytes();
        s->worker = std::thread([s] { s->run(); });
    } catch (...) {
        delete s;
        svdi::report_error("could not start the service's thread");
        return SVDI_ERR_INTERNAL;
    }
    *out = s;
    return SVDI_OK;
}

int svdi_service_destroy(svdi_service* s) {
    if (!s) return service_error("service is null");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
    delete s;
    return SVDI_OK;
}

int svdi_service_submit(svdi_service* s, uint64_t id, const svdi_snapshot* snap, size_t snapshot_bytes,
                        float* image, size_t image_size, float cell_size) {
    if (!s || !snap || !image) return service_error("service / snapshot / image is null");
    if (image_size == 0 || !(cell_size > 0.0f)) return service_error("image size and cell must be positive");
    svdi_service::Request r;
    r.id = id;
    r.snap = *snap;
    r.cost = snapshot_bytes + 4 * image_size * image_size;
    r.image = image;
    r.N = image_size;
    r.cell = cell_size;
    r.arrival = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return service_error("the service is stopping");
        if (s->workspace + r.cost > s->budget)
            return service_error("the request could never fit: its bytes and the workspace pass the budget");
        s->incoming.push_back(r);
        ++s->outstanding;
    }
    s->cv.notify_all();
    return SVDI_OK;
}

int svdi_service_collect(svdi_service* s, svdi_service_result* out, size_t capacity, size_t* count,
                         int wait_ms) {
    if (!s || !count || (capacity > 0 && !out)) return service_error("service / count / out is null");
    std::unique_lock<std::mutex> lk(s->mu);
    if (wait_ms > 0 && s->results.empty() && s->outstanding > 0)
        s->cv.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->results.empty() || s->outstanding == 0; });
    const size_t n = std::min(capacity, s->results.size());
    for (size_t i = 0; i < n; ++i) out[i] = s->results[i];
    s->results.erase(s->results.begin(), s->results.begin() + static_cast<long>(n));
    *count = n;
    return SVDI_OK;
}

int svdi_service_stats(svdi_service* s, svdi_service_stats_info* info) {
    if (!s || !info) return service_error("service / info is null");
    std::lock_guard<std::mutex> lk(s->mu);
    info->outstanding = s->outstanding;
    info->waiting = s->incoming.size();
    size_t admitted = 0;
    for (const auto& kv : s->ready) admitted += kv.second.size();
    info->admitted = admitted;
    info->inflight_calls = s->inflight.size();
    info->held_bytes = s->held;
    info->workspace_bytes = s->workspace;
    return SVDI_OK;
}

}  // extern "C"
