// Render server: continuous batching of render requests on one context.
//
// The reference renders one view per launchKernels call from its UI loop
// (Application.cpp:101-125); it has no request queue.  Here a scheduler
// thread owns the context:
//   - submit() checks and queues a request (camera, size, spp, frame base,
//     frame options); a request whose accumulator can never fit the arena is
//     refused (CT_ERR_ADMISSION);
//   - memory: the budget holds the shared path-state queues (kLanes x chunk
//     path slots of kPathSlotBytes, sized once at creation) and the
//     accumulator arena (the rest).  A request is admitted, in arrival
//     order, once its accumulator (width x height x 16 bytes) fits the arena;
//     its path samples then take queue slots chunk by chunk as slots free up,
//     so requests that arrive while others are on the GPU join the next chunk;
//   - every chunk packs path samples of admitted requests into one wavefront
//     (one generate, depth x (extend, shade, connect), one resolve, replayed
//     from a captured hipGraph), smallest remaining work first with aging
//     (remaining paths / (1 + wait / 50 ms)), so a small request that arrives
//     behind a large one is done in the next chunk and a large one is not
//     starved.  Every option of a request (depth, RR, BSDF mode, epsilons,
//     light pdf, ...) is honoured: a chunk takes the options of its first
//     request and only requests with the same options, others wait for a
//     later chunk (aging moves them to the front);
//   - a request is finished when its last path sample has been resolved; a
//     copier thread moves its accumulator to host memory on its own stream
//     while the next chunk runs, then releases its memory.
// A request's result does not depend on what it was batched with: the RNG is
// keyed by (pixel, frame_base + sample, bounce) and samples are summed in
// sample order.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "runtime.h"

using ctd::ViewDesc;
using namespace ctr;

namespace {

// Default paths per chunk (split over the two lanes): small enough that a
// request arriving while a chunk runs starts within a few milliseconds,
// large enough that both lanes stay full.
constexpr uint32_t kServerChunk = 1u << 23;

enum ReqState { PENDING = 0, ACTIVE = 1, COPYING = 2, DONE = 3 };

struct Req {
    uint64_t id = 0;
    ct_request r{};
    ct_frame_opts opts{};          // r.opts with spp = 1 (the chunk options)
    int state = PENDING;
    int status = CT_OK;
    uint64_t total = 0, next = 0;  // path samples
    uint64_t footprint = 0;        // accumulator bytes
    uint64_t acc_off = 0;          // float4 offset in the arena
    double submit_ms = 0, admit_ms = 0, done_ms = 0;
    uint64_t rays = 0;
    bool cancel = false;
    std::vector<float> result;
};

// First-fit allocator over the accumulator arena (units: float4).
struct Arena {
    std::map<uint64_t, uint64_t> free_;  // offset -> length
    explicit Arena(uint64_t n = 0) { if (n) free_[0] = n; }
    bool alloc(uint64_t n, uint64_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= n) {
                *off = it->first;
                const uint64_t rest = it->second - n, at = it->first + n;
                free_.erase(it);
                if (rest) free_[at] = rest;
                return true;
            }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        auto it = free_.emplace(off, n).first;
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
};

// Requests share a chunk only with equal options (spp normalised to 1).
bool same_opts(const ct_frame_opts& a, const ct_frame_opts& b) { return std::memcmp(&a, &b, sizeof(a)) == 0; }

}  // namespace

struct ct_server {
    ct_ctx* ctx = nullptr;
    ct_server_opts opts{};
    std::mutex mu;
    std::condition_variable cv;
    std::thread worker, copier;
    bool stop = false, worker_done = false;
    std::deque<uint64_t> to_copy;  // finished requests awaiting copy-out
    hipStream_t copy_stream = nullptr;
    uint64_t next_id = 1;
    std::map<uint64_t, Req> reqs;
    std::deque<uint64_t> pending;
    std::vector<uint64_t> active;  // admission order
    uint64_t in_use = 0;     // accumulator bytes of admitted requests
    uint64_t queue_bytes = 0;  // the shared path-state queues charged to the budget
    float4* arena_ptr = nullptr;
    uint64_t arena_len = 0;  // float4
    Arena arena;
    uint32_t chunk = 0;
    std::chrono::steady_clock::time_point t0;
    int fatal = CT_OK;
    std::string fatal_msg;
    std::string err;  // message of the last failed call on the server (under mu)

    double now() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    void loop();
    void copy_loop();
    void finish(Req& q, int status);
};

// Caller holds mu.
void ct_server::finish(Req& q, int status) {
    arena.release(q.acc_off, (uint64_t)q.r.width * q.r.height);
    in_use -= q.footprint;
    q.status = status;
    q.state = DONE;
    q.done_ms = now();
    cv.notify_all();
}

void ct_server::copy_loop() {
    (void)hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        if (to_copy.empty()) {
            if (worker_done) return;
            cv.wait(lk);
            continue;
        }
        Req& q = reqs[to_copy.front()];
        to_copy.pop_front();
        const size_t n = (size_t)q.r.width * q.r.height * 4;
        q.result.resize(n);
        const uint64_t off = q.acc_off;
        float* dst = q.result.data();  // map nodes stay put while unlocked
        lk.unlock();
        // the chunk that finished this request was synchronised by the worker
        hipError_t e = hipMemcpyAsync(dst, arena_ptr + off, n * sizeof(float), hipMemcpyDeviceToHost, copy_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(copy_stream);
        lk.lock();
        finish(q, e == hipSuccess ? CT_OK : CT_ERR_DEVICE);
    }
}

void ct_server::loop() {
    (void)hipSetDevice(ctx->device);
    std::vector<ViewDesc> views;
    std::vector<Run> runs;
    std::vector<uint64_t> ids;
    // accumulators to zero before the chunk: (float4 offset, pixels), copied
    // out under the lock (submit() may insert into reqs concurrently)
    std::vector<std::pair<uint64_t, uint64_t>> fresh;
    std::vector<std::pair<double, uint64_t>> order;  // (priority, id)
    ct_frame_opts chunk_opts{};
    while (true) {
        {
            std::unique_lock<std::mutex> lk(mu);
            // drop cancelled requests
            for (auto it = active.begin(); it != active.end();) {
                Req& q = reqs[*it];
                if (q.cancel) {
                    finish(q, CT_ERR_CANCELLED);
                    it = active.erase(it);
                } else {
                    ++it;
                }
            }
            // admit in arrival order while the footprint fits
            while (!pending.empty() && (opts.max_active <= 0 || (int)active.size() < opts.max_active)) {
                Req& q = reqs[pending.front()];
                if (q.cancel) {
                    q.state = DONE;
                    q.status = CT_ERR_CANCELLED;
                    q.done_ms = now();
                    pending.pop_front();
                    continue;
                }
                uint64_t off = 0;
                if (!arena.alloc((uint64_t)q.r.width * q.r.height, &off)) break;
                q.acc_off = off;
                q.state = ACTIVE;
                q.admit_ms = now();
                in_use += q.footprint;
                active.push_back(q.id);
                pending.pop_front();
                cv.notify_all();
            }
            if (active.empty()) {
                if (stop) {
                    worker_done = true;
                    cv.notify_all();
                    return;
                }
                cv.wait(lk);
                continue;
            }
            // pack one chunk: smallest remaining work first, with aging; the
            // first request sets the chunk's options
            views.clear();
            runs.clear();
            ids.clear();
            fresh.clear();
            order.clear();
            const double t_now = now();
            for (uint64_t id : active) {
                const Req& q = reqs[id];
                const double wait = t_now - q.submit_ms;
                order.emplace_back((double)(q.total - q.next) / (1.0 + wait / 50.0), id);
            }
            std::sort(order.begin(), order.end());
            uint64_t room = chunk;
            bool have_opts = false;
            for (const auto& oi : order) {
                const uint64_t id = oi.second;
                if (room == 0) break;
                Req& q = reqs[id];
                const uint64_t take = std::min<uint64_t>(room, q.total - q.next);
                if (take == 0) continue;
                if (!have_opts) {
                    chunk_opts = q.opts;
                    have_opts = true;
                } else if (!same_opts(q.opts, chunk_opts)) {
                    continue;
                }
                views.push_back(make_view(q.r.camera, q.r.width, q.r.height, q.r.spp, q.r.frame_base, q.acc_off));
                runs.push_back({(uint32_t)(views.size() - 1), (uint32_t)q.next, (uint32_t)take});
                ids.push_back(id);
                if (q.next == 0) fresh.emplace_back(q.acc_off, (uint64_t)q.r.width * q.r.height);
                q.next += take;
                room -= take;
            }
        }
        // zero accumulators of requests starting in this chunk, then trace
        int rc = CT_OK;
        for (const auto& f : fresh)
            if (rc == CT_OK && hipMemsetAsync(arena_ptr + f.first, 0, f.second * sizeof(float4), ctx->ln[0].stream) != hipSuccess)
                rc = CT_ERR_DEVICE;
        ChunkStats st;
        if (rc == CT_OK)
            rc = run_paths(ctx, views, runs, chunk, arena_ptr, chunk_opts, kRunGraph, &st);
        std::unique_lock<std::mutex> lk(mu);
        if (rc != CT_OK) {
            fatal = rc;
            fatal_msg = ctx->err;
            for (uint64_t id : active) finish(reqs[id], rc);
            active.clear();
            continue;
        }
        // rays are attributed to requests in proportion to their path samples
        uint64_t paths = 0;
        for (const Run& r : runs) paths += r.n;
        const uint64_t rays = st.primary + st.extension + st.shadow;
        for (size_t k = 0; k < ids.size(); ++k)
            reqs[ids[k]].rays += paths ? rays * runs[k].n / paths : 0;
        for (auto it = active.begin(); it != active.end();) {
            Req& q = reqs[*it];
            if (q.next >= q.total) {
                q.state = COPYING;
                to_copy.push_back(q.id);
                it = active.erase(it);
                cv.notify_all();
            } else {
                ++it;
            }
        }
    }
}

extern "C" {

int ct_server_create(ct_ctx* ctx, const ct_server_opts* opts, ct_server** out) {
    if (!ctx || !out) return CT_ERR_INVALID_ARG;
    *out = nullptr;
    if (!ctx->has_scene) return fail(ctx, CT_ERR_NO_SCENE, "no scene uploaded");
    ct_server* s = new ct_server();
    s->ctx = ctx;
    if (opts) s->opts = *opts;
    if (s->opts.mem_budget == 0) s->opts.mem_budget = 8ull << 30;
    s->chunk = s->opts.chunk_paths > 0 ? (uint32_t)s->opts.chunk_paths : kServerChunk;
    (void)hipSetDevice(ctx->device);
    int r = ensure_paths(ctx, s->chunk);
    if (r == CT_OK) {
        // the shared queues come out of the budget (at the size they hold for
        // this server's chunks), the accumulator arena is the rest
        s->queue_bytes = (uint64_t)kLanes * std::min<uint32_t>(ctx->lane_cap, s->chunk) * kPathSlotBytes;
        if (s->queue_bytes + sizeof(float4) > s->opts.mem_budget)
            r = fail(ctx, CT_ERR_OUT_OF_MEMORY, "server budget smaller than its shared path-state queues");
    }
    if (r == CT_OK) {
        s->arena_len = (s->opts.mem_budget - s->queue_bytes) / sizeof(float4);
        r = dalloc(ctx, (void**)&s->arena_ptr, s->arena_len * sizeof(float4), "server accumulator arena");
    }
    if (r != CT_OK) {
        delete s;
        return r;
    }
    if (r == CT_OK && hipStreamCreate(&s->copy_stream) != hipSuccess)
        r = fail(ctx, CT_ERR_DEVICE, "copy stream");
    if (r != CT_OK) {
        dfree(ctx, s->arena_ptr);
        delete s;
        return r;
    }
    s->arena = Arena(s->arena_len);
    s->t0 = std::chrono::steady_clock::now();
    s->worker = std::thread([s] { s->loop(); });
    s->copier = std::thread([s] { s->copy_loop(); });
    *out = s;
    return CT_OK;
}

int ct_server_submit(ct_server* s, const ct_request* req, uint64_t* id) {
    if (!s || !req || !id) return CT_ERR_INVALID_ARG;
    const ct_request& r = *req;
    std::lock_guard<std::mutex> lk(s->mu);
    auto refuse = [&](int code, const char* why) {
        s->err = why;
        return code;
    };
    if (r.width <= 0 || r.height <= 0 || r.spp <= 0) return refuse(CT_ERR_INVALID_ARG, "bad request size or spp");
    if (!camera_ok(&r.camera)) return refuse(CT_ERR_INVALID_ARG, "camera is not finite or degenerate");
    const uint64_t total = (uint64_t)r.width * r.height * (uint64_t)r.spp;
    if (total >= (1ull << 32)) return refuse(CT_ERR_INVALID_ARG, "width*height*spp must be < 2^32");
    const uint64_t fp = (uint64_t)r.width * r.height * sizeof(float4);
    if (fp > s->arena_len * sizeof(float4))
        return refuse(CT_ERR_ADMISSION, "request's accumulator can never fit the server's arena");
    if (s->fatal != CT_OK) return refuse(s->fatal, s->fatal_msg.c_str());
    ct_frame_opts o = r.opts;
    o.spp = 1;
    const char* why = "";
    const int v = check_opts(s->ctx, o, &why);  // scene flags only change while no server exists
    if (v != CT_OK) return refuse(v, why);
    Req q;
    q.id = s->next_id++;
    q.r = r;
    q.opts = o;
    q.total = total;
    q.footprint = fp;
    q.submit_ms = s->now();
    s->reqs[q.id] = q;
    s->pending.push_back(q.id);
    *id = q.id;
    s->cv.notify_all();
    return CT_OK;
}

static void fill_result(const Req& q, ct_result* res) {
    if (!res) return;
    res->status = q.status;
    res->width = q.r.width;
    res->height = q.r.height;
    res->spp = q.r.spp;
    res->submit_ms = q.submit_ms;
    res->admit_ms = q.admit_ms;
    res->done_ms = q.done_ms;
    res->rays = q.rays;
}

int ct_server_poll(ct_server* s, uint64_t id, int32_t* done, ct_result* res, float* accum_out, size_t n_floats) {
    if (!s || !done) return CT_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return CT_ERR_NOT_FOUND;
    Req& q = it->second;
    *done = q.state == DONE ? 1 : 0;
    if (q.state != DONE) return CT_OK;
    fill_result(q, res);
    if (accum_out && q.status == CT_OK) {
        if (n_floats < q.result.size()) return CT_ERR_INVALID_ARG;
        std::memcpy(accum_out, q.result.data(), q.result.size() * sizeof(float));
    }
    return CT_OK;
}

int ct_server_wait(ct_server* s, uint64_t id, ct_result* res, float* accum_out, size_t n_floats) {
    if (!s) return CT_ERR_INVALID_ARG;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) return CT_ERR_NOT_FOUND;
        s->cv.wait(lk, [&] { return s->reqs[id].state == DONE; });
    }
    int32_t done = 0;
    return ct_server_poll(s, id, &done, res, accum_out, n_floats);
}

int ct_server_cancel(ct_server* s, uint64_t id) {
    if (!s) return CT_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return CT_ERR_NOT_FOUND;
    it->second.cancel = true;
    s->cv.notify_all();
    return CT_OK;
}

double ct_server_clock_ms(const ct_server* s) { return s ? s->now() : 0.0; }

const char* ct_server_last_error(ct_server* s) {
    if (!s) return "null server";
    std::lock_guard<std::mutex> lk(s->mu);
    return s->err.c_str();
}

void ct_server_destroy(ct_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (uint64_t id : s->pending) s->reqs[id].cancel = true;
        for (uint64_t id : s->active) s->reqs[id].cancel = true;
        s->cv.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    if (s->copier.joinable()) s->copier.join();
    sync_lanes(s->ctx);
    if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
    dfree(s->ctx, s->arena_ptr);
    delete s;
}

}  // extern "C"
