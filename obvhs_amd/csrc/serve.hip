// Serving: CWBVHs resident in HBM, ray requests coalesced into shared traversal launches (continuous batching),
// refit / rebuild updates overlapped on a second stream, admission by an HBM budget.
//
// No single reference counterpart: each request does exactly what the crate's CwBvh::ray_traverse /
// ray_traverse_miss (cwbvh/mod.rs:169-225), Bvh2::refit_all + bvh2_to_cwbvh (bvh2/mod.rs:506-569,
// cwbvh/bvh2_to_cwbvh.rs:490-510) or build_cwbvh_from_tris (cwbvh/builder.rs:20-85) do, so a request's result is the
// same whether it ran alone or coalesced with others.
//
// Threads and streams:
//   * the scheduler thread owns all queues.  It moves submitted requests into per-scene FIFOs (submission order
//     per scene), completes finished launches (event queries, in any order), admits requests in submission order
//     while the HBM budget allows, and launches one traversal per mode over every ready request at the head of its
//     scene's FIFO (each request a segment of the launch: its own BVH, rays and outputs) on a free trace stream:
//     max_inflight trace contexts (stream + traversal spill columns each), so a launch that starts while a large
//     one runs shares the GPU with it instead of waiting for it to finish;
//   * update threads (one context and stream each) run refits and rebuilds, several scenes at a time.  An update
//     writes the scene's other version slot, after every trace stream's last launch that read that slot (event
//     waits), and becomes the scene's current version when it completes; traces submitted after it wait for it,
//     traces submitted before it keep reading the old slot.
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

enum ReqKind : uint32_t {
    REQ_TRACE_CLOSEST = OBVHS_REQ_TRACE_CLOSEST,
    REQ_TRACE_MISS = OBVHS_REQ_TRACE_MISS,
    REQ_REFIT = OBVHS_REQ_REFIT,
    REQ_REBUILD = OBVHS_REQ_REBUILD,
};

struct Req {
    uint64_t id = 0;
    uint32_t kind = 0;
    uint32_t scene = 0;
    const void* rays = nullptr;  // device
    void* out = nullptr;         // device
    const void* tris = nullptr;  // device, updates
    uint32_t n = 0;
    uint64_t bytes = 0;          // footprint counted against the budget while in flight
    bool admitted = false;
    bool done = false;
    obvhs_status status = OBVHS_OK;
    uint32_t version = 0;        // scene version the request read (traces) or produced (updates)
    uint32_t batch_requests = 0; // requests in the launch that carried it
    uint64_t batch = 0;
    int64_t t_submit = 0, t_start = 0, t_done = 0;
};

constexpr uint32_t MAX_TRACE_LANES = 8;  // trace contexts (streams) at most: max_inflight is capped to this

struct Version {
    obvhs_cwbvh* cw = nullptr;  // owns its BVH2
    obvhs::DevBuf tris;         // triangles in input order
    obvhs::DevBuf bvh_tris;     // triangles in primitive_indices order
    uint32_t n = 0;
    hipEvent_t last_use[MAX_TRACE_LANES] = {};  // per trace stream: recorded after each launch reading this slot
    bool used[MAX_TRACE_LANES] = {};
    uint64_t bytes = 0;
};

struct Scene {
    uint32_t id = 0;
    obvhs_build_params params{};
    Version v[2];
    int cur = 0;
    uint32_t version = 0;
    bool updating = false;
    std::deque<Req*> queue;
};

struct Batch {
    uint64_t id = 0;
    uint32_t lane = 0;  // trace context (stream) it runs on
    hipEvent_t done = nullptr;
    std::vector<Req*> reqs;
};

int64_t now_ns(Clock::time_point t0) {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now() - t0).count();
}

uint64_t version_bytes(const Version& v) {
    uint64_t b = v.tris.bytes + v.bvh_tris.bytes;
    if (v.cw) {
        b += v.cw->nodes.bytes + v.cw->prim_indices.bytes;
        if (v.cw->owned_bvh2) {
            const obvhs_bvh2* b2 = v.cw->owned_bvh2;
            b += b2->nodes.bytes + b2->prim_indices.bytes + b2->hist.bytes + b2->dp.bytes + b2->codes.bytes +
                 b2->sorted_index.bytes;
        }
    }
    return b;
}

}  // namespace

struct obvhs_server {
    int device = 0;
    obvhs_server_config cfg{};
    std::vector<obvhs_context*> tctx;  // trace contexts: one stream + traversal spill columns per in-flight launch
    std::vector<uint32_t> free_lanes;  // trace contexts without a launch in flight
    std::vector<obvhs_context*> uctx;  // one context (stream + build scratch) per updater thread, then one for
                                       // add_scene
    Clock::time_point t0 = Clock::now();

    std::mutex mu;
    std::mutex setup_mu;  // serialises add_scene's use of its context
    std::condition_variable cv_work, cv_done, cv_update;
    std::deque<Req*> incoming;
    std::unordered_map<uint64_t, Req*> reqs;
    std::vector<Scene*> scenes;
    uint64_t next_id = 1;
    std::atomic<bool> stop{false};

    // scheduler-owned
    std::deque<Batch*> inflight;
    std::vector<Batch*> free_batches;
    std::deque<Req*> order;  // admitted-or-waiting requests in submission order (admission is FIFO)
    uint64_t resident = 0, inflight_bytes = 0;
    uint64_t next_batch = 1;
    obvhs::DevBuf seg_dev;   // TraceSegment tables: 2 per trace context (one per mode) * max_segments
    obvhs::TraceSegment* seg_host = nullptr;
    std::vector<obvhs::TravScratch> trav;  // spill columns of each trace context

    // update thread
    std::deque<Req*> updates;
    std::deque<Req*> finished_updates;

    // statistics
    uint64_t stat_launches = 0, stat_rays = 0, stat_segments = 0;

    std::thread scheduler;
    std::vector<std::thread> updaters;
};

namespace obvhs {
namespace {

constexpr uint32_t MAX_SEGMENTS = 1024;  // requests per launch

void complete(obvhs_server* S, Req* r, obvhs_status st) {
    r->status = st;
    r->t_done = now_ns(S->t0);
    r->done = true;
}

// Runs one update on the update context (blocking).  Returns the status; on success the scene's other slot holds the
// new version.
obvhs_status run_update(obvhs_server* S, obvhs_context* ctx, Scene* sc, Req* r, int slot) {
    hipStream_t st = ctx->stream;
    Version& dst = sc->v[slot];
    Version& src = sc->v[sc->cur];
    for (uint32_t k = 0; k < MAX_TRACE_LANES; ++k) {  // every trace stream's last launch that read this slot first
        if (dst.used[k] && hipStreamWaitEvent(st, dst.last_use[k], 0) != hipSuccess) return OBVHS_ERR_DEVICE;
    }
    // a refit keeps the triangle count of the version it refits (checked when it runs: earlier rebuilds may change it)
    if (r->kind == REQ_REFIT && r->n != src.n) return OBVHS_ERR_INVALID_ARG;
    const uint32_t n = r->n;
    if (dst.tris.ensure((size_t)n * 48 + 48) != hipSuccess) return OBVHS_ERR_OOM;
    if (n && hipMemcpyAsync(dst.tris.ptr, r->tris, (size_t)n * 48, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return OBVHS_ERR_DEVICE;
    dst.n = n;
    if (!dst.cw) {
        dst.cw = new (std::nothrow) obvhs_cwbvh();
        if (!dst.cw) return OBVHS_ERR_OOM;
    }
    obvhs_status s = OBVHS_OK;
    if (r->kind == REQ_REBUILD) {
        obvhs_cwbvh* out = nullptr;
        s = obvhs_build_cwbvh_from_tris(ctx, dst.tris.as<obvhs_triangle>(), n, &sc->params, dst.cw, &out, nullptr);
    } else {
        // refit: the current BVH2 copied into this slot, leaf boxes from the moved triangles, refit_all, convert
        const obvhs_bvh2* b = src.cw->owned_bvh2;
        if (!b) return OBVHS_ERR_INVALID_ARG;
        if (!dst.cw->owned_bvh2) {
            dst.cw->owned_bvh2 = new (std::nothrow) obvhs_bvh2();
            if (!dst.cw->owned_bvh2) return OBVHS_ERR_OOM;
        }
        obvhs_bvh2* d = dst.cw->owned_bvh2;
        if (d->nodes.ensure(b->nodes.bytes) != hipSuccess ||
            d->prim_indices.ensure(b->prim_indices.bytes) != hipSuccess ||
            d->hist.ensure(b->hist.bytes ? b->hist.bytes : 4) != hipSuccess)
            return OBVHS_ERR_OOM;
        if (hipMemcpyAsync(d->nodes.ptr, b->nodes.ptr, (size_t)b->node_count * 32, hipMemcpyDeviceToDevice, st) !=
                hipSuccess ||
            hipMemcpyAsync(d->prim_indices.ptr, b->prim_indices.ptr, (size_t)b->prim_count * 4,
                           hipMemcpyDeviceToDevice, st) != hipSuccess ||
            (b->hist.bytes && hipMemcpyAsync(d->hist.ptr, b->hist.ptr, b->hist.bytes, hipMemcpyDeviceToDevice, st) !=
                                  hipSuccess))
            return OBVHS_ERR_DEVICE;
        d->node_count = b->node_count;
        d->prim_count = b->prim_count;
        d->max_depth = b->max_depth;
        d->ploc_iterations = b->ploc_iterations;
        d->cluster_sum = b->cluster_sum;
        d->pass_ordered = b->pass_ordered;
        d->dp_max_prims = 0;
        s = bvh2_refit_from_triangles(ctx, d, dst.tris.as<obvhs_triangle>());
        uint32_t mp = sc->params.max_prims_per_leaf;
        mp = mp < 1u ? 1u : (mp > 3u ? 3u : mp);
        if (s == OBVHS_OK) s = cwbvh_convert(ctx, d, mp, true, dst.cw);
    }
    if (s != OBVHS_OK) return s;
    if (dst.bvh_tris.ensure((size_t)dst.cw->prim_count * 48 + 48) != hipSuccess) return OBVHS_ERR_OOM;
    if (launch_reorder_triangles(dst.tris.as<obvhs_triangle>(), dst.cw->prim_indices.as<uint32_t>(),
                                 dst.cw->prim_count, dst.bvh_tris.as<obvhs_triangle>(), st) != hipSuccess)
        return OBVHS_ERR_DEVICE;
    if (hipStreamSynchronize(st) != hipSuccess) return OBVHS_ERR_DEVICE;
    dst.bytes = version_bytes(dst);
    return OBVHS_OK;
}

void updater_main(obvhs_server* S, obvhs_context* ctx) {
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        S->cv_update.wait(lk, [&] { return S->stop.load() || !S->updates.empty(); });
        if (S->updates.empty() && S->stop.load()) return;
        Req* r = S->updates.front();
        S->updates.pop_front();
        Scene* sc = S->scenes[r->scene];
        const int slot = 1 - sc->cur;
        const uint64_t old_bytes = sc->v[slot].bytes;
        lk.unlock();
        r->t_start = now_ns(S->t0);
        const obvhs_status st = run_update(S, ctx, sc, r, slot);
        lk.lock();
        S->resident = S->resident - std::min(S->resident, old_bytes) + sc->v[slot].bytes;
        if (st == OBVHS_OK) {
            sc->cur = slot;
            sc->version += 1;
        }
        r->version = sc->version;
        S->finished_updates.push_back(r);
        r->status = st;
        S->cv_work.notify_all();
    }
}

// Admission in submission order while the budget allows (the first request that does not fit stops admission, so
// large requests are not starved by small ones).
void admit(obvhs_server* S) {
    while (!S->order.empty()) {
        Req* r = S->order.front();
        if (!r->admitted) {
            if (S->resident + S->inflight_bytes + r->bytes > S->cfg.hbm_budget_bytes && S->inflight_bytes > 0) break;
            r->admitted = true;
            S->inflight_bytes += r->bytes;
        }
        S->order.pop_front();
    }
}

void launch_batch(obvhs_server* S, std::vector<Req*>& ready) {
    const uint32_t lane = S->free_lanes.back();
    S->free_lanes.pop_back();
    obvhs_context* tc = S->tctx[lane];
    hipStream_t st = tc->stream;
    Batch* b = nullptr;
    if (!S->free_batches.empty()) {
        b = S->free_batches.back();
        S->free_batches.pop_back();
    } else {
        b = new Batch();
        (void)hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
    }
    b->id = S->next_batch++;
    b->lane = lane;
    b->reqs = ready;
    const int64_t t = now_ns(S->t0);
    for (int mode = 0; mode < 2; ++mode) {
        const uint32_t slot = 2 * lane + (uint32_t)mode;  // this lane's table for the mode: free while it is idle
        TraceSegment* hs = S->seg_host + (size_t)slot * MAX_SEGMENTS;
        uint32_t nseg = 0, total = 0, largest = 0;
        for (Req* r : ready) {
            if ((r->kind == REQ_TRACE_MISS) != (mode == 1)) continue;
            Scene* sc = S->scenes[r->scene];
            const Version& v = sc->v[sc->cur];
            TraceSegment sg;
            sg.nodes = v.cw ? v.cw->nodes.ptr : nullptr;
            sg.tris = v.bvh_tris.ptr;
            sg.rays = r->rays;
            sg.out = r->out;
            sg.node_count = v.cw ? v.cw->node_count : 0u;
            sg.prim_count = v.cw ? v.cw->prim_count : 0u;
            largest = std::max(largest, sg.prim_count);
            sg.begin = total;
            sg.count = r->n;
            hs[nseg++] = sg;
            total += r->n;
            r->version = sc->version;
            r->t_start = t;
            r->batch = b->id;
            r->batch_requests = (uint32_t)ready.size();
        }
        if (nseg == 0) continue;
        TraceSegment* ds = S->seg_dev.as<TraceSegment>() + (size_t)slot * MAX_SEGMENTS;
        hipError_t e = hipMemcpyAsync(ds, hs, (size_t)nseg * sizeof(TraceSegment), hipMemcpyHostToDevice, st);
        TravScratch ts = S->trav[lane];
        ts.lds_depth = lds_depth_for(tc, largest);  // the depth tuned for the launch's largest BVH
        if (e == hipSuccess) e = launch_traverse_segments(ds, nseg, total, mode, ts, st);
        if (e != hipSuccess) {
            for (Req* r : ready)
                if ((r->kind == REQ_TRACE_MISS) == (mode == 1)) r->status = OBVHS_ERR_DEVICE;
        }
        S->stat_launches += 1;
        S->stat_rays += total;
        S->stat_segments += nseg;
    }
    for (Req* r : ready) {  // slots read by this launch
        Version& v = S->scenes[r->scene]->v[S->scenes[r->scene]->cur];
        (void)hipEventRecord(v.last_use[lane], st);
        v.used[lane] = true;
    }
    (void)hipEventRecord(b->done, st);
    S->inflight.push_back(b);
}

void scheduler_main(obvhs_server* S) {
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    std::vector<Req*> ready;
    while (!S->stop.load()) {
        bool progress = false;
        while (!S->incoming.empty()) {
            Req* r = S->incoming.front();
            S->incoming.pop_front();
            S->scenes[r->scene]->queue.push_back(r);
            S->order.push_back(r);
            progress = true;
        }
        // finished launches (the trace streams finish in any order)
        for (size_t i = 0; i < S->inflight.size();) {
            Batch* b = S->inflight[i];
            if (hipEventQuery(b->done) != hipSuccess) {
                ++i;
                continue;
            }
            S->inflight.erase(S->inflight.begin() + (std::ptrdiff_t)i);
            for (Req* r : b->reqs) {
                S->inflight_bytes -= std::min(S->inflight_bytes, r->bytes);
                complete(S, r, r->status);
            }
            b->reqs.clear();
            S->free_lanes.push_back(b->lane);
            S->free_batches.push_back(b);
            progress = true;
            S->cv_done.notify_all();
        }
        // finished updates
        while (!S->finished_updates.empty()) {
            Req* r = S->finished_updates.front();
            S->finished_updates.pop_front();
            Scene* sc = S->scenes[r->scene];
            sc->updating = false;
            if (!sc->queue.empty() && sc->queue.front() == r) sc->queue.pop_front();
            S->inflight_bytes -= std::min(S->inflight_bytes, r->bytes);
            complete(S, r, r->status);
            progress = true;
            S->cv_done.notify_all();
        }
        admit(S);
        // updates at the head of their scene's FIFO go to the update thread
        for (Scene* sc : S->scenes) {
            if (sc->updating || sc->queue.empty()) continue;
            Req* r = sc->queue.front();
            if ((r->kind == REQ_REFIT || r->kind == REQ_REBUILD) && r->admitted) {
                sc->updating = true;
                S->updates.push_back(r);
                S->cv_update.notify_one();
                progress = true;
            }
        }
        // one coalesced launch over every ready trace at the head of its scene's FIFO
        if (!S->free_lanes.empty()) {
            ready.clear();
            uint64_t rays = 0;
            for (Scene* sc : S->scenes) {
                while (!sc->queue.empty() && !sc->updating) {
                    Req* r = sc->queue.front();
                    if (!(r->kind == REQ_TRACE_CLOSEST || r->kind == REQ_TRACE_MISS) || !r->admitted) break;
                    if (!ready.empty() && (rays + r->n > S->cfg.max_batch_rays || ready.size() >= MAX_SEGMENTS)) break;
                    ready.push_back(r);
                    rays += r->n;
                    sc->queue.pop_front();
                }
                if (rays >= S->cfg.max_batch_rays || ready.size() >= MAX_SEGMENTS) break;
            }
            if (!ready.empty()) {
                launch_batch(S, ready);
                progress = true;
            }
        }
        if (!progress) {
            S->cv_work.wait_for(lk, std::chrono::microseconds(S->inflight.empty() ? 2000 : 20));
        } else {  // let submitters and waiters take the lock between busy iterations
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        }
    }
}

}  // namespace
}  // namespace obvhs

using namespace obvhs;

extern "C" {

obvhs_status obvhs_server_create(int device, const obvhs_server_config* cfg, obvhs_server** out) {
    if (!out) return OBVHS_ERR_INVALID_ARG;
    *out = nullptr;
    obvhs_server* S = new (std::nothrow) obvhs_server();
    if (!S) return OBVHS_ERR_OOM;
    S->device = device;
    S->cfg.hbm_budget_bytes = cfg && cfg->hbm_budget_bytes ? cfg->hbm_budget_bytes : (200ull << 30);
    S->cfg.max_batch_rays = cfg && cfg->max_batch_rays ? cfg->max_batch_rays : (1u << 23);
    S->cfg.max_inflight = cfg && cfg->max_inflight ? std::min(cfg->max_inflight, MAX_TRACE_LANES) : 2u;
    obvhs_status s = OBVHS_OK;
    S->trav.resize(S->cfg.max_inflight);
    for (uint32_t k = 0; s == OBVHS_OK && k < S->cfg.max_inflight; ++k) {
        obvhs_context* c = nullptr;
        s = obvhs_context_create(device, nullptr, &c);
        if (s == OBVHS_OK) {
            S->tctx.push_back(c);
            s = traverse_scratch(c, S->trav[k], S->cfg.max_batch_rays);
            S->free_lanes.push_back(S->cfg.max_inflight - 1 - k);
        }
    }
    S->cfg.update_streams = cfg && cfg->update_streams ? std::min(cfg->update_streams, 4u) : 2u;
    for (uint32_t u = 0; s == OBVHS_OK && u <= S->cfg.update_streams; ++u) {
        obvhs_context* c = nullptr;
        s = obvhs_context_create(device, nullptr, &c);
        if (s == OBVHS_OK) S->uctx.push_back(c);
    }
    const size_t nseg = (size_t)(2 * S->cfg.max_inflight) * MAX_SEGMENTS;
    if (s == OBVHS_OK && S->seg_dev.ensure(nseg * sizeof(TraceSegment)) != hipSuccess) s = OBVHS_ERR_OOM;
    if (s == OBVHS_OK &&
        hipHostMalloc(reinterpret_cast<void**>(&S->seg_host), nseg * sizeof(TraceSegment), hipHostMallocDefault) !=
            hipSuccess)
        s = OBVHS_ERR_OOM;
    if (s != OBVHS_OK) {
        obvhs_server_destroy(S);
        return s;
    }
    S->scheduler = std::thread(scheduler_main, S);
    for (uint32_t u = 0; u < S->cfg.update_streams; ++u) S->updaters.emplace_back(updater_main, S, S->uctx[u]);
    *out = S;
    return OBVHS_OK;
}

obvhs_status obvhs_server_destroy(obvhs_server* S) {
    if (!S) return OBVHS_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop.store(true);
    }
    S->cv_work.notify_all();
    S->cv_update.notify_all();
    if (S->scheduler.joinable()) S->scheduler.join();
    for (std::thread& t : S->updaters)
        if (t.joinable()) t.join();
    (void)hipSetDevice(S->device);
    for (obvhs_context* c : S->tctx) (void)hipStreamSynchronize(c->stream);
    for (obvhs_context* c : S->uctx) (void)hipStreamSynchronize(c->stream);
    for (Batch* b : S->inflight) {
        if (b->done) (void)hipEventDestroy(b->done);
        delete b;
    }
    for (Batch* b : S->free_batches) {
        if (b->done) (void)hipEventDestroy(b->done);
        delete b;
    }
    for (Scene* sc : S->scenes) {
        for (Version& v : sc->v) {
            if (v.cw) obvhs_cwbvh_destroy(v.cw);
            v.tris.release();
            v.bvh_tris.release();
            for (hipEvent_t& e : v.last_use)
                if (e) (void)hipEventDestroy(e);
        }
        delete sc;
    }
    for (auto& kv : S->reqs) delete kv.second;
    S->seg_dev.release();
    if (S->seg_host) (void)hipHostFree(S->seg_host);
    for (obvhs_context* c : S->tctx) obvhs_context_destroy(c);
    for (obvhs_context* c : S->uctx) obvhs_context_destroy(c);
    delete S;
    return OBVHS_OK;
}

obvhs_status obvhs_server_add_scene(obvhs_server* S, const obvhs_triangle* d_tris, uint32_t n,
                                    const obvhs_build_params* params, uint32_t* out_scene) {
    if (!S || !params || !out_scene || (n > 0 && !d_tris)) return OBVHS_ERR_INVALID_ARG;
    if (hipSetDevice(S->device) != hipSuccess) return OBVHS_ERR_DEVICE;
    Scene* sc = new (std::nothrow) Scene();
    if (!sc) return OBVHS_ERR_OOM;
    sc->params = *params;
    Req r;
    r.kind = REQ_REBUILD;
    r.tris = d_tris;
    r.n = n;
    for (Version& v : sc->v)
        for (hipEvent_t& e : v.last_use)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return OBVHS_ERR_DEVICE;
    // the initial version is built like a rebuild into slot 0 on the update context
    obvhs_status s;
    {
        std::lock_guard<std::mutex> ul(S->setup_mu);
        sc->cur = 1;  // run_update writes the other slot
        s = run_update(S, S->uctx.back(), sc, &r, 0);
        sc->cur = 0;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    if (s != OBVHS_OK) {
        for (Version& v : sc->v) {
            if (v.cw) obvhs_cwbvh_destroy(v.cw);
            v.tris.release();
            v.bvh_tris.release();
            for (hipEvent_t& e : v.last_use)
                if (e) (void)hipEventDestroy(e);
        }
        delete sc;
        return s;
    }
    sc->id = (uint32_t)S->scenes.size();
    S->resident += sc->v[0].bytes;
    S->scenes.push_back(sc);
    *out_scene = sc->id;
    return OBVHS_OK;
}

static obvhs_status submit(obvhs_server* S, Req* r, uint64_t* out_id) {
    std::lock_guard<std::mutex> lk(S->mu);
    if (r->scene >= S->scenes.size()) {
        delete r;
        return OBVHS_ERR_INVALID_ARG;
    }
    r->id = S->next_id++;
    r->t_submit = now_ns(S->t0);
    S->reqs[r->id] = r;
    S->incoming.push_back(r);
    if (out_id) *out_id = r->id;
    S->cv_work.notify_one();
    return OBVHS_OK;
}

obvhs_status obvhs_server_submit_trace(obvhs_server* S, uint32_t scene, uint32_t kind, const obvhs_ray* d_rays,
                                       uint32_t n, void* d_out, uint64_t* out_id) {
    if (!S || !out_id || (kind != OBVHS_REQ_TRACE_CLOSEST && kind != OBVHS_REQ_TRACE_MISS) ||
        (n > 0 && (!d_rays || !d_out)))
        return OBVHS_ERR_INVALID_ARG;
    Req* r = new (std::nothrow) Req();
    if (!r) return OBVHS_ERR_OOM;
    r->kind = kind;
    r->scene = scene;
    r->rays = d_rays;
    r->out = d_out;
    r->n = n;
    r->bytes = (uint64_t)n * (64 + (kind == OBVHS_REQ_TRACE_CLOSEST ? 16 : 1));
    return submit(S, r, out_id);
}

obvhs_status obvhs_server_submit_update(obvhs_server* S, uint32_t scene, uint32_t kind, const obvhs_triangle* d_tris,
                                        uint32_t n, uint64_t* out_id) {
    if (!S || !out_id || (kind != OBVHS_REQ_REFIT && kind != OBVHS_REQ_REBUILD) || (n > 0 && !d_tris))
        return OBVHS_ERR_INVALID_ARG;
    Req* r = new (std::nothrow) Req();
    if (!r) return OBVHS_ERR_OOM;
    r->kind = kind;
    r->scene = scene;
    r->tris = d_tris;
    r->n = n;
    r->bytes = (uint64_t)n * 48;
    return submit(S, r, out_id);
}

obvhs_status obvhs_server_wait(obvhs_server* S, uint64_t id, obvhs_request_info* info) {
    if (!S) return OBVHS_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return OBVHS_ERR_INVALID_ARG;
    Req* r = it->second;
    S->cv_done.wait(lk, [&] { return r->done; });
    if (info) {
        info->status = r->status;
        info->kind = r->kind;
        info->scene = r->scene;
        info->version = r->version;
        info->batch = r->batch;
        info->batch_requests = r->batch_requests;
        info->submit_ns = r->t_submit;
        info->start_ns = r->t_start;
        info->done_ns = r->t_done;
    }
    const obvhs_status st = r->status;
    S->reqs.erase(it);
    delete r;
    return st;
}

obvhs_status obvhs_server_replay(obvhs_server* S, const obvhs_server_request* list, uint32_t count,
                                 obvhs_request_info* infos) {
    if (!S || (count > 0 && (!list || !infos))) return OBVHS_ERR_INVALID_ARG;
    std::vector<uint64_t> ids(count, 0);
    std::vector<obvhs_status> sub(count, OBVHS_OK);
    const Clock::time_point start = Clock::now();
    for (uint32_t i = 0; i < count; ++i) {
        const obvhs_server_request& q = list[i];
        const Clock::time_point due = start + std::chrono::nanoseconds(q.arrival_ns);
        while (Clock::now() < due) {
            const auto left = due - Clock::now();
            if (left > std::chrono::microseconds(200))
                std::this_thread::sleep_for(left - std::chrono::microseconds(100));
        }
        if (q.kind == OBVHS_REQ_TRACE_CLOSEST || q.kind == OBVHS_REQ_TRACE_MISS)
            sub[i] = obvhs_server_submit_trace(S, q.scene, q.kind, static_cast<const obvhs_ray*>(q.rays), q.n, q.out,
                                               &ids[i]);
        else
            sub[i] = obvhs_server_submit_update(S, q.scene, q.kind, static_cast<const obvhs_triangle*>(q.tris), q.n,
                                                &ids[i]);
    }
    obvhs_status first = OBVHS_OK;
    for (uint32_t i = 0; i < count; ++i) {
        if (sub[i] != OBVHS_OK) {
            memset(&infos[i], 0, sizeof(obvhs_request_info));
            infos[i].status = sub[i];
        } else {
            obvhs_server_wait(S, ids[i], &infos[i]);
        }
        if (first == OBVHS_OK && infos[i].status != OBVHS_OK) first = infos[i].status;
    }
    return first;
}

obvhs_status obvhs_server_scene_info(obvhs_server* S, uint32_t scene, uint32_t* version, uint32_t* prim_count,
                                     uint32_t* cwbvh_nodes) {
    if (!S) return OBVHS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    if (scene >= S->scenes.size()) return OBVHS_ERR_INVALID_ARG;
    const Scene* sc = S->scenes[scene];
    const Version& v = sc->v[sc->cur];
    if (version) *version = sc->version;
    if (prim_count) *prim_count = v.cw ? v.cw->prim_count : 0u;
    if (cwbvh_nodes) *cwbvh_nodes = v.cw ? v.cw->node_count : 0u;
    return OBVHS_OK;
}

obvhs_status obvhs_server_get_stats(obvhs_server* S, obvhs_server_stats* st) {
    if (!S || !st) return OBVHS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(S->mu);
    st->launches = S->stat_launches;
    st->rays = S->stat_rays;
    st->segments = S->stat_segments;
    st->resident_bytes = S->resident;
    st->scenes = (uint32_t)S->scenes.size();
    return OBVHS_OK;
}

}  // extern "C"
