// The per-context tuning table and the measurements that fill it.
//
// Three choices of this implementation depend on the size class of the work (0: below 4,096 primitives or keys,
// 1: below 65,536, 2: below 1,048,576, 3: larger):
//   sort_items     : keys per thread of the radix sort's tiles (sort.hip) -- the order never changes;
//   trav_lds_depth : traversal stack entries kept in LDS (traverse.hip) -- the hits never change;
//   trav_chunk     : single-BVH launches as wave-owned ray rows with lanes refilled (traverse.hip) -- the hits
//                    never change;
//   ploc_radius    : the PLOC search distance of builds that ask for distance 0 ("auto").  This one changes the
//                    tree, so it is chosen from tree costs only: the smallest distance whose CWBVH cost (the
//                    conversion's own cost model at the root, bvh2_to_cwbvh.rs:220-344) is within a tolerance of a
//                    reference -- the cost at a baseline distance (the fastest build as good as the baseline's), or
//                    the lowest over all six.  The same triangles always give the same choice.
// The first two are timed: each candidate runs once to warm up and three times on the context stream between HIP
// events, and the fastest minimum wins.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <new>

#include "common.hpp"
#include "cwdp.hpp"
#include "internal.hpp"

namespace obvhs {

namespace {

// Seeded 63-bit keys (an integer hash of the index), the shape of Morton codes for the sort timing.
__global__ void k_tune_keys(uint64_t* keys, uint32_t n) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    uint64_t x = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    x ^= x >> 31;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 29;
    keys[i] = x >> 1;
}

constexpr uint32_t CANDIDATE_TILES[3] = {4u, 8u, 16u};
constexpr uint32_t CANDIDATE_DEPTHS[3] = {4u, 8u, 16u};
constexpr uint32_t CANDIDATE_RADII[6] = {1u, 2u, 6u, 14u, 24u, 32u};

bool valid_radius(uint32_t r) { return r == 1 || r == 2 || r == 6 || r == 14 || r == 24 || r == 32; }
bool valid_tile(uint32_t t) { return t == 4 || t == 8 || t == 16; }

// Minimum of three timed runs of `run` after one warm-up run.
template <class F>
obvhs_status time_min(obvhs_context* ctx, F run, float* out_ms) {
    hipStream_t st = ctx->stream;
    float best = INFINITY;
    for (int k = 0; k < 4; ++k) {
        if (hipEventRecord(ctx->ev_start, st) != hipSuccess) return set_hip_error(ctx, hipGetLastError(), "event");
        const obvhs_status s = run();
        if (s != OBVHS_OK) return s;
        if (hipEventRecord(ctx->ev_stop, st) != hipSuccess) return set_hip_error(ctx, hipGetLastError(), "event");
        if (hipEventSynchronize(ctx->ev_stop) != hipSuccess) return set_hip_error(ctx, hipGetLastError(), "sync");
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop) != hipSuccess)
            return set_hip_error(ctx, hipGetLastError(), "elapsed");
        if (k > 0) best = std::min(best, ms);
    }
    *out_ms = best;
    return OBVHS_OK;
}

#define TUNE_TRY(expr, where)                                        \
    do {                                                             \
        hipError_t _e = (expr);                                      \
        if (_e != hipSuccess) return set_hip_error(ctx, _e, where);  \
    } while (0)

}  // namespace

}  // namespace obvhs

using namespace obvhs;

extern "C" {

uint32_t obvhs_size_class(uint64_t n) { return size_class(n); }

obvhs_status obvhs_tune_get(const obvhs_context* ctx, obvhs_tune_table* out) {
    if (!ctx || !out) return OBVHS_ERR_INVALID_ARG;
    *out = ctx->tune;
    return OBVHS_OK;
}

obvhs_status obvhs_tune_set(obvhs_context* ctx, const obvhs_tune_table* table) {
    if (!ctx || !table) return OBVHS_ERR_INVALID_ARG;
    for (int c = 0; c < 4; ++c)
        if (!valid_radius(table->ploc_radius[c]) || !valid_tile(table->sort_items[c]) ||
            !valid_tile(table->trav_lds_depth[c]) || table->trav_chunk[c] > 1u)
            return OBVHS_ERR_INVALID_ARG;
    ctx->tune = *table;
    return OBVHS_OK;
}

obvhs_status obvhs_tune_sort(obvhs_context* ctx, uint32_t n, uint32_t* out_items) {
    if (!ctx || n == 0) return OBVHS_ERR_INVALID_ARG;
    TUNE_TRY(hipSetDevice(ctx->device), "set device");
    hipStream_t st = ctx->stream;
    const size_t kb = ((size_t)n * 8 + 255) & ~(size_t)255, vb = ((size_t)n * 4 + 255) & ~(size_t)255;
    TUNE_TRY(ctx->tune_scratch.ensure(2 * kb + vb + radix_sort_scratch_bytes(n)), "alloc tuning scratch");
    char* base = ctx->tune_scratch.as<char>();
    uint64_t* keys0 = reinterpret_cast<uint64_t*>(base);
    uint64_t* keys = reinterpret_cast<uint64_t*>(base + kb);
    uint32_t* vals = reinterpret_cast<uint32_t*>(base + 2 * kb);
    void* scratch = base + 2 * kb + vb;
    k_tune_keys<<<(n + 255) / 256, 256, 0, st>>>(keys0, n);
    TUNE_TRY(hipGetLastError(), "tuning keys");
    uint32_t best = 8u;
    float best_ms = INFINITY;
    for (uint32_t items : CANDIDATE_TILES) {
        float ms = 0.0f;
        const obvhs_status s = time_min(
            ctx,
            [&]() -> obvhs_status {
                // the copy belongs to every candidate alike; the sort dominates
                TUNE_TRY(hipMemcpyAsync(keys, keys0, (size_t)n * 8, hipMemcpyDeviceToDevice, st), "copy keys");
                TUNE_TRY(radix_sort_pairs(keys, vals, true, n, 63, scratch, st, (int)items), "sort");
                return OBVHS_OK;
            },
            &ms);
        if (s != OBVHS_OK) return s;
        if (ms < best_ms) {
            best_ms = ms;
            best = items;
        }
    }
    ctx->tune.sort_items[size_class(n)] = best;
    if (out_items) *out_items = best;
    return OBVHS_OK;
}

obvhs_status obvhs_tune_traversal(obvhs_context* ctx, const obvhs_cwbvh* bvh, const obvhs_triangle* d_bvh_tris,
                                  const obvhs_ray* d_rays, uint32_t n, int any_hit, uint32_t* out_depth) {
    if (!ctx || !bvh || n == 0 || !d_rays || (bvh->node_count && !d_bvh_tris)) return OBVHS_ERR_INVALID_ARG;
    TUNE_TRY(hipSetDevice(ctx->device), "set device");
    TravScratch ts;
    const obvhs_status ss = traverse_scratch(ctx, ts, n);
    if (ss != OBVHS_OK) return ss;
    TUNE_TRY(ctx->tune_scratch.ensure((size_t)n * sizeof(obvhs_ray_hit)), "alloc tuning scratch");
    obvhs_ray_hit* hits = ctx->tune_scratch.as<obvhs_ray_hit>();
    uint8_t* miss = ctx->tune_scratch.as<uint8_t>();
    uint32_t best = 8u;
    bool best_chunked = false;
    float best_ms = INFINITY;
    for (int chunked = 0; chunked < 2; ++chunked)
    for (uint32_t depth : CANDIDATE_DEPTHS) {
        ts.lds_depth = depth;
        ts.chunked = chunked != 0;
        float ms = 0.0f;
        const obvhs_status s = time_min(
            ctx,
            [&]() -> obvhs_status {
                TUNE_TRY(launch_traverse(bvh->nodes.as<CwNode>(), bvh->node_count, bvh->prim_count, d_bvh_tris,
                                         d_rays, n, any_hit ? 1 : 0, any_hit ? nullptr : hits,
                                         any_hit ? miss : nullptr, nullptr, nullptr, ts, ctx->stream),
                         "traverse");
                return OBVHS_OK;
            },
            &ms);
        if (s != OBVHS_OK) return s;
        if (ms < best_ms) {
            best_ms = ms;
            best = depth;
            best_chunked = chunked != 0;
        }
    }
    ctx->tune.trav_lds_depth[size_class(bvh->prim_count)] = best;
    ctx->tune.trav_chunk[size_class(bvh->prim_count)] = best_chunked ? 1u : 0u;
    if (out_depth) *out_depth = best;
    return OBVHS_OK;
}

obvhs_status obvhs_tune_ploc_radius(obvhs_context* ctx, const obvhs_triangle* d_tris, uint32_t n,
                                    uint32_t max_prims_per_leaf, uint32_t baseline, float tolerance,
                                    uint32_t* out_radius, float* out_costs) {
    if (!ctx || !d_tris || n < 2 || n >= (1u << 31) || !(tolerance >= 0.0f)) return OBVHS_ERR_INVALID_ARG;
    if (baseline != 0u && !valid_radius(baseline)) return OBVHS_ERR_INVALID_ARG;
    TUNE_TRY(hipSetDevice(ctx->device), "set device");
    const uint32_t mp = std::min(3u, std::max(1u, max_prims_per_leaf));  // cwbvh/builder.rs:74
    struct Held {  // the candidate trees share one handle, released on every exit
        obvhs_bvh2* b = new (std::nothrow) obvhs_bvh2();
        ~Held() { obvhs_bvh2_destroy(b); }
    } held;
    if (!held.b) return OBVHS_ERR_OOM;
    obvhs_bvh2& bvh = *held.b;
    float cost[6];
    for (int k = 0; k < 6; ++k) {
        // the conversion's cost table is evaluated by the PLOC merges themselves (fused)
        const obvhs_status s = ploc_build(ctx, d_tris, nullptr, nullptr, n, CANDIDATE_RADII[k], false, 0u, mp, &bvh);
        if (s != OBVHS_OK) return s;
        Node32 root;
        TUNE_TRY(hipMemcpy(&root, bvh.nodes.ptr, sizeof(Node32), hipMemcpyDeviceToHost), "read root");
        const DpView dp = bvh_dp_view(&bvh, mp);
        TUNE_TRY(hipMemcpy(&cost[k], &dp.rec[dkey(root.first_index)].cost[0], 4, hipMemcpyDeviceToHost),
                 "read root cost");
    }
    float ref = INFINITY;
    for (int k = 0; k < 6; ++k)
        if (baseline == 0u ? cost[k] < ref : CANDIDATE_RADII[k] == baseline) ref = cost[k];
    uint32_t pick = baseline != 0u ? baseline : 14u;
    for (int k = 0; k < 6; ++k)
        if (cost[k] <= ref * (1.0f + tolerance)) {
            pick = CANDIDATE_RADII[k];
            break;
        }
    ctx->tune.ploc_radius[size_class(n)] = pick;
    if (out_radius) *out_radius = pick;
    if (out_costs) memcpy(out_costs, cost, sizeof(cost));
    return OBVHS_OK;
}

}  // extern "C"
