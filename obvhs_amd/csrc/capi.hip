// C ABI of the obvhs MI355X library (declared in include/obvhs_hip.h): contexts, device memory helpers, the
// PLOC / CWBVH builders and the traversal entry points.  Host orchestration only; kernels live in sort.hip,
// ploc.hip, cwbvh.hip and traverse.hip.
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "common.hpp"
#include "internal.hpp"

namespace obvhs {

obvhs_status set_hip_error(obvhs_context* ctx, hipError_t e, const char* where) {
    if (ctx) {
        ctx->last_error = std::string(where) + ": " + hipGetErrorString(e);
    }
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return OBVHS_ERR_OOM;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return OBVHS_ERR_NO_DEVICE;
    return OBVHS_ERR_DEVICE;
}

obvhs_status device_error_to_status(uint32_t err) {
    if (err & DERR_NAN_INPUT) return OBVHS_ERR_NAN_INPUT;
    if (err & DERR_SELF_MERGE) return OBVHS_ERR_PLOC_SELF_MERGE;
    if (err & DERR_CW_NONUNIT_LEAF) return OBVHS_ERR_CW_NONUNIT_LEAF;
    if (err & DERR_CW_DISTRIBUTE) return OBVHS_ERR_CW_DISTRIBUTE;
    if (err & DERR_CW_ORDER) return OBVHS_ERR_CW_ORDER;
    if (err & DERR_CW_PRIMCOUNT) return OBVHS_ERR_CW_PRIMCOUNT;
    if (err & DERR_CW_DEPTH) return OBVHS_ERR_DEPTH;
    if (err & DERR_CAPACITY) return OBVHS_ERR_CAPACITY;
    return OBVHS_ERR_DEVICE;
}

}  // namespace obvhs

using namespace obvhs;

#define CTX_TRY(expr, where)                                        \
    do {                                                            \
        hipError_t _e = (expr);                                     \
        if (_e != hipSuccess) return set_hip_error(ctx, _e, where); \
    } while (0)

extern "C" {

const char* obvhs_status_string(obvhs_status s) {
    switch (s) {
        case OBVHS_OK: return "ok";
        case OBVHS_ERR_INVALID_ARG: return "invalid argument";
        case OBVHS_ERR_NAN_INPUT: return "NaN in primitive bounds";
        case OBVHS_ERR_PLOC_SELF_MERGE: return "PLOC cluster without a finite-cost neighbour";
        case OBVHS_ERR_CW_DISTRIBUTE: return "CWBVH cost table without a finite distribution";
        case OBVHS_ERR_CW_NONUNIT_LEAF: return "CWBVH conversion expects one primitive per BVH2 leaf";
        case OBVHS_ERR_CW_ORDER: return "CWBVH child ordering left a child unassigned";
        case OBVHS_ERR_CW_PRIMCOUNT: return "CWBVH leaf with an invalid primitive count";
        case OBVHS_ERR_CAPACITY: return "internal capacity exceeded";
        case OBVHS_ERR_DEPTH: return "CWBVH too deep";
        case OBVHS_ERR_OOM: return "out of device memory";
        case OBVHS_ERR_DEVICE: return "HIP runtime error";
        case OBVHS_ERR_NO_DEVICE: return "no HIP device";
    }
    return "unknown status";
}

obvhs_status obvhs_context_create(int device, void* hip_stream, obvhs_context** out) {
    if (!out) return OBVHS_ERR_INVALID_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return OBVHS_ERR_NO_DEVICE;
    if (device < 0 || device >= count) return OBVHS_ERR_INVALID_ARG;
    obvhs_context* ctx = new (std::nothrow) obvhs_context();
    if (!ctx) return OBVHS_ERR_OOM;
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) {
        if (hip_stream) {
            ctx->stream = static_cast<hipStream_t>(hip_stream);
        } else {
            e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
            ctx->own_stream = e == hipSuccess;
        }
    }
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev_start);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev_stop);
    if (e == hipSuccess) e = hipHostMalloc(&ctx->pinned, 4096, hipHostMallocDefault);
    if (e == hipSuccess) e = upload_log2_table();
    if (e != hipSuccess) {
        obvhs_status s = set_hip_error(ctx, e, "context create");
        obvhs_context_destroy(ctx);
        return s;
    }
    const obvhs_status ns = noalloc_scratch_init(ctx);
    if (ns != OBVHS_OK) {
        obvhs_context_destroy(ctx);
        return ns;
    }
    *out = ctx;
    return OBVHS_OK;
}

obvhs_status obvhs_context_destroy(obvhs_context* ctx) {
    if (!ctx) return OBVHS_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    BuildScratch& S = ctx->scratch;
    S.leaves.release();
    S.clusters_a.release();
    S.clusters_b.release();
    S.keys.release();
    S.keys_hi.release();
    S.vals.release();
    S.merge.release();
    S.blocksums.release();
    S.sort_scratch.release();
    S.misc.release();
    S.cw_scratch.release();
    S.cw_misc.release();
    S.trav_spill.release();
    ctx->quality.release_all();
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return OBVHS_OK;
}

obvhs_status obvhs_context_synchronize(obvhs_context* ctx) {
    if (!ctx) return OBVHS_ERR_INVALID_ARG;
    CTX_TRY(hipStreamSynchronize(ctx->stream), "synchronize");
    return OBVHS_OK;
}

obvhs_status obvhs_context_set_trace(obvhs_context* ctx, int enable) {
    if (!ctx) return OBVHS_ERR_INVALID_ARG;
    ctx->keep_trace = enable != 0;
    return OBVHS_OK;
}

const char* obvhs_last_error_message(const obvhs_context* ctx) { return ctx ? ctx->last_error.c_str() : ""; }

obvhs_status obvhs_device_malloc(obvhs_context* ctx, uint64_t bytes, void** dptr) {
    if (!ctx || !dptr) return OBVHS_ERR_INVALID_ARG;
    *dptr = nullptr;
    CTX_TRY(hipMalloc(dptr, bytes ? bytes : 1), "device malloc");
    return OBVHS_OK;
}

obvhs_status obvhs_device_free(obvhs_context* ctx, void* dptr) {
    if (!ctx) return OBVHS_ERR_INVALID_ARG;
    if (dptr) CTX_TRY(hipFree(dptr), "device free");
    return OBVHS_OK;
}

obvhs_status obvhs_memcpy_htod(obvhs_context* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx) return OBVHS_ERR_INVALID_ARG;
    if (!bytes) return OBVHS_OK;
    CTX_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "memcpy htod");
    CTX_TRY(hipStreamSynchronize(ctx->stream), "memcpy htod sync");
    return OBVHS_OK;
}

obvhs_status obvhs_memcpy_dtoh(obvhs_context* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx) return OBVHS_ERR_INVALID_ARG;
    if (!bytes) return OBVHS_OK;
    CTX_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "memcpy dtoh");
    CTX_TRY(hipStreamSynchronize(ctx->stream), "memcpy dtoh sync");
    return OBVHS_OK;
}

obvhs_status obvhs_memset_device(obvhs_context* ctx, void* dst, int value, uint64_t bytes) {
    if (!ctx) return OBVHS_ERR_INVALID_ARG;
    if (!bytes) return OBVHS_OK;
    CTX_TRY(hipMemsetAsync(dst, value, bytes, ctx->stream), "memset");
    return OBVHS_OK;
}

obvhs_status obvhs_memcpy_dtod(obvhs_context* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return OBVHS_ERR_INVALID_ARG;
    if (!bytes) return OBVHS_OK;
    CTX_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream), "memcpy dtod");
    return OBVHS_OK;
}

obvhs_status obvhs_timer_start(obvhs_context* ctx) {
    if (!ctx) return OBVHS_ERR_INVALID_ARG;
    CTX_TRY(hipEventRecord(ctx->ev_start, ctx->stream), "timer start");
    return OBVHS_OK;
}

obvhs_status obvhs_timer_stop(obvhs_context* ctx, float* ms) {
    if (!ctx || !ms) return OBVHS_ERR_INVALID_ARG;
    CTX_TRY(hipEventRecord(ctx->ev_stop, ctx->stream), "timer stop");
    CTX_TRY(hipEventSynchronize(ctx->ev_stop), "timer sync");
    CTX_TRY(hipEventElapsedTime(ms, ctx->ev_start, ctx->ev_stop), "timer elapsed");
    return OBVHS_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// BVH2
// ---------------------------------------------------------------------------------------------------------------

static obvhs_status ploc_entry(obvhs_context* ctx, const obvhs_triangle* d_tris, const obvhs_aabb* d_aabbs,
                               const uint32_t* d_indices, uint32_t n, uint32_t search_distance,
                               uint32_t sort_precision, uint32_t search_depth_threshold, uint32_t fused_max_prims,
                               obvhs_bvh2* reuse, obvhs_bvh2** out) {
    if (!ctx || !out) return OBVHS_ERR_INVALID_ARG;
    if (n > 0 && !d_tris && !d_aabbs) return OBVHS_ERR_INVALID_ARG;
    if (n >= (1u << 31)) return OBVHS_ERR_INVALID_ARG;
    if (sort_precision != 64u && sort_precision != 128u) return OBVHS_ERR_INVALID_ARG;  // SortPrecision
    if (search_distance == 0u) search_distance = radius_for(ctx, n);  // "auto": the tuning table's distance
    CTX_TRY(hipSetDevice(ctx->device), "set device");
    obvhs_bvh2* bvh = reuse ? reuse : new (std::nothrow) obvhs_bvh2();
    if (!bvh) return OBVHS_ERR_OOM;
    obvhs_status s = ploc_build(ctx, d_tris, d_aabbs, d_indices, n, search_distance, sort_precision == 128u,
                                search_depth_threshold, fused_max_prims, bvh);
    if (s != OBVHS_OK) {
        if (!reuse) obvhs_bvh2_destroy(bvh);
        *out = nullptr;
        return s;
    }
    *out = bvh;
    return OBVHS_OK;
}

obvhs_status obvhs_ploc_build_aabbs(obvhs_context* ctx, const obvhs_aabb* d_aabbs, const uint32_t* d_indices,
                                    uint32_t n, uint32_t search_distance, uint32_t sort_precision,
                                    uint32_t search_depth_threshold, obvhs_bvh2* reuse, obvhs_bvh2** out) {
    return ploc_entry(ctx, nullptr, d_aabbs, d_indices, n, search_distance, sort_precision, search_depth_threshold,
                      0u, reuse, out);
}

obvhs_status obvhs_ploc_build_triangles(obvhs_context* ctx, const obvhs_triangle* d_tris, const uint32_t* d_indices,
                                        uint32_t n, uint32_t search_distance, uint32_t sort_precision,
                                        uint32_t search_depth_threshold, obvhs_bvh2* reuse, obvhs_bvh2** out) {
    return ploc_entry(ctx, d_tris, nullptr, d_indices, n, search_distance, sort_precision, search_depth_threshold,
                      0u, reuse, out);
}

obvhs_status obvhs_bvh2_info(const obvhs_bvh2* bvh, uint32_t* node_count, uint32_t* prim_count,
                             uint32_t* max_depth) {
    if (!bvh) return OBVHS_ERR_INVALID_ARG;
    if (node_count) *node_count = bvh->node_count;
    if (prim_count) *prim_count = bvh->prim_count;
    if (max_depth) *max_depth = bvh->max_depth;
    return OBVHS_OK;
}

obvhs_status obvhs_bvh2_copy_nodes_small(obvhs_context* ctx, const obvhs_bvh2* bvh, obvhs_bvh2_node_small* host_out) {
    if (!ctx || !bvh || (!host_out && bvh->node_count)) return OBVHS_ERR_INVALID_ARG;
    if (!bvh->node_count) return OBVHS_OK;
    return obvhs_memcpy_dtoh(ctx, host_out, bvh->nodes.ptr, (size_t)bvh->node_count * sizeof(Node32));
}

obvhs_status obvhs_bvh2_copy_nodes(obvhs_context* ctx, const obvhs_bvh2* bvh, obvhs_bvh2_node* host_out) {
    if (!ctx || !bvh || (!host_out && bvh->node_count)) return OBVHS_ERR_INVALID_ARG;
    if (!bvh->node_count) return OBVHS_OK;
    std::vector<Node32> tmp(bvh->node_count);
    obvhs_status s = obvhs_memcpy_dtoh(ctx, tmp.data(), bvh->nodes.ptr, tmp.size() * sizeof(Node32));
    if (s != OBVHS_OK) return s;
    for (size_t i = 0; i < tmp.size(); ++i) {
        const Node32& a = tmp[i];
        obvhs_bvh2_node& o = host_out[i];
        o.aabb.min = obvhs_vec3a{a.minx, a.miny, a.minz, 0.0f};
        o.aabb.max = obvhs_vec3a{a.maxx, a.maxy, a.maxz, 0.0f};
        o.prim_count = a.prim_count;
        o.first_index = a.first_index;
        o.meta1 = 0u;
        o.meta2 = 0u;
    }
    return OBVHS_OK;
}

obvhs_status obvhs_bvh2_copy_primitive_indices(obvhs_context* ctx, const obvhs_bvh2* bvh, uint32_t* host_out) {
    if (!ctx || !bvh) return OBVHS_ERR_INVALID_ARG;
    return obvhs_memcpy_dtoh(ctx, host_out, bvh->prim_indices.ptr, (size_t)bvh->prim_count * 4);
}

obvhs_status obvhs_bvh2_copy_sort_trace(obvhs_context* ctx, const obvhs_bvh2* bvh, uint64_t* host_codes,
                                        uint32_t* host_sorted_index) {
    if (!ctx || !bvh) return OBVHS_ERR_INVALID_ARG;
    if (!bvh->prim_count) return OBVHS_OK;
    if (!bvh->codes.ptr || !bvh->sorted_index.ptr) return OBVHS_ERR_INVALID_ARG;  // build without trace
    obvhs_status s = OBVHS_OK;
    const size_t n = bvh->prim_count;
    if (host_codes && bvh->code_words == 1u) s = obvhs_memcpy_dtoh(ctx, host_codes, bvh->codes.ptr, n * 8);
    if (host_codes && bvh->code_words == 2u) {  // device: low words, then high words; host: (low, high) per leaf
        std::vector<uint64_t> words(2 * n);
        s = obvhs_memcpy_dtoh(ctx, words.data(), bvh->codes.ptr, n * 16);
        for (size_t i = 0; s == OBVHS_OK && i < n; ++i) {
            host_codes[2 * i] = words[i];
            host_codes[2 * i + 1] = words[n + i];
        }
    }
    if (s == OBVHS_OK && host_sorted_index)
        s = obvhs_memcpy_dtoh(ctx, host_sorted_index, bvh->sorted_index.ptr, (size_t)bvh->prim_count * 4);
    return s;
}

obvhs_status obvhs_bvh2_destroy(obvhs_bvh2* bvh) {
    if (!bvh) return OBVHS_OK;
    bvh->nodes.release();
    bvh->prim_indices.release();
    bvh->codes.release();
    bvh->sorted_index.release();
    bvh->hist.release();
    bvh->dp.release();
    delete bvh;
    return OBVHS_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// CWBVH
// ---------------------------------------------------------------------------------------------------------------

obvhs_status obvhs_bvh2_to_cwbvh(obvhs_context* ctx, const obvhs_bvh2* bvh2, uint32_t max_prims_per_leaf,
                                 int order_children, int include_exact_node_aabbs, obvhs_cwbvh* reuse,
                                 obvhs_cwbvh** out) {
    if (!ctx || !bvh2 || !out) return OBVHS_ERR_INVALID_ARG;
    if (include_exact_node_aabbs) return OBVHS_ERR_INVALID_ARG;
    if (max_prims_per_leaf > 3) return OBVHS_ERR_INVALID_ARG;  // bvh2_to_cwbvh.rs:489 (0..=3)
    CTX_TRY(hipSetDevice(ctx->device), "set device");
    obvhs_cwbvh* cw = reuse ? reuse : new (std::nothrow) obvhs_cwbvh();
    if (!cw) return OBVHS_ERR_OOM;
    // The handle's cached cost table is internal state: converting does not change the BVH2 itself.
    obvhs_status s = cwbvh_convert(ctx, const_cast<obvhs_bvh2*>(bvh2), max_prims_per_leaf, order_children != 0, cw);
    if (s != OBVHS_OK) {
        if (!reuse) obvhs_cwbvh_destroy(cw);
        *out = nullptr;
        return s;
    }
    *out = cw;
    return OBVHS_OK;
}

obvhs_status obvhs_build_cwbvh_from_tris(obvhs_context* ctx, const obvhs_triangle* d_tris, uint32_t n,
                                         const obvhs_build_params* params, obvhs_cwbvh* reuse, obvhs_cwbvh** out,
                                         obvhs_build_stats* stats) {
    return obvhs_build_cwbvh_from_tris_ex(ctx, d_tris, n, params, nullptr, reuse, out, stats);
}

obvhs_status obvhs_build_cwbvh_from_tris_ex(obvhs_context* ctx, const obvhs_triangle* d_tris, uint32_t n,
                                            const obvhs_build_params* params, const obvhs_build_ext* ext,
                                            obvhs_cwbvh* reuse, obvhs_cwbvh** out, obvhs_build_stats* stats) {
    if (!ctx || !params || !out) return OBVHS_ERR_INVALID_ARG;
    if (n > 0 && !d_tris) return OBVHS_ERR_INVALID_ARG;
    if (ext && ext->ratio_sequence_len > 0 && !ext->ratio_sequence) return OBVHS_ERR_INVALID_ARG;
    CTX_TRY(hipSetDevice(ctx->device), "set device");
    obvhs_cwbvh* cw = reuse ? reuse : new (std::nothrow) obvhs_cwbvh();
    if (!cw) return OBVHS_ERR_OOM;
    obvhs_bvh2* b2 = cw->owned_bvh2;
    uint32_t mp = params->max_prims_per_leaf;
    mp = mp < 1u ? 1u : (mp > 3u ? 3u : mp);  // cwbvh/builder.rs:74 clamp(1, 3)
    const bool reinsert = params->reinsertion_batch_ratio > 0.0f;
    // Without reinsertion the cost table of the conversion is evaluated by the PLOC merges themselves.
    const uint32_t fused = reinsert ? 0u : mp;
    obvhs_status s;
    if (params->pre_split && n > 0) {  // cwbvh/builder.rs:26-55
        uint32_t n_split = 0;
        s = presplit_leaves(ctx, d_tris, n, ext ? ext->split_max_total : 0ull, n_split, nullptr, nullptr);
        if (s == OBVHS_OK)
            s = ploc_entry(ctx, nullptr, ctx->quality.split_aabbs.as<obvhs_aabb>(),
                           ctx->quality.split_indices.as<uint32_t>(), n_split, params->ploc_search_distance,
                           params->sort_precision, params->search_depth_threshold, fused, b2, &b2);
    } else {
        s = ploc_entry(ctx, d_tris, nullptr, nullptr, n, params->ploc_search_distance, params->sort_precision,
                       params->search_depth_threshold, fused, b2, &b2);
    }
    if (s == OBVHS_OK) cw->owned_bvh2 = b2;
    if (s == OBVHS_OK && reinsert)  // cwbvh/builder.rs:73
        s = bvh2_reinsertion(ctx, b2, params->reinsertion_batch_ratio, ext ? ext->ratio_sequence : nullptr,
                             ext ? ext->ratio_sequence_len : 0u, mp);
    if (s == OBVHS_OK) s = cwbvh_convert(ctx, b2, mp, true, cw);
    if (s != OBVHS_OK) {
        if (!reuse) obvhs_cwbvh_destroy(cw);
        *out = nullptr;
        return s;
    }
    if (stats) {
        stats->prim_count = b2 ? b2->prim_count : n;
        stats->ploc_iterations = b2 ? b2->ploc_iterations : 0;
        stats->bvh2_node_count = b2 ? b2->node_count : 0;
        stats->cwbvh_node_count = cw->node_count;
        stats->ploc_cluster_sum = b2 ? b2->cluster_sum : 0;
    }
    *out = cw;
    return OBVHS_OK;
}

obvhs_status obvhs_cwbvh_bvh2(const obvhs_cwbvh* bvh, obvhs_bvh2** out) {
    if (!bvh || !out) return OBVHS_ERR_INVALID_ARG;
    *out = bvh->owned_bvh2;
    return OBVHS_OK;
}

obvhs_status obvhs_bvh2_reinsertion(obvhs_context* ctx, obvhs_bvh2* bvh, float batch_size_ratio,
                                    const float* ratio_sequence, uint32_t ratio_sequence_len) {
    if (!ctx || !bvh || (ratio_sequence_len > 0 && !ratio_sequence)) return OBVHS_ERR_INVALID_ARG;
    CTX_TRY(hipSetDevice(ctx->device), "set device");
    return bvh2_reinsertion(ctx, bvh, batch_size_ratio, ratio_sequence, ratio_sequence_len);
}

obvhs_status obvhs_bvh2_refit_all(obvhs_context* ctx, obvhs_bvh2* bvh) {
    if (!ctx || !bvh) return OBVHS_ERR_INVALID_ARG;
    CTX_TRY(hipSetDevice(ctx->device), "set device");
    return bvh2_refit_all(ctx, bvh);
}

obvhs_status obvhs_bvh2_refit_from_triangles(obvhs_context* ctx, obvhs_bvh2* bvh, const obvhs_triangle* d_tris) {
    if (!ctx || !bvh || (bvh->prim_count > 0 && !d_tris)) return OBVHS_ERR_INVALID_ARG;
    CTX_TRY(hipSetDevice(ctx->device), "set device");
    return bvh2_refit_from_triangles(ctx, bvh, d_tris);
}

obvhs_status obvhs_bvh2_create_from_host(obvhs_context* ctx, const obvhs_bvh2_node* nodes, uint32_t node_count,
                                         const uint32_t* prim_indices, uint32_t prim_count, obvhs_bvh2** out) {
    if (!ctx || !out || (node_count && !nodes) || (prim_count && !prim_indices)) return OBVHS_ERR_INVALID_ARG;
    std::vector<Node32> tmp(node_count);
    for (uint32_t i = 0; i < node_count; ++i) {
        const obvhs_bvh2_node& a = nodes[i];
        if (a.prim_count == 0u && (uint64_t)a.first_index + 1u >= node_count) return OBVHS_ERR_INVALID_ARG;
        if (a.prim_count != 0u && (uint64_t)a.first_index + (a.prim_count & 0x7fffffffu) > prim_count)
            return OBVHS_ERR_INVALID_ARG;
        Node32& o = tmp[i];
        o.minx = a.aabb.min.x;
        o.miny = a.aabb.min.y;
        o.minz = a.aabb.min.z;
        o.maxx = a.aabb.max.x;
        o.maxy = a.aabb.max.y;
        o.maxz = a.aabb.max.z;
        o.prim_count = a.prim_count;
        o.first_index = a.first_index;
    }
    CTX_TRY(hipSetDevice(ctx->device), "set device");
    obvhs_bvh2* b = new (std::nothrow) obvhs_bvh2();
    if (!b) return OBVHS_ERR_OOM;
    b->node_count = node_count;
    b->prim_count = prim_count;
    b->pass_ordered = false;  // node order unknown: bottom-up sweeps use level lists
    hipError_t e = b->nodes.ensure((size_t)node_count * sizeof(Node32) + 32);
    if (e == hipSuccess) e = b->prim_indices.ensure((size_t)prim_count * 4 + 4);
    if (e == hipSuccess && node_count)
        e = hipMemcpy(b->nodes.ptr, tmp.data(), tmp.size() * sizeof(Node32), hipMemcpyHostToDevice);
    if (e == hipSuccess && prim_count)
        e = hipMemcpy(b->prim_indices.ptr, prim_indices, (size_t)prim_count * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        obvhs_bvh2_destroy(b);
        return set_hip_error(ctx, e, "create bvh2 from host");
    }
    *out = b;
    return OBVHS_OK;
}

obvhs_status obvhs_bvh2_set_node_aabbs(obvhs_context* ctx, obvhs_bvh2* bvh, const uint32_t* d_node_ids,
                                       const obvhs_aabb* d_aabbs, uint32_t count) {
    if (!ctx || !bvh || (count && (!d_node_ids || !d_aabbs))) return OBVHS_ERR_INVALID_ARG;
    CTX_TRY(hipSetDevice(ctx->device), "set device");
    return bvh2_set_node_aabbs(ctx, bvh, d_node_ids, d_aabbs, count);
}

obvhs_status obvhs_bvh2_primitives_to_nodes(obvhs_context* ctx, const obvhs_bvh2* bvh, uint32_t* d_out) {
    if (!ctx || !bvh || (bvh->prim_count && !d_out)) return OBVHS_ERR_INVALID_ARG;
    CTX_TRY(hipSetDevice(ctx->device), "set device");
    return bvh2_primitives_to_nodes(ctx, bvh, d_out);
}

obvhs_status obvhs_bvh2_rebuild_path_flags(obvhs_context* ctx, obvhs_bvh2* bvh, const uint32_t* d_leaf_ids,
                                           uint32_t count, uint8_t* d_flags) {
    if (!ctx || !bvh || (count && !d_leaf_ids) || (bvh->node_count && !d_flags)) return OBVHS_ERR_INVALID_ARG;
    CTX_TRY(hipSetDevice(ctx->device), "set device");
    return bvh2_rebuild_path_flags(ctx, bvh, d_leaf_ids, count, d_flags);
}

obvhs_status obvhs_bvh2_partial_rebuild(obvhs_context* ctx, obvhs_bvh2* bvh, const uint8_t* d_flags,
                                        uint32_t search_distance, uint32_t sort_precision,
                                        uint32_t search_depth_threshold) {
    if (!ctx || !bvh || (bvh->node_count && !d_flags)) return OBVHS_ERR_INVALID_ARG;
    if (sort_precision != 64u && sort_precision != 128u) return OBVHS_ERR_INVALID_ARG;
    if (search_distance == 0u) search_distance = radius_for(ctx, bvh->prim_count);
    CTX_TRY(hipSetDevice(ctx->device), "set device");
    return bvh2_partial_rebuild(ctx, bvh, d_flags, search_distance, sort_precision == 128u, search_depth_threshold);
}

obvhs_status obvhs_bvh2_full_rebuild(obvhs_context* ctx, obvhs_bvh2* bvh, uint32_t search_distance,
                                     uint32_t sort_precision, uint32_t search_depth_threshold) {
    if (!ctx || !bvh) return OBVHS_ERR_INVALID_ARG;
    if (sort_precision != 64u && sort_precision != 128u) return OBVHS_ERR_INVALID_ARG;
    if (search_distance == 0u) search_distance = radius_for(ctx, bvh->prim_count);
    CTX_TRY(hipSetDevice(ctx->device), "set device");
    return bvh2_full_rebuild(ctx, bvh, search_distance, sort_precision == 128u, search_depth_threshold);
}

obvhs_status obvhs_split_aabbs_preset(obvhs_context* ctx, const obvhs_triangle* d_tris, uint32_t n,
                                      uint64_t max_total, obvhs_aabb* d_out_aabbs, uint32_t* d_out_indices,
                                      uint32_t out_capacity, uint32_t* out_count, float* out_avg_half_area,
                                      float* out_largest_half_area) {
    if (!ctx || !out_count || (n > 0 && !d_tris)) return OBVHS_ERR_INVALID_ARG;
    CTX_TRY(hipSetDevice(ctx->device), "set device");
    uint32_t m = 0;
    float avg = 0.0f, largest = 0.0f;
    obvhs_status s = presplit_leaves(ctx, d_tris, n, max_total, m, &avg, &largest);
    if (s != OBVHS_OK) return s;
    *out_count = m;
    if (out_avg_half_area) *out_avg_half_area = avg;
    if (out_largest_half_area) *out_largest_half_area = largest;
    if (m > out_capacity) return OBVHS_ERR_CAPACITY;
    if (m == 0) return OBVHS_OK;
    if (!d_out_aabbs || !d_out_indices) return OBVHS_ERR_INVALID_ARG;
    CTX_TRY(hipMemcpyAsync(d_out_aabbs, ctx->quality.split_aabbs.ptr, (size_t)m * 32, hipMemcpyDeviceToDevice,
                           ctx->stream),
            "copy split boxes");
    CTX_TRY(hipMemcpyAsync(d_out_indices, ctx->quality.split_indices.ptr, (size_t)m * 4, hipMemcpyDeviceToDevice,
                           ctx->stream),
            "copy split indices");
    return OBVHS_OK;
}

obvhs_status obvhs_cwbvh_info(const obvhs_cwbvh* bvh, uint32_t* node_count, uint32_t* prim_count,
                              obvhs_aabb* total_aabb) {
    if (!bvh) return OBVHS_ERR_INVALID_ARG;
    if (node_count) *node_count = bvh->node_count;
    if (prim_count) *prim_count = bvh->prim_count;
    if (total_aabb) *total_aabb = bvh->total_aabb;
    return OBVHS_OK;
}

obvhs_status obvhs_cwbvh_copy_nodes(obvhs_context* ctx, const obvhs_cwbvh* bvh, obvhs_cwbvh_node* host_out) {
    if (!ctx || !bvh) return OBVHS_ERR_INVALID_ARG;
    return obvhs_memcpy_dtoh(ctx, host_out, bvh->nodes.ptr, (size_t)bvh->node_count * sizeof(CwNode));
}

obvhs_status obvhs_cwbvh_copy_primitive_indices(obvhs_context* ctx, const obvhs_cwbvh* bvh, uint32_t* host_out) {
    if (!ctx || !bvh) return OBVHS_ERR_INVALID_ARG;
    return obvhs_memcpy_dtoh(ctx, host_out, bvh->prim_indices.ptr, (size_t)bvh->prim_count * 4);
}

obvhs_status obvhs_cwbvh_device_arrays(const obvhs_cwbvh* bvh, const obvhs_cwbvh_node** d_nodes,
                                       const uint32_t** d_primitive_indices) {
    if (!bvh) return OBVHS_ERR_INVALID_ARG;
    if (d_nodes) *d_nodes = static_cast<const obvhs_cwbvh_node*>(bvh->nodes.ptr);
    if (d_primitive_indices) *d_primitive_indices = static_cast<const uint32_t*>(bvh->prim_indices.ptr);
    return OBVHS_OK;
}

obvhs_status obvhs_cwbvh_reorder_triangles(obvhs_context* ctx, const obvhs_cwbvh* bvh, const obvhs_triangle* d_tris,
                                           obvhs_triangle* d_bvh_tris) {
    if (!ctx || !bvh || (bvh->prim_count && (!d_tris || !d_bvh_tris))) return OBVHS_ERR_INVALID_ARG;
    CTX_TRY(launch_reorder_triangles(d_tris, bvh->prim_indices.as<uint32_t>(), bvh->prim_count, d_bvh_tris,
                                     ctx->stream),
            "reorder triangles");
    return OBVHS_OK;
}

obvhs_status obvhs_cwbvh_destroy(obvhs_cwbvh* bvh) {
    if (!bvh) return OBVHS_OK;
    bvh->nodes.release();
    bvh->prim_indices.release();
    if (bvh->owned_bvh2) obvhs_bvh2_destroy(bvh->owned_bvh2);
    delete bvh;
    return OBVHS_OK;
}

// ---------------------------------------------------------------------------------------------------------------
// Traversal
// ---------------------------------------------------------------------------------------------------------------

obvhs_status obvhs_cwbvh_ray_traverse(obvhs_context* ctx, const obvhs_cwbvh* bvh, const obvhs_triangle* d_bvh_tris,
                                      const obvhs_ray* d_rays, uint32_t n, obvhs_ray_hit* d_hits) {
    if (!ctx || !bvh || (n && (!d_rays || !d_hits))) return OBVHS_ERR_INVALID_ARG;
    if (bvh->node_count && !d_bvh_tris) return OBVHS_ERR_INVALID_ARG;
    TravScratch ts;
    if (n) {
        const obvhs_status ss = traverse_scratch(ctx, ts, n);
        if (ss != OBVHS_OK) return ss;
    }
    ts.lds_depth = lds_depth_for(ctx, bvh->prim_count);
    ts.chunked = chunk_for(ctx, bvh->prim_count);
    CTX_TRY(launch_traverse(bvh->nodes.as<CwNode>(), bvh->node_count, bvh->prim_count, d_bvh_tris, d_rays, n, 0,
                            d_hits, nullptr, nullptr, nullptr, ts, ctx->stream),
            "traverse");
    return OBVHS_OK;
}

obvhs_status obvhs_cwbvh_ray_traverse_miss(obvhs_context* ctx, const obvhs_cwbvh* bvh,
                                           const obvhs_triangle* d_bvh_tris, const obvhs_ray* d_rays, uint32_t n,
                                           uint8_t* d_miss) {
    if (!ctx || !bvh || (n && (!d_rays || !d_miss))) return OBVHS_ERR_INVALID_ARG;
    if (bvh->node_count && !d_bvh_tris) return OBVHS_ERR_INVALID_ARG;
    TravScratch ts;
    if (n) {
        const obvhs_status ss = traverse_scratch(ctx, ts, n);
        if (ss != OBVHS_OK) return ss;
    }
    ts.lds_depth = lds_depth_for(ctx, bvh->prim_count);
    ts.chunked = chunk_for(ctx, bvh->prim_count);
    CTX_TRY(launch_traverse(bvh->nodes.as<CwNode>(), bvh->node_count, bvh->prim_count, d_bvh_tris, d_rays, n, 1,
                            nullptr, d_miss, nullptr, nullptr, ts, ctx->stream),
            "traverse miss");
    return OBVHS_OK;
}

obvhs_status obvhs_cwbvh_ray_traverse_counted(obvhs_context* ctx, const obvhs_cwbvh* bvh,
                                              const obvhs_triangle* d_bvh_tris, const obvhs_ray* d_rays, uint32_t n,
                                              obvhs_ray_hit* d_hits, uint32_t* d_node_visits,
                                              uint32_t* d_tri_tests) {
    if (!ctx || !bvh || (n && (!d_rays || !d_hits || !d_node_visits || !d_tri_tests))) return OBVHS_ERR_INVALID_ARG;
    if (bvh->node_count && !d_bvh_tris) return OBVHS_ERR_INVALID_ARG;
    TravScratch ts;
    if (n) {
        const obvhs_status ss = traverse_scratch(ctx, ts, n);
        if (ss != OBVHS_OK) return ss;
    }
    ts.lds_depth = lds_depth_for(ctx, bvh->prim_count);
    ts.chunked = chunk_for(ctx, bvh->prim_count);
    CTX_TRY(launch_traverse(bvh->nodes.as<CwNode>(), bvh->node_count, bvh->prim_count, d_bvh_tris, d_rays, n, 0,
                            d_hits, nullptr, d_node_visits, d_tri_tests, ts, ctx->stream),
            "traverse counted");
    return OBVHS_OK;
}

// The watertight primitive test in place of Triangle::intersect (traverse.hip: tri_intersect_wt), closest hit
// (mode 2) or occlusion (mode 3); the traversal itself is ray_traverse / ray_traverse_miss's.
static obvhs_status traverse_watertight(obvhs_context* ctx, const obvhs_cwbvh* bvh, const obvhs_triangle* d_bvh_tris,
                                        const obvhs_ray* d_rays, uint32_t n, int mode, obvhs_ray_hit* d_hits,
                                        uint8_t* d_miss) {
    if (!ctx || !bvh || (n && (!d_rays || (mode == 2 ? !d_hits : !d_miss)))) return OBVHS_ERR_INVALID_ARG;
    if (bvh->node_count && !d_bvh_tris) return OBVHS_ERR_INVALID_ARG;
    TravScratch ts;
    if (n) {
        const obvhs_status ss = traverse_scratch(ctx, ts, n);
        if (ss != OBVHS_OK) return ss;
    }
    ts.lds_depth = lds_depth_for(ctx, bvh->prim_count);
    ts.chunked = chunk_for(ctx, bvh->prim_count);
    CTX_TRY(launch_traverse(bvh->nodes.as<CwNode>(), bvh->node_count, bvh->prim_count, d_bvh_tris, d_rays, n, mode,
                            d_hits, d_miss, nullptr, nullptr, ts, ctx->stream),
            "traverse watertight");
    return OBVHS_OK;
}

obvhs_status obvhs_cwbvh_ray_traverse_watertight(obvhs_context* ctx, const obvhs_cwbvh* bvh,
                                                 const obvhs_triangle* d_bvh_tris, const obvhs_ray* d_rays,
                                                 uint32_t n, obvhs_ray_hit* d_hits) {
    return traverse_watertight(ctx, bvh, d_bvh_tris, d_rays, n, 2, d_hits, nullptr);
}

obvhs_status obvhs_cwbvh_ray_traverse_miss_watertight(obvhs_context* ctx, const obvhs_cwbvh* bvh,
                                                      const obvhs_triangle* d_bvh_tris, const obvhs_ray* d_rays,
                                                      uint32_t n, uint8_t* d_miss) {
    return traverse_watertight(ctx, bvh, d_bvh_tris, d_rays, n, 3, nullptr, d_miss);
}

obvhs_status obvhs_rays_new(obvhs_context* ctx, const obvhs_vec3a* d_origins, const obvhs_vec3a* d_directions,
                            uint32_t n, float tmin, float tmax, obvhs_ray* d_out) {
    if (!ctx || (n && (!d_origins || !d_directions || !d_out))) return OBVHS_ERR_INVALID_ARG;
    CTX_TRY(launch_rays_new(d_origins, d_directions, n, tmin, tmax, d_out, ctx->stream), "rays new");
    return OBVHS_OK;
}

obvhs_status obvhs_ray_hits_none(obvhs_context* ctx, obvhs_ray_hit* d_hits, uint32_t n) {
    if (!ctx || (n && !d_hits)) return OBVHS_ERR_INVALID_ARG;
    CTX_TRY(launch_hits_none(d_hits, n, ctx->stream), "hits none");
    return OBVHS_OK;
}

obvhs_status obvhs_rt_triangles_from_triangles(obvhs_context* ctx, const obvhs_triangle* d_tris, uint32_t n,
                                               obvhs_rt_triangle* d_out) {
    if (!ctx || (n && (!d_tris || !d_out))) return OBVHS_ERR_INVALID_ARG;
    CTX_TRY(launch_rt_triangles(d_tris, n, d_out, ctx->stream), "rt triangles");
    return OBVHS_OK;
}

// ray_traverse / ray_traverse_miss with RtTriangle primitives (mode 4: closest hit, 5: occlusion).
static obvhs_status traverse_rt(obvhs_context* ctx, const obvhs_cwbvh* bvh, const obvhs_rt_triangle* d_bvh_rt_tris,
                                const obvhs_ray* d_rays, uint32_t n, int mode, obvhs_ray_hit* d_hits,
                                uint8_t* d_miss) {
    if (!ctx || !bvh || (n && (!d_rays || (mode == 4 ? !d_hits : !d_miss)))) return OBVHS_ERR_INVALID_ARG;
    if (bvh->node_count && !d_bvh_rt_tris) return OBVHS_ERR_INVALID_ARG;
    TravScratch ts;
    if (n) {
        const obvhs_status ss = traverse_scratch(ctx, ts, n);
        if (ss != OBVHS_OK) return ss;
    }
    ts.lds_depth = lds_depth_for(ctx, bvh->prim_count);
    ts.chunked = chunk_for(ctx, bvh->prim_count);
    CTX_TRY(launch_traverse(bvh->nodes.as<CwNode>(), bvh->node_count, bvh->prim_count,
                            reinterpret_cast<const obvhs_triangle*>(d_bvh_rt_tris), d_rays, n, mode, d_hits, d_miss,
                            nullptr, nullptr, ts, ctx->stream),
            "traverse rt");
    return OBVHS_OK;
}

obvhs_status obvhs_cwbvh_ray_traverse_rt(obvhs_context* ctx, const obvhs_cwbvh* bvh,
                                         const obvhs_rt_triangle* d_bvh_rt_tris, const obvhs_ray* d_rays, uint32_t n,
                                         obvhs_ray_hit* d_hits) {
    return traverse_rt(ctx, bvh, d_bvh_rt_tris, d_rays, n, 4, d_hits, nullptr);
}

obvhs_status obvhs_cwbvh_ray_traverse_miss_rt(obvhs_context* ctx, const obvhs_cwbvh* bvh,
                                              const obvhs_rt_triangle* d_bvh_rt_tris, const obvhs_ray* d_rays,
                                              uint32_t n, uint8_t* d_miss) {
    return traverse_rt(ctx, bvh, d_bvh_rt_tris, d_rays, n, 5, nullptr, d_miss);
}

// The CWBVH exponent threshold table (host copy; built from the host libm on first use, no device needed).
obvhs_status obvhs_log2_threshold_table(uint32_t* out256) {
    if (!out256) return OBVHS_ERR_INVALID_ARG;
    memcpy(out256, host_log2_table(), 256 * sizeof(uint32_t));
    return OBVHS_OK;
}

}  // extern "C"
