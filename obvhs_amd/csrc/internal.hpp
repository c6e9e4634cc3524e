// Host-side declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <initializer_list>
#include <string>
#include <vector>

#include "../../include/obvhs_hip.h"

namespace obvhs {

struct Node32;
struct CwNode;

// Grow-only device allocation owned by a context or a handle.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t want);
    void release();
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

// PLOC per-pass state, double buffered on the device (pass k reads slot k&1, writes slot (k+1)&1).
struct PlocState {
    uint32_t count;
    uint32_t insert;
    uint32_t depth;
    uint32_t pad;
    unsigned long long cluster_sum;
};

struct BuildScratch {
    DevBuf leaves;      // Node32[n] leaf nodes in input order
    DevBuf clusters_a;  // Node32[n]
    DevBuf clusters_b;  // Node32[n]
    DevBuf keys;        // u64[n]
    DevBuf keys_hi;     // u64[n] high words of SortPrecision::U128 codes
    DevBuf vals;        // u32[n]
    DevBuf merge;       // i8[n]
    DevBuf blocksums;   // uint2[n/256+1]
    DevBuf sort_scratch;
    DevBuf misc;        // bounds keys, states, error word
    DevBuf cw_scratch;  // conversion working memory: item lists
    DevBuf cw_misc;     // conversion counters, error word and root item (kept across the item lists' growth)
    DevBuf trav_spill;  // traversal stack spill columns
};

constexpr uint32_t MAX_TREE_LEVELS = 8192;

// Working memory of the tree utilities and quality passes (quality.hip).
struct QualityScratch {
    DevBuf level_list, level_meta;                                   // levels
    DevBuf parents, owner, touched, misc;                            // reinsertion state
    DevBuf keys, vals, rkeys, rvals, rfrom, rto, conf, state, stack;  // candidates / reinsertions
    DevBuf sort_scratch;
    DevBuf split_aabbs, split_indices, ha, cand_a, cand_b, halves, flags, sums;  // pre-splits
    DevBuf marks, kept, freed;                                                   // rebuilds
    void release_all() {
        for (DevBuf* d : {&level_list, &level_meta, &parents, &owner, &touched, &misc, &keys, &vals, &rkeys, &rvals,
                          &rfrom, &rto, &conf, &state, &stack, &sort_scratch, &split_aabbs, &split_indices, &ha,
                          &cand_a, &cand_b, &halves, &flags, &sums, &marks, &kept, &freed})
            d->release();
    }
};

}  // namespace obvhs

struct obvhs_context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    obvhs::BuildScratch scratch;
    obvhs::QualityScratch quality;
    void* pinned = nullptr;  // small pinned host staging area
    bool keep_trace = false; // keep Morton codes and the sort permutation of each build (parity tests)
    uint32_t trav_lanes = 0;  // persistent traversal grid size (0 = not yet queried)
    void* noalloc_spill = nullptr;  // this device's address of the static traversal scratch (traverse.hip)
    obvhs_tune_table tune{{14u, 14u, 14u, 14u}, {8u, 8u, 8u, 8u}, {8u, 8u, 8u, 8u}, {0u, 0u, 0u, 0u}};
    obvhs::DevBuf tune_scratch;  // working memory of the tuning runs
    std::string last_error;
};

struct obvhs_bvh2 {
    obvhs::DevBuf nodes;           // Node32[2n-1]
    obvhs::DevBuf prim_indices;    // u32[n]
    obvhs::DevBuf codes;           // u64[n * code_words] Morton codes (original leaf order), for parity tests
    uint32_t code_words = 1;       // 2 after a SortPrecision::U128 build: codes = low words, then high words
    obvhs::DevBuf sorted_index;    // u32[n]
    uint32_t node_count = 0;
    uint32_t prim_count = 0;
    uint32_t max_depth = 96;
    uint32_t ploc_iterations = 0;
    uint64_t cluster_sum = 0;
    // Insert index after each PLOC pass: pass p emitted the child slots [hist[p+1], hist[p]).
    obvhs::DevBuf hist;
    // CWBVH cost table keyed by first_index >> 1 (cost[7], meta[7], nprims, cwsize per internal node), valid for
    // dp_max_prims (0 = not evaluated).
    obvhs::DevBuf dp;
    uint32_t dp_max_prims = 0;
    // Children are stored in PLOC pass order (the pass ranges in hist order the nodes bottom-up); false after
    // reinsertion, when bottom-up sweeps use the level lists instead.
    bool pass_ordered = true;
};

struct obvhs_cwbvh {
    obvhs::DevBuf nodes;          // CwNode[node_count]
    obvhs::DevBuf prim_indices;   // u32[prim_count]
    uint32_t node_count = 0;
    uint32_t prim_count = 0;
    obvhs_aabb total_aabb{};
    obvhs_bvh2* owned_bvh2 = nullptr;  // intermediate BVH2 kept for reuse by build_cwbvh_from_tris
};

namespace obvhs {

hipError_t launch_iota(uint32_t* p, uint32_t n, hipStream_t stream);

// sort.hip
size_t radix_sort_scratch_bytes(uint32_t n);
hipError_t radix_sort_pairs(uint64_t* keys, uint32_t* values, bool values_identity, uint32_t n, int bits,
                            void* scratch, hipStream_t stream, int items);
// Size class of n primitives or keys (the tuning table's index).
inline uint32_t size_class(uint64_t n) { return n < 4096u ? 0u : n < 65536u ? 1u : n < 1048576u ? 2u : 3u; }
inline int sort_items_for(const obvhs_context* ctx, uint64_t n) { return (int)ctx->tune.sort_items[size_class(n)]; }
inline uint32_t lds_depth_for(const obvhs_context* ctx, uint64_t n) { return ctx->tune.trav_lds_depth[size_class(n)]; }
inline bool chunk_for(const obvhs_context* ctx, uint64_t n) { return ctx->tune.trav_chunk[size_class(n)] != 0u; }
inline uint32_t radius_for(const obvhs_context* ctx, uint64_t n) { return ctx->tune.ploc_radius[size_class(n)]; }

// ploc.hip
obvhs_status ploc_rebuild(obvhs_context* ctx, obvhs_bvh2* bvh, const uint32_t* d_src_slots, uint32_t k,
                          const uint32_t* d_slot_map, uint32_t search_distance, bool u128,
                          uint32_t search_depth_threshold);
obvhs_status ploc_build(obvhs_context* ctx, const obvhs_triangle* d_tris, const obvhs_aabb* d_aabbs,
                        const uint32_t* d_indices, uint32_t n, uint32_t search_distance, bool u128,
                        uint32_t search_depth_threshold, uint32_t fused_max_prims, obvhs_bvh2* bvh);
struct DpView;
hipError_t bvh_dp_alloc(obvhs_bvh2* bvh);
DpView bvh_dp_view(const obvhs_bvh2* bvh, uint32_t max_prims);

// cwbvh.hip
obvhs_status cwbvh_convert(obvhs_context* ctx, obvhs_bvh2* bvh2, uint32_t max_prims_per_leaf,
                           bool order_children, obvhs_cwbvh* out);
struct MeshDesc;
obvhs_status cwbvh_convert_batch(obvhs_context* ctx, const Node32* nodes, const uint32_t* bvh2_prims,
                                 const DpView& dp, MeshDesc* d_meshes, std::vector<MeshDesc>& meshes,
                                 uint32_t n_total, bool order_children, DevBuf& out_nodes, DevBuf& out_prims,
                                 uint32_t& total_cw);
hipError_t launch_reorder_triangles_meshes(const obvhs_triangle* tris, const uint32_t* prim_indices,
                                           const uint32_t* offsets, uint32_t m, uint32_t n, obvhs_triangle* out,
                                           hipStream_t stream);
const uint32_t* host_log2_table();
hipError_t upload_log2_table();
hipError_t launch_reorder_triangles(const obvhs_triangle* tris, const uint32_t* prim_indices, uint32_t n,
                                    obvhs_triangle* out, hipStream_t stream);

// traverse.hip
// Persistent traversal grid: `lanes` lanes, each with a spill column for the deep part of its group stack.
struct TravScratch {
    void* spill = nullptr;
    uint32_t lanes = 0;
    uint32_t resident = 0;   // lanes the device keeps resident (compute units x 32 waves x 64)
    uint32_t lds_depth = 8;  // stack entries in LDS for the launch (4, 8 or 16); results do not depend on it
    bool chunked = false;    // wave-owned ray rows with lanes refilled (k_traverse_chunk); results do not depend on it
};
// Spill columns (and the lane count) for a traversal launch of n rays.
obvhs_status traverse_scratch(obvhs_context* ctx, TravScratch& ts, uint64_t n);
obvhs_status noalloc_scratch_init(obvhs_context* ctx);
hipError_t launch_rt_triangles(const obvhs_triangle* tris, uint32_t n, obvhs_rt_triangle* out, hipStream_t stream);
hipError_t launch_rays_new(const obvhs_vec3a* origins, const obvhs_vec3a* directions, uint32_t n, float tmin,
                           float tmax, obvhs_ray* out, hipStream_t stream);
hipError_t launch_hits_none(obvhs_ray_hit* hits, uint32_t n, hipStream_t stream);
hipError_t launch_traverse(const CwNode* nodes, uint32_t node_count, uint32_t prim_count, const obvhs_triangle* tris,
                           const obvhs_ray* rays, uint32_t n, int mode, obvhs_ray_hit* hits, uint8_t* miss,
                           uint32_t* node_visits, uint32_t* tri_tests, const TravScratch& ts, hipStream_t stream);
// Batched traversal: ray i is traced against mesh ray_mesh[i] of the MeshDesc table (ids >= mesh_count miss).
hipError_t launch_traverse_meshes(const CwNode* nodes, const MeshDesc* meshes, uint32_t mesh_count,
                                  const uint32_t* ray_mesh, const obvhs_triangle* tris, const obvhs_ray* rays,
                                  uint32_t n, int mode, obvhs_ray_hit* hits, uint8_t* miss, const TravScratch& ts,
                                  hipStream_t stream);

// CwBvh::ray_traverse_tlas_blas with the meshes of a batched build as BLASes.
hipError_t launch_traverse_tlas(const CwNode* tlas, uint32_t tlas_count, const uint32_t* tlas_prims,
                                uint32_t tlas_prim_count, const CwNode* blas_nodes, const MeshDesc* meshes,
                                uint32_t mesh_count, const obvhs_triangle* tris, const obvhs_ray* rays, uint32_t n,
                                obvhs_ray_hit* hits, const TravScratch& ts, hipStream_t stream);

// One request's share of a coalesced traversal launch: rays [begin, begin + count) of the launch are rays[0..count)
// of this request, traced against one CWBVH, results to out (RayHit[] in closest mode, u8[] in miss mode).
struct TraceSegment {
    const void* nodes;
    const void* tris;
    const void* rays;
    void* out;
    uint32_t node_count;
    uint32_t prim_count;
    uint32_t begin;
    uint32_t count;
};
hipError_t launch_traverse_segments(const TraceSegment* segs, uint32_t nseg, uint32_t n, int mode,
                                    const TravScratch& ts, hipStream_t stream);

// quality.hip
obvhs_status tree_levels(obvhs_context* ctx, obvhs_bvh2* bvh, std::vector<uint32_t>& meta_host, uint32_t& levels,
                         const uint8_t* expand);
obvhs_status cost_table_levels(obvhs_context* ctx, obvhs_bvh2* bvh, const DpView& dp, uint32_t* d_err);
obvhs_status bvh2_refit_all(obvhs_context* ctx, obvhs_bvh2* bvh);
obvhs_status bvh2_refit_from_triangles(obvhs_context* ctx, obvhs_bvh2* bvh, const obvhs_triangle* d_tris);
// dp_max_prims != 0: also evaluates the conversion's cost table for that leaf size over the final tree (fused with
// the last refit), so the conversion that follows skips its own pass over the levels.
obvhs_status bvh2_reinsertion(obvhs_context* ctx, obvhs_bvh2* bvh, float batch_size_ratio, const float* seq,
                              uint32_t seq_len, uint32_t dp_max_prims = 0u);
obvhs_status bvh2_rebuild_path_flags(obvhs_context* ctx, obvhs_bvh2* bvh, const uint32_t* d_leaves, uint32_t count,
                                     uint8_t* d_flags);
obvhs_status bvh2_partial_rebuild(obvhs_context* ctx, obvhs_bvh2* bvh, const uint8_t* d_flags,
                                  uint32_t search_distance, bool u128, uint32_t search_depth_threshold);
obvhs_status bvh2_full_rebuild(obvhs_context* ctx, obvhs_bvh2* bvh, uint32_t search_distance, bool u128,
                               uint32_t search_depth_threshold);
obvhs_status bvh2_set_node_aabbs(obvhs_context* ctx, obvhs_bvh2* bvh, const uint32_t* d_ids,
                                 const obvhs_aabb* d_aabbs, uint32_t count);
obvhs_status bvh2_primitives_to_nodes(obvhs_context* ctx, const obvhs_bvh2* bvh, uint32_t* d_out);
obvhs_status presplit_leaves(obvhs_context* ctx, const obvhs_triangle* d_tris, uint32_t n, uint64_t max_total,
                             uint32_t& out_n, float* out_avg, float* out_largest);

// capi.cpp
obvhs_status set_hip_error(obvhs_context* ctx, hipError_t e, const char* where);
obvhs_status device_error_to_status(uint32_t err_bits);

}  // namespace obvhs
