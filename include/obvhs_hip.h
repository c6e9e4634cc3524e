/* obvhs_hip.h -- C ABI of the MI355X (gfx950) implementation of the obvhs hot path:
 *   PLOC BVH2 construction -> BVH2-to-CWBVH (80-byte node) conversion -> CWBVH closest-hit / occlusion traversal.
 *
 * Every entry point cites the obvhs v0.3.1 interface it replaces (paths relative to the crate root).  Plain C:
 * pointers and sizes only.  All device pointers are HIP device memory on the context's device; all work is
 * enqueued on the context's stream unless the function says it synchronises.  A context is externally
 * synchronised (one host thread at a time); separate contexts may be used concurrently.  Handles returned by the
 * build calls own their device memory until destroyed.  Each reference panic site maps to one status code.
 */
#ifndef OBVHS_HIP_H
#define OBVHS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum obvhs_status {
    OBVHS_OK = 0,
    OBVHS_ERR_INVALID_ARG = 1,      /* argument outside what the reference accepts (ploc/mod.rs:559 etc.) */
    OBVHS_ERR_NAN_INPUT = 2,        /* NaN primitive bounds: ploc/mod.rs:191-192 debug_assert */
    OBVHS_ERR_PLOC_SELF_MERGE = 3,  /* a cluster without any finite-cost neighbour: ploc/mod.rs:451,466 panic */
    OBVHS_ERR_CW_DISTRIBUTE = 4,    /* bvh2_to_cwbvh.rs:365-366 debug_assert (infinite costs) */
    OBVHS_ERR_CW_NONUNIT_LEAF = 5,  /* bvh2_to_cwbvh.rs:243-247 panic */
    OBVHS_ERR_CW_ORDER = 6,         /* bvh2_to_cwbvh.rs:460-465: a child left without a slot */
    OBVHS_ERR_CW_PRIMCOUNT = 7,     /* bvh2_to_cwbvh.rs:155 panic */
    OBVHS_ERR_CAPACITY = 8,         /* an internal fixed-size structure was exceeded */
    OBVHS_ERR_DEPTH = 9,            /* CWBVH deeper than the 32-entry traversal stack: cwbvh/mod.rs:791 */
    OBVHS_ERR_OOM = 10,             /* hipMalloc failed */
    OBVHS_ERR_DEVICE = 11,          /* any other HIP runtime error; see obvhs_last_error_message */
    OBVHS_ERR_NO_DEVICE = 12        /* no HIP device present */
} obvhs_status;

/* ---- plain-data types, byte-identical to the crate's repr(C) types ---- */
typedef struct obvhs_vec3a { float x, y, z, w; } obvhs_vec3a;                       /* glam::Vec3A, 16 B */
typedef struct obvhs_aabb { obvhs_vec3a min, max; } obvhs_aabb;                     /* aabb.rs:11-16, 32 B */
typedef struct obvhs_triangle { obvhs_vec3a v0, v1, v2; } obvhs_triangle;           /* triangle.rs:8-13, 48 B */
typedef struct obvhs_rt_triangle { obvhs_vec3a v0, e1, e2, ng; } obvhs_rt_triangle; /* rt_triangle.rs:147-160, 64 B */
typedef struct obvhs_ray {                                                          /* ray.rs:15-30, 64 B */
    obvhs_vec3a origin, direction, inv_direction;
    float tmin, tmax;
    float _pad[2];
} obvhs_ray;
typedef struct obvhs_ray_hit { uint32_t primitive_id, geometry_id, instance_id; float t; } obvhs_ray_hit;
/* ray.rs:63-70 */
typedef struct obvhs_bvh2_node {                                                    /* bvh2/node.rs:38-66, 48 B */
    obvhs_aabb aabb;
    uint32_t prim_count, first_index, meta1, meta2;
} obvhs_bvh2_node;
typedef struct obvhs_bvh2_node_small {                                              /* bvh2/node.rs:12-36, 32 B */
    float min[3];
    uint32_t prim_count;
    float max[3];
    uint32_t first_index;
} obvhs_bvh2_node_small;
typedef struct obvhs_cwbvh_node {                                                   /* cwbvh/node.rs:12-54, 80 B */
    float p[3];
    uint8_t e[3];
    uint8_t imask;
    uint32_t child_base_idx;
    uint32_t primitive_base_idx;
    uint8_t child_meta[8];
    uint8_t child_min_x[8], child_max_x[8], child_min_y[8], child_max_y[8], child_min_z[8], child_max_z[8];
} obvhs_cwbvh_node;

/* BvhBuildParams (lib.rs:207-230).  ploc_search_distance is the numeric distance {1,2,6,14,24,32}
 * (PlocSearchDistance, ploc/mod.rs:534-562), or 0 for the distance of the context's tuning table (below);
 * sort_precision 64 or 128 (SortPrecision, ploc/mod.rs:657-661).  Every search_distance argument below takes 0
 * the same way. */
typedef struct obvhs_build_params {
    uint32_t pre_split;
    uint32_t ploc_search_distance;
    uint32_t search_depth_threshold;
    float reinsertion_batch_ratio;
    float post_collapse_reinsertion_batch_ratio_multiplier;
    uint32_t sort_precision;
    uint32_t max_prims_per_leaf;
    float collapse_traversal_cost;
} obvhs_build_params;

typedef struct obvhs_context obvhs_context;
typedef struct obvhs_bvh2 obvhs_bvh2;
typedef struct obvhs_cwbvh obvhs_cwbvh;
typedef struct obvhs_batch obvhs_batch;

/* Per-build statistics (the reference's `scope!` stage names, lib.rs:197-205). */
typedef struct obvhs_build_stats {
    uint32_t prim_count;
    uint32_t ploc_iterations;     /* PLOC passes (depth in ploc/mod.rs:492) */
    uint32_t bvh2_node_count;
    uint32_t cwbvh_node_count;
    uint64_t ploc_cluster_sum;    /* sum of the cluster counts over all PLOC passes */
} obvhs_build_stats;

/* ---- context, memory, timing (no reference counterpart: the crate is CPU-only) ----
 * A context keeps grow-only working memory until it is destroyed: build scratch for its largest build and 224 bytes
 * of traversal stack spill per ray of its largest traversal launch (at most 16M rays' worth, 3.8 GB). */
const char* obvhs_status_string(obvhs_status s);
obvhs_status obvhs_context_create(int device, void* hip_stream, obvhs_context** out);
obvhs_status obvhs_context_destroy(obvhs_context* ctx);
obvhs_status obvhs_context_synchronize(obvhs_context* ctx);
const char* obvhs_last_error_message(const obvhs_context* ctx);
obvhs_status obvhs_device_malloc(obvhs_context* ctx, uint64_t bytes, void** dptr);
obvhs_status obvhs_device_free(obvhs_context* ctx, void* dptr);
obvhs_status obvhs_memcpy_htod(obvhs_context* ctx, void* dst, const void* src, uint64_t bytes); /* sync */
obvhs_status obvhs_memcpy_dtoh(obvhs_context* ctx, void* dst, const void* src, uint64_t bytes); /* sync */
obvhs_status obvhs_memset_device(obvhs_context* ctx, void* dst, int value, uint64_t bytes);      /* async */
obvhs_status obvhs_memcpy_dtod(obvhs_context* ctx, void* dst, const void* src, uint64_t bytes);  /* async */
/* Keep each build's Morton codes and sort permutation for obvhs_bvh2_copy_sort_trace (parity tests). */
obvhs_status obvhs_context_set_trace(obvhs_context* ctx, int enable);

/* ---- tuning table (no reference counterpart: choices of this implementation, per context) ----
 * Size class of n primitives (or sort keys): 0 below 4,096, 1 below 65,536, 2 below 1,048,576, else 3.
 *   ploc_radius[c]    : the PlocSearchDistance a build with search distance 0 ("auto") uses for a BVH of class c;
 *                       batched builds resolve it per mesh.  Changes the tree (each result equals the build with
 *                       that distance).  Default 14.
 *   sort_items[c]     : keys per thread of the radix sort's tiles (4, 8 or 16) for a sort of class c.  Default 8.
 *   trav_lds_depth[c] : traversal stack entries kept in LDS (4, 8 or 16) for a BVH of class c.  Default 8.
 *   trav_chunk[c]     : 1: a single-BVH traversal launch of at least two rays per resident lane gives each wave
 *                       interleaved rows of rays and refills a lane as soon as its ray ends (faster for incoherent
 *                       rays, slower for coherent ones); 0: one lane per ray.  Default 0.
 * The last three never change a result, only the speed.  obvhs_tune_* fill entries by measurement. */
typedef struct obvhs_tune_table {
    uint32_t ploc_radius[4];
    uint32_t sort_items[4];
    uint32_t trav_lds_depth[4];
    uint32_t trav_chunk[4];
} obvhs_tune_table;
uint32_t obvhs_size_class(uint64_t n);
obvhs_status obvhs_tune_get(const obvhs_context* ctx, obvhs_tune_table* out);
obvhs_status obvhs_tune_set(obvhs_context* ctx, const obvhs_tune_table* table); /* rejects values out of range */
/* Times the radix sort of n seeded 63-bit keys with each tile size and stores the fastest in sort_items[class(n)]. */
obvhs_status obvhs_tune_sort(obvhs_context* ctx, uint32_t n, uint32_t* out_items);
/* Times closest-hit (any_hit = 0) or occlusion traversal of the given rays with each LDS depth, one lane per ray
 * and chunked, and stores the fastest in trav_lds_depth[class(prim count)] and trav_chunk[class(prim count)]. */
obvhs_status obvhs_tune_traversal(obvhs_context* ctx, const obvhs_cwbvh* bvh, const obvhs_triangle* d_bvh_tris,
                                  const obvhs_ray* d_rays, uint32_t n, int any_hit, uint32_t* out_depth);
/* Builds the triangles with every PlocSearchDistance (1, 2, 6, 14, 24, 32), evaluates each tree's CWBVH cost for
 * max_prims_per_leaf (the conversion's cost model, calculate_cost bvh2_to_cwbvh.rs:220-344, at the root), and stores
 * in ploc_radius[class(n)] the smallest distance whose cost is within `tolerance` (relative) of the reference cost:
 * the cost at distance `baseline` (one of the six: the fastest build as good as that distance's tree), or with
 * baseline 0 the lowest of the six (the best tree).  The choice depends on the costs only, not on timing, so it is
 * reproducible.  out_costs (6 entries) may be NULL. */
obvhs_status obvhs_tune_ploc_radius(obvhs_context* ctx, const obvhs_triangle* d_tris, uint32_t n,
                                    uint32_t max_prims_per_leaf, uint32_t baseline, float tolerance,
                                    uint32_t* out_radius, float* out_costs);
/* Stream timer: records a start/stop event on the context stream; obvhs_timer_stop synchronises. */
obvhs_status obvhs_timer_start(obvhs_context* ctx);
obvhs_status obvhs_timer_stop(obvhs_context* ctx, float* ms);

/* ---- PLOC BVH2 build ----
 * Replaces PlocBuilder::build / build_with_bvh (ploc/mod.rs:95-113, 129-159) for T = Aabb.  sort_precision 64
 * (SortPrecision::U64, 63-bit keys) or 128 (U128, 126-bit keys).  d_aabbs: n primitive AABBs (device).
 * d_indices: the `indices` vector (device), or NULL for 0..n.  `reuse` may be a handle from an earlier build
 * whose allocations are reused (build_with_bvh); pass NULL to allocate a new handle. */
obvhs_status obvhs_ploc_build_aabbs(obvhs_context* ctx, const obvhs_aabb* d_aabbs, const uint32_t* d_indices,
                                    uint32_t n, uint32_t search_distance, uint32_t sort_precision,
                                    uint32_t search_depth_threshold, obvhs_bvh2* reuse, obvhs_bvh2** out);
/* Same for T = Triangle (Boundable via Triangle::aabb, triangle.rs:28-30). */
obvhs_status obvhs_ploc_build_triangles(obvhs_context* ctx, const obvhs_triangle* d_tris, const uint32_t* d_indices,
                                        uint32_t n, uint32_t search_distance, uint32_t sort_precision,
                                        uint32_t search_depth_threshold, obvhs_bvh2* reuse, obvhs_bvh2** out);
/* Bvh2 fields (bvh2/mod.rs:31-85): node count (2n-1), primitive count, max_depth (ploc/mod.rs:501). */
obvhs_status obvhs_bvh2_info(const obvhs_bvh2* bvh, uint32_t* node_count, uint32_t* prim_count, uint32_t* max_depth);
/* Copies Bvh2::nodes in the default 48-byte layout, or the 32-byte small_bvh2_node layout (synchronous). */
obvhs_status obvhs_bvh2_copy_nodes(obvhs_context* ctx, const obvhs_bvh2* bvh, obvhs_bvh2_node* host_out);
obvhs_status obvhs_bvh2_copy_nodes_small(obvhs_context* ctx, const obvhs_bvh2* bvh, obvhs_bvh2_node_small* host_out);
obvhs_status obvhs_bvh2_copy_primitive_indices(obvhs_context* ctx, const obvhs_bvh2* bvh, uint32_t* host_out);
/* Debug/parity outputs of the last build into this handle: Morton codes in original leaf order and the sorted
 * permutation (sorted position -> leaf index), ploc/mod.rs:782-846.  host_codes holds prim_count u64 after a U64
 * build and 2 * prim_count u64 after a U128 build (each code as its low word, then its high word). */
obvhs_status obvhs_bvh2_copy_sort_trace(obvhs_context* ctx, const obvhs_bvh2* bvh, uint64_t* host_codes,
                                        uint32_t* host_sorted_index);
obvhs_status obvhs_bvh2_destroy(obvhs_bvh2* bvh);

/* ---- CWBVH ----
 * Replaces bvh2_to_cwbvh(bvh2, max_prims_per_leaf, order_children, include_exact_node_aabbs)
 * (cwbvh/bvh2_to_cwbvh.rs:490-510).  include_exact_node_aabbs must be 0 (the flag is kept for the signature). */
obvhs_status obvhs_bvh2_to_cwbvh(obvhs_context* ctx, const obvhs_bvh2* bvh2, uint32_t max_prims_per_leaf,
                                 int order_children, int include_exact_node_aabbs, obvhs_cwbvh* reuse,
                                 obvhs_cwbvh** out);
/* Replaces build_cwbvh_from_tris (cwbvh/builder.rs:20-85): [pre-splits] -> PLOC -> [reinsertion with the
 * default ratio sequence] -> bvh2_to_cwbvh(.., clamp(max prims,1,3), true, false).  d_tris is device memory.
 * stats may be NULL. */
obvhs_status obvhs_build_cwbvh_from_tris(obvhs_context* ctx, const obvhs_triangle* d_tris, uint32_t n,
                                         const obvhs_build_params* params, obvhs_cwbvh* reuse, obvhs_cwbvh** out,
                                         obvhs_build_stats* stats);
/* Options the crate's helper does not expose, used through its building blocks instead:
 * split_max_total: stop splitting once this many leaves exist (0 = no limit, the crate's behaviour; splits are
 * taken in the crate's candidate order, so the first ones are the crate's); ratio_sequence: the `ratio_sequence`
 * argument of ReinsertionOptimizer::run (NULL = the default 1, 1/3, ..., 1/31; reinsertion.rs:40, 101-110). */
typedef struct obvhs_build_ext {
    uint64_t split_max_total;
    const float* ratio_sequence;  /* host array */
    uint32_t ratio_sequence_len;
} obvhs_build_ext;
obvhs_status obvhs_build_cwbvh_from_tris_ex(obvhs_context* ctx, const obvhs_triangle* d_tris, uint32_t n,
                                            const obvhs_build_params* params, const obvhs_build_ext* ext,
                                            obvhs_cwbvh* reuse, obvhs_cwbvh** out, obvhs_build_stats* stats);
/* The BVH2 a CwBvh handle was built from by obvhs_build_cwbvh_from_tris (owned by the CwBvh handle; NULL for
 * conversions).  Refit or reinsert it, then convert it again with obvhs_bvh2_to_cwbvh. */
obvhs_status obvhs_cwbvh_bvh2(const obvhs_cwbvh* bvh, obvhs_bvh2** out);

/* ---- BVH2 quality passes and refit ----
 * ReinsertionOptimizer::default().run(bvh, batch_size_ratio, ratio_sequence) (bvh2/reinsertion.rs:40-61):
 * ratio_sequence is a host array, NULL for the default.  Ties of the crate's two unstable sorts are ordered by
 * node index (candidates) and by candidate rank (reinsertions). */
obvhs_status obvhs_bvh2_reinsertion(obvhs_context* ctx, obvhs_bvh2* bvh, float batch_size_ratio,
                                    const float* ratio_sequence, uint32_t ratio_sequence_len);
/* Bvh2::refit_all (bvh2/mod.rs:527-569). */
obvhs_status obvhs_bvh2_refit_all(obvhs_context* ctx, obvhs_bvh2* bvh);
/* The refit example of Bvh2::refit_all (bvh2/mod.rs:506-526): every one-primitive leaf takes the box of its
 * triangle in d_tris (original order, device), then refit_all.  For BVH2s built without pre-splits. */
obvhs_status obvhs_bvh2_refit_from_triangles(obvhs_context* ctx, obvhs_bvh2* bvh, const obvhs_triangle* d_tris);
/* ---- BVH2 from its parts, leaf updates and rebuilds (ploc/rebuild.rs) ----
 * A Bvh2 from host arrays (the crate's Bvh2 fields nodes / primitive_indices are public, bvh2/mod.rs:31-85): 48-byte
 * nodes (bvh2/node.rs:38-66), node_count of them, and prim_count primitive indices.  Inner nodes must point inside
 * the array and leaves inside primitive_indices (OBVHS_ERR_INVALID_ARG otherwise). */
obvhs_status obvhs_bvh2_create_from_host(obvhs_context* ctx, const obvhs_bvh2_node* nodes, uint32_t node_count,
                                         const uint32_t* prim_indices, uint32_t prim_count, obvhs_bvh2** out);
/* Bvh2Node::set_aabb (bvh2/node.rs:116-125) for count nodes: d_aabbs[i] becomes the box of node d_node_ids[i]
 * (device arrays; ids distinct).  Nothing above them is refit, as examples/physics.rs:433-439 does before a
 * partial rebuild. */
obvhs_status obvhs_bvh2_set_node_aabbs(obvhs_context* ctx, obvhs_bvh2* bvh, const uint32_t* d_node_ids,
                                       const obvhs_aabb* d_aabbs, uint32_t count);
/* Bvh2::compute_primitives_to_nodes (bvh2/mod.rs:647-663): d_out[p] (prim_count entries, device) = slot of the leaf
 * holding primitive p, INVALID (0xffffffff) if none; the highest slot when several leaves hold it. */
obvhs_status obvhs_bvh2_primitives_to_nodes(obvhs_context* ctx, const obvhs_bvh2* bvh, uint32_t* d_out);
/* compute_rebuild_path_flags (ploc/rebuild.rs:12-42): d_flags (node_count bytes, device) = 1 for every given leaf
 * and each of its ancestors, 0 elsewhere; untouched for a BVH of fewer than two nodes.  Parents are computed here
 * (the crate requires init_parents_if_uninit first). */
obvhs_status obvhs_bvh2_rebuild_path_flags(obvhs_context* ctx, obvhs_bvh2* bvh, const uint32_t* d_leaf_ids,
                                           uint32_t count, uint8_t* d_flags);
/* PlocBuilder::partial_rebuild(bvh, |i| d_flags[i] != 0, search_distance, sort_precision,
 * search_depth_threshold) (ploc/rebuild.rs:98-134), sort_precision 64 or 128: the flagged inner nodes under the
 * root are opened, the leaves and unflagged subtrees they expose are rebuilt into a new top with PLOC, and the new
 * nodes go to the freed sibling pairs, highest first.  Subtrees that were not opened keep their slots.  Clusters
 * with equal Morton codes are ordered by their old slot (the crate's unstable sort leaves that order open). */
obvhs_status obvhs_bvh2_partial_rebuild(obvhs_context* ctx, obvhs_bvh2* bvh, const uint8_t* d_flags,
                                        uint32_t search_distance, uint32_t sort_precision,
                                        uint32_t search_depth_threshold);
/* PlocBuilder::full_rebuild(bvh, search_distance, sort_precision, search_depth_threshold)
 * (ploc/rebuild.rs:53-79): every leaf, in slot order, rebuilt with PLOC; the root's stored box is the scene box. */
obvhs_status obvhs_bvh2_full_rebuild(obvhs_context* ctx, obvhs_bvh2* bvh, uint32_t search_distance,
                                     uint32_t sort_precision, uint32_t search_depth_threshold);
/* The pre-split leaf set of build_cwbvh_from_tris with pre_split = true (cwbvh/builder.rs:26-55 +
 * split_aabbs_preset, splits.rs:16-33): writes out_count boxes and triangle indices to caller device arrays of
 * out_capacity entries (OBVHS_ERR_CAPACITY with *out_count set when they do not fit).  max_total as
 * obvhs_build_ext::split_max_total.  The builder's `avg_area` and `largest_half_area` (cwbvh/builder.rs:28-43) are
 * returned through the last two pointers when they are not NULL. */
obvhs_status obvhs_split_aabbs_preset(obvhs_context* ctx, const obvhs_triangle* d_tris, uint32_t n,
                                      uint64_t max_total, obvhs_aabb* d_out_aabbs, uint32_t* d_out_indices,
                                      uint32_t out_capacity, uint32_t* out_count, float* out_avg_half_area,
                                      float* out_largest_half_area);
/* CwBvh fields (cwbvh/mod.rs:43-55). */
obvhs_status obvhs_cwbvh_info(const obvhs_cwbvh* bvh, uint32_t* node_count, uint32_t* prim_count,
                              obvhs_aabb* total_aabb);
obvhs_status obvhs_cwbvh_copy_nodes(obvhs_context* ctx, const obvhs_cwbvh* bvh, obvhs_cwbvh_node* host_out);
obvhs_status obvhs_cwbvh_copy_primitive_indices(obvhs_context* ctx, const obvhs_cwbvh* bvh, uint32_t* host_out);
/* Device pointers of CwBvh::nodes and CwBvh::primitive_indices (valid until the handle is destroyed/rebuilt). */
obvhs_status obvhs_cwbvh_device_arrays(const obvhs_cwbvh* bvh, const obvhs_cwbvh_node** d_nodes,
                                       const uint32_t** d_primitive_indices);
/* bvh_tris[i] = tris[primitive_indices[i]] (the reordering of examples/obj_cwbvh.rs:63-67), device to device. */
obvhs_status obvhs_cwbvh_reorder_triangles(obvhs_context* ctx, const obvhs_cwbvh* bvh, const obvhs_triangle* d_tris,
                                           obvhs_triangle* d_bvh_tris);
obvhs_status obvhs_cwbvh_destroy(obvhs_cwbvh* bvh);

/* ---- rays ----
 * Ray::new (ray.rs:34-52) on the device for n rays: d_out[i] = {origin[i], direction[i], inv_direction, tmin, tmax}
 * with inv_direction = safe_inverse per component (ray.rs:6-12: |x| <= f32::EPSILON -> signum(x) / EPSILON, else
 * 1 / x, correctly rounded); the fourth lane of each vec3a and the padding are written as 0. */
obvhs_status obvhs_rays_new(obvhs_context* ctx, const obvhs_vec3a* d_origins, const obvhs_vec3a* d_directions,
                            uint32_t n, float tmin, float tmax, obvhs_ray* d_out);
/* RayHit::none() (ray.rs:76-83) into n device hit records. */
obvhs_status obvhs_ray_hits_none(obvhs_context* ctx, obvhs_ray_hit* d_hits, uint32_t n);

/* ---- traversal ----
 * Replaces CwBvh::ray_traverse (cwbvh/mod.rs:169-198) with the closure |ray, id| bvh_tris[id].intersect(ray)
 * (triangle.rs:35-76), for n rays at once.  d_bvh_tris: triangles in primitive_indices order.  Each hit starts as
 * RayHit::none() (ray.rs:76-83); primitive_id is the index into CwBvh::primitive_indices like the reference's. */
obvhs_status obvhs_cwbvh_ray_traverse(obvhs_context* ctx, const obvhs_cwbvh* bvh, const obvhs_triangle* d_bvh_tris,
                                      const obvhs_ray* d_rays, uint32_t n, obvhs_ray_hit* d_hits);
/* Replaces CwBvh::ray_traverse_miss (cwbvh/mod.rs:201-225): d_miss[i] = 1 when ray i hit nothing in [tmin,tmax]. */
obvhs_status obvhs_cwbvh_ray_traverse_miss(obvhs_context* ctx, const obvhs_cwbvh* bvh,
                                           const obvhs_triangle* d_bvh_tris, const obvhs_ray* d_rays, uint32_t n,
                                           uint8_t* d_miss);
/* Instrumented closest-hit traversal: also writes per-ray node-test and triangle-test counts (for the byte
 * model of the roofline).  Same hits as obvhs_cwbvh_ray_traverse. */
obvhs_status obvhs_cwbvh_ray_traverse_counted(obvhs_context* ctx, const obvhs_cwbvh* bvh,
                                              const obvhs_triangle* d_bvh_tris, const obvhs_ray* d_rays, uint32_t n,
                                              obvhs_ray_hit* d_hits, uint32_t* d_node_visits,
                                              uint32_t* d_tri_tests);
/* Watertight mode (no counterpart in the reference: an extra primitive test, reported beside the crate's).  The
 * traversal of obvhs_cwbvh_ray_traverse / _miss with the triangle test of Woop, Benthin and Wald, "Watertight
 * Ray/Triangle Intersection" (JCGT 2013) in place of Triangle::intersect (triangle.rs:35-76): shear onto the ray's
 * dominant axis, 2D edge functions recomputed in f64 when one is exactly 0, no culling, t in [tmin, tmax].  A ray
 * through an edge or vertex shared by several triangles hits at least one of them; hit distances can differ from
 * Triangle::intersect's in the last bits. */
obvhs_status obvhs_cwbvh_ray_traverse_watertight(obvhs_context* ctx, const obvhs_cwbvh* bvh,
                                                 const obvhs_triangle* d_bvh_tris, const obvhs_ray* d_rays,
                                                 uint32_t n, obvhs_ray_hit* d_hits);
obvhs_status obvhs_cwbvh_ray_traverse_miss_watertight(obvhs_context* ctx, const obvhs_cwbvh* bvh,
                                                      const obvhs_triangle* d_bvh_tris, const obvhs_ray* d_rays,
                                                      uint32_t n, uint8_t* d_miss);
/* RtTriangle primitives (rt_triangle.rs:147-239): the crate's intersection-optimised triangle (v0, e1 = v0 - v1,
 * e2 = v2 - v0, ng = e1 x e2, one Vec3A each).  obvhs_rt_triangles_from_triangles replaces RtTriangle::from(&Triangle)
 * (rt_triangle.rs:162-181) for n triangles; the traversals replace CwBvh::ray_traverse / ray_traverse_miss with the
 * closure |ray, id| rt_tris[id].intersect(ray) (rt_triangle.rs:200-239), which computes the same bits as
 * Triangle::intersect, so the hits equal obvhs_cwbvh_ray_traverse's. */
obvhs_status obvhs_rt_triangles_from_triangles(obvhs_context* ctx, const obvhs_triangle* d_tris, uint32_t n,
                                               obvhs_rt_triangle* d_out);
obvhs_status obvhs_cwbvh_ray_traverse_rt(obvhs_context* ctx, const obvhs_cwbvh* bvh,
                                         const obvhs_rt_triangle* d_bvh_rt_tris, const obvhs_ray* d_rays, uint32_t n,
                                         obvhs_ray_hit* d_hits);
obvhs_status obvhs_cwbvh_ray_traverse_miss_rt(obvhs_context* ctx, const obvhs_cwbvh* bvh,
                                              const obvhs_rt_triangle* d_bvh_rt_tris, const obvhs_ray* d_rays,
                                              uint32_t n, uint8_t* d_miss);
/* Closest-hit traversal that allocates nothing (no counterpart in the reference; for callers that cannot allocate:
 * code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The same traversal and hits
 * as obvhs_cwbvh_ray_traverse, but its stack spill columns live in a static device array of the library (96 MiB per
 * device, held until the process ends: 2^19 lanes x 24 entries x 8 B; the LDS part keeps at least 8 entries) whose
 * address obvhs_context_create looks up; at most 2^31 rays per call (OBVHS_ERR_CAPACITY beyond), lanes stride over
 * the rays.  Every call on a device shares that array: calls that overlap on one device (other streams or threads)
 * race; nothing in it needs clearing between calls (a spill entry is written before it is read). */
obvhs_status obvhs_cwbvh_ray_traverse_noalloc(obvhs_context* ctx, const obvhs_cwbvh* bvh,
                                              const obvhs_triangle* d_bvh_tris, const obvhs_ray* d_rays, uint32_t n,
                                              obvhs_ray_hit* d_hits);

/* ---- batched build / traversal (many independent meshes in shared launches) ----
 * No single reference entry point: equivalent to calling build_cwbvh_from_tris (cwbvh/builder.rs:20-85) once per
 * mesh, with identical per-mesh results.  d_tris: concatenated triangles (device); tri_offsets: host array of
 * mesh_count + 1 prefix offsets (tri_offsets[0] == 0).  Pre-splits, reinsertion and SortPrecision::U128 are not
 * supported (OBVHS_ERR_INVALID_ARG). */
obvhs_status obvhs_batch_build_cwbvh(obvhs_context* ctx, const obvhs_triangle* d_tris, const uint32_t* tri_offsets,
                                     uint32_t mesh_count, const obvhs_build_params* params, obvhs_batch* reuse,
                                     obvhs_batch** out);
obvhs_status obvhs_batch_info(const obvhs_batch* b, uint32_t* mesh_count, uint32_t* total_nodes,
                              uint32_t* total_prims, uint32_t* ploc_passes);
/* Mesh i's CwBvh::nodes are nodes[node_offset .. node_offset + node_count) of obvhs_batch_copy_nodes, its
 * primitive_indices (mesh-local) are prims[prim_offset .. prim_offset + prim_count). */
obvhs_status obvhs_batch_mesh_info(const obvhs_batch* b, uint32_t mesh, uint32_t* node_offset, uint32_t* node_count,
                                   uint32_t* prim_offset, uint32_t* prim_count);
obvhs_status obvhs_batch_copy_nodes(obvhs_context* ctx, const obvhs_batch* b, obvhs_cwbvh_node* host_out);
obvhs_status obvhs_batch_copy_primitive_indices(obvhs_context* ctx, const obvhs_batch* b, uint32_t* host_out);
/* Per mesh, bvh_tris[prim_offset + i] = tris[prim_offset + primitive_indices[i]] (examples/obj_cwbvh.rs:63-67). */
obvhs_status obvhs_batch_reorder_triangles(obvhs_context* ctx, const obvhs_batch* b, const obvhs_triangle* d_tris,
                                           obvhs_triangle* d_bvh_tris);
/* CwBvh::ray_traverse / ray_traverse_miss (cwbvh/mod.rs:169-225) for rays of many meshes in one launch: ray i is
 * traced against mesh d_ray_mesh[i]; hits' primitive_id indexes that mesh's primitive_indices. */
obvhs_status obvhs_batch_ray_traverse(obvhs_context* ctx, const obvhs_batch* b, const obvhs_triangle* d_bvh_tris,
                                      const obvhs_ray* d_rays, const uint32_t* d_ray_mesh, uint32_t n,
                                      obvhs_ray_hit* d_hits);
obvhs_status obvhs_batch_ray_traverse_miss(obvhs_context* ctx, const obvhs_batch* b,
                                           const obvhs_triangle* d_bvh_tris, const obvhs_ray* d_rays,
                                           const uint32_t* d_ray_mesh, uint32_t n, uint8_t* d_miss);
obvhs_status obvhs_batch_destroy(obvhs_batch* b);
/* ---- TLAS over the meshes of a batch ----
 * A CwBvh whose primitives are the meshes' bounds (each mesh's BVH2 root box; every mesh must be non-empty), built
 * like obvhs_build_cwbvh_from_tris over boxes with the given parameters (pre-splits and reinsertion ignored). */
obvhs_status obvhs_batch_build_tlas(obvhs_context* ctx, const obvhs_batch* b, const obvhs_build_params* params,
                                    obvhs_cwbvh** out);
/* CwBvh::ray_traverse_tlas_blas (cwbvh/mod.rs:341-491) with the batch's meshes as BLASes and Triangle::intersect
 * on d_bvh_tris (obvhs_batch_reorder_triangles): RayHit::primitive_id is the primitive slot inside the BLAS,
 * RayHit::geometry_id the BLAS (mesh) index.  Follows the crate's function as written, including that when several
 * TLAS leaves of one node are hit only the first one (highest bit) is entered. */
obvhs_status obvhs_batch_tlas_ray_traverse(obvhs_context* ctx, const obvhs_cwbvh* tlas, const obvhs_batch* b,
                                           const obvhs_triangle* d_bvh_tris, const obvhs_ray* d_rays, uint32_t n,
                                           obvhs_ray_hit* d_hits);

/* ---- serving (resident BVHs, continuous batching) ----
 * No single reference counterpart: a server keeps scenes (CWBVH + triangles) resident in HBM, coalesces ray
 * requests into shared traversal launches, runs refit / rebuild updates on a second stream, and admits requests in
 * submission order while the HBM budget allows.  Each request's result equals the single-request call:
 * CwBvh::ray_traverse / ray_traverse_miss (cwbvh/mod.rs:169-225) on the scene version current when it was
 * submitted; a refit is Bvh2::refit_all + bvh2_to_cwbvh (bvh2/mod.rs:506-569), a rebuild build_cwbvh_from_tris
 * (cwbvh/builder.rs:20-85) with the scene's parameters.  Device buffers passed in requests must stay valid until
 * obvhs_server_wait returns for that request. */
typedef struct obvhs_server obvhs_server;
enum {
    OBVHS_REQ_TRACE_CLOSEST = 0, /* out: RayHit[n] */
    OBVHS_REQ_TRACE_MISS = 1,    /* out: u8[n], 1 = the ray hit nothing */
    OBVHS_REQ_REFIT = 2,         /* moved triangles, same count and order (else the request fails) */
    OBVHS_REQ_REBUILD = 3        /* new triangle set */
};
typedef struct obvhs_server_config {
    uint64_t hbm_budget_bytes;  /* resident scenes + in-flight request buffers (0 = 200 GiB) */
    uint32_t max_batch_rays;    /* rays per coalesced launch (0 = 8 Mi) */
    uint32_t max_inflight;      /* coalesced launches in flight on the trace stream (0 = 2, at most 8) */
    uint32_t update_streams;    /* refits / rebuilds of different scenes run concurrently (0 = 2, at most 4) */
    uint32_t pad;
} obvhs_server_config;
typedef struct obvhs_request_info {
    obvhs_status status;
    uint32_t kind;
    uint32_t scene;
    uint32_t version;         /* scene version read (traces) or produced (updates) */
    uint64_t batch;           /* coalesced launch that carried a trace */
    uint32_t batch_requests;  /* requests in that launch */
    uint32_t pad;
    int64_t submit_ns, start_ns, done_ns;  /* server clock */
} obvhs_request_info;
typedef struct obvhs_server_request {
    uint64_t arrival_ns;  /* offset from the start of obvhs_server_replay */
    uint32_t kind;
    uint32_t scene;
    const void* rays;     /* traces: device Ray[n] */
    void* out;            /* traces: device RayHit[n] or u8[n] */
    const void* tris;     /* updates: device Triangle[n] */
    uint32_t n;
    uint32_t pad;
} obvhs_server_request;
typedef struct obvhs_server_stats {
    uint64_t launches, rays, segments, resident_bytes;
    uint32_t scenes;
    uint32_t pad;
} obvhs_server_stats;
obvhs_status obvhs_server_create(int device, const obvhs_server_config* cfg, obvhs_server** out);
obvhs_status obvhs_server_destroy(obvhs_server* server);
/* Builds a scene (version 0) from device triangles with the given build parameters (blocking). */
obvhs_status obvhs_server_add_scene(obvhs_server* server, const obvhs_triangle* d_tris, uint32_t n,
                                    const obvhs_build_params* params, uint32_t* out_scene);
obvhs_status obvhs_server_submit_trace(obvhs_server* server, uint32_t scene, uint32_t kind, const obvhs_ray* d_rays,
                                       uint32_t n, void* d_out, uint64_t* out_id);
obvhs_status obvhs_server_submit_update(obvhs_server* server, uint32_t scene, uint32_t kind,
                                        const obvhs_triangle* d_tris, uint32_t n, uint64_t* out_id);
/* Blocks until the request is done; fills info (may be NULL) and forgets the request. */
obvhs_status obvhs_server_wait(obvhs_server* server, uint64_t id, obvhs_request_info* info);
/* Submits list[i] at list[i].arrival_ns after the call starts (open-loop load), then waits for all of them. */
obvhs_status obvhs_server_replay(obvhs_server* server, const obvhs_server_request* list, uint32_t count,
                                 obvhs_request_info* infos);
obvhs_status obvhs_server_scene_info(obvhs_server* server, uint32_t scene, uint32_t* version, uint32_t* prim_count,
                                     uint32_t* cwbvh_nodes);
obvhs_status obvhs_server_get_stats(obvhs_server* server, obvhs_server_stats* stats);

/* Host copy of the CWBVH exponent table: entry E is the largest 23-bit mantissa M with
 * log2f(2^(E-127) * (1 + M/2^23)) == E - 127 under the host libm, i.e. the inputs whose
 * exp2(ceil(log2(x))) stays at 2^(E-127) in bvh2_to_cwbvh.rs:85-88.  Needs no device. */
obvhs_status obvhs_log2_threshold_table(uint32_t* out256);

#ifdef __cplusplus
}
#endif

#endif /* OBVHS_HIP_H */
