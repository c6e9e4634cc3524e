// PLOC BVH2 construction on gfx950.
//
// Restates PlocBuilder::build_ploc / build_ploc_from_leaves (ploc/mod.rs:169-503) as a sequence of data-parallel
// passes whose outputs are bit-identical to the reference's serial loop:
//   k_leaf_init_*  : leaf Bvh2Node per primitive (ploc/mod.rs:187-243) + scene bounds (atomic min/max on
//                    order-preserving keys; the zero sign of the bounds cannot change a Morton code)
//   k_morton       : 63-bit Morton code of the f32 centre normalised in f64 (ploc/mod.rs:287-288, 782-785;
//                    morton.rs:35-58), or the 126-bit code of SortPrecision::U128 (morton.rs:63-89) as two words
//   radix sort     : stable (sort.hip); a U128 key is sorted by its low word, then stably by its high word
//   k_gather       : sorted leaves become the first cluster array (ploc/mod.rs:829-846)
//   per pass       : k_ploc_search_count (ploc/mod.rs:329-419, window in LDS, and the per-tile counts of emitted
//                    clusters and merges), k_ploc_scan (their prefix sums; two launches over chunks for the largest
//                    passes),
//                    k_ploc_compact (ploc/mod.rs:420-487: the parent of a mutual pair is emitted at the higher
//                    index, its children go to nodes[insert - 2(s+1)] and [+1] where s is the merge's rank in the
//                    pass); passes of at most 1,024 tiles fuse the scan into the compaction (k_ploc_scan_compact)
//   k_ploc_tail    : once at most 2048 clusters remain, one workgroup finishes every remaining pass with the
//                    clusters resident in LDS and writes the root (ploc/mod.rs:499); half its waves search while
//                    the other half evaluate the cost-table entries of the previous pass's merges
// Pass state (cluster count, insert index, depth) lives on the device so passes are enqueued without host syncs;
// the host only reads the count every few passes to shrink the grids.
//
// ploc_rebuild runs the same passes over clusters taken from an existing tree (rebuild_from_leaves,
// ploc/rebuild.rs:136-183).  For a partial rebuild the children of the g-th merge (counted over all passes in the
// serial loop's order) go to the g-th freed sibling pair from the top (`slot_map`, the REBUILD placement of
// ploc/mod.rs:449-462); the plain placement is the special case slot_map[g] = 2n-3-2g.
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>

#include "common.hpp"
#include "cwdp.hpp"
#include "device_util.hpp"
#include "internal.hpp"

namespace obvhs {

namespace {

constexpr int T256 = 256;
constexpr uint32_t TAIL_MAX = 2048;  // 4096 measured 5% slower per build (the tail's passes are LDS and ALU bound)
constexpr int TAIL_THREADS = 1024;
constexpr int TAIL_ITEMS = TAIL_MAX / TAIL_THREADS;
constexpr int TAIL_SEARCH_THREADS = 512;  // the tail's search threads; the rest evaluate the previous pass's entries
constexpr uint32_t MAX_RADIUS = 32;
// Passes enqueued between host reads of the cluster count (1.29M-triangle build: 4 -> 1.97 ms, 8 -> 1.92, 16 -> 1.93).
constexpr int PASSES_PER_SYNC = 8;

struct Misc {
    uint32_t bounds[6];  // order keys: min x,y,z (atomicMin) then max x,y,z (atomicMax)
    uint32_t err;
    uint32_t pad;
    PlocState state[2];
};

__global__ void k_init_misc(Misc* misc, uint32_t n, uint32_t* hist) {
    if (threadIdx.x == 0) {
        hist[0] = 2u * n - 1u;
        for (int a = 0; a < 3; ++a) {
            misc->bounds[a] = 0xffffffffu;
            misc->bounds[3 + a] = 0u;
        }
        misc->err = 0u;
        misc->state[0] = PlocState{n, 2u * n - 1u, 0u, 0u, 0ull};
        misc->state[1] = PlocState{n, 2u * n - 1u, 0u, 0u, 0ull};
    }
}

// Reduces the leaf bounds of one block into the scene bounds.  Both the min and the max corner of every leaf go
// into both reductions, like `local_aabb.extend(aabb.min); local_aabb.extend(aabb.max)` (ploc/mod.rs:189-190).
__device__ __forceinline__ void reduce_scene_bounds(bool valid, const Box& b, bool nan, Misc* misc,
                                                    uint32_t* __restrict__ partials) {
    __shared__ uint32_t red[6][T256 / 64];
    uint32_t k[6];
    if (valid) {
        uint32_t kl[3] = {f32_order_key(b.lx), f32_order_key(b.ly), f32_order_key(b.lz)};
        uint32_t kh[3] = {f32_order_key(b.hx), f32_order_key(b.hy), f32_order_key(b.hz)};
        for (int a = 0; a < 3; ++a) {
            k[a] = min(kl[a], kh[a]);
            k[3 + a] = max(kl[a], kh[a]);
        }
    } else {
        for (int a = 0; a < 3; ++a) {
            k[a] = 0xffffffffu;
            k[3 + a] = 0u;
        }
    }
    for (int a = 0; a < 3; ++a) {
        k[a] = wave_min_u32(k[a]);
        k[3 + a] = wave_max_u32(k[3 + a]);
    }
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    if (lane == 0)
        for (int a = 0; a < 6; ++a) red[a][wave] = k[a];
    if (__ballot(nan) != 0ull && lane == 0) atomicOr(&misc->err, DERR_NAN_INPUT);
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        uint32_t v = red[a][0];
        for (int w = 1; w < T256 / 64; ++w) v = a < 3 ? min(v, red[a][w]) : max(v, red[a][w]);
        partials[blockIdx.x * 6 + a] = v;
    }
}

// Final reduction of the per-block scene bounds (one workgroup).
__global__ void __launch_bounds__(1024) k_reduce_bounds(const uint32_t* __restrict__ partials, uint32_t nblocks,
                                                        Misc* misc) {
    __shared__ uint32_t red[6][16];
    uint32_t k[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    for (uint32_t b = threadIdx.x; b < nblocks; b += 1024) {
        for (int a = 0; a < 3; ++a) {
            k[a] = min(k[a], partials[b * 6 + a]);
            k[3 + a] = max(k[3 + a], partials[b * 6 + 3 + a]);
        }
    }
    for (int a = 0; a < 3; ++a) {
        k[a] = wave_min_u32(k[a]);
        k[3 + a] = wave_max_u32(k[3 + a]);
    }
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    if (lane == 0)
        for (int a = 0; a < 6; ++a) red[a][wave] = k[a];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        uint32_t v = red[a][0];
        for (int w = 1; w < 16; ++w) v = a < 3 ? min(v, red[a][w]) : max(v, red[a][w]);
        misc->bounds[a] = v;
    }
}

// Triangle::aabb = Aabb::from_points([v0, v1, v2]) (triangle.rs:28-30; aabb.rs:59-66).
// iota_out (when the caller gave no indices): primitive_indices = 0..n-1, written here instead of by a launch of its own.
__global__ void __launch_bounds__(T256) k_leaf_init_tris(const float4* __restrict__ tris,
                                                         const uint32_t* __restrict__ indices, uint32_t n,
                                                         Node32* __restrict__ leaves, Misc* misc,
                                                         uint32_t* __restrict__ partials,
                                                         uint32_t* __restrict__ iota_out) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    const bool valid = i < n;
    if (iota_out && valid) iota_out[i] = i;
    Box b{0, 0, 0, 0, 0, 0};
    bool nan = false;
    if (valid) {
        float4 v0 = tris[3 * (size_t)i], v1 = tris[3 * (size_t)i + 1], v2 = tris[3 * (size_t)i + 2];
        b = Box{v0.x, v0.y, v0.z, v0.x, v0.y, v0.z};
        b = Box{smin(b.lx, v1.x), smin(b.ly, v1.y), smin(b.lz, v1.z), smax(b.hx, v1.x), smax(b.hy, v1.y),
                smax(b.hz, v1.z)};
        b = Box{smin(b.lx, v2.x), smin(b.ly, v2.y), smin(b.lz, v2.z), smax(b.hx, v2.x), smax(b.hy, v2.y),
                smax(b.hz, v2.z)};
        nan = box_has_nan(b);
        store_node(&leaves[i], make_node(b, 1u, indices ? indices[i] : i));
    }
    reduce_scene_bounds(valid, b, nan, misc, partials);
}

// T = Aabb: the leaf box is the input box (aabb.rs:209-214).
__global__ void __launch_bounds__(T256) k_leaf_init_aabbs(const float4* __restrict__ aabbs,
                                                          const uint32_t* __restrict__ indices, uint32_t n,
                                                          Node32* __restrict__ leaves, Misc* misc,
                                                          uint32_t* __restrict__ partials,
                                                          uint32_t* __restrict__ iota_out) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    const bool valid = i < n;
    if (iota_out && valid) iota_out[i] = i;
    Box b{0, 0, 0, 0, 0, 0};
    bool nan = false;
    if (valid) {
        float4 lo = aabbs[2 * (size_t)i], hi = aabbs[2 * (size_t)i + 1];
        b = Box{lo.x, lo.y, lo.z, hi.x, hi.y, hi.z};
        nan = box_has_nan(b);
        store_node(&leaves[i], make_node(b, 1u, indices ? indices[i] : i));
    }
    reduce_scene_bounds(valid, b, nan, misc, partials);
}

// ploc/mod.rs:287-288 and 782-785; morton.rs:55-58.
// rebuild_from_leaves takes the root's stored box as the scene box (ploc/rebuild.rs:154).
__global__ void k_bounds_from_node(Misc* misc, const Node32* __restrict__ nodes) {
    if (threadIdx.x == 0) {
        const Node32 r = load_node(&nodes[0]);
        misc->bounds[0] = f32_order_key(r.minx);
        misc->bounds[1] = f32_order_key(r.miny);
        misc->bounds[2] = f32_order_key(r.minz);
        misc->bounds[3] = f32_order_key(r.maxx);
        misc->bounds[4] = f32_order_key(r.maxy);
        misc->bounds[5] = f32_order_key(r.maxz);
    }
}

// Child slot of a merge whose plain-placement left slot is v (see the header).  Returns false (and flags the error)
// when v is outside the map, which a well-formed pass sequence never produces.
__device__ __forceinline__ bool child_slot(uint32_t v, const uint32_t* __restrict__ slot_map, uint32_t vtop,
                                           uint32_t map_len, uint32_t& first) {
    if (!slot_map) {
        first = v;
        return true;
    }
    const uint32_t g = (vtop - v) >> 1;
    if (v > vtop || g >= map_len) return false;
    first = slot_map[g];
    return true;
}

// keys_hi = NULL: U64 codes into keys.  Otherwise U128 codes, low words into keys and high words into keys_hi.
__global__ void __launch_bounds__(T256) k_morton(const Node32* __restrict__ leaves, uint32_t n,
                                                 const Misc* __restrict__ misc, uint64_t* __restrict__ keys,
                                                 uint64_t* __restrict__ keys_hi) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i >= n) return;
    const float mnx = f32_from_order_key(misc->bounds[0]), mny = f32_from_order_key(misc->bounds[1]),
                mnz = f32_from_order_key(misc->bounds[2]);
    const float mxx = f32_from_order_key(misc->bounds[3]), mxy = f32_from_order_key(misc->bounds[4]),
                mxz = f32_from_order_key(misc->bounds[5]);
    const double sx = 1.0 / (double)(mxx - mnx), sy = 1.0 / (double)(mxy - mny), sz = 1.0 / (double)(mxz - mnz);
    const double ox = -(double)mnx * sx, oy = -(double)mny * sy, oz = -(double)mnz * sz;
    Node32 l = load_node(&leaves[i]);
    const float cx = (l.maxx + l.minx) * 0.5f, cy = (l.maxy + l.miny) * 0.5f, cz = (l.maxz + l.minz) * 0.5f;
    const double px = (double)cx * sx + ox, py = (double)cy * sy + oy, pz = (double)cz * sz + oz;
    if (keys_hi) {
        morton128_of_point(px, py, pz, keys[i], keys_hi[i]);
        return;
    }
    const double s = 2097152.0;
    keys[i] = split_by_3_u64(f64_as_u32(px * s)) | split_by_3_u64(f64_as_u32(py * s)) << 1 |
              split_by_3_u64(f64_as_u32(pz * s)) << 2;
}

__global__ void __launch_bounds__(T256) k_gather_u64(const uint64_t* __restrict__ src,
                                                     const uint32_t* __restrict__ order, uint32_t n,
                                                     uint64_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i >= n) return;
    out[i] = src[order[i]];
}

__global__ void __launch_bounds__(T256) k_gather(const Node32* __restrict__ leaves,
                                                 const uint32_t* __restrict__ order, uint32_t n,
                                                 Node32* __restrict__ out) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i >= n) return;
    store_node(&out[i], load_node(&leaves[order[i]]));
}

// Classification of cluster i (ploc/mod.rs:423-439): `emit` when it goes to the next pass (kept, or the higher
// index of a mutual pair, which becomes the parent), `mrg` when it is the higher index of a mutual pair.
// A self-pair (no finite neighbour, already flagged as an error) is treated as kept so that no slot underflows.
__device__ __forceinline__ void classify(const int8_t* merge, uint32_t i, uint32_t count, bool& emit, bool& mrg,
                                         uint32_t& b) {
    emit = false;
    mrg = false;
    b = i;
    if (i >= count) return;
    if (count == 1) {
        emit = true;
        return;
    }
    int m = merge[i];
    int bi = (int)i + m;
    b = (uint32_t)bi;
    bool mutual = bi + (int)merge[bi] == (int)i;
    if (!mutual || bi == (int)i) {
        emit = true;
    } else if (bi < (int)i) {
        emit = true;
        mrg = true;
    }
}

// The neighbour search of one pass (ploc/mod.rs:329-419) and the per-tile counts of emitted clusters and merges in
// one launch.  Each cluster's best neighbour is the LAST minimum over its window in ascending index order (`<=`,
// find_best_node_basic ploc/mod.rs:508-528, equal to the cached search).  Besides its 256 clusters, a workgroup
// searches the r clusters on either side of its tile (the halo) with threads of their own (SC_THREADS = 256 +
// 2 * MAX_RADIUS), so that it classifies each of its clusters against its neighbour's choice (ploc/mod.rs:423-439)
// without another launch; the window in LDS spans [base - 2r, base + 256 + 2r).  Every workgroup writes its (emit,
// merge) counts, zero beyond the cluster count.
constexpr int SC_THREADS = T256 + 2 * (int)MAX_RADIUS;
constexpr int SC_WINDOW = T256 + 4 * (int)MAX_RADIUS;

// Best neighbour offset of cluster c over a window starting at cluster wstart.
__device__ __forceinline__ int8_t cluster_merge(const float (*s)[SC_WINDOW], int wstart, uint32_t c, uint32_t count,
                                                uint32_t r, uint32_t dist1, bool nonan, uint32_t* __restrict__ err) {
    const int lc = (int)c - wstart;
    const Box me{s[0][lc], s[1][lc], s[2][lc], s[3][lc], s[4][lc], s[5][lc]};
    if (dist1) {
        if (c == count - 1) return (int8_t)-1;
        const float cost = half_area(box_union(me, Box{s[0][lc + 1], s[1][lc + 1], s[2][lc + 1], s[3][lc + 1],
                                                      s[4][lc + 1], s[5][lc + 1]}));
        float last = INFINITY;
        if (c > 0)
            last = half_area(box_union(Box{s[0][lc - 1], s[1][lc - 1], s[2][lc - 1], s[3][lc - 1], s[4][lc - 1],
                                           s[5][lc - 1]},
                                       me));
        return last < cost ? (int8_t)-1 : (int8_t)1;
    }
    const uint32_t begin = c - min(r, c);
    const uint32_t end = min(c + r + 1u, count);
    uint32_t best = c;
    float best_cost = INFINITY;
    for (uint32_t j = begin; j < end; ++j) {
        if (j == c) continue;
        const int o = (int)j - wstart;
        const Box ob{s[0][o], s[1][o], s[2][o], s[3][o], s[4][o], s[5][o]};
        const float cost = nonan ? union_area_nonan(me, ob) : half_area(box_union(me, ob));
        if (cost <= best_cost) {
            best = j;
            best_cost = cost;
        }
    }
    if (best == c) atomicOr(err, DERR_SELF_MERGE);
    return (int8_t)((int)best - (int)c);
}

// The same search for a distance R known at compile time, over a window whose slots outside [0, count) hold NaN
// boxes: a NaN cost never passes `cost <= best_cost`, so the unrolled offsets -R..R (ascending, as j ascends) pick
// what the bounded loop picks, with the LDS addresses as immediates and no per-neighbour bounds test.
template <int R, bool NONAN>
__device__ __forceinline__ int cluster_merge_unrolled(const float* __restrict__ x0, const float* __restrict__ y0,
                                                      const float* __restrict__ z0, const float* __restrict__ x1,
                                                      const float* __restrict__ y1, const float* __restrict__ z1,
                                                      int lc) {
    // U neighbours per unrolled block (U divides R); the blocks themselves stay a loop so that the compiler does
    // not hoist every neighbour's loads (2R boxes would not fit the registers).
    constexpr int U = R <= 8 ? R : (R % 8 == 0 ? 8 : (R % 7 == 0 ? 7 : 6));
    static_assert(R % U == 0, "block size must divide the distance");
    const Box me{x0[lc], y0[lc], z0[lc], x1[lc], y1[lc], z1[lc]};
    int best = 0;
    float best_cost = INFINITY;
    auto block = [&](int k0) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u;
            const int o = lc + k;
            const Box ob{x0[o], y0[o], z0[o], x1[o], y1[o], z1[o]};
            const float cost = NONAN ? union_area_nonan(me, ob) : half_area(box_union(me, ob));
            if (cost <= best_cost) {
                best = k;
                best_cost = cost;
            }
        }
    };
#pragma unroll 1
    for (int k0 = -R; k0 < 0; k0 += U) block(k0);
#pragma unroll 1
    for (int k0 = 1; k0 <= R; k0 += U) block(k0);
    return best;
}

// Symmetric form of the unrolled search for a NaN-free window: union_area_nonan(a, b) == union_area_nonan(b, a) bit
// for bit (IEEE minimum / maximum are commutative, signed zeros included), so slot w computes its R right-hand pair
// costs once into pc[k - 1][w] and a cluster reads its left-hand costs from the slots before it.
template <int R>
__device__ __forceinline__ void pair_costs(const float (*s)[SC_WINDOW], float (*pc)[SC_WINDOW], int w) {
    constexpr int U = R <= 8 ? R : (R % 8 == 0 ? 8 : (R % 7 == 0 ? 7 : 6));
    const Box me{s[0][w], s[1][w], s[2][w], s[3][w], s[4][w], s[5][w]};
#pragma unroll 1
    for (int k0 = 1; k0 <= R; k0 += U) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u, o = w + k;
            pc[k - 1][w] = union_area_nonan(me, Box{s[0][o], s[1][o], s[2][o], s[3][o], s[4][o], s[5][o]});
        }
    }
}

// The pick over the pair costs: offsets -R..-1 then 1..R (j ascending), the last minimum, as cluster_merge_unrolled.
template <int R>
__device__ __forceinline__ int pick_from_pairs(const float (*pc)[SC_WINDOW], int lc) {
    int best = 0;
    float best_cost = INFINITY;
#pragma unroll
    for (int k = R; k >= 1; --k) {
        const float cost = pc[k - 1][lc - k];
        if (cost <= best_cost) {
            best = -k;
            best_cost = cost;
        }
    }
#pragma unroll
    for (int k = 1; k <= R; ++k) {
        const float cost = pc[k - 1][lc];
        if (cost <= best_cost) {
            best = k;
            best_cost = cost;
        }
    }
    return best;
}

// R > 0: the search distance r, known at compile time (unrolled search over a NaN-padded window); R = 0: any r.
template <int R>
__global__ void __launch_bounds__(SC_THREADS) k_ploc_search_count(const Node32* __restrict__ cur, Misc* misc,
                                                                  uint32_t slot, uint32_t r, uint32_t dist1,
                                                                  int8_t* __restrict__ merge,
                                                                  uint2* __restrict__ blocksums) {
    __shared__ float s[6][SC_WINDOW];
    __shared__ int8_t mm[T256 + 2 * MAX_RADIUS];  // offsets of clusters [base - r, base + 256 + r)
    __shared__ uint32_t we[T256 / 64], wm[T256 / 64];
    const uint32_t count = misc->state[slot].count;
    const uint32_t base = blockIdx.x * T256;
    if (count <= 1 || base >= count) {  // workgroup-uniform
        if (threadIdx.x == 0)
            blocksums[blockIdx.x] = count == 1 && blockIdx.x == 0 ? make_uint2(1u, 0u) : make_uint2(0u, 0u);
        return;
    }
    const int wstart = (int)base - 2 * (int)r;
    const int wlen = T256 + 4 * (int)r;
    bool nan_here = false;
    for (int w = threadIdx.x; w < wlen; w += SC_THREADS) {
        const int g = wstart + w;
        if (g >= 0 && (uint32_t)g < count) {
            const Node32 nd = load_node(&cur[g]);
            s[0][w] = nd.minx;
            s[1][w] = nd.miny;
            s[2][w] = nd.minz;
            s[3][w] = nd.maxx;
            s[4][w] = nd.maxy;
            s[5][w] = nd.maxz;
            nan_here |= box_has_nan_any(node_box(nd));
        } else if (R > 0) {
#pragma unroll
            for (int a = 0; a < 6; ++a) s[a][w] = __builtin_nanf("");
        }
    }
    const bool nonan = !__syncthreads_or(nan_here);  // NaN-free window: costs from union_area_nonan
    // thread t < 256: cluster base + t; thread 256 + h, h < 2r: halo cluster base - r + h or base + 256 + h - r
    const uint32_t t = threadIdx.x;
    // NaN-free windows: every pair cost once (the searchers sit at slots [R, 256 + 3R), their partners below 256 + 4R)
    __shared__ float pcost[R > 0 ? R : 1][SC_WINDOW];
    const bool pairs = R > 0 && !dist1 && nonan;  // workgroup-uniform
    if constexpr (R > 0) {
        if (pairs) {
            for (int w = (int)t; w < T256 + 3 * R; w += SC_THREADS) pair_costs<R>(s, pcost, w);
            __syncthreads();
        }
    }
    int c = -1;
    if (t < (uint32_t)T256) c = (int)(base + t);
    else if (t - T256 < 2u * r) c = t - T256 < r ? (int)base - (int)r + (int)(t - T256) : (int)(base + t - r);
    if (c >= 0 && (uint32_t)c < count) {
        int8_t m;
        bool unrolled = false;
        if constexpr (R > 0) {
            if (!dist1) {
                const int lc = c - wstart;
                const int b = pairs ? pick_from_pairs<R>(pcost, lc)
                                    : cluster_merge_unrolled<R, false>(s[0], s[1], s[2], s[3], s[4], s[5], lc);
                if (b == 0) atomicOr(&misc->err, DERR_SELF_MERGE);
                m = (int8_t)b;
                unrolled = true;
            }
        }
        if (!unrolled) m = cluster_merge(s, wstart, (uint32_t)c, count, r, dist1, nonan, &misc->err);
        mm[c - ((int)base - (int)r)] = m;
        if (t < (uint32_t)T256) merge[c] = m;
    }
    __syncthreads();
    // classify (ploc/mod.rs:423-439) the tile's clusters from the offsets in LDS
    bool emit = false, mrg = false;
    if (t < (uint32_t)T256 && base + t < count) {
        const int i = (int)(base + t);
        const int bi = i + (int)mm[i - ((int)base - (int)r)];
        const bool mutual = bi + (int)mm[bi - ((int)base - (int)r)] == i;
        if (!mutual || bi == i) {
            emit = true;
        } else if (bi < i) {
            emit = true;
            mrg = true;
        }
    }
    const uint64_t be = __ballot(emit), bm = __ballot(mrg);
    const uint32_t wave = t >> 6;
    if ((t & 63u) == 0 && wave < (uint32_t)(T256 / 64)) {
        we[wave] = (uint32_t)__popcll(be);
        wm[wave] = (uint32_t)__popcll(bm);
    }
    __syncthreads();
    if (t == 0) {
        uint32_t E = 0, M = 0;
        for (int w = 0; w < T256 / 64; ++w) {
            E += we[w];
            M += wm[w];
        }
        blocksums[blockIdx.x] = make_uint2(E, M);
    }
}

// Single workgroup: exclusive scan of the per-block (emit, merge) counts and the next pass's state.
__global__ void __launch_bounds__(1024) k_ploc_scan(uint2* __restrict__ blocksums, uint32_t nblocks, Misc* misc,
                                                    uint32_t slot, uint32_t* __restrict__ hist) {
    __shared__ uint32_t se[16], sm[16];
    const uint32_t tid = threadIdx.x;
    const uint32_t chunk = (nblocks + 1023u) / 1024u;
    const uint32_t b0 = tid * chunk, b1 = min(nblocks, b0 + chunk);
    uint32_t E = 0, M = 0;
    // each thread's chunk is contiguous (the prefix order); unrolled so that several loads are in flight
#pragma unroll 8
    for (uint32_t b = b0; b < b1; ++b) {
        uint2 v = blocksums[b];
        E += v.x;
        M += v.y;
    }
    uint32_t ie = wave_inclusive_scan(E), im = wave_inclusive_scan(M);
    const uint32_t wave = tid >> 6, lane = tid & 63u;
    if (lane == 63) {
        se[wave] = ie;
        sm[wave] = im;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t re = 0, rm = 0;
        for (int w = 0; w < 16; ++w) {
            uint32_t te = se[w], tm = sm[w];
            se[w] = re;
            sm[w] = rm;
            re += te;
            rm += tm;
        }
        const PlocState st = misc->state[slot];
        PlocState nx = st;
        if (st.count > 1) {
            nx.count = re;
            nx.insert = st.insert - 2u * rm;
            nx.depth = st.depth + 1u;
            nx.cluster_sum = st.cluster_sum + st.count;
            hist[nx.depth] = nx.insert;
        }
        misc->state[slot ^ 1u] = nx;
    }
    __syncthreads();
    uint32_t ee = se[wave] + ie - E, em = sm[wave] + im - M;
#pragma unroll 8
    for (uint32_t b = b0; b < b1; ++b) {
        uint2 v = blocksums[b];
        blocksums[b] = make_uint2(ee, em);
        ee += v.x;
        em += v.y;
    }
}

// Two-level scan of the tile counts for passes of more than SCAN_ONE_BLOCK tiles (one workgroup reading every count
// twice takes ~50 us at 74K tiles): k_ploc_scan_chunks sums SCAN_CHUNK counts per workgroup, k_ploc_scan_apply
// offsets each chunk by the chunks before it and scans it, and its workgroup 0 writes the next pass's state.
constexpr uint32_t SCAN_ONE_BLOCK = 16384;   // passes of at most this many tiles use k_ploc_scan

__global__ void __launch_bounds__(1024) k_ploc_scan_chunks(const uint2* __restrict__ blocksums, uint32_t nblocks,
                                                           uint2* __restrict__ chunk_tot) {
    scan_chunk_total(blocksums, nblocks, chunk_tot);
}

__global__ void __launch_bounds__(1024) k_ploc_scan_apply(uint2* __restrict__ blocksums, uint32_t nblocks,
                                                          const uint2* __restrict__ chunk_tot, uint32_t nchunks,
                                                          Misc* misc, uint32_t slot, uint32_t* __restrict__ hist) {
    const uint2 all = scan_chunk_apply(blocksums, nblocks, chunk_tot, nchunks);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const PlocState st = misc->state[slot];
        PlocState nx = st;
        if (st.count > 1) {
            nx.count = all.x;
            nx.insert = st.insert - 2u * all.y;
            nx.depth = st.depth + 1u;
            nx.cluster_sum = st.cluster_sum + st.count;
            hist[nx.depth] = nx.insert;
        }
        misc->state[slot ^ 1u] = nx;
    }
}

// Compaction of one 256-cluster tile (ploc/mod.rs:420-493): kept clusters and the parents of mutual pairs go to
// `next` in order; a merge's two children go to the slots below the pass's insert index in the serial loop's order
// (`pe0` / `pm0`: the emits and merges of the tiles before this one).  The tile's merges are then
// listed in LDS and their cost-table entries evaluated by consecutive threads, so that a pass where a fifth of the
// clusters merge runs the evaluation in about one wave per tile instead of in every wave with a few lanes active.
// The caller classifies its cluster and loads its node (`a`, for i < count) at the start of the kernel, so that
// those loads overlap the tile offsets' loads.  Called by every thread of the workgroup (it holds workgroup
// barriers).
__device__ __forceinline__ void compact_tile(const Node32* __restrict__ cur, Node32* __restrict__ next,
                                             const Node32& a, bool emit, bool mrg, uint32_t b,
                                             const PlocState& st, uint32_t pe0, uint32_t pm0,
                                             Node32* __restrict__ bvh_nodes,
                                             const DpView& dp, uint32_t fused, const uint32_t* __restrict__ slot_map,
                                             uint32_t vtop, uint32_t map_len, uint32_t* __restrict__ err,
                                             uint32_t* we, uint32_t* wm) {
    // A merge is counted at its higher cluster; pairs inside the tile use two of its clusters, and at most r of its
    // clusters pair with distinct clusters in [base - r, base): at most 128 + r / 2 merges.
    constexpr uint32_t MAX_TILE_MERGES = (T256 + MAX_RADIUS) / 2;
    __shared__ Node32 s_left[MAX_TILE_MERGES], s_right[MAX_TILE_MERGES];
    __shared__ uint32_t s_first[MAX_TILE_MERGES];
    const uint64_t be = __ballot(emit), bm = __ballot(mrg);
    const uint64_t lt = lanemask_lt();
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0) {
        we[wave] = (uint32_t)__popcll(be);
        wm[wave] = (uint32_t)__popcll(bm);
    }
    __syncthreads();
    if (emit) {
        uint32_t pe = pe0, pm = pm0;
        for (uint32_t w = 0; w < wave; ++w) {
            pe += we[w];
            pm += wm[w];
        }
        const uint32_t pos = pe + (uint32_t)__popcll(be & lt);
        if (mrg) {
            const uint32_t sq = pm + (uint32_t)__popcll(bm & lt);
            const Node32 c = load_node(&cur[b]);
            uint32_t first;
            const uint32_t q = sq - pm0;  // the merge's rank in this tile
            if (q >= MAX_TILE_MERGES || !child_slot(st.insert - 2u * (sq + 1u), slot_map, vtop, map_len, first)) {
                atomicOr(err, DERR_CAPACITY);
                if (q < MAX_TILE_MERGES) s_first[q] = UINT32_MAX;
            } else {
                store_node(&bvh_nodes[first], a);
                store_node(&bvh_nodes[first + 1], c);
                store_node(&next[pos], make_node(box_union(node_box(a), node_box(c)), 0u, first));
                s_left[q] = a;
                s_right[q] = c;
                s_first[q] = first;
            }
        } else {
            store_node(&next[pos], a);
        }
    }
    if (!fused) return;  // uniform
    __syncthreads();
    const uint32_t nm = min(wm[0] + wm[1] + wm[2] + wm[3], MAX_TILE_MERGES);
    if (threadIdx.x < nm && s_first[threadIdx.x] != UINT32_MAX) {
        const Node32 l = s_left[threadIdx.x], r = s_right[threadIdx.x];
        const uint32_t e = dp_evaluate(dp, box_union(node_box(l), node_box(r)), s_first[threadIdx.x], l, r);
        if (e) atomicOr(err, e);
    }
}

__global__ void __launch_bounds__(T256) k_ploc_compact(const Node32* __restrict__ cur, Node32* __restrict__ next,
                                                       const int8_t* __restrict__ merge, Misc* misc, uint32_t slot,
                                                       const uint2* __restrict__ offsets,
                                                       Node32* __restrict__ bvh_nodes, DpView dp, uint32_t fused,
                                                       const uint32_t* __restrict__ slot_map, uint32_t vtop,
                                                       uint32_t map_len) {
    __shared__ uint32_t we[T256 / 64], wm[T256 / 64];
    const PlocState st = misc->state[slot];
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    Node32 a{};
    if (i < st.count) a = load_node(&cur[i]);
    bool emit, mrg;
    uint32_t b;
    classify(merge, i, st.count, emit, mrg, b);
    const uint2 off = offsets[blockIdx.x];
    compact_tile(cur, next, a, emit, mrg, b, st, off.x, off.y, bvh_nodes, dp, fused, slot_map, vtop, map_len,
                 &misc->err, we, wm);
}

// k_ploc_scan and k_ploc_compact in one launch for passes of at most SCAN_FUSED_BLOCKS tiles: each workgroup sums the
// (emit, merge) counts of the tiles before its own (at most 4 per thread), and workgroup 0 also sums them all and
// writes the next pass's state, as k_ploc_scan does.
constexpr uint32_t SCAN_FUSED_BLOCKS = 1024;
__global__ void __launch_bounds__(T256) k_ploc_scan_compact(const Node32* __restrict__ cur, Node32* __restrict__ next,
                                                            const int8_t* __restrict__ merge, Misc* misc,
                                                            uint32_t slot, const uint2* __restrict__ blocksums,
                                                            uint32_t nblocks, Node32* __restrict__ bvh_nodes, DpView dp,
                                                            uint32_t fused, const uint32_t* __restrict__ slot_map,
                                                            uint32_t vtop, uint32_t map_len,
                                                            uint32_t* __restrict__ hist) {
    __shared__ uint32_t we[T256 / 64], wm[T256 / 64];
    __shared__ uint2 pre;
    const PlocState st = misc->state[slot];
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    Node32 a{};
    if (i < st.count) a = load_node(&cur[i]);
    bool emit, mrg;
    uint32_t b;
    classify(merge, i, st.count, emit, mrg, b);
    const uint32_t upto = blockIdx.x == 0 ? nblocks : blockIdx.x;  // workgroup 0 sums every tile (the totals)
    uint32_t E = 0, M = 0;
    for (uint32_t t = threadIdx.x; t < upto; t += T256) {
        const uint2 v = blocksums[t];
        E += v.x;
        M += v.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
        E += (uint32_t)__shfl_xor((int)E, off);
        M += (uint32_t)__shfl_xor((int)M, off);
    }
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0) {
        we[wave] = E;
        wm[wave] = M;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t te = 0, tm = 0;
        for (int w = 0; w < T256 / 64; ++w) {
            te += we[w];
            tm += wm[w];
        }
        if (blockIdx.x == 0) {
            PlocState nx = st;
            if (st.count > 1) {
                nx.count = te;
                nx.insert = st.insert - 2u * tm;
                nx.depth = st.depth + 1u;
                nx.cluster_sum = st.cluster_sum + st.count;
                hist[nx.depth] = nx.insert;
            }
            misc->state[slot ^ 1u] = nx;
            pre = make_uint2(0u, 0u);
        } else {
            pre = make_uint2(te, tm);
        }
    }
    __syncthreads();
    const uint2 off = pre;
    __syncthreads();  // compact_tile reuses we / wm
    compact_tile(cur, next, a, emit, mrg, b, st, off.x, off.y, bvh_nodes, dp, fused, slot_map, vtop, map_len,
                 &misc->err, we, wm);
}

// Finishes the build inside one workgroup once count <= TAIL_MAX.  Clusters live in LDS as SoA.
// R > 0: the search distance, known at compile time (the unrolled search over NaN-padded arrays); R = 0: any r.
template <int R>
__global__ void __launch_bounds__(TAIL_THREADS) k_ploc_tail(const Node32* __restrict__ cur, Misc* misc,
                                                            uint32_t slot, uint32_t first_pass, uint32_t r,
                                                            uint32_t sd_is_one, uint32_t depth_threshold,
                                                            Node32* __restrict__ bvh_nodes,
                                                            uint32_t* __restrict__ hist, DpView dp, uint32_t fused,
                                                            const uint32_t* __restrict__ slot_map, uint32_t vtop,
                                                            uint32_t map_len) {
    // Box arrays with MAX_RADIUS slots on either side: NaN boxes before cluster 0 and after the last cluster, so that
    // the unrolled search needs no bounds test (refreshed after each pass's compaction).
    constexpr int PADDED = TAIL_MAX + 2 * MAX_RADIUS;
    __shared__ float lx_p[PADDED], ly_p[PADDED], lz_p[PADDED], hx_p[PADDED], hy_p[PADDED], hz_p[PADDED];
    float* const lx = lx_p + MAX_RADIUS;
    float* const ly = ly_p + MAX_RADIUS;
    float* const lz = lz_p + MAX_RADIUS;
    float* const hx = hx_p + MAX_RADIUS;
    float* const hy = hy_p + MAX_RADIUS;
    float* const hz = hz_p + MAX_RADIUS;
    __shared__ uint32_t pc[TAIL_MAX], fi[TAIL_MAX];
    __shared__ int8_t mg[TAIL_MAX];
    __shared__ uint32_t mlist[TAIL_MAX / 2];  // first child slot of every merge of the pass (cost table work list)
    __shared__ uint32_t wsum[TAIL_THREADS / 64];
    __shared__ uint32_t s_total;
    const uint32_t tid = threadIdx.x;
    PlocState st = misc->state[slot];
    uint32_t count = st.count;
    if (count > TAIL_MAX) {
        if (tid == 0) atomicOr(&misc->err, DERR_CAPACITY);
        return;
    }
    for (uint32_t i = tid; i < count; i += TAIL_THREADS) {
        Node32 nd = load_node(&cur[i]);
        lx[i] = nd.minx;
        ly[i] = nd.miny;
        lz[i] = nd.minz;
        hx[i] = nd.maxx;
        hy[i] = nd.maxy;
        hz[i] = nd.maxz;
        pc[i] = nd.prim_count;
        fi[i] = nd.first_index;
    }
    if (tid < 2u * MAX_RADIUS) {  // the pads: [-MAX_RADIUS, 0) and [count, count + MAX_RADIUS)
        const int p = tid < MAX_RADIUS ? -1 - (int)tid : (int)(count + tid - MAX_RADIUS);
        lx[p] = ly[p] = lz[p] = hx[p] = hy[p] = hz[p] = __builtin_nanf("");
    }
    bool nan_here = false;
    for (uint32_t i = tid; i < count; i += TAIL_THREADS)
        nan_here |= box_has_nan_any(Box{lx[i], ly[i], lz[i], hx[i], hy[i], hz[i]});
    // NaN-free clusters (merging keeps them so): search costs from union_area_nonan
    const bool nonan = !__syncthreads_or(nan_here);
    uint32_t insert = st.insert;
    uint32_t depth = st.depth;
    unsigned long long csum = st.cluster_sum;
    uint32_t pass = first_pass;
    bool self_merge = false;
    bool bad_slot = false;
    uint32_t dp_err = 0u;
    // Cost-table entries of a pass's merges are evaluated during the next pass's search, by the other half of the
    // workgroup: the search is ALU-bound and the entries wait on table reads, so the two overlap.
    uint32_t nm_prev = 0u;  // merges of the previous pass (their first child slots in mlist)
    auto dp_entries = [&](uint32_t j0, uint32_t stride) {
        for (uint32_t j = j0; j < nm_prev; j += stride) {
            const uint32_t first = mlist[j];
            const Node32 l = load_node(&bvh_nodes[first]), r = load_node(&bvh_nodes[first + 1]);
            dp_err |= dp_evaluate(dp, box_union(node_box(l), node_box(r)), first, l, r);
        }
    };
    while (count > 1) {
        csum += count;
        const bool dist1 = sd_is_one || pass < depth_threshold;
        if (tid >= TAIL_SEARCH_THREADS) dp_entries(tid - TAIL_SEARCH_THREADS, TAIL_THREADS - TAIL_SEARCH_THREADS);
        for (uint32_t i = tid; tid < TAIL_SEARCH_THREADS && i < count; i += TAIL_SEARCH_THREADS) {
            const Box me{lx[i], ly[i], lz[i], hx[i], hy[i], hz[i]};
            int8_t m;
            if (dist1) {
                if (i == count - 1) {
                    m = -1;
                } else {
                    const Box nx{lx[i + 1], ly[i + 1], lz[i + 1], hx[i + 1], hy[i + 1], hz[i + 1]};
                    float cost = half_area(box_union(me, nx));
                    float last = INFINITY;
                    if (i > 0) {
                        const Box pv{lx[i - 1], ly[i - 1], lz[i - 1], hx[i - 1], hy[i - 1], hz[i - 1]};
                        last = half_area(box_union(pv, me));
                    }
                    m = last < cost ? (int8_t)-1 : (int8_t)1;
                }
            } else if constexpr (R > 0) {
                const int b = nonan ? cluster_merge_unrolled<R, true>(lx, ly, lz, hx, hy, hz, (int)i)
                                    : cluster_merge_unrolled<R, false>(lx, ly, lz, hx, hy, hz, (int)i);
                if (b == 0) self_merge = true;
                m = (int8_t)b;
            } else {
                const uint32_t begin = i - min(r, i);
                const uint32_t end = min(i + r + 1u, count);
                uint32_t best = i;
                float best_cost = INFINITY;
                for (uint32_t j = begin; j < end; ++j) {
                    if (j == i) continue;
                    const Box o{lx[j], ly[j], lz[j], hx[j], hy[j], hz[j]};
                    const float c = nonan ? union_area_nonan(me, o) : half_area(box_union(me, o));
                    if (c <= best_cost) {
                        best = j;
                        best_cost = c;
                    }
                }
                if (best == i) self_merge = true;
                m = (int8_t)((int)best - (int)i);
            }
            mg[i] = m;
        }
        __syncthreads();
        // Blocked assignment for the ordered scan: thread t owns clusters [4t, 4t+4).
        bool emit[TAIL_ITEMS], mrg[TAIL_ITEMS];
        uint32_t bidx[TAIL_ITEMS];
        uint32_t local = 0;
#pragma unroll
        for (int q = 0; q < TAIL_ITEMS; ++q) {
            classify(mg, tid * TAIL_ITEMS + q, count, emit[q], mrg[q], bidx[q]);
            local += (emit[q] ? 1u : 0u) | (mrg[q] ? (1u << 16) : 0u);
        }
        uint32_t incl = wave_inclusive_scan(local);
        const uint32_t wave = tid >> 6, lane = tid & 63u;
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        if (tid == 0) {
            uint32_t run = 0;
            for (int w = 0; w < TAIL_THREADS / 64; ++w) {
                uint32_t t = wsum[w];
                wsum[w] = run;
                run += t;
            }
            s_total = run;
        }
        __syncthreads();
        uint32_t excl = wsum[wave] + incl - local;
        const uint32_t total = s_total;
        // Read phase (everything this thread writes is loaded before anyone overwrites the LDS arrays).
        Box va[TAIL_ITEMS], vb[TAIL_ITEMS];
        uint32_t vpc[TAIL_ITEMS], vfi[TAIL_ITEMS], vbpc[TAIL_ITEMS], vbfi[TAIL_ITEMS];
#pragma unroll
        for (int q = 0; q < TAIL_ITEMS; ++q) {
            const uint32_t i = tid * TAIL_ITEMS + q;
            if (emit[q]) {
                va[q] = Box{lx[i], ly[i], lz[i], hx[i], hy[i], hz[i]};
                vpc[q] = pc[i];
                vfi[q] = fi[i];
                if (mrg[q]) {
                    const uint32_t b = bidx[q];
                    vb[q] = Box{lx[b], ly[b], lz[b], hx[b], hy[b], hz[b]};
                    vbpc[q] = pc[b];
                    vbfi[q] = fi[b];
                }
            }
        }
        __syncthreads();
        uint32_t pe = excl & 0xffffu, pm = excl >> 16;
#pragma unroll
        for (int q = 0; q < TAIL_ITEMS; ++q) {
            if (!emit[q]) continue;
            Box outb = va[q];
            uint32_t opc = vpc[q], ofi = vfi[q];
            if (mrg[q]) {
                uint32_t first = 0u;
                if (child_slot(insert - 2u * (pm + 1u), slot_map, vtop, map_len, first)) {
                    const Node32 lnode = make_node(va[q], vpc[q], vfi[q]);
                    const Node32 rnode = make_node(vb[q], vbpc[q], vbfi[q]);
                    store_node(&bvh_nodes[first], lnode);
                    store_node(&bvh_nodes[first + 1], rnode);
                } else {
                    bad_slot = true;
                }
                outb = box_union(va[q], vb[q]);
                if (fused) mlist[pm] = first;
                opc = 0u;
                ofi = first;
                pm++;
            }
            lx[pe] = outb.lx;
            ly[pe] = outb.ly;
            lz[pe] = outb.lz;
            hx[pe] = outb.hx;
            hy[pe] = outb.hy;
            hz[pe] = outb.hz;
            pc[pe] = opc;
            fi[pe] = ofi;
            pe++;
        }
        if (tid < MAX_RADIUS) {  // the pad after the new count (old clusters' slots, read before the last barrier)
            const uint32_t p = (total & 0xffffu) + tid;
            lx[p] = ly[p] = lz[p] = hx[p] = hy[p] = hz[p] = __builtin_nanf("");
        }
        __syncthreads();
        nm_prev = fused ? (total >> 16) : 0u;
        count = total & 0xffffu;
        insert -= 2u * (total >> 16);
        depth += 1u;
        pass += 1u;
        if (tid == 0) hist[depth] = insert;
        if ((total >> 16) == 0u) {
            // No mutual pair in this pass: the reference would loop forever; report instead of hanging.
            self_merge = true;
            break;
        }
    }
    dp_entries(tid, TAIL_THREADS);  // the last pass's entries
    if (self_merge) atomicOr(&misc->err, DERR_SELF_MERGE);
    if (bad_slot) atomicOr(&misc->err, DERR_CAPACITY);
    if (dp_err) atomicOr(&misc->err, dp_err);
    if (tid == 0) {
        store_node(&bvh_nodes[0], make_node(Box{lx[0], ly[0], lz[0], hx[0], hy[0], hz[0]}, pc[0], fi[0]));
        PlocState out{count, insert, depth, 0u, csum};
        misc->state[slot] = out;
    }
}

}  // namespace

hipError_t DevBuf::ensure(size_t want) {
    if (want <= bytes && ptr) return hipSuccess;
    release();
    size_t sz = want < 256 ? 256 : want;
    hipError_t e = hipMalloc(&ptr, sz);
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        return e;
    }
    bytes = sz;
    return hipSuccess;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

#define OBVHS_TRY(expr, where)                                              \
    do {                                                                    \
        hipError_t _e = (expr);                                             \
        if (_e != hipSuccess) return set_hip_error(ctx, _e, where);         \
    } while (0)

// Sorts the leaf indices vals[0..n) (identity on entry) by their Morton codes, ties in leaf order
// (ploc/mod.rs:811-827).  U64: one sort of the 63-bit codes in keys.  U128 (keys_hi != NULL): LSD over the whole
// 126-bit key as a stable sort of the low words followed by a stable sort of the high words (62 bits) gathered in
// that order; keys ends up holding the sorted high words.
static obvhs_status sort_codes(obvhs_context* ctx, uint64_t* keys, const uint64_t* keys_hi, uint32_t* vals,
                               uint32_t n) {
    hipStream_t st = ctx->stream;
    void* scratch = ctx->scratch.sort_scratch.ptr;
    const int items = sort_items_for(ctx, n);
    if (!keys_hi) {
        OBVHS_TRY(radix_sort_pairs(keys, vals, true, n, 63, scratch, st, items), "radix sort");
        return OBVHS_OK;
    }
    OBVHS_TRY(radix_sort_pairs(keys, vals, true, n, 64, scratch, st, items), "radix sort (low words)");
    k_gather_u64<<<(n + T256 - 1) / T256, T256, 0, st>>>(keys_hi, vals, n, keys);
    OBVHS_TRY(hipGetLastError(), "gather high words");
    OBVHS_TRY(radix_sort_pairs(keys, vals, false, n, 62, scratch, st, items), "radix sort (high words)");
    return OBVHS_OK;
}

// The PLOC passes from the sorted clusters in `cur` to the root (ploc/mod.rs:327-499): grid passes while more than
// TAIL_MAX clusters remain, then the one-workgroup tail.  `fin`
// receives the final pass state.
// The error word and the pass state of `slot` with one copy (err, pad, state[0], state[1] are adjacent) into the
// pinned staging area, then a stream synchronisation.
static hipError_t read_state(const Misc* misc, uint32_t slot, void* pinned, hipStream_t st, PlocState& state,
                             uint32_t& err) {
    const size_t lo = offsetof(Misc, err), hi = offsetof(Misc, state) + (slot + 1u) * sizeof(PlocState);
    hipError_t e = hipMemcpyAsync(pinned, reinterpret_cast<const char*>(misc) + lo, hi - lo, hipMemcpyDeviceToHost,
                                  st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    const char* h = static_cast<const char*>(pinned);
    memcpy(&err, h, 4);
    memcpy(&state, h + (offsetof(Misc, state) + slot * sizeof(PlocState) - lo), sizeof(PlocState));
    return hipSuccess;
}

// The search-and-count kernel for the search distance (each allowed distance has its unrolled instance).
static void launch_search_count(uint32_t blocks, hipStream_t st, const Node32* cur, Misc* misc, uint32_t slot,
                                uint32_t r, uint32_t dist1, int8_t* merge, uint2* blocksums) {
    switch (r) {
        case 2: k_ploc_search_count<2><<<blocks, SC_THREADS, 0, st>>>(cur, misc, slot, r, dist1, merge, blocksums); break;
        case 6: k_ploc_search_count<6><<<blocks, SC_THREADS, 0, st>>>(cur, misc, slot, r, dist1, merge, blocksums); break;
        case 14: k_ploc_search_count<14><<<blocks, SC_THREADS, 0, st>>>(cur, misc, slot, r, dist1, merge, blocksums); break;
        case 24: k_ploc_search_count<24><<<blocks, SC_THREADS, 0, st>>>(cur, misc, slot, r, dist1, merge, blocksums); break;
        case 32: k_ploc_search_count<32><<<blocks, SC_THREADS, 0, st>>>(cur, misc, slot, r, dist1, merge, blocksums); break;
        default: k_ploc_search_count<0><<<blocks, SC_THREADS, 0, st>>>(cur, misc, slot, r, dist1, merge, blocksums); break;
    }
}

// The tail for the search distance (each allowed distance has its unrolled instance).
static void launch_tail(hipStream_t st, const Node32* cur, Misc* misc, uint32_t slot, uint32_t pass, uint32_t r,
                        uint32_t threshold, Node32* bvh_nodes, uint32_t* hist, const DpView& dp, uint32_t fused,
                        const uint32_t* slot_map, uint32_t vtop, uint32_t map_len) {
    const uint32_t one = r == 1 ? 1u : 0u;
#define OBVHS_TAIL(RR)                                                                                           \
    k_ploc_tail<RR><<<1, TAIL_THREADS, 0, st>>>(cur, misc, slot, pass, r, one, threshold, bvh_nodes, hist, dp, \
                                                fused, slot_map, vtop, map_len)
    switch (r) {
        case 2: OBVHS_TAIL(2); break;
        case 6: OBVHS_TAIL(6); break;
        case 14: OBVHS_TAIL(14); break;
        case 24: OBVHS_TAIL(24); break;
        case 32: OBVHS_TAIL(32); break;
        default: OBVHS_TAIL(0); break;
    }
#undef OBVHS_TAIL
}

static obvhs_status run_passes(obvhs_context* ctx, Node32* cur, Node32* nxt, uint32_t n, uint32_t r,
                               uint32_t search_depth_threshold, Node32* bvh_nodes, uint32_t* hist, DpView dp,
                               uint32_t fused, const uint32_t* slot_map, uint32_t vtop, uint32_t map_len,
                               PlocState& fin) {
    hipStream_t st = ctx->stream;
    BuildScratch& S = ctx->scratch;
    Misc* misc = S.misc.as<Misc>();
    int8_t* merge = S.merge.as<int8_t>();
    uint2* blocksums = S.blocksums.as<uint2>();
    uint32_t known = n;
    uint32_t pass = 0;
    uint32_t slot = 0;
    PlocState hstate{};
    while (known > TAIL_MAX) {
        const uint32_t blocks = (known + T256 - 1) / T256;
        for (int k = 0; k < PASSES_PER_SYNC; ++k) {
            const uint32_t dist1 = (r == 1 || pass < search_depth_threshold) ? 1u : 0u;
            launch_search_count(blocks, st, cur, misc, slot, r, dist1, merge, blocksums);
            if (blocks <= SCAN_FUSED_BLOCKS) {
                k_ploc_scan_compact<<<blocks, T256, 0, st>>>(cur, nxt, merge, misc, slot, blocksums, blocks, bvh_nodes,
                                                            dp, fused, slot_map, vtop, map_len, hist);
            } else {
                if (blocks <= SCAN_ONE_BLOCK) {
                    k_ploc_scan<<<1, 1024, 0, st>>>(blocksums, blocks, misc, slot, hist);
                } else {
                    // chunk totals after the counts, in the blocksums allocation's tail (dead during the passes)
                    uint2* chunk_tot = blocksums + blocks + 16;
                    const uint32_t nchunks = (blocks + SCAN_CHUNK - 1) / SCAN_CHUNK;
                    k_ploc_scan_chunks<<<nchunks, 1024, 0, st>>>(blocksums, blocks, chunk_tot);
                    k_ploc_scan_apply<<<nchunks, 1024, 0, st>>>(blocksums, blocks, chunk_tot, nchunks, misc, slot,
                                                                 hist);
                }
                k_ploc_compact<<<blocks, T256, 0, st>>>(cur, nxt, merge, misc, slot, blocksums, bvh_nodes, dp, fused,
                                                       slot_map, vtop, map_len);
            }
            std::swap(cur, nxt);
            slot ^= 1u;
            pass++;
        }
        OBVHS_TRY(hipGetLastError(), "ploc pass launch");
        uint32_t herr_loop = 0u;
        OBVHS_TRY(read_state(misc, slot, ctx->pinned, st, hstate, herr_loop), "read ploc state");
        if (herr_loop) return device_error_to_status(herr_loop);
        // Every pass merges at least one pair unless costs are NaN; a batch without progress would never end.
        if (hstate.count >= known) return OBVHS_ERR_PLOC_SELF_MERGE;
        known = hstate.count;
        // The device depth counts real passes only; later enqueued passes may have been no-ops.
        pass = hstate.depth;
    }
    launch_tail(st, cur, misc, slot, pass, r, search_depth_threshold, bvh_nodes, hist, dp, fused, slot_map, vtop,
                map_len);
    OBVHS_TRY(hipGetLastError(), "ploc tail launch");
    uint32_t herr = 0u;
    OBVHS_TRY(read_state(misc, slot, ctx->pinned, st, hstate, herr), "read final ploc state");
    if (herr) return device_error_to_status(herr);
    if (hstate.insert != 1u) return OBVHS_ERR_PLOC_SELF_MERGE;  // ploc/mod.rs:495-497
    fin = hstate;
    return OBVHS_OK;
}

obvhs_status ploc_build(obvhs_context* ctx, const obvhs_triangle* d_tris, const obvhs_aabb* d_aabbs,
                        const uint32_t* d_indices, uint32_t n, uint32_t search_distance, bool u128,
                        uint32_t search_depth_threshold, uint32_t fused_max_prims, obvhs_bvh2* bvh) {
    if (!(search_distance == 1 || search_distance == 2 || search_distance == 6 || search_distance == 14 ||
          search_distance == 24 || search_distance == 32))
        return OBVHS_ERR_INVALID_ARG;  // ploc/mod.rs:550-561
    hipStream_t st = ctx->stream;
    bvh->prim_count = n;
    bvh->node_count = n == 0 ? 0u : 2u * n - 1u;
    bvh->max_depth = 96;
    bvh->ploc_iterations = 0;
    bvh->cluster_sum = 0;
    bvh->dp_max_prims = 0;
    bvh->pass_ordered = true;
    OBVHS_TRY(bvh->prim_indices.ensure((size_t)n * 4 + 4), "alloc primitive_indices");
    if (n == 0) return OBVHS_OK;
    OBVHS_TRY(bvh->hist.ensure((size_t)(n + 1) * 4), "alloc pass history");
    DpView dp{};
    if (fused_max_prims) {
        OBVHS_TRY(bvh_dp_alloc(bvh), "alloc cost table");
        dp = bvh_dp_view(bvh, fused_max_prims);
    }
    const uint32_t fused = fused_max_prims ? 1u : 0u;
    OBVHS_TRY(bvh->nodes.ensure((size_t)bvh->node_count * sizeof(Node32)), "alloc bvh2 nodes");
    BuildScratch& S = ctx->scratch;
    OBVHS_TRY(S.leaves.ensure((size_t)n * sizeof(Node32)), "alloc leaves");
    OBVHS_TRY(S.clusters_a.ensure((size_t)n * sizeof(Node32)), "alloc clusters");
    OBVHS_TRY(S.clusters_b.ensure((size_t)n * sizeof(Node32)), "alloc clusters");
    OBVHS_TRY(S.keys.ensure((size_t)n * 8), "alloc keys");
    if (u128) OBVHS_TRY(S.keys_hi.ensure((size_t)n * 8), "alloc high key words");
    OBVHS_TRY(S.vals.ensure((size_t)n * 4), "alloc vals");
    OBVHS_TRY(S.merge.ensure((size_t)n + 64), "alloc merge");
    const uint32_t nblk = (n + T256 - 1) / T256;
    OBVHS_TRY(S.blocksums.ensure((size_t)nblk * sizeof(uint2) + 64 + (size_t)nblk * 6 * 4), "alloc blocksums");
    OBVHS_TRY(S.sort_scratch.ensure(radix_sort_scratch_bytes(n)), "alloc sort scratch");
    OBVHS_TRY(S.misc.ensure(sizeof(Misc)), "alloc misc");
    Misc* misc = S.misc.as<Misc>();
    Node32* leaves = S.leaves.as<Node32>();
    Node32* bvh_nodes = bvh->nodes.as<Node32>();
    uint64_t* keys = S.keys.as<uint64_t>();
    uint64_t* keys_hi = u128 ? S.keys_hi.as<uint64_t>() : nullptr;
    uint32_t* vals = S.vals.as<uint32_t>();
    uint32_t* hist = bvh->hist.as<uint32_t>();
    // per-block bound partials share the blocksums allocation (they are dead before the first pass)
    uint32_t* partials = reinterpret_cast<uint32_t*>(S.blocksums.as<char>() + (size_t)nblk * sizeof(uint2) + 64);

    // primitive_indices = indices (Bvh2::reset_for_reuse, bvh2/mod.rs:106-121)
    if (d_indices) {
        OBVHS_TRY(hipMemcpyAsync(bvh->prim_indices.ptr, d_indices, (size_t)n * 4, hipMemcpyDeviceToDevice, st),
                  "copy indices");
    }
    k_init_misc<<<1, 64, 0, st>>>(misc, n, hist);
    uint32_t* iota_out = d_indices ? nullptr : bvh->prim_indices.as<uint32_t>();  // primitive_indices = 0..n-1
    if (d_tris) {
        k_leaf_init_tris<<<nblk, T256, 0, st>>>(reinterpret_cast<const float4*>(d_tris), d_indices, n, leaves,
                                                misc, partials, iota_out);
    } else {
        k_leaf_init_aabbs<<<nblk, T256, 0, st>>>(reinterpret_cast<const float4*>(d_aabbs), d_indices, n, leaves,
                                                 misc, partials, iota_out);
    }
    k_reduce_bounds<<<1, 1024, 0, st>>>(partials, nblk, misc);
    k_morton<<<nblk, T256, 0, st>>>(leaves, n, misc, keys, keys_hi);
    OBVHS_TRY(hipGetLastError(), "leaf/morton launch");
    bvh->code_words = u128 ? 2u : 1u;
    if (ctx->keep_trace) {  // low words, then the high words of U128 codes
        OBVHS_TRY(bvh->codes.ensure((size_t)n * 8 * bvh->code_words), "alloc codes");
        OBVHS_TRY(hipMemcpyAsync(bvh->codes.ptr, keys, (size_t)n * 8, hipMemcpyDeviceToDevice, st), "copy codes");
        if (u128)
            OBVHS_TRY(hipMemcpyAsync(bvh->codes.as<uint64_t>() + n, keys_hi, (size_t)n * 8, hipMemcpyDeviceToDevice,
                                     st),
                      "copy codes");
    }
    obvhs_status ss = sort_codes(ctx, keys, keys_hi, vals, n);
    if (ss != OBVHS_OK) return ss;
    if (ctx->keep_trace) {
        OBVHS_TRY(bvh->sorted_index.ensure((size_t)n * 4), "alloc sorted index");
        OBVHS_TRY(hipMemcpyAsync(bvh->sorted_index.ptr, vals, (size_t)n * 4, hipMemcpyDeviceToDevice, st),
                  "copy sorted index");
    }
    Node32* cur = S.clusters_a.as<Node32>();
    Node32* nxt = S.clusters_b.as<Node32>();
    k_gather<<<nblk, T256, 0, st>>>(leaves, vals, n, cur);
    OBVHS_TRY(hipGetLastError(), "gather launch");

    PlocState fin{};
    obvhs_status ps = run_passes(ctx, cur, nxt, n, search_distance, search_depth_threshold, bvh_nodes, hist, dp,
                                 fused, nullptr, 0u, 0u, fin);
    if (ps != OBVHS_OK) return ps;
    bvh->ploc_iterations = fin.depth;
    bvh->cluster_sum = fin.cluster_sum;
    bvh->max_depth = fin.depth + 1 > 96 ? fin.depth + 1 : 96;  // ploc/mod.rs:501
    bvh->dp_max_prims = fused_max_prims;
    return OBVHS_OK;
}

// rebuild_from_leaves (ploc/rebuild.rs:136-183) over the nodes at d_src_slots[0..k) of bvh->nodes, listed in slot
// order (the order that breaks Morton-code ties).  slot_map = NULL: full rebuild, the tree becomes 2k-1 nodes with
// the plain placement (children after parents).  Otherwise slot_map[g] (k-1 entries, descending) is the left slot
// of the g-th freed sibling pair; every other slot keeps its node.  The scene box is the root's stored box.
obvhs_status ploc_rebuild(obvhs_context* ctx, obvhs_bvh2* bvh, const uint32_t* d_src_slots, uint32_t k,
                          const uint32_t* d_slot_map, uint32_t search_distance, bool u128,
                          uint32_t search_depth_threshold) {
    if (!(search_distance == 1 || search_distance == 2 || search_distance == 6 || search_distance == 14 ||
          search_distance == 24 || search_distance == 32))
        return OBVHS_ERR_INVALID_ARG;
    if (k == 0) return OBVHS_OK;
    if (!d_slot_map && 2ull * k - 1ull > bvh->node_count) return OBVHS_ERR_INVALID_ARG;
    hipStream_t st = ctx->stream;
    BuildScratch& S = ctx->scratch;
    OBVHS_TRY(bvh->hist.ensure((size_t)(k + 1) * 4), "alloc pass history");
    OBVHS_TRY(S.leaves.ensure((size_t)k * sizeof(Node32)), "alloc leaves");
    OBVHS_TRY(S.clusters_a.ensure((size_t)k * sizeof(Node32)), "alloc clusters");
    OBVHS_TRY(S.clusters_b.ensure((size_t)k * sizeof(Node32)), "alloc clusters");
    OBVHS_TRY(S.keys.ensure((size_t)k * 8), "alloc keys");
    if (u128) OBVHS_TRY(S.keys_hi.ensure((size_t)k * 8), "alloc high key words");
    OBVHS_TRY(S.vals.ensure((size_t)k * 4), "alloc vals");
    OBVHS_TRY(S.merge.ensure((size_t)k + 64), "alloc merge");
    const uint32_t nblk = (k + T256 - 1) / T256;
    OBVHS_TRY(S.blocksums.ensure((size_t)nblk * sizeof(uint2) + 64 + (size_t)nblk * 6 * 4), "alloc blocksums");
    OBVHS_TRY(S.sort_scratch.ensure(radix_sort_scratch_bytes(k)), "alloc sort scratch");
    OBVHS_TRY(S.misc.ensure(sizeof(Misc)), "alloc misc");
    Misc* misc = S.misc.as<Misc>();
    Node32* leaves = S.leaves.as<Node32>();
    Node32* bvh_nodes = bvh->nodes.as<Node32>();
    uint64_t* keys = S.keys.as<uint64_t>();
    uint64_t* keys_hi = u128 ? S.keys_hi.as<uint64_t>() : nullptr;
    uint32_t* vals = S.vals.as<uint32_t>();
    uint32_t* hist = bvh->hist.as<uint32_t>();
    // The clusters are copied out first: merges may overwrite the freed slots they came from.
    k_gather<<<nblk, T256, 0, st>>>(bvh_nodes, d_src_slots, k, leaves);
    k_init_misc<<<1, 64, 0, st>>>(misc, k, hist);
    k_bounds_from_node<<<1, 64, 0, st>>>(misc, bvh_nodes);
    k_morton<<<nblk, T256, 0, st>>>(leaves, k, misc, keys, keys_hi);
    OBVHS_TRY(hipGetLastError(), "rebuild morton launch");
    obvhs_status ss = sort_codes(ctx, keys, keys_hi, vals, k);
    if (ss != OBVHS_OK) return ss;
    Node32* cur = S.clusters_a.as<Node32>();
    Node32* nxt = S.clusters_b.as<Node32>();
    k_gather<<<nblk, T256, 0, st>>>(leaves, vals, k, cur);
    OBVHS_TRY(hipGetLastError(), "gather launch");
    PlocState fin{};
    obvhs_status ps = run_passes(ctx, cur, nxt, k, search_distance, search_depth_threshold, bvh_nodes, hist,
                                 DpView{}, 0u, d_slot_map, 2u * k - 3u, k - 1u, fin);
    if (ps != OBVHS_OK) return ps;
    if (!d_slot_map) bvh->node_count = 2u * k - 1u;  // ploc/mod.rs:283-284 (nodes.resize)
    bvh->pass_ordered = d_slot_map == nullptr;    // ploc/mod.rs:502
    bvh->ploc_iterations = fin.depth;
    bvh->cluster_sum = fin.cluster_sum;
    bvh->max_depth = fin.depth + 1 > 96 ? fin.depth + 1 : 96;  // ploc/mod.rs:501
    bvh->dp_max_prims = 0;                                      // boxes changed: any cost table is stale
    return OBVHS_OK;
}

namespace {
__global__ void k_iota(uint32_t* p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}
}  // namespace

hipError_t launch_iota(uint32_t* p, uint32_t n, hipStream_t st) {
    k_iota<<<(n + 255) / 256, 256, 0, st>>>(p, n);
    return hipGetLastError();
}

}  // namespace obvhs
