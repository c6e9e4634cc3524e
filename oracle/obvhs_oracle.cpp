// obvhs CPU oracle: a C++ restatement of the obvhs v0.3.1 hot path
//   PLOC BVH2 construction -> BVH2-to-CWBVH conversion -> CWBVH closest-hit / occlusion traversal.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is compared against.  Only the test
// suite, the smoke entry point and the CPU-baseline leg of the benchmark may load it; the product library never
// links or calls it.
//
// Every function cites the reference file:line it restates (paths relative to the obvhs crate root).
// Arithmetic follows the reference bit for bit:
//   * glam's SSE Vec3A min/max are `_mm_min_ps(a,b) = a<b?a:b` and `_mm_max_ps(a,b) = a>b?a:b`;
//   * no FMA contraction anywhere (built with -ffp-contract=off), IEEE division, denormals preserved;
//   * dot = (x*x' + y*y') + z*z', cross is the standard unfused form;
//   * CWBVH exponent goes through the host libm's log2f/exp2f exactly like Rust's f32::log2/exp2 on linux-gnu.
//
// Parity status of each function is listed in the oracle header comment of tests/test_oracle.py and in the design
// notes.  The sort tie order (reference: std unstable sorts) is fixed here to "stable by original index".

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct V3 {
    float x, y, z, w;
};
struct Aabb {
    V3 min, max;
};
struct Triangle {
    V3 v0, v1, v2;
};
struct Ray {
    V3 origin, direction, inv_direction;
    float tmin, tmax;
    float pad0, pad1;
};
struct RayHit {
    uint32_t primitive_id, geometry_id, instance_id;
    float t;
};
struct Bvh2Node {
    Aabb aabb;
    uint32_t prim_count, first_index, meta1, meta2;
};
struct CwBvhNode {
    float p[3];
    uint8_t e[3];
    uint8_t imask;
    uint32_t child_base_idx;
    uint32_t primitive_base_idx;
    uint8_t child_meta[8];
    uint8_t child_min_x[8], child_max_x[8], child_min_y[8], child_max_y[8], child_min_z[8], child_max_z[8];
};
static_assert(sizeof(V3) == 16, "Vec3A is 16 bytes");
static_assert(sizeof(Aabb) == 32, "Aabb is 32 bytes");
static_assert(sizeof(Triangle) == 48, "Triangle is 48 bytes");
static_assert(sizeof(Ray) == 64, "Ray is 64 bytes");
static_assert(sizeof(RayHit) == 16, "RayHit is 16 bytes");
static_assert(sizeof(Bvh2Node) == 48, "Bvh2Node is 48 bytes");
static_assert(sizeof(CwBvhNode) == 80, "CwBvhNode is 80 bytes");

// Error codes shared with the product library's status enum (include/obvhs_hip.h).
enum {
    ORC_OK = 0,
    ORC_ERR_INVALID_ARG = 1,
    ORC_ERR_NAN_INPUT = 2,        // ploc/mod.rs:191-192 debug_assert, :451,:466 out-of-bounds panic
    ORC_ERR_PLOC_SELF_MERGE = 3,  // ploc/mod.rs:464-467 insert_index underflow panic
    ORC_ERR_CW_DISTRIBUTE = 4,    // bvh2_to_cwbvh.rs:365-366 debug_assert distribute < 7
    ORC_ERR_CW_NONUNIT_LEAF = 5,  // bvh2_to_cwbvh.rs:243-247 panic
    ORC_ERR_CW_ORDER = 6,         // bvh2_to_cwbvh.rs:463 child left unassigned -> out-of-bounds panic
    ORC_ERR_CW_PRIMCOUNT = 7,     // bvh2_to_cwbvh.rs:155 panic
    ORC_ERR_CAPACITY = 8,
    ORC_ERR_VALIDATE = 9,
};

constexpr uint32_t INVALID32 = 0xffffffffu;

// Planted-mistake switch for the tests that show the checks catch plausible restatement errors.  0 = faithful.
enum {
    MUT_NONE = 0,
    MUT_SEARCH_STRICT = 1,     // `<` instead of `<=` in the neighbour search (first minimum wins)
    MUT_MORTON_SWAP_XY = 2,    // x and y interleaved in the wrong bit positions
    MUT_SORT_TIES_DESC = 3,    // equal codes ordered by descending index
    MUT_TRI_GE_ZERO = 4,       // barycentric test `>= 0` instead of the sign-bit test
    MUT_NODE_TMIN_ZERO = 5,    // node test clamps tmin to 0 instead of 1e-4
    MUT_PRIM_LOWBIT = 6,       // primitives of a leaf group tested lowest bit first
    MUT_CW_MAX_FLOOR = 7,      // child max quantised with floor instead of ceil
    MUT_HALF_AREA = 8,         // (dx + dy) * dz + dx * dx
    MUT_PLOC_SWAP_CHILDREN = 9,  // left/right children of a merge swapped
    MUT_PLOC_SLOT_ORDER = 10,  // merges of one pass take child slots in ascending instead of descending order
    MUT_SPLIT_UNFUSED = 11,    // split crossing point with a separate multiply and add instead of mul_add
    MUT_REINS_PUSH_ORDER = 12,  // reinsertion search pushes the right child before the left one
    MUT_REINS_NO_CONFLICT = 13,  // reinsertions applied without the conflict check
    MUT_REFIT_FULL = 14,       // refit_from_fast never stops early
    MUT_REBUILD_SLOT_PARITY = 15,  // partial rebuild checks the right (even) slot of a pair for being free
    MUT_CW_DIST_LE = 16,       // `<=` instead of `<` in the conversion's distribute searches
    MUT_CW_LEAF_LE = 17,       // `cost_leaf <= cost_internal` (a tie makes a leaf)
    MUT_CW_LUT_FLIP_X = 18,    // direction LUT with the x sign flipped
    MUT_CW_SLOT_LE = 19,       // `<=` instead of `<` in the greedy slot assignment (the last minimum wins)
    MUT_CW_SLOT_MAJOR = 20,    // greedy slot assignment scanned slot-major (slot outer, child inner)
    MUT_WT_MIN_AXIS = 21,      // watertight test shears onto the smallest direction axis instead of the largest
    MUT_WT_EDGE_STRICT = 22,   // watertight test counts an edge function of exactly 0 as outside
    MUT_WT_NODE_PLAIN = 23,    // watertight mode traverses with the crate's node test, not the widened one
    MUT_NODE_IEEE = 24,        // node test with IEEE min/max (a NaN slab value dropped) instead of SSE min/max
};
int g_mut = MUT_NONE;

// Threads of the multithreaded build (the CPU baseline's counterpart of the crate's `parallel` feature); 1 runs the
// serial restatement.  Every parallel section below mirrors a rayon section of the crate, with its size threshold,
// and produces the same result as the serial code.
int g_threads = 1;

// Runs fn(begin, end) over [0, n) split into `parts` contiguous chunks, one thread each.
template <class F>
void parallel_chunks(size_t n, int parts, F&& fn) {
    if (parts <= 1 || n < 2) {
        fn((size_t)0, n);
        return;
    }
    const size_t chunk = (n + (size_t)parts - 1) / (size_t)parts;
    std::vector<std::thread> pool;
    for (int t = 0; t < parts; ++t) {
        const size_t b = std::min(n, (size_t)t * chunk), e = std::min(n, b + chunk);
        if (b < e) pool.emplace_back([&fn, b, e] { fn(b, e); });
    }
    for (auto& th : pool) th.join();
}

// Sorts [first, last) with `less` on g_threads threads: chunks sorted in parallel, then merged pairwise in parallel
// rounds.  `less` is a strict total order here (ties broken by a unique key), so the result equals std::sort's.
template <class T, class L>
void parallel_sort(std::vector<T>& v, L less) {
    const size_t n = v.size();
    const int parts = g_threads;
    if (parts <= 1 || n < 4096) {
        std::sort(v.begin(), v.end(), less);
        return;
    }
    const size_t chunk = (n + (size_t)parts - 1) / (size_t)parts;
    std::vector<size_t> bounds;
    for (size_t b = 0; b < n; b += chunk) bounds.push_back(b);
    bounds.push_back(n);
    {
        std::vector<std::thread> pool;
        for (size_t k = 0; k + 1 < bounds.size(); ++k)
            pool.emplace_back([&, k] { std::sort(v.begin() + bounds[k], v.begin() + bounds[k + 1], less); });
        for (auto& th : pool) th.join();
    }
    std::vector<T> tmp(n);
    while (bounds.size() > 2) {
        std::vector<size_t> nb;
        std::vector<std::thread> pool;
        for (size_t k = 0; k + 1 < bounds.size(); k += 2) {
            const size_t a = bounds[k], m = bounds[k + 1], e = k + 2 < bounds.size() ? bounds[k + 2] : m;
            nb.push_back(a);
            pool.emplace_back([&, a, m, e] {
                std::merge(v.begin() + a, v.begin() + m, v.begin() + m, v.begin() + e, tmp.begin() + a, less);
            });
        }
        nb.push_back(n);
        for (auto& th : pool) th.join();
        std::swap(v, tmp);
        bounds.swap(nb);
    }
}

// glam Vec3A::min / max on x86 SSE (aabb.rs:84-89 via glam).
inline float smin(float a, float b) { return a < b ? a : b; }
inline float smax(float a, float b) { return a > b ? a : b; }

inline V3 v3(float x, float y, float z) { return V3{x, y, z, 0.0f}; }
inline V3 vsub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 vadd(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 vmin(V3 a, V3 b) { return v3(smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)); }
inline V3 vmax(V3 a, V3 b) { return v3(smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)); }
// glam dot3 on SSE: (x*x' + y*y') + z*z'
inline float vdot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// glam cross: (y*z' - z*y', z*x' - x*z', x*y' - y*x')
inline V3 vcross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

// aabb.rs:84-89
inline Aabb aabb_union(const Aabb& a, const Aabb& b) { return Aabb{vmin(a.min, b.min), vmax(a.max, b.max)}; }
// aabb.rs:151-154
inline float half_area(const Aabb& a) {
    float dx = a.max.x - a.min.x, dy = a.max.y - a.min.y, dz = a.max.z - a.min.z;
    if (g_mut == MUT_HALF_AREA) return (dx + dy) * dz + dx * dx;
    return (dx + dy) * dz + dx * dy;
}
// aabb.rs:165-170
inline Aabb aabb_empty() {
    const float M = 3.40282347e+38f;
    return Aabb{v3(M, M, M), v3(-M, -M, -M)};
}
// aabb.rs:76-79 extend(point) = union(from_point(point))
inline void aabb_extend(Aabb& a, V3 p) {
    a.min = vmin(a.min, p);
    a.max = vmax(a.max, p);
}
// triangle.rs:28-30 + aabb.rs:59-66 (from_points: start at v0, extend v1, v2)
inline Aabb tri_aabb(const Triangle& t) {
    Aabb a{t.v0, t.v0};
    aabb_extend(a, t.v1);
    aabb_extend(a, t.v2);
    a.min.w = 0.0f;
    a.max.w = 0.0f;
    return a;
}
inline bool aabb_eq(const Aabb& a, const Aabb& b) {
    return a.min.x == b.min.x && a.min.y == b.min.y && a.min.z == b.min.z && a.max.x == b.max.x &&
           a.max.y == b.max.y && a.max.z == b.max.z;
}
inline bool has_nan(const Aabb& a) {
    return std::isnan(a.min.x) || std::isnan(a.min.y) || std::isnan(a.min.z) || std::isnan(a.max.x) ||
           std::isnan(a.max.y) || std::isnan(a.max.z);
}

// ---------------------------------------------------------------------------------------------------------------
// Morton (ploc/morton.rs)
// ---------------------------------------------------------------------------------------------------------------

// morton.rs:35-43
inline uint64_t split_by_3_u64(uint32_t a) {
    uint64_t x = (uint64_t)a & 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}
// morton.rs:48-50
inline uint64_t morton_encode_u64(uint32_t x, uint32_t y, uint32_t z) {
    if (g_mut == MUT_MORTON_SWAP_XY) return split_by_3_u64(y) | split_by_3_u64(x) << 1 | split_by_3_u64(z) << 2;
    return split_by_3_u64(x) | split_by_3_u64(y) << 1 | split_by_3_u64(z) << 2;
}
// Rust `f64 as u32`: saturating, NaN -> 0.
inline uint32_t f64_as_u32(double v) {
    if (!(v >= 1.0)) return 0u;
    if (v >= 4294967296.0) return 0xffffffffu;
    return (uint32_t)v;
}
// morton.rs:55-58
inline uint64_t morton_encode_u64_unorm(double px, double py, double pz) {
    const double s = (double)(1u << 21);
    return morton_encode_u64(f64_as_u32(px * s), f64_as_u32(py * s), f64_as_u32(pz * s));
}

typedef unsigned __int128 uint128;
inline uint128 u128_from(uint64_t hi, uint64_t lo) { return (uint128)hi << 64 | (uint128)lo; }
// morton.rs:66-75 (42 bits per axis; each mask is the reference's 128-bit constant as high and low words)
inline uint128 split_by_3_u128(uint64_t a) {
    uint128 x = (uint128)a & 0x3ffffffffffull;
    x = (x | x << 64) & u128_from(0x3ff00000000ull, 0xffffffffull);
    x = (x | x << 32) & u128_from(0x3ff00000000ull, 0xffff00000000ffffull);
    x = (x | x << 16) & u128_from(0x30000ff0000ff00ull, 0xff0000ff0000ffull);
    x = (x | x << 8) & u128_from(0x300f00f00f00f00ull, 0xf00f00f00f00f00full);
    x = (x | x << 4) & u128_from(0x30c30c30c30c30cull, 0x30c30c30c30c30c3ull);
    x = (x | x << 2) & u128_from(0x924924924924924ull, 0x9249249249249249ull);
    return x;
}
// morton.rs:79-81
inline uint128 morton_encode_u128(uint64_t x, uint64_t y, uint64_t z) {
    return split_by_3_u128(x) | split_by_3_u128(y) << 1 | split_by_3_u128(z) << 2;
}
// Rust `f64 as u64`: saturating, NaN -> 0.
inline uint64_t f64_as_u64(double v) {
    if (!(v >= 1.0)) return 0u;
    if (v >= 18446744073709551616.0) return ~0ull;
    return (uint64_t)v;
}
// morton.rs:85-88
inline uint128 morton_encode_u128_unorm(double px, double py, double pz) {
    const double s = (double)(1ull << 42);
    return morton_encode_u128(f64_as_u64(px * s), f64_as_u64(py * s), f64_as_u64(pz * s));
}

// ---------------------------------------------------------------------------------------------------------------
// PLOC (ploc/mod.rs)
// ---------------------------------------------------------------------------------------------------------------

// ploc/mod.rs:508-528 (reference brute-force neighbour search)
int8_t find_best_node_basic(size_t index, const Bvh2Node* nodes, size_t count, size_t sd) {
    size_t best_node = index;
    float best_cost = INFINITY;
    size_t begin = index - std::min(sd, index);
    size_t end = std::min(index + sd + 1, count);
    const Aabb& our = nodes[index].aabb;
    for (size_t other = begin; other < end; ++other) {
        if (other == index) continue;
        float cost = half_area(aabb_union(our, nodes[other].aabb));
        if (g_mut == MUT_SEARCH_STRICT ? cost < best_cost : cost <= best_cost) {
            best_node = other;
            best_cost = cost;
        }
    }
    return (int8_t)((int64_t)best_node - (int64_t)index);
}

// ploc/mod.rs:567-651 (SearchCache, the default non-parallel search path)
struct SearchCache {
    size_t sd;
    std::vector<float> c;  // sd x sd, zero initialised (ploc/mod.rs:569-573)
    explicit SearchCache(size_t s) : sd(s), c(s * s, 0.0f) {}
    float back(size_t index, size_t other) const { return c[(other % sd) * sd + (index % sd)]; }
    float& front(size_t index, size_t other) { return c[(index % sd) * sd + (other % sd)]; }
    // ploc/mod.rs:624-650
    int8_t find_best_node(size_t index, const Bvh2Node* nodes, size_t count) {
        size_t best_node = index;
        float best_cost = INFINITY;
        size_t begin = index - std::min(sd, index);
        size_t end = std::min(index + sd + 1, count);
        for (size_t other = begin; other < index; ++other) {
            float area = back(index, other);
            if (g_mut == MUT_SEARCH_STRICT ? area < best_cost : area <= best_cost) {
                best_node = other;
                best_cost = area;
            }
        }
        const Aabb& our = nodes[index].aabb;
        for (size_t other = index + 1; other < end; ++other) {
            float cost = half_area(aabb_union(our, nodes[other].aabb));
            front(index, other) = cost;
            if (g_mut == MUT_SEARCH_STRICT ? cost < best_cost : cost <= best_cost) {
                best_node = other;
                best_cost = cost;
            }
        }
        return (int8_t)((int64_t)best_node - (int64_t)index);
    }
};

// find_best_node_parallel for the first search-distance clusters of a chunk: the areas towards lower indices are
// computed (no earlier cluster of this chunk cached them), the ones towards higher indices cached as
// find_best_node caches them; the same `<=` rule in the same order.
int8_t find_best_node_basic_fill(size_t index, const Bvh2Node* nodes, size_t count, SearchCache& cache) {
    const size_t sd = cache.sd;
    size_t best_node = index;
    float best_cost = INFINITY;
    size_t begin = index - std::min(sd, index);
    size_t end = std::min(index + sd + 1, count);
    const Aabb& our = nodes[index].aabb;
    for (size_t other = begin; other < end; ++other) {
        if (other == index) continue;
        float cost = half_area(aabb_union(our, nodes[other].aabb));
        if (other > index) cache.front(index, other) = cost;
        if (cost <= best_cost) {
            best_node = other;
            best_cost = cost;
        }
    }
    return (int8_t)((int64_t)best_node - (int64_t)index);
}

// Morton64 (ploc/mod.rs:705-709) and Morton128 (ploc/mod.rs:686-692) as one record: U64 keys leave `hi` zero.
struct MortonKey {
    uint64_t hi;
    uint64_t lo;
    uint64_t index;
};

struct PlocTrace {
    uint64_t* codes;           // [n] Morton codes (the low 64 bits for U128) in original leaf order (optional)
    uint64_t* codes_hi;        // [n] high 64 bits of the U128 codes (optional; untouched for U64)
    uint32_t* sorted_index;    // [n] sorted position -> original leaf index (optional)
    uint32_t* iter_counts;     // [cap] cluster count at the start of each iteration (optional)
    uint32_t iter_cap;
    uint32_t num_iters;
    uint32_t depth;
    uint32_t max_offset_abs;   // largest |merge offset| seen (diagnostic)
};

constexpr uint32_t NODE_INVALID_BIT = 0x80000000u;  // Bvh2Node::set_invalid (bvh2/node.rs:137-151)
inline bool node_valid(const Bvh2Node& n) { return (n.prim_count & NODE_INVALID_BIT) == 0u; }

// ploc/mod.rs:258-503 (build_ploc_from_leaves) + 771-847 (sort) over the clusters in `current`.
//   tie      : sort key that orders equal Morton codes (the reference's sorts are unstable there); null = position
//              in `current`.
//   rebuild  : the REBUILD = true placement (ploc/mod.rs:449-462): every merge scans down from the array end for
//              the next sibling pair whose left slot is invalid.  Otherwise children fill nodes[2n-3 ..] downwards
//              and the array is exactly 2n-1 long.
//   u128     : SortPrecision::U128 (126-bit keys, morton_encode_u128_unorm) instead of U64.
// `nodes` has nodes_len entries; out_depth receives the pass count.
int ploc_from_leaves(std::vector<Bvh2Node>& current, const uint32_t* tie, const Aabb& total, size_t sd,
                     size_t search_depth_threshold, bool basic_search, bool rebuild, bool u128, Bvh2Node* out_nodes,
                     size_t nodes_len, PlocTrace* tr, size_t* out_depth) {
    const size_t n = current.size();
    if (out_depth) *out_depth = 0;
    if (n == 0) return ORC_OK;  // ploc/mod.rs:268-270
    std::vector<Bvh2Node> next(n);
    const size_t nodes_count = 2 * n - 1;  // ploc/mod.rs:274
    size_t insert_index;
    if (rebuild) {  // ploc/mod.rs:276-282
        if (nodes_len == 0) return ORC_OK;
        if (nodes_len < nodes_count) return ORC_ERR_CAPACITY;
        insert_index = nodes_len - 1;
    } else {
        if (nodes_len != nodes_count) return ORC_ERR_CAPACITY;
        insert_index = nodes_count;  // ploc/mod.rs:283-284
    }

    // ploc/mod.rs:287-288: scale = 1 / diagonal (f64), offset = -min * scale
    double diag[3] = {(double)(total.max.x - total.min.x), (double)(total.max.y - total.min.y),
                      (double)(total.max.z - total.min.z)};
    double scale[3] = {1.0 / diag[0], 1.0 / diag[1], 1.0 / diag[2]};
    double offset[3] = {-(double)total.min.x * scale[0], -(double)total.min.y * scale[1],
                        -(double)total.min.z * scale[2]};

    // ploc/mod.rs:782-809 gen_mort
    std::vector<MortonKey> mortons(n);
    parallel_chunks(n, n > 100000 ? g_threads : 1, [&](size_t i0, size_t i1) {
    for (size_t i = i0; i < i1; ++i) {
        const Aabb& a = current[i].aabb;
        float cx = (a.max.x + a.min.x) * 0.5f, cy = (a.max.y + a.min.y) * 0.5f, cz = (a.max.z + a.min.z) * 0.5f;
        double px = (double)cx * scale[0] + offset[0];
        double py = (double)cy * scale[1] + offset[1];
        double pz = (double)cz * scale[2] + offset[2];
        if (u128) {  // Morton128::new (ploc/mod.rs:727-733)
            const uint128 c = morton_encode_u128_unorm(px, py, pz);
            mortons[i] = MortonKey{(uint64_t)(c >> 64), (uint64_t)c, (uint64_t)i};
            if (tr && tr->codes_hi) tr->codes_hi[i] = mortons[i].hi;
        } else {     // Morton64::new (ploc/mod.rs:749-755)
            mortons[i] = MortonKey{0u, morton_encode_u64_unorm(px, py, pz), (uint64_t)i};
        }
        if (tr && tr->codes) tr->codes[i] = mortons[i].lo;
    }
    });
    // ploc/mod.rs:811-827. The reference uses unstable sorts; ties are ordered by the tie key here (the original
    // index for a build, the node slot for a partial rebuild).  Multithreaded: par_sort_unstable_by_key
    // (ploc/mod.rs:814-815) as a parallel merge sort, with the same total order.
    const bool ties_desc = g_mut == MUT_SORT_TIES_DESC;
    parallel_sort(mortons, [ties_desc, tie](const MortonKey& a, const MortonKey& b) {
        const uint64_t ta = tie ? tie[a.index] : a.index, tb = tie ? tie[b.index] : b.index;
        if (a.hi != b.hi) return a.hi < b.hi;
        return a.lo < b.lo || (a.lo == b.lo && (ties_desc ? ta > tb : ta < tb));
    });
    // ploc/mod.rs:829-846 remap, then swap (ploc/mod.rs:305)
    for (size_t k = 0; k < n; ++k) {
        next[k] = current[mortons[k].index];
        if (tr && tr->sorted_index) tr->sorted_index[k] = (uint32_t)mortons[k].index;
    }
    std::swap(current, next);

    std::vector<int8_t> merge(n);
    SearchCache cache(sd);  // ploc/mod.rs:311-312
    size_t depth = 0;
    size_t next_nodes_idx = 0;
    size_t count = n;
    while (count > 1) {  // ploc/mod.rs:327
        if (tr && tr->iter_counts && tr->num_iters < tr->iter_cap) tr->iter_counts[tr->num_iters] = (uint32_t)count;
        if (tr) tr->num_iters++;
        if (sd == 1 || depth < search_depth_threshold) {
            // ploc/mod.rs:329-382 distance-1 rule
            // (multithreaded above 300k clusters, ploc/mod.rs:340-375: each chunk starts from its left
            // neighbour's cost)
            parallel_chunks(count - 1, count >= 300000 ? g_threads : 1, [&](size_t i0, size_t i1) {
                float last_cost = i0 == 0 ? INFINITY
                                          : half_area(aabb_union(current[i0 - 1].aabb, current[i0].aabb));
                for (size_t i = i0; i < i1; ++i) {
                    float cost = half_area(aabb_union(current[i].aabb, current[i + 1].aabb));
                    merge[i] = last_cost < cost ? -1 : 1;
                    last_cost = cost;
                }
            });
            merge[count - 1] = -1;
        } else if (g_threads > 1 && count >= 4000 && !basic_search) {
            // ploc/mod.rs:389-417: chunks in parallel, each with its own cache; the first sd clusters of a chunk
            // compute the areas the cache would have held (find_best_node_parallel), so results are the serial ones
            parallel_chunks(count, g_threads, [&](size_t i0, size_t i1) {
                SearchCache local(sd);
                for (size_t i = i0; i < i1; ++i)
                    merge[i] = i - i0 <= sd ? find_best_node_basic_fill(i, current.data(), count, local)
                                           : local.find_best_node(i, current.data(), count);
            });
        } else {
            // ploc/mod.rs:384-387
            for (size_t i = 0; i < count; ++i) {
                merge[i] = basic_search ? find_best_node_basic(i, current.data(), count, sd)
                                        : cache.find_best_node(i, current.data(), count);
            }
        }
        if (tr) {
            for (size_t i = 0; i < count; ++i) {
                uint32_t a = (uint32_t)std::abs((int)merge[i]);
                if (a > tr->max_offset_abs) tr->max_offset_abs = a;
            }
        }
        // ploc/mod.rs:420-487 merge / compaction scan
        size_t pass_merges = 0, pass_slot_top = insert_index;
        if (g_mut == MUT_PLOC_SLOT_ORDER) {
            for (size_t i = 0; i < count; ++i) {
                size_t b = (size_t)((int64_t)i + merge[i]);
                if ((int64_t)b + merge[b] == (int64_t)i && b < i) pass_merges++;
            }
        }
        size_t ascending_slot = pass_slot_top - 2 * pass_merges;
        size_t index = 0;
        while (index < count) {
            int64_t index_offset = merge[index];
            size_t best_index = (size_t)((int64_t)index + index_offset);
            if ((int64_t)best_index + (int64_t)merge[best_index] != (int64_t)index) {
                next[next_nodes_idx++] = current[index];
                index += 1;
                continue;
            }
            if (best_index > index) {
                index += 1;
                continue;
            }
            Bvh2Node left = current[index];
            Bvh2Node right = current[best_index];
            if (g_mut == MUT_PLOC_SWAP_CHILDREN) std::swap(left, right);
            size_t first_child;
            if (rebuild) {
                // ploc/mod.rs:449-462: the left slot of the next free sibling pair, scanning down.
                for (;;) {
                    if (insert_index < 2) return ORC_ERR_CAPACITY;  // reference panics out of bounds here
                    const size_t left_slot = g_mut == MUT_REBUILD_SLOT_PARITY ? insert_index : insert_index - 1;
                    if (!node_valid(out_nodes[left_slot])) {
                        first_child = left_slot;
                        insert_index -= 2;
                        break;
                    }
                    insert_index -= 2;
                }
            } else {
                if (insert_index < 2) return ORC_ERR_PLOC_SELF_MERGE;  // reference panics out of bounds here
                insert_index -= 2;
                first_child = insert_index;
                if (g_mut == MUT_PLOC_SLOT_ORDER) {
                    first_child = ascending_slot;
                    ascending_slot += 2;
                }
            }
            if (first_child + 1 >= nodes_len) return ORC_ERR_CAPACITY;
            out_nodes[first_child] = left;
            out_nodes[first_child + 1] = right;
            Aabb u = aabb_union(left.aabb, right.aabb);
            next[next_nodes_idx++] = Bvh2Node{u, 0u, (uint32_t)first_child, 0u, 0u};
            if (sd == 1 && index_offset == 1) {
                index += 2;  // unreachable in the reference as well (best_index > index was handled above)
            } else {
                index += 1;
            }
        }
        std::swap(next, current);
        count = next_nodes_idx;
        next_nodes_idx = 0;
        depth += 1;
    }
    if (!rebuild && insert_index != 1) return ORC_ERR_PLOC_SELF_MERGE;  // ploc/mod.rs:495-497 debug_assert
    out_nodes[0] = current[0];                                          // ploc/mod.rs:499
    if (tr) tr->depth = (uint32_t)depth;
    if (out_depth) *out_depth = depth;
    return ORC_OK;
}

// ploc/mod.rs:169-251 (build_ploc): leaves from the boxes, scene bounds, then build_ploc_from_leaves.
int ploc_build(const Aabb* aabbs, const uint32_t* indices, size_t n, size_t sd, size_t search_depth_threshold,
               bool basic_search, bool u128, Bvh2Node* out_nodes, uint32_t* out_prim_indices, PlocTrace* tr) {
    if (sd < 1 || sd > 32) return ORC_ERR_INVALID_ARG;
    for (size_t i = 0; i < n; ++i) out_prim_indices[i] = indices[i];  // reset_for_reuse (bvh2/mod.rs:106-121)
    if (tr) {
        tr->num_iters = 0;
        tr->depth = 0;
        tr->max_offset_abs = 0;
    }
    if (n == 0) return ORC_OK;

    std::vector<Bvh2Node> current(n);
    Aabb total = aabb_empty();
    // ploc/mod.rs:187-194, 233-243 (multithreaded above 500k leaves, ploc/mod.rs:199-231: per-chunk bounds, then
    // their union in chunk order)
    const int parts = n >= 500000 ? g_threads : 1;
    std::vector<Aabb> part_total((size_t)std::max(parts, 1), aabb_empty());
    std::vector<uint8_t> part_nan((size_t)std::max(parts, 1), 0);
    const size_t chunk = (n + (size_t)std::max(parts, 1) - 1) / (size_t)std::max(parts, 1);
    parallel_chunks(n, parts, [&](size_t i0, size_t i1) {
        const size_t k = i0 / chunk;
        Aabb t = aabb_empty();
        for (size_t i = i0; i < i1; ++i) {
            Aabb a = aabbs[i];
            a.min.w = a.max.w = 0.0f;
            if (has_nan(a)) part_nan[k] = 1;
            aabb_extend(t, a.min);
            aabb_extend(t, a.max);
            current[i] = Bvh2Node{a, 1u, indices[i], 0u, 0u};
        }
        part_total[k] = t;
    });
    for (size_t k = 0; k < part_total.size(); ++k) {
        if (part_nan[k]) return ORC_ERR_NAN_INPUT;
        aabb_extend(total, part_total[k].min);
        aabb_extend(total, part_total[k].max);
    }
    const size_t nodes_count = 2 * n - 1;
    for (size_t i = 0; i < nodes_count; ++i) out_nodes[i] = Bvh2Node{};
    return ploc_from_leaves(current, nullptr, total, sd, search_depth_threshold, basic_search, false, u128,
                            out_nodes, nodes_count, tr, nullptr);
}

// ---------------------------------------------------------------------------------------------------------------
// BVH2 -> CWBVH (cwbvh/bvh2_to_cwbvh.rs)
// ---------------------------------------------------------------------------------------------------------------

enum : uint8_t { KIND_LEAF = 0, KIND_INTERNAL = 1, KIND_DISTRIBUTE = 2 };  // bvh2_to_cwbvh.rs:470-476
struct Decision {                                                          // bvh2_to_cwbvh.rs:478-484
    float cost = 0.0f;
    uint8_t kind = KIND_DISTRIBUTE;
    uint8_t distribute_left = 0;
    uint8_t distribute_right = 0;
};
constexpr uint8_t INVALID8 = 0xff;
constexpr float PRIM_COST = 0.3f;          // bvh2_to_cwbvh.rs:30
constexpr float DENOM = 1.0f / 255.0f;     // cwbvh/mod.rs:36-38

struct Converter {
    const Bvh2Node* nodes;
    size_t nnodes;
    const uint32_t* bvh2_prim_indices;
    std::vector<CwBvhNode> out;
    std::vector<uint32_t> prims;
    std::vector<Decision> decisions;
    std::vector<uint32_t> cw_to_bvh2;  // which BVH2 node each CW node was emitted from (diagnostic)
    bool order_children;
    V3 lut[8];
    int err = ORC_OK;

    // bvh2_to_cwbvh.rs:220-344
    uint32_t cost_node(size_t node_index, uint32_t max_prims_per_leaf, uint32_t num_child_prims) {
        const Bvh2Node& node = nodes[node_index];
        float ha = half_area(node.aabb);
        uint32_t first_index = node.first_index;
        size_t dec = node_index * 7;
        size_t f7 = (size_t)first_index * 7, n7 = ((size_t)first_index + 1) * 7;
        if (node.prim_count != 0) {
            uint32_t np = node.prim_count;
            if (np != 1) {
                err = ORC_ERR_CW_NONUNIT_LEAF;
                return np;
            }
            float cost_leaf = ha * (float)np * PRIM_COST;
            for (int i = 0; i < 7; ++i) {
                decisions[dec + i].kind = KIND_LEAF;
                decisions[dec + i].cost = cost_leaf;
            }
            return np;
        }
        uint32_t num_primitives = num_child_prims;
        {
            float cost_leaf = num_primitives <= max_prims_per_leaf ? ((float)num_primitives) * ha * PRIM_COST
                                                                   : INFINITY;
            float cost_distribute = INFINITY;
            uint8_t dl = INVALID8, dr = INVALID8;
            for (int k = 0; k < 7; ++k) {
                float c = decisions[f7 + k].cost + decisions[n7 + 6 - k].cost;
                if (g_mut == MUT_CW_DIST_LE ? c <= cost_distribute : c < cost_distribute) {
                    cost_distribute = c;
                    dl = (uint8_t)k;
                    dr = (uint8_t)(6 - k);
                }
            }
            float cost_internal = cost_distribute + ha;
            Decision& d = decisions[dec];
            if (g_mut == MUT_CW_LEAF_LE ? cost_leaf <= cost_internal : cost_leaf < cost_internal) {
                d.kind = KIND_LEAF;
                d.cost = cost_leaf;
            } else {
                d.kind = KIND_INTERNAL;
                d.cost = cost_internal;
            }
            d.distribute_left = dl;
            d.distribute_right = dr;
        }
        for (int i = 1; i < 7; ++i) {
            size_t ni = dec + i;
            float cost_distribute = decisions[ni - 1].cost;
            uint8_t dl = INVALID8, dr = INVALID8;
            for (int k = 0; k < i; ++k) {
                float c = decisions[f7 + k].cost + decisions[n7 + i - k - 1].cost;
                if (g_mut == MUT_CW_DIST_LE ? c <= cost_distribute : c < cost_distribute) {
                    cost_distribute = c;
                    dl = (uint8_t)k;
                    dr = (uint8_t)(i - k - 1);
                }
            }
            Decision& d = decisions[ni];
            d.cost = cost_distribute;
            if (dl != INVALID8) {
                d.kind = KIND_DISTRIBUTE;
                d.distribute_left = dl;
                d.distribute_right = dr;
            } else {
                decisions[ni] = decisions[ni - 1];
            }
        }
        return num_primitives;
    }

    // Post-order evaluation of cost_node.  The reference recurses (bvh2_to_cwbvh.rs:258-266); the DP value of a
    // node only depends on its children, so an explicit post-order stack gives the same table without deep
    // recursion on degenerate trees.
    void calculate_cost(uint32_t max_prims_per_leaf) {
        std::vector<uint32_t> nprims(nnodes, 0);
        std::vector<std::pair<uint32_t, bool>> st;
        st.push_back({0u, false});
        while (!st.empty() && err == ORC_OK) {
            auto [ni, expanded] = st.back();
            st.pop_back();
            const Bvh2Node& nd = nodes[ni];
            if (nd.prim_count != 0) {
                nprims[ni] = cost_node(ni, max_prims_per_leaf, 0);
                continue;
            }
            if (!expanded) {
                st.push_back({ni, true});
                st.push_back({nd.first_index + 1, false});
                st.push_back({nd.first_index, false});
            } else {
                nprims[ni] =
                    cost_node(ni, max_prims_per_leaf, nprims[nd.first_index] + nprims[nd.first_index + 1]);
            }
        }
    }

    // bvh2_to_cwbvh.rs:346-397
    void get_children(size_t node_index, uint32_t* children, uint32_t& child_count, size_t i) {
        const Bvh2Node& node = nodes[node_index];
        if (node.prim_count != 0) {
            children[child_count++] = (uint32_t)node_index;
            return;
        }
        const Decision& d = decisions[node_index * 7 + i];
        uint8_t dl = d.distribute_left, dr = d.distribute_right;
        if (dl >= 7 || dr >= 7) {
            err = ORC_ERR_CW_DISTRIBUTE;
            return;
        }
        if (decisions[(size_t)node.first_index * 7 + dl].kind == KIND_DISTRIBUTE) {
            get_children(node.first_index, children, child_count, dl);
        } else {
            children[child_count++] = node.first_index;
        }
        if (err) return;
        if (decisions[((size_t)node.first_index + 1) * 7 + dr].kind == KIND_DISTRIBUTE) {
            get_children(node.first_index + 1, children, child_count, dr);
        } else {
            children[child_count++] = node.first_index + 1;
        }
    }

    // bvh2_to_cwbvh.rs:402-467
    void order(size_t node_index, uint32_t* children, uint32_t child_count) {
        const Aabb& na = nodes[node_index].aabb;
        V3 p = v3((na.max.x + na.min.x) * 0.5f, (na.max.y + na.min.y) * 0.5f, (na.max.z + na.min.z) * 0.5f);
        const float FMAX = 3.40282347e+38f;
        float cost[8][8];
        for (int c = 0; c < 8; ++c)
            for (int s = 0; s < 8; ++s) cost[c][s] = FMAX;
        for (int s = 0; s < 8; ++s) {
            for (uint32_t c = 0; c < child_count; ++c) {
                const Aabb& ca = nodes[children[c]].aabb;
                V3 cc = v3((ca.max.x + ca.min.x) * 0.5f, (ca.max.y + ca.min.y) * 0.5f, (ca.max.z + ca.min.z) * 0.5f);
                cost[c][s] = vdot(lut[s], vsub(cc, p));
            }
        }
        size_t assignment[8];
        bool slot_filled[8];
        for (int i = 0; i < 8; ++i) {
            assignment[i] = INVALID32;
            slot_filled[i] = false;
        }
        for (;;) {
            float min_cost = FMAX;
            size_t min_slot = INVALID32, min_index = INVALID32;
            if (g_mut == MUT_CW_SLOT_MAJOR) {
                for (int s = 0; s < 8; ++s) {
                    if (slot_filled[s]) continue;
                    for (uint32_t c = 0; c < child_count; ++c) {
                        if (assignment[c] == INVALID32 && cost[c][s] < min_cost) {
                            min_cost = cost[c][s];
                            min_slot = s;
                            min_index = c;
                        }
                    }
                }
            } else {
                for (uint32_t c = 0; c < child_count; ++c) {
                    if (assignment[c] == INVALID32) {
                        for (int s = 0; s < 8; ++s) {
                            float cst = cost[c][s];
                            if (!slot_filled[s] && (g_mut == MUT_CW_SLOT_LE ? cst <= min_cost : cst < min_cost)) {
                                min_cost = cst;
                                min_slot = s;
                                min_index = c;
                            }
                        }
                    }
                }
            }
            if (min_slot == INVALID32) break;
            slot_filled[min_slot] = true;
            assignment[min_index] = min_slot;
        }
        uint32_t orig[8];
        for (int i = 0; i < 8; ++i) {
            orig[i] = children[i];
            children[i] = INVALID32;
        }
        for (uint32_t i = 0; i < child_count; ++i) {
            if (assignment[i] == INVALID32) {
                err = ORC_ERR_CW_ORDER;
                return;
            }
            children[assignment[i]] = orig[i];
        }
    }

    // bvh2_to_cwbvh.rs:197-211
    uint32_t count_primitives(size_t node_index) {
        const Bvh2Node& node = nodes[node_index];
        if (node.prim_count != 0) {
            prims.push_back(bvh2_prim_indices[node.first_index]);
            return node.prim_count;
        }
        return count_primitives(node.first_index) + count_primitives((size_t)node.first_index + 1);
    }

    // bvh2_to_cwbvh.rs:75-193
    void convert(size_t node_index_bvh8, size_t node_index_bvh2) {
        CwBvhNode node = out[node_index_bvh8];
        const Aabb& aabb = nodes[node_index_bvh2].aabb;
        node.p[0] = aabb.min.x;
        node.p[1] = aabb.min.y;
        node.p[2] = aabb.min.z;
        // bvh2_to_cwbvh.rs:85-88: exp2(ceil(log2(max(ext, 1e-20) * DENOM)))
        float ext[3] = {aabb.max.x - aabb.min.x, aabb.max.y - aabb.min.y, aabb.max.z - aabb.min.z};
        float e[3], rcp_e[3];
        for (int a = 0; a < 3; ++a) {
            float m = smax(ext[a], 1e-20f) * DENOM;
            e[a] = exp2f(ceilf(log2f(m)));
            rcp_e[a] = 1.0f / e[a];
            uint32_t bits;
            std::memcpy(&bits, &e[a], 4);
            node.e[a] = (uint8_t)(bits >> 23);
        }
        uint32_t children[8];
        for (int i = 0; i < 8; ++i) children[i] = INVALID32;
        uint32_t child_count = 0;
        get_children(node_index_bvh2, children, child_count, 0);
        if (err) return;
        if (order_children) {
            order(node_index_bvh2, children, child_count);
            if (err) return;
        }
        node.imask = 0;
        node.primitive_base_idx = (uint32_t)prims.size();
        node.child_base_idx = (uint32_t)out.size();
        uint32_t num_internal = 0, num_primitives = 0;
        float pn[3] = {node.p[0], node.p[1], node.p[2]};
        for (int i = 0; i < 8; ++i) {
            uint32_t ci = children[i];
            if (ci == INVALID32) continue;
            const Aabb& ca = nodes[ci].aabb;
            float cmin[3] = {ca.min.x, ca.min.y, ca.min.z}, cmax[3] = {ca.max.x, ca.max.y, ca.max.z};
            uint8_t qmin[3], qmax[3];
            for (int a = 0; a < 3; ++a) {
                // bvh2_to_cwbvh.rs:128-132 (PAD = 0.0; clamp = max(min).min(max) in glam)
                float lo = floorf(((cmin[a] - pn[a]) - 0.0f) * rcp_e[a]);
                float hi = g_mut == MUT_CW_MAX_FLOOR ? floorf(((cmax[a] - pn[a]) + 0.0f) * rcp_e[a])
                                                     : ceilf(((cmax[a] - pn[a]) + 0.0f) * rcp_e[a]);
                lo = smin(smax(lo, 0.0f), 255.0f);
                hi = smin(smax(hi, 0.0f), 255.0f);
                qmin[a] = (uint8_t)lo;
                qmax[a] = (uint8_t)hi;
            }
            node.child_min_x[i] = qmin[0];
            node.child_min_y[i] = qmin[1];
            node.child_min_z[i] = qmin[2];
            node.child_max_x[i] = qmax[0];
            node.child_max_y[i] = qmax[1];
            node.child_max_z[i] = qmax[2];
            uint8_t kind = decisions[(size_t)ci * 7].kind;
            if (kind == KIND_LEAF) {
                uint32_t pc = count_primitives(ci);
                uint8_t unary;
                if (pc == 1) unary = 0x20;
                else if (pc == 2) unary = 0x60;
                else if (pc == 3) unary = 0xe0;
                else {
                    err = ORC_ERR_CW_PRIMCOUNT;
                    return;
                }
                node.child_meta[i] = (uint8_t)(num_primitives | unary);
                num_primitives += pc;
            } else if (kind == KIND_INTERNAL) {
                node.imask |= (uint8_t)(1u << i);
                node.child_meta[i] = (uint8_t)((24 + i) | 0x20);
                num_internal += 1;
            } else {
                err = ORC_ERR_CW_DISTRIBUTE;  // bvh2_to_cwbvh.rs:168 unreachable!()
                return;
            }
        }
        out.resize(out.size() + num_internal, CwBvhNode{});
        cw_to_bvh2.resize(out.size(), INVALID32);
        out[node_index_bvh8] = node;
        cw_to_bvh2[node_index_bvh8] = (uint32_t)node_index_bvh2;
        uint32_t off = 0;
        for (int i = 0; i < 8; ++i) {
            if (children[i] != INVALID32 && (node.imask & (1u << i)) != 0) {
                convert((size_t)node.child_base_idx + off, children[i]);
                if (err) return;
                off += 1;
            }
        }
    }
};

// ---------------------------------------------------------------------------------------------------------------
// Traversal (cwbvh/mod.rs, cwbvh/traverse_macro.rs, cwbvh/node.rs, cwbvh/simd.rs, triangle.rs, faststack.rs)
// ---------------------------------------------------------------------------------------------------------------

constexpr float NODE_EPSILON = 0.0001f;  // cwbvh/node.rs:82

// cwbvh/mod.rs:996-998
inline uint32_t firstbithigh(uint32_t v) { return 31u - (uint32_t)__builtin_clz(v); }

// cwbvh/mod.rs:1001-1010
inline uint32_t ray_get_octant_inv4(V3 d) {
    return (d.x < 0.0f ? 0u : 0x04040404u) | (d.y < 0.0f ? 0u : 0x02020202u) | (d.z < 0.0f ? 0u : 0x01010101u);
}

inline float bits_to_f32(uint32_t b) {
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
inline uint32_t f32_to_bits(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

// cwbvh/node.rs:207-231
inline void get_child_and_index_bits(const CwBvhNode& n, uint32_t oct_inv4, uint64_t& child_bits8,
                                     uint64_t& bit_index8) {
    uint64_t oct_inv8 = (uint64_t)oct_inv4;
    oct_inv8 |= oct_inv8 << 32;
    uint64_t meta8;
    std::memcpy(&meta8, n.child_meta, 8);
    const uint64_t inner_mask = 0x1010101010101010ull;
    uint64_t is_inner8 = (meta8 & (meta8 << 1)) & inner_mask;
    uint64_t inner_mask8 = (is_inner8 >> 4) * 0xffull;
    const uint64_t index_mask = 0x1f1f1f1f1f1f1f1full;
    bit_index8 = (meta8 ^ (oct_inv8 & inner_mask8)) & index_mask;
    const uint64_t child_mask = 0x0707070707070707ull;
    child_bits8 = (meta8 >> 5) & child_mask;
}

// cwbvh/simd.rs:17-100 (the x86 SSE path the reference takes by default)
// `robust` (the watertight mode only): each quantised child box widened by one quantum per side and the slab
// interval by a relative 2^-20, so rounding in the quantisation and in the slab arithmetic cannot cull a child the
// ray touches (the crate's test is exact only in exact arithmetic).
inline uint32_t intersect_node(const CwBvhNode& n, const Ray& ray, float ray_tmax, uint32_t oct_inv4,
                               bool robust = false) {
    V3 ext = v3(bits_to_f32((uint32_t)n.e[0] << 23), bits_to_f32((uint32_t)n.e[1] << 23),
                bits_to_f32((uint32_t)n.e[2] << 23));
    float adx = ext.x * ray.inv_direction.x, ady = ext.y * ray.inv_direction.y, adz = ext.z * ray.inv_direction.z;
    float aox = (n.p[0] - ray.origin.x) * ray.inv_direction.x;
    float aoy = (n.p[1] - ray.origin.y) * ray.inv_direction.y;
    float aoz = (n.p[2] - ray.origin.z) * ray.inv_direction.z;
    bool rdx = ray.direction.x < 0.0f, rdy = ray.direction.y < 0.0f, rdz = ray.direction.z < 0.0f;
    uint64_t child_bits8, bit_index8;
    get_child_and_index_bits(n, oct_inv4, child_bits8, bit_index8);
    uint32_t hit_mask = 0;
    for (int c = 0; c < 8; ++c) {
        float qlx = (float)n.child_min_x[c], qhx = (float)n.child_max_x[c];
        float qly = (float)n.child_min_y[c], qhy = (float)n.child_max_y[c];
        float qlz = (float)n.child_min_z[c], qhz = (float)n.child_max_z[c];
        if (robust) {
            qlx -= 1.0f, qly -= 1.0f, qlz -= 1.0f;
            qhx += 1.0f, qhy += 1.0f, qhz += 1.0f;
        }
        float xmin = rdx ? qhx : qlx, xmax = rdx ? qlx : qhx;
        float ymin = rdy ? qhy : qly, ymax = rdy ? qly : qhy;
        float zmin = rdz ? qhz : qlz, zmax = rdz ? qlz : qhz;
        float tminx = xmin * adx + aox, tmaxx = xmax * adx + aox;
        float tminy = ymin * ady + aoy, tmaxy = ymax * ady + aoy;
        float tminz = zmin * adz + aoz, tmaxz = zmax * adz + aoz;
        float tmin = smax(tminx, smax(tminy, tminz));
        float tmax = smin(tmaxx, smin(tmaxy, tmaxz));
        if (g_mut == MUT_NODE_IEEE) {
            tmin = std::fmax(tminx, std::fmax(tminy, tminz));
            tmax = std::fmin(tmaxx, std::fmin(tmaxy, tmaxz));
        }
        if (robust) {
            const float k = 9.5367431640625e-07f;  // 2^-20
            tmin = tmin - std::fabs(tmin) * k;
            tmax = tmax + std::fabs(tmax) * k;
        }
        if (g_mut == MUT_NODE_IEEE) {
            tmin = std::fmax(tmin, NODE_EPSILON);
            tmax = std::fmin(tmax, ray_tmax);
        } else {
            tmin = smax(tmin, g_mut == MUT_NODE_TMIN_ZERO ? 0.0f : NODE_EPSILON);
            tmax = smin(tmax, ray_tmax);
        }
        if (tmin <= tmax) {
            uint32_t cb = (uint32_t)(child_bits8 >> (c * 8)) & 0xffu;
            uint32_t bi = (uint32_t)(bit_index8 >> (c * 8)) & 0xffu;
            hit_mask |= cb << bi;
        }
    }
    return hit_mask;
}

// triangle.rs:35-76
inline float tri_intersect(const Triangle& tri, const Ray& ray, float tmin, float tmax) {
    V3 e1 = vsub(tri.v0, tri.v1);
    V3 e2 = vsub(tri.v2, tri.v0);
    V3 n = vcross(e1, e2);
    V3 c = vsub(tri.v0, ray.origin);
    V3 r = vcross(ray.direction, c);
    float inv_det = 1.0f / vdot(n, ray.direction);
    float u = vdot(r, e2) * inv_det;
    float v = vdot(r, e1) * inv_det;
    float w = 1.0f - u - v;
    uint32_t hit = f32_to_bits(u) | f32_to_bits(v) | f32_to_bits(w);
    bool valid = inv_det != 0.0f && (hit & 0x80000000u) == 0;
    if (g_mut == MUT_TRI_GE_ZERO) valid = inv_det != 0.0f && u >= 0.0f && v >= 0.0f && w >= 0.0f;
    if (valid) {
        float t = vdot(n, c) * inv_det;
        if (t >= tmin && t <= tmax) return t;
    }
    return INFINITY;
}

// Watertight ray/triangle test (no counterpart in the crate: the library's extra primitive test), restated from
// Woop, Benthin, Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1) 2013: dominant axis kz of |d| (ties to
// the lower axis), kx, ky the next two cyclically and swapped when d[kz] < 0; shear Sx = d[kx]/d[kz],
// Sy = d[ky]/d[kz], Sz = 1/d[kz]; edge functions of the sheared, translated vertices, recomputed in f64 when one is
// exactly 0; no culling; t = T/det in [tmin, tmax].
struct WtRay {
    uint32_t kx, ky, kz;
    float sx, sy, sz;
};

inline float comp3(const V3& v, uint32_t k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

inline WtRay wt_ray(const Ray& ray) {
    const V3& d = ray.direction;
    const float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    WtRay w;
    if (g_mut == MUT_WT_MIN_AXIS)
        w.kz = (ax <= ay && ax <= az) ? 0u : (ay <= az ? 1u : 2u);
    else
        w.kz = (ax >= ay && ax >= az) ? 0u : (ay >= az ? 1u : 2u);
    w.kx = (w.kz + 1u) % 3u;
    w.ky = (w.kx + 1u) % 3u;
    const float dz = comp3(d, w.kz);
    if (dz < 0.0f) std::swap(w.kx, w.ky);
    w.sx = comp3(d, w.kx) / dz;
    w.sy = comp3(d, w.ky) / dz;
    w.sz = 1.0f / dz;
    return w;
}

inline float tri_intersect_wt(const Triangle& tri, const Ray& ray, const WtRay& w, float tmin, float tmax) {
    const V3 a = vsub(tri.v0, ray.origin), b = vsub(tri.v1, ray.origin), c = vsub(tri.v2, ray.origin);
    const float az0 = comp3(a, w.kz), bz0 = comp3(b, w.kz), cz0 = comp3(c, w.kz);
    const float ax = comp3(a, w.kx) - w.sx * az0, ay = comp3(a, w.ky) - w.sy * az0;
    const float bx = comp3(b, w.kx) - w.sx * bz0, by = comp3(b, w.ky) - w.sy * bz0;
    const float cx = comp3(c, w.kx) - w.sx * cz0, cy = comp3(c, w.ky) - w.sy * cz0;
    float u = cx * by - cy * bx;
    float v = ax * cy - ay * cx;
    float e = bx * ay - by * ax;
    if (u == 0.0f || v == 0.0f || e == 0.0f) {
        const double uu = (double)cx * (double)by - (double)cy * (double)bx;
        const double vv = (double)ax * (double)cy - (double)ay * (double)cx;
        const double ee = (double)bx * (double)ay - (double)by * (double)ax;
        u = (float)uu;
        v = (float)vv;
        e = (float)ee;
    }
    const bool any_neg = u < 0.0f || v < 0.0f || e < 0.0f;
    const bool any_pos = u > 0.0f || v > 0.0f || e > 0.0f;
    if (any_neg && any_pos) return INFINITY;
    if (g_mut == MUT_WT_EDGE_STRICT && (u == 0.0f || v == 0.0f || e == 0.0f)) return INFINITY;
    const float det = (u + v) + e;
    if (det == 0.0f) return INFINITY;
    const float az = w.sz * az0, bz = w.sz * bz0, cz = w.sz * cz0;
    float tt = (u * az + v * bz) + e * cz;
    float adet = det;
    if (det < 0.0f) {
        tt = -tt;
        adet = -det;
    }
    const float t = tt / adet;
    return (t >= tmin && t <= tmax) ? t : INFINITY;
}

struct TravStats {
    uint32_t nodes;
    uint32_t tris;
};

// cwbvh/mod.rs:169-198 (closest) and 201-225 (miss), both through traverse! (traverse_macro.rs:59-126) with a
// StackStack<UVec2, 32> (faststack.rs:282-336).  `mode` 0 = closest hit, 1 = miss/occlusion; 2 and 3 the same with
// the watertight primitive test (tri_intersect_wt) in place of Triangle::intersect.
// Returns hit (closest) or miss flag.  Triangles are indexed by position in primitive_indices.
inline void traverse_one(const CwBvhNode* nodes, uint32_t nnodes, const Triangle* tris, const Ray& ray, int mode,
                         RayHit* hit, uint8_t* miss, TravStats* st) {
    struct G {
        uint32_t x, y;
    };
    const int STACK = 32;
    G stack[STACK];
    int sp = 0;
    G current = nnodes == 0 ? G{0, 0} : G{0, 0x80000000u};
    G prim{0, 0};
    uint32_t oct_inv4 = ray_get_octant_inv4(ray.direction);
    const WtRay wr = (mode & 2) ? wt_ray(ray) : WtRay{};
    float tmax = ray.tmax;
    RayHit h{INVALID32, INVALID32, INVALID32, INFINITY};
    bool is_miss = true;
    uint32_t n_nodes = 0, n_tris = 0;
    for (;;) {
        bool stop = false;
        while (prim.y != 0) {
            uint32_t local = g_mut == MUT_PRIM_LOWBIT ? (uint32_t)__builtin_ctz(prim.y) : firstbithigh(prim.y);
            prim.y &= ~(1u << local);
            uint32_t pid = prim.x + local;
            n_tris++;
            float t = (mode & 2) ? tri_intersect_wt(tris[pid], ray, wr, ray.tmin, (mode & 1) ? ray.tmax : tmax)
                                 : tri_intersect(tris[pid], ray, ray.tmin, tmax);
            if ((mode & 1) == 0) {
                if (t < tmax) {
                    h.primitive_id = pid;
                    h.t = t;
                    tmax = t;
                }
            } else {
                if (t < ray.tmax) {
                    is_miss = false;
                    stop = true;
                    break;
                }
            }
        }
        if (stop) break;
        prim = G{0, 0};
        if ((current.y & 0xff000000u) != 0) {
            uint32_t hits_imask = current.y;
            uint32_t child_index_offset = firstbithigh(hits_imask);
            uint32_t child_index_base = current.x;
            current.y &= ~(1u << child_index_offset);
            if ((current.y & 0xff000000u) != 0) {
                stack[sp] = current;  // faststack.rs:299-303
                sp = std::min(sp + 1, STACK - 1);
            }
            uint32_t slot_index = (child_index_offset - 24) ^ (oct_inv4 & 0xffu);
            uint32_t relative_index = (uint32_t)__builtin_popcount(hits_imask & ~(0xffffffffu << slot_index));
            uint32_t child = child_index_base + relative_index;
            const CwBvhNode& nd = nodes[child];
            n_nodes++;
            uint32_t hitmask = intersect_node(nd, ray, tmax, oct_inv4, (mode & 2) != 0 && g_mut != MUT_WT_NODE_PLAIN);
            current = G{nd.child_base_idx, (hitmask & 0xff000000u) | (uint32_t)nd.imask};
            prim = G{nd.primitive_base_idx, hitmask & 0x00ffffffu};
        } else {
            current = G{0, 0};
        }
        if (prim.y == 0 && (current.y & 0xff000000u) == 0) {
            if (sp == 0) break;
            sp = sp > 0 ? sp - 1 : 0;  // faststack.rs:306-310
            current = stack[sp];
        }
    }
    if (hit) *hit = h;
    if (miss) *miss = is_miss ? 1 : 0;
    if (st) {
        st->nodes = n_nodes;
        st->tris = n_tris;
    }
}

}  // namespace

// ---------------------------------------------------------------------------------------------------------------
// Spatial pre-splits (splits.rs) and the pre-split leaf set of build_cwbvh_from_tris (cwbvh/builder.rs:26-55)
// ---------------------------------------------------------------------------------------------------------------

inline float vget(const V3& v, uint32_t d) { return d == 0 ? v.x : (d == 1 ? v.y : v.z); }
inline void vset(V3& v, uint32_t d, float f) {
    if (d == 0) v.x = f;
    else if (d == 1) v.y = f;
    else v.z = f;
}
// aabb.rs:125-134
inline uint32_t largest_axis(const Aabb& a) {
    const float dx = a.max.x - a.min.x, dy = a.max.y - a.min.y, dz = a.max.z - a.min.z;
    if (dx < dy) return dy < dz ? 2u : 1u;
    return dx < dz ? 2u : 0u;
}
// aabb.rs:98-103: min = self.min.max(other.min), max = self.max.min(other.max)
inline Aabb aabb_intersection(const Aabb& a, const Aabb& b) { return Aabb{vmax(a.min, b.min), vmin(a.max, b.max)}; }
// aabb.rs:23-26
inline Aabb aabb_invalid() {
    const float M = 3.40282347e+38f;
    return Aabb{v3(M, M, M), v3(-M, -M, -M)};
}

// splits.rs:137-158.  The crossing point is Vec3A::mul_add(splat(t), v1 - v0, v0), a fused multiply-add per lane.
void split_triangle(uint32_t dim, float pos, const V3 v[4], Aabb& left, Aabb& right) {
    left = aabb_invalid();
    right = aabb_invalid();
    for (int i = 0; i < 3; ++i) {
        const V3 v0 = v[i], v1 = v[i + 1];
        const float v0d = vget(v0, dim), v1d = vget(v1, dim);
        if (v0d <= pos) aabb_extend(left, v0);
        if (v0d >= pos) aabb_extend(right, v0);
        if ((v0d < pos && pos < v1d) || (v1d < pos && pos < v0d)) {
            const float inv_length = 1.0f / (v1d - v0d);
            const float t = (pos - v0d) * inv_length;
            const V3 e = vsub(v1, v0);
            V3 c;
            if (g_mut == MUT_SPLIT_UNFUSED)
                c = v3(t * e.x + v0.x, t * e.y + v0.y, t * e.z + v0.z);
            else
                c = v3(std::fmaf(t, e.x, v0.x), std::fmaf(t, e.y, v0.y), std::fmaf(t, e.z, v0.z));
            aabb_extend(left, c);
            aabb_extend(right, c);
        }
    }
}

// splits.rs:48-118.  `max_total` (0 = unlimited) is an extension: once the leaf list holds max_total entries no
// further split is applied (candidates are still visited in the reference order, so the first splits are the
// reference's).
void split_aabbs_precise(std::vector<Aabb>& aabbs, std::vector<uint32_t>& indices, const Triangle* tris,
                         float area_thresh_low, float area_thresh_high, float split_factor_low,
                         float split_factor_high, uint32_t max_iterations, uint32_t split_tests, size_t max_total) {
    std::vector<size_t> candidates;
    for (size_t i = 0; i < aabbs.size(); ++i)
        if (half_area(aabbs[i]) > area_thresh_low) candidates.push_back(i);
    size_t old_len = candidates.size();
    for (uint32_t it = 0; it < max_iterations; ++it) {
        const size_t count = candidates.size();
        for (size_t ci = 0; ci < count; ++ci) {
            const size_t c = candidates[ci];
            const Aabb aabb = aabbs[c];
            const uint32_t index = indices[c];
            const uint32_t axis = largest_axis(aabb);
            const Triangle& tri = tris[index];
            float best_cost = 3.40282347e+38f;
            Aabb left = aabb, right = aabb;
            for (uint32_t k = 1; k < split_tests; ++k) {
                const float n = (float)k / (float)split_tests;
                const float pos = vget(aabb.min, axis) * n + vget(aabb.max, axis) * (1.0f - n);
                Aabb tl = aabb, tr = aabb;
                vset(tl.max, axis, pos);
                vset(tr.min, axis, pos);
                const V3 verts[4] = {tri.v0, tri.v1, tri.v2, tri.v0};
                Aabb sl, sr;
                split_triangle(axis, pos, verts, sl, sr);
                tl = aabb_intersection(sl, tl);
                tr = aabb_intersection(sr, tr);
                const float area = half_area(tl) + half_area(tr);
                if (area < best_cost) {
                    best_cost = area;
                    left = tl;
                    right = tr;
                }
            }
            const float old_cost = half_area(aabb);
            if ((area_thresh_high > old_cost && best_cost * split_factor_high < old_cost) ||
                best_cost * split_factor_low < old_cost) {
                if (max_total != 0 && aabbs.size() >= max_total) continue;
                aabbs[c] = left;
                candidates.push_back(aabbs.size());
                aabbs.push_back(right);
                indices.push_back(index);
            }
        }
        if (old_len == candidates.size()) break;
        size_t w = 0;
        for (size_t ci = 0; ci < candidates.size(); ++ci)
            if (half_area(aabbs[candidates[ci]]) > area_thresh_low) candidates[w++] = candidates[ci];
        candidates.resize(w);
        old_len = candidates.size();
    }
}

// cwbvh/builder.rs:26-46 (leaf boxes, largest and average half area, summed in triangle order) + splits.rs:16-33.
void presplit_leaves(const Triangle* tris, size_t n, size_t max_total, std::vector<Aabb>& aabbs,
                     std::vector<uint32_t>& indices, float* out_avg, float* out_largest) {
    float largest = 0.0f, avg = 0.0f;
    aabbs.resize(n);
    indices.resize(n);
    for (size_t i = 0; i < n; ++i) {
        aabbs[i] = tri_aabb(tris[i]);
        const float ha = half_area(aabbs[i]);
        largest = std::fmax(ha, largest);  // f32::max (IEEE maxNum)
        avg += ha;
        indices[i] = (uint32_t)i;
    }
    avg /= (float)n;
    if (out_avg) *out_avg = avg;
    if (out_largest) *out_largest = largest;
    split_aabbs_precise(aabbs, indices, tris, avg * 3.0f, std::fmax(avg * 4.0f, avg * 0.9f + largest * 0.1f), 1.8f,
                        1.6f, 12, 12, max_total);
}

// ---------------------------------------------------------------------------------------------------------------
// Reinsertion optimizer (bvh2/reinsertion.rs) on a Bvh2 with parents (bvh2/mod.rs:572-619, 722-751)
// ---------------------------------------------------------------------------------------------------------------

inline size_t sibling_id(size_t id) { return (id % 2 == 1) ? id + 1 : id - 1; }      // bvh2/node.rs:158-164
inline size_t left_sibling_id(size_t id) { return (id % 2 == 1) ? id : id - 1; }     // bvh2/node.rs:166-172

// bvh2/mod.rs:586-609 (compute_parents)
void compute_parents(const std::vector<Bvh2Node>& nodes, std::vector<uint32_t>& parents) {
    parents.assign(nodes.size(), 0u);
    for (size_t i = 0; i < nodes.size(); ++i) {
        if (nodes[i].prim_count == 0) {
            parents[nodes[i].first_index] = (uint32_t)i;
            parents[nodes[i].first_index + 1] = (uint32_t)i;
        }
    }
}

struct Reinsertion {
    uint32_t from, to;
    float area_diff;
};

// bvh2/mod.rs:722-751 (refit_from_fast, release build: stops at the second unchanged node)
// A corrupted tree (only reachable through a planted mistake) could hold a parent cycle: the walks below are bounded
// by the node count and report it through g_tree_corrupt instead of looping.
bool g_tree_corrupt = false;

void refit_from_fast(std::vector<Bvh2Node>& nodes, const std::vector<uint32_t>& parents, size_t index) {
    int same_count = 0;
    for (size_t steps = 0;; ++steps) {
        if (steps > nodes.size()) {
            g_tree_corrupt = true;
            return;
        }
        Bvh2Node& node = nodes[index];
        if (node.prim_count == 0) {
            const Aabb nb = aabb_union(nodes[node.first_index].aabb, nodes[node.first_index + 1].aabb);
            if (aabb_eq(node.aabb, nb)) {
                same_count += 1;
                if (same_count == 2 && g_mut != MUT_REFIT_FULL) return;
            }
            node.aabb = nb;
        }
        if (index == 0) break;
        index = parents[index];
    }
}

// reinsertion.rs:247-340 (find_reinsertion)
Reinsertion find_reinsertion(const std::vector<Bvh2Node>& nodes, const std::vector<uint32_t>& parents,
                             size_t node_id, std::vector<std::pair<float, uint32_t>>& stack) {
    Reinsertion best{(uint32_t)node_id, 0u, 0.0f};
    const float node_area = half_area(nodes[node_id].aabb);
    const float parent_area = half_area(nodes[parents[node_id]].aabb);
    float area_diff = parent_area;
    size_t sib = sibling_id(node_id);
    Aabb pivot_bbox = nodes[sib].aabb;
    const size_t parent_id = parents[node_id];
    size_t pivot_id = parent_id;
    const Aabb aabb = nodes[node_id].aabb;
    stack.clear();
    size_t budget = 4 * nodes.size() + 64;
    for (;;) {
        stack.push_back({area_diff, (uint32_t)sib});
        while (!stack.empty()) {
            if (budget-- == 0) {
                g_tree_corrupt = true;
                return Reinsertion{0u, 0u, 0.0f};
            }
            const auto top = stack.back();
            stack.pop_back();
            if (top.first - node_area <= best.area_diff) continue;
            const Bvh2Node& dst = nodes[top.second];
            const float merged_area = half_area(aabb_union(dst.aabb, aabb));
            const float reinsert_area = top.first - merged_area;
            if (reinsert_area > best.area_diff) {
                best.to = top.second;
                best.area_diff = reinsert_area;
            }
            if (dst.prim_count == 0) {
                const float child_area = reinsert_area + half_area(dst.aabb);
                if (g_mut == MUT_REINS_PUSH_ORDER) {
                    stack.push_back({child_area, dst.first_index + 1});
                    stack.push_back({child_area, dst.first_index});
                } else {
                    stack.push_back({child_area, dst.first_index});
                    stack.push_back({child_area, dst.first_index + 1});
                }
            }
        }
        if (pivot_id != parent_id) {
            pivot_bbox = aabb_union(pivot_bbox, nodes[sib].aabb);
            area_diff += half_area(nodes[pivot_id].aabb) - half_area(pivot_bbox);
        }
        if (pivot_id == 0) break;
        sib = sibling_id(pivot_id);
        pivot_id = parents[pivot_id];
    }
    if (best.to == sibling_id(best.from) || best.to == parents[best.from]) best = Reinsertion{0u, 0u, 0.0f};
    return best;
}

// reinsertion.rs:342-382 (reinsert_node; primitives_to_nodes is not kept here)
void reinsert_node(std::vector<Bvh2Node>& nodes, std::vector<uint32_t>& parents, size_t from, size_t to) {
    const size_t sib = sibling_id(from);
    const size_t parent_id = parents[from];
    const Bvh2Node sibling_node = nodes[sib];
    const Bvh2Node dst_node = nodes[to];
    nodes[to].prim_count = 0;  // make_inner (bvh2/node.rs:211-214)
    nodes[to].first_index = (uint32_t)left_sibling_id(from);
    nodes[sib] = dst_node;
    nodes[parent_id] = sibling_node;
    if (nodes[sib].prim_count == 0) {
        parents[nodes[sib].first_index] = (uint32_t)sib;
        parents[nodes[sib].first_index + 1] = (uint32_t)sib;
    }
    if (nodes[parent_id].prim_count == 0) {
        parents[nodes[parent_id].first_index] = (uint32_t)parent_id;
        parents[nodes[parent_id].first_index + 1] = (uint32_t)parent_id;
    }
    parents[sib] = (uint32_t)to;
    parents[from] = (uint32_t)to;
    refit_from_fast(nodes, parents, to);
    refit_from_fast(nodes, parents, parent_id);
}

// The candidate order key (reinsertion.rs:210-218): rdst's f32 radix key of -cost (sign-flip total order),
// ascending.  rdst's sort is unstable; equal keys are ordered by node index here.
inline uint32_t rdst_f32_key(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return (b >> 31) ? ~b : (b ^ 0x80000000u);
}

// reinsertion.rs:40-61 (run), 101-119 (optimize_impl), 128-147 (find_candidates), 149-204 (optimize_candidates).
// Ties of the two unstable sorts (candidates by cost, reinsertions by area_diff) are broken by node index and by
// candidate rank respectively.
void reinsertion_run(std::vector<Bvh2Node>& nodes, float batch_size_ratio, const float* seq, size_t seq_len) {
    if (nodes.empty() || nodes[0].prim_count != 0 || !(batch_size_ratio > 0.0f)) return;
    std::vector<uint32_t> parents;
    compute_parents(nodes, parents);
    std::vector<float> default_seq;
    if (!seq) {
        for (int k = 1; k < 32; k += 2) default_seq.push_back(1.0f / (float)k);
        seq = default_seq.data();
        seq_len = default_seq.size();
    }
    std::vector<std::pair<uint32_t, uint32_t>> cand;  // (key, node id)
    std::vector<std::pair<float, uint32_t>> stack;
    std::vector<Reinsertion> reins;
    std::vector<uint8_t> touched(nodes.size());
    g_tree_corrupt = false;
    for (size_t si = 0; si < seq_len && !g_tree_corrupt; ++si) {
        const float prod = ((float)nodes.size() * batch_size_ratio) * seq[si];
        size_t batch_size = prod >= 18446744073709551615.0f ? SIZE_MAX : (prod > 0.0f ? (size_t)prod : 0);
        batch_size = std::max<size_t>(batch_size, 1);
        const size_t node_count = std::min(nodes.size(), batch_size + 1);
        // find_candidates
        cand.clear();
        const size_t take = std::min(nodes.size(), node_count * 2);
        for (size_t i = 1; i < take; ++i) cand.push_back({rdst_f32_key(-half_area(nodes[i].aabb)), (uint32_t)i});
        std::sort(cand.begin(), cand.end());
        // optimize_candidates(node_count - 1)
        const size_t count = node_count - 1;
        std::fill(touched.begin(), touched.end(), 0);
        reins.clear();
        if (g_threads > 1) {
            // reinsertion.rs:145-155: every candidate's search in parallel (par_iter), then the positive ones
            const size_t m = std::min(count, cand.size());
            std::vector<Reinsertion> all(m);
            parallel_chunks(m, g_threads, [&](size_t i0, size_t i1) {
                std::vector<std::pair<float, uint32_t>> st;
                for (size_t i = i0; i < i1; ++i) all[i] = find_reinsertion(nodes, parents, cand[i].second, st);
            });
            for (const Reinsertion& r : all)
                if (r.area_diff > 0.0f) reins.push_back(r);
        } else {
            for (size_t i = 0; i < count && i < cand.size(); ++i) {
                Reinsertion r = find_reinsertion(nodes, parents, cand[i].second, stack);
                if (r.area_diff > 0.0f) reins.push_back(r);
            }
        }
        std::stable_sort(reins.begin(), reins.end(),
                         [](const Reinsertion& a, const Reinsertion& b) { return a.area_diff > b.area_diff; });
        for (const Reinsertion& r : reins) {
            const size_t conf[5] = {r.to, r.from, sibling_id(r.from), parents[r.to], parents[r.from]};
            bool any = false;
            if (g_mut != MUT_REINS_NO_CONFLICT)
                for (size_t c : conf) any |= touched[c] != 0;
            if (any) continue;
            for (size_t c : conf) touched[c] = 1;
            reinsert_node(nodes, parents, r.from, r.to);
            if (g_tree_corrupt) break;
        }
    }
}

// bvh2/mod.rs:527-569 (refit_all, general order: reverse of a depth-first visit)
void refit_all(std::vector<Bvh2Node>& nodes) {
    if (nodes.empty()) return;
    std::vector<uint32_t> stack{0u}, rev{0u};
    while (!stack.empty()) {
        const uint32_t c = stack.back();
        stack.pop_back();
        if (nodes[c].prim_count == 0) {
            rev.push_back(nodes[c].first_index);
            rev.push_back(nodes[c].first_index + 1);
            stack.push_back(nodes[c].first_index);
            stack.push_back(nodes[c].first_index + 1);
        }
    }
    for (size_t k = rev.size(); k-- > 0;) {
        Bvh2Node& nd = nodes[rev[k]];
        if (nd.prim_count == 0) nd.aabb = aabb_union(nodes[nd.first_index].aabb, nodes[nd.first_index + 1].aabb);
    }
}

// cwbvh/mod.rs:341-491 (CwBvh::ray_traverse_tlas_blas) with the intersection closure fixed to Triangle::intersect
// on the BLAS's triangles in primitive_indices order.  One shared StackStack<UVec2, 32> for both levels; TLAS
// primitives are BLAS indices through the TLAS's primitive_indices.  Restated as written, including that a TLAS
// primitive group popped after a BLAS finishes is treated as a node group and dropped (mod.rs:396-401, 469).
struct BlasRef {
    const CwBvhNode* nodes;
    uint32_t nnodes;
    const Triangle* tris;
};

inline void traverse_tlas_one(const CwBvhNode* tlas, uint32_t tlas_nnodes, const uint32_t* tlas_prims,
                              const BlasRef* blas, uint32_t nblas, Ray ray, RayHit* hit) {
    struct G {
        uint32_t x, y;
    };
    const int STACK = 32;
    G stack[STACK];
    int index = 0;  // StackStack::index
    auto push = [&](G v) {
        stack[index] = v;
        index = std::min(index + 1, STACK - 1);
    };
    auto pop_fast = [&]() {
        index = index > 0 ? index - 1 : 0;
        return stack[index];
    };
    uint32_t tlas_stack_size = INVALID32;
    uint32_t current_mesh = INVALID32;
    const CwBvhNode* nodes = tlas;
    uint32_t nnodes = tlas_nnodes;
    const uint32_t oct_inv4 = ray_get_octant_inv4(ray.direction);
    G current = tlas_nnodes == 0 ? G{0, 0} : G{0, 0x80000000u};
    RayHit h{INVALID32, INVALID32, INVALID32, INFINITY};
    for (;;) {
        G prim{0, 0};
        if ((current.y & 0xff000000u) != 0) {
            const uint32_t hits_imask = current.y;
            const uint32_t off = firstbithigh(hits_imask);
            const uint32_t base = current.x;
            current.y &= ~(1u << off);
            if ((current.y & 0xff000000u) != 0) push(current);
            const uint32_t slot_index = (off - 24) ^ (oct_inv4 & 0xffu);
            const uint32_t rel = (uint32_t)__builtin_popcount(hits_imask & ~(0xffffffffu << slot_index));
            const uint32_t child = base + rel;
            if (child >= nnodes) break;  // out of bounds in the reference (panic); never for valid inputs
            const CwBvhNode& nd = nodes[child];
            const uint32_t hitmask = intersect_node(nd, ray, ray.tmax, oct_inv4);
            current = G{nd.child_base_idx, (hitmask & 0xff000000u) | (uint32_t)nd.imask};
            prim = G{nd.primitive_base_idx, hitmask & 0x00ffffffu};
        } else {
            current = G{0, 0};
        }
        while (prim.y != 0) {
            if (tlas_stack_size == INVALID32) {
                const uint32_t local = firstbithigh(prim.y);
                prim.y &= ~(1u << local);
                const uint32_t global = prim.x + local;
                if (prim.y != 0) push(prim);
                if ((current.y & 0xff000000u) != 0) push(current);
                tlas_stack_size = (uint32_t)index;
                const uint32_t blas_index = tlas_prims[global];
                if (blas_index >= nblas) {  // out of bounds in the reference; never for valid inputs
                    current = G{0, 0};
                    break;
                }
                nodes = blas[blas_index].nodes;
                nnodes = blas[blas_index].nnodes;
                current_mesh = blas_index;
                current = nnodes == 0 ? G{0, 0} : G{0, 0x80000000u};
                break;
            } else {
                const uint32_t local = firstbithigh(prim.y);
                prim.y &= ~(1u << local);
                const uint32_t global = prim.x + local;
                const float t = tri_intersect(blas[current_mesh].tris[global], ray, ray.tmin, ray.tmax);
                if (t < ray.tmax) {
                    h.primitive_id = global;
                    h.geometry_id = current_mesh;
                    ray.tmax = t;
                }
            }
        }
        if ((current.y & 0xff000000u) == 0) {
            if (index == 0) break;
            if ((uint32_t)index == tlas_stack_size) {
                tlas_stack_size = INVALID32;
                current_mesh = INVALID32;
                nodes = tlas;
                nnodes = tlas_nnodes;
            }
            current = pop_fast();
        }
    }
    if (h.primitive_id != INVALID32) h.t = ray.tmax;
    *hit = h;
}

// ---------------------------------------------------------------------------------------------------------------
// Partial / full rebuild (ploc/rebuild.rs) and primitive -> leaf map (bvh2/mod.rs:647-663)
// ---------------------------------------------------------------------------------------------------------------

// ploc/rebuild.rs:12-42 (compute_rebuild_path_flags): flag each given leaf and its ancestors, stopping at the first
// ancestor already flagged.  Nothing is written for a tree of fewer than two nodes.
int rebuild_path_flags(const std::vector<Bvh2Node>& nodes, const uint32_t* leaves, size_t n_leaves, uint8_t* flags) {
    if (nodes.size() < 2) return ORC_OK;
    std::vector<uint32_t> parents;
    compute_parents(nodes, parents);
    std::fill(flags, flags + nodes.size(), (uint8_t)0);
    for (size_t k = 0; k < n_leaves; ++k) {
        size_t index = leaves[k];
        if (index >= nodes.size()) return ORC_ERR_INVALID_ARG;  // reference: index out of bounds
        flags[index] = 1;
        size_t steps = 0;
        while (index > 0) {
            index = parents[index];
            if (flags[index]) break;
            flags[index] = 1;
            if (++steps > nodes.size()) return ORC_ERR_INVALID_ARG;  // not a tree
        }
    }
    return ORC_OK;
}

// ploc/rebuild.rs:98-134 (partial_rebuild) + 136-183 (rebuild_from_leaves::<true>): the root's subtree is walked
// top-down with a stack; a node that is not flagged, or is a leaf, becomes a cluster (its subtree stays where it
// is), a flagged inner node is opened.  Every visited slot is marked invalid, and build_ploc_from_leaves::<REBUILD>
// refills exactly those slots.  The scene box is the root's stored box.  Clusters with equal Morton codes are
// ordered by their old slot.
int partial_rebuild(std::vector<Bvh2Node>& nodes, const uint8_t* flags, size_t sd, bool u128, size_t sdt,
                    size_t* out_depth) {
    if (out_depth) *out_depth = 0;
    if (nodes.size() < 2) return ORC_OK;
    std::vector<Bvh2Node> current;
    std::vector<uint32_t> slot_of;
    std::vector<uint32_t> stack{nodes[0].first_index};
    size_t visited = 0;
    while (!stack.empty()) {
        const uint32_t left = stack.back();
        stack.pop_back();
        for (size_t node_index : {(size_t)left, (size_t)left + 1}) {
            if (node_index >= nodes.size() || ++visited > nodes.size()) return ORC_ERR_INVALID_ARG;
            Bvh2Node& node = nodes[node_index];
            if (!flags[node_index] || node.prim_count != 0u) {
                current.push_back(node);
                slot_of.push_back((uint32_t)node_index);
            } else {
                stack.push_back(node.first_index);
            }
            node.prim_count |= NODE_INVALID_BIT;
        }
    }
    const Aabb total = nodes[0].aabb;  // rebuild.rs:154
    return ploc_from_leaves(current, slot_of.data(), total, sd, sdt, false, true, u128, nodes.data(), nodes.size(),
                            nullptr, out_depth);
}

// ploc/rebuild.rs:53-79 (full_rebuild) + rebuild_from_leaves::<false>: every leaf, in slot order, rebuilt with the
// root's stored box as the scene box.  The node array becomes 2k-1 long for k leaves.
int full_rebuild(std::vector<Bvh2Node>& nodes, size_t sd, bool u128, size_t sdt, size_t* out_depth) {
    if (out_depth) *out_depth = 0;
    if (nodes.size() < 2) return ORC_OK;
    std::vector<Bvh2Node> current;
    for (const Bvh2Node& node : nodes)
        if (node.prim_count != 0u) current.push_back(node);
    const Aabb total = nodes[0].aabb;
    if (current.empty()) return ORC_OK;
    nodes.resize(2 * current.size() - 1);
    return ploc_from_leaves(current, nullptr, total, sd, sdt, false, false, u128, nodes.data(), nodes.size(),
                            nullptr, out_depth);
}

// bvh2/mod.rs:647-663 (compute_primitives_to_nodes): leaf slot of every primitive, INVALID where none.
int primitives_to_nodes(const Bvh2Node* nodes, size_t nnodes, const uint32_t* prim_indices, size_t nprims,
                        uint32_t* out) {
    std::fill(out, out + nprims, INVALID32);
    for (size_t node_id = 0; node_id < nnodes; ++node_id) {
        const Bvh2Node& node = nodes[node_id];
        if (node.prim_count == 0u) continue;
        const uint64_t start = node.first_index, end = start + node.prim_count;
        for (uint64_t k = start; k < end; ++k) {
            if (k >= nprims) return ORC_ERR_INVALID_ARG;
            const uint32_t prim_id = prim_indices[k];
            if (prim_id >= nprims) return ORC_ERR_INVALID_ARG;
            out[prim_id] = (uint32_t)node_id;
        }
    }
    return ORC_OK;
}

// ===============================================================================================================
// C ABI for the tests (ctypes)
// ===============================================================================================================
extern "C" {

int orc_sizeof(int which) {
    switch (which) {
        case 0: return (int)sizeof(Aabb);
        case 1: return (int)sizeof(Triangle);
        case 2: return (int)sizeof(Ray);
        case 3: return (int)sizeof(RayHit);
        case 4: return (int)sizeof(Bvh2Node);
        case 5: return (int)sizeof(CwBvhNode);
        default: return -1;
    }
}

void orc_tri_aabbs(const Triangle* tris, uint64_t n, Aabb* out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = tri_aabb(tris[i]);
}

uint64_t orc_morton_encode_u64_unorm(double x, double y, double z) { return morton_encode_u64_unorm(x, y, z); }
uint64_t orc_split_by_3_u64(uint32_t a) { return split_by_3_u64(a); }
// 128-bit results as {low 64 bits, high 64 bits}.
void orc_split_by_3_u128(uint64_t a, uint64_t* out) {
    const uint128 c = split_by_3_u128(a);
    out[0] = (uint64_t)c;
    out[1] = (uint64_t)(c >> 64);
}
void orc_morton_encode_u128_unorm(double x, double y, double z, uint64_t* out) {
    const uint128 c = morton_encode_u128_unorm(x, y, z);
    out[0] = (uint64_t)c;
    out[1] = (uint64_t)(c >> 64);
}
uint32_t orc_f64_as_u32(double v) { return f64_as_u32(v); }
float orc_half_area(const Aabb* a) { return half_area(*a); }

// Full PLOC build (ploc/mod.rs:95-113 with indices 0..n supplied by the caller); sort_precision 64 or 128.
// out_nodes: 2n-1 Bvh2Node (48 B layout), out_prim_indices: n.  out_codes gets the U64 codes, or the low 64 bits
// of the U128 codes with the high 64 bits in out_codes_hi.
// Trace outputs may be null.  Returns an ORC_* code.
int orc_ploc_build(const Aabb* aabbs, const uint32_t* indices, uint64_t n, uint32_t search_distance,
                   uint32_t search_depth_threshold, int basic_search, Bvh2Node* out_nodes,
                   uint32_t* out_prim_indices, uint64_t* out_codes, uint32_t* out_sorted_index,
                   uint32_t* out_iter_counts, uint32_t iter_cap, uint32_t* out_num_iters, uint32_t* out_depth,
                   uint32_t* out_max_offset, uint32_t sort_precision, uint64_t* out_codes_hi) {
    if (sort_precision != 64u && sort_precision != 128u) return ORC_ERR_INVALID_ARG;
    PlocTrace tr{out_codes, out_codes_hi, out_sorted_index, out_iter_counts, iter_cap, 0, 0, 0};
    int rc = ploc_build(aabbs, indices, (size_t)n, search_distance, search_depth_threshold, basic_search != 0,
                        sort_precision == 128u, out_nodes, out_prim_indices, &tr);
    if (out_num_iters) *out_num_iters = tr.num_iters;
    if (out_depth) *out_depth = tr.depth;
    if (out_max_offset) *out_max_offset = tr.max_offset_abs;
    return rc;
}

// bvh2_to_cwbvh(bvh2, max_prims_per_leaf, order_children, include_exact_node_aabbs = false)
// (bvh2_to_cwbvh.rs:490-510).  out_nodes has capacity `cap`; out_prim_indices receives the CWBVH primitive
// index list (same length as the BVH2 primitive list for non-split builds).
int orc_bvh2_to_cwbvh(const Bvh2Node* nodes, uint64_t nnodes, const uint32_t* prim_indices, uint64_t nprims,
                      uint32_t max_prims_per_leaf, int order_children, CwBvhNode* out_nodes, uint64_t cap,
                      uint32_t* out_prim_indices, uint64_t prim_cap, uint64_t* out_num_nodes,
                      uint64_t* out_num_prims, uint32_t* out_cw_to_bvh2) {
    *out_num_nodes = 0;
    *out_num_prims = 0;
    if (nnodes == 0) return ORC_OK;
    Converter cv;
    cv.nodes = nodes;
    cv.nnodes = (size_t)nnodes;
    cv.bvh2_prim_indices = prim_indices;
    cv.order_children = order_children != 0;
    for (int s = 0; s < 8; ++s)  // bvh2_to_cwbvh.rs:40-50
        cv.lut[s] = v3(((s & 4) != 0) != (g_mut == MUT_CW_LUT_FLIP_X) ? -1.0f : 1.0f, (s & 2) ? -1.0f : 1.0f,
                       (s & 1) ? -1.0f : 1.0f);
    cv.decisions.assign((size_t)nnodes * 7, Decision{});
    cv.out.reserve((size_t)nprims);
    cv.out.push_back(CwBvhNode{});
    cv.cw_to_bvh2.push_back(INVALID32);
    cv.prims.reserve((size_t)nprims);
    (void)nprims;
    cv.calculate_cost(max_prims_per_leaf);
    if (cv.err) return cv.err;
    cv.convert(0, 0);
    if (cv.err) return cv.err;
    if (cv.out.size() > cap || cv.prims.size() > prim_cap) return ORC_ERR_CAPACITY;
    std::memcpy(out_nodes, cv.out.data(), cv.out.size() * sizeof(CwBvhNode));
    std::memcpy(out_prim_indices, cv.prims.data(), cv.prims.size() * sizeof(uint32_t));
    if (out_cw_to_bvh2) std::memcpy(out_cw_to_bvh2, cv.cw_to_bvh2.data(), cv.out.size() * sizeof(uint32_t));
    *out_num_nodes = cv.out.size();
    *out_num_prims = cv.prims.size();
    return ORC_OK;
}

// The conversion's cost table (calculate_cost, bvh2_to_cwbvh.rs:220-344) for each BVH2 node and i = 0..6:
// cost, kind (0 LEAF, 1 INTERNAL, 2 DISTRIBUTE), distribute_left, distribute_right; arrays of 7 * nnodes.
int orc_cw_cost_table(const Bvh2Node* nodes, uint64_t nnodes, uint32_t max_prims_per_leaf, float* costs,
                      uint8_t* kinds, uint8_t* dist_left, uint8_t* dist_right) {
    if (nnodes == 0) return ORC_OK;
    Converter cv;
    cv.nodes = nodes;
    cv.nnodes = (size_t)nnodes;
    cv.decisions.assign((size_t)nnodes * 7, Decision{});
    cv.calculate_cost(max_prims_per_leaf);
    if (cv.err) return cv.err;
    for (size_t i = 0; i < (size_t)nnodes * 7; ++i) {
        costs[i] = cv.decisions[i].cost;
        kinds[i] = (uint8_t)cv.decisions[i].kind;
        dist_left[i] = cv.decisions[i].distribute_left;
        dist_right[i] = cv.decisions[i].distribute_right;
    }
    return ORC_OK;
}

// Pre-split leaf set of build_cwbvh_from_tris with pre_split = true (cwbvh/builder.rs:26-55): out_aabbs and
// out_indices have capacity `cap`; max_total = 0 follows the reference (no limit).
int orc_presplit(const Triangle* tris, uint64_t n, uint64_t max_total, Aabb* out_aabbs, uint32_t* out_indices,
                 uint64_t cap, uint64_t* out_count, float* out_avg, float* out_largest) {
    std::vector<Aabb> aabbs;
    std::vector<uint32_t> indices;
    *out_count = 0;
    if (n == 0) return ORC_OK;
    presplit_leaves(tris, (size_t)n, (size_t)max_total, aabbs, indices, out_avg, out_largest);
    if (aabbs.size() > cap) return ORC_ERR_CAPACITY;
    std::memcpy(out_aabbs, aabbs.data(), aabbs.size() * sizeof(Aabb));
    std::memcpy(out_indices, indices.data(), indices.size() * 4);
    *out_count = aabbs.size();
    return ORC_OK;
}

// split_triangle (splits.rs:137-158) for one triangle and plane.
void orc_split_triangle(const Triangle* tri, uint32_t dim, float pos, Aabb* out_left, Aabb* out_right) {
    const V3 v[4] = {tri->v0, tri->v1, tri->v2, tri->v0};
    split_triangle(dim, pos, v, *out_left, *out_right);
}

// ReinsertionOptimizer::default().run(bvh, ratio, seq) in place on Bvh2Node[nnodes] (seq = NULL: default sequence).
void orc_reinsertion(Bvh2Node* nodes, uint64_t nnodes, float ratio, const float* seq, uint64_t seq_len) {
    std::vector<Bvh2Node> v(nodes, nodes + nnodes);
    reinsertion_run(v, ratio, seq, (size_t)seq_len);
    std::memcpy(nodes, v.data(), nnodes * sizeof(Bvh2Node));
}

// find_reinsertion (reinsertion.rs:247-340) for one node of a Bvh2 (parents computed here).
void orc_find_reinsertion(const Bvh2Node* nodes, uint64_t nnodes, uint32_t node_id, uint32_t* out_from,
                          uint32_t* out_to, float* out_area_diff) {
    std::vector<Bvh2Node> v(nodes, nodes + nnodes);
    std::vector<uint32_t> parents;
    compute_parents(v, parents);
    std::vector<std::pair<float, uint32_t>> stack;
    const Reinsertion r = find_reinsertion(v, parents, node_id, stack);
    *out_from = r.from;
    *out_to = r.to;
    *out_area_diff = r.area_diff;
}

// reinsert_node (reinsertion.rs:342-382) in place (parents computed here).
void orc_reinsert_node(Bvh2Node* nodes, uint64_t nnodes, uint32_t from, uint32_t to) {
    std::vector<Bvh2Node> v(nodes, nodes + nnodes);
    std::vector<uint32_t> parents;
    compute_parents(v, parents);
    reinsert_node(v, parents, from, to);
    std::memcpy(nodes, v.data(), nnodes * sizeof(Bvh2Node));
}

// compute_rebuild_path_flags (ploc/rebuild.rs:12-42); flags has nnodes bytes.
int orc_rebuild_path_flags(const Bvh2Node* nodes, uint64_t nnodes, const uint32_t* leaves, uint64_t n_leaves,
                           uint8_t* flags) {
    std::vector<Bvh2Node> v(nodes, nodes + nnodes);
    return rebuild_path_flags(v, leaves, (size_t)n_leaves, flags);
}

// PlocBuilder::partial_rebuild(bvh, |i| flags[i] != 0, search_distance, sort_precision (64 or 128),
// search_depth_threshold) in place.
int orc_partial_rebuild(Bvh2Node* nodes, uint64_t nnodes, const uint8_t* flags, uint32_t search_distance,
                        uint32_t search_depth_threshold, uint32_t* out_depth, uint32_t sort_precision) {
    if (sort_precision != 64u && sort_precision != 128u) return ORC_ERR_INVALID_ARG;
    std::vector<Bvh2Node> v(nodes, nodes + nnodes);
    size_t depth = 0;
    int rc = partial_rebuild(v, flags, search_distance, sort_precision == 128u, search_depth_threshold, &depth);
    std::memcpy(nodes, v.data(), nnodes * sizeof(Bvh2Node));
    if (out_depth) *out_depth = (uint32_t)depth;
    return rc;
}

// PlocBuilder::full_rebuild(bvh, search_distance, sort_precision (64 or 128), search_depth_threshold) in place;
// *out_nnodes receives the new node count (2k-1 for k leaves, equal to nnodes for a well-formed tree).
int orc_full_rebuild(Bvh2Node* nodes, uint64_t nnodes, uint32_t search_distance, uint32_t search_depth_threshold,
                     uint64_t* out_nnodes, uint32_t* out_depth, uint32_t sort_precision) {
    if (sort_precision != 64u && sort_precision != 128u) return ORC_ERR_INVALID_ARG;
    std::vector<Bvh2Node> v(nodes, nodes + nnodes);
    size_t depth = 0;
    int rc = full_rebuild(v, search_distance, sort_precision == 128u, search_depth_threshold, &depth);
    if (v.size() > nnodes) return ORC_ERR_CAPACITY;
    std::memcpy(nodes, v.data(), v.size() * sizeof(Bvh2Node));
    if (out_nnodes) *out_nnodes = v.size();
    if (out_depth) *out_depth = (uint32_t)depth;
    return rc;
}

int orc_primitives_to_nodes(const Bvh2Node* nodes, uint64_t nnodes, const uint32_t* prim_indices, uint64_t nprims,
                            uint32_t* out) {
    return primitives_to_nodes(nodes, (size_t)nnodes, prim_indices, (size_t)nprims, out);
}

// Bvh2::refit_all (bvh2/mod.rs:527-569) in place.
void orc_refit_all(Bvh2Node* nodes, uint64_t nnodes) {
    std::vector<Bvh2Node> v(nodes, nodes + nnodes);
    refit_all(v);
    std::memcpy(nodes, v.data(), nnodes * sizeof(Bvh2Node));
}

// Closest-hit (mode 0) or miss (mode 1) traversal of every ray.  `tris` is indexed by primitive_indices
// position (the reference's examples reorder triangles this way: obj_cwbvh.rs:63-67).
// Optional per-ray visit counters (node tests and triangle tests).  threads <= 1 runs serially.
void orc_cwbvh_traverse(const CwBvhNode* nodes, uint64_t nnodes, const Triangle* tris, const Ray* rays,
                        uint64_t nrays, int mode, RayHit* out_hits, uint8_t* out_miss, uint32_t* out_node_visits,
                        uint32_t* out_tri_tests, int threads) {
    auto work = [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            TravStats st;
            traverse_one(nodes, (uint32_t)nnodes, tris, rays[i], mode, out_hits ? &out_hits[i] : nullptr,
                         out_miss ? &out_miss[i] : nullptr, &st);
            if (out_node_visits) out_node_visits[i] = st.nodes;
            if (out_tri_tests) out_tri_tests[i] = st.tris;
        }
    };
    if (threads <= 1 || nrays < 1024) {
        work(0, nrays);
        return;
    }
    // Dynamic chunking like rayon's work stealing over rays (examples/obj_cwbvh.rs:88-94).
    std::atomic<uint64_t> next{0};
    const uint64_t chunk = 256;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        pool.emplace_back([&]() {
            for (;;) {
                uint64_t b = next.fetch_add(chunk);
                if (b >= nrays) break;
                work(b, std::min(nrays, b + chunk));
            }
        });
    }
    for (auto& th : pool) th.join();
}

// ray_traverse_tlas_blas for every ray.  BLAS i: nodes[blas_node_offsets[i] .. + blas_node_counts[i]), triangles
// tris[blas_tri_offsets[i] ..] in that BLAS's primitive_indices order.
void orc_tlas_traverse(const CwBvhNode* tlas, uint64_t tlas_nnodes, const uint32_t* tlas_prims,
                       const CwBvhNode* blas_nodes, const uint32_t* blas_node_offsets,
                       const uint32_t* blas_node_counts, const Triangle* tris, const uint32_t* blas_tri_offsets,
                       uint32_t nblas, const Ray* rays, uint64_t nrays, RayHit* out) {
    std::vector<BlasRef> refs(nblas);
    for (uint32_t i = 0; i < nblas; ++i)
        refs[i] = BlasRef{blas_nodes + blas_node_offsets[i], blas_node_counts[i], tris + blas_tri_offsets[i]};
    for (uint64_t r = 0; r < nrays; ++r)
        traverse_tlas_one(tlas, (uint32_t)tlas_nnodes, tlas_prims, refs.data(), nblas, rays[r], &out[r]);
}

// Brute force: min t over all triangles with Triangle::intersect; ties keep the lowest index.
void orc_brute_force(const Triangle* tris, uint64_t ntris, const Ray* rays, uint64_t nrays, RayHit* out) {
    for (uint64_t r = 0; r < nrays; ++r) {
        RayHit h{INVALID32, INVALID32, INVALID32, INFINITY};
        float tmax = rays[r].tmax;
        for (uint64_t i = 0; i < ntris; ++i) {
            float t = tri_intersect(tris[i], rays[r], rays[r].tmin, tmax);
            if (t < h.t) {
                h.t = t;
                h.primitive_id = (uint32_t)i;
            }
        }
        out[r] = h;
    }
}

// Brute force with the watertight test: the closest t over all triangles (first index on equal t).
void orc_brute_force_wt(const Triangle* tris, uint64_t ntris, const Ray* rays, uint64_t nrays, RayHit* out) {
    for (uint64_t r = 0; r < nrays; ++r) {
        RayHit h{INVALID32, INVALID32, INVALID32, INFINITY};
        const WtRay w = wt_ray(rays[r]);
        for (uint64_t i = 0; i < ntris; ++i) {
            float t = tri_intersect_wt(tris[i], rays[r], w, rays[r].tmin, rays[r].tmax);
            if (t < h.t) {
                h.t = t;
                h.primitive_id = (uint32_t)i;
            }
        }
        out[r] = h;
    }
}

// The watertight test alone, for unit tests.
float orc_tri_intersect_wt(const Triangle* tri, const Ray* ray) {
    return tri_intersect_wt(*tri, *ray, wt_ray(*ray), ray->tmin, ray->tmax);
}

// The raw ray-triangle test (triangle.rs:35-76) for unit tests.
float orc_tri_intersect(const Triangle* tri, const Ray* ray) { return tri_intersect(*tri, *ray, ray->tmin, ray->tmax); }

// The raw node test (cwbvh/simd.rs:17-100) for unit tests.
uint32_t orc_intersect_node(const CwBvhNode* n, const Ray* ray, uint32_t oct_inv4) {
    return intersect_node(*n, *ray, ray->tmax, oct_inv4);
}
uint32_t orc_ray_octant_inv4(const Ray* ray) { return ray_get_octant_inv4(ray->direction); }

// ----- validation (bvh2/mod.rs:786-981; cwbvh/mod.rs:747-908), restated as checks returning a failure code -----

// Returns 0 when valid; otherwise a positive code naming the first failed check.  flags: bit 0 = tight_fit as
// the reference; bit 1 = children need not be stored after parents (after reinsertion, reinsertion.rs:102);
// bit 2 = spatial splits (a primitive may sit in several leaves whose boxes are parts of its box:
// bvh2/mod.rs:823, 928, 945 skip those checks).  Also reports the max depth.
int orc_bvh2_validate(const Bvh2Node* nodes, uint64_t nnodes, const uint32_t* prim_indices, uint64_t nprims,
                      const Aabb* prim_aabbs, int flags, uint32_t* out_max_depth) {
    if (nnodes == 0) return 0;
    const bool tight_fit = (flags & 1) != 0, unordered = (flags & 2) != 0, splits = (flags & 4) != 0;
    std::vector<uint8_t> seen_node(nnodes, 0);
    std::vector<uint8_t> seen_prim(nprims, 0);
    uint64_t node_count = 0, prim_count = 0;
    uint32_t max_depth = 0;
    struct Item {
        uint32_t node, parent, depth;
    };
    std::vector<Item> st{{0, 0, 0}};
    std::vector<uint32_t> parents(nnodes, 0);
    for (uint64_t i = 0; i < nnodes; ++i)
        if (nodes[i].prim_count == 0) {
            if ((uint64_t)nodes[i].first_index + 1 >= nnodes) return 10;
            parents[nodes[i].first_index] = (uint32_t)i;
            parents[nodes[i].first_index + 1] = (uint32_t)i;
        }
    while (!st.empty()) {
        Item it = st.back();
        st.pop_back();
        if (it.node >= nnodes) return 11;
        if (seen_node[it.node]) return 12;
        seen_node[it.node] = 1;
        node_count++;
        max_depth = std::max(max_depth, it.depth);
        const Bvh2Node& nd = nodes[it.node];
        const Aabb& pa = nodes[it.parent].aabb;
        if (!(nd.aabb.min.x >= pa.min.x && nd.aabb.min.y >= pa.min.y && nd.aabb.min.z >= pa.min.z &&
              nd.aabb.max.x <= pa.max.x && nd.aabb.max.y <= pa.max.y && nd.aabb.max.z <= pa.max.z))
            return 13;  // bvh2/mod.rs:905-913 child fits in parent
        if (nd.prim_count != 0) {
            Aabb tmp = aabb_empty();
            for (uint32_t i = 0; i < nd.prim_count; ++i) {
                uint64_t slot = (uint64_t)nd.first_index + i;
                if (slot >= nprims) return 14;
                uint32_t p = prim_indices[slot];
                if (p >= nprims) return 15;
                prim_count++;
                if (splits) continue;
                if (seen_prim[p]) return 16;
                seen_prim[p] = 1;
                const Aabb& pb = prim_aabbs[p];
                tmp = aabb_union(tmp, pb);
                if (!(pb.min.x >= nd.aabb.min.x && pb.min.y >= nd.aabb.min.y && pb.min.z >= nd.aabb.min.z &&
                      pb.max.x <= nd.aabb.max.x && pb.max.y <= nd.aabb.max.y && pb.max.z <= nd.aabb.max.z))
                    return 17;  // bvh2/mod.rs:944-954
            }
            if (tight_fit && !splits && !aabb_eq(tmp, nd.aabb)) return 18;  // bvh2/mod.rs:957-963
        } else {
            if (tight_fit) {
                Aabb u = aabb_union(nodes[nd.first_index].aabb, nodes[nd.first_index + 1].aabb);
                if (!aabb_eq(u, nd.aabb)) return 19;  // bvh2/mod.rs:965-976
            }
            st.push_back({nd.first_index + 1, it.node, it.depth + 1});
            st.push_back({nd.first_index, it.node, it.depth + 1});
        }
    }
    if (node_count != nnodes) return 20;  // bvh2/mod.rs:816-817
    if (prim_count != nprims) return 21;  // bvh2/mod.rs:825-827
    if (!unordered)
        for (uint64_t i = 1; i < nnodes; ++i)
            if (!(parents[i] < i)) return 22;  // bvh2/mod.rs:861-874 children ordered after parents
    if (out_max_depth) *out_max_depth = max_depth;
    return 0;
}

// cwbvh/mod.rs:747-908.  prim_aabbs are in the original primitive order (direct_layout = false).
int orc_cwbvh_validate(const CwBvhNode* nodes, uint64_t nnodes, const uint32_t* prim_indices, uint64_t nprims,
                       const Aabb* prim_aabbs, uint32_t* out_max_depth) {
    if (nnodes == 0) return 0;
    std::vector<uint8_t> seen_node(nnodes, 0);
    std::vector<uint8_t> seen_prim(nprims, 0);
    uint64_t discovered_nodes = 0, discovered_prims = 0;
    uint32_t max_depth = 0;
    const float L = 3.40282347e+38f;
    struct Item {
        uint32_t node;
        Aabb bounds;
        uint32_t depth;
    };
    std::vector<Item> st;
    st.push_back({0, Aabb{v3(-L, -L, -L), v3(L, L, L)}, 0});
    while (!st.empty()) {
        Item it = st.back();
        st.pop_back();
        if (it.node >= nnodes) return 30;
        if (seen_node[it.node]) return 31;
        seen_node[it.node] = 1;
        discovered_nodes++;
        max_depth = std::max(max_depth, it.depth);
        const CwBvhNode& n = nodes[it.node];
        const Aabb& pb = it.bounds;
        if (!(std::isfinite(n.p[0]) && std::isfinite(n.p[1]) && std::isfinite(n.p[2]))) return 32;
        if (!(n.p[0] >= pb.min.x - 1.0e-5f && n.p[1] >= pb.min.y - 1.0e-5f && n.p[2] >= pb.min.z - 1.0e-5f))
            return 33;
        if (!(n.p[0] <= pb.max.x + 1.0e-5f && n.p[1] <= pb.max.y + 1.0e-5f && n.p[2] <= pb.max.z + 1.0e-5f))
            return 34;
        float e[3] = {bits_to_f32((uint32_t)n.e[0] << 23), bits_to_f32((uint32_t)n.e[1] << 23),
                      bits_to_f32((uint32_t)n.e[2] << 23)};
        for (int ch = 0; ch < 8; ++ch) {
            uint8_t meta = n.child_meta[ch];
            if (meta == 0) continue;
            float qmin[3] = {(float)n.child_min_x[ch] * e[0] + n.p[0], (float)n.child_min_y[ch] * e[1] + n.p[1],
                             (float)n.child_min_z[ch] * e[2] + n.p[2]};
            float qmax[3] = {(float)n.child_max_x[ch] * e[0] + n.p[0], (float)n.child_max_y[ch] * e[1] + n.p[1],
                             (float)n.child_max_z[ch] * e[2] + n.p[2]};
            bool inner = (n.imask & (1u << ch)) != 0;
            if (inner != ((meta & 0x1f) >= 24)) return 35;
            if (inner) {
                uint32_t slot_index = (uint32_t)(meta & 0x1f) - 24;
                uint32_t rel = (uint32_t)__builtin_popcount((uint32_t)n.imask & ~(0xffffffffu << slot_index));
                uint32_t child = n.child_base_idx + rel;
                Aabb q{v3(qmin[0], qmin[1], qmin[2]), v3(qmax[0], qmax[1], qmax[2])};
                // Aabb::intersection (aabb.rs:98-103)
                Aabb b{vmax(q.min, pb.min), vmin(q.max, pb.max)};
                st.push_back({child, b, it.depth + 1});
            } else {
                uint32_t first = n.primitive_base_idx + (meta & 0x1f);
                uint32_t cnt = 0;
                for (uint32_t i = 0; i < 3; ++i) {
                    if ((meta & (0x20u << i)) == 0) continue;
                    cnt++;
                    uint32_t slot = first + i;
                    if (slot >= nprims) return 36;
                    if (seen_prim[slot]) return 37;
                    seen_prim[slot] = 1;
                    discovered_prims++;
                    uint32_t p = prim_indices[slot];
                    const Aabb& a = prim_aabbs[p];
                    if (!(a.min.x >= pb.min.x - 1.0e-5f && a.min.y >= pb.min.y - 1.0e-5f &&
                          a.min.z >= pb.min.z - 1.0e-5f && a.max.x <= pb.max.x + 1.0e-5f &&
                          a.max.y <= pb.max.y + 1.0e-5f && a.max.z <= pb.max.z + 1.0e-5f))
                        return 38;  // cwbvh/mod.rs:884-893
                }
                if (cnt != (uint32_t)__builtin_popcount((uint32_t)(meta & 0xe0))) return 39;
            }
        }
    }
    if (discovered_nodes != nnodes) return 40;  // cwbvh/mod.rs:786
    if (discovered_prims != nprims) return 41;  // cwbvh/mod.rs:787-790
    if (out_max_depth) *out_max_depth = max_depth;
    if (max_depth >= 32) return 42;  // cwbvh/mod.rs:791
    return 0;
}

// Wall-clock helper for the CPU baseline (seconds, monotonic).
double orc_now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // extern "C"

extern "C" {
void orc_set_mutation(int m) { g_mut = m; }
// Threads of the multithreaded build (1: the serial restatement).
void orc_set_threads(int t) { g_threads = t < 1 ? 1 : t; }
// Host libm log2f (the function the reference's f32::log2 calls on linux-gnu), for the exponent-table tests.
float orc_log2f(float x) { return log2f(x); }
}
