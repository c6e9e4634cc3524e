// Runs every stage of the CPU oracle (oracle/obvhs_oracle.cpp) once on a small seeded scene, for a build of the
// oracle under AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_oracle_sanitize.py).  Prints "ok" and
// exits 0 when every call succeeds and every produced tree validates; a sanitizer report aborts the process.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

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
    float tmin, tmax, pad0, pad1;
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
    uint8_t bytes[80];
};
static_assert(sizeof(Triangle) == 48 && sizeof(Ray) == 64 && sizeof(Bvh2Node) == 48, "oracle layouts");

extern "C" {
void orc_tri_aabbs(const Triangle* tris, uint64_t n, Aabb* out);
int orc_ploc_build(const Aabb* aabbs, const uint32_t* indices, uint64_t n, uint32_t search_distance,
                   uint32_t search_depth_threshold, int basic_search, Bvh2Node* out_nodes, uint32_t* out_prim_indices,
                   uint64_t* out_codes, uint32_t* out_sorted_index, uint32_t* out_iter_counts, uint32_t iter_cap,
                   uint32_t* out_num_iters, uint32_t* out_depth, uint32_t* out_max_offset, uint32_t sort_precision,
                   uint64_t* out_codes_hi);
int orc_bvh2_to_cwbvh(const Bvh2Node* nodes, uint64_t nnodes, const uint32_t* prim_indices, uint64_t nprims,
                      uint32_t max_prims_per_leaf, int order_children, CwBvhNode* out_nodes, uint64_t cap,
                      uint32_t* out_prim_indices, uint64_t prim_cap, uint64_t* out_num_nodes, uint64_t* out_num_prims,
                      uint32_t* out_cw_to_bvh2);
int orc_presplit(const Triangle* tris, uint64_t n, uint64_t max_total, Aabb* out_aabbs, uint32_t* out_indices,
                 uint64_t cap, uint64_t* out_count, float* out_avg, float* out_largest);
void orc_reinsertion(Bvh2Node* nodes, uint64_t nnodes, float ratio, const float* seq, uint64_t seq_len);
int orc_rebuild_path_flags(const Bvh2Node* nodes, uint64_t nnodes, const uint32_t* leaves, uint64_t n_leaves,
                           uint8_t* flags);
int orc_partial_rebuild(Bvh2Node* nodes, uint64_t nnodes, const uint8_t* flags, uint32_t search_distance,
                        uint32_t search_depth_threshold, uint32_t* out_depth, uint32_t sort_precision);
int orc_full_rebuild(Bvh2Node* nodes, uint64_t nnodes, uint32_t search_distance, uint32_t search_depth_threshold,
                     uint64_t* out_nnodes, uint32_t* out_depth, uint32_t sort_precision);
int orc_primitives_to_nodes(const Bvh2Node* nodes, uint64_t nnodes, const uint32_t* prim_indices, uint64_t nprims,
                            uint32_t* out);
void orc_refit_all(Bvh2Node* nodes, uint64_t nnodes);
void orc_cwbvh_traverse(const CwBvhNode* nodes, uint64_t nnodes, const Triangle* tris, const Ray* rays, uint64_t nrays,
                        int mode, RayHit* out_hits, uint8_t* out_miss, uint32_t* out_node_visits,
                        uint32_t* out_tri_tests, int threads);
void orc_brute_force(const Triangle* tris, uint64_t ntris, const Ray* rays, uint64_t nrays, RayHit* out);
int orc_bvh2_validate(const Bvh2Node* nodes, uint64_t nnodes, const uint32_t* prim_indices, uint64_t nprims,
                      const Aabb* prim_aabbs, int flags, uint32_t* out_max_depth);
int orc_cwbvh_validate(const CwBvhNode* nodes, uint64_t nnodes, const uint32_t* prim_indices, uint64_t nprims,
                       const Aabb* prim_aabbs, uint32_t* out_max_depth);
}

namespace {

int failures = 0;

void check(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAILED: %s\n", what);
        ++failures;
    }
}

// Integer hash mapped to [0, 1) (a seeded scene, not the test generators' stream).
float unit(uint32_t i, uint32_t salt) {
    uint32_t x = i * 0x9E3779B1u ^ salt * 0x85EBCA77u;
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return (float)(x >> 8) * (1.0f / 16777216.0f);
}

struct Tree {
    std::vector<Bvh2Node> nodes;
    std::vector<uint32_t> prims;
};

Tree build(const std::vector<Aabb>& boxes, const std::vector<uint32_t>& idx, uint32_t sd, uint32_t precision) {
    const size_t n = boxes.size();
    Tree t;
    t.nodes.resize(2 * n - 1);
    t.prims.resize(n);
    std::vector<uint64_t> codes(n), codes_hi(n);
    std::vector<uint32_t> order(n), iters(4096);
    uint32_t ni = 0, depth = 0, maxoff = 0;
    const int rc = orc_ploc_build(boxes.data(), idx.data(), n, sd, 0, 0, t.nodes.data(), t.prims.data(),
                                  codes.data(), order.data(), iters.data(), (uint32_t)iters.size(), &ni, &depth,
                                  &maxoff, precision, codes_hi.data());
    check(rc == 0, "ploc build");
    return t;
}

}  // namespace

int main() {
    const uint32_t n = 3000;
    std::vector<Triangle> tris(n + 40);
    for (uint32_t i = 0; i < n; ++i) {  // a soup of small triangles
        const float cx = unit(i, 1), cy = unit(i, 2), cz = unit(i, 3);
        V3 v[3];
        for (int k = 0; k < 3; ++k)
            v[k] = V3{cx + (unit(i, 4 + 3 * k) - 0.5f) * 0.02f, cy + (unit(i, 5 + 3 * k) - 0.5f) * 0.02f,
                      cz + (unit(i, 6 + 3 * k) - 0.5f) * 0.02f, 0.0f};
        tris[i] = Triangle{v[0], v[1], v[2]};
    }
    for (uint32_t i = 0; i < 40; ++i) {  // long thin triangles for the splits
        const float y = unit(i, 20), z = unit(i, 21);
        tris[n + i] = Triangle{V3{0.0f, y, z, 0.0f}, V3{1.0f, y + 0.001f, z, 0.0f}, V3{0.5f, y, z + 0.002f, 0.0f}};
    }
    const uint64_t nt = tris.size();
    std::vector<Aabb> boxes(nt);
    orc_tri_aabbs(tris.data(), nt, boxes.data());
    std::vector<uint32_t> idx(nt);
    for (uint32_t i = 0; i < nt; ++i) idx[i] = i;

    // PLOC with both key widths
    Tree a = build(boxes, idx, 14, 64), b = build(boxes, idx, 1, 128);
    uint32_t md = 0;
    check(orc_bvh2_validate(a.nodes.data(), a.nodes.size(), a.prims.data(), nt, boxes.data(), 1, &md) == 0,
          "validate U64 build");
    check(orc_bvh2_validate(b.nodes.data(), b.nodes.size(), b.prims.data(), nt, boxes.data(), 1, &md) == 0,
          "validate U128 build");

    // conversion and traversal
    std::vector<CwBvhNode> cw(a.nodes.size());
    std::vector<uint32_t> cwp(nt), cw2b(a.nodes.size());
    uint64_t ncw = 0, ncp = 0;
    check(orc_bvh2_to_cwbvh(a.nodes.data(), a.nodes.size(), a.prims.data(), nt, 3, 1, cw.data(), cw.size(),
                            cwp.data(), cwp.size(), &ncw, &ncp, cw2b.data()) == 0,
          "bvh2_to_cwbvh");
    check(orc_cwbvh_validate(cw.data(), ncw, cwp.data(), ncp, boxes.data(), &md) == 0, "validate CWBVH");
    std::vector<Triangle> bvh_tris(ncp);
    for (uint64_t i = 0; i < ncp; ++i) bvh_tris[i] = tris[cwp[i]];
    const uint32_t nr = 256;
    std::vector<Ray> rays(nr);
    for (uint32_t i = 0; i < nr; ++i) {
        const float dx = unit(i, 30) - 0.5f, dy = unit(i, 31) - 0.5f, dz = 1.0f;
        const float l = 1.0f / std::sqrt(dx * dx + dy * dy + dz * dz);
        const V3 d{dx * l, dy * l, dz * l, 0.0f};
        rays[i] = Ray{V3{0.5f, 0.5f, -1.0f, 0.0f}, d, V3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z, 0.0f}, 0.0f, 1e30f,
                      0.0f, 0.0f};
    }
    std::vector<RayHit> hits(nr), bf(nr);
    std::vector<uint8_t> miss(nr);
    std::vector<uint32_t> nv(nr), tv(nr);
    orc_cwbvh_traverse(cw.data(), ncw, bvh_tris.data(), rays.data(), nr, 0, hits.data(), nullptr, nv.data(),
                       tv.data(), 2);
    orc_cwbvh_traverse(cw.data(), ncw, bvh_tris.data(), rays.data(), nr, 1, nullptr, miss.data(), nullptr, nullptr,
                       2);
    orc_brute_force(tris.data(), nt, rays.data(), nr, bf.data());
    for (uint32_t i = 0; i < nr; ++i) {
        check((hits[i].primitive_id == 0xffffffffu) == (bf[i].primitive_id == 0xffffffffu), "hit vs brute force");
        check((miss[i] != 0) == (bf[i].primitive_id == 0xffffffffu), "occlusion vs brute force");
    }

    // pre-splits, then a build over the split leaf set
    const uint64_t cap = nt * 13 / 10;
    std::vector<Aabb> sboxes(cap);
    std::vector<uint32_t> sidx(cap);
    uint64_t ns = 0;
    float avg = 0.0f, largest = 0.0f;
    check(orc_presplit(tris.data(), nt, cap, sboxes.data(), sidx.data(), cap, &ns, &avg, &largest) == 0, "presplit");
    sboxes.resize(ns);
    sidx.resize(ns);
    Tree s = build(sboxes, sidx, 14, 64);
    check(orc_bvh2_validate(s.nodes.data(), s.nodes.size(), s.prims.data(), nt, boxes.data(), 5, &md) == 0,
          "validate split build");

    // reinsertion, refit, leaf map, rebuilds
    const float seq[2] = {1.0f, 0.5f};
    orc_reinsertion(a.nodes.data(), a.nodes.size(), 0.1f, seq, 2);
    check(orc_bvh2_validate(a.nodes.data(), a.nodes.size(), a.prims.data(), nt, boxes.data(), 3, &md) == 0,
          "validate after reinsertion");
    orc_refit_all(a.nodes.data(), a.nodes.size());
    std::vector<uint32_t> p2n(nt);
    check(orc_primitives_to_nodes(a.nodes.data(), a.nodes.size(), a.prims.data(), nt, p2n.data()) == 0,
          "primitives_to_nodes");
    std::vector<uint32_t> leaves;
    for (uint32_t p = 0; p < nt; p += 7) leaves.push_back(p2n[p]);
    std::vector<uint8_t> flags(a.nodes.size());
    check(orc_rebuild_path_flags(a.nodes.data(), a.nodes.size(), leaves.data(), leaves.size(), flags.data()) == 0,
          "rebuild path flags");
    uint32_t depth = 0;
    check(orc_partial_rebuild(a.nodes.data(), a.nodes.size(), flags.data(), 14, 0, &depth, 128) == 0,
          "partial rebuild");
    check(orc_bvh2_validate(a.nodes.data(), a.nodes.size(), a.prims.data(), nt, boxes.data(), 3, &md) == 0,
          "validate after partial rebuild");
    uint64_t nn = 0;
    check(orc_full_rebuild(a.nodes.data(), a.nodes.size(), 6, 2, &nn, &depth, 64) == 0, "full rebuild");
    check(orc_bvh2_validate(a.nodes.data(), nn, a.prims.data(), nt, boxes.data(), 3, &md) == 0,
          "validate after full rebuild");

    if (failures) return 1;
    std::printf("ok\n");
    return 0;
}
