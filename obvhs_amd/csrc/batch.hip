// Batched build and traversal: many independent meshes of uneven size in shared launches.
//
// Every mesh gets exactly the CWBVH build_cwbvh_from_tris would give it alone (cwbvh/builder.rs:61-74): the
// kernels are the single-mesh ones (ploc.hip, cwbvh.hip) made segment-aware.
//   * leaf init / Morton: one pass over the concatenated triangles; each triangle finds its mesh by binary search
//     over the offsets; per-mesh scene bounds by wave-reduced atomics (ploc/mod.rs:187-243, 287-288, 782-809).
//   * sort: stable radix sort by Morton code, then stable radix sort by mesh id -- every mesh's leaves end up in
//     its own segment in (code, leaf index) order, exactly the single-mesh order (ploc/mod.rs:811-846).
//   * PLOC passes over all segments at once: the neighbour window is clipped to the cluster's segment
//     (ploc/mod.rs:508-528), prefix sums run over the whole array, and a per-mesh pass computes each segment's
//     new start, count, insert index and merge base so that child slots are insert_m - 2(s_m + 1) with s_m the
//     merge's rank inside its mesh (ploc/mod.rs:420-487).  Finished meshes keep one cluster that is copied along.
//     The cost table is evaluated by the merges (cwdp.hpp) with the mesh's key base.
//   * CWBVH emit: one launch per level over all meshes (cwbvh.hip).
//   * traversal: one lane per ray, the ray's mesh selects the CWBVH and triangle slices.
#include <string.h>

#include <algorithm>
#include <new>
#include <vector>

#include "common.hpp"
#include "cwdp.hpp"
#include "device_util.hpp"
#include "internal.hpp"

struct obvhs_batch {
    obvhs::DevBuf tri_offsets;    // u32[m+1]
    obvhs::DevBuf mesh_table;     // MeshDesc[m]
    obvhs::DevBuf bvh2_nodes;     // Node32[2 * n_total], mesh i at 2 * offset[i]
    obvhs::DevBuf bvh2_prims;     // u32[n_total] (mesh-local 0..n_i)
    obvhs::DevBuf dp;             // cost table, keys prim_base + first_index / 2
    obvhs::DevBuf cw_nodes;       // CwNode[total]
    obvhs::DevBuf cw_prims;       // u32[n_total] (mesh-local indices)
    std::vector<obvhs::MeshDesc> meshes;
    uint32_t n_total = 0;
    uint32_t cw_total = 0;
    uint32_t passes = 0;
};

namespace obvhs {

namespace {

constexpr int T256 = 256;
constexpr uint32_t MAX_RADIUS = 32;
// Passes between host reads; 8 and 16 measured slower here (16.1 -> 16.8 and 18.3 ms): grids sized at the last read.
constexpr int PASSES_PER_SYNC = 4;

struct SegState {
    uint32_t start;   // first cluster of the mesh in the current cluster array
    uint32_t count;   // clusters of the mesh
    uint32_t insert;  // next free child slot + 1 (mesh-local), ploc/mod.rs:283-284
    uint32_t mbase;   // global exclusive merge count at the segment's first cluster (current pass)
};

// The cluster count is double-buffered by pass parity q: pass p reads total[p & 1] and its scan writes the next
// pass's count into total[(p + 1) & 1], so no launch is needed between passes to move it.
struct BatchMisc {
    uint32_t err;
    uint32_t total[2];   // clusters in the current array (by pass parity)
    uint32_t active;     // 1 if a mesh has more than one cluster after the last pass, else 0
    uint32_t totals[2];  // emits, merges of the pass
};

__device__ __forceinline__ uint32_t mesh_in(const uint32_t* __restrict__ offsets, uint32_t lo, uint32_t hi,
                                            uint32_t g) {
    // offsets[lo] <= g < offsets[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (offsets[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t mesh_of(const uint32_t* __restrict__ offsets, uint32_t m, uint32_t g) {
    return mesh_in(offsets, 0u, m, g);
}

__global__ void __launch_bounds__(T256) k_b_init(uint32_t* __restrict__ bounds, const uint32_t* __restrict__ offsets,
                                                 uint32_t m, SegState* __restrict__ seg, BatchMisc* misc,
                                                 uint32_t n_total) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i == 0) {
        misc->err = 0u;
        misc->total[0] = n_total;
        misc->total[1] = 0u;
        misc->active = 0u;
    }
    if (i >= m) return;
    for (int a = 0; a < 3; ++a) {
        bounds[i * 6 + a] = 0xffffffffu;
        bounds[i * 6 + 3 + a] = 0u;
    }
    const uint32_t n = offsets[i + 1] - offsets[i];
    seg[i] = SegState{offsets[i], n, n == 0 ? 0u : 2u * n - 1u, 0u};
}

// Leaf boxes (triangle.rs:28-30) and per-mesh bounds from both corners (ploc/mod.rs:189-190).
// Also records every leaf's mesh (leaf_mesh) and its mesh-local primitive index (prims, the BVH2's
// primitive_indices before the build).  The meshes of a block's triangles lie between those of its first and last
// triangle, found once per block, so each thread searches only that range.  A block inside one mesh reduces its
// bounds in LDS and updates them with one set of atomics (the updates of one mesh serialise on its six words).
constexpr int LEAF_T = 1024;
__global__ void __launch_bounds__(LEAF_T) k_b_leaf_init(const float4* __restrict__ tris,
                                                        const uint32_t* __restrict__ offsets, uint32_t m,
                                                        uint32_t n_total, Node32* __restrict__ leaves,
                                                        uint32_t* __restrict__ bounds, BatchMisc* misc,
                                                        uint32_t* __restrict__ leaf_mesh,
                                                        uint32_t* __restrict__ prims) {
    __shared__ uint32_t s_range[2];
    __shared__ uint32_t s_k[LEAF_T / 64][6];
    const uint32_t g = blockIdx.x * LEAF_T + threadIdx.x;
    const bool valid = g < n_total;
    if (threadIdx.x == 0) s_range[0] = mesh_of(offsets, m, blockIdx.x * LEAF_T);  // the grid has no empty block
    if (threadIdx.x == 64) s_range[1] = mesh_of(offsets, m, min(blockIdx.x * LEAF_T + LEAF_T, n_total) - 1u);
    __syncthreads();
    uint32_t mesh = 0xffffffffu;
    uint32_t k[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    bool nan = false;
    if (valid) {
        mesh = mesh_in(offsets, s_range[0], s_range[1] + 1u, g);
        leaf_mesh[g] = mesh;
        prims[g] = g - offsets[mesh];
        const float4 v0 = tris[3 * (size_t)g], v1 = tris[3 * (size_t)g + 1], v2 = tris[3 * (size_t)g + 2];
        Box b{v0.x, v0.y, v0.z, v0.x, v0.y, v0.z};
        b = Box{smin(b.lx, v1.x), smin(b.ly, v1.y), smin(b.lz, v1.z), smax(b.hx, v1.x), smax(b.hy, v1.y),
                smax(b.hz, v1.z)};
        b = Box{smin(b.lx, v2.x), smin(b.ly, v2.y), smin(b.lz, v2.z), smax(b.hx, v2.x), smax(b.hy, v2.y),
                smax(b.hz, v2.z)};
        nan = box_has_nan(b);
        store_node(&leaves[g], make_node(b, 1u, g - offsets[mesh]));
        const uint32_t kl[3] = {f32_order_key(b.lx), f32_order_key(b.ly), f32_order_key(b.lz)};
        const uint32_t kh[3] = {f32_order_key(b.hx), f32_order_key(b.hy), f32_order_key(b.hz)};
        for (int a = 0; a < 3; ++a) {
            k[a] = min(kl[a], kh[a]);
            k[3 + a] = max(kl[a], kh[a]);
        }
    }
    if (__ballot(nan) != 0ull && __lane_id() == 0) atomicOr(&misc->err, DERR_NAN_INPUT);
    if (s_range[0] == s_range[1]) {  // the block lies in one mesh (block-uniform branch)
        for (int a = 0; a < 3; ++a) {
            k[a] = wave_min_u32(k[a]);
            k[3 + a] = wave_max_u32(k[3 + a]);
        }
        const uint32_t wave = threadIdx.x >> 6;
        if (__lane_id() == 0)
            for (int a = 0; a < 6; ++a) s_k[wave][a] = k[a];
        __syncthreads();
        if (threadIdx.x < 6) {
            const uint32_t a = threadIdx.x;
            uint32_t v = s_k[0][a];
            for (int w = 1; w < LEAF_T / 64; ++w) v = a < 3 ? min(v, s_k[w][a]) : max(v, s_k[w][a]);
            if (a < 3) atomicMin(&bounds[s_range[0] * 6 + a], v);
            else atomicMax(&bounds[s_range[0] * 6 + a], v);
        }
        return;
    }
    // Wave-reduce when the whole wave lies in one mesh, else per-lane atomics.
    const uint32_t m0 = (uint32_t)__shfl((int)mesh, 0);
    const bool uniform = __ballot(mesh == m0) == __ballot(true);
    if (uniform) {
        for (int a = 0; a < 3; ++a) {
            k[a] = wave_min_u32(k[a]);
            k[3 + a] = wave_max_u32(k[3 + a]);
        }
        if (__lane_id() == 0 && m0 != 0xffffffffu) {
            for (int a = 0; a < 3; ++a) {
                atomicMin(&bounds[m0 * 6 + a], k[a]);
                atomicMax(&bounds[m0 * 6 + 3 + a], k[3 + a]);
            }
        }
    } else if (valid) {
        for (int a = 0; a < 3; ++a) {
            atomicMin(&bounds[mesh * 6 + a], k[a]);
            atomicMax(&bounds[mesh * 6 + 3 + a], k[3 + a]);
        }
    }
}

__global__ void __launch_bounds__(T256) k_b_morton(const Node32* __restrict__ leaves,
                                                   const uint32_t* __restrict__ leaf_mesh, uint32_t n_total,
                                                   const uint32_t* __restrict__ bounds, uint64_t* __restrict__ keys) {
    const uint32_t g = blockIdx.x * T256 + threadIdx.x;
    if (g >= n_total) return;
    const uint32_t mesh = leaf_mesh[g];
    const uint32_t* bk = bounds + mesh * 6;
    const float mnx = f32_from_order_key(bk[0]), mny = f32_from_order_key(bk[1]), mnz = f32_from_order_key(bk[2]);
    const float mxx = f32_from_order_key(bk[3]), mxy = f32_from_order_key(bk[4]), mxz = f32_from_order_key(bk[5]);
    const double sx = 1.0 / (double)(mxx - mnx), sy = 1.0 / (double)(mxy - mny), sz = 1.0 / (double)(mxz - mnz);
    const double ox = -(double)mnx * sx, oy = -(double)mny * sy, oz = -(double)mnz * sz;
    keys[g] = morton_of_box(node_box(load_node(&leaves[g])), sx, sy, sz, ox, oy, oz);
}

__global__ void __launch_bounds__(T256) k_b_mesh_keys(const uint32_t* __restrict__ vals,
                                                      const uint32_t* __restrict__ leaf_mesh, uint32_t n_total,
                                                      uint64_t* __restrict__ keys) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i >= n_total) return;
    keys[i] = leaf_mesh[vals[i]];
}

// After the mesh sort, keys[i] is the mesh of the leaf order[i].
__global__ void __launch_bounds__(T256) k_b_gather(const Node32* __restrict__ leaves,
                                                   const uint32_t* __restrict__ order,
                                                   const uint64_t* __restrict__ mesh_keys, uint32_t n_total,
                                                   Node32* __restrict__ out, uint32_t* __restrict__ out_mesh) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i >= n_total) return;
    const uint32_t g = order[i];
    store_node(&out[i], load_node(&leaves[g]));
    out_mesh[i] = (uint32_t)mesh_keys[i];
}

// Neighbour pass (ploc/mod.rs:329-419) with windows clipped to the cluster's mesh segment.
// Each mesh searches with its own distance (MeshDesc::radius, at most rmax, the width of the LDS window); the
// distance-1 rule applies to a mesh whose distance is 1 and to every mesh during the first passes (before_sdt).
__global__ void __launch_bounds__(T256) k_b_search(const Node32* __restrict__ cur,
                                                   const uint32_t* __restrict__ cmesh,
                                                   const SegState* __restrict__ seg, BatchMisc* misc,
                                                   const MeshDesc* __restrict__ meshes, uint32_t rmax,
                                                   uint32_t before_sdt, int8_t* __restrict__ merge, uint32_t q) {
    __shared__ float s[6][T256 + 2 * MAX_RADIUS];
    const uint32_t total = misc->total[q];
    const uint32_t base = blockIdx.x * T256;
    if (base >= total) return;
    const int wstart = (int)base - (int)rmax;
    const int wlen = T256 + 2 * (int)rmax;
    bool nan_here = false;
    for (int w = threadIdx.x; w < wlen; w += T256) {
        const int g = wstart + w;
        if (g >= 0 && (uint32_t)g < total) {
            const Node32 nd = load_node(&cur[g]);
            s[0][w] = nd.minx;
            s[1][w] = nd.miny;
            s[2][w] = nd.minz;
            s[3][w] = nd.maxx;
            s[4][w] = nd.maxy;
            s[5][w] = nd.maxz;
            nan_here |= box_has_nan_any(node_box(nd));
        }
    }
    const bool nonan = !__syncthreads_or(nan_here);  // NaN-free window: costs from union_area_nonan
    const uint32_t i = base + threadIdx.x;
    if (i >= total) return;
    const uint32_t mi = cmesh[i];
    const SegState sg = seg[mi];
    if (sg.count <= 1u) return;
    const uint32_t r = meshes[mi].radius;
    const bool dist1 = r == 1u || before_sdt != 0u;
    const uint32_t li = i - sg.start;
    const int lw = (int)threadIdx.x + (int)rmax;
    const Box me{s[0][lw], s[1][lw], s[2][lw], s[3][lw], s[4][lw], s[5][lw]};
    int8_t mv;
    if (dist1) {
        if (li == sg.count - 1) {
            mv = -1;
        } else {
            const Box nx{s[0][lw + 1], s[1][lw + 1], s[2][lw + 1], s[3][lw + 1], s[4][lw + 1], s[5][lw + 1]};
            const float cost = half_area(box_union(me, nx));
            float last = INFINITY;
            if (li > 0) {
                const Box pv{s[0][lw - 1], s[1][lw - 1], s[2][lw - 1], s[3][lw - 1], s[4][lw - 1], s[5][lw - 1]};
                last = half_area(box_union(pv, me));
            }
            mv = last < cost ? (int8_t)-1 : (int8_t)1;
        }
    } else {
        // offsets k in [-min(r, li), -1] then [1, min(r, count - 1 - li)] (j ascending, no test for j == li), the
        // NaN-free / NaN branch outside the loops
        const int klo = -(int)min(r, li), khi = (int)min(r, sg.count - 1u - li);
        int best = 0;
        float best_cost = INFINITY;
        auto scan = [&](auto cost_of) {
            for (int k = klo; k < 0; ++k) {
                const float c = cost_of(lw + k);
                if (c <= best_cost) {
                    best = k;
                    best_cost = c;
                }
            }
            for (int k = 1; k <= khi; ++k) {
                const float c = cost_of(lw + k);
                if (c <= best_cost) {
                    best = k;
                    best_cost = c;
                }
            }
        };
        if (nonan)
            scan([&](int w) { return union_area_nonan(me, Box{s[0][w], s[1][w], s[2][w], s[3][w], s[4][w], s[5][w]}); });
        else
            scan([&](int w) { return half_area(box_union(me, Box{s[0][w], s[1][w], s[2][w], s[3][w], s[4][w], s[5][w]})); });
        if (best == 0) atomicOr(&misc->err, DERR_SELF_MERGE);
        mv = (int8_t)best;
    }
    merge[i] = mv;
}

// ploc/mod.rs:423-439 within the cluster's segment (indices are global, offsets stay inside the segment).
__device__ __forceinline__ void b_classify(const int8_t* __restrict__ merge, const uint32_t* __restrict__ cmesh,
                                           const SegState* __restrict__ seg, uint32_t i, uint32_t total, bool& emit,
                                           bool& mrg, uint32_t& b) {
    emit = false;
    mrg = false;
    b = i;
    if (i >= total) return;
    const SegState sg = seg[cmesh[i]];
    if (sg.count <= 1u) {
        emit = true;
        return;
    }
    const int bi = (int)i + (int)merge[i];
    b = (uint32_t)bi;
    const bool mutual = bi + (int)merge[bi] == (int)i;
    if (!mutual || bi == (int)i) {
        emit = true;
    } else if (bi < (int)i) {
        emit = true;
        mrg = true;
    }
}

__global__ void __launch_bounds__(T256) k_b_count(const int8_t* __restrict__ merge, const uint32_t* __restrict__ cmesh,
                                                  const SegState* __restrict__ seg, const BatchMisc* misc,
                                                  uint2* __restrict__ blocksums, uint32_t q) {
    __shared__ uint32_t we[T256 / 64], wm[T256 / 64];
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    bool emit, mrg;
    uint32_t b;
    b_classify(merge, cmesh, seg, i, misc->total[q], emit, mrg, b);
    const uint32_t e = (uint32_t)__popcll(__ballot(emit)), m = (uint32_t)__popcll(__ballot(mrg));
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0) {
        we[wave] = e;
        wm[wave] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t E = 0, M = 0;
        for (int w = 0; w < T256 / 64; ++w) {
            E += we[w];
            M += wm[w];
        }
        blocksums[blockIdx.x] = make_uint2(E, M);
    }
}

// Also writes the next pass's cluster count (total[q ^ 1]) and clears the active-mesh flag k_b_segments sets.
__global__ void __launch_bounds__(1024) k_b_scan(uint2* __restrict__ blocksums, uint32_t nblocks, BatchMisc* misc,
                                                 uint32_t q) {
    __shared__ uint32_t se[16], sm[16];
    const uint32_t tid = threadIdx.x;
    const uint32_t chunk = (nblocks + 1023u) / 1024u;
    const uint32_t b0 = tid * chunk, b1 = min(nblocks, b0 + chunk);
    uint32_t E = 0, M = 0;
    // each thread's chunk is contiguous (the prefix order); unrolled so that several loads are in flight
#pragma unroll 8
    for (uint32_t b = b0; b < b1; ++b) {
        const uint2 v = blocksums[b];
        E += v.x;
        M += v.y;
    }
    const uint32_t ie = wave_inclusive_scan(E), im = wave_inclusive_scan(M);
    const uint32_t wave = tid >> 6, lane = tid & 63u;
    if (lane == 63) {
        se[wave] = ie;
        sm[wave] = im;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t re = 0, rm = 0;
        for (int w = 0; w < 16; ++w) {
            const uint32_t te = se[w], tm = sm[w];
            se[w] = re;
            sm[w] = rm;
            re += te;
            rm += tm;
        }
        misc->totals[0] = re;
        misc->totals[1] = rm;
        misc->total[q ^ 1u] = re;
        misc->active = 0u;
    }
    __syncthreads();
    uint32_t ee = se[wave] + ie - E, em = sm[wave] + im - M;
#pragma unroll 8
    for (uint32_t b = b0; b < b1; ++b) {
        const uint2 v = blocksums[b];
        blocksums[b] = make_uint2(ee, em);
        ee += v.x;
        em += v.y;
    }
}

// Two-level form of k_b_scan for passes of more than B_SCAN_ONE_BLOCK tiles (the first passes hold ~78K): chunk
// totals by many workgroups, then each chunk offset and scanned by its own; workgroup 0 writes what k_b_scan writes.
constexpr uint32_t B_SCAN_ONE_BLOCK = 16384;

__global__ void __launch_bounds__(1024) k_b_scan_chunks(const uint2* __restrict__ blocksums, uint32_t nblocks,
                                                        uint2* __restrict__ chunk_tot) {
    scan_chunk_total(blocksums, nblocks, chunk_tot);
}

__global__ void __launch_bounds__(1024) k_b_scan_apply(uint2* __restrict__ blocksums, uint32_t nblocks,
                                                       const uint2* __restrict__ chunk_tot, uint32_t nchunks,
                                                       BatchMisc* misc, uint32_t q) {
    const uint2 all = scan_chunk_apply(blocksums, nblocks, chunk_tot, nchunks);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        misc->totals[0] = all.x;
        misc->totals[1] = all.y;
        misc->total[q ^ 1u] = all.x;
        misc->active = 0u;
    }
}

// Exclusive (emit, merge) prefixes at the segment's first cluster i0 and one past its last, i1: the scanned offset of
// the index's block plus the block's clusters before it, counted by the whole wave 64 clusters per step (one ballot
// each), as b_classify decides.  The up to 4 steps of both prefixes are independent, so all their loads go out in
// two rounds (cluster mesh and merge offset; segment count and partner's offset) before the first ballot.
__device__ __forceinline__ void b_prefix_pair(const int8_t* __restrict__ merge, const uint32_t* __restrict__ cmesh,
                                              const SegState* __restrict__ seg, const uint2* __restrict__ offsets,
                                              uint32_t i0, uint32_t i1, uint32_t total, const BatchMisc* misc,
                                              uint32_t lane, uint2& p0, uint2& p1) {
    uint32_t idx[8], cm[8];
    int mv[8];
    bool valid[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const uint32_t i = s < 4 ? i0 : i1;
        const uint32_t j = (i / T256) * T256 + (uint32_t)(s & 3) * 64u + lane;
        idx[s] = j;
        valid[s] = i < total && j < i;
        cm[s] = valid[s] ? cmesh[j] : 0u;
        mv[s] = valid[s] ? (int)merge[j] : 0;
    }
    uint32_t cnt[8];
    int mb[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        cnt[s] = valid[s] ? seg[cm[s]].count : 0u;
        // The partner's offset, read before the count says whether it matters: clamped into the array (a mesh
        // with one cluster left keeps a stale offset from an earlier pass).
        const int bi = min(max((int)idx[s] + mv[s], 0), (int)total - 1);
        mb[s] = valid[s] ? (int)merge[bi] : 0;
    }
    p0 = i0 >= total ? make_uint2(misc->totals[0], misc->totals[1]) : offsets[i0 / T256];
    p1 = i1 >= total ? make_uint2(misc->totals[0], misc->totals[1]) : offsets[i1 / T256];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        bool emit = false, mrg = false;
        if (valid[s]) {
            const int i = (int)idx[s], bi = i + mv[s];
            if (cnt[s] <= 1u) {
                emit = true;
            } else if (bi + mb[s] != i || bi == i) {  // not mutual, or a self merge (b_classify)
                emit = true;
            } else if (bi < i) {
                emit = true;
                mrg = true;
            }
        }
        const uint32_t ce = (uint32_t)__popcll(__ballot(emit)), cg = (uint32_t)__popcll(__ballot(mrg));
        if (s < 4) {
            p0.x += ce;
            p0.y += cg;
        } else {
            p1.x += ce;
            p1.y += cg;
        }
    }
}

// Per mesh (one wave each): the segment of the next pass, its insert index and this pass's merge base.
__global__ void __launch_bounds__(T256) k_b_segments(const int8_t* __restrict__ merge,
                                                     const uint32_t* __restrict__ cmesh, SegState* __restrict__ seg,
                                                     SegState* __restrict__ seg_next, const uint2* __restrict__ offsets,
                                                     uint32_t m, BatchMisc* misc, uint32_t q) {
    const uint32_t k = (blockIdx.x * T256 + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    bool active = false;
    if (k < m) {  // wave-uniform
        const SegState sg = seg[k];
        if (sg.count == 0u) {
            if (lane == 0u) seg_next[k] = sg;
        } else {
            const uint32_t total = misc->total[q];
            uint2 head, end;
            b_prefix_pair(merge, cmesh, seg, offsets, sg.start, sg.start + sg.count, total, misc, lane, head, end);
            if (lane == 0u) {
                const uint32_t merges = end.y - head.y;
                seg[k].mbase = head.y;
                const SegState nx{head.x, end.x - head.x, sg.insert - 2u * merges, 0u};
                seg_next[k] = nx;
                active = nx.count > 1u;
            }
        }
    }
    // One flag write per block: the host only asks whether any mesh is still active, and 4,096 waves adding to
    // one counter serialise on it.
    if (__syncthreads_or(active) && threadIdx.x == 0) misc->active = 1u;
}

__global__ void __launch_bounds__(T256) k_b_compact(const Node32* __restrict__ cur, Node32* __restrict__ next,
                                                    const uint32_t* __restrict__ cmesh, uint32_t* __restrict__ nmesh,
                                                    const int8_t* __restrict__ merge,
                                                    const SegState* __restrict__ seg, BatchMisc* misc,
                                                    const uint2* __restrict__ offsets,
                                                    const MeshDesc* __restrict__ meshes, Node32* __restrict__ nodes,
                                                    DpView dp, uint32_t q) {
    __shared__ uint32_t we[T256 / 64], wm[T256 / 64];
    // The tile's merges, their cost-table entries evaluated by consecutive threads afterwards (a merge is counted at
    // its higher cluster; a partner lies within its mesh's distance, at most MAX_RADIUS: at most 128 + r / 2).
    constexpr uint32_t MAX_TILE_MERGES = (T256 + MAX_RADIUS) / 2;
    __shared__ Node32 s_left[MAX_TILE_MERGES], s_right[MAX_TILE_MERGES];
    __shared__ uint32_t s_first[MAX_TILE_MERGES], s_kbase[MAX_TILE_MERGES];
    const uint32_t total = misc->total[q];
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    Node32 a{};
    if (i < total) a = load_node(&cur[i]);  // before the classification's dependent loads and the barrier
    bool emit, mrg;
    uint32_t b;
    b_classify(merge, cmesh, seg, i, total, emit, mrg, b);
    const uint64_t be = __ballot(emit), bm = __ballot(mrg);
    const uint64_t lt = lanemask_lt();
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0) {
        we[wave] = (uint32_t)__popcll(be);
        wm[wave] = (uint32_t)__popcll(bm);
    }
    __syncthreads();
    if (emit) {
        uint32_t pe = offsets[blockIdx.x].x, pm = 0u;
        for (uint32_t w = 0; w < wave; ++w) {
            pe += we[w];
            pm += wm[w];
        }
        const uint32_t pos = pe + (uint32_t)__popcll(be & lt);
        const uint32_t mesh = cmesh[i];
        nmesh[pos] = mesh;
        if (mrg) {
            const SegState sg = seg[mesh];
            const MeshDesc md = meshes[mesh];
            const uint32_t tq = pm + (uint32_t)__popcll(bm & lt);  // the merge's rank in this tile
            const uint32_t s = offsets[blockIdx.x].y + tq - sg.mbase;
            const Node32 c = load_node(&cur[b]);
            const uint32_t first = sg.insert - 2u * (s + 1u);
            Node32* mn = nodes + md.node_base;
            store_node(&mn[first], a);
            store_node(&mn[first + 1], c);
            store_node(&next[pos], make_node(box_union(node_box(a), node_box(c)), 0u, first));
            if (tq < MAX_TILE_MERGES) {
                s_left[tq] = a;
                s_right[tq] = c;
                s_first[tq] = first;
                s_kbase[tq] = md.prim_base;
            } else {
                atomicOr(&misc->err, DERR_CAPACITY);
            }
        } else {
            store_node(&next[pos], a);
        }
    }
    __syncthreads();
    const uint32_t nm = min(wm[0] + wm[1] + wm[2] + wm[3], MAX_TILE_MERGES);
    if (threadIdx.x < nm) {
        const Node32 l = s_left[threadIdx.x], r = s_right[threadIdx.x];
        DpView dpm = dp;
        dpm.kbase = s_kbase[threadIdx.x];
        const uint32_t e = dp_evaluate(dpm, box_union(node_box(l), node_box(r)), s_first[threadIdx.x], l, r);
        if (e) atomicOr(&misc->err, e);
    }
}

// Roots (ploc/mod.rs:499): the mesh's last cluster goes to its node 0.
__global__ void __launch_bounds__(T256) k_b_roots(const Node32* __restrict__ cur, const SegState* __restrict__ seg,
                                                  const MeshDesc* __restrict__ meshes, uint32_t m,
                                                  Node32* __restrict__ nodes, uint32_t* __restrict__ prims,
                                                  BatchMisc* misc) {
    const uint32_t k = blockIdx.x * T256 + threadIdx.x;
    if (k >= m) return;
    const SegState sg = seg[k];
    const MeshDesc md = meshes[k];
    if (md.prim_count == 0u) return;
    if (sg.count != 1u || sg.insert != 1u) {
        atomicOr(&misc->err, DERR_SELF_MERGE);  // ploc/mod.rs:496
        return;
    }
    store_node(&nodes[md.node_base], load_node(&cur[sg.start]));
}

// Root box of every mesh's BVH2 (the BLAS bounds a TLAS is built over), as Aabb records.
__global__ void __launch_bounds__(T256) k_b_root_boxes(const Node32* __restrict__ nodes,
                                                       const MeshDesc* __restrict__ meshes, uint32_t m,
                                                       float4* __restrict__ out) {
    const uint32_t k = blockIdx.x * T256 + threadIdx.x;
    if (k >= m) return;
    const Node32 r = load_node(&nodes[meshes[k].node_base]);
    out[2 * (size_t)k] = make_float4(r.minx, r.miny, r.minz, 0.0f);
    out[2 * (size_t)k + 1] = make_float4(r.maxx, r.maxy, r.maxz, 0.0f);
}

}  // namespace

#define OBVHS_TRY(expr, where)                                      \
    do {                                                            \
        hipError_t _e = (expr);                                     \
        if (_e != hipSuccess) return set_hip_error(ctx, _e, where); \
    } while (0)

static obvhs_status batch_build(obvhs_context* ctx, const obvhs_triangle* d_tris, const uint32_t* h_offsets,
                                uint32_t m, uint32_t r, uint32_t sdt, uint32_t max_prims, obvhs_batch* B) {
    hipStream_t st = ctx->stream;
    const uint32_t n = h_offsets[m];
    B->n_total = n;
    B->meshes.assign(m, MeshDesc{});
    uint32_t rmax = 1u;  // the widest search window of any mesh
    for (uint32_t k = 0; k < m; ++k) {
        if (h_offsets[k + 1] < h_offsets[k]) return OBVHS_ERR_INVALID_ARG;
        MeshDesc& md = B->meshes[k];
        md.prim_base = h_offsets[k];
        md.node_base = 2u * h_offsets[k];
        md.prim_count = h_offsets[k + 1] - h_offsets[k];
        md.radius = r != 0u ? r : radius_for(ctx, md.prim_count);
        rmax = std::max(rmax, md.radius);
    }
    OBVHS_TRY(B->tri_offsets.ensure((size_t)(m + 1) * 4), "alloc offsets");
    OBVHS_TRY(B->mesh_table.ensure((size_t)m * sizeof(MeshDesc)), "alloc mesh table");
    OBVHS_TRY(hipMemcpyAsync(B->tri_offsets.ptr, h_offsets, (size_t)(m + 1) * 4, hipMemcpyHostToDevice, st),
              "upload offsets");
    OBVHS_TRY(hipMemcpyAsync(B->mesh_table.ptr, B->meshes.data(), (size_t)m * sizeof(MeshDesc),
                             hipMemcpyHostToDevice, st),
              "upload mesh table");
    OBVHS_TRY(hipStreamSynchronize(st), "upload sync");  // the host vectors may change below
    if (n == 0) {
        B->cw_total = 0;
        for (auto& md : B->meshes) md.cw_count = 0;
        return OBVHS_OK;
    }
    BuildScratch& S = ctx->scratch;
    const uint32_t nblk = (n + T256 - 1) / T256;
    OBVHS_TRY(S.leaves.ensure((size_t)n * sizeof(Node32)), "alloc leaves");
    OBVHS_TRY(S.clusters_a.ensure((size_t)n * sizeof(Node32)), "alloc clusters");
    OBVHS_TRY(S.clusters_b.ensure((size_t)n * sizeof(Node32)), "alloc clusters");
    OBVHS_TRY(S.keys.ensure((size_t)n * 8), "alloc keys");
    OBVHS_TRY(S.vals.ensure((size_t)n * 8), "alloc vals + cluster meshes");
    const size_t cm_off = ((size_t)n + 64 + 255) & ~size_t(255);
    OBVHS_TRY(S.merge.ensure(cm_off + (size_t)n * 4), "alloc merge + cluster meshes");
    // per-tile counts, then the two-level scan's chunk totals
    OBVHS_TRY(S.blocksums.ensure((size_t)(nblk + 16 + nblk / SCAN_CHUNK + 1) * sizeof(uint2) + 64),
              "alloc blocksums");
    OBVHS_TRY(S.sort_scratch.ensure(radix_sort_scratch_bytes(n)), "alloc sort scratch");
    const size_t seg_bytes = (size_t)m * sizeof(SegState);
    OBVHS_TRY(S.misc.ensure(2 * seg_bytes + (size_t)m * 24 + sizeof(BatchMisc) + 1024), "alloc batch misc");
    OBVHS_TRY(B->bvh2_nodes.ensure((size_t)2 * n * sizeof(Node32)), "alloc bvh2 nodes");
    OBVHS_TRY(B->bvh2_prims.ensure((size_t)n * 4), "alloc bvh2 prims");
    const size_t k = n;
    OBVHS_TRY(B->dp.ensure(dp_table_bytes(k)), "alloc cost table");
    char* mb = S.misc.as<char>();
    SegState* seg[2] = {reinterpret_cast<SegState*>(mb), reinterpret_cast<SegState*>(mb + seg_bytes)};
    uint32_t* bounds = reinterpret_cast<uint32_t*>(mb + 2 * seg_bytes);
    BatchMisc* misc = reinterpret_cast<BatchMisc*>(mb + 2 * seg_bytes + (size_t)m * 24);
    const uint32_t* offsets = B->tri_offsets.as<uint32_t>();
    MeshDesc* d_meshes = B->mesh_table.as<MeshDesc>();
    Node32* leaves = S.leaves.as<Node32>();
    uint64_t* keys = S.keys.as<uint64_t>();
    uint32_t* vals = S.vals.as<uint32_t>();
    uint32_t* cmesh[2] = {S.vals.as<uint32_t>() + n, reinterpret_cast<uint32_t*>(S.merge.as<char>() + cm_off)};
    int8_t* merge = S.merge.as<int8_t>();
    uint2* blocksums = S.blocksums.as<uint2>();
    Node32* nodes = B->bvh2_nodes.as<Node32>();
    DpView dp = dp_view_at(B->dp.as<char>(), k, max_prims);

    const uint32_t mblk = (m + T256 - 1) / T256;
    k_b_init<<<mblk, T256, 0, st>>>(bounds, offsets, m, seg[0], misc, n);
    // every leaf's mesh, in the second cluster-mesh buffer (free until the first pass)
    uint32_t* leaf_mesh = cmesh[1];
    k_b_leaf_init<<<(n + LEAF_T - 1) / LEAF_T, LEAF_T, 0, st>>>(reinterpret_cast<const float4*>(d_tris), offsets, m,
                                                                n, leaves, bounds, misc, leaf_mesh,
                                                                B->bvh2_prims.as<uint32_t>());
    k_b_morton<<<nblk, T256, 0, st>>>(leaves, leaf_mesh, n, bounds, keys);
    OBVHS_TRY(hipGetLastError(), "batch leaf launch");
    const int items = sort_items_for(ctx, n);
    OBVHS_TRY(radix_sort_pairs(keys, vals, true, n, 63, S.sort_scratch.ptr, st, items), "batch morton sort");
    k_b_mesh_keys<<<nblk, T256, 0, st>>>(vals, leaf_mesh, n, keys);
    int mbits = 1;
    while (mbits < 32 && (1ull << mbits) < m) mbits++;
    OBVHS_TRY(radix_sort_pairs(keys, vals, false, n, mbits, S.sort_scratch.ptr, st, items), "batch mesh sort");
    Node32* cur = S.clusters_a.as<Node32>();
    Node32* nxt = S.clusters_b.as<Node32>();
    k_b_gather<<<nblk, T256, 0, st>>>(leaves, vals, keys, n, cur, cmesh[0]);
    OBVHS_TRY(hipGetLastError(), "batch gather launch");

    uint32_t known = n;
    uint32_t pass = 0;
    int cm = 0;    // which cluster-mesh buffer is current
    int cs = 0;    // which segment table is current
    uint32_t* h = static_cast<uint32_t*>(ctx->pinned);
    for (;;) {
        const uint32_t blocks = (known + T256 - 1) / T256;
        for (int rep = 0; rep < PASSES_PER_SYNC; ++rep) {
            const uint32_t q = pass & 1u;
            k_b_search<<<blocks, T256, 0, st>>>(cur, cmesh[cm], seg[cs], misc, d_meshes, rmax,
                                                pass < sdt ? 1u : 0u, merge, q);
            k_b_count<<<blocks, T256, 0, st>>>(merge, cmesh[cm], seg[cs], misc, blocksums, q);
            if (blocks <= B_SCAN_ONE_BLOCK) {
                k_b_scan<<<1, 1024, 0, st>>>(blocksums, blocks, misc, q);
            } else {
                uint2* chunk_tot = blocksums + blocks + 16;
                const uint32_t nchunks = (blocks + SCAN_CHUNK - 1) / SCAN_CHUNK;
                k_b_scan_chunks<<<nchunks, 1024, 0, st>>>(blocksums, blocks, chunk_tot);
                k_b_scan_apply<<<nchunks, 1024, 0, st>>>(blocksums, blocks, chunk_tot, nchunks, misc, q);
            }
            k_b_segments<<<(m + 3u) / 4u, T256, 0, st>>>(merge, cmesh[cm], seg[cs], seg[cs ^ 1], blocksums, m,
                                                          misc, q);
            k_b_compact<<<blocks, T256, 0, st>>>(cur, nxt, cmesh[cm], cmesh[cm ^ 1], merge, seg[cs], misc, blocksums,
                                                 d_meshes, nodes, dp, q);
            std::swap(cur, nxt);
            cm ^= 1;
            cs ^= 1;
            pass++;
        }
        OBVHS_TRY(hipGetLastError(), "batch pass launch");
        OBVHS_TRY(hipMemcpyAsync(h, misc, 16, hipMemcpyDeviceToHost, st), "read batch state");
        OBVHS_TRY(hipStreamSynchronize(st), "batch pass sync");
        if (h[0]) return device_error_to_status(h[0]);
        if (h[3] == 0u) break;  // every mesh is down to its root
        // Every active mesh merges at least one pair per pass unless costs are NaN.
        const uint32_t now = h[1 + (pass & 1u)];
        if (now >= known) return OBVHS_ERR_PLOC_SELF_MERGE;
        known = now;
    }
    B->passes = pass;
    k_b_roots<<<mblk, T256, 0, st>>>(cur, seg[cs], d_meshes, m, nodes, B->bvh2_prims.as<uint32_t>(), misc);
    OBVHS_TRY(hipGetLastError(), "batch roots launch");
    OBVHS_TRY(hipMemcpyAsync(h, misc, 4, hipMemcpyDeviceToHost, st), "read batch err");
    OBVHS_TRY(hipStreamSynchronize(st), "batch roots sync");
    if (h[0]) return device_error_to_status(h[0]);
    return cwbvh_convert_batch(ctx, nodes, B->bvh2_prims.as<uint32_t>(), dp, d_meshes, B->meshes, n, true,
                               B->cw_nodes, B->cw_prims, B->cw_total);
}

}  // namespace obvhs

using namespace obvhs;

extern "C" {

obvhs_status obvhs_batch_build_cwbvh(obvhs_context* ctx, const obvhs_triangle* d_tris, const uint32_t* tri_offsets,
                                     uint32_t mesh_count, const obvhs_build_params* params, obvhs_batch* reuse,
                                     obvhs_batch** out) {
    if (!ctx || !params || !out || !tri_offsets) return OBVHS_ERR_INVALID_ARG;
    if (params->pre_split || params->reinsertion_batch_ratio > 0.0f || params->sort_precision != 64u)
        return OBVHS_ERR_INVALID_ARG;
    const uint32_t r = params->ploc_search_distance;  // 0: per mesh from the tuning table
    if (!(r == 0 || r == 1 || r == 2 || r == 6 || r == 14 || r == 24 || r == 32)) return OBVHS_ERR_INVALID_ARG;
    if (tri_offsets[0] != 0u || tri_offsets[mesh_count] >= (1u << 30)) return OBVHS_ERR_INVALID_ARG;
    if (tri_offsets[mesh_count] > 0 && !d_tris) return OBVHS_ERR_INVALID_ARG;
    uint32_t mp = params->max_prims_per_leaf;
    mp = mp < 1u ? 1u : (mp > 3u ? 3u : mp);  // cwbvh/builder.rs:74
    if (hipSetDevice(ctx->device) != hipSuccess) return OBVHS_ERR_DEVICE;
    obvhs_batch* B = reuse ? reuse : new (std::nothrow) obvhs_batch();
    if (!B) return OBVHS_ERR_OOM;
    const obvhs_status s =
        batch_build(ctx, d_tris, tri_offsets, mesh_count, r, params->search_depth_threshold, mp, B);
    if (s != OBVHS_OK) {
        if (!reuse) obvhs_batch_destroy(B);
        *out = nullptr;
        return s;
    }
    *out = B;
    return OBVHS_OK;
}

obvhs_status obvhs_batch_info(const obvhs_batch* b, uint32_t* mesh_count, uint32_t* total_nodes,
                              uint32_t* total_prims, uint32_t* ploc_passes) {
    if (!b) return OBVHS_ERR_INVALID_ARG;
    if (mesh_count) *mesh_count = (uint32_t)b->meshes.size();
    if (total_nodes) *total_nodes = b->cw_total;
    if (total_prims) *total_prims = b->n_total;
    if (ploc_passes) *ploc_passes = b->passes;
    return OBVHS_OK;
}

obvhs_status obvhs_batch_mesh_info(const obvhs_batch* b, uint32_t mesh, uint32_t* node_offset, uint32_t* node_count,
                                   uint32_t* prim_offset, uint32_t* prim_count) {
    if (!b || mesh >= b->meshes.size()) return OBVHS_ERR_INVALID_ARG;
    const MeshDesc& md = b->meshes[mesh];
    if (node_offset) *node_offset = md.cw_base;
    if (node_count) *node_count = md.cw_count;
    if (prim_offset) *prim_offset = md.prim_base;
    if (prim_count) *prim_count = md.prim_count;
    return OBVHS_OK;
}

obvhs_status obvhs_batch_copy_nodes(obvhs_context* ctx, const obvhs_batch* b, obvhs_cwbvh_node* host_out) {
    if (!ctx || !b) return OBVHS_ERR_INVALID_ARG;
    return obvhs_memcpy_dtoh(ctx, host_out, b->cw_nodes.ptr, (size_t)b->cw_total * sizeof(CwNode));
}

obvhs_status obvhs_batch_copy_primitive_indices(obvhs_context* ctx, const obvhs_batch* b, uint32_t* host_out) {
    if (!ctx || !b) return OBVHS_ERR_INVALID_ARG;
    return obvhs_memcpy_dtoh(ctx, host_out, b->cw_prims.ptr, (size_t)b->n_total * 4);
}

obvhs_status obvhs_batch_reorder_triangles(obvhs_context* ctx, const obvhs_batch* b, const obvhs_triangle* d_tris,
                                           obvhs_triangle* d_bvh_tris) {
    if (!ctx || !b) return OBVHS_ERR_INVALID_ARG;
    if (b->n_total == 0) return OBVHS_OK;
    hipError_t e = launch_reorder_triangles_meshes(d_tris, b->cw_prims.as<uint32_t>(), b->tri_offsets.as<uint32_t>(),
                                                   (uint32_t)b->meshes.size(), b->n_total, d_bvh_tris, ctx->stream);
    if (e != hipSuccess) return set_hip_error(ctx, e, "batch reorder");
    return OBVHS_OK;
}

obvhs_status obvhs_batch_ray_traverse(obvhs_context* ctx, const obvhs_batch* b, const obvhs_triangle* d_bvh_tris,
                                      const obvhs_ray* d_rays, const uint32_t* d_ray_mesh, uint32_t n,
                                      obvhs_ray_hit* d_hits) {
    if (!ctx || !b || (n && (!d_rays || !d_ray_mesh || !d_hits))) return OBVHS_ERR_INVALID_ARG;
    if (b->n_total && !d_bvh_tris) return OBVHS_ERR_INVALID_ARG;
    TravScratch ts;
    if (n) {
        const obvhs_status ss = traverse_scratch(ctx, ts, n);
        if (ss != OBVHS_OK) return ss;
    }
    ts.lds_depth = lds_depth_for(ctx, b->meshes.empty() ? 0u : b->n_total / b->meshes.size());  // mean mesh
    ts.chunked = chunk_for(ctx, b->meshes.empty() ? 0u : b->n_total / b->meshes.size());
    hipError_t e = launch_traverse_meshes(b->cw_nodes.as<CwNode>(), b->mesh_table.as<MeshDesc>(),
                                          (uint32_t)b->meshes.size(), d_ray_mesh,
                                          d_bvh_tris, d_rays, n, 0, d_hits, nullptr, ts, ctx->stream);
    if (e != hipSuccess) return set_hip_error(ctx, e, "batch traverse");
    return OBVHS_OK;
}

obvhs_status obvhs_batch_ray_traverse_miss(obvhs_context* ctx, const obvhs_batch* b,
                                           const obvhs_triangle* d_bvh_tris, const obvhs_ray* d_rays,
                                           const uint32_t* d_ray_mesh, uint32_t n, uint8_t* d_miss) {
    if (!ctx || !b || (n && (!d_rays || !d_ray_mesh || !d_miss))) return OBVHS_ERR_INVALID_ARG;
    if (b->n_total && !d_bvh_tris) return OBVHS_ERR_INVALID_ARG;
    TravScratch ts;
    if (n) {
        const obvhs_status ss = traverse_scratch(ctx, ts, n);
        if (ss != OBVHS_OK) return ss;
    }
    ts.lds_depth = lds_depth_for(ctx, b->meshes.empty() ? 0u : b->n_total / b->meshes.size());  // mean mesh
    ts.chunked = chunk_for(ctx, b->meshes.empty() ? 0u : b->n_total / b->meshes.size());
    hipError_t e = launch_traverse_meshes(b->cw_nodes.as<CwNode>(), b->mesh_table.as<MeshDesc>(),
                                          (uint32_t)b->meshes.size(), d_ray_mesh,
                                          d_bvh_tris, d_rays, n, 1, nullptr, d_miss, ts, ctx->stream);
    if (e != hipSuccess) return set_hip_error(ctx, e, "batch traverse miss");
    return OBVHS_OK;
}

obvhs_status obvhs_batch_build_tlas(obvhs_context* ctx, const obvhs_batch* b, const obvhs_build_params* params,
                                    obvhs_cwbvh** out) {
    if (!ctx || !b || !params || !out) return OBVHS_ERR_INVALID_ARG;
    *out = nullptr;
    const uint32_t m = (uint32_t)b->meshes.size();
    for (const MeshDesc& md : b->meshes)
        if (md.prim_count == 0u) return OBVHS_ERR_INVALID_ARG;  // an empty BLAS has no bounds
    if (hipSetDevice(ctx->device) != hipSuccess) return OBVHS_ERR_DEVICE;
    DevBuf boxes;
    hipError_t e = boxes.ensure((size_t)std::max(m, 1u) * 32);
    if (e != hipSuccess) return set_hip_error(ctx, e, "alloc tlas boxes");
    if (m) {
        k_b_root_boxes<<<(m + T256 - 1) / T256, T256, 0, ctx->stream>>>(b->bvh2_nodes.as<Node32>(),
                                                                       b->mesh_table.as<MeshDesc>(), m,
                                                                       boxes.as<float4>());
        e = hipGetLastError();
        if (e != hipSuccess) {
            boxes.release();
            return set_hip_error(ctx, e, "tlas boxes");
        }
    }
    obvhs_cwbvh* cw = new (std::nothrow) obvhs_cwbvh();
    if (!cw) {
        boxes.release();
        return OBVHS_ERR_OOM;
    }
    uint32_t mp = params->max_prims_per_leaf;
    mp = mp < 1u ? 1u : (mp > 3u ? 3u : mp);
    obvhs_bvh2* b2 = nullptr;
    obvhs_status s = obvhs_ploc_build_aabbs(ctx, boxes.as<obvhs_aabb>(), nullptr, m, params->ploc_search_distance,
                                            params->sort_precision, params->search_depth_threshold, nullptr, &b2);
    if (s == OBVHS_OK) {
        cw->owned_bvh2 = b2;
        s = cwbvh_convert(ctx, b2, mp, true, cw);
    }
    (void)hipStreamSynchronize(ctx->stream);
    boxes.release();
    if (s != OBVHS_OK) {
        obvhs_cwbvh_destroy(cw);
        return s;
    }
    *out = cw;
    return OBVHS_OK;
}

obvhs_status obvhs_batch_tlas_ray_traverse(obvhs_context* ctx, const obvhs_cwbvh* tlas, const obvhs_batch* b,
                                           const obvhs_triangle* d_bvh_tris, const obvhs_ray* d_rays, uint32_t n,
                                           obvhs_ray_hit* d_hits) {
    if (!ctx || !tlas || !b || (n && (!d_rays || !d_hits))) return OBVHS_ERR_INVALID_ARG;
    if (b->n_total && !d_bvh_tris) return OBVHS_ERR_INVALID_ARG;
    TravScratch ts;
    if (n) {
        const obvhs_status ss = traverse_scratch(ctx, ts, n);
        if (ss != OBVHS_OK) return ss;
    }
    hipError_t e = launch_traverse_tlas(tlas->nodes.as<CwNode>(), tlas->node_count, tlas->prim_indices.as<uint32_t>(),
                                        tlas->prim_count, b->cw_nodes.as<CwNode>(), b->mesh_table.as<MeshDesc>(),
                                        (uint32_t)b->meshes.size(), d_bvh_tris, d_rays, n, d_hits, ts, ctx->stream);
    if (e != hipSuccess) return set_hip_error(ctx, e, "tlas traverse");
    return OBVHS_OK;
}

obvhs_status obvhs_batch_destroy(obvhs_batch* b) {
    if (!b) return OBVHS_OK;
    b->tri_offsets.release();
    b->mesh_table.release();
    b->bvh2_nodes.release();
    b->bvh2_prims.release();
    b->dp.release();
    b->cw_nodes.release();
    b->cw_prims.release();
    delete b;
    return OBVHS_OK;
}

}  // extern "C"
