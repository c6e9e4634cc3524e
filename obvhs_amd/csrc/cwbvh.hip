// BVH2 -> CWBVH (8-wide, 80-byte compressed nodes) conversion on gfx950.
//
// Restates Bvh2Converter (cwbvh/bvh2_to_cwbvh.rs:15-510) with outputs byte-identical to the reference's two
// recursive passes:
//   k_cw_parents : parent links of the BVH2 (bvh2/mod.rs:586-619) and arrival counters.
//   k_cw_dp      : bottom-up cost table (calculate_cost_impl, bvh2_to_cwbvh.rs:220-344).  One thread per leaf
//                  climbs; the second thread to arrive at a node (acq_rel atomic) evaluates it.  The same thread
//                  also evaluates get_children(node, 0) to know the size of the CWBVH subtree that the node would
//                  root, which lets the top-down pass number nodes exactly like the reference's pre-order DFS.
//   k_cw_emit    : one launch per CWBVH level.  Each node computes its children (get_children, order_children),
//                  its exponents and quantised child boxes, writes the 80-byte node and its leaf primitive
//                  indices, and hands its internal children their DFS-derived node index, child base and
//                  primitive base (convert_to_cwbvh_impl, bvh2_to_cwbvh.rs:75-193; count_primitives 197-211).
// The exponent exp2(ceil(log2(x))) (bvh2_to_cwbvh.rs:85-88) goes through the host libm's log2f in the reference;
// a 256-entry table of per-exponent mantissa thresholds, generated from that same log2f when the context is
// created, reproduces it with integer compares.
#include <math.h>
#include <string.h>

#include "common.hpp"
#include "internal.hpp"

namespace obvhs {

namespace {

constexpr int T256 = 256;
constexpr uint32_t KIND_LEAF = 0, KIND_INTERNAL = 1, KIND_DISTRIBUTE = 2;  // bvh2_to_cwbvh.rs:470-476
constexpr uint32_t DINVALID = 0xffu;
constexpr float PRIM_COST = 0.3f;  // bvh2_to_cwbvh.rs:30
constexpr uint32_t MAX_CW_LEVELS = 64;

__constant__ uint32_t c_log2_thresh[256];

struct CwItem {
    uint32_t cw_index;
    uint32_t bvh2;
    uint32_t a;  // CW nodes allocated before this node's child block, minus the root
    uint32_t b;  // primitives appended before this node
};

struct CwMisc {
    uint32_t err;
    uint32_t level_count[MAX_CW_LEVELS + 1];
};

// Decision packing: kind | distribute_left << 8 | distribute_right << 16 (bvh2_to_cwbvh.rs:478-484).
__device__ __forceinline__ uint32_t dkind(uint32_t m) { return m & 0xffu; }
__device__ __forceinline__ uint32_t dleft(uint32_t m) { return (m >> 8) & 0xffu; }
__device__ __forceinline__ uint32_t dright(uint32_t m) { return (m >> 16) & 0xffu; }
__device__ __forceinline__ uint32_t dpack(uint32_t k, uint32_t l, uint32_t r) { return k | (l << 8) | (r << 16); }

__global__ void __launch_bounds__(T256) k_cw_parents(const Node32* __restrict__ nodes, uint32_t N,
                                                     uint32_t* __restrict__ parent, uint32_t* __restrict__ arrive,
                                                     CwMisc* misc) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i == 0) {
        misc->err = 0u;
        for (uint32_t l = 0; l <= MAX_CW_LEVELS; ++l) misc->level_count[l] = 0u;
        parent[0] = 0u;
    }
    if (i >= N) return;
    arrive[i] = 0u;
    const Node32 nd = load_node(&nodes[i]);
    if (nd.prim_count == 0u) {
        parent[nd.first_index] = i;
        parent[nd.first_index + 1] = i;
    }
}

// get_children(node, 0) (bvh2_to_cwbvh.rs:346-397) with an explicit stack; left before right.
// Returns the number of children or -1 on a malformed decision.
__device__ __forceinline__ int get_children(const Node32* __restrict__ nodes, const uint32_t* __restrict__ dmeta,
                                            uint32_t node, uint32_t* children) {
    uint32_t stk_node[8];
    uint32_t stk_i[8];
    int sp = 0;
    int cc = 0;
    const Node32 root = load_node(&nodes[node]);
    if (root.prim_count != 0u) {
        children[cc++] = node;
        return cc;
    }
    stk_node[sp] = node;
    stk_i[sp] = 0u;
    sp++;
    while (sp > 0) {
        sp--;
        const uint32_t x = stk_node[sp], i = stk_i[sp];
        if (i == 7u) {  // emit as a child
            if (cc >= 8) return -1;
            children[cc++] = x;
            continue;
        }
        const uint32_t first = nodes[x].first_index;
        const uint32_t m = dmeta[(size_t)x * 7 + i];
        const uint32_t dl = dleft(m), dr = dright(m);
        if (dl >= 7u || dr >= 7u) return -2;
        if (sp + 2 > 8) return -1;
        // right pushed first so that the left subtree is produced first
        const uint32_t kr = dkind(dmeta[(size_t)(first + 1) * 7 + dr]);
        stk_node[sp] = first + 1;
        stk_i[sp] = kr == KIND_DISTRIBUTE ? dr : 7u;
        sp++;
        const uint32_t kl = dkind(dmeta[(size_t)first * 7 + dl]);
        stk_node[sp] = first;
        stk_i[sp] = kl == KIND_DISTRIBUTE ? dl : 7u;
        sp++;
    }
    return cc;
}

// Evaluates calculate_cost_impl for an internal node whose children are finished (bvh2_to_cwbvh.rs:268-340).
__device__ __forceinline__ void dp_internal(uint32_t node, float ha, uint32_t first, uint32_t num_primitives,
                                            uint32_t max_prims, float* __restrict__ dcost,
                                            uint32_t* __restrict__ dmeta) {
    float lc[7], rc[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        lc[k] = dcost[(size_t)first * 7 + k];
        rc[k] = dcost[(size_t)(first + 1) * 7 + k];
    }
    float cost[7];
    uint32_t meta[7];
    {
        const float cost_leaf = num_primitives <= max_prims ? ((float)num_primitives) * ha * PRIM_COST : INFINITY;
        float cost_distribute = INFINITY;
        uint32_t dl = DINVALID, dr = DINVALID;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const float c = lc[k] + rc[6 - k];
            if (c < cost_distribute) {
                cost_distribute = c;
                dl = (uint32_t)k;
                dr = (uint32_t)(6 - k);
            }
        }
        const float cost_internal = cost_distribute + ha;
        if (cost_leaf < cost_internal) {
            cost[0] = cost_leaf;
            meta[0] = dpack(KIND_LEAF, dl, dr);
        } else {
            cost[0] = cost_internal;
            meta[0] = dpack(KIND_INTERNAL, dl, dr);
        }
    }
#pragma unroll
    for (int i = 1; i < 7; ++i) {
        float cost_distribute = cost[i - 1];
        uint32_t dl = DINVALID, dr = DINVALID;
#pragma unroll
        for (int k = 0; k < i; ++k) {
            const float c = lc[k] + rc[i - k - 1];
            if (c < cost_distribute) {
                cost_distribute = c;
                dl = (uint32_t)k;
                dr = (uint32_t)(i - k - 1);
            }
        }
        if (dl != DINVALID) {
            cost[i] = cost_distribute;
            meta[i] = dpack(KIND_DISTRIBUTE, dl, dr);
        } else {
            cost[i] = cost[i - 1];
            meta[i] = meta[i - 1];
        }
    }
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        dcost[(size_t)node * 7 + k] = cost[k];
        dmeta[(size_t)node * 7 + k] = meta[k];
    }
}

__global__ void __launch_bounds__(T256) k_cw_dp(const Node32* __restrict__ nodes, uint32_t N,
                                                const uint32_t* __restrict__ parent, uint32_t* arrive,
                                                float* dcost, uint32_t* dmeta, uint32_t* nprims, uint32_t* cwsize,
                                                uint32_t max_prims, CwMisc* misc) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i >= N) return;
    const Node32 leaf = load_node(&nodes[i]);
    if (leaf.prim_count == 0u) return;
    if (leaf.prim_count != 1u) {  // bvh2_to_cwbvh.rs:241-247
        atomicOr(&misc->err, DERR_CW_NONUNIT_LEAF);
        return;
    }
    const float cost_leaf = half_area(node_box(leaf)) * 1.0f * PRIM_COST;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        dcost[(size_t)i * 7 + k] = cost_leaf;
        dmeta[(size_t)i * 7 + k] = dpack(KIND_LEAF, 0u, 0u);
    }
    nprims[i] = 1u;
    cwsize[i] = 1u;
    uint32_t x = i;
    while (x != 0u) {
        const uint32_t p = parent[x];
        const uint32_t old = __hip_atomic_fetch_add(&arrive[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0u) return;  // the sibling's thread finishes this node
        const Node32 pn = load_node(&nodes[p]);
        const uint32_t np = nprims[pn.first_index] + nprims[pn.first_index + 1];
        dp_internal(p, half_area(node_box(pn)), pn.first_index, np, max_prims, dcost, dmeta);
        nprims[p] = np;
        uint32_t ch[8];
        const int cc = get_children(nodes, dmeta, p, ch);
        if (cc < 0) {
            atomicOr(&misc->err, cc == -2 ? DERR_CW_DISTRIBUTE : DERR_CAPACITY);
            cwsize[p] = 1u;
        } else {
            uint32_t sz = 1u;
            for (int c = 0; c < cc; ++c)
                if (dkind(dmeta[(size_t)ch[c] * 7]) == KIND_INTERNAL) sz += cwsize[ch[c]];
            cwsize[p] = sz;
        }
        x = p;
    }
}

__global__ void k_cw_root(CwItem* items, CwMisc* misc) {
    if (threadIdx.x == 0) {
        items[0] = CwItem{0u, 0u, 0u, 0u};
        misc->level_count[0] = 1u;
    }
}

// exp2(ceil(log2(x))) for x = max(extent, 1e-20) * (1/255) as the reference computes it with libm; returns the
// biased exponent byte (bvh2_to_cwbvh.rs:85-98).
__device__ __forceinline__ uint32_t cw_exponent(float ext) {
    const float x = smax(ext, 1e-20f) * (1.0f / 255.0f);
    const uint32_t bits = __float_as_uint(x);
    const uint32_t E = bits >> 23;  // x > 0
    if (E >= 255u) return 255u;     // +inf stays inf
    const uint32_t M = bits & 0x7fffffu;
    uint32_t k = (M <= c_log2_thresh[E]) ? E : E + 1u;  // ceil(log2 x) + 127
    return k >= 255u ? 255u : k;
}

__global__ void __launch_bounds__(T256) k_cw_emit(const Node32* __restrict__ nodes,
                                                  const uint32_t* __restrict__ bvh2_prims,
                                                  const uint32_t* __restrict__ dmeta,
                                                  const uint32_t* __restrict__ nprims,
                                                  const uint32_t* __restrict__ cwsize,
                                                  const CwItem* __restrict__ in_items, CwItem* __restrict__ out_items,
                                                  uint32_t level, uint32_t order_children, CwNode* __restrict__ out,
                                                  uint32_t* __restrict__ out_prims, uint32_t cw_cap, uint32_t prim_cap,
                                                  CwMisc* misc) {
    const uint32_t t = blockIdx.x * T256 + threadIdx.x;
    const uint32_t count = misc->level_count[level];
    if (t >= count) return;
    const CwItem it = in_items[t];
    const Node32 nd = load_node(&nodes[it.bvh2]);
    const Box nb = node_box(nd);

    uint32_t ch[8];
    int cc = get_children(nodes, dmeta, it.bvh2, ch);
    if (cc < 0) {
        atomicOr(&misc->err, cc == -2 ? DERR_CW_DISTRIBUTE : DERR_CAPACITY);
        return;
    }
    Box cb[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        if (c < cc) cb[c] = node_box(load_node(&nodes[ch[c]]));
        else cb[c] = Box{0, 0, 0, 0, 0, 0};
    }
    // order_children (bvh2_to_cwbvh.rs:402-467): greedy octant assignment, child-major / slot-minor, strict <.
    int slot_of[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) slot_of[c] = c;
    if (order_children) {
        const float px = (nb.hx + nb.lx) * 0.5f, py = (nb.hy + nb.ly) * 0.5f, pz = (nb.hz + nb.lz) * 0.5f;
        const float FMAX = 3.40282347e+38f;
        float cost[8][8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const float vx = (cb[c].hx + cb[c].lx) * 0.5f - px;
            const float vy = (cb[c].hy + cb[c].ly) * 0.5f - py;
            const float vz = (cb[c].hz + cb[c].lz) * 0.5f - pz;
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const float dx = (s & 4) ? -1.0f : 1.0f, dy = (s & 2) ? -1.0f : 1.0f, dz = (s & 1) ? -1.0f : 1.0f;
                cost[c][s] = c < cc ? (dx * vx + dy * vy) + dz * vz : FMAX;
            }
        }
        uint32_t assigned = 0u, filled = 0u;
#pragma unroll
        for (int c = 0; c < 8; ++c) slot_of[c] = -1;
        for (int round = 0; round < 8; ++round) {
            float min_cost = FMAX;
            int min_slot = -1, min_index = -1;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                if (c < cc && !((assigned >> c) & 1u)) {
#pragma unroll
                    for (int s = 0; s < 8; ++s) {
                        const float v = cost[c][s];
                        if (!((filled >> s) & 1u) && v < min_cost) {
                            min_cost = v;
                            min_slot = s;
                            min_index = c;
                        }
                    }
                }
            }
            if (min_slot < 0) break;
            filled |= 1u << min_slot;
            assigned |= 1u << min_index;
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (c == min_index) slot_of[c] = min_slot;
        }
        for (int c = 0; c < cc; ++c) {
            if (slot_of[c] < 0) {
                atomicOr(&misc->err, DERR_CW_ORDER);
                return;
            }
        }
    }
    // slots[s] = child index placed in slot s (or -1)
    int slot_child[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        slot_child[s] = -1;
#pragma unroll
        for (int c = 0; c < 8; ++c)
            if (c < cc && slot_of[c] == s) slot_child[s] = c;
    }

    CwNode o;
    o.p[0] = nb.lx;
    o.p[1] = nb.ly;
    o.p[2] = nb.lz;
    const uint32_t ex = cw_exponent(nb.hx - nb.lx), ey = cw_exponent(nb.hy - nb.ly), ez = cw_exponent(nb.hz - nb.lz);
    o.e[0] = (uint8_t)ex;
    o.e[1] = (uint8_t)ey;
    o.e[2] = (uint8_t)ez;
    const float rx = 1.0f / __uint_as_float(ex << 23), ry = 1.0f / __uint_as_float(ey << 23),
                rz = 1.0f / __uint_as_float(ez << 23);
    uint32_t imask = 0u;
    uint32_t num_internal = 0u, num_primitives = 0u;
    uint32_t internal_slots[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        o.child_meta[s] = 0;
        o.child_min_x[s] = o.child_max_x[s] = 0;
        o.child_min_y[s] = o.child_max_y[s] = 0;
        o.child_min_z[s] = o.child_max_z[s] = 0;
    }
    const uint32_t child_base = 1u + it.a;
    // First sweep: leaves append primitives in slot order before any recursion (bvh2_to_cwbvh.rs:117-170).
    for (int s = 0; s < 8; ++s) {
        const int c = slot_child[s];
        if (c < 0) continue;
        Box b = cb[0];
#pragma unroll
        for (int q = 1; q < 8; ++q)
            if (q == c) b = cb[q];
        const uint32_t ci = ch[c];
        float lo, hi;
        lo = floorf(((b.lx - o.p[0]) - 0.0f) * rx);
        hi = ceilf(((b.hx - o.p[0]) + 0.0f) * rx);
        o.child_min_x[s] = (uint8_t)smin(smax(lo, 0.0f), 255.0f);
        o.child_max_x[s] = (uint8_t)smin(smax(hi, 0.0f), 255.0f);
        lo = floorf(((b.ly - o.p[1]) - 0.0f) * ry);
        hi = ceilf(((b.hy - o.p[1]) + 0.0f) * ry);
        o.child_min_y[s] = (uint8_t)smin(smax(lo, 0.0f), 255.0f);
        o.child_max_y[s] = (uint8_t)smin(smax(hi, 0.0f), 255.0f);
        lo = floorf(((b.lz - o.p[2]) - 0.0f) * rz);
        hi = ceilf(((b.hz - o.p[2]) + 0.0f) * rz);
        o.child_min_z[s] = (uint8_t)smin(smax(lo, 0.0f), 255.0f);
        o.child_max_z[s] = (uint8_t)smin(smax(hi, 0.0f), 255.0f);
        const uint32_t kind = dkind(dmeta[(size_t)ci * 7]);
        if (kind == KIND_LEAF) {
            // count_primitives: left-first DFS of the (at most 3-leaf) subtree.
            uint32_t stk[8];
            int sp = 0;
            stk[sp++] = ci;
            uint32_t pc = 0;
            while (sp > 0) {
                const uint32_t x = stk[--sp];
                const Node32 xn = load_node(&nodes[x]);
                if (xn.prim_count != 0u) {
                    const uint32_t dst = it.b + num_primitives + pc;
                    if (dst < prim_cap) out_prims[dst] = bvh2_prims[xn.first_index];
                    pc += xn.prim_count;
                } else {
                    if (sp + 2 > 8) {
                        atomicOr(&misc->err, DERR_CW_PRIMCOUNT);
                        return;
                    }
                    stk[sp++] = xn.first_index + 1;
                    stk[sp++] = xn.first_index;
                }
            }
            uint32_t unary;
            if (pc == 1u) unary = 0x20u;
            else if (pc == 2u) unary = 0x60u;
            else if (pc == 3u) unary = 0xe0u;
            else {
                atomicOr(&misc->err, DERR_CW_PRIMCOUNT);
                return;
            }
            o.child_meta[s] = (uint8_t)(num_primitives | unary);
            num_primitives += pc;
        } else if (kind == KIND_INTERNAL) {
            imask |= 1u << s;
            o.child_meta[s] = (uint8_t)((24u + (uint32_t)s) | 0x20u);
            internal_slots[num_internal++] = (uint32_t)s;
        } else {
            atomicOr(&misc->err, DERR_CW_DISTRIBUTE);
            return;
        }
    }
    o.imask = (uint8_t)imask;
    o.child_base_idx = child_base;
    o.primitive_base_idx = it.b;
    if (it.cw_index < cw_cap) {
        const uint4* src = reinterpret_cast<const uint4*>(&o);
        uint4* dst = reinterpret_cast<uint4*>(&out[it.cw_index]);
#pragma unroll
        for (int q = 0; q < 5; ++q) dst[q] = src[q];
    } else {
        atomicOr(&misc->err, DERR_CAPACITY);
    }
    if (num_internal == 0u) return;
    if (level + 1 >= MAX_CW_LEVELS) {
        atomicOr(&misc->err, DERR_CW_DEPTH);
        return;
    }
    // Children's DFS numbering (pre-order, child blocks appended at visit time, bvh2_to_cwbvh.rs:172-191).
    const uint32_t base = atomicAdd(&misc->level_count[level + 1], num_internal);
    uint32_t a_run = it.a + num_internal;
    uint32_t b_run = it.b + num_primitives;
    for (uint32_t j = 0; j < num_internal; ++j) {
        const uint32_t s = internal_slots[j];
        const int c = slot_child[s];
        uint32_t ci = ch[0];
#pragma unroll
        for (int q = 1; q < 8; ++q)
            if (q == c) ci = ch[q];
        out_items[base + j] = CwItem{child_base + j, ci, a_run, b_run};
        a_run += cwsize[ci] - 1u;
        b_run += nprims[ci];
    }
}

__global__ void __launch_bounds__(T256) k_reorder_tris(const float4* __restrict__ tris,
                                                       const uint32_t* __restrict__ prims, uint32_t n,
                                                       float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i >= n) return;
    const size_t s = (size_t)prims[i] * 3, d = (size_t)i * 3;
    out[d] = tris[s];
    out[d + 1] = tris[s + 1];
    out[d + 2] = tris[s + 2];
}

}  // namespace

// Builds the exponent table from the host libm's log2f (the function Rust's f32::log2 calls on linux-gnu):
// thresh[E] = largest 23-bit mantissa M such that log2f(2^(E-127) * (1 + M/2^23)) == E - 127 exactly.
const uint32_t* host_log2_table() {
    static uint32_t table[256];
    static bool built = false;
    if (!built) {
        for (int E = 0; E < 256; ++E) {
            table[E] = 0u;
            if (E == 0 || E == 255) continue;
            const float k = (float)(E - 127);
            uint32_t best = 0u;
            int misses = 0;
            for (uint32_t M = 0; M < (1u << 23) && misses < 256; ++M) {
                const uint32_t bits = ((uint32_t)E << 23) | M;
                float x;
                memcpy(&x, &bits, 4);
                if (log2f(x) == k) {
                    best = M;
                    misses = 0;
                } else {
                    misses++;
                }
            }
            table[E] = best;
        }
        built = true;
    }
    return table;
}

hipError_t upload_log2_table() {
    return hipMemcpyToSymbol(HIP_SYMBOL(c_log2_thresh), host_log2_table(), 256 * sizeof(uint32_t));
}

#define OBVHS_TRY(expr, where)                                      \
    do {                                                            \
        hipError_t _e = (expr);                                     \
        if (_e != hipSuccess) return set_hip_error(ctx, _e, where); \
    } while (0)

obvhs_status cwbvh_convert(obvhs_context* ctx, const obvhs_bvh2* bvh2, uint32_t max_prims_per_leaf,
                           bool order_children, obvhs_cwbvh* out) {
    hipStream_t st = ctx->stream;
    const uint32_t N = bvh2->node_count;
    out->node_count = 0;
    out->prim_count = 0;
    memset(&out->total_aabb, 0, sizeof(out->total_aabb));
    if (N == 0) return OBVHS_OK;  // bvh2_to_cwbvh.rs:496-498
    const uint32_t n = bvh2->prim_count;
    // Scratch layout.
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~size_t(255);
        return o;
    };
    const size_t o_parent = carve((size_t)N * 4), o_arrive = carve((size_t)N * 4), o_cost = carve((size_t)N * 28),
                 o_meta = carve((size_t)N * 28), o_np = carve((size_t)N * 4), o_sz = carve((size_t)N * 4),
                 o_items0 = carve((size_t)N * sizeof(CwItem)), o_items1 = carve((size_t)N * sizeof(CwItem)),
                 o_misc = carve(sizeof(CwMisc));
    OBVHS_TRY(ctx->scratch.cw_scratch.ensure(off), "alloc cwbvh scratch");
    char* base = ctx->scratch.cw_scratch.as<char>();
    uint32_t* parent = reinterpret_cast<uint32_t*>(base + o_parent);
    uint32_t* arrive = reinterpret_cast<uint32_t*>(base + o_arrive);
    float* dcost = reinterpret_cast<float*>(base + o_cost);
    uint32_t* dmeta = reinterpret_cast<uint32_t*>(base + o_meta);
    uint32_t* nprims = reinterpret_cast<uint32_t*>(base + o_np);
    uint32_t* cwsize = reinterpret_cast<uint32_t*>(base + o_sz);
    CwItem* items[2] = {reinterpret_cast<CwItem*>(base + o_items0), reinterpret_cast<CwItem*>(base + o_items1)};
    CwMisc* misc = reinterpret_cast<CwMisc*>(base + o_misc);
    const Node32* nodes = bvh2->nodes.as<Node32>();
    const uint32_t blocks = (N + T256 - 1) / T256;
    k_cw_parents<<<blocks, T256, 0, st>>>(nodes, N, parent, arrive, misc);
    k_cw_dp<<<blocks, T256, 0, st>>>(nodes, N, parent, arrive, dcost, dmeta, nprims, cwsize, max_prims_per_leaf,
                                     misc);
    k_cw_root<<<1, 64, 0, st>>>(items[0], misc);
    OBVHS_TRY(hipGetLastError(), "cwbvh dp launch");
    uint32_t* h = static_cast<uint32_t*>(ctx->pinned);
    OBVHS_TRY(hipMemcpyAsync(h, cwsize, 4, hipMemcpyDeviceToHost, st), "read cw size");
    OBVHS_TRY(hipMemcpyAsync(h + 1, &misc->err, 4, hipMemcpyDeviceToHost, st), "read cw err");
    OBVHS_TRY(hipStreamSynchronize(st), "cwbvh dp sync");
    if (h[1]) return device_error_to_status(h[1]);
    const uint32_t ncw = h[0];
    OBVHS_TRY(out->nodes.ensure((size_t)ncw * sizeof(CwNode)), "alloc cwbvh nodes");
    OBVHS_TRY(out->prim_indices.ensure((size_t)n * 4 + 4), "alloc cwbvh primitive indices");
    uint32_t level_count = 1;
    uint32_t level = 0;
    while (level_count > 0) {
        if (level >= MAX_CW_LEVELS) return OBVHS_ERR_DEPTH;
        const uint32_t g = (level_count + T256 - 1) / T256;
        k_cw_emit<<<g, T256, 0, st>>>(nodes, bvh2->prim_indices.as<uint32_t>(), dmeta, nprims, cwsize,
                                      items[level & 1], items[(level + 1) & 1], level, order_children ? 1u : 0u,
                                      out->nodes.as<CwNode>(), out->prim_indices.as<uint32_t>(), ncw, n, misc);
        OBVHS_TRY(hipGetLastError(), "cwbvh emit launch");
        OBVHS_TRY(hipMemcpyAsync(h, &misc->level_count[level + 1], 4, hipMemcpyDeviceToHost, st), "read level");
        OBVHS_TRY(hipMemcpyAsync(h + 1, &misc->err, 4, hipMemcpyDeviceToHost, st), "read cw err");
        OBVHS_TRY(hipStreamSynchronize(st), "cwbvh level sync");
        if (h[1]) return device_error_to_status(h[1]);
        level_count = h[0];
        level++;
    }
    // total_aabb = bvh2.nodes[0].aabb (bvh2_to_cwbvh.rs:506)
    Node32 root;
    OBVHS_TRY(hipMemcpy(&root, nodes, sizeof(Node32), hipMemcpyDeviceToHost), "read root");
    out->total_aabb.min = obvhs_vec3a{root.minx, root.miny, root.minz, 0.0f};
    out->total_aabb.max = obvhs_vec3a{root.maxx, root.maxy, root.maxz, 0.0f};
    out->node_count = ncw;
    out->prim_count = n;
    // The CWBVH is a tree of depth `level`; traversal's 32-entry stack bounds it (cwbvh/mod.rs:791).
    return OBVHS_OK;
}

hipError_t launch_reorder_triangles(const obvhs_triangle* tris, const uint32_t* prim_indices, uint32_t n,
                                    obvhs_triangle* out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    k_reorder_tris<<<(n + T256 - 1) / T256, T256, 0, stream>>>(reinterpret_cast<const float4*>(tris), prim_indices,
                                                                n, reinterpret_cast<float4*>(out));
    return hipGetLastError();
}

}  // namespace obvhs
