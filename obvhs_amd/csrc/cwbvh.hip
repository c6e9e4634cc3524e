// BVH2 -> CWBVH (8-wide, 80-byte compressed nodes) conversion on gfx950.
//
// Restates Bvh2Converter (cwbvh/bvh2_to_cwbvh.rs:15-510) with outputs byte-identical to the reference's two
// recursive passes:
//   cost table : calculate_cost_impl (bvh2_to_cwbvh.rs:220-344), evaluated by the PLOC merge that creates each node
//                (cwdp.hpp) or, for a BVH2 built without it, by k_dp_range over the PLOC pass ranges: pass p emitted
//                the child slots [hist[p+1], hist[p]) and every internal node's children were emitted by an
//                earlier pass, so ranges processed in pass order see finished children (no atomics, no fences).
//                Alongside the seven decisions each node stores the size of the CWBVH subtree it roots, which
//                lets the top-down pass number nodes exactly like the reference's pre-order DFS.
//   k_cw_emit  : one launch per CWBVH level, 8 lanes per node.  The node's own box, first_index and table words
//                come with its work item from its parent; the group walks get_children breadth first (lane j holds
//                the j-th entry of the frontier and loads its node and table entry, one dependent load per
//                expansion level), the octant assignment (order_children) runs as a per-round 8-lane argmin, then lane s
//                quantises the child in slot s; prefix sums over the slots give the leaf primitive offsets and the
//                internal children's DFS-derived node index, child base and primitive base, and the 80-byte node
//                is assembled with lane shuffles (convert_to_cwbvh_impl, bvh2_to_cwbvh.rs:75-193;
//                count_primitives 197-211).
// The exponent exp2(ceil(log2(x))) (bvh2_to_cwbvh.rs:85-88) goes through the host libm's log2f in the reference;
// a 256-entry table of per-exponent mantissa thresholds, generated from that same log2f when the context is
// created, reproduces it with integer compares.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "cwdp.hpp"
#include "internal.hpp"

namespace obvhs {

namespace {

constexpr int T256 = 256;
constexpr uint32_t MAX_CW_LEVELS = 64;
constexpr uint32_t DP_SMALL_RANGE = 4096;  // ranges at most this large are finished by one workgroup

__constant__ uint32_t c_log2_thresh[256];

struct CwItem {
    uint32_t cw_index;
    uint32_t bvh2;
    uint32_t a;  // CW nodes allocated before this node's child block, minus the root
    uint32_t b;  // primitives appended before this node
    uint32_t mesh;
};

// What the 8-lane emit of a node needs of the node itself, handed down by its parent (whose walk loaded it) so
// that the node's emit starts with its own cost-table walk: the box, first_index and prim_count, and its entry's
// decision-0 byte and children words.  An array of these runs beside each item list of that form.
struct CwPre {
    float lx, ly, lz, hx, hy, hz;
    uint32_t first, pc;
    uint32_t meta0;   // decision byte 0 of the node's entry (internal nodes)
    uint32_t kx, ky;  // the entry's children words (first_index | KID_LEAF)
    uint32_t pad;
};

struct CwMisc {
    uint32_t err;
    uint32_t level_count[MAX_CW_LEVELS + 1];
    uint32_t info[10];  // k_cw_info: the BVH2 root (8 words), the CWBVH node count and depth
    CwItem root_item;   // level 0 of a single conversion
    CwPre root_pre;
};

__global__ void k_cw_init(CwMisc* misc) {
    if (threadIdx.x == 0) {
        misc->err = 0u;
        for (uint32_t l = 0; l <= MAX_CW_LEVELS; ++l) misc->level_count[l] = 0u;
        misc->level_count[0] = 1u;
    }
}

// The CwPre of BVH2 node x (slot x of nds; its entry keyed through d).
__device__ __forceinline__ CwPre pre_of(const Node32* __restrict__ nds, const DpView& d, uint32_t x) {
    const Node32 n = load_node(&nds[x]);
    CwPre p{n.minx, n.miny, n.minz, n.maxx, n.maxy, n.maxz, n.first_index, n.prim_count, 0u, 0u, 0u, 0u};
    if (n.prim_count == 0u) {
        const DpRec& e = d.rec[dp_key(d, n.first_index)];
        p.meta0 = dbyte(e.meta, 0u);
        p.kx = e.kids.x;
        p.ky = e.kids.y;
    }
    return p;
}

// What the host needs before the emit, gathered so that one copy brings it with the error word: the BVH2 root, and
// the size and depth of the CWBVH subtree it roots (1 and 1 for a single-leaf BVH2).
__global__ void k_cw_info(const Node32* __restrict__ nodes, DpView dp, CwMisc* misc) {
    if (threadIdx.x != 0) return;
    const Node32 root = load_node(&nodes[0]);
    const uint4* w = reinterpret_cast<const uint4*>(&root);
    const uint4 a = w[0], b = w[1];
    misc->info[0] = a.x;
    misc->info[1] = a.y;
    misc->info[2] = a.z;
    misc->info[3] = a.w;
    misc->info[4] = b.x;
    misc->info[5] = b.y;
    misc->info[6] = b.z;
    misc->info[7] = b.w;
    uint32_t size = 1u, depth = 1u;
    if (root.prim_count == 0u) {
        const DpRec& e = dp.rec[dkey(root.first_index)];
        size = e.cwsize;
        depth = e.cwdepth;
    }
    misc->info[8] = size;
    misc->info[9] = depth;
    misc->root_item = CwItem{0u, 0u, 0u, 0u, 0u};
    misc->root_pre = pre_of(nodes, dp, 0u);
}

__device__ __forceinline__ void dp_node(const Node32* __restrict__ nodes, const DpView& dp, uint32_t x,
                                        uint32_t* err) {
    const Node32 nx = load_node(&nodes[x]);
    if (nx.prim_count != 0u) return;
    const Node32 l = load_node(&nodes[nx.first_index]);
    const Node32 r = load_node(&nodes[nx.first_index + 1]);
    const uint32_t e = dp_evaluate(dp, node_box(nx), nx.first_index, l, r);
    if (e) atomicOr(err, e);
}

// Cost table of the nodes a large PLOC pass emitted (slots [lo, hi)).
__global__ void __launch_bounds__(T256) k_dp_range(const Node32* __restrict__ nodes, DpView dp, uint32_t lo,
                                                   uint32_t hi, CwMisc* misc) {
    const uint32_t x = lo + blockIdx.x * T256 + threadIdx.x;
    if (x >= hi) return;
    dp_node(nodes, dp, x, &misc->err);
}

// Cost table of every node in slots [1, hi) -- the small late passes, processed pass by pass inside one
// workgroup -- and of the root.  hist[p] for p in [p0, passes] bounds the ranges.
__global__ void __launch_bounds__(1024) k_dp_small(const Node32* __restrict__ nodes, DpView dp,
                                                   const uint32_t* __restrict__ hist, uint32_t p0,
                                                   uint32_t passes, CwMisc* misc) {
    for (uint32_t p = p0; p < passes; ++p) {
        const uint32_t lo = hist[p + 1], hi = hist[p];
        for (uint32_t x = lo + threadIdx.x; x < hi; x += 1024) dp_node(nodes, dp, x, &misc->err);
        __syncthreads();
    }
    if (threadIdx.x == 0) dp_node(nodes, dp, 0u, &misc->err);
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

// Group helpers: the 8 lanes r = 0..7 of an aligned lane group work on one CWBVH node (lane r first owns the
// node's child r, then slot r).  The fixed lane patterns inside a group (xor 1, 2, 4 and shifts up by 1, 2, 4) are
// DPP moves, which stay in the VALU instead of a round trip through the LDS crossbar: quad_perm for xor 1 and 2,
// and within a row of 16 lanes, row_shl:4 (lane i reads lane i + 4) and row_shr:d (lane i reads lane i - d), used
// only where the source lies in the same group.
constexpr int DPP_XOR1 = 0xB1;      // quad_perm [1, 0, 3, 2]
constexpr int DPP_XOR2 = 0x4E;      // quad_perm [2, 3, 0, 1]
constexpr int DPP_ROW_SHL4 = 0x104;
constexpr int DPP_ROW_SHR1 = 0x111;
constexpr int DPP_ROW_SHR2 = 0x112;
constexpr int DPP_ROW_SHR4 = 0x114;
template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t g8_xor1(uint32_t v) { return dpp<DPP_XOR1>(v); }
__device__ __forceinline__ uint32_t g8_xor2(uint32_t v) { return dpp<DPP_XOR2>(v); }
__device__ __forceinline__ uint32_t g8_xor4(uint32_t v, uint32_t r) {
    const uint32_t up = dpp<DPP_ROW_SHL4>(v), dn = dpp<DPP_ROW_SHR4>(v);
    return (r & 4u) ? dn : up;
}
__device__ __forceinline__ float g8_xor1f(float v) { return __uint_as_float(g8_xor1(__float_as_uint(v))); }
__device__ __forceinline__ float g8_xor2f(float v) { return __uint_as_float(g8_xor2(__float_as_uint(v))); }
__device__ __forceinline__ float g8_xor4f(float v, uint32_t r) { return __uint_as_float(g8_xor4(__float_as_uint(v), r)); }
__device__ __forceinline__ uint32_t g8_or(uint32_t v) {
    const uint32_t r = threadIdx.x & 7u;
    v |= g8_xor1(v);
    v |= g8_xor2(v);
    v |= g8_xor4(v, r);
    return v;
}
__device__ __forceinline__ uint32_t g8_inclusive_sum(uint32_t v, uint32_t r) {
    uint32_t t = dpp<DPP_ROW_SHR1>(v);
    if (r >= 1u) v += t;
    t = dpp<DPP_ROW_SHR2>(v);
    if (r >= 2u) v += t;
    t = dpp<DPP_ROW_SHR4>(v);
    if (r >= 4u) v += t;
    return v;
}
// One 8-byte child array of the node from one byte per slot lane: (bytes of slots 0-3, bytes of slots 4-7).
__device__ __forceinline__ uint2 g8_bytes(uint32_t byte, uint32_t r) {
    uint32_t w = byte << (8u * (r & 3u));
    w |= g8_xor1(w);
    w |= g8_xor2(w);
    const uint32_t x = g8_xor4(w, r);
    return r < 4u ? make_uint2(w, x) : make_uint2(x, w);
}
__device__ __forceinline__ uint32_t quantise_lo(float b, float p, float rcp) {
    return (uint32_t)(uint8_t)smin(smax(floorf(((b - p) - 0.0f) * rcp), 0.0f), 255.0f);
}
__device__ __forceinline__ uint32_t quantise_hi(float b, float p, float rcp) {
    return (uint32_t)(uint8_t)smin(smax(ceilf(((b - p) + 0.0f) * rcp), 0.0f), 255.0f);
}

constexpr uint32_t EMIT_MAX_BLOCKS = 2048;   // 8 waves per SIMD on 256 CUs; larger levels loop

// What one node hands to the next level: its internal child count (0 on error), and per slot lane whether the
// slot holds an internal child, its rank among them and its work item.
struct EmitOut {
    uint32_t num_internal;
    uint32_t isint;
    uint32_t rank;
    CwItem item;
    CwPre pre;
};

// One CWBVH node by the 8-lane group of lane r (every lane of the group calls with the same item).
__device__ __forceinline__ void emit_node(const Node32* __restrict__ nodes, const uint32_t* __restrict__ bvh2_prims,
                                          const DpView& dp, const CwItem it, const CwPre pre, uint32_t r,
                                          uint32_t level, uint32_t order_children, CwNode* __restrict__ out,
                                          uint32_t* __restrict__ out_prims, uint32_t cw_cap, uint32_t prim_cap,
                                          CwMisc* misc, const MeshDesc* __restrict__ meshes, EmitOut& eo) {
    const Node32* nds = nodes;
    const uint32_t* bprims = bvh2_prims;
    CwNode* onodes = out;
    uint32_t* oprims = out_prims;
    uint32_t ccap = cw_cap, pcap = prim_cap;
    DpView d = dp;
    if (meshes) {  // batched conversions: every array is the item's mesh's slice
        const MeshDesc md = meshes[it.mesh];
        nds += md.node_base;
        bprims += md.prim_base;
        onodes += md.cw_base;
        oprims += md.prim_base;
        ccap = md.cw_count;
        pcap = md.prim_count;
        d.kbase = md.prim_base;
    }
    const Box nb{pre.lx, pre.ly, pre.lz, pre.hx, pre.hy, pre.hz};

    // get_children (bvh2_to_cwbvh.rs:346-397) breadth first over the group: lane j holds entry j, left to right, of
    // the expansion frontier (slot, the parent's children word for it, its distribute index).  Each round every
    // lane loads its entry's node and cost-table entry at once (one dependent load per round, not per visited
    // entry), DISTRIBUTE entries split in place into their two children, and the final frontier read left to right
    // is the reference's left-first recursion order.  The loads of the last round serve the rest of the emit.
    uint32_t fx = 0u, finfo = KID_LEAF, fdi = 0u;
    int F;
    uint32_t werr = 0u;
    if (pre.pc != 0u) {  // a single-leaf BVH2: the root is its own only child
        F = 1;
        fx = it.bvh2;
    } else {
        const uint32_t dl = dleft(pre.meta0), dr = dright(pre.meta0);
        if (dl == DIDX_INVALID || dr == DIDX_INVALID) werr = DERR_CW_DISTRIBUTE;
        F = 2;
        fx = pre.first + (r == 1u ? 1u : 0u);
        finfo = r == 1u ? pre.ky : pre.kx;
        fdi = r == 1u ? dr : dl;
    }
    Node32 cn{0.0f, 0.0f, 0.0f, 0u, 0.0f, 0.0f, 0.0f, 0u};
    uint2 rmeta = make_uint2(0u, 0u), rkids = make_uint2(0u, 0u);
    uint32_t rcw = 0u, rnp = 0u;
    uint4 rpl = make_uint4(0u, 0u, 0u, 0u);
    while (werr == 0u) {  // at most 8 rounds: every round but the last grows the frontier, which stops at 8
        const bool mine = (int)r < F;
        const bool internal = mine && !(finfo & KID_LEAF);
        if (mine) cn = load_node(&nds[fx]);
        if (internal) {
            const DpRec& e = d.rec[dp_key(d, finfo & ~KID_LEAF)];
            rmeta = e.meta;
            rkids = e.kids;
            rcw = e.cwsize;
            rnp = e.nprims;
            rpl = e.plist;
        }
        const uint32_t b = internal ? dbyte(rmeta, fdi) : 0u;
        const bool dist = internal && dkind(b) == KIND_DISTRIBUTE;
        if (dist && (dleft(b) == DIDX_INVALID || dright(b) == DIDX_INVALID)) werr = DERR_CW_DISTRIBUTE;
        werr = g8_or(werr);
        if (werr != 0u || g8_or(dist ? 1u : 0u) == 0u) break;
        const uint32_t w = mine ? (dist ? 2u : 1u) : 0u;
        // every lane's width (2 bits each) in one word; lane r of the new frontier takes entry part of the entry of
        // source lane q whose range [acc_q, acc_q + w_q) holds r (acc_q: the widths of the lanes before q)
        const uint32_t widths = g8_or(w << (2u * r));
        int src = 0;
        uint32_t part = 0u, acc = 0u;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t wq = (widths >> (2 * q)) & 3u;
            if (r >= acc && r < acc + wq) {
                src = q;
                part = r - acc;
            }
            acc += wq;
        }
        const int F2 = (int)acc;
        if (F2 > 8) {  // more than 8 children
            werr = DERR_CAPACITY;
            break;
        }
        const uint32_t sx = (uint32_t)__shfl((int)fx, src, 8), sinfo = (uint32_t)__shfl((int)finfo, src, 8);
        const uint32_t sdi = (uint32_t)__shfl((int)fdi, src, 8), sb = (uint32_t)__shfl((int)b, src, 8);
        const uint32_t skx = (uint32_t)__shfl((int)rkids.x, src, 8), sky = (uint32_t)__shfl((int)rkids.y, src, 8);
        const uint32_t sw = (widths >> (2 * src)) & 3u;
        if (sw == 2u) {  // a DISTRIBUTE entry's children: slots first, first + 1
            fx = (sinfo & ~KID_LEAF) + part;
            finfo = part != 0u ? sky : skx;
            fdi = part != 0u ? dright(sb) : dleft(sb);
        } else {
            fx = sx;
            finfo = sinfo;
            fdi = sdi;
        }
        F = F2;
    }
    if (werr != 0u) {
        if (r == 0u) atomicOr(&misc->err, werr);
        return;
    }
    const int cc = F;
    // lane r: child r (its node and, unless it is a BVH2 leaf, its table entry, from the walk's last round)
    const bool has = (int)r < cc;
    const uint32_t my_ch = fx;
    Box cb{0, 0, 0, 0, 0, 0};
    uint32_t cfirst = 0u, cpc = 0u;
    if (has) {
        cb = node_box(cn);
        cfirst = cn.first_index;
        cpc = cn.prim_count;
    }
    const uint32_t ckd = (!has || (finfo & KID_LEAF)) ? KIND_LEAF : dkind(rmeta.x & 0xffu);  // decision-0 kind

    // order_children (bvh2_to_cwbvh.rs:402-467): greedy octant assignment, the first strict minimum in
    // child-major / slot-minor order per round.  Lane r finds its own first minimum over the free slots; the
    // group keeps the lowest cost, ties to the lowest child.
    int my_slot = has ? (int)r : -1;
    if (order_children) {
        const float px = (nb.hx + nb.lx) * 0.5f, py = (nb.hy + nb.ly) * 0.5f, pz = (nb.hz + nb.lz) * 0.5f;
        const float vx = (cb.hx + cb.lx) * 0.5f - px, vy = (cb.hy + cb.ly) * 0.5f - py,
                    vz = (cb.hz + cb.lz) * 0.5f - pz;
        const float FMAX = 3.40282347e+38f;
        uint32_t filled = 0u;
        bool assigned = false;
        my_slot = -1;
        for (int round = 0; round < 8; ++round) {
            float bc = FMAX;
            int bs = -1;
            if (has && !assigned) {
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    const float dx = (s & 4) ? -1.0f : 1.0f, dy = (s & 2) ? -1.0f : 1.0f,
                                dz = (s & 1) ? -1.0f : 1.0f;
                    const float v = (dx * vx + dy * vy) + dz * vz;  // exact for d in {+-1}^3
                    if (!((filled >> s) & 1u) && v < bc) {
                        bc = v;
                        bs = s;
                    }
                }
            }
            int bl = (int)r;
#pragma unroll
            for (int m = 1; m < 8; m <<= 1) {
                const float oc = m == 1 ? g8_xor1f(bc) : (m == 2 ? g8_xor2f(bc) : g8_xor4f(bc, r));
                const int os = (int)(m == 1 ? g8_xor1((uint32_t)bs) : (m == 2 ? g8_xor2((uint32_t)bs) : g8_xor4((uint32_t)bs, r)));
                const int ol = (int)(m == 1 ? g8_xor1((uint32_t)bl) : (m == 2 ? g8_xor2((uint32_t)bl) : g8_xor4((uint32_t)bl, r)));
                if (os >= 0 && (bs < 0 || oc < bc || (oc == bc && ol < bl))) {
                    bc = oc;
                    bs = os;
                    bl = ol;
                }
            }
            if (bs < 0) break;  // the same (bs, bl) in all 8 lanes
            filled |= 1u << bs;
            if ((uint32_t)bl == r) {
                assigned = true;
                my_slot = bs;
            }
        }
        if (g8_or(has && my_slot < 0 ? 1u : 0u)) {
            if (r == 0u) atomicOr(&misc->err, DERR_CW_ORDER);
            return;
        }
    }

    // lane r: slot r (the child placed there, if any), from every lane's slot + 1 (4 bits each) in one word
    const uint32_t slots = g8_or((uint32_t)(my_slot + 1) << (4u * r));
    int src = -1;
#pragma unroll
    for (int q = 0; q < 8; ++q)
        if (((slots >> (4 * q)) & 0xfu) == r + 1u) src = q;
    const bool occ = src >= 0;
    const int sl = occ ? src : (int)r;
    const Box sb{__shfl(cb.lx, sl, 8), __shfl(cb.ly, sl, 8), __shfl(cb.lz, sl, 8),
                 __shfl(cb.hx, sl, 8), __shfl(cb.hy, sl, 8), __shfl(cb.hz, sl, 8)};
    const uint32_t s_first = (uint32_t)__shfl((int)cfirst, sl, 8), s_pc = (uint32_t)__shfl((int)cpc, sl, 8);
    const uint32_t s_kind = (uint32_t)__shfl((int)ckd, sl, 8), s_ch = (uint32_t)__shfl((int)my_ch, sl, 8);
    const uint4 s_pl = make_uint4((uint32_t)__shfl((int)rpl.x, sl, 8), (uint32_t)__shfl((int)rpl.y, sl, 8),
                                  (uint32_t)__shfl((int)rpl.z, sl, 8), (uint32_t)__shfl((int)rpl.w, sl, 8));

    // Leaves append primitives in slot order (bvh2_to_cwbvh.rs:117-170; count_primitives 197-211): a BVH2 leaf
    // its first primitive with prim_count, a subtree of at most 3 leaves its leaves left first, as recorded in
    // the cost table.
    uint32_t lpc = 0u, isint = 0u, bad = 0u;
    uint4 pl = make_uint4(0u, 0u, 0u, 0u);
    if (occ) {
        if (s_kind == KIND_LEAF) {
            pl = s_pc != 0u ? make_uint4(s_first, 0u, 0u, 1u) : s_pl;
            lpc = s_pc != 0u ? s_pc : pl.w;
            if (lpc < 1u || lpc > 3u) bad |= DERR_CW_PRIMCOUNT;
        } else if (s_kind == KIND_INTERNAL) {
            isint = 1u;
        } else {
            bad |= DERR_CW_DISTRIBUTE;
        }
    }
    const uint32_t packed = (bad ? 0u : lpc) | isint << 8;
    const uint32_t incl = g8_inclusive_sum(packed, r);
    const uint32_t tot = (uint32_t)__shfl((int)incl, 7, 8);
    const uint32_t prim_off = (incl - packed) & 0xffu, int_rank = (incl - packed) >> 8;
    const uint32_t num_primitives = tot & 0xffu, num_internal = tot >> 8;
    const uint32_t imask = g8_or(isint << r);
    if (it.cw_index >= ccap) bad |= DERR_CAPACITY;
    if (num_internal != 0u && level + 1 >= MAX_CW_LEVELS) bad |= DERR_CW_DEPTH;
    bad = g8_or(bad);
    if (bad) {
        if (r == 0u) atomicOr(&misc->err, bad);
        return;
    }
    if (lpc != 0u) {
        const uint32_t dst = it.b + prim_off;
        if (dst < pcap) oprims[dst] = bprims[pl.x];
        if (lpc >= 2u && dst + 1u < pcap) oprims[dst + 1u] = bprims[pl.y];
        if (lpc >= 3u && dst + 2u < pcap) oprims[dst + 2u] = bprims[pl.z];
    }

    // The 80-byte node (bvh2_to_cwbvh.rs:85-170): exponents, quantised child boxes, meta bytes.
    const uint32_t ex = cw_exponent(nb.hx - nb.lx), ey = cw_exponent(nb.hy - nb.ly), ez = cw_exponent(nb.hz - nb.lz);
    const float rx = 1.0f / __uint_as_float(ex << 23), ry = 1.0f / __uint_as_float(ey << 23),
                rz = 1.0f / __uint_as_float(ez << 23);
    uint32_t meta = 0u, qlx = 0u, qhx = 0u, qly = 0u, qhy = 0u, qlz = 0u, qhz = 0u;
    if (occ) {
        qlx = quantise_lo(sb.lx, nb.lx, rx);
        qhx = quantise_hi(sb.hx, nb.lx, rx);
        qly = quantise_lo(sb.ly, nb.ly, ry);
        qhy = quantise_hi(sb.hy, nb.ly, ry);
        qlz = quantise_lo(sb.lz, nb.lz, rz);
        qhz = quantise_hi(sb.hz, nb.lz, rz);
        meta = isint ? ((24u + r) | 0x20u) : (prim_off | (lpc == 1u ? 0x20u : (lpc == 2u ? 0x60u : 0xe0u)));
    }
    const uint2 bm = g8_bytes(meta, r), blx = g8_bytes(qlx, r), bhx = g8_bytes(qhx, r), bly = g8_bytes(qly, r),
                bhy = g8_bytes(qhy, r), blz = g8_bytes(qlz, r), bhz = g8_bytes(qhz, r);
    const uint32_t child_base = 1u + it.a;
    uint4 chunk;
    if (r == 0u)
        chunk = make_uint4(__float_as_uint(nb.lx), __float_as_uint(nb.ly), __float_as_uint(nb.lz),
                           ex | ey << 8 | ez << 16 | imask << 24);
    else if (r == 1u)
        chunk = make_uint4(child_base, it.b, bm.x, bm.y);
    else if (r == 2u)
        chunk = make_uint4(blx.x, blx.y, bhx.x, bhx.y);
    else if (r == 3u)
        chunk = make_uint4(bly.x, bly.y, bhy.x, bhy.y);
    else
        chunk = make_uint4(blz.x, blz.y, bhz.x, bhz.y);
    if (r < 5u) reinterpret_cast<uint4*>(&onodes[it.cw_index])[r] = chunk;

    // Children's DFS numbering (pre-order, child blocks appended at visit time, bvh2_to_cwbvh.rs:172-191); the
    // caller places the items.
    if (num_internal == 0u) return;
    const uint32_t s_cw = (uint32_t)__shfl((int)rcw, sl, 8), s_np = (uint32_t)__shfl((int)rnp, sl, 8);
    const uint32_t s_meta0 = (uint32_t)__shfl((int)(rmeta.x & 0xffu), sl, 8);
    const uint32_t s_kx = (uint32_t)__shfl((int)rkids.x, sl, 8), s_ky = (uint32_t)__shfl((int)rkids.y, sl, 8);
    const uint32_t csz = isint ? s_cw - 1u : 0u;
    const uint32_t cnp = isint ? s_np : 0u;
    const uint32_t a_ex = g8_inclusive_sum(csz, r) - csz, b_ex = g8_inclusive_sum(cnp, r) - cnp;
    eo.num_internal = num_internal;
    eo.isint = isint;
    eo.rank = int_rank;
    eo.item = CwItem{child_base + int_rank, s_ch, it.a + num_internal + a_ex, it.b + num_primitives + b_ex, it.mesh};
    eo.pre = CwPre{sb.lx, sb.ly, sb.lz, sb.hx, sb.hy, sb.hz, s_first, s_pc, s_meta0, s_kx, s_ky, 0u};
}

// One CWBVH level.  Each iteration of the block-uniform loop gives every 8-lane group one node; the groups' next-level
// items are counted in LDS and placed with one global atomic per block (a counter shared by the whole level would
// serialise one atomic per node).
__global__ void __launch_bounds__(T256) k_cw_emit(const Node32* __restrict__ nodes,
                                                  const uint32_t* __restrict__ bvh2_prims,
                                                  DpView dp,
                                                  const CwItem* __restrict__ in_items, CwItem* __restrict__ out_items,
                                                  const CwPre* __restrict__ in_pre, CwPre* __restrict__ out_pre,
                                                  uint32_t level, uint32_t order_children, CwNode* __restrict__ out,
                                                  uint32_t* __restrict__ out_prims, uint32_t cw_cap, uint32_t prim_cap,
                                                  CwMisc* misc, const MeshDesc* __restrict__ meshes) {
    constexpr uint32_t GROUPS = T256 / 8;
    __shared__ uint32_t s_count, s_base;
    const uint32_t r = threadIdx.x & 7u;
    const uint32_t count = misc->level_count[level];
    for (uint32_t first = blockIdx.x * GROUPS; first < count; first += gridDim.x * GROUPS) {
        const uint32_t t = first + (threadIdx.x >> 3);
        EmitOut eo;
        eo.num_internal = 0u;
        eo.isint = 0u;
        eo.rank = 0u;
        if (t < count)
            emit_node(nodes, bvh2_prims, dp, in_items[t], in_pre[t], r, level, order_children, out, out_prims,
                      cw_cap, prim_cap, misc, meshes, eo);
        if (threadIdx.x == 0) s_count = 0u;
        __syncthreads();
        uint32_t off = 0u;
        if (r == 0u && eo.num_internal != 0u) off = atomicAdd(&s_count, eo.num_internal);
        __syncthreads();
        if (threadIdx.x == 0 && s_count != 0u) s_base = atomicAdd(&misc->level_count[level + 1], s_count);
        __syncthreads();
        off = (uint32_t)__shfl((int)off, 0, 8);
        if (eo.isint) {
            out_items[s_base + off + eo.rank] = eo.item;
            out_pre[s_base + off + eo.rank] = eo.pre;
        }
        __syncthreads();  // s_count and s_base are reused by the next iteration
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

// Batched conversions: each mesh's CWBVH node count and depth (from its root's cost-table entry) and its
// level-0 work item.
__global__ void __launch_bounds__(T256) k_b_cw_sizes(const Node32* __restrict__ nodes, DpView dp,
                                                     MeshDesc* __restrict__ meshes, uint32_t m) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i >= m) return;
    MeshDesc md = meshes[i];
    if (md.prim_count == 0u) {
        md.cw_count = 0u;
        md.cw_depth = 0u;
    } else {
        const Node32 root = load_node(&nodes[md.node_base]);
        if (root.prim_count != 0u) {
            md.cw_count = 1u;
            md.cw_depth = 1u;
        } else {
            dp.kbase = md.prim_base;
            md.cw_count = dp.rec[dp_key(dp, root.first_index)].cwsize;
            md.cw_depth = dp.rec[dp_key(dp, root.first_index)].cwdepth;
        }
    }
    meshes[i] = md;
}

__global__ void __launch_bounds__(T256) k_b_cw_roots(const MeshDesc* __restrict__ meshes, uint32_t m,
                                                     CwItem* __restrict__ items, CwMisc* misc,
                                                     const Node32* __restrict__ nodes, DpView dp,
                                                     CwPre* __restrict__ pre) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i == 0) {
        misc->err = 0u;
        for (uint32_t l = 1; l <= MAX_CW_LEVELS; ++l) misc->level_count[l] = 0u;
    }
    if (i >= m) return;
    if (meshes[i].prim_count == 0u) return;
    const uint32_t slot = atomicAdd(&misc->level_count[0], 1u);
    items[slot] = CwItem{0u, 0u, 0u, 0u, i};
    const MeshDesc md = meshes[i];
    DpView d = dp;
    d.kbase = md.prim_base;
    pre[slot] = pre_of(nodes + md.node_base, d, 0u);
}

__global__ void k_zero_level0(CwMisc* misc) {
    if (threadIdx.x == 0) misc->level_count[0] = 0u;
}

// Batched reordering: bvh_tris[g] = tris[prim_base(mesh of g) + primitive_indices[g]] (indices are mesh-local).
__global__ void __launch_bounds__(T256) k_reorder_tris_meshes(const float4* __restrict__ tris,
                                                              const uint32_t* __restrict__ prims,
                                                              const uint32_t* __restrict__ offsets, uint32_t m,
                                                              uint32_t n, float4* __restrict__ out) {
    const uint32_t g = blockIdx.x * T256 + threadIdx.x;
    if (g >= n) return;
    uint32_t lo = 0, hi = m;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (offsets[mid] <= g) lo = mid;
        else hi = mid;
    }
    const size_t s = ((size_t)offsets[lo] + prims[g]) * 3, d = (size_t)g * 3;
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

hipError_t bvh_dp_alloc(obvhs_bvh2* bvh) {
    const size_t k = bvh->prim_count > 1 ? bvh->prim_count - 1 : 1;
    return bvh->dp.ensure(dp_table_bytes(k));
}

DpView bvh_dp_view(const obvhs_bvh2* bvh, uint32_t max_prims) {
    const size_t k = bvh->prim_count > 1 ? bvh->prim_count - 1 : 1;
    return dp_view_at(bvh->dp.as<char>(), k, max_prims);
}

obvhs_status cwbvh_convert(obvhs_context* ctx, obvhs_bvh2* bvh2, uint32_t max_prims_per_leaf,
                           bool order_children, obvhs_cwbvh* out) {
    hipStream_t st = ctx->stream;
    const uint32_t N = bvh2->node_count;
    out->node_count = 0;
    out->prim_count = 0;
    memset(&out->total_aabb, 0, sizeof(out->total_aabb));
    if (N == 0) return OBVHS_OK;  // bvh2_to_cwbvh.rs:496-498
    const uint32_t n = bvh2->prim_count;
    OBVHS_TRY(ctx->scratch.cw_misc.ensure(sizeof(CwMisc)), "alloc cwbvh misc");
    CwMisc* misc = ctx->scratch.cw_misc.as<CwMisc>();
    const Node32* nodes = bvh2->nodes.as<Node32>();
    k_cw_init<<<1, 64, 0, st>>>(misc);
    uint32_t* h = static_cast<uint32_t*>(ctx->pinned);
    if (bvh2->dp_max_prims != max_prims_per_leaf && n > 1 && !bvh2->pass_ordered) {
        // children moved by reinsertion: evaluate the table over the level lists, deepest level first
        OBVHS_TRY(bvh_dp_alloc(bvh2), "alloc cost table");
        const obvhs_status ls = cost_table_levels(ctx, bvh2, bvh_dp_view(bvh2, max_prims_per_leaf), &misc->err);
        if (ls != OBVHS_OK) return ls;
        bvh2->dp_max_prims = max_prims_per_leaf;
    }
    if (bvh2->dp_max_prims != max_prims_per_leaf && n > 1) {
        OBVHS_TRY(bvh_dp_alloc(bvh2), "alloc cost table");
        const DpView dpv = bvh_dp_view(bvh2, max_prims_per_leaf);
        const uint32_t passes = bvh2->ploc_iterations;
        std::vector<uint32_t> hist(passes + 1);
        OBVHS_TRY(hipMemcpyAsync(hist.data(), bvh2->hist.ptr, hist.size() * 4, hipMemcpyDeviceToHost, st),
                  "read pass history");
        OBVHS_TRY(hipStreamSynchronize(st), "pass history sync");
        uint32_t p = 0;
        for (; p < passes; ++p) {
            const uint32_t lo = hist[p + 1], hi = hist[p];
            if (hi - lo <= DP_SMALL_RANGE) break;
            k_dp_range<<<(hi - lo + T256 - 1) / T256, T256, 0, st>>>(nodes, dpv, lo, hi, misc);
        }
        k_dp_small<<<1, 1024, 0, st>>>(nodes, dpv, bvh2->hist.as<uint32_t>(), p, passes, misc);
        OBVHS_TRY(hipGetLastError(), "cost table launch");
        bvh2->dp_max_prims = max_prims_per_leaf;
    }
    const DpView dp = bvh_dp_view(bvh2, max_prims_per_leaf);
    // Root, CWBVH node count and depth (the root's subtree size and levels) and device errors: one copy.
    k_cw_info<<<1, 64, 0, st>>>(nodes, dp, misc);
    OBVHS_TRY(hipGetLastError(), "cwbvh info launch");
    OBVHS_TRY(hipMemcpyAsync(h, misc, sizeof(CwMisc), hipMemcpyDeviceToHost, st), "read cw info");
    OBVHS_TRY(hipStreamSynchronize(st), "cwbvh dp sync");
    CwMisc hm;
    memcpy(&hm, h, sizeof(CwMisc));
    if (hm.err) return device_error_to_status(hm.err);
    Node32 root;
    memcpy(&root, hm.info, sizeof(Node32));
    const uint32_t ncw = hm.info[8], levels = hm.info[9];
    if (ncw == 0u) return OBVHS_ERR_CW_DISTRIBUTE;  // some CWBVH node's children cannot be distributed
    if (ncw > N) return OBVHS_ERR_CAPACITY;
    if (levels > MAX_CW_LEVELS) return OBVHS_ERR_DEPTH;
    OBVHS_TRY(out->nodes.ensure((size_t)ncw * sizeof(CwNode)), "alloc cwbvh nodes");
    OBVHS_TRY(out->prim_indices.ensure((size_t)n * 4 + 4), "alloc cwbvh primitive indices");
    // Item lists (and their node data) for two levels; a level holds at most ncw nodes.  Level 0 reads the root's
    // from the misc block.
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~size_t(255);
        return o;
    };
    const size_t o_items0 = carve((size_t)ncw * sizeof(CwItem)), o_items1 = carve((size_t)ncw * sizeof(CwItem)),
                 o_pre0 = carve((size_t)ncw * sizeof(CwPre)), o_pre1 = carve((size_t)ncw * sizeof(CwPre));
    OBVHS_TRY(ctx->scratch.cw_scratch.ensure(off), "alloc cwbvh scratch");
    char* base = ctx->scratch.cw_scratch.as<char>();
    CwItem* items[2] = {reinterpret_cast<CwItem*>(base + o_items0), reinterpret_cast<CwItem*>(base + o_items1)};
    CwPre* pres[2] = {reinterpret_cast<CwPre*>(base + o_pre0), reinterpret_cast<CwPre*>(base + o_pre1)};
    // One launch per level, no host round trips: a level holds at most min(8^level, ncw) nodes.
    uint64_t bound = 1;
    for (uint32_t level = 0; level < levels; ++level) {
        const uint64_t cap = bound < ncw ? bound : ncw;
        const uint32_t g = (uint32_t)std::min<uint64_t>((cap * 8 + T256 - 1) / T256, EMIT_MAX_BLOCKS);
        const CwItem* in_items = level == 0 ? &misc->root_item : items[level & 1];
        const CwPre* in_pre = level == 0 ? &misc->root_pre : pres[level & 1];
        k_cw_emit<<<g, T256, 0, st>>>(nodes, bvh2->prim_indices.as<uint32_t>(), dp, in_items,
                                      items[(level + 1) & 1], in_pre, pres[(level + 1) & 1], level,
                                      order_children ? 1u : 0u, out->nodes.as<CwNode>(),
                                      out->prim_indices.as<uint32_t>(), ncw, n, misc, nullptr);
        bound *= 8;
    }
    OBVHS_TRY(hipGetLastError(), "cwbvh emit launch");
    // the error word and the count of the level below the last (err and level_count are adjacent): one copy
    OBVHS_TRY(hipMemcpyAsync(h, misc, (2 + (size_t)levels) * 4, hipMemcpyDeviceToHost, st), "read cw err");
    OBVHS_TRY(hipStreamSynchronize(st), "cwbvh emit sync");
    if (h[0]) return device_error_to_status(h[0]);
    if (h[1 + levels] != 0u) return OBVHS_ERR_DEPTH;  // deeper than the depth table said (cannot happen)
    // total_aabb = bvh2.nodes[0].aabb (bvh2_to_cwbvh.rs:506)
    out->total_aabb.min = obvhs_vec3a{root.minx, root.miny, root.minz, 0.0f};
    out->total_aabb.max = obvhs_vec3a{root.maxx, root.maxy, root.maxz, 0.0f};
    out->node_count = ncw;
    out->prim_count = n;
    return OBVHS_OK;
}

obvhs_status cwbvh_convert_batch(obvhs_context* ctx, const Node32* nodes, const uint32_t* bvh2_prims,
                                 const DpView& dp, MeshDesc* d_meshes, std::vector<MeshDesc>& meshes,
                                 uint32_t n_total, bool order_children, DevBuf& out_nodes, DevBuf& out_prims,
                                 uint32_t& total_cw) {
    hipStream_t st = ctx->stream;
    const uint32_t m = (uint32_t)meshes.size();
    total_cw = 0;
    if (m == 0) return OBVHS_OK;
    k_b_cw_sizes<<<(m + T256 - 1) / T256, T256, 0, st>>>(nodes, dp, d_meshes, m);
    OBVHS_TRY(hipGetLastError(), "batch cw sizes");
    OBVHS_TRY(hipMemcpyAsync(meshes.data(), d_meshes, m * sizeof(MeshDesc), hipMemcpyDeviceToHost, st),
              "read mesh table");
    OBVHS_TRY(hipStreamSynchronize(st), "mesh table sync");
    uint64_t total = 0;
    uint32_t levels = 0, active = 0;
    for (auto& md : meshes) {
        if (md.prim_count > 0 && md.cw_count == 0) return OBVHS_ERR_CW_DISTRIBUTE;
        md.cw_base = (uint32_t)total;
        total += md.cw_count;
        levels = md.cw_depth > levels ? md.cw_depth : levels;
        active += md.prim_count > 0 ? 1u : 0u;
    }
    if (total >= (1ull << 32)) return OBVHS_ERR_CAPACITY;
    if (levels > MAX_CW_LEVELS) return OBVHS_ERR_DEPTH;
    total_cw = (uint32_t)total;
    OBVHS_TRY(hipMemcpyAsync(d_meshes, meshes.data(), m * sizeof(MeshDesc), hipMemcpyHostToDevice, st),
              "write mesh table");
    OBVHS_TRY(out_nodes.ensure((size_t)total * sizeof(CwNode) + 80), "alloc batch cwbvh nodes");
    OBVHS_TRY(out_prims.ensure((size_t)n_total * 4 + 4), "alloc batch cwbvh primitive indices");
    // Level-0 items: every non-empty mesh's root.
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~size_t(255);
        return o;
    };
    // a level holds at most `total` nodes
    const size_t cap = (size_t)total;
    const size_t o_items0 = carve(cap * sizeof(CwItem)), o_items1 = carve(cap * sizeof(CwItem)),
                 o_pre0 = carve(cap * sizeof(CwPre)), o_pre1 = carve(cap * sizeof(CwPre));
    OBVHS_TRY(ctx->scratch.cw_scratch.ensure(off), "alloc cwbvh scratch");
    OBVHS_TRY(ctx->scratch.cw_misc.ensure(sizeof(CwMisc)), "alloc cwbvh misc");
    char* base = ctx->scratch.cw_scratch.as<char>();
    CwItem* items[2] = {reinterpret_cast<CwItem*>(base + o_items0), reinterpret_cast<CwItem*>(base + o_items1)};
    CwPre* pres[2] = {reinterpret_cast<CwPre*>(base + o_pre0), reinterpret_cast<CwPre*>(base + o_pre1)};
    CwMisc* misc = ctx->scratch.cw_misc.as<CwMisc>();
    k_zero_level0<<<1, 64, 0, st>>>(misc);
    k_b_cw_roots<<<(m + T256 - 1) / T256, T256, 0, st>>>(d_meshes, m, items[0], misc, nodes, dp, pres[0]);
    uint64_t bound = active;
    for (uint32_t level = 0; level < levels; ++level) {
        const uint64_t capl = bound < total ? bound : total;
        if (capl == 0) continue;
        const uint32_t g = (uint32_t)std::min<uint64_t>((capl * 8 + T256 - 1) / T256, EMIT_MAX_BLOCKS);
        k_cw_emit<<<g, T256, 0, st>>>(nodes, bvh2_prims, dp, items[level & 1], items[(level + 1) & 1],
                                      pres[level & 1], pres[(level + 1) & 1], level, order_children ? 1u : 0u,
                                      out_nodes.as<CwNode>(), out_prims.as<uint32_t>(), 0u, 0u, misc,
                                      d_meshes);
        bound *= 8;
    }
    OBVHS_TRY(hipGetLastError(), "batch cwbvh emit launch");
    uint32_t* h = static_cast<uint32_t*>(ctx->pinned);
    OBVHS_TRY(hipMemcpyAsync(h, misc, (2 + (size_t)levels) * 4, hipMemcpyDeviceToHost, st), "read cw err");
    OBVHS_TRY(hipStreamSynchronize(st), "batch cwbvh emit sync");
    if (h[0]) return device_error_to_status(h[0]);
    if (h[1 + levels] != 0u) return OBVHS_ERR_DEPTH;
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

namespace obvhs {
hipError_t launch_reorder_triangles_meshes(const obvhs_triangle* tris, const uint32_t* prim_indices,
                                           const uint32_t* offsets, uint32_t m, uint32_t n, obvhs_triangle* out,
                                           hipStream_t stream) {
    if (n == 0) return hipSuccess;
    k_reorder_tris_meshes<<<(n + T256 - 1) / T256, T256, 0, stream>>>(
        reinterpret_cast<const float4*>(tris), prim_indices, offsets, m, n, reinterpret_cast<float4*>(out));
    return hipGetLastError();
}
}  // namespace obvhs
