// The BVH2 -> CWBVH cost table (calculate_cost_impl, cwbvh/bvh2_to_cwbvh.rs:220-344) and get_children
// (bvh2_to_cwbvh.rs:346-397) over a BVH2 whose internal nodes are identified by their first child slot.
//
// A PLOC internal node's children always exist before the node does, and the node's own array slot is assigned
// only later, when it is merged in turn.  Its first_index (the slot of its left child, always odd) is known at
// creation, so the table is keyed by first_index >> 1.  That lets the merge that creates a node evaluate the
// node's entry immediately (children finished earlier, no cross-workgroup synchronisation), and the same
// functions serve a standalone pass over a finished BVH2.  Leaves need no entry: all seven of their decisions are
// LEAF with cost half_area * 1 * 0.3 (bvh2_to_cwbvh.rs:249-256).
//
// Besides the reference's seven (cost, kind, distribute_left, distribute_right) decisions, each entry holds, per
// decision i, sub[i] = the number of CWBVH nodes in the subtrees of the INTERNAL children that get_children(node, i)
// would produce.  It obeys the same recurrence as the costs (children only), so no tree walk is needed, and
// cwsize = 1 + sub[0] is the size of the CWBVH subtree the node roots if it becomes a CWBVH node -- what the
// top-down pass needs to number nodes in the reference's pre-order (bvh2_to_cwbvh.rs:172-191).  The same
// recurrence with max instead of + gives the depth of that CWBVH subtree, so the top-down pass knows its level
// count without reading anything back.  Each entry also keeps the children's first_index words and, for a subtree
// of at most 3 primitives, its leaf list: the top-down pass walks get_children and collects a leaf child's
// primitives from the table, one dependent load per level instead of a node and then its decisions.
#pragma once

#include "common.hpp"

namespace obvhs {

constexpr uint32_t KIND_LEAF = 0, KIND_INTERNAL = 1, KIND_DISTRIBUTE = 2;  // bvh2_to_cwbvh.rs:470-476
constexpr uint32_t DIDX_INVALID = 7u;      // distribute index "INVALID" (u8::MAX in the reference) in 3 bits
constexpr uint32_t SUB_INVALID = 0xffffffffu;
constexpr float PRIM_COST = 0.3f;  // bvh2_to_cwbvh.rs:30
constexpr uint32_t KID_LEAF = 0x80000000u;  // leaf flag in DpView::kids

// One table entry, a 128-byte cache line: every read of a node's entry (a parent's evaluation, a step of the
// get_children walk, the emit of a leaf list or a child's sizes) touches a single line.
// Decision byte: kind (2 bits) | distribute_left << 2 (3 bits) | distribute_right << 5 (3 bits).
struct __align__(128) DpRec {
    float cost[7];
    uint32_t nprims;    // primitives under the node
    uint32_t sub[7];
    uint32_t cwsize;    // 1 + sub[0], or 0 when the node's CWBVH subtree cannot be built
    uint2 meta;         // seven decision bytes (byte 7 unused)
    uint2 subdepth;     // seven bytes: CWBVH depth below the children get_children(node, i) produces
    uint2 kids;         // the children's first_index (bit 31 set for a leaf), left then right
    uint32_t cwdepth;   // 1 + subdepth[0]: levels of the CWBVH subtree the node roots
    uint32_t pad0;
    uint4 plist;        // for a subtree of at most 3 primitives: its leaves' first_index, left first, and count
    uint32_t pad1[4];
};
static_assert(sizeof(DpRec) == 128, "one cache line per entry");

struct DpView {
    DpRec* rec;         // [k]
    uint32_t max_prims; // max_prims_per_leaf the table was evaluated for
    uint32_t kbase;     // key offset of the BVH2 being processed (batched builds: the mesh's first primitive)
};

// One BVH of a batched build: slices of the shared node / primitive / cost-table / CWBVH arrays.
struct MeshDesc {
    uint32_t node_base;   // first BVH2 slot (2 * prim_base, so cost-table keys stay disjoint)
    uint32_t prim_base;   // first primitive; also the mesh's cost-table key base
    uint32_t prim_count;
    uint32_t cw_base;     // first CWBVH node
    uint32_t cw_count;
    uint32_t cw_depth;
    uint32_t radius;      // the mesh's PLOC search distance
    uint32_t pad;
};

// Table of k entries (the base must be 128-byte aligned; device allocations are 256-byte aligned).
__host__ __device__ __forceinline__ size_t dp_table_bytes(size_t k) { return k * sizeof(DpRec) + 1024; }
__host__ __device__ __forceinline__ DpView dp_view_at(char* base, size_t k, uint32_t max_prims) {
    (void)k;
    DpView v;
    v.rec = reinterpret_cast<DpRec*>(base);
    v.max_prims = max_prims;
    v.kbase = 0u;
    return v;
}

__host__ __device__ __forceinline__ uint32_t dkey(uint32_t first_index) { return first_index >> 1; }
__device__ __forceinline__ size_t dp_key(const DpView& dp, uint32_t first_index) {
    return (size_t)dp.kbase + (first_index >> 1);
}
__device__ __forceinline__ uint32_t dbyte(uint2 m, uint32_t i) {
    return ((i < 4u ? m.x : m.y) >> (8u * (i & 3u))) & 0xffu;
}
__device__ __forceinline__ uint32_t dkind(uint32_t b) { return b & 3u; }
__device__ __forceinline__ uint32_t dleft(uint32_t b) { return (b >> 2) & 7u; }
__device__ __forceinline__ uint32_t dright(uint32_t b) { return (b >> 5) & 7u; }
__device__ __forceinline__ uint32_t dpack(uint32_t k, uint32_t l, uint32_t r) { return k | (l << 2) | (r << 5); }

// Decision i of an internal node's entry.
__device__ __forceinline__ uint32_t node_decision(const DpView& dp, uint32_t first_index, uint32_t i) {
    return dbyte(dp.rec[dp_key(dp, first_index)].meta, i);
}
__device__ __forceinline__ uint32_t node_kind(const DpView& dp, const Node32& nx, uint32_t i) {
    return nx.prim_count != 0u ? KIND_LEAF : dkind(node_decision(dp, nx.first_index, i));
}

// Everything the parent's evaluation needs from one child.  (Per-decision values selected by a runtime index are
// read from memory, not from a register array, so the evaluation stays free of private-memory spills.)
struct ChildDp {
    float cost[7];
    uint2 meta;       // seven decision bytes
    size_t key;
    uint2 subdepth;
    uint32_t nprims;
    uint32_t cwsize;
    uint32_t cwdepth;
    bool leaf;
};

__device__ __forceinline__ void load_child_dp(const DpView& dp, const Node32& c, ChildDp& o) {
    o.leaf = c.prim_count != 0u;
    if (o.leaf) {
        const float leaf = half_area(node_box(c)) * (float)c.prim_count * PRIM_COST;
#pragma unroll
        for (int k = 0; k < 7; ++k) o.cost[k] = leaf;
        o.meta = make_uint2(0u, 0u);
        o.subdepth = make_uint2(0u, 0u);
        o.key = 0;
        o.nprims = c.prim_count;
        o.cwsize = 1u;
        o.cwdepth = 1u;
    } else {
        o.key = dp_key(dp, c.first_index);
        const DpRec& e = dp.rec[o.key];
        const float4 c0 = *reinterpret_cast<const float4*>(&e.cost[0]);
        const float4 c1 = *reinterpret_cast<const float4*>(&e.cost[4]);  // cost[4..6], nprims
        const uint4 m = *reinterpret_cast<const uint4*>(&e.meta);         // meta, subdepth
        o.cost[0] = c0.x;
        o.cost[1] = c0.y;
        o.cost[2] = c0.z;
        o.cost[3] = c0.w;
        o.cost[4] = c1.x;
        o.cost[5] = c1.y;
        o.cost[6] = c1.z;
        o.nprims = __float_as_uint(c1.w);
        o.meta = make_uint2(m.x, m.y);
        o.subdepth = make_uint2(m.z, m.w);
        o.cwsize = e.cwsize;
        o.cwdepth = e.cwdepth;
    }
}

// Contribution of child c taken through its decision d: its own CWBVH subtree when it becomes a CWBVH child,
// the contributions of its distributed descendants when it is DISTRIBUTE, nothing when it is a leaf child.
__device__ __forceinline__ uint32_t child_contribution(const DpView& dp, const ChildDp& c, uint32_t d) {
    if (c.leaf) return 0u;
    if (dkind(dbyte(c.meta, d)) == KIND_DISTRIBUTE) return dp.rec[c.key].sub[d];
    if (dkind(dbyte(c.meta, 0u)) == KIND_INTERNAL) return c.cwsize == 0u ? SUB_INVALID : c.cwsize;
    return 0u;
}

// Depth counterpart of child_contribution.
__device__ __forceinline__ uint32_t child_depth(const ChildDp& c, uint32_t d) {
    if (c.leaf) return 0u;
    if (dkind(dbyte(c.meta, d)) == KIND_DISTRIBUTE) return dbyte(c.subdepth, d);
    if (dkind(dbyte(c.meta, 0u)) == KIND_INTERNAL) return c.cwdepth;
    return 0u;
}

// Evaluates the entry of the internal node with box `box` whose children records are `l` and `r` (slots first,
// first + 1) and stores it under its key.  Returns a DERR_* bit or 0.
__device__ __forceinline__ uint32_t dp_evaluate(const DpView& dp, const Box& box, uint32_t first, const Node32& l,
                                                const Node32& r) {
    if ((l.prim_count != 0u && l.prim_count != 1u) || (r.prim_count != 0u && r.prim_count != 1u))
        return DERR_CW_NONUNIT_LEAF;  // bvh2_to_cwbvh.rs:243-247
    ChildDp L, R;
    load_child_dp(dp, l, L);
    load_child_dp(dp, r, R);
    const uint32_t num_primitives = L.nprims + R.nprims;
    const float ha = half_area(box);
    float cost[7];
    uint32_t meta[7];
    {
        const float cost_leaf =
            num_primitives <= dp.max_prims ? ((float)num_primitives) * ha * PRIM_COST : INFINITY;
        float cost_distribute = INFINITY;
        uint32_t dl = DIDX_INVALID, dr = DIDX_INVALID;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const float c = L.cost[k] + R.cost[6 - k];
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
        uint32_t dl = DIDX_INVALID, dr = DIDX_INVALID;
#pragma unroll
        for (int k = 0; k < i; ++k) {
            const float c = L.cost[k] + R.cost[i - k - 1];
            if (c < cost_distribute) {
                cost_distribute = c;
                dl = (uint32_t)k;
                dr = (uint32_t)(i - k - 1);
            }
        }
        if (dl != DIDX_INVALID) {
            cost[i] = cost_distribute;
            meta[i] = dpack(KIND_DISTRIBUTE, dl, dr);
        } else {
            cost[i] = cost[i - 1];
            meta[i] = meta[i - 1];
        }
    }
    uint32_t sub[7], sd[7];
#pragma unroll
    for (int i = 0; i < 7; ++i) {
        const uint32_t dl = dleft(meta[i]), dr = dright(meta[i]);
        if (dl == DIDX_INVALID || dr == DIDX_INVALID) {
            sub[i] = SUB_INVALID;
            sd[i] = 0u;
        } else {
            const uint32_t a = child_contribution(dp, L, dl), b = child_contribution(dp, R, dr);
            sub[i] = (a == SUB_INVALID || b == SUB_INVALID) ? SUB_INVALID : a + b;
            sd[i] = min(255u, max(child_depth(L, dl), child_depth(R, dr)));
        }
    }
    uint4 pl = make_uint4(0u, 0u, 0u, 0u);
    if (num_primitives <= 3u) {
        // The leaves count_primitives (bvh2_to_cwbvh.rs:197-211) visits, left subtree first.  Each child holds at
        // least one leaf, so the left list has one or two entries.
        uint4 a = make_uint4(l.first_index, 0u, 0u, 1u), c = make_uint4(r.first_index, 0u, 0u, 1u);
        if (!L.leaf) a = dp.rec[L.key].plist;
        if (!R.leaf) c = dp.rec[R.key].plist;
        pl = a.w >= 2u ? make_uint4(a.x, a.y, c.x, num_primitives) : make_uint4(a.x, c.x, c.y, num_primitives);
    }
    // The whole 128-byte entry in eight 16-byte stores (the padding too, so that no 32-byte sector of the line is
    // left partly written).
    uint4* e = reinterpret_cast<uint4*>(&dp.rec[dp_key(dp, first)]);
    e[0] = make_uint4(__float_as_uint(cost[0]), __float_as_uint(cost[1]), __float_as_uint(cost[2]),
                      __float_as_uint(cost[3]));
    e[1] = make_uint4(__float_as_uint(cost[4]), __float_as_uint(cost[5]), __float_as_uint(cost[6]), num_primitives);
    e[2] = make_uint4(sub[0], sub[1], sub[2], sub[3]);
    e[3] = make_uint4(sub[4], sub[5], sub[6], sub[0] == SUB_INVALID ? 0u : 1u + sub[0]);
    e[4] = make_uint4(meta[0] | meta[1] << 8 | meta[2] << 16 | meta[3] << 24, meta[4] | meta[5] << 8 | meta[6] << 16,
                      sd[0] | sd[1] << 8 | sd[2] << 16 | sd[3] << 24, sd[4] | sd[5] << 8 | sd[6] << 16);
    e[5] = make_uint4(l.first_index | (L.leaf ? KID_LEAF : 0u), r.first_index | (R.leaf ? KID_LEAF : 0u), 1u + sd[0],
                      0u);
    e[6] = pl;
    e[7] = make_uint4(0u, 0u, 0u, 0u);
    return 0u;
}

// get_children(node, 0) from the cost table alone: the node's children are the slots first and first + 1 and their
// first_index values are in its `kids` entry, so each level of the walk loads one entry (decisions and children
// together) instead of a node and then its decisions.  Children are appended left-first like the reference's
// recursion (bvh2_to_cwbvh.rs:346-397).  Outputs per child: its slot, and its decision-0 kind in 2 bits of *kinds.
// Returns the count, -1 when more than 8 children would be produced, -2 on a malformed decision (distribute index
// invalid).
__device__ __forceinline__ int get_children_kids(const DpView& dp, uint32_t first, uint32_t meta0, uint32_t* children,
                                                 uint32_t* kinds) {
    uint32_t stk_slot[8], stk_info[8], stk_i[8];
    int sp = 0;
    int cc = 0;
    uint32_t kb = 0u;
    const uint32_t dl = dleft(meta0), dr = dright(meta0);
    if (dl == DIDX_INVALID || dr == DIDX_INVALID) return -2;
    const uint2 k0 = dp.rec[dp_key(dp, first)].kids;
    stk_slot[0] = first + 1u;
    stk_info[0] = k0.y;
    stk_i[0] = dr;
    stk_slot[1] = first;
    stk_info[1] = k0.x;
    stk_i[1] = dl;
    sp = 2;
    while (sp > 0) {
        sp--;
        const uint32_t x = stk_slot[sp], info = stk_info[sp], i = stk_i[sp];
        uint32_t kind0 = KIND_LEAF;
        if (!(info & KID_LEAF)) {  // a leaf is never DISTRIBUTE
            const uint32_t f = info & ~KID_LEAF;
            const size_t key = dp_key(dp, f);
            const uint2 m = dp.rec[key].meta;
            const uint2 kx = dp.rec[key].kids;  // same line as the decisions; used when the child is distributed
            const uint32_t b = dbyte(m, i);
            if (dkind(b) == KIND_DISTRIBUTE) {
                const uint32_t cl = dleft(b), cr = dright(b);
                if (cl == DIDX_INVALID || cr == DIDX_INVALID) return -2;
                if (sp + 2 > 8) return -1;
                stk_slot[sp] = f + 1u;
                stk_info[sp] = kx.y;
                stk_i[sp] = cr;
                sp++;
                stk_slot[sp] = f;
                stk_info[sp] = kx.x;
                stk_i[sp] = cl;
                sp++;
                continue;
            }
            kind0 = dkind(dbyte(m, 0u));
        }
        if (cc >= 8) return -1;
        children[cc] = x;
        kb |= kind0 << (2 * cc);
        cc++;
    }
    *kinds = kb;
    return cc;
}

}  // namespace obvhs
