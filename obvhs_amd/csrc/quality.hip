// Tree utilities and the build-quality passes on a device-resident BVH2 (Node32 layout).
//
//   levels       : every node in exactly one top-down frontier list (root = level 0), built one launch per level;
//                  bottom-up sweeps then walk the lists from the deepest level to the root.
//   refit_all    : bvh2/mod.rs:527-569 -- each internal box becomes the union of its children, deepest level first.
//   cost table   : the CWBVH cost table (cwdp.hpp) for a BVH2 whose children are no longer in PLOC pass order.
//   reinsertion  : ReinsertionOptimizer::run (bvh2/reinsertion.rs:40-204) in parallel form.  Per ratio step:
//                  candidates = nodes 1..2*node_count sorted by descending half area (stable radix sort of the
//                  crate's f32 key of -cost, ties by node index); find_reinsertion (reinsertion.rs:247-340) for the
//                  first node_count-1 candidates, one lane each, with its stack in HBM; the reinsertions sorted by
//                  descending area_diff (ties by candidate rank); the crate's greedy conflict filter
//                  (reinsertion.rs:180-203) as a parallel greedy maximal independent set over the five conflict
//                  nodes of each reinsertion -- identical accept decisions, because a conflict node's parent can
//                  only have changed through an accepted reinsertion that already touched it; the accepted moves
//                  (reinsertion.rs:342-382) applied in parallel (their node sets are disjoint); then one full
//                  refit, which leaves every internal box equal to the union of its children, as the crate's
//                  refit_from_fast sequence does.  Where that sequence stops early a box keeps its old bits, which
//                  differ from the union's only in the sign of a zero; so when a leaf holds a -0 coordinate the
//                  accepted moves and their walks run as the crate runs them, one after another (k_reins_replay).
//   pre-splits   : cwbvh/builder.rs:26-55 + splits.rs:16-158.  Largest and average half area (the average summed in
//                  triangle order by one wave, as the crate sums it); per iteration every candidate's best of 11
//                  split planes in parallel, appended right halves numbered by an order-preserving scan, the
//                  candidate list extended and filtered in order.
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.hpp"
#include "cwdp.hpp"
#include "device_util.hpp"
#include "internal.hpp"

namespace obvhs {

namespace {

constexpr int T256 = 256;
constexpr uint32_t LEVELS_PER_SYNC = 16;
constexpr uint32_t MIS_ROUNDS_PER_SYNC = 8;
constexpr uint32_t FIND_THREADS = 1u << 19;  // concurrent find_reinsertion lanes (each owns a stack column)
constexpr uint32_t FIND_REFILL = 32;  // idle lanes of a wave that take new candidates together (16: 0.4 ms slower)
constexpr uint32_t FIND_LDS = 8;  // top stack entries in LDS: 16 KB per block, 32 waves per CU (16 entries: 16% slower)

__device__ __forceinline__ uint32_t sibling_of(uint32_t id) { return (id & 1u) ? id + 1u : id - 1u; }
__device__ __forceinline__ uint32_t left_sibling_of(uint32_t id) { return (id & 1u) ? id : id - 1u; }

// ---------------------------------------------------------------------------------------------------------------
// levels
// ---------------------------------------------------------------------------------------------------------------

// meta layout: start[L] = meta[2L], count[L] = meta[2L + 1].
__global__ void __launch_bounds__(T256) k_levels_init(uint32_t* __restrict__ meta, uint32_t words,
                                                      uint32_t* __restrict__ list) {
    for (uint32_t i = blockIdx.x * T256 + threadIdx.x; i < words; i += gridDim.x * T256) meta[i] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        list[0] = 0u;
        meta[1] = 1u;
    }
}

// Appends the children of level `level` as level + 1.  Each block takes a contiguous chunk of the level and
// reserves its output with one atomic (the order inside a level does not matter).  A node whose children would fall
// outside the node array, or more than n_nodes entries in total (not a tree), sets the error word meta[err_word]
// instead of writing.
// Children of an inner node sit at (first, first + 1) inside the array.  The partial rebuild's slot reuse also needs
// left children at odd slots (Bvh2Node::is_left_sibling, bvh2/node.rs:154-156), as every builder lays them out.
__device__ __forceinline__ bool child_pair_ok(uint32_t first, uint32_t n_nodes, bool odd_left) {
    return first != 0u && first < n_nodes - 1u && (!odd_left || (first & 1u) != 0u);
}

// With `expand` set, only the root and the inner nodes x with expand[x] != 0 are opened (the top-down walk of
// PlocBuilder::partial_rebuild, ploc/rebuild.rs:113-126); the lists then hold the slots that walk visits.
__global__ void __launch_bounds__(T256) k_levels_expand(const Node32* __restrict__ nodes, uint32_t n_nodes,
                                                        uint32_t* __restrict__ list, uint32_t* __restrict__ meta,
                                                        uint32_t level, uint32_t err_word,
                                                        const uint8_t* __restrict__ expand) {
    __shared__ uint32_t wcount[T256 / 64];
    __shared__ uint32_t block_base;
    const uint32_t s = meta[2 * level], c = meta[2 * level + 1];
    const uint32_t out = s + c;
    if (blockIdx.x == 0 && threadIdx.x == 0) meta[2 * (level + 1)] = out;
    const uint32_t per_block = ((c + gridDim.x - 1) / gridDim.x + T256 - 1) / T256 * T256;
    const uint32_t b0 = blockIdx.x * per_block;
    if (b0 >= c) return;
    const uint32_t b1 = min(c, b0 + per_block);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    // pass 1: count the block's internal nodes
    uint32_t mine = 0;
    for (uint32_t i = b0 + threadIdx.x; i < b1; i += T256) {
        const uint32_t x = list[s + i];
        const Node32 nd = load_node(&nodes[x]);
        if (nd.prim_count == 0u && (!expand || x == 0u || expand[x] != 0u)) {
            if (!child_pair_ok(nd.first_index, n_nodes, expand != nullptr)) atomicOr(&meta[err_word], 1u);
            else mine++;
        }
    }
    uint32_t wsum = mine;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) wsum += (uint32_t)__shfl_xor((int)wsum, off);
    if (lane == 0) wcount[wave] = wsum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < T256 / 64; ++w) tot += wcount[w];
        block_base = atomicAdd(&meta[2 * (level + 1) + 1], 2u * tot);
    }
    __syncthreads();
    // pass 2: write the children at block_base + 2 * (rank among the block's internal nodes)
    uint32_t base = block_base;
    for (uint32_t i0 = b0; i0 < b1; i0 += T256) {
        const uint32_t i = i0 + threadIdx.x;
        bool inner = false;
        uint32_t first = 0u;
        if (i < b1) {
            const uint32_t x = list[s + i];
            const Node32 nd = load_node(&nodes[x]);
            first = nd.first_index;
            inner = nd.prim_count == 0u && (!expand || x == 0u || expand[x] != 0u) &&
                    child_pair_ok(first, n_nodes, expand != nullptr);
        }
        const uint64_t m = __ballot(inner);
        __syncthreads();
        if (lane == 0) wcount[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int w = 0; w < T256 / 64; ++w) {
            if (w < (int)wave) before += wcount[w];
            total += wcount[w];
        }
        if (inner) {
            const uint64_t at = (uint64_t)out + base + 2u * (before + (uint32_t)__popcll(m & lanemask_lt()));
            if (at + 1 < n_nodes) {
                list[at] = first;
                list[at + 1] = first + 1u;
            } else {
                atomicOr(&meta[err_word], 1u);
            }
        }
        base += 2u * total;
    }
}

__global__ void __launch_bounds__(T256) k_refit_level(Node32* __restrict__ nodes, const uint32_t* __restrict__ list,
                                                      uint32_t start, uint32_t count) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i >= count) return;
    const uint32_t x = list[start + i];
    Node32 nd = load_node(&nodes[x]);
    if (nd.prim_count != 0u) return;
    const Box u =
        box_union(node_box(load_node(&nodes[nd.first_index])), node_box(load_node(&nodes[nd.first_index + 1])));
    store_node(&nodes[x], make_node(u, 0u, nd.first_index));
}

// The last refit of the reinsertion passes fused with the conversion's cost table (k_refit_level, then k_dp_level,
// in one pass): the same union of the two children, in the same operand order, and the same table evaluation.
template <bool REFIT>
__global__ void __launch_bounds__(T256) k_refit_dp_level(Node32* __restrict__ nodes, DpView dp,
                                                         const uint32_t* __restrict__ list, uint32_t start,
                                                         uint32_t count, uint32_t* __restrict__ err) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i >= count) return;
    const uint32_t x = list[start + i];
    const Node32 nx = load_node(&nodes[x]);
    if (nx.prim_count != 0u) return;
    const Node32 l = load_node(&nodes[nx.first_index]);
    const Node32 r = load_node(&nodes[nx.first_index + 1]);
    Box b = node_box(nx);
    if (REFIT) {
        b = box_union(node_box(l), node_box(r));
        store_node(&nodes[x], make_node(b, 0u, nx.first_index));
    }
    const uint32_t e = dp_evaluate(dp, b, nx.first_index, l, r);
    if (e) atomicOr(err, e);
}

__global__ void __launch_bounds__(T256) k_dp_level(const Node32* __restrict__ nodes, DpView dp,
                                                   const uint32_t* __restrict__ list, uint32_t start, uint32_t count,
                                                   uint32_t* __restrict__ err) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i >= count) return;
    const uint32_t x = list[start + i];
    const Node32 nx = load_node(&nodes[x]);
    if (nx.prim_count != 0u) return;
    const Node32 l = load_node(&nodes[nx.first_index]);
    const Node32 r = load_node(&nodes[nx.first_index + 1]);
    const uint32_t e = dp_evaluate(dp, node_box(nx), nx.first_index, l, r);
    if (e) atomicOr(err, e);
}

// Leaf boxes from (moved) triangles: the primitive of a one-primitive leaf is primitive_indices[first_index].
__global__ void __launch_bounds__(T256) k_leaf_refit(Node32* __restrict__ nodes, uint32_t n_nodes,
                                                     const uint32_t* __restrict__ prims, uint32_t n_prims,
                                                     const float4* __restrict__ tris, uint32_t* __restrict__ err) {
    const uint32_t x = blockIdx.x * T256 + threadIdx.x;
    if (x >= n_nodes) return;
    const Node32 nd = load_node(&nodes[x]);
    if (nd.prim_count == 0u) return;
    if (nd.prim_count != 1u || nd.first_index >= n_prims) {
        atomicOr(err, DERR_CW_NONUNIT_LEAF);
        return;
    }
    const uint32_t p = prims[nd.first_index];
    if (p >= n_prims) {
        atomicOr(err, DERR_CAPACITY);
        return;
    }
    const float4 v0 = tris[3 * (size_t)p], v1 = tris[3 * (size_t)p + 1], v2 = tris[3 * (size_t)p + 2];
    Box b{v0.x, v0.y, v0.z, v0.x, v0.y, v0.z};
    b = Box{smin(b.lx, v1.x), smin(b.ly, v1.y), smin(b.lz, v1.z), smax(b.hx, v1.x), smax(b.hy, v1.y), smax(b.hz, v1.z)};
    b = Box{smin(b.lx, v2.x), smin(b.ly, v2.y), smin(b.lz, v2.z), smax(b.hx, v2.x), smax(b.hy, v2.y), smax(b.hz, v2.z)};
    store_node(&nodes[x], make_node(b, 1u, nd.first_index));
}

// ---------------------------------------------------------------------------------------------------------------
// reinsertion
// ---------------------------------------------------------------------------------------------------------------

// bvh2/mod.rs:586-609.  With `negzero` set, also raises *negzero when a leaf box holds a -0 coordinate.
__global__ void __launch_bounds__(T256) k_parents(const Node32* __restrict__ nodes, uint32_t n,
                                                  uint32_t* __restrict__ parents, uint32_t* __restrict__ negzero) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i == 0) parents[0] = 0u;
    if (i >= n) return;
    const Node32 nd = load_node(&nodes[i]);
    if (nd.prim_count == 0u && nd.first_index + 1u < n) {
        parents[nd.first_index] = i;
        parents[nd.first_index + 1] = i;
    } else if (negzero && nd.prim_count != 0u) {
        const uint32_t nz = 0x80000000u;
        if (__float_as_uint(nd.minx) == nz || __float_as_uint(nd.miny) == nz || __float_as_uint(nd.minz) == nz ||
            __float_as_uint(nd.maxx) == nz || __float_as_uint(nd.maxy) == nz || __float_as_uint(nd.maxz) == nz)
            atomicOr(negzero, 1u);
    }
}

// reinsertion.rs:128-147 + 210-218: key of node i (1 <= i < take) is the crate's radix key of -half_area.
__global__ void __launch_bounds__(T256) k_cand_keys(const Node32* __restrict__ nodes, uint32_t take,
                                                    uint64_t* __restrict__ keys) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x + 1u;
    if (i >= take) return;
    keys[i - 1] = f32_order_key(-half_area(node_box(load_node(&nodes[i]))));
}

struct ReinsMisc {
    uint32_t err;
    uint32_t positive;               // reinsertions with area_diff > 0
    uint32_t negzero;                // a leaf box holds a -0 coordinate (k_parents)
    uint32_t next;                   // candidates handed out beyond the grid's first span
    uint32_t undecided[MIS_ROUNDS_PER_SYNC];
};

// reinsertion.rs:247-340, one lane per candidate walk, stack column `tid` of an [sd][nthreads] array; a walk's
// steps are the crate's in the crate's order, one per loop iteration (start, search step, or climb to the next
// pivot), so lanes of a wave refill with new candidates independently.
__global__ void __launch_bounds__(T256) k_find_reinsertion(const Node32* __restrict__ nodes, uint32_t n_nodes,
                                                           const uint32_t* __restrict__ parents,
                                                           const uint32_t* __restrict__ cand, uint32_t count,
                                                           uint2* __restrict__ stack, uint32_t sd,
                                                           uint32_t* __restrict__ r_from, uint32_t* __restrict__ r_to,
                                                           uint64_t* __restrict__ keys, ReinsMisc* misc) {
    // The DFS stack of each lane: its top FIND_LDS entries in an LDS ring, older entries in the lane's HBM column
    // (entry k at stack[k * nthreads + tid]).  Entries [lo, sp) are in the ring, [0, lo) in HBM: a push into a full
    // ring first writes the ring's oldest entry out (off the walk's critical path), a pop of an empty ring reads
    // HBM.  The walk's pops then cost an LDS access instead of an L2 round trip.
    __shared__ uint2 ring[FIND_LDS][T256];
    const uint32_t tid = blockIdx.x * T256 + threadIdx.x;
    const uint32_t nthreads = gridDim.x * T256;
    uint32_t sp = 0, lo = 0;
    auto push = [&](uint2 v) {
        if (sp - lo == FIND_LDS) {
            stack[(size_t)lo * nthreads + tid] = ring[lo % FIND_LDS][threadIdx.x];
            ++lo;
        }
        ring[sp % FIND_LDS][threadIdx.x] = v;
        ++sp;
    };
    auto pop = [&]() -> uint2 {
        --sp;
        if (sp < lo) {
            lo = sp;
            return stack[(size_t)sp * nthreads + tid];
        }
        return ring[sp % FIND_LDS][threadIdx.x];
    };
    // Each lane walks one candidate at a time: the grid first takes candidates [0, nthreads) (lane tid the
    // candidate tid); a lane that finishes waits until FIND_REFILL lanes of its wave are idle (or none is walking),
    // and the wave then takes that many further candidates from the counter with one atomic.  Lanes walk
    // independently, so a long walk no longer holds the other 63 lanes idle until it ends.
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t step_budget = 2u * n_nodes + 64u;  // walks of a valid tree end long before these budgets
    uint32_t t = tid;
    bool done = t >= count, live = !done, fresh = live;
    uint32_t node_id = 0u, parent_id = 0u, sib = 0u, pivot_id = 0u, best_to = 0u, climb = 0u, steps = 0u;
    float node_area = 0.0f, area_diff = 0.0f, best_area = 0.0f;
    Box aabb{}, pivot_bbox{};
    bool overflow = false, held = false;
    uint2 top = make_uint2(0u, 0u);
    for (;;) {
        const uint64_t want = __ballot(!live && !done);
        const uint64_t walking = __ballot(live);
        if (want != 0ull && ((uint32_t)__popcll(want) >= FIND_REFILL || walking == 0ull)) {
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)want) - 1u;
            uint32_t base = 0u;
            if (lane == leader) base = nthreads + atomicAdd(&misc->next, (uint32_t)__popcll(want));
            base = (uint32_t)__shfl((int)base, (int)leader);
            if (!live && !done) {
                t = base + (uint32_t)__popcll(want & lanemask_lt());
                if (t < count) {
                    live = true;
                    fresh = true;
                } else {
                    done = true;
                }
            }
        }
        if (__ballot(live) == 0ull) break;  // every lane of the wave is done (wave-uniform)
        bool fin = false;
        if (live) {
            if (fresh) {  // the candidate's start (reinsertion.rs:250-262)
                fresh = false;
                node_id = cand[t] + 1u;  // candidates were numbered from node 1
                best_to = 0u;
                best_area = 0.0f;
                aabb = node_box(load_node(&nodes[node_id]));
                node_area = half_area(aabb);
                parent_id = parents[node_id];
                area_diff = half_area(node_box(load_node(&nodes[parent_id])));
                sib = sibling_of(node_id);
                pivot_bbox = node_box(load_node(&nodes[sib]));
                pivot_id = parent_id;
                overflow = false;
                climb = 0u;
                steps = 0u;
                sp = 0;
                lo = 0;
                // The second child of an expanded node is the next entry popped, so it is kept in registers
                // (`held`) instead of pushed and popped: the same visit order without an LDS round trip.
                top = make_uint2(__float_as_uint(area_diff), sib);
                held = true;
            }
            if (!held && sp == 0) {
                // this level's search is over: climb to the next pivot (or end the walk)
                fin = overflow;
                if (!fin) {
                    if (pivot_id != parent_id) {
                        pivot_bbox = box_union(pivot_bbox, node_box(load_node(&nodes[sib])));
                        area_diff += half_area(node_box(load_node(&nodes[pivot_id]))) - half_area(pivot_bbox);
                    }
                    if (pivot_id == 0u) {
                        fin = true;
                    } else if (++climb > sd) {
                        overflow = true;
                        fin = true;
                    } else {
                        sib = sibling_of(pivot_id);
                        pivot_id = parents[pivot_id];
                        sp = 0;
                        lo = 0;
                        top = make_uint2(__float_as_uint(area_diff), sib);
                        held = true;
                    }
                }
            } else if (++steps > step_budget) {
                overflow = true;  // ends the walk at the next iteration
                held = false;
                sp = 0;
            } else {
                if (!held) top = pop();
                held = false;
                const float top_area = __uint_as_float(top.x);
                if (!(top_area - node_area <= best_area)) {
                    const Node32 dst = load_node(&nodes[top.y]);
                    const Box db = node_box(dst);
                    const float merged_area = half_area(box_union(db, aabb));
                    const float reinsert_area = top_area - merged_area;
                    if (reinsert_area > best_area) {
                        best_to = top.y;
                        best_area = reinsert_area;
                    }
                    const float child_area = reinsert_area + half_area(db);
                    // children whose bound already fails would fail it when popped (best_area only grows)
                    if (dst.prim_count == 0u && !(child_area - node_area <= best_area)) {
                        if (sp + 2 > sd) {
                            overflow = true;
                            held = false;
                            sp = 0;
                        } else {
                            push(make_uint2(__float_as_uint(child_area), dst.first_index));
                            top = make_uint2(__float_as_uint(child_area), dst.first_index + 1u);
                            held = true;
                        }
                    }
                }
            }
        }
        bool pos = false;
        if (fin) {
            if (overflow) atomicOr(&misc->err, DERR_CAPACITY);
            uint32_t from = node_id;
            if (best_to == sibling_of(node_id) || best_to == parent_id) {
                from = 0u;
                best_to = 0u;
                best_area = 0.0f;
            }
            r_from[t] = from;
            r_to[t] = best_to;
            pos = best_area > 0.0f;
            keys[t] = pos ? (uint64_t)(~f32_order_key(best_area)) : 0xffffffffull;
            live = false;
        }
        const uint64_t m = __ballot(pos);
        if (m != 0ull && lane == (uint32_t)__ffsll((unsigned long long)m) - 1u) atomicAdd(&misc->positive, (uint32_t)__popcll(m));
    }
}

constexpr uint32_t ST_UNDECIDED = 0u, ST_ACCEPTED_NEW = 1u, ST_ACCEPTED = 2u, ST_REJECTED = 3u;

// Conflict nodes of the k-th reinsertion in priority order (reinsertion.rs:206-216), with the parents of the step.
__global__ void __launch_bounds__(T256) k_mis_init(const uint32_t* __restrict__ order, const ReinsMisc* misc,
                                                   const uint32_t* __restrict__ r_from,
                                                   const uint32_t* __restrict__ r_to,
                                                   const uint32_t* __restrict__ parents, uint32_t* __restrict__ conf,
                                                   uint8_t* __restrict__ state) {
    const uint32_t k = blockIdx.x * T256 + threadIdx.x;
    if (k >= misc->positive) return;
    const uint32_t t = order[k];
    const uint32_t f = r_from[t], to = r_to[t];
    conf[5 * (size_t)k + 0] = to;
    conf[5 * (size_t)k + 1] = f;
    conf[5 * (size_t)k + 2] = sibling_of(f);
    conf[5 * (size_t)k + 3] = parents[to];
    conf[5 * (size_t)k + 4] = parents[f];
    state[k] = (uint8_t)ST_UNDECIDED;
}

__global__ void __launch_bounds__(T256) k_mis_claim(const uint32_t* __restrict__ conf,
                                                    const uint8_t* __restrict__ state, ReinsMisc* misc, uint32_t round,
                                                    uint32_t* __restrict__ owner) {
    const uint32_t k = blockIdx.x * T256 + threadIdx.x;
    if (k == 0) misc->undecided[round] = 0u;
    if (k >= misc->positive || state[k] != ST_UNDECIDED) return;
    for (int j = 0; j < 5; ++j) atomicMin(&owner[conf[5 * (size_t)k + j]], k);
}

__global__ void __launch_bounds__(T256) k_mis_decide(const uint32_t* __restrict__ conf, uint8_t* __restrict__ state,
                                                     const ReinsMisc* misc, const uint32_t* __restrict__ owner,
                                                     uint8_t* __restrict__ touched) {
    const uint32_t k = blockIdx.x * T256 + threadIdx.x;
    if (k >= misc->positive || state[k] != ST_UNDECIDED) return;
    bool mine = true;
    for (int j = 0; j < 5; ++j) mine &= owner[conf[5 * (size_t)k + j]] == k;
    if (!mine) return;
    state[k] = (uint8_t)ST_ACCEPTED_NEW;
    for (int j = 0; j < 5; ++j) touched[conf[5 * (size_t)k + j]] = 1u;
}

__global__ void __launch_bounds__(T256) k_mis_settle(const uint32_t* __restrict__ conf, uint8_t* __restrict__ state,
                                                     ReinsMisc* misc, uint32_t round, uint32_t* __restrict__ owner,
                                                     const uint8_t* __restrict__ touched) {
    const uint32_t k = blockIdx.x * T256 + threadIdx.x;
    bool still = false;
    if (k < misc->positive) {
        const uint32_t s = state[k];
        if (s == ST_UNDECIDED || s == ST_ACCEPTED_NEW) {
            bool hit = false;
            for (int j = 0; j < 5; ++j) {
                const uint32_t c = conf[5 * (size_t)k + j];
                owner[c] = 0xffffffffu;
                hit |= touched[c] != 0u;
            }
            if (s == ST_ACCEPTED_NEW) state[k] = (uint8_t)ST_ACCEPTED;
            else if (hit) state[k] = (uint8_t)ST_REJECTED;
            else still = true;
        }
    }
    const uint64_t m = __ballot(still);
    if ((threadIdx.x & 63u) == 0u && m) atomicAdd(&misc->undecided[round], (uint32_t)__popcll(m));
}

// reinsertion.rs:342-382 without the refits (one full refit follows).  Accepted moves touch disjoint node sets.
__global__ void __launch_bounds__(T256) k_reins_apply(const uint32_t* __restrict__ conf,
                                                      const uint8_t* __restrict__ state, const ReinsMisc* misc,
                                                      Node32* __restrict__ nodes) {
    const uint32_t k = blockIdx.x * T256 + threadIdx.x;
    if (k >= misc->positive || state[k] != ST_ACCEPTED) return;
    const uint32_t to = conf[5 * (size_t)k + 0], from = conf[5 * (size_t)k + 1];
    const uint32_t sib = conf[5 * (size_t)k + 2], parent_id = conf[5 * (size_t)k + 4];
    const Node32 sibling_node = load_node(&nodes[sib]);
    const Node32 dst_node = load_node(&nodes[to]);
    store_node(&nodes[to], make_node(node_box(dst_node), 0u, left_sibling_of(from)));
    store_node(&nodes[sib], dst_node);
    store_node(&nodes[parent_id], sibling_node);
}

__device__ __forceinline__ bool box_value_eq(const Box& a, const Box& b) {
    return a.lx == b.lx && a.ly == b.ly && a.lz == b.lz && a.hx == b.hx && a.hy == b.hy && a.hz == b.hz;
}

// refit_from_fast (bvh2/mod.rs:722-751, release build): every internal box on the way up from `index` is set to the
// union of its children, except that the walk returns, without writing, at the second box whose recomputed value
// equals the stored one (by value: -0 == +0).
__device__ void refit_from_fast_walk(Node32* __restrict__ nodes, const uint32_t* __restrict__ parents,
                                     uint32_t index) {
    uint32_t same = 0;
    for (;;) {
        const Node32 nd = load_node(&nodes[index]);
        if (nd.prim_count == 0u) {
            const Box u = box_union(node_box(load_node(&nodes[nd.first_index])),
                                    node_box(load_node(&nodes[nd.first_index + 1])));
            if (box_value_eq(node_box(nd), u) && ++same == 2u) return;
            store_node(&nodes[index], make_node(u, 0u, nd.first_index));
        }
        if (index == 0u) break;
        index = parents[index];
    }
}

// The crate's apply loop for one step (reinsertion.rs:175-195, 342-382), one accepted move after another in
// priority order, each followed by its two refit_from_fast walks, with the parents kept as reinsert_node keeps them.
// Where a walk stops early, boxes above keep the bits they had, which differ from a full refit's only in the sign
// of a zero coordinate; the host runs this instead of k_reins_apply + a full refit when a leaf holds a -0.  One lane.
__global__ void __launch_bounds__(64) k_reins_replay(const uint32_t* __restrict__ conf,
                                                     const uint8_t* __restrict__ state, const ReinsMisc* misc,
                                                     Node32* __restrict__ nodes, uint32_t* __restrict__ parents) {
    if (threadIdx.x != 0u) return;
    const uint32_t npos = misc->positive;
    for (uint32_t k = 0; k < npos; ++k) {
        if (state[k] != ST_ACCEPTED) continue;
        const uint32_t to = conf[5 * (size_t)k + 0], from = conf[5 * (size_t)k + 1];
        const uint32_t sib = conf[5 * (size_t)k + 2], parent_id = conf[5 * (size_t)k + 4];
        const Node32 sibling_node = load_node(&nodes[sib]);
        const Node32 dst_node = load_node(&nodes[to]);
        store_node(&nodes[to], make_node(node_box(dst_node), 0u, left_sibling_of(from)));
        store_node(&nodes[sib], dst_node);
        store_node(&nodes[parent_id], sibling_node);
        if (dst_node.prim_count == 0u) {
            parents[dst_node.first_index] = sib;
            parents[dst_node.first_index + 1] = sib;
        }
        if (sibling_node.prim_count == 0u) {
            parents[sibling_node.first_index] = parent_id;
            parents[sibling_node.first_index + 1] = parent_id;
        }
        parents[sib] = to;
        parents[from] = to;
        refit_from_fast_walk(nodes, parents, to);
        refit_from_fast_walk(nodes, parents, parent_id);
    }
}

__global__ void __launch_bounds__(T256) k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * T256 + threadIdx.x; i < n; i += (size_t)gridDim.x * T256) p[i] = v;
}

// ---------------------------------------------------------------------------------------------------------------
// pre-splits
// ---------------------------------------------------------------------------------------------------------------

struct SplitMisc {
    uint32_t largest_key;  // order key of the largest half area (atomicMax)
    float avg;
    float thresh_low;
    float thresh_high;
    uint32_t splits;       // splitters of the current iteration
    uint32_t kept;         // candidates kept by the filter
    uint32_t err;
    uint32_t pad;
    float sum;         // the running half-area sum at element sum_pos
    uint32_t sum_pos;
};

__device__ __forceinline__ float vget(const float4& v, uint32_t d) { return d == 0 ? v.x : (d == 1 ? v.y : v.z); }
__device__ __forceinline__ float bget_lo(const Box& b, uint32_t d) { return d == 0 ? b.lx : (d == 1 ? b.ly : b.lz); }
__device__ __forceinline__ float bget_hi(const Box& b, uint32_t d) { return d == 0 ? b.hx : (d == 1 ? b.hy : b.hz); }
__device__ __forceinline__ void bset_lo(Box& b, uint32_t d, float f) {
    if (d == 0) b.lx = f;
    else if (d == 1) b.ly = f;
    else b.lz = f;
}
__device__ __forceinline__ void bset_hi(Box& b, uint32_t d, float f) {
    if (d == 0) b.hx = f;
    else if (d == 1) b.hy = f;
    else b.hz = f;
}
__device__ __forceinline__ Box box_invalid() {
    const float M = 3.40282347e+38f;
    return Box{M, M, M, -M, -M, -M};
}
__device__ __forceinline__ void box_extend(Box& b, float x, float y, float z) {
    b = Box{smin(b.lx, x), smin(b.ly, y), smin(b.lz, z), smax(b.hx, x), smax(b.hy, y), smax(b.hz, z)};
}
// aabb.rs:98-103 (self = the clipped triangle box, other = the halved leaf box)
__device__ __forceinline__ Box box_intersection(const Box& a, const Box& b) {
    return Box{smax(a.lx, b.lx), smax(a.ly, b.ly), smax(a.lz, b.lz),
               smin(a.hx, b.hx), smin(a.hy, b.hy), smin(a.hz, b.hz)};
}
// aabb.rs:125-134
__device__ __forceinline__ uint32_t largest_axis(const Box& a) {
    const float dx = a.hx - a.lx, dy = a.hy - a.ly, dz = a.hz - a.lz;
    if (dx < dy) return dy < dz ? 2u : 1u;
    return dx < dz ? 2u : 0u;
}

// splits.rs:137-158; the crossing point is a fused multiply-add per lane (Vec3A::mul_add).
__device__ __forceinline__ void split_triangle(uint32_t dim, float pos, const float4 v[4], Box& left, Box& right) {
    left = box_invalid();
    right = box_invalid();
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float4 v0 = v[i], v1 = v[i + 1];
        const float v0d = vget(v0, dim), v1d = vget(v1, dim);
        if (v0d <= pos) box_extend(left, v0.x, v0.y, v0.z);
        if (v0d >= pos) box_extend(right, v0.x, v0.y, v0.z);
        if ((v0d < pos && pos < v1d) || (v1d < pos && pos < v0d)) {
            const float inv_length = 1.0f / (v1d - v0d);
            const float t = (pos - v0d) * inv_length;
            const float cx = __builtin_fmaf(t, v1.x - v0.x, v0.x);
            const float cy = __builtin_fmaf(t, v1.y - v0.y, v0.y);
            const float cz = __builtin_fmaf(t, v1.z - v0.z, v0.z);
            box_extend(left, cx, cy, cz);
            box_extend(right, cx, cy, cz);
        }
    }
}

__device__ __forceinline__ Box load_box(const float4* __restrict__ a, uint32_t i) {
    const float4 lo = a[2 * (size_t)i], hi = a[2 * (size_t)i + 1];
    return Box{lo.x, lo.y, lo.z, hi.x, hi.y, hi.z};
}
__device__ __forceinline__ void store_box(float4* __restrict__ a, uint32_t i, const Box& b) {
    a[2 * (size_t)i] = make_float4(b.lx, b.ly, b.lz, 0.0f);
    a[2 * (size_t)i + 1] = make_float4(b.hx, b.hy, b.hz, 0.0f);
}

// Leaf boxes (triangle.rs:28-30), indices 0..n, half areas, largest half area.  Grid-stride over at most
// PREPARE_MAX_BLOCKS workgroups, so the largest half area takes one global atomic per workgroup (one per wave made
// the launch wait on about 270K atomics to one address at 17M triangles).
constexpr uint32_t PREPARE_MAX_BLOCKS = 2048;
__global__ void __launch_bounds__(T256) k_split_prepare(const float4* __restrict__ tris, uint32_t n,
                                                        float4* __restrict__ aabbs, uint32_t* __restrict__ indices,
                                                        float* __restrict__ ha, SplitMisc* misc) {
    __shared__ uint32_t wmax[T256 / 64];
    uint32_t k = 0u;
    for (uint32_t i = blockIdx.x * T256 + threadIdx.x; i < n; i += gridDim.x * T256) {
        const float4 v0 = tris[3 * (size_t)i], v1 = tris[3 * (size_t)i + 1], v2 = tris[3 * (size_t)i + 2];
        Box b{v0.x, v0.y, v0.z, v0.x, v0.y, v0.z};
        box_extend(b, v1.x, v1.y, v1.z);
        box_extend(b, v2.x, v2.y, v2.z);
        store_box(aabbs, i, b);
        indices[i] = i;
        const float h = half_area(b);
        ha[i] = h;
        // f32::max keeps the larger value; half areas of triangle boxes are never negative, so the order key's max
        // is it
        k = max(k, f32_order_key(h));
    }
    k = wave_max_u32(k);
    if (__lane_id() == 0) wmax[threadIdx.x >> 6] = k;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < T256 / 64; ++w) k = max(k, wmax[w]);
        atomicMax(&misc->largest_key, k);
    }
}

// The crate's running sum `avg_area += half_area` in triangle order (cwbvh/builder.rs:34-41), reproduced exactly
// without a sequential walk.  While the running sum s stays in one binade (ulp u), the f32 step fl(s + x) equals
// s + u * round(x / u) for every x >= 0 -- unless x / u has fraction exactly 1/2 (a tie, decided by the parity of
// s / u) or s + x leaves the binade.  So the increments round(x / u) are prefix-summed as integers; the first tie,
// binade crossing or non-finite value is then added with one IEEE f32 add (the sequential step itself, from the
// exact s before it) and the scan resumes after it.  Stops crowd the start of the sum (every binade crossing while s
// is small) and are rare after it, so one workgroup scans the first SUM_SERIAL_SPAN elements tile by tile
// (k_sum_serial), and the rest goes in passes over all CUs (k_sum_blocks: per-block increment totals; k_sum_pass:
// one workgroup that jumps over the blocks without a stop, and steps through ties in place, since a tie keeps the
// binade and so the block totals).  A pass ends at a binade crossing, and the host starts the next one.  The split
// thresholds of split_aabbs_preset follow.
constexpr int SUM_THREADS = 1024;
constexpr int SUM_ITEMS = 16;  // serial tile: 16K elements
constexpr uint32_t SUM_TILE = SUM_THREADS * SUM_ITEMS;
constexpr uint32_t SUM_CAP = 1u << 24;
constexpr uint32_t SUM_NONE = 0xffffffffu;
constexpr uint32_t SUM_SERIAL_SPAN = 1u << 18;
constexpr int PASS_ITEMS = 16;
constexpr uint32_t PASS_BLOCK = T256 * PASS_ITEMS;  // elements per k_sum_blocks workgroup
constexpr int PASS_TILE_ITEMS = PASS_BLOCK / SUM_THREADS;

__device__ __forceinline__ unsigned long long wave_inclusive_scan_u64(unsigned long long v) {
    const int lane = (int)__lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned long long t = __shfl_up(v, off);
        if (lane >= off) v += t;
    }
    return v;
}

// round(x / u) for the ulp u = 2^(be - 150) of the running sum's binade and a finite x >= 0 given by its bits, in
// integer arithmetic; `tie` is set when x / u has fraction exactly 1/2.  Results of 2^24 and more are clamped to
// 2^24: one such step alone carries the sum out of the binade, which stops the scan at it.
__device__ __forceinline__ uint32_t ulp_steps(uint32_t xb, uint32_t be, bool& tie) {
    const uint32_t ex = xb >> 23;
    const uint32_t mx = (xb & 0x7fffffu) | (ex != 0u ? 0x800000u : 0u);
    const int d = (int)be - (int)(ex != 0u ? ex : 1u);  // x / u = mx / 2^d
    tie = false;
    if (d < 0) return SUM_CAP;  // x is normal here (mx >= 2^23), so x / u >= 2^24
    if (d == 0) return mx;
    if (d > 24) return 0u;  // x / u < 2^24 / 2^25
    const uint32_t half = 1u << (d - 1);
    const uint32_t rem = mx & ((half << 1) - 1u);
    tie = rem == half;
    return (mx >> d) + (rem > half ? 1u : 0u);
}

// The running sum as the integer scan sees it: s = (m0 + k) * u after k more ulps.
struct SumMode {
    float s;
    bool zero;  // s == 0: the first non-zero element starts the sum
    bool scan;  // finite s > 0: the integer scan applies; otherwise every element that changes s is a stop
    uint32_t be, m0;
};
__device__ __forceinline__ SumMode sum_mode(float s) {
    const uint32_t bits = __float_as_uint(s);
    const uint32_t be_raw = (bits >> 23) & 0xffu;
    SumMode m;
    m.s = s;
    m.zero = s == 0.0f;
    m.scan = isfinite(s) && s > 0.0f;
    m.be = be_raw == 0u ? 1u : be_raw;  // subnormals share the first normal binade's ulp
    m.m0 = m.zero ? 0u : (bits & 0x7fffffu) | (be_raw ? 0x800000u : 0u);
    return m;
}
// x's increment in ulps (0 for a stop) and whether x stops the integer scan.
__device__ __forceinline__ bool sum_rule(const SumMode& md, float x, uint32_t& q) {
    const uint32_t xb = __float_as_uint(x);
    bool tie;
    const uint32_t steps = ulp_steps(xb, md.be, tie);
    bool st;
    if (md.zero)
        st = (xb & 0x7fffffffu) != 0u;  // 0 + x = x: the first non-zero value starts the sum
    else if (!md.scan)
        st = !(md.s + x == md.s);  // e.g. an infinite sum stays itself until it turns NaN
    else
        st = xb >= 0x7f800000u || tie;  // a non-finite or negative x, or a tie: one IEEE add
    q = !md.scan || st ? 0u : steps;
    return st;
}
__device__ __forceinline__ bool sum_crosses(const SumMode& md, unsigned long long k) {
    return md.scan && md.m0 + k >= SUM_CAP;
}
__device__ __forceinline__ float sum_value(const SumMode& md, unsigned long long k) {  // k below SUM_CAP - m0
    if (!md.scan) return md.s;
    return ldexpf((float)(md.m0 + k), (int)md.be - 150);
}

struct SumShared {
    unsigned long long wtot[SUM_THREADS / 64];
    unsigned long long total;
    uint32_t first;
    float s_before, s_after;
};
struct SumStep {
    uint32_t idx;              // first stop in the tile, SUM_NONE for none
    unsigned long long total;  // no stop: base + the tile's increments
    float s_before, s_after;   // a stop: the exact sum before it and the IEEE step that adds it
};

// One block-wide step over `len` elements (thread tid holds elements tid * ITEMS + r; get(e) reads element e) from
// `base` ulps above md.m0: the first stop and the sums around it, or the total.  Ends with a barrier, after which
// the shared state may be reused.
template <int ITEMS, class Get>
__device__ __forceinline__ SumStep sum_step(const SumMode& md, unsigned long long base, uint32_t len, Get get,
                                            SumShared& sh) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint32_t q[ITEMS];
    uint32_t stops = 0u;  // bit r: element r stops the integer scan
    uint32_t tsum = 0u;   // at most ITEMS * 2^24
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const uint32_t e = tid * ITEMS + (uint32_t)r;
        q[r] = 0u;
        if (e < len && sum_rule(md, get(e), q[r])) stops |= 1u << r;
        tsum += q[r];
    }
    // block-wide exclusive prefix of the thread totals (element order)
    const unsigned long long incl = wave_inclusive_scan_u64(tsum);
    if (tid == 0) sh.first = SUM_NONE;
    if (lane == 63u) sh.wtot[wave] = incl;
    __syncthreads();
    unsigned long long excl = base + incl - tsum;
    for (uint32_t w = 0; w < wave; ++w) excl += sh.wtot[w];
    unsigned long long run = excl;
    uint32_t my_stop = SUM_NONE;
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        run += q[r];
        if ((((stops >> r) & 1u) || sum_crosses(md, run)) && my_stop == SUM_NONE) my_stop = tid * ITEMS + (uint32_t)r;
    }
    if (my_stop != SUM_NONE) atomicMin(&sh.first, my_stop);
    __syncthreads();
    const uint32_t idx = sh.first;
    if (idx == SUM_NONE) {
        if (tid == SUM_THREADS - 1) sh.total = run;  // below SUM_CAP - m0: no crossing
    } else if (idx / ITEMS == tid) {
        // every increment before the first stop is exact (a clamped one would cross, so stop, itself)
        const int r0 = (int)(idx % ITEMS);
        unsigned long long before = excl;
#pragma unroll
        for (int r = 0; r < ITEMS; ++r)
            if (r < r0) before += q[r];
        const float sb = sum_value(md, before);
        sh.s_before = sb;
        sh.s_after = sb + get(idx);  // the sequential step
    }
    __syncthreads();
    SumStep out;
    out.idx = idx;
    out.total = sh.total;
    out.s_before = sh.s_before;
    out.s_after = sh.s_after;
    __syncthreads();
    return out;
}

__device__ __forceinline__ void split_finish(SplitMisc* misc, float s, uint32_t n) {
    const float avg = s / (float)n;
    const float largest = f32_from_order_key(misc->largest_key);
    misc->avg = avg;
    misc->thresh_low = avg * 3.0f;                                       // splits.rs:27
    misc->thresh_high = fmaxf(avg * 4.0f, avg * 0.9f + largest * 0.1f);  // splits.rs:28
}

// From the state (misc->sum, misc->sum_pos), 16K-element tiles until the position reaches `limit` (or the sum turns
// NaN).  The tiles go through LDS, one pad word after every SUM_ITEMS elements: filled by lane-contiguous loads and
// read back as each thread's SUM_ITEMS consecutive elements (thread stride SUM_ITEMS + 1 words); the next tile is
// loaded while the current one is scanned.
__global__ void __launch_bounds__(SUM_THREADS) k_sum_serial(const float* __restrict__ ha, uint32_t n, uint32_t limit,
                                                           SplitMisc* misc) {
    __shared__ float tile[SUM_TILE + SUM_TILE / SUM_ITEMS];
    __shared__ SumShared sh;
    const uint32_t tid = threadIdx.x;
    float nxt[SUM_ITEMS];
    auto load_tile = [&](uint32_t base) {
#pragma unroll
        for (int r = 0; r < SUM_ITEMS; ++r) {
            const uint64_t k = (uint64_t)base + (uint32_t)r * SUM_THREADS + tid;
            nxt[r] = k < n ? ha[k] : 0.0f;
        }
    };
    auto store_tile = [&]() {
#pragma unroll
        for (int r = 0; r < SUM_ITEMS; ++r) {
            const uint32_t e = (uint32_t)r * SUM_THREADS + tid;
            tile[e + e / SUM_ITEMS] = nxt[r];
        }
    };
    float s = misc->sum;  // the running sum, identical in every thread
    uint32_t pos = misc->sum_pos;
    if (pos < n) {
        load_tile(pos);
        store_tile();
    }
    __syncthreads();
    while (pos < limit && !isnan(s)) {
        const uint32_t tile_len = min(SUM_TILE, n - pos);
        if (tile_len < n - pos) load_tile(pos + SUM_TILE);  // the next tile when this one has no stop
        const SumMode md = sum_mode(s);
        const SumStep o =
            sum_step<SUM_ITEMS>(md, 0ull, tile_len, [&](uint32_t e) { return tile[e + e / SUM_ITEMS]; }, sh);
        const uint32_t prev = pos;
        if (o.idx == SUM_NONE) {
            s = sum_value(md, o.total);
            pos += tile_len;
        } else {
            s = o.s_after;
            pos += o.idx + 1u;
        }
        if (pos < n) {
            if (pos != prev + SUM_TILE) load_tile(pos);  // a stop moved the window: load it afresh
            store_tile();
        }
        __syncthreads();
    }
    if (tid == 0) {
        misc->sum = s;
        misc->sum_pos = pos;
        if (pos >= n || isnan(s)) split_finish(misc, s, n);
    }
}

// Per-block increments and whether the block holds a stop, for elements [pos, n) from the running sum s.
__global__ void __launch_bounds__(T256) k_sum_blocks(const float* __restrict__ ha, uint32_t n, uint32_t pos, float s,
                                                     unsigned long long* __restrict__ bsum,
                                                     uint32_t* __restrict__ bstop) {
    __shared__ unsigned long long ws[T256 / 64];
    __shared__ uint32_t wf[T256 / 64];
    const uint32_t tid = threadIdx.x;
    const SumMode md = sum_mode(s);
    const uint32_t base = pos + blockIdx.x * PASS_BLOCK;
    uint32_t tsum = 0u, first = SUM_NONE;
#pragma unroll
    for (int r = 0; r < PASS_ITEMS; ++r) {
        const uint32_t e = (uint32_t)r * T256 + tid;  // lane-contiguous
        if ((uint64_t)base + e < n) {
            uint32_t q;
            if (sum_rule(md, ha[base + e], q)) first = min(first, e);
            tsum += q;
        }
    }
    unsigned long long v = tsum;
    for (int off = 32; off >= 1; off >>= 1) {
        v += __shfl_xor(v, off);
        first = min(first, (uint32_t)__shfl_xor((int)first, off));
    }
    if ((tid & 63u) == 0u) {
        ws[tid >> 6] = v;
        wf[tid >> 6] = first;
    }
    __syncthreads();
    if (tid == 0) {
        unsigned long long t = 0ull;
        uint32_t f = SUM_NONE;
        for (int w = 0; w < T256 / 64; ++w) {
            t += ws[w];
            f = min(f, wf[w]);
        }
        bsum[blockIdx.x] = t;
        bstop[blockIdx.x] = f;
    }
}

// One pass from (pos, s) over the k_sum_blocks results of its nb blocks: block totals carry the sum over blocks
// without a stop; the block with the next stop is scanned element by element; a tie keeps the binade, so the pass
// goes on after it; any other stop ends the pass with the new state.
__global__ void __launch_bounds__(SUM_THREADS) k_sum_pass(const float* __restrict__ ha, uint32_t n, uint32_t pos,
                                                         float s, uint32_t nb,
                                                         const unsigned long long* __restrict__ bsum,
                                                         const uint32_t* __restrict__ bstop, SplitMisc* misc) {
    __shared__ SumShared sh;
    __shared__ uint32_t s_block;
    __shared__ unsigned long long s_pre;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const SumMode md = sum_mode(s);
    unsigned long long acc = 0ull;  // ulps above md.m0 at element `at`
    uint32_t b = 0, at = pos;       // block b holds element `at`
    float s_end = s;
    uint32_t pos_end = n;
    bool scan_block = true;  // block b needs an element scan from `at`
    for (;;) {
        if (scan_block) {
            const uint32_t bend = (uint32_t)min((uint64_t)n, (uint64_t)pos + (uint64_t)(b + 1u) * PASS_BLOCK);
            const uint32_t start = at;
            const SumStep o = sum_step<PASS_TILE_ITEMS>(md, acc, bend - start,
                                                        [&](uint32_t e) { return ha[start + e]; }, sh);
            if (o.idx == SUM_NONE) {
                acc = o.total;
                at = bend;
                ++b;
            } else {
                const float sa = o.s_after;
                at = start + o.idx + 1u;
                // a tie inside the binade: the block totals still hold; anything else ends the pass
                const uint32_t eb = __float_as_uint(s) >> 23;  // sign and exponent of the pass's sum
                const bool stay = md.scan && !(sa < o.s_before) && (__float_as_uint(sa) >> 23) == eb;
                if (!stay) {
                    s_end = sa;
                    pos_end = at;
                    break;
                }
                acc = ((__float_as_uint(sa) & 0x7fffffu) | (eb != 0u ? 0x800000u : 0u)) - md.m0;  // same binade
                if (at == bend) ++b;
            }
            if (at >= n) {
                s_end = sum_value(md, acc);
                pos_end = n;
                break;
            }
            scan_block = at != pos + b * PASS_BLOCK;  // inside a block after a tie: scan the rest of it
            if (scan_block) continue;
        }
        // block b starts at `at`: find the first block from b with a flagged stop or a crossing
        const uint32_t rem = nb - b;
        const uint32_t c = (rem + SUM_THREADS - 1u) / SUM_THREADS;
        const uint32_t k0 = b + min(rem, tid * c), k1 = b + min(rem, (tid + 1u) * c);
        unsigned long long tsum = 0ull;
        for (uint32_t k = k0; k < k1; ++k) tsum += bsum[k];
        const unsigned long long incl = wave_inclusive_scan_u64(tsum);
        if (tid == 0) s_block = SUM_NONE;
        if (lane == 63u) sh.wtot[wave] = incl;
        __syncthreads();
        unsigned long long run = acc + incl - tsum;
        for (uint32_t w = 0; w < wave; ++w) run += sh.wtot[w];
        uint32_t mine = SUM_NONE;
        for (uint32_t k = k0; k < k1; ++k) {
            if (bstop[k] != SUM_NONE || sum_crosses(md, run + bsum[k])) {
                mine = k;
                break;
            }
            run += bsum[k];
        }
        if (mine != SUM_NONE) atomicMin(&s_block, mine);
        __syncthreads();
        const uint32_t found = s_block;
        if (found == SUM_NONE) {
            if (tid == SUM_THREADS - 1) s_pre = run;  // every remaining block: the final total
        } else if (mine == found) {
            s_pre = run;  // the sum before block `found`
        }
        __syncthreads();
        const unsigned long long pre = s_pre;
        __syncthreads();
        if (found == SUM_NONE) {
            s_end = sum_value(md, pre);
            pos_end = n;
            break;
        }
        acc = pre;
        b = found;
        at = pos + b * PASS_BLOCK;
        scan_block = true;
    }
    if (tid == 0) {
        misc->sum = s_end;
        misc->sum_pos = pos_end;
        if (pos_end >= n || isnan(s_end)) split_finish(misc, s_end, n);
    }
}

// Order-preserving selection helpers: per-block counts then one scan (blocksums[b] = exclusive prefix).
__device__ __forceinline__ uint32_t block_count(bool f) {
    __shared__ uint32_t w[T256 / 64];
    const uint32_t c = (uint32_t)__popcll(__ballot(f));
    if ((threadIdx.x & 63u) == 0) w[threadIdx.x >> 6] = c;
    __syncthreads();
    uint32_t s = 0;
    for (int k = 0; k < T256 / 64; ++k) s += w[k];
    return s;
}
__device__ __forceinline__ uint32_t block_rank(bool f, uint32_t block_base) {
    __shared__ uint32_t w[T256 / 64];
    const uint64_t m = __ballot(f);
    if ((threadIdx.x & 63u) == 0) w[threadIdx.x >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t r = block_base;
    for (uint32_t k = 0; k < (threadIdx.x >> 6); ++k) r += w[k];
    return r + (uint32_t)__popcll(m & lanemask_lt());
}

__global__ void __launch_bounds__(1024) k_scan_blocks(uint32_t* __restrict__ sums, uint32_t nblocks,
                                                      uint32_t* __restrict__ total) {
    __shared__ uint32_t sw[16];
    const uint32_t tid = threadIdx.x;
    const uint32_t chunk = (nblocks + 1023u) / 1024u;
    const uint32_t b0 = tid * chunk, b1 = min(nblocks, b0 + chunk);
    uint32_t s = 0;
    for (uint32_t b = b0; b < b1; ++b) s += sums[b];
    const uint32_t inc = wave_inclusive_scan(s);
    if ((tid & 63u) == 63u) sw[tid >> 6] = inc;
    __syncthreads();
    if (tid == 0) {
        uint32_t r = 0;
        for (int w = 0; w < 16; ++w) {
            const uint32_t t = sw[w];
            sw[w] = r;
            r += t;
        }
        *total = r;
    }
    __syncthreads();
    uint32_t e = sw[tid >> 6] + inc - s;
    for (uint32_t b = b0; b < b1; ++b) {
        const uint32_t v = sums[b];
        sums[b] = e;
        e += v;
    }
}

// Initial candidates: leaves whose half area exceeds the low threshold (splits.rs:63-69).
__global__ void __launch_bounds__(T256) k_split_cand_count(const float* __restrict__ ha, uint32_t n,
                                                           const SplitMisc* misc, uint32_t* __restrict__ sums) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    const bool f = i < n && ha[i] > misc->thresh_low;
    const uint32_t c = block_count(f);
    if (threadIdx.x == 0) sums[blockIdx.x] = c;
}
__global__ void __launch_bounds__(T256) k_split_cand_write(const float* __restrict__ ha, uint32_t n,
                                                           const SplitMisc* misc, const uint32_t* __restrict__ sums,
                                                           uint32_t* __restrict__ cand) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    const bool f = i < n && ha[i] > misc->thresh_low;
    const uint32_t r = block_rank(f, sums[blockIdx.x]);
    if (f) cand[r] = i;
}

// splits.rs:72-112 for one candidate: the best of split_tests - 1 planes and the split decision.
__global__ void __launch_bounds__(T256) k_split_eval(const float4* __restrict__ tris, const float4* __restrict__ aabbs,
                                                     const uint32_t* __restrict__ indices,
                                                     const uint32_t* __restrict__ cand, uint32_t count,
                                                     const SplitMisc* misc, float4* __restrict__ halves,
                                                     uint8_t* __restrict__ flag, uint32_t* __restrict__ sums) {
    const uint32_t k = blockIdx.x * T256 + threadIdx.x;
    bool split = false;
    if (k < count) {
        const uint32_t c = cand[k];
        const Box aabb = load_box(aabbs, c);
        const uint32_t index = indices[c];
        const uint32_t axis = largest_axis(aabb);
        const float4 v[4] = {tris[3 * (size_t)index], tris[3 * (size_t)index + 1], tris[3 * (size_t)index + 2],
                             tris[3 * (size_t)index]};
        float best_cost = 3.40282347e+38f;
        Box left = aabb, right = aabb;
        const float amin = bget_lo(aabb, axis), amax = bget_hi(aabb, axis);
        for (uint32_t i = 1; i < 12u; ++i) {
            const float nf = (float)i / 12.0f;
            const float pos = amin * nf + amax * (1.0f - nf);
            Box tl = aabb, tr = aabb;
            bset_hi(tl, axis, pos);
            bset_lo(tr, axis, pos);
            Box sl, sr;
            split_triangle(axis, pos, v, sl, sr);
            tl = box_intersection(sl, tl);
            tr = box_intersection(sr, tr);
            const float area = half_area(tl) + half_area(tr);
            if (area < best_cost) {
                best_cost = area;
                left = tl;
                right = tr;
            }
        }
        const float old_cost = half_area(aabb);
        split = (misc->thresh_high > old_cost && best_cost * 1.6f < old_cost) || best_cost * 1.8f < old_cost;
        flag[k] = split ? 1u : 0u;
        if (split) {
            store_box(halves, 2 * k, left);
            store_box(halves, 2 * k + 1, right);
        }
    }
    const uint32_t c = block_count(split);
    if (threadIdx.x == 0) sums[blockIdx.x] = c;
}

// Applies the splits in candidate order: the r-th splitter's right half becomes leaf len + r (while the optional
// budget allows), and the extended candidate list cand ++ [len, len + accepted) is written to ext.
__global__ void __launch_bounds__(T256) k_split_apply(const uint32_t* __restrict__ cand, uint32_t count,
                                                      const uint8_t* __restrict__ flag,
                                                      const uint32_t* __restrict__ sums,
                                                      const float4* __restrict__ halves, uint32_t len,
                                                      uint32_t accept, float4* __restrict__ aabbs,
                                                      uint32_t* __restrict__ indices, uint32_t* __restrict__ ext) {
    const uint32_t k = blockIdx.x * T256 + threadIdx.x;
    const bool f = k < count && flag[k] != 0u;
    const uint32_t r = block_rank(f, sums[blockIdx.x]);
    if (k >= count) return;
    const uint32_t c = cand[k];
    ext[k] = c;
    if (f && r < accept) {
        store_box(aabbs, c, load_box(halves, 2 * k));
        store_box(aabbs, len + r, load_box(halves, 2 * k + 1));
        indices[len + r] = indices[c];
        ext[count + r] = len + r;
    }
}

// splits.rs:113-116: keep the candidates whose box still exceeds the low threshold, in order.
__global__ void __launch_bounds__(T256) k_split_keep_count(const uint32_t* __restrict__ ext, uint32_t count,
                                                           const float4* __restrict__ aabbs, const SplitMisc* misc,
                                                           uint32_t* __restrict__ sums) {
    const uint32_t k = blockIdx.x * T256 + threadIdx.x;
    const bool f = k < count && half_area(load_box(aabbs, ext[k])) > misc->thresh_low;
    const uint32_t c = block_count(f);
    if (threadIdx.x == 0) sums[blockIdx.x] = c;
}
__global__ void __launch_bounds__(T256) k_split_keep_write(const uint32_t* __restrict__ ext, uint32_t count,
                                                           const float4* __restrict__ aabbs, const SplitMisc* misc,
                                                           const uint32_t* __restrict__ sums,
                                                           uint32_t* __restrict__ cand) {
    const uint32_t k = blockIdx.x * T256 + threadIdx.x;
    const bool f = k < count && half_area(load_box(aabbs, ext[k])) > misc->thresh_low;
    const uint32_t r = block_rank(f, sums[blockIdx.x]);
    if (f) cand[r] = ext[k];
}


// ---- partial / full rebuild (ploc/rebuild.rs) ----

// compute_rebuild_path_flags (ploc/rebuild.rs:12-42), one lane per given leaf: flag it and walk up, stopping at the
// first ancestor that is already flagged.  The lane that set a flag always carries on upwards from it, so every
// ancestor ends up flagged whatever order the lanes run in (the result is the set of ancestors-or-self).
__global__ void __launch_bounds__(T256) k_path_flags(const uint32_t* __restrict__ parents, uint32_t n_nodes,
                                                     const uint32_t* __restrict__ leaves, uint32_t count,
                                                     uint8_t* __restrict__ flags, uint32_t* __restrict__ err) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i >= count) return;
    uint32_t index = leaves[i];
    if (index >= n_nodes) {
        atomicOr(err, 1u);  // the reference indexes out of bounds here
        return;
    }
    flags[index] = 1u;
    for (uint32_t steps = 0; index > 0u; ++steps) {
        index = parents[index];
        if (index >= n_nodes || steps > n_nodes) {
            atomicOr(err, 2u);  // not a tree
            return;
        }
        if (flags[index] != 0u) break;
        flags[index] = 1u;
    }
}

// Slot marks of a rebuild: bit 0 = the slot is freed (visited by the walk), bit 1 = its node becomes a cluster.
// Partial: the visited slots are level-list entries 1..count (the root stays); kept = a leaf or not flagged
// (ploc/rebuild.rs:118-123).
__global__ void __launch_bounds__(T256) k_rb_mark_partial(const Node32* __restrict__ nodes,
                                                          const uint32_t* __restrict__ list, uint32_t count,
                                                          const uint8_t* __restrict__ flags,
                                                          uint8_t* __restrict__ marks) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i >= count) return;
    const uint32_t x = list[1u + i];
    const Node32 nd = load_node(&nodes[x]);
    const bool kept = nd.prim_count != 0u || flags[x] == 0u;
    marks[x] = kept ? 3u : 1u;
}
// Full: every leaf (Bvh2Node::is_leaf, ploc/rebuild.rs:67-71) is a cluster.
__global__ void __launch_bounds__(T256) k_rb_mark_full(const Node32* __restrict__ nodes, uint32_t n,
                                                       uint8_t* __restrict__ marks) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i >= n) return;
    marks[i] = load_node(&nodes[i]).prim_count != 0u ? 2u : 0u;
}

// Two order-preserving selections over the slots in one pass: clusters (ascending slot) and freed left slots (odd
// slots, listed in descending order).  sums[b] / sums[nblk + b] are the per-block counts, then their prefixes.
__device__ __forceinline__ void rb_flags(const uint8_t* marks, uint32_t i, uint32_t n, bool& kept, bool& freed) {
    const uint32_t m = i < n ? marks[i] : 0u;
    kept = (m & 2u) != 0u;
    freed = (m & 1u) != 0u && (i & 1u) != 0u;
}
__global__ void __launch_bounds__(T256) k_rb_count(const uint8_t* __restrict__ marks, uint32_t n,
                                                   uint32_t* __restrict__ sums, uint32_t nblk) {
    __shared__ uint32_t wk[T256 / 64], wf[T256 / 64];
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    bool kept, freed;
    rb_flags(marks, i, n, kept, freed);
    const uint32_t ck = (uint32_t)__popcll(__ballot(kept)), cf = (uint32_t)__popcll(__ballot(freed));
    if ((threadIdx.x & 63u) == 0u) {
        wk[threadIdx.x >> 6] = ck;
        wf[threadIdx.x >> 6] = cf;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t a = 0, b = 0;
        for (int w = 0; w < T256 / 64; ++w) {
            a += wk[w];
            b += wf[w];
        }
        sums[blockIdx.x] = a;
        sums[nblk + blockIdx.x] = b;
    }
}
__global__ void __launch_bounds__(T256) k_rb_write(const uint8_t* __restrict__ marks, uint32_t n,
                                                   const uint32_t* __restrict__ sums, uint32_t nblk,
                                                   const uint32_t* __restrict__ totals,
                                                   uint32_t* __restrict__ kept_list, uint32_t* __restrict__ free_desc) {
    __shared__ uint32_t wk[T256 / 64], wf[T256 / 64];
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    bool kept, freed;
    rb_flags(marks, i, n, kept, freed);
    const uint64_t mk = __ballot(kept), mf = __ballot(freed);
    if ((threadIdx.x & 63u) == 0u) {
        wk[threadIdx.x >> 6] = (uint32_t)__popcll(mk);
        wf[threadIdx.x >> 6] = (uint32_t)__popcll(mf);
    }
    __syncthreads();
    uint32_t rk = sums[blockIdx.x], rf = sums[nblk + blockIdx.x];
    for (uint32_t w = 0; w < (threadIdx.x >> 6); ++w) {
        rk += wk[w];
        rf += wf[w];
    }
    rk += (uint32_t)__popcll(mk & lanemask_lt());
    rf += (uint32_t)__popcll(mf & lanemask_lt());
    if (kept) kept_list[rk] = i;
    if (freed) free_desc[totals[1] - 1u - rf] = i;
}

// Bvh2Node::set_aabb (bvh2/node.rs:116-125) for a list of nodes.
__global__ void __launch_bounds__(T256) k_set_node_aabbs(Node32* __restrict__ nodes, uint32_t n_nodes,
                                                         const uint32_t* __restrict__ ids,
                                                         const float4* __restrict__ aabbs, uint32_t count,
                                                         uint32_t* __restrict__ err) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i >= count) return;
    const uint32_t id = ids[i];
    if (id >= n_nodes) {
        atomicOr(err, 1u);
        return;
    }
    const float4 lo = aabbs[2 * (size_t)i], hi = aabbs[2 * (size_t)i + 1];
    Node32 nd = load_node(&nodes[id]);
    nd.minx = lo.x;
    nd.miny = lo.y;
    nd.minz = lo.z;
    nd.maxx = hi.x;
    nd.maxy = hi.y;
    nd.maxz = hi.z;
    store_node(&nodes[id], nd);
}

// compute_primitives_to_nodes (bvh2/mod.rs:647-663): out[prim] = slot of the leaf holding it.  The reference's
// serial loop leaves the highest such slot when a primitive sits in several leaves (pre-splits): atomicMax.
__global__ void __launch_bounds__(T256) k_prims_to_nodes(const Node32* __restrict__ nodes, uint32_t n_nodes,
                                                         const uint32_t* __restrict__ prim_indices,
                                                         uint32_t n_prims, uint32_t* __restrict__ out,
                                                         uint32_t* __restrict__ err) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i >= n_nodes) return;
    const Node32 nd = load_node(&nodes[i]);
    if (nd.prim_count == 0u) return;
    const uint64_t end = (uint64_t)nd.first_index + nd.prim_count;
    if (end > n_prims) {
        atomicOr(err, 1u);
        return;
    }
    for (uint32_t k = nd.first_index; k < (uint32_t)end; ++k) {
        const uint32_t prim = prim_indices[k];
        if (prim >= n_prims) {
            atomicOr(err, 1u);
            continue;
        }
        atomicMax(&out[prim], i + 1u);  // 0 = no leaf yet
    }
}
__global__ void __launch_bounds__(T256) k_p2n_finish(uint32_t* __restrict__ out, uint32_t n) {
    const uint32_t i = blockIdx.x * T256 + threadIdx.x;
    if (i < n) out[i] = out[i] == 0u ? 0xffffffffu : out[i] - 1u;  // INVALID where no leaf holds it
}

}  // namespace

#define OBVHS_TRY(expr, where)                                      \
    do {                                                            \
        hipError_t _e = (expr);                                     \
        if (_e != hipSuccess) return set_hip_error(ctx, _e, where); \
    } while (0)

// Top-down level lists of the BVH2 in bvh->nodes.  On return meta_host[2L], meta_host[2L+1] = start, count of
// level L for L < levels.
// With `expand` set, the lists hold what the partial rebuild's top-down walk visits (see k_levels_expand).
obvhs_status tree_levels(obvhs_context* ctx, obvhs_bvh2* bvh, std::vector<uint32_t>& meta_host, uint32_t& levels,
                         const uint8_t* expand) {
    hipStream_t st = ctx->stream;
    const uint32_t N = bvh->node_count;
    levels = 0;
    meta_host.clear();
    if (N == 0) return OBVHS_OK;
    QualityScratch& Q = ctx->quality;
    const uint32_t words = 2 * (MAX_TREE_LEVELS + 2) + 1;  // + the error word
    OBVHS_TRY(Q.level_list.ensure((size_t)N * 4), "alloc level list");
    OBVHS_TRY(Q.level_meta.ensure((size_t)words * 4), "alloc level meta");
    uint32_t* list = Q.level_list.as<uint32_t>();
    uint32_t* meta = Q.level_meta.as<uint32_t>();
    const Node32* nodes = bvh->nodes.as<Node32>();
    k_levels_init<<<8, T256, 0, st>>>(meta, words, list);
    uint32_t* h = static_cast<uint32_t*>(ctx->pinned);
    uint32_t level = 0;
    uint64_t known = 1;  // nodes at `level`, as last read (a level holds at most twice the nodes of the one above)
    for (;;) {
        for (uint32_t q = 0; q < LEVELS_PER_SYNC && level + q < MAX_TREE_LEVELS; ++q) {
            uint64_t bound = q < 31 ? std::min<uint64_t>(known << q, N) : N;
            if (level + q < 31) bound = std::min<uint64_t>(bound, 1ull << (level + q));
            const uint32_t g = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((bound + T256 - 1) / T256, 2048));
            k_levels_expand<<<g, T256, 0, st>>>(nodes, N, list, meta, level + q, words - 1, expand);
        }
        OBVHS_TRY(hipGetLastError(), "levels launch");
        level = std::min(level + LEVELS_PER_SYNC, MAX_TREE_LEVELS);
        OBVHS_TRY(hipMemcpyAsync(h, meta + 2 * level, 8, hipMemcpyDeviceToHost, st), "read level count");
        OBVHS_TRY(hipMemcpyAsync(h + 2, meta + words - 1, 4, hipMemcpyDeviceToHost, st), "read level error");
        OBVHS_TRY(hipStreamSynchronize(st), "levels sync");
        if (h[2] != 0u) return OBVHS_ERR_CAPACITY;  // not a tree
        if (h[1] == 0u) break;
        known = h[1];
        if (level >= MAX_TREE_LEVELS) return OBVHS_ERR_DEPTH;
    }
    meta_host.resize(2 * (size_t)(level + 1));
    OBVHS_TRY(hipMemcpy(meta_host.data(), meta, meta_host.size() * 4, hipMemcpyDeviceToHost), "read levels");
    uint32_t total = 0;
    levels = 0;
    for (uint32_t l = 0; l <= level; ++l) {
        if (meta_host[2 * l + 1] == 0u) break;
        total += meta_host[2 * l + 1];
        levels = l + 1;
    }
    if (!expand && total != N) return OBVHS_ERR_CAPACITY;  // not a tree: some node unreachable or reached twice
    return OBVHS_OK;
}

static obvhs_status refit_levels(obvhs_context* ctx, obvhs_bvh2* bvh, const std::vector<uint32_t>& meta,
                                 uint32_t levels) {
    hipStream_t st = ctx->stream;
    const uint32_t* list = ctx->quality.level_list.as<uint32_t>();
    Node32* nodes = bvh->nodes.as<Node32>();
    for (uint32_t l = levels; l-- > 0;) {
        const uint32_t s = meta[2 * l], c = meta[2 * l + 1];
        k_refit_level<<<(c + T256 - 1) / T256, T256, 0, st>>>(nodes, list, s, c);
    }
    OBVHS_TRY(hipGetLastError(), "refit launch");
    return OBVHS_OK;
}

// Cost table over the level lists (any node order), deepest level first.
obvhs_status cost_table_levels(obvhs_context* ctx, obvhs_bvh2* bvh, const DpView& dp, uint32_t* d_err) {
    std::vector<uint32_t> meta;
    uint32_t levels = 0;
    obvhs_status s = tree_levels(ctx, bvh, meta, levels, nullptr);
    if (s != OBVHS_OK) return s;
    hipStream_t st = ctx->stream;
    const uint32_t* list = ctx->quality.level_list.as<uint32_t>();
    const Node32* nodes = bvh->nodes.as<Node32>();
    for (uint32_t l = levels; l-- > 0;) {
        const uint32_t b = meta[2 * l], c = meta[2 * l + 1];
        k_dp_level<<<(c + T256 - 1) / T256, T256, 0, st>>>(nodes, dp, list, b, c, d_err);
    }
    OBVHS_TRY(hipGetLastError(), "cost table launch");
    return OBVHS_OK;
}

obvhs_status bvh2_refit_all(obvhs_context* ctx, obvhs_bvh2* bvh) {
    std::vector<uint32_t> meta;
    uint32_t levels = 0;
    obvhs_status s = tree_levels(ctx, bvh, meta, levels, nullptr);
    if (s != OBVHS_OK) return s;
    s = refit_levels(ctx, bvh, meta, levels);
    bvh->dp_max_prims = 0;  // boxes changed: the cost table is stale
    return s;
}

obvhs_status bvh2_refit_from_triangles(obvhs_context* ctx, obvhs_bvh2* bvh, const obvhs_triangle* d_tris) {
    if (bvh->node_count == 0) return OBVHS_OK;
    uint32_t* err = ctx->quality.misc.as<uint32_t>();
    OBVHS_TRY(ctx->quality.misc.ensure(4096), "alloc quality misc");
    err = ctx->quality.misc.as<uint32_t>();
    hipStream_t st = ctx->stream;
    k_fill_u32<<<1, T256, 0, st>>>(err, 0u, 1);
    k_leaf_refit<<<(bvh->node_count + T256 - 1) / T256, T256, 0, st>>>(
        bvh->nodes.as<Node32>(), bvh->node_count, bvh->prim_indices.as<uint32_t>(), bvh->prim_count,
        reinterpret_cast<const float4*>(d_tris), err);
    OBVHS_TRY(hipGetLastError(), "leaf refit launch");
    uint32_t* h = static_cast<uint32_t*>(ctx->pinned);
    OBVHS_TRY(hipMemcpyAsync(h, err, 4, hipMemcpyDeviceToHost, st), "read leaf refit err");
    OBVHS_TRY(hipStreamSynchronize(st), "leaf refit sync");
    if (h[0]) return device_error_to_status(h[0]);
    return bvh2_refit_all(ctx, bvh);
}

// ---- partial / full rebuild (ploc/rebuild.rs) ----

obvhs_status bvh2_rebuild_path_flags(obvhs_context* ctx, obvhs_bvh2* bvh, const uint32_t* d_leaves, uint32_t count,
                                     uint8_t* d_flags) {
    const uint32_t N = bvh->node_count;
    if (N < 2) return OBVHS_OK;  // rebuild.rs:17-19
    hipStream_t st = ctx->stream;
    QualityScratch& Q = ctx->quality;
    OBVHS_TRY(Q.parents.ensure((size_t)N * 4), "alloc parents");
    OBVHS_TRY(Q.misc.ensure(4096), "alloc quality misc");
    uint32_t* err = Q.misc.as<uint32_t>();
    uint32_t* parents = Q.parents.as<uint32_t>();
    k_fill_u32<<<1, 64, 0, st>>>(err, 0u, 1);
    OBVHS_TRY(hipMemsetAsync(d_flags, 0, N, st), "clear flags");
    k_parents<<<(N + T256 - 1) / T256, T256, 0, st>>>(bvh->nodes.as<Node32>(), N, parents, nullptr);
    if (count) k_path_flags<<<(count + T256 - 1) / T256, T256, 0, st>>>(parents, N, d_leaves, count, d_flags, err);
    OBVHS_TRY(hipGetLastError(), "path flags launch");
    uint32_t* h = static_cast<uint32_t*>(ctx->pinned);
    OBVHS_TRY(hipMemcpyAsync(h, err, 4, hipMemcpyDeviceToHost, st), "read path flags err");
    OBVHS_TRY(hipStreamSynchronize(st), "path flags sync");
    return h[0] ? OBVHS_ERR_INVALID_ARG : OBVHS_OK;
}

// Cluster list (ascending slot) and freed left slots (descending) from the slot marks; returns their sizes.
static obvhs_status rebuild_collect(obvhs_context* ctx, uint32_t N, uint32_t& kept, uint32_t& freed) {
    hipStream_t st = ctx->stream;
    QualityScratch& Q = ctx->quality;
    const uint32_t nblk = (N + T256 - 1) / T256;
    OBVHS_TRY(Q.sums.ensure((size_t)nblk * 8 + 16), "alloc rebuild sums");
    OBVHS_TRY(Q.misc.ensure(4096), "alloc quality misc");
    uint32_t* sums = Q.sums.as<uint32_t>();
    uint32_t* totals = Q.misc.as<uint32_t>() + 8;
    const uint8_t* marks = Q.marks.as<uint8_t>();
    k_rb_count<<<nblk, T256, 0, st>>>(marks, N, sums, nblk);
    k_scan_blocks<<<1, 1024, 0, st>>>(sums, nblk, totals);
    k_scan_blocks<<<1, 1024, 0, st>>>(sums + nblk, nblk, totals + 1);
    OBVHS_TRY(hipGetLastError(), "rebuild count launch");
    uint32_t* h = static_cast<uint32_t*>(ctx->pinned);
    OBVHS_TRY(hipMemcpyAsync(h, totals, 8, hipMemcpyDeviceToHost, st), "read rebuild counts");
    OBVHS_TRY(hipStreamSynchronize(st), "rebuild count sync");
    kept = h[0];
    freed = h[1];
    OBVHS_TRY(Q.kept.ensure((size_t)kept * 4 + 4), "alloc cluster list");
    OBVHS_TRY(Q.freed.ensure((size_t)freed * 4 + 4), "alloc free list");
    k_rb_write<<<nblk, T256, 0, st>>>(marks, N, sums, nblk, totals, Q.kept.as<uint32_t>(), Q.freed.as<uint32_t>());
    OBVHS_TRY(hipGetLastError(), "rebuild write launch");
    return OBVHS_OK;
}

static bool valid_search_distance(uint32_t r) {
    return r == 1 || r == 2 || r == 6 || r == 14 || r == 24 || r == 32;  // ploc/mod.rs:550-561
}

// PlocBuilder::partial_rebuild (ploc/rebuild.rs:98-134) with should_remove(i) = d_flags[i] != 0.
obvhs_status bvh2_partial_rebuild(obvhs_context* ctx, obvhs_bvh2* bvh, const uint8_t* d_flags,
                                  uint32_t search_distance, bool u128, uint32_t search_depth_threshold) {
    if (!valid_search_distance(search_distance)) return OBVHS_ERR_INVALID_ARG;
    const uint32_t N = bvh->node_count;
    if (N < 2) return OBVHS_OK;  // rebuild.rs:106-108
    std::vector<uint32_t> meta;
    uint32_t levels = 0;
    obvhs_status s = tree_levels(ctx, bvh, meta, levels, d_flags);
    if (s == OBVHS_ERR_CAPACITY) return OBVHS_ERR_INVALID_ARG;  // not a tree, or a left child at an even slot
    if (s != OBVHS_OK) return s;
    uint64_t total = 0;
    for (uint32_t l = 0; l < levels; ++l) total += meta[2 * l + 1];
    const uint32_t visited = (uint32_t)(total - 1);  // the root is not visited
    hipStream_t st = ctx->stream;
    QualityScratch& Q = ctx->quality;
    OBVHS_TRY(Q.marks.ensure(N), "alloc slot marks");
    OBVHS_TRY(hipMemsetAsync(Q.marks.ptr, 0, N, st), "clear slot marks");
    if (visited)
        k_rb_mark_partial<<<(visited + T256 - 1) / T256, T256, 0, st>>>(
            bvh->nodes.as<Node32>(), Q.level_list.as<uint32_t>(), visited, d_flags, Q.marks.as<uint8_t>());
    OBVHS_TRY(hipGetLastError(), "rebuild mark launch");
    uint32_t k = 0, f = 0;
    s = rebuild_collect(ctx, N, k, f);
    if (s != OBVHS_OK) return s;
    // The visited slots form whole sibling pairs at (odd, odd + 1), one pair per merge to come.
    if (k < 2 || visited != 2 * k - 2 || f != k - 1) return OBVHS_ERR_INVALID_ARG;
    return ploc_rebuild(ctx, bvh, Q.kept.as<uint32_t>(), k, Q.freed.as<uint32_t>(), search_distance, u128,
                        search_depth_threshold);
}

// PlocBuilder::full_rebuild (ploc/rebuild.rs:53-79).
obvhs_status bvh2_full_rebuild(obvhs_context* ctx, obvhs_bvh2* bvh, uint32_t search_distance, bool u128,
                               uint32_t search_depth_threshold) {
    if (!valid_search_distance(search_distance)) return OBVHS_ERR_INVALID_ARG;
    const uint32_t N = bvh->node_count;
    if (N < 2) return OBVHS_OK;  // rebuild.rs:60-62
    hipStream_t st = ctx->stream;
    QualityScratch& Q = ctx->quality;
    OBVHS_TRY(Q.marks.ensure(N), "alloc slot marks");
    k_rb_mark_full<<<(N + T256 - 1) / T256, T256, 0, st>>>(bvh->nodes.as<Node32>(), N, Q.marks.as<uint8_t>());
    OBVHS_TRY(hipGetLastError(), "rebuild mark launch");
    uint32_t k = 0, f = 0;
    obvhs_status s = rebuild_collect(ctx, N, k, f);
    if (s != OBVHS_OK) return s;
    return ploc_rebuild(ctx, bvh, Q.kept.as<uint32_t>(), k, nullptr, search_distance, u128,
                        search_depth_threshold);
}

obvhs_status bvh2_set_node_aabbs(obvhs_context* ctx, obvhs_bvh2* bvh, const uint32_t* d_ids,
                                 const obvhs_aabb* d_aabbs, uint32_t count) {
    if (count == 0) return OBVHS_OK;
    hipStream_t st = ctx->stream;
    OBVHS_TRY(ctx->quality.misc.ensure(4096), "alloc quality misc");
    uint32_t* err = ctx->quality.misc.as<uint32_t>();
    k_fill_u32<<<1, 64, 0, st>>>(err, 0u, 1);
    k_set_node_aabbs<<<(count + T256 - 1) / T256, T256, 0, st>>>(
        bvh->nodes.as<Node32>(), bvh->node_count, d_ids, reinterpret_cast<const float4*>(d_aabbs), count, err);
    OBVHS_TRY(hipGetLastError(), "set aabbs launch");
    uint32_t* h = static_cast<uint32_t*>(ctx->pinned);
    OBVHS_TRY(hipMemcpyAsync(h, err, 4, hipMemcpyDeviceToHost, st), "read set aabbs err");
    OBVHS_TRY(hipStreamSynchronize(st), "set aabbs sync");
    bvh->dp_max_prims = 0;
    return h[0] ? OBVHS_ERR_INVALID_ARG : OBVHS_OK;
}

obvhs_status bvh2_primitives_to_nodes(obvhs_context* ctx, const obvhs_bvh2* bvh, uint32_t* d_out) {
    const uint32_t P = bvh->prim_count, N = bvh->node_count;
    if (P == 0) return OBVHS_OK;
    hipStream_t st = ctx->stream;
    OBVHS_TRY(ctx->quality.misc.ensure(4096), "alloc quality misc");
    uint32_t* err = ctx->quality.misc.as<uint32_t>();
    k_fill_u32<<<1, 64, 0, st>>>(err, 0u, 1);
    k_fill_u32<<<std::min((P + T256 - 1) / T256, 4096u), T256, 0, st>>>(d_out, 0u, P);
    if (N)
        k_prims_to_nodes<<<(N + T256 - 1) / T256, T256, 0, st>>>(bvh->nodes.as<Node32>(), N,
                                                                 bvh->prim_indices.as<uint32_t>(), P, d_out, err);
    k_p2n_finish<<<(P + T256 - 1) / T256, T256, 0, st>>>(d_out, P);
    OBVHS_TRY(hipGetLastError(), "primitives to nodes launch");
    uint32_t* h = static_cast<uint32_t*>(ctx->pinned);
    OBVHS_TRY(hipMemcpyAsync(h, err, 4, hipMemcpyDeviceToHost, st), "read primitives to nodes err");
    OBVHS_TRY(hipStreamSynchronize(st), "primitives to nodes sync");
    return h[0] ? OBVHS_ERR_INVALID_ARG : OBVHS_OK;
}

// ReinsertionOptimizer::run (reinsertion.rs:40-61, 101-119).
obvhs_status bvh2_reinsertion(obvhs_context* ctx, obvhs_bvh2* bvh, float batch_size_ratio, const float* seq,
                              uint32_t seq_len, uint32_t dp_max_prims) {
    const uint32_t N = bvh->node_count;
    if (N == 0 || !(batch_size_ratio > 0.0f)) return OBVHS_OK;
    hipStream_t st = ctx->stream;
    uint32_t* h = static_cast<uint32_t*>(ctx->pinned);
    {
        uint32_t root_pc = 0;
        OBVHS_TRY(hipMemcpyAsync(h, reinterpret_cast<const char*>(bvh->nodes.ptr) + 12, 4, hipMemcpyDeviceToHost,
                                 st),
                  "read root");
        OBVHS_TRY(hipStreamSynchronize(st), "root sync");
        root_pc = h[0];
        if (root_pc != 0u) return OBVHS_OK;  // a leaf root (reinsertion.rs:45)
    }
    std::vector<float> default_seq;
    if (!seq || seq_len == 0) {
        for (int k = 1; k < 32; k += 2) default_seq.push_back(1.0f / (float)k);  // reinsertion.rs:105-110
        seq = default_seq.data();
        seq_len = (uint32_t)default_seq.size();
    }
    QualityScratch& Q = ctx->quality;
    OBVHS_TRY(Q.parents.ensure((size_t)N * 4), "alloc parents");
    OBVHS_TRY(Q.owner.ensure((size_t)N * 4), "alloc owner");
    OBVHS_TRY(Q.touched.ensure((size_t)N), "alloc touched");
    OBVHS_TRY(Q.misc.ensure(4096), "alloc quality misc");
    uint32_t* parents = Q.parents.as<uint32_t>();
    uint32_t* owner = Q.owner.as<uint32_t>();
    uint8_t* touched = Q.touched.as<uint8_t>();
    ReinsMisc* misc = Q.misc.as<ReinsMisc>();
    Node32* nodes = bvh->nodes.as<Node32>();
    const uint32_t nblk = (N + T256 - 1) / T256;
    k_fill_u32<<<std::min(nblk, 4096u), T256, 0, st>>>(owner, 0xffffffffu, N);

    std::vector<uint32_t> meta;
    uint32_t levels = 0;
    bool have_levels = false;  // meta holds the current tree's level lists
    obvhs_status s = OBVHS_OK;
    if (bvh->pass_ordered && bvh->ploc_iterations > 0) {
        // a tree fresh from PLOC: a pass raises a cluster's height by at most one, so its depth is at most the pass
        // count + 1 -- all the first step needs (its stack bound); no level listing
        levels = bvh->ploc_iterations + 1u;
    } else {
        s = tree_levels(ctx, bvh, meta, levels, nullptr);
        if (s != OBVHS_OK) return s;
        have_levels = true;
    }
    // After a step's moves the level lists are rebuilt at once (the next step's stack bound is the depth) and the
    // refit waits for the next step, which reads the boxes; the last one runs after the loop, fused with the
    // conversion's cost table when one is asked for.
    bool refit_pending = false;
    for (uint32_t si = 0; si < seq_len; ++si) {
        if (refit_pending) {
            s = refit_levels(ctx, bvh, meta, levels);
            if (s != OBVHS_OK) return s;
            refit_pending = false;
        }
        // reinsertion.rs:112-117 in f32, `as usize` saturating
        const float prod = ((float)N * batch_size_ratio) * seq[si];
        uint64_t batch = prod >= 1.8446744e19f ? UINT64_MAX : (prod > 0.0f ? (uint64_t)prod : 0ull);
        batch = std::max<uint64_t>(batch, 1);
        const uint32_t node_count = (uint32_t)std::min<uint64_t>(N, batch + 1);
        const uint32_t take = (uint32_t)std::min<uint64_t>(N, 2ull * node_count);
        const uint32_t ncand = take - 1;
        const uint32_t count = node_count - 1;
        if (count == 0) continue;
        OBVHS_TRY(Q.keys.ensure((size_t)ncand * 8), "alloc candidate keys");
        OBVHS_TRY(Q.vals.ensure((size_t)ncand * 4), "alloc candidate values");
        OBVHS_TRY(Q.rkeys.ensure((size_t)count * 8), "alloc reinsertion keys");
        OBVHS_TRY(Q.rvals.ensure((size_t)count * 4), "alloc reinsertion order");
        OBVHS_TRY(Q.rfrom.ensure((size_t)count * 4), "alloc reinsertion from");
        OBVHS_TRY(Q.rto.ensure((size_t)count * 4), "alloc reinsertion to");
        OBVHS_TRY(Q.conf.ensure((size_t)count * 20), "alloc conflicts");
        OBVHS_TRY(Q.state.ensure((size_t)count), "alloc states");
        OBVHS_TRY(Q.sort_scratch.ensure(radix_sort_scratch_bytes(std::max(ncand, count))), "alloc sort scratch");
        const uint32_t sd = levels + 2;
        uint32_t fthreads = std::min<uint32_t>(FIND_THREADS, (count + T256 - 1) / T256 * T256);
        const uint64_t stack_limit = 1ull << 31;  // bytes of stack columns at most
        while (fthreads > T256 && (uint64_t)fthreads * sd * 8 > stack_limit) fthreads = fthreads / 2 / T256 * T256;
        OBVHS_TRY(Q.stack.ensure((size_t)fthreads * sd * 8), "alloc reinsertion stacks");

        k_fill_u32<<<1, 64, 0, st>>>(reinterpret_cast<uint32_t*>(misc), 0u, sizeof(ReinsMisc) / 4);
        k_parents<<<nblk, T256, 0, st>>>(nodes, N, parents, &misc->negzero);
        k_cand_keys<<<(ncand + T256 - 1) / T256, T256, 0, st>>>(nodes, take, Q.keys.as<uint64_t>());
        OBVHS_TRY(hipGetLastError(), "candidate launch");
        OBVHS_TRY(radix_sort_pairs(Q.keys.as<uint64_t>(), Q.vals.as<uint32_t>(), true, ncand, 32,
                                   Q.sort_scratch.ptr, st, sort_items_for(ctx, ncand)),
                  "candidate sort");
        k_find_reinsertion<<<fthreads / T256, T256, 0, st>>>(nodes, N, parents, Q.vals.as<uint32_t>(), count,
                                                              Q.stack.as<uint2>(), sd, Q.rfrom.as<uint32_t>(),
                                                              Q.rto.as<uint32_t>(), Q.rkeys.as<uint64_t>(), misc);
        OBVHS_TRY(hipGetLastError(), "find reinsertion launch");
        OBVHS_TRY(radix_sort_pairs(Q.rkeys.as<uint64_t>(), Q.rvals.as<uint32_t>(), true, count, 32,
                                   Q.sort_scratch.ptr, st, sort_items_for(ctx, count)),
                  "reinsertion sort");
        OBVHS_TRY(hipMemcpyAsync(h, misc, 12, hipMemcpyDeviceToHost, st), "read reinsertion count");
        OBVHS_TRY(hipStreamSynchronize(st), "reinsertion count sync");
        if (h[0]) return device_error_to_status(h[0]);
        const uint32_t npos = h[1];
        const bool replay = h[2] != 0u;
        if (npos == 0) continue;
        const uint32_t pblk = (npos + T256 - 1) / T256;
        k_fill_u32<<<std::min(nblk, 4096u), T256, 0, st>>>(reinterpret_cast<uint32_t*>(touched), 0u,
                                                          ((size_t)N + 3) / 4);
        k_mis_init<<<pblk, T256, 0, st>>>(Q.rvals.as<uint32_t>(), misc, Q.rfrom.as<uint32_t>(),
                                          Q.rto.as<uint32_t>(), parents, Q.conf.as<uint32_t>(),
                                          Q.state.as<uint8_t>());
        // every round accepts at least the best-ranked undecided reinsertion, so npos rounds always suffice
        for (uint32_t batch = 0;; ++batch) {
            if ((uint64_t)batch * MIS_ROUNDS_PER_SYNC > (uint64_t)npos + MIS_ROUNDS_PER_SYNC) return OBVHS_ERR_DEVICE;
            for (uint32_t r = 0; r < MIS_ROUNDS_PER_SYNC; ++r) {
                k_mis_claim<<<pblk, T256, 0, st>>>(Q.conf.as<uint32_t>(), Q.state.as<uint8_t>(), misc, r, owner);
                k_mis_decide<<<pblk, T256, 0, st>>>(Q.conf.as<uint32_t>(), Q.state.as<uint8_t>(), misc, owner,
                                                    touched);
                k_mis_settle<<<pblk, T256, 0, st>>>(Q.conf.as<uint32_t>(), Q.state.as<uint8_t>(), misc, r, owner,
                                                    touched);
            }
            OBVHS_TRY(hipGetLastError(), "conflict rounds launch");
            OBVHS_TRY(hipMemcpyAsync(h, &misc->undecided[MIS_ROUNDS_PER_SYNC - 1], 4, hipMemcpyDeviceToHost, st),
                      "read undecided");
            OBVHS_TRY(hipStreamSynchronize(st), "conflict rounds sync");
            if (h[0] == 0u) break;
        }
        if (replay) {
            // leaves hold a -0: the crate's walks decide which zero bits each box keeps, so they run as written
            k_reins_replay<<<1, 64, 0, st>>>(Q.conf.as<uint32_t>(), Q.state.as<uint8_t>(), misc, nodes, parents);
        } else {
            // every zero coordinate is +0, so a value-equal box has equal bits and one full refit (deferred to the
            // next step or fused with the cost table) leaves exactly the boxes the crate's walks leave
            k_reins_apply<<<pblk, T256, 0, st>>>(Q.conf.as<uint32_t>(), Q.state.as<uint8_t>(), misc, nodes);
        }
        OBVHS_TRY(hipGetLastError(), "apply launch");
        s = tree_levels(ctx, bvh, meta, levels, nullptr);
        if (s != OBVHS_OK) return s;
        have_levels = true;
        refit_pending = !replay;
    }
    bvh->pass_ordered = false;  // children are no longer stored in PLOC pass order (reinsertion.rs:102)
    bvh->dp_max_prims = 0;
    if (dp_max_prims != 0u && bvh->prim_count > 1) {
        if (!have_levels) {  // no step moved anything: the level lists of the tree as it is
            s = tree_levels(ctx, bvh, meta, levels, nullptr);
            if (s != OBVHS_OK) return s;
        }
        OBVHS_TRY(bvh_dp_alloc(bvh), "alloc cost table");
        const DpView dp = bvh_dp_view(bvh, dp_max_prims);
        k_fill_u32<<<1, 64, 0, st>>>(&misc->err, 0u, 1);
        for (uint32_t l = levels; l-- > 0;) {
            const uint32_t b = meta[2 * l], c = meta[2 * l + 1];
            if (refit_pending)
                k_refit_dp_level<true><<<(c + T256 - 1) / T256, T256, 0, st>>>(nodes, dp, Q.level_list.as<uint32_t>(),
                                                                              b, c, &misc->err);
            else
                k_refit_dp_level<false><<<(c + T256 - 1) / T256, T256, 0, st>>>(nodes, dp, Q.level_list.as<uint32_t>(),
                                                                               b, c, &misc->err);
        }
        OBVHS_TRY(hipGetLastError(), "refit and cost table launch");
        OBVHS_TRY(hipMemcpyAsync(h, &misc->err, 4, hipMemcpyDeviceToHost, st), "read cost table error");
        OBVHS_TRY(hipStreamSynchronize(st), "cost table sync");
        if (h[0]) return device_error_to_status(h[0]);
        bvh->dp_max_prims = dp_max_prims;
    } else if (refit_pending) {
        s = refit_levels(ctx, bvh, meta, levels);
        if (s != OBVHS_OK) return s;
    }
    return OBVHS_OK;
}

// Pre-split leaf set (cwbvh/builder.rs:26-55).  max_total = 0: no limit (the crate's behaviour).  On return
// Q.split_aabbs / Q.split_indices hold `out_n` leaves.
obvhs_status presplit_leaves(obvhs_context* ctx, const obvhs_triangle* d_tris, uint32_t n, uint64_t max_total,
                             uint32_t& out_n, float* out_avg, float* out_largest) {
    hipStream_t st = ctx->stream;
    QualityScratch& Q = ctx->quality;
    out_n = n;
    if (n == 0) return OBVHS_OK;
    uint64_t cap = std::max<uint64_t>((uint64_t)n + (uint64_t)n / 2, 1024);
    if (max_total) cap = std::max<uint64_t>(std::min<uint64_t>(cap, max_total), n);
    if (cap >= (1ull << 31)) return OBVHS_ERR_CAPACITY;
    OBVHS_TRY(Q.split_aabbs.ensure(cap * 32), "alloc split boxes");
    OBVHS_TRY(Q.split_indices.ensure(cap * 4), "alloc split indices");
    OBVHS_TRY(Q.ha.ensure((size_t)n * 4), "alloc half areas");
    OBVHS_TRY(Q.misc.ensure(4096), "alloc quality misc");
    SplitMisc* misc = reinterpret_cast<SplitMisc*>(Q.misc.as<char>() + 1024);
    uint32_t* h = static_cast<uint32_t*>(ctx->pinned);
    const float4* tris = reinterpret_cast<const float4*>(d_tris);
    const uint32_t nblk = (n + T256 - 1) / T256;
    const uint32_t max_nb = (n + PASS_BLOCK - 1) / PASS_BLOCK;  // the sum's passes use the block sums first
    OBVHS_TRY(Q.sums.ensure(std::max<size_t>((size_t)(nblk + 1) * 4, (size_t)max_nb * 12) + 64), "alloc block sums");
    k_fill_u32<<<1, 64, 0, st>>>(reinterpret_cast<uint32_t*>(misc), 0u, sizeof(SplitMisc) / 4);
    k_split_prepare<<<std::min(nblk, PREPARE_MAX_BLOCKS), T256, 0, st>>>(tris, n, Q.split_aabbs.as<float4>(), Q.split_indices.as<uint32_t>(),
                                           Q.ha.as<float>(), misc);
    {  // the exact running sum: one workgroup over its first SUM_SERIAL_SPAN elements, then passes over all CUs
        k_sum_serial<<<1, SUM_THREADS, 0, st>>>(Q.ha.as<float>(), n, std::min(n, SUM_SERIAL_SPAN), misc);
        OBVHS_TRY(hipGetLastError(), "sum launch");
        unsigned long long* bsum = Q.sums.as<unsigned long long>();
        uint32_t* bstop = reinterpret_cast<uint32_t*>(bsum + max_nb);
        uint32_t start = 0;
        bool last_pass = false;
        while (n > SUM_SERIAL_SPAN) {
            OBVHS_TRY(hipMemcpyAsync(h, &misc->sum, 8, hipMemcpyDeviceToHost, st), "read sum");
            OBVHS_TRY(hipStreamSynchronize(st), "sum sync");
            float sum;
            uint32_t pos;
            memcpy(&sum, h, 4);
            memcpy(&pos, h + 1, 4);
            if (pos >= n || std::isnan(sum)) break;
            const bool dense = last_pass && pos - start < SUM_SERIAL_SPAN / 4;  // the last pass ended soon
            start = pos;
            if (dense) {
                k_sum_serial<<<1, SUM_THREADS, 0, st>>>(
                    Q.ha.as<float>(), n, (uint32_t)std::min<uint64_t>(n, (uint64_t)pos + SUM_SERIAL_SPAN), misc);
            } else {
                const uint32_t nb = (n - pos + PASS_BLOCK - 1) / PASS_BLOCK;
                k_sum_blocks<<<nb, T256, 0, st>>>(Q.ha.as<float>(), n, pos, sum, bsum, bstop);
                k_sum_pass<<<1, SUM_THREADS, 0, st>>>(Q.ha.as<float>(), n, pos, sum, nb, bsum, bstop, misc);
            }
            last_pass = !dense;
            OBVHS_TRY(hipGetLastError(), "sum pass launch");
        }
    }
    // candidate lists hold distinct leaf indices (at most cap); the extended list at most twice that
    OBVHS_TRY(Q.cand_a.ensure(cap * 4), "alloc candidates");
    OBVHS_TRY(Q.cand_b.ensure(cap * 8), "alloc extended candidates");
    k_split_cand_count<<<nblk, T256, 0, st>>>(Q.ha.as<float>(), n, misc, Q.sums.as<uint32_t>());
    k_scan_blocks<<<1, 1024, 0, st>>>(Q.sums.as<uint32_t>(), nblk, &misc->kept);
    k_split_cand_write<<<nblk, T256, 0, st>>>(Q.ha.as<float>(), n, misc, Q.sums.as<uint32_t>(),
                                              Q.cand_a.as<uint32_t>());
    OBVHS_TRY(hipGetLastError(), "split prepare launch");
    OBVHS_TRY(hipMemcpyAsync(h, misc, sizeof(SplitMisc), hipMemcpyDeviceToHost, st), "read candidates");
    OBVHS_TRY(hipStreamSynchronize(st), "split prepare sync");
    SplitMisc hm;
    memcpy(&hm, h, sizeof(SplitMisc));
    if (out_avg) *out_avg = hm.avg;
    if (out_largest) *out_largest = f32_from_order_key(hm.largest_key);
    uint32_t count = hm.kept;
    uint64_t len = n;
    for (uint32_t it = 0; it < 12 && count > 0; ++it) {  // max_iterations = 12 (splits.rs:31)
        // every candidate may append one leaf: grow the leaf arrays (keeping their contents) when needed
        if (len + count > cap) {
            uint64_t ncap = std::max<uint64_t>(cap * 2, len + count);
            if (max_total) ncap = std::max<uint64_t>(std::min<uint64_t>(ncap, max_total), len);
            if (ncap >= (1ull << 31)) return OBVHS_ERR_CAPACITY;
            DevBuf na, ni, nc;
            OBVHS_TRY(na.ensure(ncap * 32), "grow split boxes");
            OBVHS_TRY(ni.ensure(ncap * 4), "grow split indices");
            OBVHS_TRY(nc.ensure(ncap * 4), "grow candidates");
            OBVHS_TRY(hipMemcpyAsync(na.ptr, Q.split_aabbs.ptr, len * 32, hipMemcpyDeviceToDevice, st), "copy boxes");
            OBVHS_TRY(hipMemcpyAsync(ni.ptr, Q.split_indices.ptr, len * 4, hipMemcpyDeviceToDevice, st),
                      "copy indices");
            OBVHS_TRY(hipMemcpyAsync(nc.ptr, Q.cand_a.ptr, (size_t)count * 4, hipMemcpyDeviceToDevice, st),
                      "copy candidates");
            OBVHS_TRY(hipStreamSynchronize(st), "grow sync");
            Q.split_aabbs.release();
            Q.split_indices.release();
            Q.cand_a.release();
            Q.split_aabbs = na;
            Q.split_indices = ni;
            Q.cand_a = nc;
            OBVHS_TRY(Q.cand_b.ensure(ncap * 8), "grow extended candidates");
            cap = ncap;
        }
        const uint32_t cblk = (count + T256 - 1) / T256;
        OBVHS_TRY(Q.halves.ensure((size_t)count * 64), "alloc split halves");
        OBVHS_TRY(Q.flags.ensure((size_t)count), "alloc split flags");
        OBVHS_TRY(Q.sums.ensure((size_t)(2 * cblk + 1) * 4 + 64), "alloc block sums");
        k_split_eval<<<cblk, T256, 0, st>>>(tris, Q.split_aabbs.as<float4>(), Q.split_indices.as<uint32_t>(),
                                            Q.cand_a.as<uint32_t>(), count, misc, Q.halves.as<float4>(),
                                            Q.flags.as<uint8_t>(), Q.sums.as<uint32_t>());
        k_scan_blocks<<<1, 1024, 0, st>>>(Q.sums.as<uint32_t>(), cblk, &misc->splits);
        OBVHS_TRY(hipGetLastError(), "split eval launch");
        OBVHS_TRY(hipMemcpyAsync(h, &misc->splits, 4, hipMemcpyDeviceToHost, st), "read splits");
        OBVHS_TRY(hipStreamSynchronize(st), "split eval sync");
        const uint32_t splits = h[0];
        uint32_t accept = splits;
        if (max_total) accept = (uint32_t)std::min<uint64_t>(splits, max_total > len ? max_total - len : 0);
        k_split_apply<<<cblk, T256, 0, st>>>(Q.cand_a.as<uint32_t>(), count, Q.flags.as<uint8_t>(),
                                             Q.sums.as<uint32_t>(), Q.halves.as<float4>(), (uint32_t)len, accept,
                                             Q.split_aabbs.as<float4>(), Q.split_indices.as<uint32_t>(),
                                             Q.cand_b.as<uint32_t>());
        OBVHS_TRY(hipGetLastError(), "split apply launch");
        len += accept;
        if (accept == 0) break;  // splits.rs:113: no candidate was added
        const uint32_t ext = count + accept;
        const uint32_t eblk = (ext + T256 - 1) / T256;
        OBVHS_TRY(Q.sums.ensure((size_t)(eblk + 1) * 4 + 64), "alloc block sums");
        k_split_keep_count<<<eblk, T256, 0, st>>>(Q.cand_b.as<uint32_t>(), ext, Q.split_aabbs.as<float4>(), misc,
                                                  Q.sums.as<uint32_t>());
        k_scan_blocks<<<1, 1024, 0, st>>>(Q.sums.as<uint32_t>(), eblk, &misc->kept);
        k_split_keep_write<<<eblk, T256, 0, st>>>(Q.cand_b.as<uint32_t>(), ext, Q.split_aabbs.as<float4>(), misc,
                                                  Q.sums.as<uint32_t>(), Q.cand_a.as<uint32_t>());
        OBVHS_TRY(hipGetLastError(), "split filter launch");
        OBVHS_TRY(hipMemcpyAsync(h, &misc->kept, 4, hipMemcpyDeviceToHost, st), "read kept");
        OBVHS_TRY(hipStreamSynchronize(st), "split filter sync");
        count = h[0];
    }
    if (len >= (1ull << 30)) return OBVHS_ERR_CAPACITY;
    out_n = (uint32_t)len;
    return OBVHS_OK;
}

}  // namespace obvhs
