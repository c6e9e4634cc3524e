"""Pins of the oracle's BVH2 -> CWBVH conversion choices (cwbvh/bvh2_to_cwbvh.rs) and of safe_inverse (ray.rs:6-12).

What pins them (none of these re-types the oracle's recurrence):
  * the cost DP (calculate_cost, bvh2_to_cwbvh.rs:220-344): an exhaustive enumeration of every collapse of a small
    BVH2 (<= 12 leaves) into 8-wide nodes -- each internal CWBVH node takes an antichain of 2..8 descendants
    covering its leaves, each taken as a leaf (<= max_prims_per_leaf primitives) or as a further internal node --
    under the crate's cost model (internal: half area; leaf: primitives x half area x 0.3); the DP's root cost
    equals the minimum and the CWBVH the oracle emits has exactly that cost;
  * the DP's strict `<` tie rules (bvh2_to_cwbvh.rs:285, 296, 321): primitives of zero area make every cost 0, so
    every comparison is a tie; with strict `<` no later option ever replaces an earlier one, so each internal BVH2
    node keeps INTERNAL with the first distribution at every i, and the CWBVH is binary: one CWBVH node per
    internal BVH2 node, two children each (hand-derived; `<=` in either rule is caught);
  * the greedy octant assignment (bvh2_to_cwbvh.rs:402-467) and its direction LUT (bvh2_to_cwbvh.rs:40-49): a
    hand-derived three-child node with two children of identical boxes competing for one slot (a flipped LUT sign
    and a last-minimum `<=` are caught); a slot-major scan is provably equivalent to the child-major one (see
    test_slot_major_scan_is_equivalent), so no test can or needs to tell them apart;
  * safe_inverse: hand-derived values at +-0, +-EPSILON (inclusive bound), just above it, +-2 EPSILON, infinities,
    NaN and a denormal."""
import numpy as np
import pytest

import oracle_lib as O
from obvhs_amd.types import BVH2_NODE, F32_EPS, aabbs_from_minmax, safe_inverse

PRIM_COST = 0.3
MUT_CW_DIST_LE, MUT_CW_LEAF_LE, MUT_CW_LUT_FLIP_X, MUT_CW_SLOT_LE, MUT_CW_SLOT_MAJOR = 16, 17, 18, 19, 20


@pytest.fixture
def mutation():
    yield lambda m: O.lib().orc_set_mutation(m)
    O.lib().orc_set_mutation(0)


def _half_area(lo, hi):
    d = np.asarray(hi, dtype=np.float64) - np.asarray(lo, dtype=np.float64)
    return (d[0] + d[1]) * d[2] + d[0] * d[1]


class Collapses:
    """Exhaustive minimum of the CWBVH cost model over every collapse of a BVH2 into 8-wide nodes."""

    def __init__(self, nodes, max_prims):
        self.nodes, self.maxp = nodes, max_prims
        n = len(nodes)
        self.leaf = nodes["prim_count"] != 0
        self.ha = [_half_area(nodes["min"][i, :3], nodes["max"][i, :3]) for i in range(n)]
        self.nprims = [0] * n
        self._count(0)
        self._cuts, self._best_any, self._best_internal = {}, {}, {}

    def _count(self, i):
        if self.leaf[i]:
            self.nprims[i] = int(self.nodes["prim_count"][i])
        else:
            f = int(self.nodes["first_index"][i])
            self._count(f)
            self._count(f + 1)
            self.nprims[i] = self.nprims[f] + self.nprims[f + 1]

    def cuts(self, i):
        """Every antichain of i's subtree (i included) that covers its leaves, as tuples of at most 8 nodes."""
        if i not in self._cuts:
            out = [(i,)]
            if not self.leaf[i]:
                f = int(self.nodes["first_index"][i])
                for a in self.cuts(f):
                    for b in self.cuts(f + 1):
                        if len(a) + len(b) <= 8:
                            out.append(a + b)
            self._cuts[i] = out
        return self._cuts[i]

    def best_internal(self, i):
        if i not in self._best_internal:
            f = int(self.nodes["first_index"][i])
            best = min(sum(self.best_any(c) for c in a + b)
                       for a in self.cuts(f) for b in self.cuts(f + 1) if len(a) + len(b) <= 8)
            self._best_internal[i] = self.ha[i] + best
        return self._best_internal[i]

    def best_any(self, i):
        if i not in self._best_any:
            opts = []
            if self.nprims[i] <= self.maxp:
                opts.append(self.nprims[i] * self.ha[i] * PRIM_COST)
            if not self.leaf[i]:
                opts.append(self.best_internal(i))
            self._best_any[i] = min(opts)
        return self._best_any[i]


def _emitted_cost(cw, cw_prims, cw2b, bvh_nodes, prim_aabbs):
    """Cost of the emitted CWBVH under the same model: the half area of each node's BVH2 node, plus for each leaf
    child its primitive count x the half area of its primitives' union x 0.3 (the BVH2 is tight)."""
    total = 0.0
    for k in range(len(cw)):
        b = int(cw2b[k])
        total += _half_area(bvh_nodes["min"][b, :3], bvh_nodes["max"][b, :3])
        for s in range(8):
            meta = int(cw[k]["child_meta"][s])
            if meta == 0 or (cw[k]["imask"] >> s) & 1:
                continue
            count = {0b001: 1, 0b011: 2, 0b111: 3}[meta >> 5]
            start = int(cw[k]["primitive_base_idx"]) + (meta & 0x1F)
            p = cw_prims[start:start + count]
            lo = prim_aabbs["min"][p, :3].min(axis=0)
            hi = prim_aabbs["max"][p, :3].max(axis=0)
            total += count * _half_area(lo, hi) * PRIM_COST
    return total


def _random_aabbs(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.random((n, 3)).astype(np.float32)
    h = (rng.random((n, 3)) * 0.2 + 0.01).astype(np.float32)
    return aabbs_from_minmax(c - h, c + h)


@pytest.mark.parametrize("max_prims", [1, 2, 3])
def test_cost_dp_equals_exhaustive_minimum(max_prims):
    checked = 0
    for seed in range(40):
        n = 2 + seed % 11  # 2..12 leaves
        aabbs = _random_aabbs(n, 1000 + seed)
        b = O.ploc_build(aabbs, (1, 2, 6, 14)[seed % 4], 0)
        nodes = b["nodes"]
        cost, kind, _, _ = O.cw_cost_table(nodes, max_prims)
        bf = Collapses(nodes, max_prims)
        want_root = min(bf.best_internal(0), bf.nprims[0] * bf.ha[0] * PRIM_COST if n <= max_prims else np.inf)
        assert cost[0, 0] == pytest.approx(want_root, rel=1e-5), (seed, n)
        cw, prims, cw2b = O.bvh2_to_cwbvh(nodes, b["prim_indices"], max_prims, True)
        got = _emitted_cost(cw, prims, cw2b, nodes, aabbs)
        assert got == pytest.approx(bf.best_internal(0), rel=1e-5), (seed, n)
        # cases where the collapse is a real choice: more than one CWBVH node, or a leaf of several primitives
        multi = ((cw["child_meta"] >> 5) > 1).any()
        checked += len(cw) > 1 or bool(multi)
    assert checked >= 10


def _point_aabbs(n):
    p = np.full((n, 3), 0.5, dtype=np.float32)
    return aabbs_from_minmax(p, p)


@pytest.mark.parametrize("max_prims", [1, 3])
def test_zero_area_ties_keep_first_choice(mutation, max_prims):
    aabbs = _point_aabbs(11)
    b = O.ploc_build(aabbs, 14, 0)
    nodes = b["nodes"]
    internal = set(np.nonzero(nodes["prim_count"] == 0)[0].tolist())

    def binary_tree(cw, cw2b):
        return (len(cw) == len(internal) and set(cw2b.tolist()) == internal
                and all(bin(int(x)).count("1") == 2 for x in (cw["child_meta"] != 0).dot(1 << np.arange(8))))

    cw, _, cw2b = O.bvh2_to_cwbvh(nodes, b["prim_indices"], max_prims, True)
    assert binary_tree(cw, cw2b)
    _, kind, dl, dr = O.cw_cost_table(nodes, max_prims)
    inner = sorted(internal)
    assert (kind[inner] == 1).all() and (dl[inner] == 0).all() and (dr[inner] == 6).all()
    for planted in (MUT_CW_DIST_LE, MUT_CW_LEAF_LE):
        if planted == MUT_CW_LEAF_LE and max_prims == 1:
            continue  # no internal node holds a single primitive: the leaf option never exists
        mutation(planted)
        cw_m, _, cw2b_m = O.bvh2_to_cwbvh(nodes, b["prim_indices"], max_prims, True)
        mutation(0)
        assert not binary_tree(cw_m, cw2b_m), planted


def _node(lo, hi, prim_count, first_index):
    n = np.zeros(1, dtype=BVH2_NODE)
    n["min"][0, :3] = lo
    n["max"][0, :3] = hi
    n["prim_count"], n["first_index"] = prim_count, first_index
    return n[0]


def _octant_tree():
    """root(X(A, B), C): A and B the same box [0.5, 1]^3 (primitives 0 and 1), C = [-1, -0.5]^3 (primitive 2)."""
    a, c, r = ((0.5,) * 3, (1.0,) * 3), ((-1.0,) * 3, (-0.5,) * 3), ((-1.0,) * 3, (1.0,) * 3)
    nodes = np.array([_node(*r, 0, 1), _node(*a, 0, 3), _node(*c, 1, 2), _node(*a, 1, 0), _node(*a, 1, 1)],
                     dtype=BVH2_NODE)
    return nodes, np.arange(3, dtype=np.uint32)


def test_octant_assignment_hand_derived(mutation):
    """max_prims_per_leaf 1.  DP: X (two primitives) distributes at every i >= 1, so the root's children are A, B, C
    in get_children order.  Node centre p = 0; child centres v = +-0.75 per axis; cost[c][s] = d_s . v with
    d_s = (s&4 ? -1 : 1, s&2 ? -1 : 1, s&1 ? -1 : 1).  A and B cost -2.25 at slot 7, C -2.25 at slot 0: a
    three-way tie.  Child-major scan, strict `<`: round 1 takes (A, 7); round 2 (C, 0) (-2.25 beats B's best
    remaining, -0.75); round 3 B takes the first slot at -0.75, slot 3.  Primitives are appended in slot order:
    [C, B, A] = [2, 1, 0]; child_meta = running primitive offset | 0b0010_0000 (one primitive)."""
    nodes, prims = _octant_tree()
    cw, cw_prims, _ = O.bvh2_to_cwbvh(nodes, prims, 1, True)
    assert len(cw) == 1 and cw[0]["imask"] == 0
    assert cw_prims.tolist() == [2, 1, 0]
    assert cw[0]["child_meta"].tolist() == [0x20, 0, 0, 0x21, 0, 0, 0, 0x22]
    # e = 2^ceil(log2(2/255)) = 2^-6, p = (-1,-1,-1): A, B -> [96, 128], C -> [0, 32] on every axis
    assert cw[0]["e"].tolist() == [127 - 6] * 3
    for ax in "xyz":
        assert cw[0][f"child_min_{ax}"].tolist() == [0, 0, 0, 96, 0, 0, 0, 96]
        assert cw[0][f"child_max_{ax}"].tolist() == [32, 0, 0, 128, 0, 0, 0, 128]
    # a flipped x sign in the LUT sends A to slot 3, C to slot 4 and B to slot 1 (its first slot at -0.75), so the
    # slot order of the primitives becomes B, A, C; the last minimum (`<=`) gives C slot 0, B slot 7, A slot 6
    for planted, want_prims in ((MUT_CW_LUT_FLIP_X, None), (MUT_CW_SLOT_LE, [2, 0, 1])):
        mutation(planted)
        cw_m, prims_m, _ = O.bvh2_to_cwbvh(nodes, prims, 1, True)
        mutation(0)
        assert cw_m.tobytes() != cw.tobytes(), planted
        if want_prims is not None:
            assert prims_m.tolist() == want_prims
    mutation(MUT_CW_LUT_FLIP_X)
    cw_m, _, _ = O.bvh2_to_cwbvh(nodes, prims, 1, True)
    mutation(0)
    assert cw_m[0]["child_meta"].tolist() == [0, 0x20, 0, 0x21, 0x22, 0, 0, 0]


def _tie_heavy_aabbs(n, seed):
    """Boxes on a coarse lattice with many duplicate centres: ties everywhere in the slot costs."""
    rng = np.random.default_rng(seed)
    c = (rng.integers(0, 3, size=(n, 3)) * 0.5).astype(np.float32)
    h = (rng.integers(1, 3, size=(n, 3)) * 0.125).astype(np.float32)
    return aabbs_from_minmax(c - h, c + h)


def test_slot_major_scan_is_equivalent(mutation):
    """Why a slot-major scan cannot be told apart: each round takes the first pair (child, slot) of minimum cost in
    scan order.  Let T be the minimum-cost pairs of a round.  Child-major takes (ca, sx): ca the smallest child in
    T, sx its smallest slot in T; slot-major takes (cb, sy): sy the smallest slot in T, cb its smallest child in T.
    If ca == cb then ca has a pair at sy, so sx <= sy <= sx; if sx == sy then cb <= ca <= cb.  So the two picks
    are equal or share neither child nor slot, and then the other pick is still a minimum-cost pair of the next
    round (its child unassigned, its slot free, the minimum cannot fall).  Induction over rounds: the scans agree
    on every assignment.  Checked here on tie-heavy inputs, where a last-minimum scan (`<=`) does differ."""
    differs_le = 0
    for seed in range(30):
        aabbs = _tie_heavy_aabbs(60, seed)
        b = O.ploc_build(aabbs, 14, 0)
        cw, prims, _ = O.bvh2_to_cwbvh(b["nodes"], b["prim_indices"], 3, True)
        mutation(MUT_CW_SLOT_MAJOR)
        cw_m, prims_m, _ = O.bvh2_to_cwbvh(b["nodes"], b["prim_indices"], 3, True)
        mutation(MUT_CW_SLOT_LE)
        cw_le, prims_le, _ = O.bvh2_to_cwbvh(b["nodes"], b["prim_indices"], 3, True)
        mutation(0)
        assert cw_m.tobytes() == cw.tobytes() and prims_m.tobytes() == prims.tobytes()
        differs_le += cw_le.tobytes() != cw.tobytes() or prims_le.tobytes() != prims.tobytes()
    assert differs_le >= 10


def test_safe_inverse_hand_derived():
    eps = F32_EPS  # 2^-23
    above = np.nextafter(eps, np.float32(1.0))
    below = np.nextafter(eps, np.float32(0.0))
    x = np.array([0.0, -0.0, eps, -eps, below, above, -above, 2 * eps, -2 * eps, 3.0, np.inf, -np.inf, 1e-40],
                 dtype=np.float32)
    got = safe_inverse(x)
    want = np.array([2.0 ** 23, -(2.0 ** 23), 2.0 ** 23, -(2.0 ** 23), 2.0 ** 23,
                     8388607.0, -8388607.0,  # 1 / (2^-23 (1 + 2^-23)) rounds to 2^23 - 1
                     2.0 ** 22, -(2.0 ** 22), np.float32(1.0) / np.float32(3.0), 0.0, -0.0, 2.0 ** 23],
                    dtype=np.float32)
    assert got.view(np.uint32).tolist() == want.view(np.uint32).tolist()
    assert np.isnan(safe_inverse(np.array([np.nan], dtype=np.float32)))[0]
