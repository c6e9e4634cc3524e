"""GPU parity of partial and full rebuilds (ploc/rebuild.rs) against the oracle, bit for bit.

* compute_rebuild_path_flags (rebuild.rs:12-42) and compute_primitives_to_nodes (bvh2/mod.rs:647-663);
* partial_rebuild (rebuild.rs:98-134) over random leaf subsets, with nothing and with everything flagged, at every
  search distance, with and without search_depth_threshold, on trees reshaped by reinsertion and on trees uploaded
  from the host; enough clusters that the grid passes (not only the one-workgroup tail) place into freed slots;
* full_rebuild (rebuild.rs:53-79);
* the physics-example flow (examples/physics.rs:426-448): move some triangles, set their leaves' boxes, flag the
  paths, partial rebuild, re-emit the CWBVH and trace: nodes, primitive indices and hits equal the oracle's;
* at the 1.29M-triangle heightfield size, where the oracle's rebuild is too slow to run per case: partial with
  every leaf flagged equals the GPU's own full rebuild, the tree validates tight, and traced hits equal brute force
  on sampled rays."""
import numpy as np
import pytest

import obvhs_amd as ob
import oracle_lib as O
from obvhs_amd import scenes
from obvhs_amd.test_util import hash_noise

pytestmark = pytest.mark.gpu


def same(a, b, what):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    if a.size == 0:
        return
    av, bv = a.view(np.uint8).reshape(len(a), -1), b.view(np.uint8).reshape(len(b), -1)
    bad = np.nonzero((av != bv).any(axis=1))[0]
    assert len(bad) == 0, f"{what}: {len(bad)} of {len(a)} records differ, first at {bad[:5]}"


def leaf_ids(nodes):
    return np.nonzero(nodes["prim_count"] != 0)[0].astype(np.uint32)


def pick_leaves(nodes, frac, seed):
    leaves = leaf_ids(nodes)
    return leaves[hash_noise(np.arange(len(leaves), dtype=np.uint32), 7, seed) < frac]


SCENES = {
    "soup20k": lambda: scenes.soup(20000, seed=3),
    "heightfield": lambda: scenes.heightfield_scene(96, 300, seed=1),
    "thin": lambda: np.concatenate([scenes.heightfield_scene(48, 60, seed=2), scenes.thin_triangles(800, seed=2)]),
}


@pytest.mark.parametrize("scene", list(SCENES))
@pytest.mark.parametrize("frac", [0.001, 0.05, 0.5])
def test_path_flags_and_prims_to_nodes_match_oracle(ctx, scene, frac):
    print(f"[case] path flags {scene} frac={frac}", flush=True)
    tris = SCENES[scene]()
    bvh = ob.PlocBuilder(ctx).build(14, tris)
    nodes = bvh.nodes
    leaves = pick_leaves(nodes, frac, seed=5)
    flags = ob.compute_rebuild_path_flags(bvh, leaves).to_host()
    rc, want = O.rebuild_path_flags(nodes, leaves)
    assert rc == 0
    np.testing.assert_array_equal(flags, want)
    p2n = bvh.primitives_to_nodes().to_host()
    rc, want = O.primitives_to_nodes(nodes, bvh.primitive_indices)
    assert rc == 0
    np.testing.assert_array_equal(p2n, want)


@pytest.mark.parametrize("scene", list(SCENES))
@pytest.mark.parametrize("frac,sd,sdt,reins", [(0.002, 1, 0, 0.0), (0.05, 1, 0, 0.1), (0.3, 14, 0, 0.0),
                                               (0.3, 6, 2, 0.1), (0.9, 32, 0, 0.0), (1.0, 2, 0, 0.0)])
def test_partial_rebuild_matches_oracle(ctx, scene, frac, sd, sdt, reins):
    print(f"[case] partial rebuild {scene} frac={frac} sd={sd} sdt={sdt} reinsertion={reins}", flush=True)
    tris = SCENES[scene]()
    bvh = ob.PlocBuilder(ctx).build(sd, tris)
    if reins:
        bvh.reinsertion(reins, [1.0])
    before = bvh.nodes
    leaves = pick_leaves(before, frac, seed=int(frac * 1000) + sd)
    flags = ob.compute_rebuild_path_flags(bvh, leaves)
    ob.PlocBuilder(ctx).partial_rebuild(bvh, flags, sd, ob.SortPrecision.U64, sdt)
    rc, want, _ = O.partial_rebuild(before, flags.to_host(), sd, sdt)
    assert rc == 0
    same(bvh.nodes, want, "BVH2 after partial rebuild")


@pytest.mark.parametrize("scene", list(SCENES))
@pytest.mark.parametrize("frac", [0.05, 1.0])
def test_rebuilds_u128_match_oracle(ctx, scene, frac):
    """partial_rebuild and full_rebuild with SortPrecision::U128."""
    print(f"[case] U128 rebuilds {scene} frac={frac}", flush=True)
    tris = SCENES[scene]()
    bvh = ob.PlocBuilder(ctx).build(14, tris, None, ob.SortPrecision.U128)
    before = bvh.nodes
    flags = ob.compute_rebuild_path_flags(bvh, pick_leaves(before, frac, seed=11))
    ob.PlocBuilder(ctx).partial_rebuild(bvh, flags, 14, ob.SortPrecision.U128, 0)
    rc, want, _ = O.partial_rebuild(before, flags.to_host(), 14, 0, sort_precision=128)
    assert rc == 0
    same(bvh.nodes, want, "BVH2 after a U128 partial rebuild")
    before = bvh.nodes
    ob.PlocBuilder(ctx).full_rebuild(bvh, 6, ob.SortPrecision.U128, 2)
    rc, want, _ = O.full_rebuild(before, 6, 2, sort_precision=128)
    assert rc == 0
    same(bvh.nodes, want, "BVH2 after a U128 full rebuild")


@pytest.mark.parametrize("scene", list(SCENES))
@pytest.mark.parametrize("sd,sdt", [(1, 0), (14, 0), (24, 3)])
def test_full_rebuild_matches_oracle(ctx, scene, sd, sdt):
    print(f"[case] full rebuild {scene} sd={sd} sdt={sdt}", flush=True)
    tris = SCENES[scene]()
    bvh = ob.PlocBuilder(ctx).build(14, tris)
    bvh.reinsertion(0.1, [1.0])
    before = bvh.nodes
    ob.PlocBuilder(ctx).full_rebuild(bvh, sd, ob.SortPrecision.U64, sdt)
    rc, want, _ = O.full_rebuild(before, sd, sdt)
    assert rc == 0
    same(bvh.nodes, want, "BVH2 after full rebuild")


def test_partial_rebuild_edge_cases(ctx):
    for n in (1, 2, 3, 5, 64):
        tris = scenes.soup(n, seed=n)
        print(f"[case] partial rebuild edge n={n}", flush=True)
        bvh = ob.PlocBuilder(ctx).build(1, tris)
        before = bvh.nodes
        for flags_host in (np.zeros(len(before), np.uint8), np.ones(len(before), np.uint8)):
            b2 = ob.Bvh2.from_host(ctx, before, bvh.primitive_indices)
            ob.PlocBuilder(ctx).partial_rebuild(b2, flags_host.astype(bool), 1)
            rc, want, _ = O.partial_rebuild(before, flags_host, 1, 0)
            assert rc == 0
            same(b2.nodes, want, f"n={n} flags={int(flags_host[0]) if len(flags_host) else '-'}")
        b3 = ob.Bvh2.from_host(ctx, before, bvh.primitive_indices)
        ob.PlocBuilder(ctx).full_rebuild(b3, 1)
        rc, want, _ = O.full_rebuild(before, 1, 0)
        assert rc == 0
        same(b3.nodes, want, f"full n={n}")
    bvh = ob.PlocBuilder(ctx).build(1, scenes.soup(100, seed=1))
    with pytest.raises(ob.ObvhsError):
        ob.compute_rebuild_path_flags(bvh, [bvh.node_count])  # the reference indexes out of bounds
    with pytest.raises(ValueError):
        ob.PlocBuilder(ctx).partial_rebuild(bvh, np.zeros(bvh.node_count, bool), 3)  # not a PlocSearchDistance


def test_partial_rebuild_of_uploaded_oracle_tree(ctx):
    tris = SCENES["heightfield"]()
    ref = O.ploc_build(O.tri_aabbs(tris), 1, 0)
    nodes = O.reinsertion(ref["nodes"], 0.3, [1.0, 0.5])
    bvh = ob.Bvh2.from_host(ctx, nodes, ref["prim_indices"])
    same(bvh.nodes, nodes, "upload round trip")
    leaves = pick_leaves(nodes, 0.1, seed=2)
    flags = ob.compute_rebuild_path_flags(bvh, leaves)
    ob.PlocBuilder(ctx).partial_rebuild(bvh, flags, 1)
    rc, want, _ = O.partial_rebuild(nodes, flags.to_host(), 1, 0)
    assert rc == 0
    same(bvh.nodes, want, "partial rebuild of an uploaded tree")


@pytest.mark.parametrize("n_moved", [50, 5000])
def test_physics_flow_matches_oracle(ctx, n_moved):
    print(f"[case] physics flow moved={n_moved}", flush=True)
    tris = scenes.heightfield_scene(128, 400, seed=4)
    moved_ids = np.unique((hash_noise(np.arange(n_moved, dtype=np.uint32), 1, 9) * len(tris)).astype(np.int64))
    moved = tris.copy()
    moved[moved_ids] = scenes.jitter_triangles(tris[moved_ids], seed=6, frac=0.02)
    bvh = ob.PlocBuilder(ctx).build(1, tris)
    prims = bvh.primitive_indices
    before = bvh.nodes
    # GPU: primitive -> leaf, set the leaves' boxes, flags, partial rebuild, convert, trace
    p2n = bvh.primitives_to_nodes().to_host()
    ids = p2n[moved_ids].astype(np.uint32)
    boxes = O.tri_aabbs(moved[moved_ids])
    bvh.set_node_aabbs(ids, boxes)
    flags = ob.compute_rebuild_path_flags(bvh, ids)
    ob.PlocBuilder(ctx).partial_rebuild(bvh, flags, 1)
    cw = ob.bvh2_to_cwbvh(bvh, 3, True, False)
    # oracle
    upd = before.copy()
    upd["min"][ids] = boxes["min"]
    upd["max"][ids] = boxes["max"]
    rc, oflags = O.rebuild_path_flags(upd, ids)
    assert rc == 0
    rc, want, _ = O.partial_rebuild(upd, oflags, 1, 0)
    assert rc == 0
    same(bvh.nodes, want, "BVH2 after the physics-flow partial rebuild")
    rcw, rprims, _ = O.bvh2_to_cwbvh(want, prims, 3, True)
    same(cw.nodes, rcw, "CWBVH re-emitted after the partial rebuild")
    np.testing.assert_array_equal(cw.primitive_indices, rprims)
    rays = scenes.pinhole_rays(160, 120, eye=(0.5, 0.6, -0.4), target=(0.5, 0.0, 0.5))
    bvh_tris = moved[rprims]
    d_bvh_tris = ctx.upload(bvh_tris)
    hits = cw.ray_traverse(ctx.upload(rays), d_bvh_tris)
    want_hits = O.traverse(rcw, bvh_tris, rays, 0, threads=8)
    same(hits, want_hits, "hits after the partial rebuild")


def test_heightfield_full_size_properties(ctx):
    tris, rays = scenes.heightfield_workload(724, 20000, 256, 144, seed=0)
    print(f"[case] full-size rebuild n={len(tris)}", flush=True)
    a = ob.PlocBuilder(ctx).build(14, tris)
    b = ob.PlocBuilder(ctx).build(14, tris)
    leaves = leaf_ids(a.nodes)
    ob.PlocBuilder(ctx).partial_rebuild(a, ob.compute_rebuild_path_flags(a, leaves), 14)
    ob.PlocBuilder(ctx).full_rebuild(b, 14)
    nodes = a.nodes
    same(nodes, b.nodes, "partial(all leaves) vs full rebuild")
    prims = a.primitive_indices
    rc, _ = O.bvh2_validate(nodes, prims, O.tri_aabbs(tris), tight_fit=True, ordered=False)
    assert rc == 0
    # a partial rebuild of 1% of the leaves, then trace against brute force on a sample of rays
    c = ob.PlocBuilder(ctx).build(14, tris)
    few = pick_leaves(c.nodes, 0.01, seed=3)
    ob.PlocBuilder(ctx).partial_rebuild(c, ob.compute_rebuild_path_flags(c, few), 14)
    cn = c.nodes
    rc, _ = O.bvh2_validate(cn, prims, O.tri_aabbs(tris), tight_fit=True, ordered=False)
    assert rc == 0
    cw = ob.bvh2_to_cwbvh(c, 3, True, False)
    cwp = cw.primitive_indices
    sample = np.ascontiguousarray(rays[::149])
    hits = cw.ray_traverse(ctx.upload(sample), ctx.upload(tris[cwp]))
    bf = O.brute_force(tris, sample)
    hit = bf["primitive_id"] != 0xFFFFFFFF
    np.testing.assert_array_equal(hits["primitive_id"] != 0xFFFFFFFF, hit)
    np.testing.assert_array_equal(hits["t"][hit], bf["t"][hit])
