"""The tuning table and its measurements (tune.hip).

* Sort tile and traversal LDS depth change speed only: every tile gives the same Morton order and tree, every depth
  the same hits and occlusion flags, for single and batched traversal; the tuners return a candidate and store it
  for the size class.
* The PLOC distance for Auto builds changes the tree, so it is pinned three ways: the tuner's choice is the
  smallest distance whose returned cost is within the tolerance of the reference (the lowest cost, or the cost at
  the baseline distance), tuning twice gives the same costs and choice, and an Auto build equals the build with the
  chosen distance byte for byte.
* Batched Auto builds resolve the distance per mesh (each mesh by its own size class, all in one launch sequence):
  every mesh equals its single build with its class's distance, and the oracle's build with that distance."""
import numpy as np
import pytest

import obvhs_amd as ob
import oracle_lib as O
from obvhs_amd import scenes

pytestmark = pytest.mark.gpu

DEFAULT = {"ploc_radius": [14] * 4, "sort_items": [8] * 4, "trav_lds_depth": [8] * 4, "trav_chunk": [0] * 4}


@pytest.fixture
def fresh(ctx):
    ctx.set_tune_table(DEFAULT)
    yield ctx
    ctx.set_tune_table(DEFAULT)


def with_table(ctx, **entries):
    t = {k: list(v) for k, v in DEFAULT.items()}
    for k, v in entries.items():
        t[k] = [v] * 4 if np.isscalar(v) else list(v)
    ctx.set_tune_table(t)


def test_table_roundtrip_and_validation(fresh):
    assert fresh.tune_table() == DEFAULT
    assert [ob.size_class(n) for n in (0, 4095, 4096, 65535, 65536, 1048575, 1048576)] == [0, 0, 1, 1, 2, 2, 3]
    with pytest.raises(ob.ObvhsError):
        with_table(fresh, sort_items=5)
    with pytest.raises(ob.ObvhsError):
        with_table(fresh, ploc_radius=3)
    with pytest.raises(ob.ObvhsError):
        with_table(fresh, trav_chunk=2)
    assert fresh.tune_table() == DEFAULT


@pytest.mark.parametrize("sp", [64, 128])
def test_sort_tiles_give_identical_builds(fresh, sp):
    scenes_ = {"soup5k": scenes.soup(5000, seed=2), "heightfield": scenes.heightfield_scene(256, 2000, seed=1)}
    for name, tris in scenes_.items():
        print(f"[case] sort tiles {name} U{sp}", flush=True)
        ref = None
        for items in (8, 4, 16):
            with_table(fresh, sort_items=items)
            b = ob.PlocBuilder(fresh).build(14, tris, None, ob.SortPrecision(sp))
            codes, order = b.sort_trace()
            got = (b.nodes.tobytes(), order.tobytes(), codes.tobytes())
            if ref is None:
                ref = got
            assert got == ref, f"{name}: tile {items} differs"
    fresh.set_tune_table(DEFAULT)
    items = ob.tune_sort(fresh, 300000)
    assert items in (4, 8, 16)
    assert fresh.tune_table()["sort_items"][ob.size_class(300000)] == items


def test_traversal_depths_give_identical_hits(fresh):
    tris, rays = scenes.heightfield_workload(256, 2000, 480, 270, seed=3)
    cw = ob.build_cwbvh_from_tris(fresh, tris, ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3))
    bt = cw.reorder_triangles(tris)
    d_rays = fresh.upload(rays)
    ref_h, ref_m = cw.ray_traverse(d_rays, bt), cw.ray_traverse_miss(d_rays, bt)
    for depth in (4, 16):
        print(f"[case] traversal LDS depth {depth}", flush=True)
        with_table(fresh, trav_lds_depth=depth)
        assert cw.ray_traverse(d_rays, bt).tobytes() == ref_h.tobytes()
        assert np.array_equal(cw.ray_traverse_miss(d_rays, bt), ref_m)
    fresh.set_tune_table(DEFAULT)
    d = ob.tune_traversal(fresh, cw, d_rays, bt)
    assert d in (4, 8, 16)
    assert fresh.tune_table()["trav_lds_depth"][ob.size_class(len(tris))] == d
    assert fresh.tune_table()["trav_chunk"][ob.size_class(len(tris))] in (0, 1)
    assert cw.ray_traverse(d_rays, bt).tobytes() == ref_h.tobytes()


def test_batched_traversal_depths_give_identical_hits(fresh):
    meshes = [scenes.soup(s, seed=i) for i, s in enumerate([64, 3000, 20000, 500, 9000])]
    offsets = np.zeros(len(meshes) + 1, dtype=np.uint32)
    offsets[1:] = np.cumsum([len(m) for m in meshes])
    tris = np.concatenate(meshes)
    params = ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3)
    batch = ob.build_cwbvh_batch(fresh, tris, offsets, params)
    bt = batch.reorder_triangles(tris)
    rays = scenes.pinhole_rays(96, 96, eye=(0.5, 0.5, -1.5), target=(0.5, 0.5, 0.5))
    ray_mesh = (np.arange(len(rays)) % len(meshes)).astype(np.uint32)
    ref = batch.ray_traverse(rays, ray_mesh, bt)
    for depth in (4, 16):
        print(f"[case] batched traversal LDS depth {depth}", flush=True)
        with_table(fresh, trav_lds_depth=depth)
        assert batch.ray_traverse(rays, ray_mesh, bt).tobytes() == ref.tobytes()


def test_ploc_radius_tuner_is_reproducible_and_auto_builds_match(fresh):
    tris = scenes.heightfield_scene(128, 300, seed=6)
    print(f"[case] radius tuner n={len(tris)}", flush=True)
    r, costs = ob.tune_ploc_radius(fresh, tris, 3, 0.01, baseline=0)
    r2, costs2 = ob.tune_ploc_radius(fresh, tris, 3, 0.01, baseline=0)
    assert r == r2 and costs.tobytes() == costs2.tobytes()
    cands = [1, 2, 6, 14, 24, 32]
    assert np.all(np.isfinite(costs)) and np.all(costs > 0)
    assert int(r) == next(c for c, v in zip(cands, costs) if v <= costs.min() * np.float32(1.01))
    rb, costs_b = ob.tune_ploc_radius(fresh, tris, 3, 0.02, baseline=14)  # the fastest as good as distance 14
    assert costs_b.tobytes() == costs.tobytes()
    assert int(rb) == next(c for c, v in zip(cands, costs) if v <= costs[3] * np.float32(1.02)) and int(rb) <= 14
    r, _ = ob.tune_ploc_radius(fresh, tris, 3, 0.01, baseline=0)
    assert fresh.tune_table()["ploc_radius"][ob.size_class(len(tris))] == int(r)
    p_auto = ob.BvhBuildParams(ploc_search_distance=ob.PlocSearchDistance.Auto, max_prims_per_leaf=3)
    p_expl = ob.BvhBuildParams(ploc_search_distance=int(r), max_prims_per_leaf=3)
    a, e = ob.build_cwbvh_from_tris(fresh, tris, p_auto), ob.build_cwbvh_from_tris(fresh, tris, p_expl)
    assert a.nodes.tobytes() == e.nodes.tobytes() and a.primitive_indices.tobytes() == e.primitive_indices.tobytes()
    # a stricter tolerance can only pick the same or a larger distance
    r0, _ = ob.tune_ploc_radius(fresh, tris, 3, 0.0, baseline=0)
    assert int(r0) >= int(r) and costs[cands.index(int(r0))] == costs.min()


def test_batched_auto_radius_is_per_mesh(fresh):
    table = {"ploc_radius": [2, 6, 24, 14], "sort_items": [8] * 4, "trav_lds_depth": [8] * 4}
    fresh.set_tune_table(table)
    sizes = [64, 5000, 70000, 1, 3000, 40000, 0, 200]
    meshes = [scenes.soup(s, seed=10 + i) for i, s in enumerate(sizes)]
    offsets = np.zeros(len(meshes) + 1, dtype=np.uint32)
    offsets[1:] = np.cumsum([len(m) for m in meshes])
    tris = np.concatenate(meshes)
    print("[case] batched auto radius", flush=True)
    params = ob.BvhBuildParams(ploc_search_distance=0, max_prims_per_leaf=3)
    batch = ob.build_cwbvh_batch(fresh, tris, offsets, params)
    nodes, prims = batch.all_nodes(), batch.all_primitive_indices()
    for i, m in enumerate(meshes):
        r = table["ploc_radius"][ob.size_class(len(m))]
        no, nc, po, pc = batch.mesh_info(i)
        single = ob.build_cwbvh_from_tris(fresh, m, ob.BvhBuildParams(ploc_search_distance=r, max_prims_per_leaf=3))
        assert nodes[no:no + nc].tobytes() == single.nodes.tobytes(), f"mesh {i} ({len(m)} tris, r={r})"
        assert prims[po:po + pc].tobytes() == single.primitive_indices.tobytes()
        if 0 < len(m) <= 5000:
            _, rcw, rprims = O.build_cwbvh(m, r, 0, 3)
            assert nodes[no:no + nc].tobytes() == rcw.tobytes()
            assert np.array_equal(prims[po:po + pc], rprims)


def test_auto_rebuilds_use_the_table(fresh):
    with_table(fresh, ploc_radius=6)
    tris = scenes.soup(20000, seed=4)
    a = ob.PlocBuilder(fresh).build(0, tris)
    b = ob.PlocBuilder(fresh).build(6, tris)
    assert a.nodes.tobytes() == b.nodes.tobytes()
    leaves = np.nonzero(a.nodes["prim_count"] != 0)[0].astype(np.uint32)[::13]
    ob.PlocBuilder(fresh).partial_rebuild(a, ob.compute_rebuild_path_flags(a, leaves), 0)
    ob.PlocBuilder(fresh).partial_rebuild(b, ob.compute_rebuild_path_flags(b, leaves), 6)
    assert a.nodes.tobytes() == b.nodes.tobytes()
    ob.PlocBuilder(fresh).full_rebuild(a, 0)
    ob.PlocBuilder(fresh).full_rebuild(b, 6)
    assert a.nodes.tobytes() == b.nodes.tobytes()
