"""Race check by determinism: the GPU path's outputs are bit-identical across repeated runs, across contexts, and
across launch shapes that change which lanes and workgroups handle which items.

The kernels share state through atomics (scene bounds, level and work counters, the reinsertion optimizer's
candidate counter and conflict marks, the batched build's per-mesh segments) and their workgroups finish in no
fixed order, so a missing barrier or an order-dependent write shows up as a run-to-run difference:
* full-quality builds (pre-splits, reinsertion with a ratio sequence) run three times in one context and once in a
  second context give identical CWBVH nodes and primitive indices;
* a batched build of mixed-size meshes run twice gives identical nodes, and equals the single builds;
* closest-hit and occlusion traversal give identical results when repeated and when the rays are split into uneven
  launches (every ray then runs on a different lane and spill column), also between launches large enough for
  wave-owned ray chunks and the one-ray-per-lane launches of their pieces;
* partial rebuilds and refits repeated from the same starting tree give identical trees."""
import numpy as np
import pytest

import obvhs_amd as ob
import oracle_lib as O
from obvhs_amd import scenes
from obvhs_amd.test_util import hash_noise

pytestmark = pytest.mark.gpu


def quality_scene():
    return np.concatenate([scenes.heightfield_scene(256, 2000, seed=3), scenes.thin_triangles(20000, seed=3)])


def test_quality_builds_repeat_bit_for_bit(ctx):
    tris = quality_scene()
    p = ob.BvhBuildParams(pre_split=True, ploc_search_distance=14, reinsertion_batch_ratio=0.05,
                          max_prims_per_leaf=3)
    outs = []
    for k in range(3):
        print(f"[case] quality build run {k} n={len(tris)}", flush=True)
        cw = ob.build_cwbvh_from_tris(ctx, tris, p, ratio_sequence=[1.0, 1.0 / 3.0])
        outs.append((cw.nodes.tobytes(), cw.primitive_indices.tobytes()))
    print("[case] quality build in a second context", flush=True)
    with ob.Context(0) as c2:
        cw = ob.build_cwbvh_from_tris(c2, tris, p, ratio_sequence=[1.0, 1.0 / 3.0])
        outs.append((cw.nodes.tobytes(), cw.primitive_indices.tobytes()))
    for k, o in enumerate(outs[1:], 1):
        assert o[0] == outs[0][0], f"CWBVH nodes differ in run {k}"
        assert o[1] == outs[0][1], f"primitive indices differ in run {k}"


def test_batched_build_repeats_bit_for_bit(ctx):
    sizes = [64, 65536, 100, 5000, 1, 0, 20000, 777] * 8
    meshes = [scenes.soup(s, seed=i) for i, s in enumerate(sizes)]
    offsets = np.zeros(len(meshes) + 1, dtype=np.uint32)
    offsets[1:] = np.cumsum([len(m) for m in meshes])
    tris = np.concatenate(meshes)
    params = ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3)
    print(f"[case] batched build x2, {len(meshes)} meshes", flush=True)
    a = ob.build_cwbvh_batch(ctx, tris, offsets, params)
    na, pa = a.all_nodes(), a.all_primitive_indices()
    b = ob.build_cwbvh_batch(ctx, tris, offsets, params)
    assert b.all_nodes().tobytes() == na.tobytes()
    assert b.all_primitive_indices().tobytes() == pa.tobytes()
    for i in (1, 3, 6):  # a mesh equals its own single build
        no, nc, po, pc = a.mesh_info(i)
        cw = ob.build_cwbvh_from_tris(ctx, meshes[i], params)
        assert na[no:no + nc].tobytes() == cw.nodes.tobytes()
        assert pa[po:po + pc].tobytes() == cw.primitive_indices.tobytes()


def test_traversal_repeats_and_launch_shapes(ctx):
    tris, rays = scenes.heightfield_workload(256, 2000, 480, 270, seed=2)
    cw = ob.build_cwbvh_from_tris(ctx, tris, ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3))
    bt = cw.reorder_triangles(tris)
    d_rays = ctx.upload(rays)
    print(f"[case] traversal x2 + uneven launches, {len(rays)} rays", flush=True)
    h1, h2 = cw.ray_traverse(d_rays, bt), cw.ray_traverse(d_rays, bt)
    assert h1.tobytes() == h2.tobytes()
    cuts = [0, 1, 63, 65, 4097, 50001, len(rays)]
    parts = [cw.ray_traverse(ctx.upload(np.ascontiguousarray(rays[a:b])), bt) for a, b in zip(cuts[:-1], cuts[1:])]
    assert np.concatenate(parts).tobytes() == h1.tobytes()
    m1 = cw.ray_traverse_miss(d_rays, bt)
    parts = [cw.ray_traverse_miss(ctx.upload(np.ascontiguousarray(rays[a:b])), bt)
             for a, b in zip(cuts[:-1], cuts[1:])]
    assert np.array_equal(np.concatenate(parts), m1)
    assert np.array_equal(cw.ray_traverse_miss(d_rays, bt), m1)


def test_rebuilds_and_refits_repeat_bit_for_bit(ctx):
    tris = scenes.heightfield_scene(192, 1000, seed=5)
    base = ob.PlocBuilder(ctx).build(14, tris)
    base.reinsertion(0.1, [1.0])
    nodes0, prims0 = base.nodes, base.primitive_indices
    leaves = np.nonzero(nodes0["prim_count"] != 0)[0].astype(np.uint32)
    few = leaves[hash_noise(np.arange(len(leaves), dtype=np.uint32), 7, 4) < 0.05]
    moved = scenes.jitter_triangles(tris, seed=9, frac=0.01)
    outs = []
    for k in range(3):
        print(f"[case] partial rebuild + refit run {k}", flush=True)
        b = ob.Bvh2.from_host(ctx, nodes0, prims0)
        ob.PlocBuilder(ctx).partial_rebuild(b, ob.compute_rebuild_path_flags(b, few), 14)
        rebuilt = b.nodes.tobytes()
        b.refit_from_triangles(moved)
        outs.append((rebuilt, b.nodes.tobytes()))
    assert outs[1] == outs[0] and outs[2] == outs[0]


@pytest.mark.parametrize("depth", [4, 8, 16])
def test_chunked_launches_match_one_ray_per_lane(ctx, depth):
    """With the tuning table's trav_chunk set, launches of at least two rays per resident lane run as wave-owned
    ray rows with lanes refilled (k_traverse_chunk); smaller launches still run one ray per lane (k_traverse).  The
    same rays traced both ways give identical hits and occlusion flags, for every LDS stack depth (4 puts the most
    entries in the spill columns), on coherent primary rays and on incoherent bounce rays, and a strided sample
    equals the oracle."""
    saved = ctx.tune_table()
    t = ctx.tune_table()
    t["trav_chunk"] = [1] * 4
    t["trav_lds_depth"] = [depth] * 4
    ctx.set_tune_table(t)
    try:
        _chunked_vs_pieces(ctx, depth)
    finally:
        ctx.set_tune_table(saved)


def test_chunked_batched_launch_matches_one_ray_per_lane(ctx):
    """The batched traversal's chunked launch (each refilled ray selects its mesh) gives the hits of the
    one-ray-per-lane launch."""
    meshes = [scenes.soup(s, seed=20 + i) for i, s in enumerate([64, 3000, 20000, 500, 9000, 1, 0, 40000])]
    offsets = np.zeros(len(meshes) + 1, dtype=np.uint32)
    offsets[1:] = np.cumsum([len(m) for m in meshes])
    tris = np.concatenate(meshes)
    batch = ob.build_cwbvh_batch(ctx, tris, offsets, ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3))
    bt = batch.reorder_triangles(tris)
    rays = scenes.pinhole_rays(1920, 1100, eye=(0.5, 0.5, -1.5), target=(0.5, 0.5, 0.5))
    ray_mesh = (hash_noise(np.arange(len(rays), dtype=np.uint32), 3, 9) * (len(meshes) + 1)).astype(np.uint32)
    print(f"[case] chunked batched launch n={len(rays)}", flush=True)
    saved = ctx.tune_table()
    try:
        outs, misses = [], []
        for chunk in (0, 1):
            t = ctx.tune_table()
            t["trav_chunk"] = [chunk] * 4
            ctx.set_tune_table(t)
            outs.append(batch.ray_traverse(rays, ray_mesh, bt))
            misses.append(batch.ray_traverse_miss(rays, ray_mesh, bt))
    finally:
        ctx.set_tune_table(saved)
    assert outs[0].tobytes() == outs[1].tobytes()
    assert np.array_equal(misses[0], misses[1])
    # a strided sample of the chunked launch against the oracle, mesh by mesh
    nodes, prims = batch.all_nodes(), batch.all_primitive_indices()
    host_bt = bt.to_host()
    sample = np.arange(0, len(rays), 101)
    for i in range(len(meshes)):
        sel = sample[ray_mesh[sample] == i]
        if len(sel) == 0 or len(meshes[i]) == 0:
            continue
        no, nc, po, pc = batch.mesh_info(i)
        mesh_bt = host_bt[offsets[i]:offsets[i + 1]]
        ref = O.traverse(nodes[no:no + nc], mesh_bt, np.ascontiguousarray(rays[sel]), 0)
        assert outs[1][sel].tobytes() == ref.tobytes(), f"mesh {i}"
        refm = O.traverse(nodes[no:no + nc], mesh_bt, np.ascontiguousarray(rays[sel]), 1)
        assert np.array_equal(misses[1][sel], refm), f"mesh {i} occlusion"


def _chunked_vs_pieces(ctx, depth):
    tris, rays = scenes.heightfield_workload(256, 2000, 1920, 1080, seed=4)
    cw = ob.build_cwbvh_from_tris(ctx, tris, ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3))
    bt = cw.reorder_triangles(tris)
    shadow = np.concatenate([rays, rays])  # 4.1M rays: four or more per resident lane on any current device
    # incoherent bounce rays (deep stacks reach the spill columns), repeated to at least two per resident lane
    bounce = scenes.bounce_rays(rays, cw.ray_traverse(ctx.upload(rays), bt), bt.to_host(), seed=5)
    bounce = np.concatenate([bounce] * max(1, -(-(1 << 21) // max(1, len(bounce)))))
    for name, rs in (("primary", rays), ("doubled", shadow), ("bounce", bounce)):
        print(f"[case] chunked launch lds={depth} {name} n={len(rs)}", flush=True)
        whole = cw.ray_traverse(ctx.upload(rs), bt)
        cuts = list(range(0, len(rs), 500000)) + [len(rs)]
        parts = [cw.ray_traverse(ctx.upload(np.ascontiguousarray(rs[a:b])), bt) for a, b in zip(cuts[:-1], cuts[1:])]
        assert np.concatenate(parts).tobytes() == whole.tobytes()
        m_whole = cw.ray_traverse_miss(ctx.upload(rs), bt)
        m_parts = [cw.ray_traverse_miss(ctx.upload(np.ascontiguousarray(rs[a:b])), bt)
                   for a, b in zip(cuts[:-1], cuts[1:])]
        assert np.array_equal(np.concatenate(m_parts), m_whole)
    for rs, lo in ((rays, 0.2), (bounce, 0.01)):
        sample = np.ascontiguousarray(rs[::997])
        ref = O.traverse(cw.nodes, tris[cw.primitive_indices], sample, 0, threads=8)
        got = cw.ray_traverse(ctx.upload(rs), bt)[::997]
        assert got.tobytes() == ref.tobytes()
        assert lo < np.isfinite(got["t"]).mean() < 1.0
