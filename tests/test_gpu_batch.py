"""GPU parity of the batched form: many meshes built and traced in shared launches.

Each mesh's CWBVH nodes and primitive indices must be byte-identical to the oracle's build_cwbvh of that mesh alone
(cwbvh/builder.rs:20-85 with the same parameters), and batched closest-hit / occlusion results must equal the
oracle's traversal of that mesh's own CWBVH."""
import numpy as np
import pytest

import obvhs_amd as ob
import oracle_lib as O
from obvhs_amd import scenes

pytestmark = pytest.mark.gpu


def assert_same_bytes(a, b, what):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    if len(a) == 0:
        return
    av, bv = a.view(np.uint8).reshape(len(a), -1), b.view(np.uint8).reshape(len(b), -1)
    bad = np.nonzero((av != bv).any(axis=1))[0]
    assert len(bad) == 0, f"{what}: {len(bad)} of {len(a)} records differ, first at {bad[:5]}"


def concat(meshes):
    offsets = np.zeros(len(meshes) + 1, dtype=np.uint32)
    offsets[1:] = np.cumsum([len(m) for m in meshes])
    tris = np.concatenate(meshes) if offsets[-1] else np.zeros(0, dtype=ob.TRIANGLE)
    return tris, offsets


def mesh_slices(batch, nodes, prims, i):
    no, nc, po, pc = batch.mesh_info(i)
    return nodes[no:no + nc], prims[po:po + pc]


def check_batch_build(ctx, meshes, sd=14, sdt=0, mp=3):
    tris, offsets = concat(meshes)
    params = ob.BvhBuildParams(ploc_search_distance=sd, search_depth_threshold=sdt, max_prims_per_leaf=mp)
    batch = ob.build_cwbvh_batch(ctx, tris, offsets, params)
    info = batch.info()
    assert info["mesh_count"] == len(meshes) and info["total_prims"] == len(tris)
    nodes, prims = batch.all_nodes(), batch.all_primitive_indices()
    for i, m in enumerate(meshes):
        n_i, p_i = mesh_slices(batch, nodes, prims, i)
        _, rcw, rprims = O.build_cwbvh(m, sd, sdt, mp)
        assert_same_bytes(n_i, rcw, f"mesh {i} ({len(m)} tris) CWBVH nodes")
        np.testing.assert_array_equal(p_i, rprims, err_msg=f"mesh {i} primitive indices")
    return batch, tris, offsets, nodes, prims


MIXED_SIZES = [0, 1, 2, 3, 5, 17, 100, 0, 640, 1, 2000, 4096, 4097, 33, 9000]


@pytest.mark.parametrize("sd,sdt,mp", [(14, 0, 3), (1, 0, 1), (2, 0, 2), (6, 2, 3), (24, 3, 1), (32, 0, 3)])
def test_batch_soups_match_single_builds(ctx, sd, sdt, mp):
    print(f"[case] batch soups sd={sd} sdt={sdt} mp={mp}", flush=True)
    meshes = [scenes.soup(n, seed=100 + i) for i, n in enumerate(MIXED_SIZES)]
    check_batch_build(ctx, meshes, sd, sdt, mp)


def test_batch_single_mesh_and_all_empty(ctx):
    print("[case] batch single mesh / empty meshes", flush=True)
    check_batch_build(ctx, [scenes.soup(30000, seed=3)])
    batch, *_ = check_batch_build(ctx, [scenes.soup(0), scenes.soup(0)])
    assert batch.info()["total_nodes"] == 0


def test_batch_many_tiny_meshes(ctx):
    """Hundreds of segments, most finishing in the first passes while a large one keeps merging."""
    print("[case] batch tiny meshes", flush=True)
    sizes = [1 + (i * 7) % 13 for i in range(300)] + [20000]
    meshes = [scenes.soup(n, seed=i) for i, n in enumerate(sizes)]
    check_batch_build(ctx, meshes, 14, 0, 3)


def test_batch_icospheres_trace(ctx):
    print("[case] batch icospheres build + trace", flush=True)
    levels = np.array([0, 1, 2, 3, 4, 0, 2, 3, 1, 4, 3, 2])
    tris, offsets, rays, ray_mesh = scenes.icosphere_batch(len(levels), seed=5, rays_per_mesh=512, levels=levels)
    meshes = [tris[offsets[i]:offsets[i + 1]] for i in range(len(levels))]
    batch, tris, offsets, nodes, prims = check_batch_build(ctx, meshes, 14, 0, 3)
    bvh_tris = batch.reorder_triangles(tris)
    hits = batch.ray_traverse(rays, ray_mesh, bvh_tris)
    miss = batch.ray_traverse_miss(rays, ray_mesh, bvh_tris)
    host_bvh_tris = bvh_tris.to_host()
    total_hits = 0
    for i in range(len(levels)):
        n_i, p_i = mesh_slices(batch, nodes, prims, i)
        bt = host_bvh_tris[offsets[i]:offsets[i + 1]]
        np.testing.assert_array_equal(bt.view(np.uint8), meshes[i][p_i].view(np.uint8))
        sel = ray_mesh == i
        ref = O.traverse(n_i, bt, rays[sel], mode=0)
        np.testing.assert_array_equal(hits["primitive_id"][sel], ref["primitive_id"], err_msg=f"mesh {i} ids")
        np.testing.assert_array_equal(hits["t"][sel].view(np.uint32), ref["t"].view(np.uint32),
                                      err_msg=f"mesh {i} t")
        refm = O.traverse(n_i, bt, rays[sel], mode=1)
        np.testing.assert_array_equal(miss[sel], refm, err_msg=f"mesh {i} miss")
        total_hits += int(np.sum(ref["primitive_id"] != ob.INVALID_ID))
    assert total_hits > len(rays) // 2  # rays aim into each mesh's AABB: most hit


def test_batch_equals_gpu_single_builds_and_traversals(ctx):
    """Batched results equal the single-mesh GPU path mesh by mesh (larger meshes than the oracle checks above)."""
    print("[case] batch vs single GPU builds", flush=True)
    levels = scenes.icosphere_batch_counts(96, seed=11)
    tris, offsets, rays, ray_mesh = scenes.icosphere_batch(96, seed=11, rays_per_mesh=256, levels=levels)
    params = ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3)
    batch = ob.build_cwbvh_batch(ctx, tris, offsets, params)
    nodes, prims = batch.all_nodes(), batch.all_primitive_indices()
    bvh_tris = batch.reorder_triangles(tris)
    hits = batch.ray_traverse(rays, ray_mesh, bvh_tris)
    for i in range(96):
        m = tris[offsets[i]:offsets[i + 1]]
        cw = ob.build_cwbvh_from_tris(ctx, m, params)
        n_i, p_i = mesh_slices(batch, nodes, prims, i)
        assert_same_bytes(n_i, cw.nodes, f"mesh {i} nodes")
        np.testing.assert_array_equal(p_i, cw.primitive_indices)
        sel = ray_mesh == i
        h = cw.ray_traverse(rays[sel], cw.reorder_triangles(m))
        np.testing.assert_array_equal(hits["primitive_id"][sel], h["primitive_id"])
        np.testing.assert_array_equal(hits["t"][sel].view(np.uint32), h["t"].view(np.uint32))


def test_batch_reuse_and_out_of_range_mesh_ids(ctx):
    print("[case] batch reuse", flush=True)
    params = ob.BvhBuildParams(ploc_search_distance=6, max_prims_per_leaf=2)
    batch = None
    for k, sizes in enumerate(([5000, 10, 300], [7, 9000], [1, 2, 3, 4, 50000])):
        meshes = [scenes.soup(n, seed=k * 10 + i) for i, n in enumerate(sizes)]
        tris, offsets = concat(meshes)
        batch = ob.build_cwbvh_batch(ctx, tris, offsets, params, reuse=batch)
        nodes, prims = batch.all_nodes(), batch.all_primitive_indices()
        for i, m in enumerate(meshes):
            n_i, p_i = mesh_slices(batch, nodes, prims, i)
            _, rcw, rprims = O.build_cwbvh(m, 6, 0, 2)
            assert_same_bytes(n_i, rcw, f"reuse {k} mesh {i}")
            np.testing.assert_array_equal(p_i, rprims)
    t0 = meshes[4][0]
    c = (t0["v0"][:3].astype(np.float64) + t0["v1"][:3] + t0["v2"][:3]) / 3.0
    rays = ob.make_rays([[c[0], c[1], -1.0]] * 3, [[0, 0, 1]] * 3)  # aimed at a triangle of mesh 4
    ray_mesh = np.array([4, 5, 1 << 31], dtype=np.uint32)  # 5 and 2^31 are not meshes: they miss
    hits = batch.ray_traverse(rays, ray_mesh, batch.reorder_triangles(tris))
    assert hits["primitive_id"][0] != ob.INVALID_ID
    assert (hits["primitive_id"][1:] == ob.INVALID_ID).all() and np.isinf(hits["t"][1:]).all()


def test_batch_invalid_arguments(ctx):
    tris, offsets = concat([scenes.soup(10), scenes.soup(10)])
    with pytest.raises(ob.ObvhsError) as e:
        ob.build_cwbvh_batch(ctx, tris, offsets,
                             ob.BvhBuildParams(ploc_search_distance=14, reinsertion_batch_ratio=0.1))
    assert e.value.name == "OBVHS_ERR_INVALID_ARG"
    with pytest.raises(ValueError):
        ob.build_cwbvh_batch(ctx, tris, offsets[:-1], ob.BvhBuildParams())
    # a non-empty batch traced without its triangles: rejected before any launch
    batch = ob.build_cwbvh_batch(ctx, tris, offsets, ob.BvhBuildParams(ploc_search_distance=14))
    rays = ctx.upload(ob.make_rays([[0.5, 0.5, -1.0]] * 4, [[0, 0, 1]] * 4))
    mesh = ctx.upload(np.zeros(4, dtype=np.uint32))
    out = ob.DeviceArray(ctx, ob.RAY_HIT, 4)
    miss = ob.DeviceArray(ctx, np.uint8, 4)
    lib = ob.native.lib()
    bad = ob.native.STATUS_CODE["OBVHS_ERR_INVALID_ARG"]
    assert lib.obvhs_batch_ray_traverse(ctx.handle, batch.handle, None, rays.ptr, mesh.ptr, 4, out.ptr) == bad
    assert lib.obvhs_batch_ray_traverse_miss(ctx.handle, batch.handle, None, rays.ptr, mesh.ptr, 4, miss.ptr) == bad


def test_batch_full_icosphere_workload_sampled(ctx):
    """4,096 icospheres (log-uniform 64..65,536 triangles, ~20 M in total): every mesh's node count and root
    bounds are checked for structure, and a sample of meshes (including the largest) is compared with single-mesh
    GPU builds and traversals."""
    print("[case] batch full icosphere workload", flush=True)
    tris, offsets, rays, ray_mesh = scenes.icosphere_batch(4096, seed=0, rays_per_mesh=1024)
    params = ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3)
    dtris = ctx.upload(tris)
    batch = ob.build_cwbvh_batch(ctx, dtris, offsets, params)
    bvh_tris = batch.reorder_triangles(dtris)
    hits = batch.ray_traverse(rays, ray_mesh, bvh_tris)
    nodes = batch.all_nodes()
    prims = batch.all_primitive_indices()
    counts = np.diff(offsets.astype(np.int64))
    for i in range(4096):  # every mesh: a permutation of its primitives
        no, nc, po, pc = batch.mesh_info(i)
        assert pc == counts[i] and po == offsets[i] and nc >= 1
    sample = sorted(set(np.linspace(0, 4095, 12).astype(int)) | {int(np.argmax(counts))})
    for i in sample:
        m = tris[offsets[i]:offsets[i + 1]]
        cw = ob.build_cwbvh_from_tris(ctx, m, params)
        n_i, p_i = mesh_slices(batch, nodes, prims, i)
        assert_same_bytes(n_i, cw.nodes, f"mesh {i} nodes")
        np.testing.assert_array_equal(p_i, cw.primitive_indices)
        sel = ray_mesh == i
        h = cw.ray_traverse(rays[sel], cw.reorder_triangles(m))
        np.testing.assert_array_equal(hits["primitive_id"][sel], h["primitive_id"])
        np.testing.assert_array_equal(hits["t"][sel].view(np.uint32), h["t"].view(np.uint32))
    perm_ok = np.zeros(len(prims), dtype=bool)
    seg = np.repeat(np.arange(4096), counts)
    perm_ok[offsets[seg] + prims] = True
    assert perm_ok.all()


@pytest.mark.parametrize("tlas_mp", [1, 3])
def test_tlas_over_batch_matches_oracle(ctx, tlas_mp):
    """CwBvh::ray_traverse_tlas_blas (cwbvh/mod.rs:341-491) with the batch's meshes as BLASes."""
    print(f"[case] tlas mp={tlas_mp}", flush=True)
    rng = np.random.default_rng(8)
    meshes = []
    for i in range(24):
        t = scenes.soup(int(rng.integers(50, 3000)), seed=200 + i)
        off = rng.uniform(-1.5, 1.5, 3).astype(np.float32)  # overlapping bounds: several leaves per node get hit
        for v in ("v0", "v1", "v2"):
            t[v][:, :3] += off
        meshes.append(t)
    tris, offsets = concat(meshes)
    params = ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3)
    batch = ob.build_cwbvh_batch(ctx, tris, offsets, params)
    tlas = batch.build_tlas(ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=tlas_mp))
    roots = []
    blas = []
    nodes, prims = batch.all_nodes(), batch.all_primitive_indices()
    for i, m in enumerate(meshes):
        b, rcw, rprims = O.build_cwbvh(m, 14, 0, 3)
        roots.append(b["nodes"][0])
        blas.append((rcw, m[rprims]))
    boxes = ob.aabbs_from_minmax(np.array([r["min"][:3] for r in roots]), np.array([r["max"][:3] for r in roots]))
    t = O.ploc_build(boxes, 14, 0)
    rtlas, rtprims, _ = O.bvh2_to_cwbvh(t["nodes"], t["prim_indices"], tlas_mp, True)
    assert_same_bytes(tlas.nodes, rtlas, "TLAS nodes")
    np.testing.assert_array_equal(tlas.primitive_indices, rtprims)
    n = 20000
    o = rng.uniform(-3, 4, (n, 3))
    pick = rng.integers(0, len(tris), n)  # aimed at triangle centroids of random meshes
    c = (tris["v0"][pick, :3].astype(np.float64) + tris["v1"][pick, :3] + tris["v2"][pick, :3]) / 3.0
    d = c - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = ob.make_rays(o.astype(np.float32), d.astype(np.float32), 0.0, 1e30)
    hits = batch.tlas_ray_traverse(tlas, rays, batch.reorder_triangles(tris))
    ref = O.tlas_traverse(rtlas, rtprims, blas, rays)
    np.testing.assert_array_equal(hits["primitive_id"], ref["primitive_id"])
    np.testing.assert_array_equal(hits["geometry_id"], ref["geometry_id"])
    np.testing.assert_array_equal(hits["t"].view(np.uint32), ref["t"].view(np.uint32))
    assert np.isfinite(ref["t"]).mean() > 0.05
