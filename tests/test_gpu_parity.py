"""GPU parity: the HIP path against the CPU oracle on identical seeded inputs, bit for bit.

Morton codes and the sort permutation, BVH2 nodes (2n-1 x 48 B), CWBVH nodes (80 B) and primitive indices must be
byte-identical; closest-hit RayHits (primitive_id and t) and occlusion flags must be identical."""
import numpy as np
import pytest

import obvhs_amd as ob
import oracle_lib as O
from obvhs_amd import scenes
from obvhs_amd.types import triangles_from_vertices

pytestmark = pytest.mark.gpu


def gpu_build(ctx, tris, sd=14, sdt=0, max_prims=3, sp=64):
    bvh2 = ob.PlocBuilder(ctx).build(sd, tris, None, ob.SortPrecision(sp), sdt)
    cw = ob.bvh2_to_cwbvh(bvh2, max_prims, True, False)
    return bvh2, cw


def assert_same_bytes(a, b, what):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    av, bv = a.view(np.uint8).reshape(len(a), -1), b.view(np.uint8).reshape(len(b), -1)
    bad = np.nonzero((av != bv).any(axis=1))[0]
    assert len(bad) == 0, f"{what}: {len(bad)} of {len(a)} records differ, first at {bad[:5]}"


def check_build(ctx, tris, sd=14, sdt=0, max_prims=3, sp=64):
    print(f"[case] build n={len(tris)} sd={sd} sdt={sdt} max_prims={max_prims} sort=U{sp}", flush=True)
    aabbs = O.tri_aabbs(tris)
    ref = O.ploc_build(aabbs, sd, sdt, sort_precision=sp)
    bvh2, cw = gpu_build(ctx, tris, sd, sdt, max_prims, sp)
    if len(tris):
        codes, order = bvh2.sort_trace()
        np.testing.assert_array_equal(codes, ref["codes"], err_msg="Morton codes")
        np.testing.assert_array_equal(order, ref["sorted_index"], err_msg="sort permutation")
    assert_same_bytes(bvh2.nodes, ref["nodes"], "BVH2 nodes")
    np.testing.assert_array_equal(bvh2.primitive_indices, ref["prim_indices"])
    assert bvh2.max_depth == max(96, ref["depth"] + 1)
    rcw, rprims, _ = O.bvh2_to_cwbvh(ref["nodes"], ref["prim_indices"], max_prims, True)
    assert_same_bytes(cw.nodes, rcw, "CWBVH nodes")
    np.testing.assert_array_equal(cw.primitive_indices, rprims)
    return bvh2, cw, rcw, rprims


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 17, 100, 1000, 2048, 2049, 4096, 4097, 9000, 40000])
def test_build_soup_sizes(ctx, n):
    check_build(ctx, scenes.soup(n, seed=n))


@pytest.mark.parametrize("n", [1, 2, 3, 100, 2049, 9000, 40000])
@pytest.mark.parametrize("sd", [1, 14])
def test_build_u128(ctx, n, sd):
    """SortPrecision::U128 (ploc/mod.rs:664-683): 126-bit codes, sorted permutation, nodes and CWBVH."""
    check_build(ctx, scenes.soup(n, seed=n + 5), sd, 0, 3, 128)


def test_build_u128_orders_below_the_u64_resolution(ctx):
    """Groups of tiny triangles inside one 2^-21 Morton cell: U64 keys tie (leaf order), U128 keys do not."""
    rng = np.random.default_rng(8)
    cells = rng.random((700, 3)) * 0.9 + 0.05
    c = (np.repeat(cells, 16, axis=0) + rng.random((700 * 16, 3)) * 2.0 ** -24).astype(np.float32)
    c = np.concatenate([c, np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0]], np.float32)])  # the unit scene box
    d = np.float32(2.0 ** -30)
    tris = triangles_from_vertices(c, c + np.array([d, 0, 0], np.float32), c + np.array([0, d, 0], np.float32))
    check_build(ctx, tris, 14, 0, 3, 128)
    a = ob.PlocBuilder(ctx).build(14, tris, None, ob.SortPrecision.U64, 0).sort_trace()[1]
    b = ob.PlocBuilder(ctx).build(14, tris, None, ob.SortPrecision.U128, 0).sort_trace()[1]
    assert a.tolist() != b.tolist()


@pytest.mark.parametrize("sd", [1, 2, 6, 14, 24, 32])
@pytest.mark.parametrize("sdt", [0, 3])
def test_build_search_distances(ctx, sd, sdt):
    check_build(ctx, scenes.soup(20000, seed=7), sd, sdt)


@pytest.mark.parametrize("mp", [1, 2, 3])
def test_build_max_prims(ctx, mp):
    check_build(ctx, scenes.soup(12000, seed=3), 14, 0, mp)


def test_build_plain_workload_and_trace(ctx):
    tris, rays = scenes.plain_workload()
    bvh2, cw, rcw, rprims = check_build(ctx, tris)
    bvh_tris = tris[rprims]
    print("[case] trace closest plain workload", flush=True)
    hits = cw.ray_traverse(rays, bvh_tris)
    ref = O.traverse(rcw, bvh_tris, rays, 0, threads=8)
    assert_same_bytes(hits, ref, "RayHit")
    print("[case] trace miss plain workload", flush=True)
    miss = cw.ray_traverse_miss(rays, bvh_tris)
    np.testing.assert_array_equal(miss, O.traverse(rcw, bvh_tris, rays, 1, threads=8))
    hits2, nv, tv = cw.ray_traverse(rays, bvh_tris, counted=True)
    _, rnv, rtv = O.traverse(rcw, bvh_tris, rays, 0, threads=8, counts=True)
    assert_same_bytes(hits2, ref, "RayHit (counted)")
    np.testing.assert_array_equal(nv, rnv)
    np.testing.assert_array_equal(tv, rtv)


def test_heightfield_build_and_trace(ctx):
    tris, rays = scenes.heightfield_workload(grid=200, boxes=2000, width=320, height=180)
    bvh2, cw, rcw, rprims = check_build(ctx, tris)
    bvh_tris = cw.reorder_triangles(tris)
    np.testing.assert_array_equal(bvh_tris.to_host().view(np.uint8), tris[rprims].view(np.uint8))
    print("[case] trace heightfield primary", flush=True)
    hits = cw.ray_traverse(rays, bvh_tris)
    ref = O.traverse(rcw, tris[rprims], rays, 0, threads=8)
    assert_same_bytes(hits, ref, "RayHit primary")
    b = scenes.bounce_rays(rays, hits, tris[rprims], seed=1)
    print(f"[case] trace heightfield bounce n={len(b)}", flush=True)
    h2 = cw.ray_traverse(b, bvh_tris)
    assert_same_bytes(h2, O.traverse(rcw, tris[rprims], b, 0, threads=8), "RayHit bounce")
    m = cw.ray_traverse_miss(b, bvh_tris)
    np.testing.assert_array_equal(m, O.traverse(rcw, tris[rprims], b, 1, threads=8))


def test_empty_build(ctx):
    tris = scenes.soup(0)
    bvh2, cw = gpu_build(ctx, tris)
    assert bvh2.node_count == 0 and cw.node_count == 0
    rays = ob.make_rays([[0, 0, 1]], [[0, 0, -1]])
    hits = cw.ray_traverse(rays, np.zeros(1, dtype=ob.TRIANGLE))
    assert hits["primitive_id"][0] == ob.INVALID_ID and np.isinf(hits["t"][0])


@pytest.mark.parametrize("n,sd,sdt,mp", [(1, 14, 0, 3), (2, 14, 0, 3), (3, 1, 0, 1), (100, 6, 2, 2), (5000, 14, 3, 3),
                                         (40000, 14, 0, 3), (40000, 32, 0, 1), (70000, 2, 0, 3), (70000, 1, 0, 2)])
def test_build_cwbvh_from_tris_fused(ctx, n, sd, sdt, mp):
    """build_cwbvh_from_tris evaluates the conversion's cost table inside the PLOC merges; same bytes."""
    print(f"[case] fused build n={n} sd={sd} sdt={sdt} mp={mp}", flush=True)
    tris = scenes.soup(n, seed=n + sd)
    params = ob.BvhBuildParams(ploc_search_distance=sd, search_depth_threshold=sdt, max_prims_per_leaf=mp)
    cw = ob.build_cwbvh_from_tris(ctx, tris, params)
    _, rcw, rprims = O.build_cwbvh(tris, sd, sdt, mp)
    assert_same_bytes(cw.nodes, rcw, "CWBVH nodes (fused)")
    np.testing.assert_array_equal(cw.primitive_indices, rprims)


def test_build_reuse_handles(ctx):
    """build_with_bvh-style reuse (ploc/mod.rs:129-159): a handle rebuilt with other inputs equals a fresh build."""
    params = ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3)
    cw = None
    for n, seed in ((30000, 1), (9000, 2), (50000, 3)):
        print(f"[case] reuse n={n}", flush=True)
        tris = scenes.soup(n, seed=seed)
        cw = ob.build_cwbvh_from_tris(ctx, tris, params, reuse=cw)
        _, rcw, rprims = O.build_cwbvh(tris, 14, 0, 3)
        assert_same_bytes(cw.nodes, rcw, "CWBVH nodes (reuse)")
        np.testing.assert_array_equal(cw.primitive_indices, rprims)


def test_conversion_of_same_bvh2_with_other_leaf_sizes(ctx):
    """One BVH2 converted with max_prims_per_leaf 3, then 1, then 3 again (cost table re-evaluated each time)."""
    tris = scenes.soup(25000, seed=9)
    bvh2 = ob.PlocBuilder(ctx).build(14, tris)
    ref = O.ploc_build(O.tri_aabbs(tris), 14, 0)
    for mp in (3, 1, 3, 2):
        print(f"[case] reconvert mp={mp}", flush=True)
        cw = ob.bvh2_to_cwbvh(bvh2, mp, True, False)
        rcw, rprims, _ = O.bvh2_to_cwbvh(ref["nodes"], ref["prim_indices"], mp, True)
        assert_same_bytes(cw.nodes, rcw, f"CWBVH nodes mp={mp}")


def test_full_heightfield_parity(ctx):
    """The benchmark's scene at full size: 1,288,352 triangles, bit-identical CWBVH; a strided ray sample traced."""
    tris, rays = scenes.heightfield_workload()
    print(f"[case] full heightfield n={len(tris)}", flush=True)
    params = ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3)
    cw = ob.build_cwbvh_from_tris(ctx, tris, params)
    b, rcw, rprims = O.build_cwbvh(tris, 14, 0, 3)
    assert_same_bytes(cw.nodes, rcw, "CWBVH nodes")
    np.testing.assert_array_equal(cw.primitive_indices, rprims)
    sample = rays[::7]
    bvh_tris = cw.reorder_triangles(tris)
    hits = cw.ray_traverse(sample, bvh_tris)
    ref = O.traverse(rcw, tris[rprims], sample, 0, threads=16)
    assert_same_bytes(hits, ref, "RayHit heightfield sample")
