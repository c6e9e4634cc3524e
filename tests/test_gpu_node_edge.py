"""GPU traversal on rays whose node-test slab terms are NaN or infinite (scenes.slab_edge_rays: zero and tiny
direction components with origins on node planes, infinite and NaN inverse directions, overflowing origins, NaN / 0 /
negative ray tmax) equals the oracle byte for byte: closest hit and occlusion (CwBvh::ray_traverse and
ray_traverse_miss, cwbvh/mod.rs:169-225), the watertight mode, the per-ray node and triangle test counts, and both
launch forms (one lane per ray, and chunked rows over about a million rays, the unique rays tiled) at LDS stack
depths 4 and 8.  The GPU node test uses IEEE max3/min3 with the crate's NaN rule restated; a planted IEEE node test
in the oracle changes thousands of these rays' visit counts (test_oracle_node_edge.py)."""
import numpy as np
import pytest

import obvhs_amd as ob
import oracle_lib as O
from obvhs_amd import scenes

pytestmark = pytest.mark.gpu

PARAMS = ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3)
UNIQUE = 16384
TILES = 64  # 1M rays: chunked when the tuning table asks for it


@pytest.fixture(scope="module")
def edge(ctx):
    tris = np.concatenate([scenes.heightfield_scene(24, 20, seed=2), scenes.signed_zero_soup(300, seed=1)])
    _, ocw, prims = O.build_cwbvh(tris, 14, 0, 3)
    cw = ob.build_cwbvh_from_tris(ctx, tris, PARAMS)
    assert cw.nodes.tobytes() == ocw.tobytes()
    bt = cw.reorder_triangles(ctx.upload(tris))
    rays = scenes.slab_edge_rays(tris, ocw["p"], UNIQUE, seed=9)
    obt = tris[prims]
    ref = {m: O.traverse(ocw, obt, rays, m) for m in (0, 1, 2, 3)}
    ref["counts"] = O.traverse(ocw, obt, rays, 0, counts=True)
    yield ctx, cw, bt, rays, ref, ob.size_class(len(tris))
    cw.close()


def test_counts_and_hits_one_lane_per_ray(edge):
    ctx, cw, bt, rays, ref, _ = edge
    print(f"[case] slab edge rays: {len(rays)} counted", flush=True)
    hits, nv, tv = cw.ray_traverse(ctx.upload(rays), bt, counted=True)
    rh, rnv, rtv = ref["counts"]
    assert hits.tobytes() == rh.tobytes()
    np.testing.assert_array_equal(nv, rnv)
    np.testing.assert_array_equal(tv, rtv)


@pytest.mark.parametrize("chunk,depth", [(0, 8), (1, 4), (1, 8)])
def test_all_modes_both_launch_forms(edge, chunk, depth):
    ctx, cw, bt, rays, ref, cls = edge
    saved = ctx.tune_table()
    t = ctx.tune_table()
    t["trav_chunk"][cls], t["trav_lds_depth"][cls] = chunk, depth
    ctx.set_tune_table(t)
    try:
        tiled = np.tile(rays, TILES)
        d_rays = ctx.upload(tiled)
        print(f"[case] slab edge rays x{TILES}: chunk={chunk} lds={depth}", flush=True)
        got = cw.ray_traverse(d_rays, bt)
        assert got.tobytes() == np.tile(ref[0], TILES).tobytes()
        np.testing.assert_array_equal(cw.ray_traverse_miss(d_rays, bt), np.tile(ref[1], TILES))
        got = cw.ray_traverse_watertight(d_rays, bt)
        assert got.tobytes() == np.tile(ref[2], TILES).tobytes()
        np.testing.assert_array_equal(cw.ray_traverse_watertight(d_rays, bt, miss=True), np.tile(ref[3], TILES))
    finally:
        ctx.set_tune_table(saved)
