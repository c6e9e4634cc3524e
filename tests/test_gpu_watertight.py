"""The watertight traversal mode on the GPU (obvhs_cwbvh_ray_traverse_watertight / _miss_watertight; an extra mode
beside the crate's Triangle::intersect): hits (id and t bits) and occlusion flags equal the oracle's watertight
traversal of the same CWBVH, and on a closed grid mesh every ray straight down through a vertex and every ray
through a shared edge's midpoint hits, where the crate's test lets some through."""
import numpy as np
import pytest

import obvhs_amd as ob
import oracle_lib as O
from obvhs_amd import scenes
from test_oracle_watertight import grid_mesh, watertight_rays

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scene", ["heightfield", "soup", "grid"])
def test_watertight_traversal_matches_oracle(ctx, scene):
    print(f"[case] watertight traversal {scene}", flush=True)
    tris = {"heightfield": lambda: scenes.heightfield_scene(96, 300, seed=5),
            "soup": lambda: scenes.soup(20000, seed=6),
            "grid": lambda: grid_mesh(64, 2)[0]}[scene]()
    cw = ob.build_cwbvh_from_tris(ctx, tris, ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3))
    bt = cw.reorder_triangles(tris)
    primary = scenes.pinhole_rays(256, 192, eye=(0.5, 0.6, -0.6), target=(0.5, 0.0, 0.5))
    hits = cw.ray_traverse(primary, bt)
    bounce = scenes.bounce_rays(primary, hits, bt.to_host(), seed=3)
    host_bt = bt.to_host()
    for name, rays in (("primary", primary), ("bounce", bounce)):
        got = cw.ray_traverse_watertight(rays, bt)
        want = O.traverse(cw.nodes, host_bt, rays, 2, threads=8)
        assert got.tobytes() == want.tobytes(), name
        got_m = cw.ray_traverse_watertight(rays, bt, miss=True)
        want_m = O.traverse(cw.nodes, host_bt, rays, 3, threads=8)
        np.testing.assert_array_equal(got_m, want_m, err_msg=name)
        np.testing.assert_array_equal(got_m, got["primitive_id"] == ob.INVALID_ID, err_msg=name)


def test_watertight_on_closed_mesh_gpu(ctx):
    print("[case] watertight closed grid", flush=True)
    tris, down, edge = watertight_rays(64, 3)
    cw = ob.build_cwbvh_from_tris(ctx, tris, ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3))
    bt = cw.reorder_triangles(tris)
    crate_missed = 0
    for rays in (down, edge):
        wt = cw.ray_traverse_watertight(rays, bt)
        assert (wt["primitive_id"] != ob.INVALID_ID).all()
        assert not cw.ray_traverse_watertight(rays, bt, miss=True).any()
        crate_missed += int((cw.ray_traverse(rays, bt)["primitive_id"] == ob.INVALID_ID).sum())
    assert crate_missed > 0
