"""RtTriangle primitives on the GPU (rt_triangle.rs:147-239, the layout of the crate's demoscene example): the
conversion writes v0, e1 = v0 - v1, e2 = v2 - v0, ng = e1 x e2 with the f32 operations of RtTriangle::new, and
traversal with RtTriangle::intersect gives exactly obvhs_cwbvh_ray_traverse's hits and occlusion flags (the same
arithmetic once e1, e2, ng exist), which equal the oracle's."""
import numpy as np
import pytest

import obvhs_amd as ob
import oracle_lib as O
from obvhs_amd import scenes

pytestmark = pytest.mark.gpu


def rt_fields_numpy(tris):
    v0, v1, v2 = (tris[k][:, :3].astype(np.float32) for k in ("v0", "v1", "v2"))
    e1 = (v0 - v1).astype(np.float32)
    e2 = (v2 - v0).astype(np.float32)
    ng = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1], e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                   e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], axis=1).astype(np.float32)
    out = np.zeros(len(tris), dtype=ob.RT_TRIANGLE)
    out["v0"][:, :3], out["e1"][:, :3], out["e2"][:, :3], out["ng"][:, :3] = v0, e1, e2, ng
    return out


def test_rt_conversion_bits(ctx):
    print("[case] rt triangle conversion", flush=True)
    tris = np.concatenate([scenes.soup(5000, seed=4), scenes.thin_triangles(2000, seed=4)])
    got = ob.rt_triangles(ctx, tris).to_host()
    assert got.tobytes() == rt_fields_numpy(tris).tobytes()


@pytest.mark.parametrize("scene", ["heightfield", "soup", "thin"])
def test_rt_traversal_equals_triangle_traversal(ctx, scene):
    print(f"[case] rt traversal {scene}", flush=True)
    tris = {"heightfield": lambda: scenes.heightfield_scene(128, 500, seed=6),
            "soup": lambda: scenes.soup(30000, seed=7),
            "thin": lambda: np.concatenate([scenes.heightfield_scene(64, 0, seed=8),
                                            scenes.thin_triangles(5000, seed=8)])}[scene]()
    cw = ob.build_cwbvh_from_tris(ctx, tris, ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3))
    bt = cw.reorder_triangles(tris)
    rt = ob.rt_triangles(ctx, bt)
    primary = scenes.pinhole_rays(320, 240, eye=(0.5, 0.6, -0.6), target=(0.5, 0.0, 0.5))
    hits = cw.ray_traverse(primary, bt)
    bounce = scenes.bounce_rays(primary, hits, bt.to_host(), seed=3)
    big = np.concatenate([bounce] * 8)  # enough rays per resident lane for the chunked launch form too
    for name, rays in (("primary", primary), ("bounce", bounce), ("bounce x8", big)):
        for chunk in (0, 1):
            t = ctx.tune_table()
            t["trav_chunk"] = [chunk] * 4
            ctx.set_tune_table(t)
            want = cw.ray_traverse(rays, bt)
            got = cw.ray_traverse_rt(rays, rt)
            assert got.tobytes() == want.tobytes(), (name, chunk)
            np.testing.assert_array_equal(cw.ray_traverse_rt(rays, rt, miss=True), cw.ray_traverse_miss(rays, bt))
    t = ctx.tune_table()
    t["trav_chunk"] = [0] * 4
    ctx.set_tune_table(t)
    sample = np.ascontiguousarray(bounce[::7])
    ref = O.traverse(cw.nodes, bt.to_host(), sample, 0)
    assert cw.ray_traverse_rt(sample, rt).tobytes() == ref.tobytes()
