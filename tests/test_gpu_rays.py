"""Ray::new and RayHit::none on the device (obvhs_rays_new, obvhs_ray_hits_none; ray.rs:6-83) against the host
restatement in obvhs_amd.types (make_rays / ray_hit_none, whose safe_inverse is pinned on hand-derived values in
test_oracle_cwbvh.py): the same bytes for directions at +-0, +-EPSILON, one ulp either side of it, +-2 EPSILON,
denormals, infinities and NaN, and for random rays; and the rays trace to the same hits."""
import numpy as np
import pytest

import obvhs_amd as ob
from obvhs_amd import scenes
from obvhs_amd.types import F32_EPS

pytestmark = pytest.mark.gpu


def special_directions():
    eps = np.float32(F32_EPS)
    vals = np.array([0.0, -0.0, eps, -eps, np.nextafter(eps, np.float32(0)), np.nextafter(eps, np.float32(1)),
                     2 * eps, -2 * eps, 1e-40, -1e-40, 3.0, -0.5, np.inf, -np.inf, np.nan, 1e30, -1e-30],
                    dtype=np.float32)
    g = np.stack(np.meshgrid(vals, vals, vals[:6], indexing="ij"), axis=-1).reshape(-1, 3)
    return g.astype(np.float32)


def test_rays_new_equals_host(ctx):
    print("[case] rays_new: special and random directions", flush=True)
    rng = np.random.default_rng(4)
    d = np.concatenate([special_directions(), rng.normal(size=(5000, 3)).astype(np.float32)])
    o = rng.random((len(d), 3)).astype(np.float32)
    for tmin, tmax in ((0.0, np.inf), (1e-3, 1e30), (0.0, np.nan)):
        got = ob.rays_new(ctx, o, d, tmin, tmax).to_host()
        want = ob.make_rays(o, d, tmin, tmax)
        assert got.tobytes() == want.tobytes()
    assert ob.rays_new(ctx, o[:0], d[:0]).count == 0


def test_hits_none_equals_host(ctx):
    for n in (1, 1000, 65537):
        assert ob.ray_hits_none(ctx, n).to_host().tobytes() == ob.ray_hit_none(n).tobytes()


def test_device_rays_trace_like_host_rays(ctx):
    tris = scenes.heightfield_scene(40, 30, seed=5)
    cw = ob.build_cwbvh_from_tris(ctx, tris, ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3))
    bt = cw.reorder_triangles(ctx.upload(tris))
    rng = np.random.default_rng(6)
    o = (rng.random((20000, 3)) * [1.0, 0.6, 1.0] + [0.0, 0.2, 0.0]).astype(np.float32)
    d = rng.normal(size=(20000, 3)).astype(np.float32)
    d[::7, 0] = 0.0  # zero components: safe_inverse's guarded branch
    a = cw.ray_traverse(ob.rays_new(ctx, o, d, 0.0, 1e30), bt)
    b = cw.ray_traverse(ctx.upload(ob.make_rays(o, d, 0.0, 1e30)), bt)
    assert a.tobytes() == b.tobytes()
    assert (a["primitive_id"] != ob.INVALID_ID).mean() > 0.1  # about a sixth of these rays hit
    cw.close()
