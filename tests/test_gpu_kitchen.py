"""The crate's known-answer render (examples/obj_cwbvh.rs:142-182) on the GPU: the kitchen asset built on the device
with each of the five presets of the crate's test, traced on the device, the hash of the hit normals equal to
1343358762; CWBVH nodes, primitive indices and hits byte-identical to the oracle's.

The kitchen has vertex coordinates of exactly -0.0 and 0.0, so after reinsertion the bytes depend on where the
crate's refit_from_fast walks (bvh2/mod.rs:722-751) stop: they keep the older zero sign above the second box that
compares equal.  The GPU runs those walks as the crate does when a leaf holds a -0."""
import numpy as np
import pytest

import kitchen_scene as K
import obvhs_amd as ob
import oracle_lib as O

pytestmark = pytest.mark.gpu

PRESETS = {
    "fastest": ob.BvhBuildParams.fastest_build,
    "fast": ob.BvhBuildParams.fast_build,
    "medium": ob.BvhBuildParams.medium_build,
    "slow": ob.BvhBuildParams.slow_build,
    "very_slow": ob.BvhBuildParams.very_slow_build,
}


@pytest.mark.parametrize("preset", list(PRESETS))
def test_kitchen_known_answer(ctx, preset):
    print(f"[case] kitchen {preset}", flush=True)
    tris = K.triangles()
    rays, dirs = K.camera_rays(32)
    p = PRESETS[preset]()  # slow and very_slow sort with SortPrecision::U128
    cw = ob.build_cwbvh_from_tris(ctx, tris, p)
    bvh_tris = cw.reorder_triangles(tris).to_host()
    hits = cw.ray_traverse(rays, bvh_tris)
    assert K.render_hash(bvh_tris, hits, dirs) == K.KNOWN_HASH
    _, rcw, rprims = O.build_cwbvh(tris, int(p.ploc_search_distance), p.search_depth_threshold, p.max_prims_per_leaf,
                                   pre_split=p.pre_split, reinsertion_batch_ratio=p.reinsertion_batch_ratio,
                                   sort_precision=int(p.sort_precision))
    assert cw.nodes.shape == rcw.shape
    assert cw.nodes.tobytes() == rcw.tobytes(), f"CWBVH nodes ({preset})"
    np.testing.assert_array_equal(cw.primitive_indices, rprims)
    want = O.traverse(rcw, tris[rprims], rays, 0)
    assert hits.tobytes() == want.tobytes()


def test_kitchen_full_resolution(ctx):
    """The example's default 1280 x 720 render: GPU hits equal the oracle's traversal of the same CWBVH."""
    print("[case] kitchen 1280x720", flush=True)
    tris = K.triangles()
    rays, dirs = K.camera_rays(1280)
    cw = ob.build_cwbvh_from_tris(ctx, tris, ob.BvhBuildParams.fast_build())
    bvh_tris = cw.reorder_triangles(tris).to_host()
    hits = cw.ray_traverse(rays, bvh_tris)
    want = O.traverse(cw.nodes, bvh_tris, rays, 0, threads=8)
    assert hits.tobytes() == want.tobytes()
    assert (hits["primitive_id"] != 0xFFFFFFFF).mean() > 0.9
