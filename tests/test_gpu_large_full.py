"""The large scene at its full size (2896 x 2896 heightfield + 500,000 thin triangles = 17,273,632 triangles, pre-splits
capped at 1.3n, reinsertion 0.05 with the ratio sequence [1, 1/3]), where the oracle's build would take minutes, checked
through properties that do not depend on the size:

* two builds of the quality pipeline are identical byte for byte (the kernels' atomics and workgroup order do not
  leak into the result);
* the closest hit's t is a property of the scene, not of the tree: for 16K incoherent rays it equals, bit for bit,
  the t found through a plain build of the same triangles (no pre-splits, no reinsertion), and brute force over all
  17.3M triangles on a sample; occlusion flags of shadow rays agree the same way.  Hits nearer than 1e-3 are left
  out of the tree-to-tree comparison: the node test clamps its entry distance to 1e-4 (cwbvh/node.rs:82), so a
  triangle that close can be culled under one tree's boxes and not under another's."""
import numpy as np
import pytest

import obvhs_amd as ob
import oracle_lib as O
from obvhs_amd import scenes

pytestmark = pytest.mark.gpu


def test_large_scene_full_size_properties(ctx):
    tris, rays, shadow, _ = scenes.large_workload(2896, 500000, 1 << 14, 1 << 14, seed=0)
    print(f"[case] large scene full size n={len(tris)}", flush=True)
    quality = ob.BvhBuildParams(pre_split=True, ploc_search_distance=14, reinsertion_batch_ratio=0.05,
                                max_prims_per_leaf=3)
    cap = int(len(tris) * 1.3)
    seq = [1.0, 1.0 / 3.0]
    d_tris = ctx.upload(tris)
    # the pre-split average is the crate's one-at-a-time f32 sum of the leaf half areas (cwbvh/builder.rs:34-41)
    v = [tris[k][:, :3] for k in ("v0", "v1", "v2")]
    ext = np.maximum(np.maximum(v[0], v[1]), v[2]) - np.minimum(np.minimum(v[0], v[1]), v[2])
    ha = (ext[:, 0] * ext[:, 1] + ext[:, 1] * ext[:, 2]) + ext[:, 2] * ext[:, 0]
    seq_avg = np.add.accumulate(ha, dtype=np.float32)[-1] / np.float32(len(ha))
    _, _, avg, _ = ob.split_aabbs_preset(ctx, d_tris, cap, with_stats=True)
    assert np.float32(avg).tobytes() == np.float32(seq_avg).tobytes(), (avg, seq_avg)
    del v, ext, ha
    a = ob.build_cwbvh_from_tris(ctx, d_tris, quality, split_max_total=cap, ratio_sequence=seq)
    na, pa = a.nodes, a.primitive_indices
    assert len(pa) > len(tris)  # the splits added leaves
    b = ob.build_cwbvh_from_tris(ctx, d_tris, quality, split_max_total=cap, ratio_sequence=seq)
    assert b.nodes.tobytes() == na.tobytes() and b.primitive_indices.tobytes() == pa.tobytes()
    b.close()
    plain = ob.build_cwbvh_from_tris(ctx, d_tris, ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3))
    print("[case] large scene traversal comparisons", flush=True)
    bt_a, bt_p = a.reorder_triangles(d_tris), plain.reorder_triangles(d_tris)
    d_rays, d_shadow = ctx.upload(rays), ctx.upload(shadow)
    ha, hp = a.ray_traverse(d_rays, bt_a), plain.ray_traverse(d_rays, bt_p)
    hit_a, hit_p = ha["primitive_id"] != 0xFFFFFFFF, hp["primitive_id"] != 0xFFFFFFFF
    far = ~((hit_a & (ha["t"] < 1e-3)) | (hit_p & (hp["t"] < 1e-3)))
    assert np.array_equal(hit_a[far], hit_p[far])
    assert np.array_equal(ha["t"][far].view(np.uint32), hp["t"][far].view(np.uint32))
    assert 0.1 < hit_a.mean() < 0.9
    # the shadow rays' nearest occluder (closest hit under the plain tree) tells which flags may depend on the tree
    occ_t = plain.ray_traverse(d_shadow, bt_p)["t"]
    ma, mp = a.ray_traverse_miss(d_shadow, bt_a), plain.ray_traverse_miss(d_shadow, bt_p)
    keep = ~(occ_t < 1e-3)
    assert np.array_equal(ma[keep], mp[keep])
    assert 0.5 < (~ma).mean()
    # brute force over every triangle on a sample of rays (t bits; ids may differ only at exactly equal t)
    sample = np.ascontiguousarray(rays[::683][:24])
    bf = O.brute_force(tris, sample)
    hs = ha[::683][:24]
    ok = ~(bf["t"] < 1e-3)
    assert np.array_equal((hs["primitive_id"] != 0xFFFFFFFF)[ok], (bf["primitive_id"] != 0xFFFFFFFF)[ok])
    hit = ok & (bf["primitive_id"] != 0xFFFFFFFF)
    assert np.array_equal(hs["t"][hit].view(np.uint32), bf["t"][hit].view(np.uint32)) and hit.sum() > 4
