"""GPU parity of the build-quality passes and refit against the oracle, bit for bit.

* reinsertion (ReinsertionOptimizer::run, bvh2/reinsertion.rs:40-204): BVH2 nodes after the optimizer;
* spatial pre-splits (cwbvh/builder.rs:26-55, splits.rs): the split leaf boxes and indices;
* build_cwbvh_from_tris with the crate's presets (pre-splits, reinsertion): CWBVH nodes and primitive indices;
* refit (bvh2/mod.rs:506-569) after a seeded vertex jitter, and the CWBVH re-emitted from the refit BVH2."""
import numpy as np
import pytest

import obvhs_amd as ob
import oracle_lib as O
from obvhs_amd import scenes, test_util

pytestmark = pytest.mark.gpu


def same(a, b, what):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    if a.size == 0:
        return
    av, bv = a.view(np.uint8).reshape(len(a), -1), b.view(np.uint8).reshape(len(b), -1)
    bad = np.nonzero((av != bv).any(axis=1))[0]
    assert len(bad) == 0, f"{what}: {len(bad)} of {len(a)} records differ, first at {bad[:5]}"


def hf_scene(grid=64, boxes=200, seed=0):
    return scenes.heightfield_scene(grid, boxes, seed=seed)


SCENES = {
    "soup5k": lambda: scenes.soup(5000, seed=3),
    "heightfield": lambda: hf_scene(64, 200, 1),
    "thin": lambda: np.concatenate([hf_scene(40, 50, 2), scenes.thin_triangles(400, seed=2)]),
    # exact +0 and -0 coordinates: refit_from_fast's early exit decides which zero bits a box keeps
    "signed_zero": lambda: scenes.signed_zero_soup(3000, seed=1),
}


@pytest.mark.parametrize("scene", list(SCENES))
@pytest.mark.parametrize("ratio,seq", [(0.01, None), (0.05, [1.0, 1.0 / 3.0]), (0.2, None), (1.0, [1.0])])
def test_reinsertion_matches_oracle(ctx, scene, ratio, seq):
    print(f"[case] reinsertion {scene} ratio={ratio} seq={seq}", flush=True)
    tris = SCENES[scene]()
    bvh = ob.PlocBuilder(ctx).build(14, tris)
    ref = O.ploc_build(O.tri_aabbs(tris), 14, 0)
    same(bvh.nodes, ref["nodes"], "BVH2 before reinsertion")
    bvh.reinsertion(ratio, seq)
    expect = O.reinsertion(ref["nodes"], ratio, seq)
    same(bvh.nodes, expect, "BVH2 after reinsertion")
    cw = ob.bvh2_to_cwbvh(bvh, 3, True, False)
    rcw, rprims, _ = O.bvh2_to_cwbvh(expect, ref["prim_indices"], 3, True)
    same(cw.nodes, rcw, "CWBVH of the reinserted BVH2")
    np.testing.assert_array_equal(cw.primitive_indices, rprims)


@pytest.mark.parametrize("max_total", [0, 1])
def test_presplit_matches_oracle(ctx, max_total):
    tris = SCENES["thin"]()
    cap = 0 if max_total == 0 else int(len(tris) * 1.05)
    print(f"[case] presplit n={len(tris)} cap={cap}", flush=True)
    aabbs, idx = ob.split_aabbs_preset(ctx, tris, cap)
    ra, ri, _, _ = O.presplit(tris, cap)
    assert len(ra) > len(tris)
    if cap:
        assert len(ra) == cap
    same(aabbs, ra, "split leaf boxes")
    np.testing.assert_array_equal(idx, ri)


PRESETS = {
    "very_fast": ob.BvhBuildParams.very_fast_build,
    "fast": ob.BvhBuildParams.fast_build,
    "medium": ob.BvhBuildParams.medium_build,
    "slow": ob.BvhBuildParams.slow_build,
    "very_slow": ob.BvhBuildParams.very_slow_build,
}


@pytest.mark.parametrize("preset", list(PRESETS))
@pytest.mark.parametrize("scene", ["soup5k", "thin"])
def test_presets_match_oracle(ctx, preset, scene):
    """The crate's presets (lib.rs:243-306), each with its own SortPrecision (U128 for slow and very_slow)."""
    print(f"[case] preset {preset} {scene}", flush=True)
    tris = SCENES[scene]()
    p = PRESETS[preset]()
    cw = ob.build_cwbvh_from_tris(ctx, tris, p)
    _, rcw, rprims = O.build_cwbvh(tris, int(p.ploc_search_distance), p.search_depth_threshold,
                                   p.max_prims_per_leaf, pre_split=p.pre_split,
                                   reinsertion_batch_ratio=p.reinsertion_batch_ratio,
                                   sort_precision=int(p.sort_precision))
    same(cw.nodes, rcw, f"CWBVH nodes ({preset})")
    np.testing.assert_array_equal(cw.primitive_indices, rprims)
    rays = scenes.pinhole_rays(64, 64, eye=(0.5, 0.45, -0.35), target=(0.5, 0.0, 0.55))
    h = cw.ray_traverse(rays, cw.reorder_triangles(tris))
    r = O.traverse(rcw, tris[rprims], rays)
    np.testing.assert_array_equal(h["primitive_id"], r["primitive_id"])
    np.testing.assert_array_equal(h["t"].view(np.uint32), r["t"].view(np.uint32))


def test_split_budget_and_ratio_sequence(ctx):
    tris = SCENES["thin"]()
    p = ob.BvhBuildParams(pre_split=True, ploc_search_distance=14, reinsertion_batch_ratio=0.05,
                          max_prims_per_leaf=3)
    cap = int(len(tris) * 1.3)
    print(f"[case] split budget {cap} + ratio sequence", flush=True)
    cw = ob.build_cwbvh_from_tris(ctx, tris, p, split_max_total=cap, ratio_sequence=[1.0, 1.0 / 3.0])
    _, rcw, rprims = O.build_cwbvh(tris, 14, 0, 3, pre_split=True, reinsertion_batch_ratio=0.05,
                                   ratio_sequence=[1.0, 1.0 / 3.0], split_max_total=cap)
    same(cw.nodes, rcw, "CWBVH nodes (budget)")
    np.testing.assert_array_equal(cw.primitive_indices, rprims)


def jitter(tris, seed, amount):
    j = tris.copy()
    i = np.arange(len(j), dtype=np.uint32)
    for k, name in enumerate(("v0", "v1", "v2")):
        for a in range(3):
            j[name][:, a] += (test_util.hash_noise(i, 50 + 3 * k + a, seed) - np.float32(0.5)) * np.float32(amount)
    return j


@pytest.mark.parametrize("reinsert", [False, True])
def test_refit_after_jitter_matches_oracle(ctx, reinsert):
    print(f"[case] refit reinsert={reinsert}", flush=True)
    tris = SCENES["heightfield"]()
    p = ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3,
                          reinsertion_batch_ratio=0.05 if reinsert else 0.0)
    cw = ob.build_cwbvh_from_tris(ctx, tris, p)
    b2 = cw.bvh2
    ref_b, _, _ = O.build_cwbvh(tris, 14, 0, 3, reinsertion_batch_ratio=p.reinsertion_batch_ratio)
    same(b2.nodes, ref_b["nodes"], "BVH2 before refit")
    moved = jitter(tris, 5, 0.01)
    b2.refit_from_triangles(moved)
    expect = ref_b["nodes"].copy()
    leaves = np.nonzero(expect["prim_count"] != 0)[0]
    prim = ref_b["prim_indices"][expect["first_index"][leaves]]
    ja = O.tri_aabbs(moved)
    expect["min"][leaves] = ja["min"][prim]
    expect["max"][leaves] = ja["max"][prim]
    expect = O.refit_all(expect)
    same(b2.nodes, expect, "BVH2 after refit")
    re = ob.bvh2_to_cwbvh(b2, 3, True, False)
    rcw, rprims, _ = O.bvh2_to_cwbvh(expect, ref_b["prim_indices"], 3, True)
    same(re.nodes, rcw, "CWBVH re-emitted after refit")
    np.testing.assert_array_equal(re.primitive_indices, rprims)
    rays = scenes.pinhole_rays(64, 64, eye=(0.5, 0.45, -0.35), target=(0.5, 0.0, 0.55))
    h = re.ray_traverse(rays, re.reorder_triangles(moved))
    bf = O.brute_force(moved, rays)
    np.testing.assert_array_equal(h["t"].view(np.uint32), bf["t"].view(np.uint32))


def test_large_pipeline_scaled(ctx):
    """The large-scene pipeline at 1/64 of its size (362 x 362 heightfield + 7,812 thin triangles): pre-splits capped
    at 1.3n, reinsertion 0.05 with the sequence [1, 1/3], closest-hit and shadow rays, all against the oracle."""
    tris, rays, shadow, _ = scenes.large_workload(362, 7812, 1 << 16, 1 << 16, seed=0)
    print(f"[case] scaled large pipeline n={len(tris)}", flush=True)
    p = ob.BvhBuildParams(pre_split=True, ploc_search_distance=14, reinsertion_batch_ratio=0.05,
                          max_prims_per_leaf=3)
    cap = int(len(tris) * 1.3)
    seq = [1.0, 1.0 / 3.0]
    cw = ob.build_cwbvh_from_tris(ctx, tris, p, split_max_total=cap, ratio_sequence=seq)
    _, rcw, rprims = O.build_cwbvh(tris, 14, 0, 3, pre_split=True, reinsertion_batch_ratio=0.05,
                                   ratio_sequence=seq, split_max_total=cap)
    assert len(rprims) > len(tris)
    same(cw.nodes, rcw, "CWBVH nodes (scaled large)")
    np.testing.assert_array_equal(cw.primitive_indices, rprims)
    bt = cw.reorder_triangles(tris)
    h = cw.ray_traverse(rays, bt)
    r = O.traverse(rcw, tris[rprims], rays, threads=8)
    np.testing.assert_array_equal(h["primitive_id"], r["primitive_id"])
    np.testing.assert_array_equal(h["t"].view(np.uint32), r["t"].view(np.uint32))
    m = cw.ray_traverse_miss(shadow, bt)
    rm = O.traverse(rcw, tris[rprims], shadow, mode=1, threads=8)
    np.testing.assert_array_equal(m, rm)
    assert 0.05 < (~m).mean() < 0.95 and np.isfinite(h["t"]).mean() > 0.05


def _area_triangles(values):
    """Triangles (0,0,0), (v,0,0), (0,1,0): their boxes have half area exactly v."""
    v = np.asarray(values, dtype=np.float32)
    z = np.zeros_like(v)
    return ob.triangles_from_vertices(np.stack([z, z, z], 1), np.stack([v, z, z], 1),
                                      np.stack([z, np.ones_like(v), z], 1))


def _sum_cases():
    rng = np.random.default_rng(3)
    half_ulp_of_one = np.float32(2.0 ** -24)
    return {
        "ties_at_one": np.concatenate([[1.0], np.full(20000, half_ulp_of_one)]),
        "ties_then_carry": np.concatenate([[1.0], np.full(5000, half_ulp_of_one), np.full(5000, 3 * half_ulp_of_one)]),
        "zeros_first": np.concatenate([np.zeros(7000), rng.uniform(0, 1, 9000)]),
        "log_uniform": np.exp(rng.uniform(np.log(1e-30), np.log(1e3), 200000)),
        "subnormal": np.concatenate([np.full(5000, np.float32(1e-45)), np.full(3, np.float32(1e-38))]),
        "powers_of_two": 2.0 ** rng.integers(-40, 20, 50000),
        "one_value": [0.25],
        # the sum reaches 2.0 at element 16384, the last of the 16K tile that starts after the first element
        "stop_on_tile_end": np.concatenate([[1.0], np.full(49252, 2.0 ** -14)]),
        "two_full_tiles": rng.uniform(0, 1, 32768),
        # past the first 2^18 elements the sum runs in passes over all CUs; from s = 300000 (ulp 2^-5), 2^-6 is a tie
        "passes_with_ties": np.concatenate([np.ones(300000), np.where(rng.random(400000) < 1e-4, 2.0 ** -6,
                                                                      rng.uniform(0, 0.1, 400000))]),
        "passes_cross_binades": np.concatenate([np.ones(300000), rng.uniform(0, 4, 600000)]),
        "late_overflow": np.concatenate([np.ones(300000), [3e38, 3e38], np.ones(100000)]),
    }


@pytest.mark.parametrize("case", list(_sum_cases()))
def test_split_average_is_the_sequential_f32_sum(ctx, case):
    """cwbvh/builder.rs:34-41 sums half areas one f32 add at a time; the GPU reproduces every rounding (ties to
    even, binade changes, subnormals) with a parallel integer scan."""
    print(f"[case] ordered sum {case}", flush=True)
    tris = _area_triangles(_sum_cases()[case])
    _, _, avg, big = ob.split_aabbs_preset(ctx, tris, 0, with_stats=True)
    _, _, ravg, rbig = O.presplit(tris, 0)
    assert np.float32(avg).tobytes() == np.float32(ravg).tobytes(), (avg, ravg)
    assert np.float32(big).tobytes() == np.float32(rbig).tobytes()
