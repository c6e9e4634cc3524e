"""CPU checks of the seeded workload generators (no GPU)."""
import numpy as np

from obvhs_amd import scenes


def test_icosphere_batch_counts_rule():
    """Log-uniform counts in [64, 65536] rounded down to 20 * 4^s; the 4,096-mesh total is pinned."""
    s = scenes.icosphere_batch_counts(4096, seed=0)
    counts = 20 * 4 ** s
    assert s.min() == 0 and s.max() == 5
    assert int(counts.sum()) == 20174540
    # rounding down: 20 * 4^s <= count < 20 * 4^(s+1) for the continuous draw
    u = np.linspace(0.0, 1.0, 1001)
    c = 64.0 * (65536.0 / 64.0) ** u
    lv = np.clip(np.floor(np.log(c / 20.0) / np.log(4.0)), 0, 5)
    assert ((20 * 4 ** lv <= c) | (lv == 0)).all() and (20 * 4 ** (lv + 1) > c).all()


def test_icosphere_batch_geometry_and_rays():
    levels = np.array([0, 1, 2, 3, 2, 0])
    tris, offsets, rays, ray_mesh = scenes.icosphere_batch(len(levels), seed=3, rays_per_mesh=64, levels=levels)
    assert list(np.diff(offsets)) == list(20 * 4 ** levels)
    for i, lvl in enumerate(levels):
        t = tris[offsets[i]:offsets[i + 1]]
        v = np.concatenate([t["v0"][:, :3], t["v1"][:, :3], t["v2"][:, :3]])
        # shared vertices are displaced identically: the mesh stays closed (10 * 4^s + 2 distinct vertices)
        assert len(np.unique(v, axis=0)) == 10 * 4 ** lvl + 2
        r = np.linalg.norm(v.astype(np.float64), axis=1)
        assert r.min() >= 0.9 - 1e-6 and r.max() <= 1.1 + 1e-6
        sel = ray_mesh == i
        assert sel.sum() == 64
        o = rays["origin"][sel, :3].astype(np.float64)
        d = rays["direction"][sel, :3].astype(np.float64)
        np.testing.assert_allclose(np.linalg.norm(o, axis=1), 2.0, rtol=1e-5)
        np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, rtol=1e-5)
        # every ray is aimed at a point of the mesh AABB: the slab test along the ray passes
        lo, hi = v.min(axis=0).astype(np.float64), v.max(axis=0).astype(np.float64)
        with np.errstate(divide="ignore"):
            inv = 1.0 / d
        t0, t1 = (lo - o) * inv, (hi - o) * inv
        tn = np.minimum(t0, t1).max(axis=1)
        tf = np.maximum(t0, t1).min(axis=1)
        assert (tn <= tf * (1 + 1e-6) + 1e-6).all() and (tf > 0).all()
    # same seed, same bytes
    tris2, *_ = scenes.icosphere_batch(len(levels), seed=3, rays_per_mesh=64, levels=levels)
    assert tris.tobytes() == tris2.tobytes()
