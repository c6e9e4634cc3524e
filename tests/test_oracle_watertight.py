"""Pins of the oracle's watertight triangle test (the library's extra primitive test; Woop, Benthin, Wald,
"Watertight Ray/Triangle Intersection", JCGT 2(1) 2013) and of the traversal with it.

What pins it:
  * hand-derived known answers: an axis-aligned ray hits the unit right triangle at t = 1 from either side (the
    second shears onto a negative dominant axis, which swaps kx and ky), and a ray through a vertex exactly hits
    (its edge functions are exactly 0, so the f64 recomputation runs and 0 counts as inside) where the crate's
    Triangle::intersect misses (its barycentric u is -0.0, triangle.rs:60); a planted smallest-axis shear and a
    planted "exactly 0 is outside" rule are caught;
  * agreement with Triangle::intersect on random rays whose hit point is clear of every edge: the same hit or miss,
    t within 1e-5 relative for 99% of the hits (grazing triangles round differently in the two tests);
  * watertightness: on a grid mesh closed over its square, rays straight down through every grid vertex and rays
    from random points above the square aimed just below every interior edge's midpoint all hit (brute force over
    all triangles, and the CWBVH traversal with the watertight test);
  * traversal of a CWBVH with the watertight test equals the watertight brute force (t bits) on a scene, and on the
    closed meshes (its node test is widened; with the crate's node test a planted run culls some edge rays)."""
import numpy as np
import pytest

import oracle_lib as O
from obvhs_amd import scenes
from obvhs_amd.test_util import hash_noise
from obvhs_amd.types import make_rays, triangles_from_vertices

MUT_WT_MIN_AXIS, MUT_WT_EDGE_STRICT, MUT_WT_NODE_PLAIN = 21, 22, 23


@pytest.fixture
def mutation():
    yield lambda m: O.lib().orc_set_mutation(m)
    O.lib().orc_set_mutation(0)


UNIT = triangles_from_vertices([[0, 0, 1]], [[1, 0, 1]], [[0, 1, 1]])[0]


def ray(o, d):
    return make_rays(np.array([o], np.float32), np.array([d], np.float32), 0.0, 1e30)[0]


def test_known_answers(mutation):
    # U, V, W = -0.5, -0.25, -0.25 (all inside, det = -1), T = -1: t = 1
    assert O.tri_intersect_wt(UNIT, ray((0.25, 0.25, 0.0), (0, 0, 1))) == 1.0
    # from the other side: d[kz] < 0 swaps kx and ky, shear (-0, -0, -1): t = 1
    assert O.tri_intersect_wt(UNIT, ray((0.25, 0.25, 2.0), (0, 0, -1))) == 1.0
    # misses: outside the triangle, behind the origin
    assert O.tri_intersect_wt(UNIT, ray((0.75, 0.75, 0.0), (0, 0, 1))) == np.inf
    assert O.tri_intersect_wt(UNIT, ray((0.25, 0.25, 2.0), (0, 0, 1))) == np.inf
    # through vertex v0 exactly: V = W = 0 (recomputed in f64, still 0), U = -1: inside, t = 1
    through_vertex = ray((0.0, 0.0, 0.0), (0, 0, 1))
    assert O.tri_intersect_wt(UNIT, through_vertex) == 1.0
    # Triangle::intersect misses it: r = d x c = 0, u = 0 * inv_det(-1) = -0.0, whose sign bit fails the test
    assert O.lib().orc_tri_intersect(O._ptr(np.ascontiguousarray(UNIT)),
                                     O._ptr(np.ascontiguousarray(through_vertex))) == np.inf
    mutation(MUT_WT_MIN_AXIS)  # shears onto d.x = 0: 0/0 shear, every test fails
    assert O.tri_intersect_wt(UNIT, ray((0.25, 0.25, 0.0), (0, 0, 1))) == np.inf
    mutation(MUT_WT_EDGE_STRICT)
    assert O.tri_intersect_wt(UNIT, through_vertex) == np.inf
    mutation(0)


def test_agrees_with_triangle_intersect_away_from_edges():
    rng = np.random.default_rng(3)
    n = 20000
    v = rng.random((n, 3, 3)).astype(np.float32)
    tris = triangles_from_vertices(v[:, 0], v[:, 1], v[:, 2])
    # aim at a point with barycentrics clear of 0 (>= 0.05) or clearly outside
    b = rng.dirichlet([1.0, 1.0, 1.0], size=n)
    inside = (b > 0.05).all(axis=1)
    p = (b[:, 0:1] * v[:, 0] + b[:, 1:2] * v[:, 1] + b[:, 2:3] * v[:, 2])
    outside = rng.random(n) < 0.3
    p[outside] += (rng.random((outside.sum(), 3)) - 0.5) * 4.0
    o = (rng.random((n, 3)) * 4.0 - 2.0).astype(np.float32)
    d = p - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = make_rays(o, d.astype(np.float32), 0.0, 1e30)
    checked, rel = 0, []
    for i in range(n):
        if not inside[i] and not outside[i]:
            continue
        t_wt = O.tri_intersect_wt(tris[i], rays[i])
        t_mt = O.lib().orc_tri_intersect(O._ptr(np.ascontiguousarray(tris[i])), O._ptr(np.ascontiguousarray(rays[i])))
        if outside[i]:
            # a point pushed outside may still land inside by chance: then both must agree on the hit
            if np.isinf(t_wt) != np.isinf(t_mt):
                continue
        assert np.isinf(t_wt) == np.isinf(t_mt), i
        if np.isfinite(t_wt):
            rel.append(abs(t_wt - t_mt) / max(1.0, abs(t_mt)))
        checked += 1
    assert checked > 15000
    # both tests round differently; grazing triangles (det near 0) amplify it
    rel = np.array(rel)
    assert np.percentile(rel, 99) <= 1e-5 and rel.max() <= 1e-3


def grid_mesh(k=16, seed=0):
    """A k x k grid over [0,1]^2 (x, z) with dyadic heights, two triangles per cell: a closed surface over the
    square (every interior edge shared by two triangles, every interior vertex by six)."""
    xs = np.arange(k + 1, dtype=np.float32) / np.float32(k)
    h = (np.floor(hash_noise(np.arange((k + 1) ** 2, dtype=np.uint32), 0, seed) * 8.0) / 16.0).astype(np.float32)
    h = h.reshape(k + 1, k + 1)
    v0, v1, v2 = [], [], []
    for i in range(k):
        for j in range(k):
            a = (xs[i], h[i, j], xs[j])
            b = (xs[i + 1], h[i + 1, j], xs[j])
            c = (xs[i], h[i, j + 1], xs[j + 1])
            d = (xs[i + 1], h[i + 1, j + 1], xs[j + 1])
            v0 += [a, b]
            v1 += [b, d]
            v2 += [c, c]
    return triangles_from_vertices(np.array(v0), np.array(v1), np.array(v2)), xs, h


def watertight_rays(k=16, seed=0):
    """Rays straight down through every interior grid vertex, and rays from random points above the square through
    the midpoint of every interior edge."""
    tris, xs, h = grid_mesh(k, seed)
    ii, jj = np.meshgrid(np.arange(1, k), np.arange(1, k), indexing="ij")
    vert = np.stack([xs[ii.ravel()], np.full(ii.size, 2.0, np.float32), xs[jj.ravel()]], axis=1)
    down = make_rays(vert, np.tile(np.array([[0, -1, 0]], np.float32), (len(vert), 1)), 0.0, 1e30)
    mids = []
    for t in tris:
        for a, b in ((t["v0"], t["v1"]), (t["v1"], t["v2"]), (t["v2"], t["v0"])):
            m = (a[:3].astype(np.float64) + b[:3]) / 2.0
            if 0.0 < m[0] < 1.0 and 0.0 < m[2] < 1.0:
                mids.append(m)
    mids = np.unique(np.array(mids), axis=0)
    rng = np.random.default_rng(seed + 7)
    o = mids + np.stack([rng.uniform(-0.3, 0.3, len(mids)), rng.uniform(1.0, 2.0, len(mids)),
                         rng.uniform(-0.3, 0.3, len(mids))], axis=1)
    # origins above the square: the ray then stays over the mesh from above every height to the midpoint, so it
    # crosses the surface (from outside the square it could pass under the open border)
    o[:, 0] = np.clip(o[:, 0], 0.001, 0.999)
    o[:, 2] = np.clip(o[:, 2], 0.001, 0.999)
    # aimed a hair below the midpoint: the segment runs from above every height to below the surface over the square,
    # so it crosses the surface, within about 1e-6 of the edge (aimed exactly at it, a rounded ray can pass just
    # above and leave the square over its open border)
    d = (mids - np.array([0.0, 1e-6, 0.0])) - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    edge = make_rays(o.astype(np.float32), d.astype(np.float32), 0.0, 1e30)
    return tris, down, edge


def test_watertight_on_closed_mesh(mutation):
    tris, down, edge = watertight_rays(16, 0)
    for rays in (down, edge):
        wt = O.brute_force_wt(tris, rays)
        assert (wt["primitive_id"] != 0xFFFFFFFF).all()
    mt = O.brute_force(tris, down)
    # the crate's test is not watertight: some vertex rays slip through (-0.0 barycentrics)
    mt_missed = int((mt["primitive_id"] == 0xFFFFFFFF).sum())
    mutation(MUT_WT_EDGE_STRICT)
    strict = O.brute_force_wt(tris, down)
    mutation(0)
    assert (strict["primitive_id"] == 0xFFFFFFFF).any()
    assert mt_missed > 0  # 106 of the 225 vertex rays on this mesh


@pytest.mark.parametrize("k,seed", [(16, 0), (64, 3)])
def test_watertight_traversal_on_closed_mesh(mutation, k, seed):
    """The CWBVH traversal in the watertight mode misses nothing either: its node test is widened (one quantum per
    side, 2^-20 relative on the slab interval), since the crate's quantised boxes and slab arithmetic are
    conservative only in exact arithmetic; with the crate's node test some edge rays are culled (k = 64: 2 of
    12,160)."""
    tris, down, edge = watertight_rays(k, seed)
    _, cw, prims = O.build_cwbvh(tris, 14, 0, 3)
    bt = tris[prims]
    culled = 0
    for rays in (down, edge):
        got = O.traverse(cw, bt, rays, 2)
        assert (got["primitive_id"] != 0xFFFFFFFF).all()
        bf = O.brute_force_wt(tris, rays)
        np.testing.assert_array_equal(got["t"].view(np.uint32), bf["t"].view(np.uint32))
        assert not O.traverse(cw, bt, rays, 3).any()
        mutation(MUT_WT_NODE_PLAIN)
        culled += int((O.traverse(cw, bt, rays, 2)["primitive_id"] == 0xFFFFFFFF).sum())
        mutation(0)
    if k == 64:
        assert culled > 0


def test_watertight_traversal_equals_brute_force():
    tris = np.concatenate([scenes.heightfield_scene(24, 20, seed=2), grid_mesh(16, 1)[0]])
    _, cw, prims = O.build_cwbvh(tris, 14, 0, 3)
    rays = scenes.pinhole_rays(64, 64, eye=(0.5, 0.6, -0.4), target=(0.5, 0.0, 0.5))
    got = O.traverse(cw, tris[prims], rays, 2)
    bf = O.brute_force_wt(tris, rays)
    np.testing.assert_array_equal(got["t"].view(np.uint32), bf["t"].view(np.uint32))
    hit = bf["primitive_id"] != 0xFFFFFFFF
    assert hit.mean() > 0.3
    # the hit triangles are the same ones (the traversal reports slots of primitive_indices)
    np.testing.assert_array_equal(prims[got["primitive_id"][hit]], bf["primitive_id"][hit])
    miss = O.traverse(cw, tris[prims], rays, 3)
    np.testing.assert_array_equal(miss, ~hit)
