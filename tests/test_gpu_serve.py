"""Serving: each request's result is the same alone and in the mix.

A seeded request list (closest-hit and occlusion ray batches of 1 .. 20,000 rays over four resident scenes,
interleaved with refit and rebuild updates) is replayed through the server in a burst (everything submitted at
once, maximal coalescing), spaced out, and under a tiny HBM budget (admission one request at a time).  Every trace
must equal, byte for byte, the single-request call of the plain API on the scene version that was current when it
was submitted (CwBvh::ray_traverse / ray_traverse_miss, cwbvh/mod.rs:169-225; refit = Bvh2::refit_all +
bvh2_to_cwbvh, bvh2/mod.rs:506-569; rebuild = build_cwbvh_from_tris, cwbvh/builder.rs:20-85)."""
import numpy as np
import pytest

import obvhs_amd as ob
from obvhs_amd import scenes, test_util

pytestmark = pytest.mark.gpu

PARAMS = [ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3),
          ob.BvhBuildParams(ploc_search_distance=6, search_depth_threshold=2, max_prims_per_leaf=3,
                            reinsertion_batch_ratio=0.05),
          ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=2),
          ob.BvhBuildParams(ploc_search_distance=1, max_prims_per_leaf=1)]


def initial_scenes():
    return [scenes.soup(3000, seed=1), scenes.heightfield_scene(48, 50, seed=2), scenes.soup(20000, seed=3),
            scenes.soup(1, seed=4)]


def aimed_rays(tris, n, seed):
    i = np.arange(n, dtype=np.uint32)
    o = np.stack([test_util.hash_noise(i, a, seed) for a in range(3)], axis=1) * np.float32(3.0) - np.float32(1.0)
    pick = (test_util.uhash2(i, np.uint32(seed + 77)) % np.uint32(len(tris))).astype(np.int64)
    c = (tris["v0"][pick, :3].astype(np.float64) + tris["v1"][pick, :3] + tris["v2"][pick, :3]) / 3.0
    d = c - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return ob.make_rays(o, d.astype(np.float32), 0.0, 1e30)


def request_list(count=90, seed=5):
    """(kind, scene, payload) in submission order; payload = ray count for traces, seed for updates."""
    out = []
    rng = np.random.default_rng(seed)
    for k in range(count):
        scene = int(rng.integers(0, 4))
        if k % 9 == 8:
            out.append(("rebuild" if k % 18 == 17 else "refit", scene, k))
        else:
            n = int(np.exp(rng.uniform(0.0, np.log(20000.0))))
            out.append(("miss" if k % 3 == 2 else "closest", scene, max(1, n)))
    return out


class Expected:
    """The plain API applied request by request (the 'alone' results)."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.tris = initial_scenes()
        self.cw = [ob.build_cwbvh_from_tris(ctx, t, p) for t, p in zip(self.tris, PARAMS)]
        self.bt = [c.reorder_triangles(t) for c, t in zip(self.cw, self.tris)]
        self.version = [0, 0, 0, 0]

    def update(self, scene, kind, seed):
        if kind == "refit":
            new = scenes.jitter_triangles(self.tris[scene], seed=seed, frac=0.01)
            b2 = self.cw[scene].bvh2
            b2.refit_from_triangles(new)
            mp = max(1, min(3, PARAMS[scene].max_prims_per_leaf))
            cw = ob.bvh2_to_cwbvh(b2, mp, True, False)
            cw._owner = self.cw[scene]  # keep the BVH2's owner alive
        else:
            base = initial_scenes()[scene]
            new = scenes.jitter_triangles(base[: max(1, len(base) - seed)], seed=seed, frac=0.01)
            cw = ob.build_cwbvh_from_tris(self.ctx, new, PARAMS[scene])
        self.tris[scene] = new
        self.cw[scene] = cw if kind == "rebuild" else _RefitView(cw)
        self.bt[scene] = cw.reorder_triangles(new)
        self.version[scene] += 1
        return new

    def trace(self, scene, rays, miss):
        cw = self.cw[scene]
        return cw.ray_traverse_miss(rays, self.bt[scene]) if miss else cw.ray_traverse(rays, self.bt[scene])


class _RefitView:
    """A CWBVH converted from a refit BVH2 whose later refits start from the same BVH2 (like the server's slots)."""

    def __init__(self, cw):
        self._cw = cw
        self.bvh2 = cw._owner.bvh2

    def __getattr__(self, name):
        return getattr(self._cw, name)


def build_workload(ctx):
    exp = Expected(ctx)
    reqs, checks = [], []
    for k, (kind, scene, payload) in enumerate(request_list()):
        if kind in ("refit", "rebuild"):
            new = exp.update(scene, kind, payload)
            reqs.append(dict(kind=ob.api.REQ_REFIT if kind == "refit" else ob.api.REQ_REBUILD, scene=scene,
                             tris=ctx.upload(new)))
            checks.append(("update", exp.version[scene], None))
        else:
            rays = aimed_rays(exp.tris[scene], payload, seed=k)
            miss = kind == "miss"
            d_rays = ctx.upload(rays)
            d_out = ob.DeviceArray(ctx, np.uint8 if miss else ob.RAY_HIT, len(rays))
            reqs.append(dict(kind=ob.api.REQ_TRACE_MISS if miss else ob.api.REQ_TRACE_CLOSEST, scene=scene,
                             rays=d_rays, out=d_out))
            checks.append(("trace", exp.version[scene], exp.trace(scene, rays, miss)))
    return reqs, checks


def run_and_check(server, ctx, reqs, checks, spacing_ns):
    scene_tris = initial_scenes()
    for t, p in zip(scene_tris, PARAMS):
        server.add_scene(ctx.upload(t), p)
    for i, q in enumerate(reqs):
        q["arrival_ns"] = i * spacing_ns
    infos = server.replay(reqs)
    for i, (q, (what, version, expect), info) in enumerate(zip(reqs, checks, infos)):
        assert info["status"] == "OBVHS_OK", (i, info)
        assert info["version"] == version, (i, info, version)
        if what == "trace":
            got = q["out"].to_host()
            if q["kind"] == ob.api.REQ_TRACE_MISS:
                np.testing.assert_array_equal(got.astype(bool), expect, err_msg=f"request {i}")
            else:
                assert got.tobytes() == expect.tobytes(), f"request {i}: hits differ"
    return infos


def test_burst_mix_equals_alone(ctx):
    print("[case] serving burst", flush=True)
    reqs, checks = build_workload(ctx)
    with ob.Server(0) as server:
        infos = run_and_check(server, ctx, reqs, checks, 0)
        st = server.stats()
    traces = [x for x in infos if x["kind"] in (0, 1)]
    assert max(x["batch_requests"] for x in traces) > 1  # coalesced
    assert st["segments"] == len(traces) and st["launches"] < len(traces)


def test_spaced_mix_equals_alone(ctx):
    print("[case] serving spaced", flush=True)
    reqs, checks = build_workload(ctx)
    with ob.Server(0, max_inflight=4) as server:
        run_and_check(server, ctx, reqs, checks, 150_000)


def test_tiny_budget_admits_one_at_a_time(ctx):
    print("[case] serving tiny budget", flush=True)
    reqs, checks = build_workload(ctx)
    with ob.Server(0, hbm_budget_bytes=1, max_batch_rays=4096) as server:
        infos = run_and_check(server, ctx, reqs, checks, 0)
    traces = [x for x in infos if x["kind"] in (0, 1)]
    assert max(x["batch_requests"] for x in traces) == 1


def test_invalid_requests(ctx):
    with ob.Server(0) as server:
        sid = server.add_scene(ctx.upload(scenes.soup(100, seed=1)), PARAMS[0])
        rays = ctx.upload(aimed_rays(scenes.soup(100, seed=1), 8, 1))
        out = ob.DeviceArray(ctx, ob.RAY_HIT, 8)
        with pytest.raises(ob.ObvhsError):
            server.submit_trace(sid + 5, rays, out)
        bad = ctx.upload(scenes.soup(50, seed=2))
        uid = server.submit_update(sid, bad, rebuild=False)  # a refit must keep the triangle count
        assert server.wait(uid)["status"] == "OBVHS_ERR_INVALID_ARG"
        assert server.scene_info(sid)["version"] == 0
        rid = server.submit_trace(sid, rays, out)
        assert server.wait(rid)["status"] == "OBVHS_OK"
