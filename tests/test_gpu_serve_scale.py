"""Serving at its configured scale: all 64 resident scenes of 100K-4M triangles (seeded heightfields and soups,
about 72M triangles), about 200 requests drawn from the serving mix (ray batches of log-uniform size in [1K, 1M]
rays, half coherent pinhole and half incoherent, a quarter of them occlusion queries; one request in ten a refit or
a rebuild after a +-1% vertex jitter), arriving open-loop at Poisson times, with a scheduler batch limit below the
largest requests.

Checked:
  * every served trace equals, byte for byte, the single-request call of the plain API on the scene version current
    when it was submitted (CwBvh::ray_traverse / ray_traverse_miss, cwbvh/mod.rs:169-225), and every request
    reports that version;
  * for every scene of at most 1M triangles, each version's CWBVH equals the oracle's bytes (the initial build and
    rebuilds: build_cwbvh_from_tris, cwbvh/builder.rs:20-85; refits: leaf boxes from the moved triangles,
    Bvh2::refit_all, bvh2_to_cwbvh, bvh2/mod.rs:506-569), and a strided sample of each trace on it equals the
    oracle's traversal of that CWBVH."""
import numpy as np
import pytest

import obvhs_amd as ob
import oracle_lib as O
from obvhs_amd import scenes
from obvhs_amd.test_util import hash_noise, uniform_sample_sphere

pytestmark = pytest.mark.gpu

N_REQUESTS = 200
SMALL = 1_000_000
POOL = 1 << 20
PARAMS = ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3)


def ray_pools():
    side = 1024
    k = np.arange(POOL, dtype=np.uint32)
    o = np.stack([hash_noise(k, a, 21000) for a in range(3)], axis=1)
    d = uniform_sample_sphere(hash_noise(k, 3, 21000), hash_noise(k, 4, 21000))
    incoherent = ob.make_rays(o, d, 0.0, 1e30)
    return {("heightfield", 0): scenes.pinhole_rays(side, side, eye=(0.5, 0.45, -0.35), target=(0.5, 0.0, 0.55)),
            ("soup", 0): scenes.pinhole_rays(side, side, eye=(0.5, 0.5, -1.5)),
            ("heightfield", 1): incoherent, ("soup", 1): incoherent}


class PlainScene:
    """The plain API's current version of one scene: its CWBVH, the BVH2 later refits start from (the built one,
    kept across refits as the server's version slots keep it), and the triangles in primitive order."""

    def __init__(self, ctx, tris):
        self.ctx = ctx
        self.built = ob.build_cwbvh_from_tris(ctx, tris, PARAMS)
        self.cw, self.bvh2 = self.built, self.built.bvh2
        self.bt = self.cw.reorder_triangles(ctx.upload(tris))

    def refit(self, tris):
        self.bvh2.refit_from_triangles(tris)
        self.cw = ob.bvh2_to_cwbvh(self.bvh2, 3, True, False)
        self.bt = self.cw.reorder_triangles(self.ctx.upload(tris))

    def trace(self, d_rays, miss):
        return self.cw.ray_traverse_miss(d_rays, self.bt) if miss else self.cw.ray_traverse(d_rays, self.bt)


class OracleScene:
    """The oracle's BVH2 and CWBVH of one scene version."""

    def __init__(self, tris):
        with O.threads(16):
            self.b, self.cw, self.prims = O.build_cwbvh(tris, 14, 0, 3)
        self.bvh_tris = tris[self.prims]

    def refit(self, tris):
        nodes = self.b["nodes"].copy()
        leaves = np.nonzero(nodes["prim_count"] != 0)[0]
        boxes = O.tri_aabbs(tris)
        p = self.b["prim_indices"][nodes["first_index"][leaves]]
        nodes["min"][leaves], nodes["max"][leaves] = boxes["min"][p], boxes["max"][p]
        self.b = dict(self.b, nodes=O.refit_all(nodes))
        self.cw, self.prims, _ = O.bvh2_to_cwbvh(self.b["nodes"], self.b["prim_indices"], 3, True)
        self.bvh_tris = tris[self.prims]


def test_serving_at_scale_equals_alone_and_oracle(ctx):
    print("[case] serving at scale: 64 scenes, 200 requests", flush=True)
    rng = np.random.default_rng(11)
    sc = scenes.serving_scenes(64, seed=0)
    tris = [t for t, _ in sc]
    kinds = [k for _, k in sc]
    assert len(tris) == 64 and min(map(len, tris)) >= 90_000 and max(map(len, tris)) >= 2_000_000
    pools = ray_pools()

    # the plain API's state per scene ("alone" results) and the oracle's for the small scenes
    plain = [PlainScene(ctx, t) for t in tris]
    version = [0] * 64
    small = [i for i in range(64) if len(tris[i]) <= SMALL]
    assert len(small) >= 20
    orc = {}
    for i in small:
        orc[i] = OracleScene(tris[i])
        assert plain[i].cw.nodes.tobytes() == orc[i].cw.tobytes(), f"scene {i}: version 0 differs from the oracle"

    reqs, checks = [], []
    t = 0.0
    for k in range(N_REQUESTS):
        t += rng.exponential(1.0 / 2000.0)
        s = int(rng.integers(0, 64))
        if rng.random() < 0.1:
            rebuild = rng.random() < 0.5
            new = scenes.jitter_triangles(tris[s], seed=1000 + k, frac=0.01)
            if rebuild:
                plain[s] = PlainScene(ctx, new)
            else:
                plain[s].refit(new)
            tris[s] = new
            version[s] += 1
            if s in orc:
                if rebuild:
                    orc[s] = OracleScene(new)
                else:
                    orc[s].refit(new)
                assert plain[s].cw.nodes.tobytes() == orc[s].cw.tobytes(), f"request {k}: scene {s} CWBVH differs"
            reqs.append(dict(arrival_ns=int(t * 1e9), kind=ob.api.REQ_REBUILD if rebuild else ob.api.REQ_REFIT,
                             scene=s, tris=ctx.upload(new)))
            checks.append((version[s], None, None))
            continue
        n = int(np.exp(rng.uniform(np.log(1000.0), np.log(float(POOL)))))
        which = int(rng.integers(0, 2))
        miss = rng.random() < 0.25
        off = int(rng.integers(0, POOL - n + 1))
        rays = pools[(kinds[s], which)][off:off + n]
        d_rays = ctx.upload(rays)
        out = ob.DeviceArray(ctx, np.uint8 if miss else ob.RAY_HIT, n)
        expect = plain[s].trace(d_rays, miss)
        if s in orc:
            stride = max(1, n // 256)
            sample = np.ascontiguousarray(rays[::stride])
            if miss:
                ref = O.traverse(orc[s].cw, orc[s].bvh_tris, sample, 1)
                np.testing.assert_array_equal(expect[::stride], ref,
                                              err_msg=f"request {k}: occlusion differs from the oracle")
            else:
                ref = O.traverse(orc[s].cw, orc[s].bvh_tris, sample, 0)
                assert expect[::stride].tobytes() == ref.tobytes(), f"request {k}: hits differ from the oracle"
        reqs.append(dict(arrival_ns=int(t * 1e9), kind=ob.api.REQ_TRACE_MISS if miss else ob.api.REQ_TRACE_CLOSEST,
                         scene=s, rays=d_rays, out=out))
        checks.append((version[s], expect, miss))

    n_traces = sum(1 for _, e, _ in checks if e is not None)
    big = sum(1 for q in reqs if q.get("rays") is not None and q["rays"].count > 262_144)
    assert n_traces >= 150 and big >= 5
    print(f"[case] replaying {len(reqs)} requests ({n_traces} traces, {big} above the batch limit)", flush=True)
    with ob.Server(0, hbm_budget_bytes=200 << 30, max_batch_rays=262_144, max_inflight=2) as server:
        for t0 in sc:
            server.add_scene(ctx.upload(t0[0]), PARAMS)
        infos = server.replay(reqs)
        st = server.stats()
    for i, (q, (ver, expect, miss), info) in enumerate(zip(reqs, checks, infos)):
        assert info["status"] == "OBVHS_OK", (i, info)
        assert info["version"] == ver, (i, info, ver)
        if expect is None:
            continue
        got = q["out"].to_host()
        if miss:
            np.testing.assert_array_equal(got.astype(bool), expect, err_msg=f"request {i}")
        else:
            assert got.tobytes() == expect.tobytes(), f"request {i}: served hits differ from the alone call"
    traces = [x for x in infos if x["kind"] in (ob.api.REQ_TRACE_CLOSEST, ob.api.REQ_TRACE_MISS)]
    assert st["segments"] == len(traces)
    assert st["scenes"] == 64
