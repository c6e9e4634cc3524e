"""Benchmark of the obvhs hot path on one MI355X.

Workload (heightfield scene, closest hit): a 724 x 724 grid over [0,1]^2 (2 triangles per cell, height = sum of 8
seeded sines) plus 20,000 seeded 12-triangle boxes = 1,288,352 f32 triangles; PLOC BVH2 (search distance 14, no
pre-splits) -> CWBVH8; 1920 x 1080 coherent primary rays plus one cosine-weighted bounce ray per primary hit.

A step is one closest-hit trace of the whole ray batch (primary + bounce) against the device-resident CWBVH, with
rays, triangles and hit buffers resident in HBM.  value = rays per step * steps / wall time of the timed steps.
The PLOC -> CWBVH build is timed in the same run (device-resident triangles to device-resident CWBVH + primitive
indices) and reported beside it as Mtris/s.  The roofline entry is for the traversal kernel (the dominant kernel),
with the algorithmic bytes of every ray (40 B of ray/hit + 80 B per node test + 36 B per triangle test, counts taken
from an instrumented run of the same kernel) divided by its event-timed duration.  The CPU baseline is the C++
restatement of the crate's CPU path (oracle/), timed on this host's cores in this process.

The "batched" object reports the batched form on 4,096 seeded icospheres (~20 M triangles, 1,024 rays each): one
shared build and one traversal launch per step, next to the plain form (a build_cwbvh_from_tris call and a
traversal launch per mesh) on the same inputs, with a check that both give identical hits.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
METRIC = "Mrays/s CWBVH closest-hit; Mtris/s PLOC build; p50/p99 batch latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(tris, rays_all, threads, sample_rays):
    """Times the oracle's traversal (all `threads` host threads) and its serial build on this host."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as O
    t0 = time.perf_counter()
    b, cw, prims = O.build_cwbvh(tris, 14, 0, 3)
    build_s = time.perf_counter() - t0
    bvh_tris = tris[prims]
    idx = np.linspace(0, len(rays_all) - 1, num=min(sample_rays, len(rays_all))).astype(np.int64)
    sample = np.ascontiguousarray(rays_all[idx])
    O.traverse(cw, bvh_tris, sample[:4096], 0, threads=threads)  # warm-up
    t0 = time.perf_counter()
    O.traverse(cw, bvh_tris, sample, 0, threads=threads)
    trace_s = time.perf_counter() - t0
    return dict(build_s=build_s, trace_s=trace_s, n_rays=len(sample), n_tris=len(tris))


def bench_batch(args, ob, ctx):
    """Batched form: 4,096 icospheres (log-uniform 64..65,536 triangles rounded down to 20*4^s) built in shared
    launches and traced with 1,024 rays per mesh in one launch per step; beside it the plain form on the same
    workload (one build_cwbvh_from_tris call and one traversal launch per mesh)."""
    import ctypes as C
    from obvhs_amd import scenes
    lib = ob.native.lib()
    m = args.batch_meshes
    log(f"generating {m} icospheres")
    tris, offsets, rays, ray_mesh = scenes.icosphere_batch(m, seed=0, rays_per_mesh=args.batch_rays_per_mesh)
    n_tris, n_rays = len(tris), len(rays)
    d_tris, d_rays, d_mesh = ctx.upload(tris), ctx.upload(rays), ctx.upload(ray_mesh)
    params = ob.BvhBuildParams(ploc_search_distance=14, search_depth_threshold=0, max_prims_per_leaf=3)
    log(f"batched build: {m} meshes, {n_tris} triangles")
    batch = ob.build_cwbvh_batch(ctx, d_tris, offsets, params)
    ms = []
    for _ in range(max(1, args.build_reps)):
        ctx.timer_start()
        ob.build_cwbvh_batch(ctx, d_tris, offsets, params, reuse=batch)
        ms.append(ctx.timer_stop())
    build_ms = float(np.median(ms))
    d_bvh_tris = batch.reorder_triangles(d_tris)
    hits = ob.DeviceArray(ctx, ob.RAY_HIT, n_rays)

    def trace_batch():
        ob.native.check(lib.obvhs_batch_ray_traverse(ctx.handle, batch.handle, d_bvh_tris.ptr, d_rays.ptr,
                                                     d_mesh.ptr, n_rays, hits.ptr), ctx.handle)

    for _ in range(args.warmup):
        trace_batch()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        trace_batch()
    ctx.synchronize()
    trace_s = (time.perf_counter() - t0) / args.steps
    batch_hits = hits.to_host()

    # Plain form: per-mesh handles, builds and traversal launches through the single-mesh entry points.
    log("plain per-mesh builds")
    cp = params.to_c()
    tri_sz, ray_sz, hit_sz = ob.TRIANGLE.itemsize, ob.RAY.itemsize, ob.RAY_HIT.itemsize
    handles = [C.c_void_p() for _ in range(m)]
    base_t, base_r, base_h = d_tris.ptr.value, d_rays.ptr.value, hits.ptr.value
    d_plain_tris = ob.DeviceArray(ctx, ob.TRIANGLE, n_tris)
    base_pt = d_plain_tris.ptr.value
    rpm = args.batch_rays_per_mesh

    def plain_build(reuse):
        for i in range(m):
            o, n = int(offsets[i]), int(offsets[i + 1] - offsets[i])
            ob.native.check(lib.obvhs_build_cwbvh_from_tris(ctx.handle, C.c_void_p(base_t + o * tri_sz), n,
                                                            C.byref(cp), handles[i] if reuse else None,
                                                            C.byref(handles[i]), None), ctx.handle)

    plain_build(False)
    ctx.timer_start()
    plain_build(True)
    plain_build_ms = ctx.timer_stop()
    for i in range(m):
        o = int(offsets[i])
        ob.native.check(lib.obvhs_cwbvh_reorder_triangles(ctx.handle, handles[i], C.c_void_p(base_t + o * tri_sz),
                                                          C.c_void_p(base_pt + o * tri_sz)), ctx.handle)

    def trace_plain():
        for i in range(m):
            o = int(offsets[i])
            ob.native.check(lib.obvhs_cwbvh_ray_traverse(ctx.handle, handles[i], C.c_void_p(base_pt + o * tri_sz),
                                                         C.c_void_p(base_r + i * rpm * ray_sz), rpm,
                                                         C.c_void_p(base_h + i * rpm * hit_sz)), ctx.handle)

    trace_plain()
    ctx.synchronize()
    plain_hits = hits.to_host()
    same = bool(plain_hits.tobytes() == batch_hits.tobytes())
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps // 4)):
        trace_plain()
    ctx.synchronize()
    plain_trace_s = (time.perf_counter() - t0) / max(1, args.steps // 4)
    for h in handles:
        lib.obvhs_cwbvh_destroy(h)
    info = batch.info()
    batch.close()
    return {"workload": f"{m} icospheres (log-uniform 64..65536 tris, rounded down to 20*4^s) = {n_tris} tris; "
                        f"{rpm} rays per mesh = {n_rays} closest-hit rays",
            "build_ms": round(build_ms, 3), "build_mtris_per_s": round(n_tris / build_ms / 1e3, 2),
            "trace_mrays_per_s": round(n_rays / trace_s / 1e6, 2), "trace_ms": round(trace_s * 1e3, 4),
            "plain_build_ms": round(plain_build_ms, 3),
            "plain_build_mtris_per_s": round(n_tris / plain_build_ms / 1e3, 2),
            "plain_trace_mrays_per_s": round(n_rays / plain_trace_s / 1e6, 2),
            "trace_speedup_vs_plain": round(plain_trace_s / trace_s, 2),
            "build_speedup_vs_plain": round(plain_build_ms / build_ms, 2),
            "hits_identical_to_plain": same, "ploc_passes": info["ploc_passes"],
            "cwbvh_nodes": info["total_nodes"], "hit_fraction": round(float(np.isfinite(batch_hits["t"]).mean()), 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--build-reps", type=int, default=5)
    ap.add_argument("--grid", type=int, default=724)
    ap.add_argument("--boxes", type=int, default=20000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-rays", type=int, default=1 << 20)
    ap.add_argument("--batch-meshes", type=int, default=4096, help="0 skips the batched leg")
    ap.add_argument("--batch-rays-per-mesh", type=int, default=1024)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "design-only, see the design notes)"}))
        sys.exit(2)

    import obvhs_amd as ob
    from obvhs_amd import scenes

    log("generating scene")
    tris, prim_rays = scenes.heightfield_workload(args.grid, args.boxes, args.width, args.height, seed=0)
    n_tris = len(tris)
    ctx = ob.Context(0)
    d_tris = ctx.upload(tris)
    params = ob.BvhBuildParams(ploc_search_distance=14, search_depth_threshold=0, max_prims_per_leaf=3)

    log(f"building CWBVH over {n_tris} triangles")
    cw = ob.build_cwbvh_from_tris(ctx, d_tris, params)
    ctx.synchronize()
    # Build timing (device-resident input -> device-resident CWBVH + primitive indices).
    build_ms = []
    for _ in range(max(1, args.build_reps)):
        ctx.timer_start()
        ob.build_cwbvh_from_tris(ctx, d_tris, params, reuse=cw)
        build_ms.append(ctx.timer_stop())
    stats = cw.stats
    build_med = float(np.median(build_ms))

    d_bvh_tris = cw.reorder_triangles(d_tris)
    bvh_tris_host = d_bvh_tris.to_host()
    d_prim = ctx.upload(prim_rays)
    hits = ob.DeviceArray(ctx, ob.RAY_HIT, len(prim_rays))
    lib = ob.native.lib()

    def trace(d_rays, n, out):
        ob.native.check(lib.obvhs_cwbvh_ray_traverse(ctx.handle, cw.handle, d_bvh_tris.ptr, d_rays.ptr, n, out.ptr),
                        ctx.handle)

    trace(d_prim, len(prim_rays), hits)
    prim_hits = hits.to_host()
    bounce = scenes.bounce_rays(prim_rays, prim_hits, bvh_tris_host, seed=1)
    d_bounce = ctx.upload(bounce)
    bhits = ob.DeviceArray(ctx, ob.RAY_HIT, max(1, len(bounce)))
    n_rays = len(prim_rays) + len(bounce)
    log(f"{len(prim_rays)} primary + {len(bounce)} bounce rays")

    # Per-ray node/triangle test counts for the byte model (instrumented launch, not timed).
    _, nv_p, tv_p = cw.ray_traverse(d_prim, d_bvh_tris, counted=True)
    _, nv_b, tv_b = cw.ray_traverse(d_bounce, d_bvh_tris, counted=True)
    node_tests = int(nv_p.sum()) + int(nv_b.sum())
    tri_tests = int(tv_p.sum()) + int(tv_b.sum())
    algo_bytes = 40 * n_rays + 80 * node_tests + 36 * tri_tests

    def step():
        trace(d_prim, len(prim_rays), hits)
        trace(d_bounce, len(bounce), bhits)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    wall = time.perf_counter() - t0

    # Kernel durations (events around each single-kernel launch on the context stream).
    kp, kb = [], []
    for _ in range(max(3, min(args.steps, 10))):
        ctx.timer_start()
        trace(d_prim, len(prim_rays), hits)
        kp.append(ctx.timer_stop())
        ctx.timer_start()
        trace(d_bounce, len(bounce), bhits)
        kb.append(ctx.timer_stop())
    kernel_s = (np.mean(kp) + np.mean(kb)) * 1e-3
    achieved_gbs = algo_bytes / kernel_s / 1e9

    value = n_rays * args.steps / wall / 1e6
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded heightfield + boxes, pinhole primary + cosine bounce rays)",
        "config": {"workload": f"heightfield {args.grid}x{args.grid}x2 + {args.boxes} boxes = {n_tris} tris; "
                               f"{len(prim_rays)} primary + {len(bounce)} bounce closest-hit rays",
                   "ploc_search_distance": 14, "max_prims_per_leaf": 3},
        "roofline": {"bound": "hbm", "achieved": round(achieved_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_traverse<closest>", "avg_kernel_ms": {"primary": round(float(np.mean(kp)), 4),
                                                                       "bounce": round(float(np.mean(kb)), 4)},
                     "bytes_per_step": algo_bytes, "node_tests": node_tests, "tri_tests": tri_tests},
        "build": {"ms": round(build_med, 3), "mtris_per_s": round(n_tris / build_med / 1e3, 2),
                  "all_ms": [round(x, 3) for x in build_ms], "ploc_passes": int(stats.ploc_iterations),
                  "ploc_cluster_sum": int(stats.ploc_cluster_sum), "cwbvh_nodes": int(stats.cwbvh_node_count)},
    }
    if args.batch_meshes > 0:
        out["batched"] = bench_batch(args, ob, ctx)
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        log(f"cpu baseline with {threads} threads")
        cb = cpu_baseline(tris, np.concatenate([prim_rays, bounce]), threads, args.cpu_sample_rays)
        out["cpu_baseline"] = {"value": round(cb["n_rays"] / cb["trace_s"] / 1e6, 3), "unit": "Mrays/s",
                               "cores": threads, "kind": "port",
                               "sample": f"{cb['n_rays']} of the {n_rays} rays (evenly strided), closest hit, "
                                         f"SSE-semantics C++ restatement of the crate's traversal",
                               "build_mtris_per_s_serial": round(cb["n_tris"] / cb["build_s"] / 1e6, 3)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
