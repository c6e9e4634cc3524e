"""Benchmark of the obvhs hot path on one MI355X.

Headline workload (the largest single-GPU configuration, closest hit): the large scene -- a 2896 x 2896 heightfield
(16,773,632 triangles) plus 500,000 seeded long thin triangles = 17,273,632 f32 triangles; spatial pre-splits capped
at 1.3x the triangle count and two reinsertion passes (ratio 0.05, sequence [1, 1/3]) -> CWBVH8 (max 3 primitives
per leaf); 4,194,304 incoherent closest-hit rays (uniform sphere origins, uniform directions).

A step is one closest-hit trace of the 4M rays against the device-resident CWBVH, with rays, triangles and hit
buffers resident in HBM.  value = rays per step * steps / wall time of the timed steps.  Before the timed steps the
traversal's LDS stack depth and launch form (one lane per ray, or wave-owned ray rows with lanes refilled) are
chosen by timing the step itself with each choice (none changes a hit).  The roofline entry is for the step's only
kernel: algorithmic bytes per launch (40 B of ray and hit + 80 B per node test + 36 B per triangle test, counts
from an instrumented launch of the same traversal, which visits the same nodes and triangles) divided by its
HIP-event duration on the context stream; "traffic" is the HBM bytes per launch that rocprofv3's FETCH_SIZE and
WRITE_SIZE counters read for the same launch (profiles/), scaled by their reading for a copy of known size.

Beside the headline, in the same run: the large scene's build (device triangles -> device CWBVH + primitive indices,
Mtris/s over the primitives after pre-splits) and any-hit shadow rays; the heightfield scene (1,288,352 triangles,
1080p primary + bounce rays) with its build and a physics-style partial rebuild; the plain 65,536-triangle soup the
crate's CPU path runs (build + 256 x 256 primary rays, GPU and CPU); the batched form on 4,096 icospheres; serving.
The CPU baseline is the C++ restatement of the crate's CPU path (oracle/), timed on this host's cores in this
process: traversal of the large scene's CWBVH with all its rays, and serial and multithreaded builds of the
large-scene recipe at 1/16 scale.
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
PMC_COPY_BYTES = 256 << 20
# HBM bytes per traversal launch from the FETCH_SIZE / WRITE_SIZE counter runs (bench.py --pmc-probe under
# rocprofv3 --pmc, then --traffic-from), scaled by the copy kernel's reading; null until those runs exist.
TRAFFIC_FILES = {"large": os.path.join(REPO, "profiles", "large_17p3M_closest_hit_hbm_traffic.json"),
                 "heightfield": os.path.join(REPO, "profiles", "heightfield_traversal_hbm_traffic.json")}
TRAFFIC_LAUNCHES = {"large": ("closest",), "heightfield": ("primary", "bounce")}
# Issue and cache counters of a leg's dominant traversal launch (rocprofv3 --pmc runs of --pmc-probe): what limits it.
COUNTER_FILES = {"large": os.path.join(REPO, "profiles", "large_17p3M_closest_hit_counters.json"),
                 "heightfield": os.path.join(REPO, "profiles", "heightfield_closest_hit_counters.json")}
METRIC = "Mrays/s CWBVH closest-hit; Mtris/s PLOC build; p50/p99 batch latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_per_launch(leg):
    """Measured HBM bytes per traversal launch of a leg (mean over its launches), or None."""
    try:
        with open(TRAFFIC_FILES[leg]) as f:
            t = json.load(f)
        return int(round(np.mean([t[k]["bytes"] for k in TRAFFIC_LAUNCHES[leg]])))
    except (OSError, KeyError, ValueError):
        return None


def write_traffic(leg, fetch_csv, write_csv):
    """Turns the two `--pmc-probe` counter runs (FETCH_SIZE, then WRITE_SIZE; rocprofv3 CSV, KiB per dispatch) into
    the leg's traffic file.  Each counter is scaled by what it reads for the copy of PMC_COPY_BYTES that the probe
    runs first (that copy reads and writes exactly PMC_COPY_BYTES); the leg's traversal launches are the first
    traversal dispatches after it."""
    import csv

    def per_dispatch(path, counter):
        rows = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                d = int(r["Dispatch_Id"])
                name, val = r["Kernel_Name"], float(r["Counter_Value"])
                prev = rows.get(d, (name, 0.0))
                rows[d] = (name, prev[1] + val)  # summed over agents/dimensions if split
        return sorted((d, n, v) for d, (n, v) in rows.items())

    launches = TRAFFIC_LAUNCHES[leg]
    out = {"copy_bytes": PMC_COPY_BYTES}
    for key, path, counter in (("fetch", fetch_csv, "FETCH_SIZE"), ("write", write_csv, "WRITE_SIZE")):
        rows = per_dispatch(path, counter)
        copies = [(d, v) for d, n, v in rows if "copyBuffer" in n]
        if not copies:
            raise SystemExit(f"{path}: no copy dispatch")
        copy_d, copy_kib = max(copies, key=lambda t: t[1])
        trav = [(d, n, v) for d, n, v in rows if d > copy_d and "k_traverse" in n][:len(launches)]
        if len(trav) != len(launches):
            raise SystemExit(f"{path}: expected {len(launches)} traversal dispatches after the copy")
        scale = PMC_COPY_BYTES / (copy_kib * 1024.0)
        out[key] = {"copy_kib": copy_kib, "scale": scale}
        for name, (_, kname, v) in zip(launches, trav):
            out[key][f"{name}_kib"] = v
            out[key][f"{name}_kernel"] = kname
    for launch in launches:
        fb = out["fetch"][f"{launch}_kib"] * 1024.0 * out["fetch"]["scale"]
        wb = out["write"][f"{launch}_kib"] * 1024.0 * out["write"]["scale"]
        out[launch] = {"bytes": int(round(fb + wb)), "fetch_bytes": int(round(fb)), "write_bytes": int(round(wb))}
    with open(TRAFFIC_FILES[leg], "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out))


def pmc_copy(ob, ctx):
    """One device-to-device copy of PMC_COPY_BYTES (a blit kernel: the counters' scale) before the probed launches."""
    lib = ob.native.lib()
    src = ob.DeviceArray(ctx, np.uint8, PMC_COPY_BYTES)
    dst = ob.DeviceArray(ctx, np.uint8, PMC_COPY_BYTES)
    ob.native.check(lib.obvhs_memset_device(ctx.handle, src.ptr, 1, PMC_COPY_BYTES), ctx.handle)
    ctx.synchronize()
    ob.native.check(lib.obvhs_memcpy_dtod(ctx.handle, dst.ptr, src.ptr, PMC_COPY_BYTES), ctx.handle)
    ctx.synchronize()


def tune_step(ob, ctx, n_prims, step, reps=3, layouts=("triangle",)):
    """Chooses the traversal LDS stack depth (4, 8, 16), launch form (one lane per ray / chunked) and, when the step
    takes one, the primitive layout (Triangle / RtTriangle rows, the same hits) for the size class of n_prims by
    timing the timed step itself with each (one warm-up, best of `reps`); keeps the fastest depth and form in the
    context's tuning table and returns the layout.  No choice changes a hit."""
    cls = ob.size_class(n_prims)
    ms_by = {}
    for layout in layouts:
        for chunk in (0, 1):
            for depth in (4, 8, 16):
                t = ctx.tune_table()
                t["trav_lds_depth"][cls], t["trav_chunk"][cls] = depth, chunk
                ctx.set_tune_table(t)
                step(layout) if len(layouts) > 1 else step()
                ms = []
                for _ in range(reps):
                    ctx.timer_start()
                    step(layout) if len(layouts) > 1 else step()
                    ms.append(ctx.timer_stop())
                ms_by[(depth, chunk, layout)] = min(ms)
    best = min(ms_by, key=ms_by.get)
    t = ctx.tune_table()
    t["trav_lds_depth"][cls], t["trav_chunk"][cls] = best[0], best[1]
    ctx.set_tune_table(t)
    return {"traversal_lds_depth": best[0], "traversal_chunked": best[1], "layout": best[2], "size_class": cls,
            "step_ms_by_lds_depth_chunked_layout": {f"{d}/{c}/{l}": round(v, 4) for (d, c, l), v in ms_by.items()}}


def kernel_name(ctx, ob, n_prims, mode="closest"):
    t = ctx.tune_table()
    cls = ob.size_class(n_prims)
    form = "k_traverse_chunk" if t["trav_chunk"][cls] else "k_traverse"
    return f"{form}<{mode}, lds {t['trav_lds_depth'][cls]}>"


def leg_counters(leg):
    """The committed issue/cache counters of a leg's traversal launch, or None."""
    try:
        with open(COUNTER_FILES[leg]) as f:
            return json.load(f)
    except (OSError, KeyError, ValueError):
        return None


def roofline(algo_bytes_per_launch, kernel_ms, traffic, kernel, counters=None, **extra):
    """Roofline object of a traversal kernel: achieved algorithmic GB/s against the HBM peak (the only roofline that
    applies: no MFMA work); with the measured traffic and counters, what actually limits the launch."""
    kernel_s = kernel_ms * 1e-3
    achieved = algo_bytes_per_launch / kernel_s / 1e9
    r = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
         "avg_kernel_ms": round(kernel_ms, 4), "algorithmic_bytes_per_launch": int(algo_bytes_per_launch)}
    hbm = traffic / kernel_s / 1e9 if traffic else None
    if hbm is not None:
        r["measured_hbm_gbs"] = round(hbm, 1)
    if counters and hbm is not None:
        r["counters"] = {k: counters[k] for k in ("VALUBusy_pct", "VALUUtilization_pct", "OccupancyPercent",
                                                  "l2_hit_rate") if k in counters}
        r["limiter"] = (f"VALU issue: busy {counters['VALUBusy_pct']:.0f}% of the time with "
                        f"{counters['VALUUtilization_pct']:.0f}% of lanes active (divergent walks); L2 misses "
                        f"move {hbm:.0f} GB/s ({hbm / HBM_PEAK_GBS:.2f} of HBM peak, L2 hit rate "
                        f"{counters['l2_hit_rate']:.2f})")
    elif hbm is not None:
        r["limiter"] = (f"latency: the counted HBM traffic moves {hbm:.0f} GB/s ({hbm / HBM_PEAK_GBS:.2f} of peak); "
                        "node and triangle reads mostly hit L2/MALL" if hbm < 0.5 * HBM_PEAK_GBS
                        else f"HBM bandwidth: {hbm:.0f} GB/s counted")
    else:
        r["limiter"] = "unmeasured (no counter run for this launch yet)"
    r.update(extra)
    return r


def cpu_threads():
    """Host threads for the CPU baseline: the cores this process may run on, at most 16 (one GPU's share of the
    host on the benchmark machines)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def oracle():
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as O
    return O


def cpu_build_baseline(args, threads):
    """Serial and multithreaded oracle builds (the crate's default features and its `parallel` sections) of the
    large-scene recipe at 1/16 scale: a 724 x 724 heightfield + 31,250 thin triangles, pre-splits capped at 1.3n,
    reinsertion 0.05 x [1, 1/3] -> CWBVH."""
    from obvhs_amd import scenes
    O = oracle()
    grid, thin = max(2, args.large_grid // 4), max(1, args.large_thin // 16)
    tris = np.concatenate([scenes.heightfield_scene(grid, 0, 0), scenes.thin_triangles(thin, 0)])
    kw = dict(pre_split=True, reinsertion_batch_ratio=0.05, ratio_sequence=[1.0, 1.0 / 3.0],
              split_max_total=int(len(tris) * 1.3))
    t0 = time.perf_counter()
    b, cw, _ = O.build_cwbvh(tris, 14, 0, 3, **kw)
    serial_s = time.perf_counter() - t0
    with O.threads(threads):
        t0 = time.perf_counter()
        b2, cw2, _ = O.build_cwbvh(tris, 14, 0, 3, **kw)
        mt_s = time.perf_counter() - t0
    leaves = len(b["prim_indices"])
    return {"sample": f"large-scene recipe at 1/16 scale: heightfield {grid}x{grid}x2 + {thin} thin = {len(tris)} "
                      f"tris ({leaves} leaves after pre-splits), pre-splits 1.3n, reinsertion 0.05 x [1, 1/3]",
            "serial_s": round(serial_s, 3), "serial_mtris_per_s": round(leaves / serial_s / 1e6, 3),
            "mt_s": round(mt_s, 3), "mt_mtris_per_s": round(leaves / mt_s / 1e6, 3), "mt_threads": threads,
            "mt_equals_serial": bool(cw2.tobytes() == cw.tobytes())}


def autotune_batch(ob, ctx, tris, offsets, rays, ray_mesh):
    """The tuning run before the batched leg (tune.hip): the radix sort tile for the batch's key count; the PLOC
    search distance of each size class from the class's median-size mesh (the smallest distance whose tree cost is
    within 1% of distance 14's, the preset; Auto builds then resolve it per mesh); the traversal LDS depth on the
    mesh nearest the mean size with its rays repeated 16 times (bench_batch then re-times the depth on the batched
    launch itself, whose class is the mean mesh size's)."""
    t0 = time.perf_counter()
    counts = np.diff(offsets).astype(np.int64)
    sort_items = ob.tune_sort(ctx, int(counts.sum()))
    radius = {}
    for c in range(4):
        idx = np.nonzero([(ob.size_class(k) == c and k >= 2) for k in counts])[0]
        if len(idx):
            i = int(idx[np.argsort(counts[idx])[len(idx) // 2]])
            r, costs = ob.tune_ploc_radius(ctx, tris[offsets[i]:offsets[i + 1]], 3, 0.01, baseline=14)
            radius[str(c)] = {"distance": int(r), "mesh_tris": int(counts[i]),
                              "cost_by_distance": {str(d): round(float(v), 6)
                                                   for d, v in zip((1, 2, 6, 14, 24, 32), costs)}}
    i = int(np.argmin(np.abs(counts - counts.mean())))
    sub = np.ascontiguousarray(tris[offsets[i]:offsets[i + 1]])
    cw = ob.build_cwbvh_from_tris(ctx, sub, ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3))
    sub_rays = np.tile(rays[ray_mesh == i], 16)
    depth = ob.tune_traversal(ctx, cw, sub_rays, cw.reorder_triangles(sub))
    cw.close()
    return {"sort_items": sort_items, "ploc_search_distance_by_size_class": radius,
            "traversal_lds_depth": depth, "traversal_size_class": ob.size_class(int(counts[i])),
            "tuning_s": round(time.perf_counter() - t0, 2)}


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
    tuned = None if args.no_tune else autotune_batch(ob, ctx, tris, offsets, rays, ray_mesh)
    fixed = ob.BvhBuildParams(ploc_search_distance=14, search_depth_threshold=0, max_prims_per_leaf=3)
    params = ob.BvhBuildParams(ploc_search_distance=ob.PlocSearchDistance.Auto, search_depth_threshold=0,
                               max_prims_per_leaf=3)
    log(f"batched build: {m} meshes, {n_tris} triangles")

    def time_build(p):
        bt = ob.build_cwbvh_batch(ctx, d_tris, offsets, p)
        ms = []
        for _ in range(max(1, args.build_reps)):
            ctx.timer_start()
            ob.build_cwbvh_batch(ctx, d_tris, offsets, p, reuse=bt)
            ms.append(ctx.timer_stop())
        return bt, float(np.median(ms))

    batch14, build_ms_r14 = time_build(fixed)
    batch14.close()
    batch, build_ms = time_build(params)
    d_bvh_tris = batch.reorder_triangles(d_tris)
    hits = ob.DeviceArray(ctx, ob.RAY_HIT, n_rays)

    def trace_batch():
        ob.native.check(lib.obvhs_batch_ray_traverse(ctx.handle, batch.handle, d_bvh_tris.ptr, d_rays.ptr,
                                                     d_mesh.ptr, n_rays, hits.ptr), ctx.handle)

    if tuned is not None and not args.lds_depth:
        # The batched launch takes the LDS depth and launch form of its mean mesh size's class: time the launch
        # itself with each (one warm-up and three timed launches each) and keep the fastest for the timed steps.
        cls = ob.size_class(n_tris // m)
        by_depth = {}
        for depth in [(d, c) for c in (0, 1) for d in (4, 8, 16)]:
            t = ctx.tune_table()
            t["trav_lds_depth"][cls], t["trav_chunk"][cls] = depth
            ctx.set_tune_table(t)
            trace_batch()
            ms = []
            for _ in range(3):
                ctx.timer_start()
                trace_batch()
                ms.append(ctx.timer_stop())
            by_depth[depth] = min(ms)
        best = min(by_depth, key=by_depth.get)
        t = ctx.tune_table()
        t["trav_lds_depth"][cls], t["trav_chunk"][cls] = best
        ctx.set_tune_table(t)
        tuned["traversal_lds_depth"], tuned["traversal_chunked"] = best
        tuned["traversal_size_class"] = cls
        tuned["trace_ms_by_lds_depth_and_chunked"] = {f"{d}/{c}": round(v, 4) for (d, c), v in by_depth.items()}
    for _ in range(args.warmup):
        trace_batch()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        trace_batch()
    ctx.synchronize()
    trace_s = (time.perf_counter() - t0) / args.steps
    batch_hits = hits.to_host()
    info = batch.info()
    result = {"workload": f"{m} icospheres (log-uniform 64..65536 tris, rounded down to 20*4^s) = {n_tris} tris; "
                          f"{args.batch_rays_per_mesh} rays per mesh = {n_rays} closest-hit rays",
              "build_ms": round(build_ms, 3), "build_mtris_per_s": round(n_tris / build_ms / 1e3, 2),
              "build_ms_search_distance_14": round(build_ms_r14, 3), "tuning": tuned,
              "trace_mrays_per_s": round(n_rays / trace_s / 1e6, 2), "trace_ms": round(trace_s * 1e3, 4),
              "ploc_passes": info["ploc_passes"], "cwbvh_nodes": info["total_nodes"],
              "hit_fraction": round(float(np.isfinite(batch_hits["t"]).mean()), 4)}
    if args.batch_no_plain:
        batch.close()
        return result

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
    # node and triangle test counts of every ray (the same walks as the batched launch's: identical CWBVHs and
    # hits), for the batched trace's roofline
    nv = ob.DeviceArray(ctx, np.uint32, n_rays)
    tv = ob.DeviceArray(ctx, np.uint32, n_rays)
    for i in range(m):
        o = int(offsets[i])
        ob.native.check(lib.obvhs_cwbvh_ray_traverse_counted(
            ctx.handle, handles[i], C.c_void_p(base_pt + o * tri_sz), C.c_void_p(base_r + i * rpm * ray_sz), rpm,
            C.c_void_p(base_h + i * rpm * hit_sz), C.c_void_p(nv.ptr.value + i * rpm * 4),
            C.c_void_p(tv.ptr.value + i * rpm * 4)), ctx.handle)
    node_tests, tri_tests = int(nv.to_host().astype(np.int64).sum()), int(tv.to_host().astype(np.int64).sum())
    result["roofline"] = roofline(40 * n_rays + 80 * node_tests + 36 * tri_tests, trace_s * 1e3, None,
                                  kernel_name(ctx, ob, n_tris // m), node_tests=node_tests, tri_tests=tri_tests,
                                  bytes_model="40 B per ray + 80 B per node test + 36 B per triangle test",
                                  timing="wall time per batched launch (one kernel), not HIP events")
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps // 4)):
        trace_plain()
    ctx.synchronize()
    plain_trace_s = (time.perf_counter() - t0) / max(1, args.steps // 4)
    for h in handles:
        lib.obvhs_cwbvh_destroy(h)
    batch.close()
    result.update({"plain_build_ms": round(plain_build_ms, 3),
                   "plain_build_mtris_per_s": round(n_tris / plain_build_ms / 1e3, 2),
                   "plain_trace_mrays_per_s": round(n_rays / plain_trace_s / 1e6, 2),
                   "trace_speedup_vs_plain": round(plain_trace_s / trace_s, 2),
                   "build_speedup_vs_plain": round(plain_build_ms / build_ms, 2),
                   "hits_identical_to_plain": same})
    return result


def bench_rebuild(args, ob, ctx, tris, d_tris, fresh_build_ms):
    """Partial rebuild as examples/physics.rs:426-448 runs it, on the heightfield scene: a fraction of the triangles
    moves, their leaves take the new boxes (Bvh2Node::set_aabb), the paths to the root are flagged
    (compute_rebuild_path_flags), the flagged top is rebuilt with PLOC (partial_rebuild) and the CWBVH re-emitted.
    Compared with full_rebuild + re-emit of the same BVH2 and with a fresh build_cwbvh_from_tris."""
    import ctypes as C
    from obvhs_amd import scenes
    from obvhs_amd.test_util import hash_noise
    lib = ob.native.lib()
    n = len(tris)
    moved_ids = np.nonzero(hash_noise(np.arange(n, dtype=np.uint32), 5, 21) < args.rebuild_frac)[0]
    moved = scenes.jitter_triangles(tris[moved_ids], seed=8, frac=0.002)
    verts = np.stack([moved["v0"][:, :3], moved["v1"][:, :3], moved["v2"][:, :3]])
    boxes = np.zeros(len(moved_ids), ob.AABB)
    boxes["min"][:, :3] = verts.min(axis=0)
    boxes["max"][:, :3] = verts.max(axis=0)
    sd = 14
    base = ob.PlocBuilder(ctx).build(sd, d_tris)
    base_nodes, prims = base.nodes, base.primitive_indices
    leaf_of = base.primitives_to_nodes().to_host()
    d_ids = ctx.upload(leaf_of[moved_ids].astype(np.uint32))
    d_boxes = ctx.upload(boxes)
    flags = ob.DeviceArray(ctx, np.uint8, len(base_nodes))
    k = len(moved_ids)

    def fresh():
        b = ob.Bvh2.from_host(ctx, base_nodes, prims)
        ctx.synchronize()
        return b

    part_ms, part_emit_ms, full_emit_ms = [], [], []
    for _ in range(max(2, args.build_reps)):
        b = fresh()
        t0 = time.perf_counter()
        ob.native.check(lib.obvhs_bvh2_set_node_aabbs(ctx.handle, b.handle, d_ids.ptr, d_boxes.ptr, k), ctx.handle)
        ob.native.check(lib.obvhs_bvh2_rebuild_path_flags(ctx.handle, b.handle, d_ids.ptr, k, flags.ptr), ctx.handle)
        ob.native.check(lib.obvhs_bvh2_partial_rebuild(ctx.handle, b.handle, flags.ptr, sd, 64, 0), ctx.handle)
        ctx.synchronize()
        t1 = time.perf_counter()
        cw = ob.bvh2_to_cwbvh(b, 3, True, False)
        ctx.synchronize()
        t2 = time.perf_counter()
        part_ms.append((t1 - t0) * 1e3)
        part_emit_ms.append((t2 - t0) * 1e3)
        cw.close()
        b.close()
        b = fresh()
        t0 = time.perf_counter()
        ob.native.check(lib.obvhs_bvh2_set_node_aabbs(ctx.handle, b.handle, d_ids.ptr, d_boxes.ptr, k), ctx.handle)
        ob.native.check(lib.obvhs_bvh2_full_rebuild(ctx.handle, b.handle, sd, 64, 0), ctx.handle)
        cw = ob.bvh2_to_cwbvh(b, 3, True, False)
        ctx.synchronize()
        full_emit_ms.append((time.perf_counter() - t0) * 1e3)
        cw.close()
        b.close()
    flagged = int(flags.to_host().astype(np.int64).sum())
    med = lambda v: round(float(np.median(v)), 3)
    return {"workload": f"heightfield {n} tris, PLOC r={sd}; {k} triangles ({args.rebuild_frac:.1%}) moved, their "
                        f"leaves re-boxed, {flagged} of {len(base_nodes)} nodes flagged",
            "partial_rebuild_ms": med(part_ms), "partial_rebuild_and_emit_ms": med(part_emit_ms),
            "full_rebuild_and_emit_ms": med(full_emit_ms), "fresh_build_ms": round(fresh_build_ms, 3),
            "timing": "host wall clock per call sequence, synchronized (includes the host reads of level and "
                      "cluster counts)"}


def bench_large(args, ob, ctx):
    """The headline: the large scene (heightfield 2896 x 2896 + 500,000 thin triangles, pre-splits capped at 1.3n,
    reinsertion 0.05 x [1, 1/3]) traced with 4M incoherent closest-hit rays per step; its build and 4M any-hit
    shadow rays beside it; the oracle's traversal of the same CWBVH on the host cores."""
    from obvhs_amd import scenes
    log("generating the large scene")
    tris, rays, shadow, light = scenes.large_workload(args.large_grid, args.large_thin, args.large_rays,
                                                      args.large_rays, seed=0)
    n = len(tris)
    d_tris, d_rays, d_shadow = ctx.upload(tris), ctx.upload(rays), ctx.upload(shadow)
    del tris
    params = ob.BvhBuildParams(pre_split=True, ploc_search_distance=14, search_depth_threshold=0,
                               reinsertion_batch_ratio=0.05, max_prims_per_leaf=3)
    cap = int(n * 1.3)
    seq = [1.0, 1.0 / 3.0]
    tuning = {}
    log(f"large build over {n} triangles")
    cw = ob.build_cwbvh_from_tris(ctx, d_tris, params, split_max_total=cap, ratio_sequence=seq)
    leaves = int(cw.stats.prim_count)
    if not args.no_tune and not args.sort_items:  # the radix sort tile for this key count (same order for all)
        tuning["sort_items"] = ob.tune_sort(ctx, leaves)
    ms = []
    for _ in range(max(1, min(args.build_reps, 3))):
        ctx.timer_start()
        ob.build_cwbvh_from_tris(ctx, d_tris, params, reuse=cw, split_max_total=cap, ratio_sequence=seq)
        ms.append(ctx.timer_stop())
    build_ms = float(np.median(ms))
    # the same build without the quality passes, for reference
    plain_params = ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3)
    plain_cw = ob.build_cwbvh_from_tris(ctx, d_tris, plain_params)
    ctx.timer_start()
    ob.build_cwbvh_from_tris(ctx, d_tris, plain_params, reuse=plain_cw)
    plain_ms = ctx.timer_stop()
    plain_cw.close()
    d_bvh_tris = cw.reorder_triangles(d_tris)
    lib = ob.native.lib()
    nr, ns = len(rays), len(shadow)
    hits = ob.DeviceArray(ctx, ob.RAY_HIT, nr)
    miss = ob.DeviceArray(ctx, np.uint8, ns)

    d_rt = ob.rt_triangles(ctx, d_bvh_tris)  # RtTriangle rows (the crate's demoscene layout): the same hits
    layout = {"value": args.layout if args.layout != "auto" else "triangle"}

    def closest(which=None):
        if (which or layout["value"]) == "rt":
            ob.native.check(lib.obvhs_cwbvh_ray_traverse_rt(ctx.handle, cw.handle, d_rt.ptr, d_rays.ptr, nr,
                                                            hits.ptr), ctx.handle)
        else:
            ob.native.check(lib.obvhs_cwbvh_ray_traverse(ctx.handle, cw.handle, d_bvh_tris.ptr, d_rays.ptr, nr,
                                                         hits.ptr), ctx.handle)

    def anyhit():
        if layout["value"] == "rt":
            ob.native.check(lib.obvhs_cwbvh_ray_traverse_miss_rt(ctx.handle, cw.handle, d_rt.ptr, d_shadow.ptr, ns,
                                                                 miss.ptr), ctx.handle)
        else:
            ob.native.check(lib.obvhs_cwbvh_ray_traverse_miss(ctx.handle, cw.handle, d_bvh_tris.ptr, d_shadow.ptr,
                                                              ns, miss.ptr), ctx.handle)

    if not args.no_tune and not args.lds_depth and args.trav_chunk < 0:
        tuning["closest"] = tune_step(ob, ctx, leaves, closest,
                                      layouts=("triangle", "rt") if args.layout == "auto" else (layout["value"],))
        layout["value"] = tuning["closest"]["layout"]
    if args.pmc_probe == "large":
        pmc_copy(ob, ctx)
        closest()
        ctx.synchronize()
        print(json.dumps({"pmc_probe": "large", "copy_bytes": PMC_COPY_BYTES, "closest_rays": nr,
                          "kernel": kernel_name(ctx, ob, leaves)}), flush=True)
        return None
    # per-ray node / triangle test counts for the byte model (instrumented launch, not timed)
    _, nv, tv = cw.ray_traverse(d_rays, d_bvh_tris, counted=True)
    node_tests, tri_tests = int(nv.sum()), int(tv.sum())
    algo_bytes = 40 * nr + 80 * node_tests + 36 * tri_tests

    for _ in range(args.warmup):
        closest()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        closest()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    # kernel durations: HIP events around each single-kernel launch on the context stream
    kms = []
    for _ in range(max(3, min(args.steps, 10))):
        ctx.timer_start()
        closest()
        kms.append(ctx.timer_stop())
    kernel_ms = float(np.mean(kms))
    closest_kernel = kernel_name(ctx, ob, leaves, "closest, RtTriangle" if layout["value"] == "rt" else "closest")

    for _ in range(args.warmup):
        anyhit()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        anyhit()
    ctx.synchronize()
    ta = (time.perf_counter() - t0) / args.steps
    h, m = hits.to_host(), miss.to_host()

    # the watertight mode (an extra primitive test, reported separately): the same rays and launch choices
    wt_hits = ob.DeviceArray(ctx, ob.RAY_HIT, nr)

    def closest_wt():
        ob.native.check(lib.obvhs_cwbvh_ray_traverse_watertight(ctx.handle, cw.handle, d_bvh_tris.ptr, d_rays.ptr,
                                                                nr, wt_hits.ptr), ctx.handle)

    for _ in range(args.warmup):
        closest_wt()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        closest_wt()
    ctx.synchronize()
    tw = (time.perf_counter() - t0) / args.steps
    hw = wt_hits.to_host()
    st = cw.stats
    result = {
        "value": nr * args.steps / wall / 1e6, "wall": wall,
        "workload": f"large scene: heightfield {args.large_grid}x{args.large_grid}x2 + {args.large_thin} thin = {n} "
                    f"tris ({leaves} leaves after pre-splits capped at 1.3n), reinsertion 0.05 x [1, 1/3]; "
                    f"{nr} incoherent closest-hit rays per step ({layout['value']} primitives)",
        "roofline": roofline(algo_bytes, kernel_ms, traffic_per_launch("large"), closest_kernel,
                             counters=leg_counters("large"),
                             node_tests=node_tests, tri_tests=tri_tests,
                             bytes_model="40 B per ray + 80 B per node test + 36 B per triangle test"),
        "large": {"triangles": n, "leaves_after_splits": leaves, "build_ms": round(build_ms, 3),
                  "build_mtris_per_s": round(leaves / build_ms / 1e3, 2),
                  "build_mtris_per_s_input_triangles": round(n / build_ms / 1e3, 2),
                  "build_ms_without_quality_passes": round(plain_ms, 3),
                  "closest_mrays_per_s": round(nr * args.steps / wall / 1e6, 2),
                  "anyhit_mrays_per_s": round(ns / ta / 1e6, 2), "anyhit_rays": ns,
                  "closest_watertight_mrays_per_s": round(nr / tw / 1e6, 2),
                  "watertight_hits_differing_from_crate_test": int((hw["primitive_id"] != h["primitive_id"]).sum()),
                  "closest_hit_fraction": round(float(np.isfinite(h["t"]).mean()), 4),
                  "shadowed_fraction": round(float((m == 0).mean()), 4), "cwbvh_nodes": int(st.cwbvh_node_count),
                  "tuning": tuning},
    }
    if not args.no_cpu_baseline:
        threads = cpu_threads()
        log(f"cpu baseline: oracle traversal of the large CWBVH, {threads} threads")
        O = oracle()
        cw_host, bvh_tris_host = cw.nodes, d_bvh_tris.to_host()
        sample = rays[:: max(1, args.cpu_ray_stride)]
        O.traverse(cw_host, bvh_tris_host, sample[:4096], 0, threads=threads)  # warm-up
        t0 = time.perf_counter()
        cpu_hits = O.traverse(cw_host, bvh_tris_host, sample, 0, threads=threads)
        cpu_s = time.perf_counter() - t0
        gpu_same = bool(cpu_hits.tobytes() == h[:: max(1, args.cpu_ray_stride)].tobytes())
        result["cpu_trace"] = {"value": len(sample) / cpu_s / 1e6, "threads": threads, "rays": len(sample),
                               "seconds": cpu_s, "hits_equal_gpu": gpu_same}
    cw.close()
    return result


def bench_heightfield(args, ob, ctx):
    """The heightfield scene: 724 x 724 grid + 20,000 boxes = 1,288,352 triangles, PLOC r=14 -> CWBVH; a step is the
    1080p primary launch then the bounce launch (closest hit); the build timed beside it."""
    from obvhs_amd import scenes
    log("generating the heightfield scene")
    tris, prim_rays = scenes.heightfield_workload(args.grid, args.boxes, args.width, args.height, seed=0)
    n_tris = len(tris)
    d_tris = ctx.upload(tris)
    params = ob.BvhBuildParams(ploc_search_distance=14, search_depth_threshold=0, max_prims_per_leaf=3)
    tuning = {}
    if not args.no_tune and not args.sort_items:
        tuning["sort_items"] = ob.tune_sort(ctx, n_tris)
    log(f"building CWBVH over {n_tris} triangles")
    cw = ob.build_cwbvh_from_tris(ctx, d_tris, params)
    ctx.synchronize()
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

    def step():
        trace(d_prim, len(prim_rays), hits)
        trace(d_bounce, len(bounce), bhits)

    if not args.no_tune and not args.lds_depth and args.trav_chunk < 0:
        tuning["step"] = tune_step(ob, ctx, n_tris, step)
    if args.pmc_probe == "heightfield":
        pmc_copy(ob, ctx)
        step()
        ctx.synchronize()
        print(json.dumps({"pmc_probe": "heightfield", "copy_bytes": PMC_COPY_BYTES, "primary_rays": len(prim_rays),
                          "bounce_rays": len(bounce), "kernel": kernel_name(ctx, ob, n_tris)}), flush=True)
        return None
    _, nv_p, tv_p = cw.ray_traverse(d_prim, d_bvh_tris, counted=True)
    _, nv_b, tv_b = cw.ray_traverse(d_bounce, d_bvh_tris, counted=True)
    node_tests = int(nv_p.sum()) + int(nv_b.sum())
    tri_tests = int(tv_p.sum()) + int(tv_b.sum())
    algo_bytes = 40 * n_rays + 80 * node_tests + 36 * tri_tests
    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    kp, kb = [], []
    for _ in range(max(3, min(args.steps, 10))):
        ctx.timer_start()
        trace(d_prim, len(prim_rays), hits)
        kp.append(ctx.timer_stop())
        ctx.timer_start()
        trace(d_bounce, len(bounce), bhits)
        kb.append(ctx.timer_stop())
    kernel_ms = float(np.mean(kp) + np.mean(kb)) / 2.0
    out = {"workload": f"heightfield {args.grid}x{args.grid}x2 + {args.boxes} boxes = {n_tris} tris; "
                       f"{len(prim_rays)} primary + {len(bounce)} bounce closest-hit rays per step",
           "closest_mrays_per_s": round(n_rays * args.steps / wall / 1e6, 2),
           "ms_per_step": round(wall / args.steps * 1e3, 4),
           "roofline": roofline(algo_bytes / 2.0, kernel_ms, traffic_per_launch("heightfield"),
                                kernel_name(ctx, ob, n_tris), counters=leg_counters("heightfield"),
                                avg_kernel_ms_by_launch={"primary": round(float(np.mean(kp)), 4),
                                                         "bounce": round(float(np.mean(kb)), 4)},
                                node_tests=node_tests, tri_tests=tri_tests),
           "build": {"ms": round(build_med, 3), "mtris_per_s": round(n_tris / build_med / 1e3, 2),
                     "all_ms": [round(x, 3) for x in build_ms], "ploc_passes": int(stats.ploc_iterations),
                     "ploc_cluster_sum": int(stats.ploc_cluster_sum), "cwbvh_nodes": int(stats.cwbvh_node_count)},
           "tuning": tuning}
    if args.rebuild_frac > 0:
        out["rebuild"] = bench_rebuild(args, ob, ctx, tris, d_tris, build_med)
    cw.close()
    return out


def bench_plain(args, ob, ctx):
    """The plain 65,536-triangle soup the crate's CPU path runs: PLOC r=14 -> CWBVH (max 3 per leaf), 256 x 256
    pinhole rays, closest hit; GPU and the oracle (serial build, serial and multithreaded traversal)."""
    from obvhs_amd import scenes
    O = oracle()
    tris, rays = scenes.plain_workload()
    params = ob.BvhBuildParams(ploc_search_distance=14, search_depth_threshold=0, max_prims_per_leaf=3)
    d_tris, d_rays = ctx.upload(tris), ctx.upload(rays)
    cw = ob.build_cwbvh_from_tris(ctx, d_tris, params)
    ms = []
    for _ in range(max(1, args.build_reps)):
        ctx.timer_start()
        ob.build_cwbvh_from_tris(ctx, d_tris, params, reuse=cw)
        ms.append(ctx.timer_stop())
    build_ms = float(np.median(ms))
    d_bvh_tris = cw.reorder_triangles(d_tris)
    hits = ob.DeviceArray(ctx, ob.RAY_HIT, len(rays))
    lib = ob.native.lib()

    def step():
        ob.native.check(lib.obvhs_cwbvh_ray_traverse(ctx.handle, cw.handle, d_bvh_tris.ptr, d_rays.ptr, len(rays),
                                                     hits.ptr), ctx.handle)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    trace_s = (time.perf_counter() - t0) / args.steps
    gpu_hits = hits.to_host()
    out = {"workload": f"soup {len(tris)} tris (PLOC r=14, max 3 per leaf); {len(rays)} pinhole closest-hit rays",
           "gpu_build_ms": round(build_ms, 4), "gpu_build_mtris_per_s": round(len(tris) / build_ms / 1e3, 2),
           "gpu_trace_mrays_per_s": round(len(rays) / trace_s / 1e6, 2)}
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        _, rcw, rprims = O.build_cwbvh(tris, 14, 0, 3)
        cpu_build_s = time.perf_counter() - t0
        bt = tris[rprims]
        t0 = time.perf_counter()
        cpu_hits = O.traverse(rcw, bt, rays, 0)
        cpu_trace_s = time.perf_counter() - t0
        threads = cpu_threads()
        t0 = time.perf_counter()
        O.traverse(rcw, bt, rays, 0, threads=threads)
        cpu_mt_s = time.perf_counter() - t0
        out.update({"cpu_build_mtris_per_s_serial": round(len(tris) / cpu_build_s / 1e6, 3),
                    "cpu_trace_mrays_per_s_serial": round(len(rays) / cpu_trace_s / 1e6, 3),
                    "cpu_trace_mrays_per_s_mt": round(len(rays) / cpu_mt_s / 1e6, 3), "cpu_threads": threads,
                    "cwbvh_equal_oracle": bool(cw.nodes.tobytes() == rcw.tobytes()),
                    "hits_equal_oracle": bool(gpu_hits.tobytes() == cpu_hits.tobytes())})
    cw.close()
    return out


def bench_serve(args, ob, ctx):
    """Serving: 64 resident scenes (100K-4M triangles, heightfields and soups), an open-loop Poisson request stream
    (90% ray batches of log-uniform size in [1K, 1M] rays, half coherent pinhole and half incoherent; 10% refit or
    rebuild after a +-1% vertex jitter), replayed through the server; beside it the plain form on the same trace
    requests (one launch and one synchronisation per request, back to back)."""
    import ctypes as C
    from obvhs_amd import scenes
    rng = np.random.default_rng(7)
    log(f"generating {args.serve_scenes} serving scenes")
    sc = scenes.serving_scenes(args.serve_scenes, seed=0)
    params = ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3)
    server = ob.Server(0, hbm_budget_bytes=200 << 30, max_inflight=args.serve_inflight,
                       update_streams=args.serve_update_streams)
    d_jit = []
    plain = []
    for i, (tris, kind) in enumerate(sc):
        d = ctx.upload(tris)
        server.add_scene(d, params)
        cw = ob.build_cwbvh_from_tris(ctx, d, params)
        plain.append((cw, cw.reorder_triangles(d)))
        d_jit.append(ctx.upload(scenes.jitter_triangles(tris, seed=i, frac=0.01)))
        del d
    # ray pools per scene kind: 1M coherent pinhole rays, 1M incoherent rays
    pool = 1 << 20
    side = 1024
    hf_pin = scenes.pinhole_rays(side, side, eye=(0.5, 0.45, -0.35), target=(0.5, 0.0, 0.55))
    soup_pin = scenes.pinhole_rays(side, side, eye=(0.5, 0.5, -1.5))
    k = np.arange(pool, dtype=np.uint32)
    from obvhs_amd.test_util import hash_noise, uniform_sample_sphere
    o = np.stack([hash_noise(k, a, 13000) for a in range(3)], axis=1)
    d = uniform_sample_sphere(hash_noise(k, 3, 13000), hash_noise(k, 4, 13000))
    incoherent = ob.make_rays(o, d, 0.0, 1e30)
    hf0 = plain[0]
    bhits = hf0[0].ray_traverse(hf_pin, hf0[1])
    hf_bounce = scenes.bounce_rays(hf_pin, bhits, hf0[1].to_host(), seed=3)
    hf_bounce = np.concatenate([hf_bounce, incoherent])[:pool]
    pools = {("heightfield", 0): ctx.upload(hf_pin), ("heightfield", 1): ctx.upload(hf_bounce),
             ("soup", 0): ctx.upload(soup_pin), ("soup", 1): ctx.upload(incoherent)}
    arena_n = 1 << 25
    out_arena = ob.DeviceArray(ctx, ob.RAY_HIT, arena_n)
    ray_sz, hit_sz = ob.RAY.itemsize, ob.RAY_HIT.itemsize

    def make_requests(count, rate):
        reqs, t, cursor = [], 0.0, 0
        for q in range(count):
            t += rng.exponential(1.0 / rate) if rate > 0 else 0.0
            s = int(rng.integers(0, len(sc)))
            if rng.random() < 0.1:
                reqs.append(dict(arrival_ns=int(t * 1e9), kind=ob.api.REQ_REFIT if rng.random() < 0.5
                                 else ob.api.REQ_REBUILD, scene=s, tris=d_jit[s]))
                continue
            n = int(np.exp(rng.uniform(np.log(1000.0), np.log(1 << 20))))
            which = int(rng.integers(0, 2))
            src = pools[(sc[s][1], which)]
            off = int(rng.integers(0, pool - n + 1))
            if cursor + n > arena_n:
                cursor = 0
            rays = _Slice(src.ptr.value + off * ray_sz, n)
            out = _Slice(out_arena.ptr.value + cursor * hit_sz, n)
            cursor += n
            reqs.append(dict(arrival_ns=int(t * 1e9), kind=ob.api.REQ_TRACE_CLOSEST, scene=s, rays=rays, out=out,
                             n=n, which=which))
        return reqs

    def summarize(reqs, infos):
        tr = [(q, x) for q, x in zip(reqs, infos) if q["kind"] == ob.api.REQ_TRACE_CLOSEST]
        up = [(q, x) for q, x in zip(reqs, infos) if q["kind"] != ob.api.REQ_TRACE_CLOSEST]
        lat = np.array([(x["done_ns"] - x["submit_ns"]) * 1e-6 for _, x in tr])
        ulat = np.array([(x["done_ns"] - x["submit_ns"]) * 1e-6 for _, x in up]) if up else np.zeros(1)
        rays = sum(q["n"] for q, _ in tr)
        t0 = min(x["submit_ns"] for x in infos)
        t1 = max(x["done_ns"] for x in infos)
        return {"requests": len(reqs), "trace_requests": len(tr), "updates": len(up),
                "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "update_p50_ms": round(float(np.percentile(ulat, 50)), 3),
                "update_p99_ms": round(float(np.percentile(ulat, 99)), 3),
                "mrays_per_s": round(rays / ((t1 - t0) * 1e-9) / 1e6, 2),
                "mean_requests_per_launch": round(float(np.mean([x["batch_requests"] for _, x in tr])), 2),
                "errors": int(sum(1 for x in infos if x["status"] != "OBVHS_OK"))}

    log("serving: Poisson replay")
    reqs = make_requests(args.serve_requests, args.serve_rate)
    poisson = summarize(reqs, server.replay(reqs))
    log("serving: burst replay (capacity)")
    burst_reqs = make_requests(args.serve_requests, 0.0)
    burst = summarize(burst_reqs, server.replay(burst_reqs))
    # rate sweep: the same request list with its Poisson arrival times scaled to each rate
    rates = [float(x) for x in args.serve_sweep.split(",")] if args.serve_sweep else []
    sweep_served = {}
    for rate in rates:
        log(f"serving: sweep at {rate:g} requests/s")
        scaled = [dict(q, arrival_ns=int(q["arrival_ns"] * args.serve_rate / rate)) for q in reqs]
        sweep_served[rate] = summarize(scaled, server.replay(scaled))
    st = server.stats()
    server.close()
    # plain form: the same requests one at a time through the single-request API (traces: one launch and a
    # synchronisation each; refits: refit the BVH2 + convert + reorder; rebuilds: build + reorder)
    lib = ob.native.lib()
    base = [cw for cw, _ in plain]
    cur = [cw for cw, _ in plain]
    btris = [bt for _, bt in plain]
    service = []
    ctx.synchronize()
    for q in reqs:
        s_ = q["scene"]
        t0 = time.perf_counter()
        if q["kind"] == ob.api.REQ_TRACE_CLOSEST:
            ob.native.check(lib.obvhs_cwbvh_ray_traverse(ctx.handle, cur[s_].handle, btris[s_].ptr,
                                                         C.c_void_p(q["rays"].ptr.value), q["n"],
                                                         C.c_void_p(q["out"].ptr.value)), ctx.handle)
        elif q["kind"] == ob.api.REQ_REFIT:
            base[s_].bvh2.refit_from_triangles(q["tris"])
            cur[s_] = ob.bvh2_to_cwbvh(base[s_].bvh2, 3, True, False)
            btris[s_] = cur[s_].reorder_triangles(q["tris"])
        else:
            base[s_] = ob.build_cwbvh_from_tris(ctx, q["tris"], params, reuse=base[s_])
            cur[s_] = base[s_]
            btris[s_] = cur[s_].reorder_triangles(q["tris"])
        ctx.synchronize()
        service.append(time.perf_counter() - t0)
    traces = [(q, sv) for q, sv in zip(reqs, service) if q["kind"] == ob.api.REQ_TRACE_CLOSEST]
    plain_rays = sum(q["n"] for q, _ in traces)

    def plain_at(scale):
        """The plain form under the request list's arrivals scaled by `scale` (one request at a time, first come
        first served, each taking its measured service time): trace p50 / p99 and rays per second of the span."""
        free_at, lat, first = 0.0, [], None
        for q, sv in zip(reqs, service):
            a = q["arrival_ns"] * 1e-9 * scale
            first = a if first is None else first
            start = max(a, free_at)
            free_at = start + sv
            if q["kind"] == ob.api.REQ_TRACE_CLOSEST:
                lat.append((free_at - a) * 1e3)
        return {"p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "mrays_per_s": round(plain_rays / (free_at - first) / 1e6, 2)}

    same = plain_at(1.0)
    plain = {"mrays_per_s": round(plain_rays / sum(service) / 1e6, 2),
             "trace_only_mrays_per_s": round(plain_rays / sum(sv for _, sv in traces) / 1e6, 2),
             "p50_ms_same_arrivals": same["p50_ms"], "p99_ms_same_arrivals": same["p99_ms"]}
    out = {"workload": f"{len(sc)} resident scenes ({sum(len(t) for t, _ in sc)} tris, 100K-4M each); "
                       f"{args.serve_requests} requests, 90% ray batches of 1K-1M rays, 10% refit/rebuild",
           "poisson_rate_per_s": args.serve_rate, "served": poisson, "served_burst": burst, "plain": plain,
           "served_burst_vs_plain_mrays": round(burst["mrays_per_s"] / plain["mrays_per_s"], 2),
           "launches": st["launches"], "resident_gb": round(st["resident_bytes"] / 2 ** 30, 2)}
    if rates:
        slo = args.serve_slo_ms
        rows = []
        for rate in rates:
            pl = plain_at(args.serve_rate / rate)
            sv = sweep_served[rate]
            rows.append({"rate_per_s": rate, "served_p50_ms": sv["p50_ms"], "served_p99_ms": sv["p99_ms"],
                         "served_mrays_per_s": sv["mrays_per_s"],
                         "served_requests_per_launch": sv["mean_requests_per_launch"],
                         "plain_p50_ms": pl["p50_ms"], "plain_p99_ms": pl["p99_ms"],
                         "plain_mrays_per_s": pl["mrays_per_s"]})
        best_s = max((r["served_mrays_per_s"] for r in rows if r["served_p99_ms"] < slo), default=0.0)
        best_p = max((r["plain_mrays_per_s"] for r in rows if r["plain_p99_ms"] < slo), default=0.0)
        out["sweep"] = {"slo_p99_ms": slo, "rows": rows,
                        "served_best_mrays_per_s_at_slo": best_s, "plain_best_mrays_per_s_at_slo": best_p,
                        "served_vs_plain_at_slo": round(best_s / best_p, 2) if best_p > 0 else None,
                        "note": "the same request list at each rate (Poisson arrival times scaled); plain = one "
                                "request at a time, first come first served, with each request's measured service "
                                "time"}
    return out


class _Slice:
    """A device pointer + length for requests that name part of a pool."""

    def __init__(self, ptr, count):
        import ctypes as C
        self.ptr = C.c_void_p(ptr)
        self.count = count


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
    ap.add_argument("--no-tune", action="store_true", help="keep the default tuning table (no tuning runs)")
    ap.add_argument("--sort-items", type=int, default=0,
                    help="set the sort tile (4, 8, 16) for every size class instead of measuring it")
    ap.add_argument("--lds-depth", type=int, default=0,
                    help="set the traversal LDS depth (4, 8, 16) for every size class instead of measuring it")
    ap.add_argument("--trav-chunk", type=int, default=-1, choices=(-1, 0, 1),
                    help="set chunked single-BVH traversal launches (0 or 1) for every size class instead of "
                         "measuring it (-1)")
    ap.add_argument("--layout", choices=("auto", "triangle", "rt"), default="auto",
                    help="primitive rows of the large scene's traversal: Triangle, RtTriangle, or timed (auto)")
    ap.add_argument("--pmc-probe", choices=("large", "heightfield"), default=None,
                    help="one copy kernel, then the leg's traversal launches of one step, then exit (counter runs)")
    ap.add_argument("--traffic-from", nargs=3, metavar=("LEG", "FETCH_CSV", "WRITE_CSV"),
                    help="write a leg's HBM traffic file from its two --pmc-probe counter runs, then exit")
    ap.add_argument("--cpu-ray-stride", type=int, default=1,
                    help="the CPU baseline traces every k-th ray of the large scene (1: all of them)")
    ap.add_argument("--batch-meshes", type=int, default=4096, help="0 skips the batched leg")
    ap.add_argument("--batch-rays-per-mesh", type=int, default=1024)
    ap.add_argument("--batch-no-plain", action="store_true", help="skip the per-mesh comparison of the batched leg")
    ap.add_argument("--rebuild-frac", type=float, default=0.01, help="0 skips the partial-rebuild leg")
    ap.add_argument("--large-grid", type=int, default=2896)
    ap.add_argument("--large-thin", type=int, default=500000)
    ap.add_argument("--large-rays", type=int, default=1 << 22)
    ap.add_argument("--no-heightfield", action="store_true", help="skip the heightfield leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain 65,536-triangle leg")
    ap.add_argument("--serve-scenes", type=int, default=64, help="0 skips the serving leg")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-sweep", default="250,500,1000,1500,2000,3000,4000,4500,5000,6000",
                    help="request rates of the serving sweep (empty: no sweep)")
    ap.add_argument("--serve-slo-ms", type=float, default=5.0)
    ap.add_argument("--serve-inflight", type=int, default=3)
    ap.add_argument("--serve-update-streams", type=int, default=3)
    args = ap.parse_args()
    if args.traffic_from:
        write_traffic(*args.traffic_from)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "design-only, see the design notes)"}))
        sys.exit(2)

    import obvhs_amd as ob

    ctx = ob.Context(0)
    fixed = {}
    if args.sort_items or args.lds_depth or args.trav_chunk >= 0:  # fixed entries (profiling runs repeat a choice)
        t = ctx.tune_table()
        if args.trav_chunk >= 0:
            t["trav_chunk"] = [args.trav_chunk] * 4
            fixed["traversal_chunked"] = args.trav_chunk
        if args.sort_items:
            t["sort_items"] = [args.sort_items] * 4
            fixed["sort_items"] = args.sort_items
        if args.lds_depth:
            t["trav_lds_depth"] = [args.lds_depth] * 4
            fixed["traversal_lds_depth"] = args.lds_depth
        ctx.set_tune_table(t)
    if args.pmc_probe == "heightfield":
        bench_heightfield(args, ob, ctx)
        return

    head = bench_large(args, ob, ctx)
    if head is None:  # counter probe
        return
    out = {
        "metric": METRIC,
        "value": round(head["value"], 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["wall"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded heightfield + long thin triangles; uniform-sphere origins and directions)",
        "config": {"workload": head["workload"], "ploc_search_distance": 14, "max_prims_per_leaf": 3,
                   "pre_splits": "1.3n", "reinsertion": "0.05 x [1, 1/3]"},
        "roofline": head["roofline"],
        "large": head["large"],
    }
    if fixed:
        out["tuning_fixed"] = fixed
    if not args.no_heightfield:
        out["heightfield"] = bench_heightfield(args, ob, ctx)
    if not args.no_plain:
        out["plain"] = bench_plain(args, ob, ctx)
    if args.batch_meshes > 0:
        out["batched"] = bench_batch(args, ob, ctx)
    if args.serve_scenes > 0:
        out["serving"] = bench_serve(args, ob, ctx)
    if not args.no_cpu_baseline:
        ct = head["cpu_trace"]
        threads = cpu_threads()
        log(f"cpu baseline: oracle builds, serial and {threads} threads")
        cb = cpu_build_baseline(args, threads)
        out["cpu_baseline"] = {
            "value": round(ct["value"], 3), "unit": "Mrays/s", "cores": ct["threads"], "kind": "port",
            "sample": f"all {ct['rays']} closest-hit rays of the headline step (large scene), traced through the "
                      f"GPU-built CWBVH by the oracle (C++ restatement of the crate's traversal, SSE semantics) on "
                      f"{ct['threads']} host threads in {ct['seconds']:.2f} s; hits equal the GPU's: "
                      f"{ct['hits_equal_gpu']}",
            "build": cb}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
