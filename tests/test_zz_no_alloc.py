"""obvhs_cwbvh_ray_traverse_noalloc: closest-hit traversal whose stack spill lives in the library's static device
array instead of memory the context allocates.  Its hits must equal obvhs_cwbvh_ray_traverse's byte for byte, on
two calls in a row (the second smaller), for both launch forms (one lane per ray, and chunked ray rows), with a
tuning table entry below its LDS minimum (it then keeps 8 entries in LDS), and the device's free memory must not
change across its calls."""
import ctypes as C

import numpy as np
import pytest

import obvhs_amd as ob
from obvhs_amd import native, scenes
from obvhs_amd.test_util import hash_noise, uniform_sample_sphere

pytestmark = pytest.mark.gpu


def _noalloc():
    f = native.lib().obvhs_cwbvh_ray_traverse_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]
    return f


def _device_free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _incoherent_rays(n, seed):
    k = np.arange(n, dtype=np.uint32)
    o = np.stack([hash_noise(k, a, seed) for a in range(3)], axis=1) * np.float32(0.6) + np.float32(0.2)
    d = uniform_sample_sphere(hash_noise(k, 3, seed), hash_noise(k, 4, seed))
    return ob.make_rays(o, d, 0.0, 1e30)


def test_noalloc_traverse_equals_allocating_entry():
    print("[case] noalloc traverse: 1.5M chunked rays then 100K one-lane rays", flush=True)
    fn = _noalloc()
    lib = native.lib()
    with ob.Context(0) as ctx:
        tris = np.concatenate([scenes.heightfield_scene(180, 300, seed=4), scenes.thin_triangles(2000, seed=4)])
        cw = ob.build_cwbvh_from_tris(ctx, tris, ob.BvhBuildParams(ploc_search_distance=14, max_prims_per_leaf=3))
        d_bvh_tris = cw.reorder_triangles(ctx.upload(tris))
        cls = ob.size_class(len(tris))
        cases = [(_incoherent_rays(1_500_000, 31), 1, 4), (_incoherent_rays(100_000, 32), 0, 8)]
        for rays, chunk, depth in cases:
            print(f"[case] n={len(rays)} chunked={chunk} lds={depth}", flush=True)
            t = ctx.tune_table()
            t["trav_chunk"][cls], t["trav_lds_depth"][cls] = chunk, depth
            ctx.set_tune_table(t)
            d_rays = ctx.upload(rays)
            got = ob.DeviceArray(ctx, ob.RAY_HIT, len(rays))
            want = ob.DeviceArray(ctx, ob.RAY_HIT, len(rays))
            ctx.synchronize()
            free0 = _device_free_bytes()
            native.check(fn(ctx.handle, cw.handle, d_bvh_tris.ptr, d_rays.ptr, len(rays), got.ptr), ctx.handle)
            ctx.synchronize()
            assert _device_free_bytes() == free0
            native.check(lib.obvhs_cwbvh_ray_traverse(ctx.handle, cw.handle, d_bvh_tris.ptr, d_rays.ptr, len(rays),
                                                      want.ptr), ctx.handle)
            g, w = got.to_host(), want.to_host()
            assert g.tobytes() == w.tobytes()
            assert np.isfinite(g["t"]).mean() > 0.15  # rays from inside the scene box, about a quarter hit
        # argument checks: no launch
        assert fn(ctx.handle, cw.handle, d_bvh_tris.ptr, None, 5, None) == native.STATUS_CODE["OBVHS_ERR_INVALID_ARG"]
        assert fn(ctx.handle, cw.handle, d_bvh_tris.ptr, d_rays.ptr, 0, got.ptr) == 0
        assert (fn(ctx.handle, cw.handle, d_bvh_tris.ptr, d_rays.ptr, (1 << 31) + 1, got.ptr)
                == native.STATUS_CODE["OBVHS_ERR_CAPACITY"])
        cw.close()
