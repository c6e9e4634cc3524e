"""Host-side mirror of the obvhs hot-path API over the HIP library.

Same names and argument meaning as the crate (PlocBuilder, PlocSearchDistance, SortPrecision, BvhBuildParams, Bvh2,
bvh2_to_cwbvh, build_cwbvh_from_tris, CwBvh.ray_traverse / ray_traverse_miss), batched over arrays: the reference
traverses one ray per call with a closure (cwbvh/mod.rs:169-225); here a call traverses a whole ray array with the
intersector fixed to Triangle::intersect (triangle.rs:35-76).  Reference panics surface as ObvhsError with the
matching status.  Every call runs on the GPU through libobvhs_hip.so; there is no CPU fallback.
"""
import ctypes as C
import enum

import numpy as np

from . import native
from .native import BuildParams, BuildStats, ObvhsError, check, lib
from .types import (AABB, BVH2_NODE, BVH2_NODE_SMALL, CWBVH_NODE, RAY, RAY_HIT, RT_TRIANGLE, TRIANGLE)


class PlocSearchDistance(enum.IntEnum):
    """ploc/mod.rs:534-562.  Auto (0) is not a distance of the crate: builds given it use the distance of the
    context's tuning table for the BVH's size class (see tune_ploc_radius); from_u32 keeps the crate's conversion
    and rejects 0."""
    Auto = 0
    Minimum = 1
    VeryLow = 2
    Low = 6
    Medium = 14
    High = 24
    VeryHigh = 32

    @classmethod
    def from_u32(cls, v):
        try:
            if int(v) == 0:
                raise ValueError
            return cls(int(v))
        except ValueError:
            raise ValueError(f"Invalid value for PlocSearchDistance: {v}") from None  # ploc/mod.rs:559

    @classmethod
    def for_build(cls, v):
        """A builder argument: a distance of the crate, or Auto."""
        return cls.Auto if int(v) == 0 else cls.from_u32(v)


class SortPrecision(enum.IntEnum):
    """ploc/mod.rs:657-661 (value = key bits)."""
    U128 = 128
    U64 = 64


class BvhBuildParams:
    """lib.rs:207-306 with its six presets."""

    def __init__(self, pre_split=False, ploc_search_distance=PlocSearchDistance.Medium, search_depth_threshold=0,
                 reinsertion_batch_ratio=0.0, post_collapse_reinsertion_batch_ratio_multiplier=0.0,
                 sort_precision=SortPrecision.U64, max_prims_per_leaf=1, collapse_traversal_cost=1.0):
        self.pre_split = bool(pre_split)
        self.ploc_search_distance = PlocSearchDistance(ploc_search_distance)  # Auto (0) allowed
        self.search_depth_threshold = int(search_depth_threshold)
        self.reinsertion_batch_ratio = float(reinsertion_batch_ratio)
        self.post_collapse_reinsertion_batch_ratio_multiplier = float(post_collapse_reinsertion_batch_ratio_multiplier)
        self.sort_precision = SortPrecision(sort_precision)
        self.max_prims_per_leaf = int(max_prims_per_leaf)
        self.collapse_traversal_cost = float(collapse_traversal_cost)

    def to_c(self):
        return BuildParams(int(self.pre_split), int(self.ploc_search_distance), self.search_depth_threshold,
                           self.reinsertion_batch_ratio, self.post_collapse_reinsertion_batch_ratio_multiplier,
                           int(self.sort_precision), self.max_prims_per_leaf, self.collapse_traversal_cost)

    def __repr__(self):
        return f"BvhBuildParams({self.__dict__})"

    @classmethod
    def fastest_build(cls):
        return cls(False, PlocSearchDistance.Minimum, 0, 0.0, 0.0, SortPrecision.U64, 1, 1.0)

    @classmethod
    def very_fast_build(cls):
        return cls(False, PlocSearchDistance.Minimum, 0, 0.01, 0.0, SortPrecision.U64, 8, 3.0)

    @classmethod
    def fast_build(cls):
        return cls(False, PlocSearchDistance.Low, 2, 0.02, 0.0, SortPrecision.U64, 8, 3.0)

    @classmethod
    def medium_build(cls):
        return cls(False, PlocSearchDistance.Medium, 3, 0.05, 2.0, SortPrecision.U64, 8, 3.0)

    @classmethod
    def slow_build(cls):
        return cls(True, PlocSearchDistance.High, 2, 0.2, 2.0, SortPrecision.U128, 8, 3.0)

    @classmethod
    def very_slow_build(cls):
        return cls(True, PlocSearchDistance.Medium, 1, 1.0, 1.0, SortPrecision.U128, 8, 3.0)


class Context:
    """A device + stream.  All builds and traversals of this context are ordered on its stream."""

    def __init__(self, device=0, stream=None):
        h = C.c_void_p()
        check(lib().obvhs_context_create(int(device), stream, C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().obvhs_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self):
        check(lib().obvhs_context_synchronize(self.handle), self.handle)

    def set_trace(self, enable=True):
        check(lib().obvhs_context_set_trace(self.handle, int(bool(enable))), self.handle)

    def tune_table(self):
        """The context's tuning table: per size class (see size_class) the PLOC distance of Auto builds, the radix
        sort tile (keys per thread), the traversal LDS stack depth and whether large single-BVH traversal launches
        run as wave-owned ray rows with lanes refilled (trav_chunk, 0 or 1)."""
        t = native.TuneTable()
        check(lib().obvhs_tune_get(self.handle, C.byref(t)), self.handle)
        return {"ploc_radius": list(t.ploc_radius), "sort_items": list(t.sort_items),
                "trav_lds_depth": list(t.trav_lds_depth), "trav_chunk": list(t.trav_chunk)}

    def set_tune_table(self, table):
        t = native.TuneTable()
        for k in ("ploc_radius", "sort_items", "trav_lds_depth", "trav_chunk"):
            getattr(t, k)[:] = [int(v) for v in table.get(k, [0] * 4)]
        check(lib().obvhs_tune_set(self.handle, C.byref(t)), self.handle)

    def upload(self, array):
        return DeviceArray.from_host(self, array)

    def empty(self, dtype, count):
        return DeviceArray(self, dtype, count)

    def timer_start(self):
        check(lib().obvhs_timer_start(self.handle), self.handle)

    def timer_stop(self):
        ms = C.c_float()
        check(lib().obvhs_timer_stop(self.handle, C.byref(ms)), self.handle)
        return float(ms.value)


class DeviceArray:
    """A typed device allocation owned by a context (keeps inputs resident in HBM)."""

    def __init__(self, ctx, dtype, count):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.dtype.itemsize * self.count
        p = C.c_void_p()
        check(lib().obvhs_device_malloc(ctx.handle, max(self.nbytes, 1), C.byref(p)), ctx.handle)
        self.ptr = p

    @classmethod
    def from_host(cls, ctx, array):
        a = np.ascontiguousarray(array)
        d = cls(ctx, a.dtype, a.size)
        if a.nbytes:
            check(lib().obvhs_memcpy_htod(ctx.handle, d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes), ctx.handle)
        return d

    def to_host(self):
        out = np.empty(self.count, dtype=self.dtype)
        if self.nbytes:
            check(lib().obvhs_memcpy_dtoh(self.ctx.handle, out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes),
                  self.ctx.handle)
        return out

    def free(self):
        if self.ptr:
            lib().obvhs_device_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.handle:
                self.free()
        except Exception:
            pass


def _as_device(ctx, data, dtype):
    if isinstance(data, DeviceArray):
        if data.dtype != np.dtype(dtype):
            raise TypeError(f"expected device array of {dtype}, got {data.dtype}")
        return data, data.count
    a = np.ascontiguousarray(data, dtype=dtype)
    return DeviceArray.from_host(ctx, a), a.size


class Bvh2:
    """Handle of a device-resident Bvh2 (bvh2/mod.rs:31-85)."""

    def __init__(self, ctx, handle, owned=True, keepalive=None):
        self.ctx = ctx
        self.handle = handle
        self.owned = owned
        self._keepalive = keepalive  # the CwBvh owning a borrowed BVH2
        self.sort_precision = SortPrecision.U64  # key width of the last PlocBuilder build into this handle

    def _info(self):
        nc, pc, md = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().obvhs_bvh2_info(self.handle, C.byref(nc), C.byref(pc), C.byref(md)))
        return nc.value, pc.value, md.value

    @property
    def node_count(self):
        return self._info()[0]

    @property
    def max_depth(self):
        return self._info()[2]

    @property
    def nodes(self):
        n = self._info()[0]
        out = np.zeros(n, dtype=BVH2_NODE)
        check(lib().obvhs_bvh2_copy_nodes(self.ctx.handle, self.handle, out.ctypes.data_as(C.c_void_p)),
              self.ctx.handle)
        return out

    @property
    def nodes_small(self):
        n = self._info()[0]
        out = np.zeros(n, dtype=BVH2_NODE_SMALL)
        check(lib().obvhs_bvh2_copy_nodes_small(self.ctx.handle, self.handle, out.ctypes.data_as(C.c_void_p)),
              self.ctx.handle)
        return out

    @property
    def primitive_indices(self):
        n = self._info()[1]
        out = np.zeros(n, dtype=np.uint32)
        check(lib().obvhs_bvh2_copy_primitive_indices(self.ctx.handle, self.handle, out.ctypes.data_as(C.c_void_p)),
              self.ctx.handle)
        return out

    def sort_trace(self):
        """(Morton codes in leaf order, sorted position -> leaf index) of the last build (context trace on).
        The codes are u64[n] after a SortPrecision::U64 build and u64[n, 2] (low, high words) after U128."""
        n = self._info()[1]
        wide = self.sort_precision == SortPrecision.U128
        codes = np.zeros((n, 2) if wide else n, dtype=np.uint64)
        order = np.zeros(n, dtype=np.uint32)
        check(lib().obvhs_bvh2_copy_sort_trace(self.ctx.handle, self.handle, codes.ctypes.data_as(C.c_void_p),
                                               order.ctypes.data_as(C.c_void_p)), self.ctx.handle)
        return codes, order

    def reinsertion(self, batch_size_ratio, ratio_sequence=None):
        """ReinsertionOptimizer::default().run(self, batch_size_ratio, ratio_sequence) (reinsertion.rs:40-61)."""
        seq = None if ratio_sequence is None else np.ascontiguousarray(ratio_sequence, dtype=np.float32)
        check(lib().obvhs_bvh2_reinsertion(self.ctx.handle, self.handle, float(batch_size_ratio),
                                           None if seq is None else seq.ctypes.data_as(C.c_void_p),
                                           0 if seq is None else len(seq)), self.ctx.handle)
        return self

    def refit_all(self):
        """Bvh2::refit_all (bvh2/mod.rs:527-569)."""
        check(lib().obvhs_bvh2_refit_all(self.ctx.handle, self.handle), self.ctx.handle)
        return self

    def refit_from_triangles(self, triangles):
        """Leaf boxes from the (moved) triangles, then refit_all (the example at bvh2/mod.rs:506-526)."""
        d, _ = _as_device(self.ctx, triangles, TRIANGLE)
        check(lib().obvhs_bvh2_refit_from_triangles(self.ctx.handle, self.handle, d.ptr), self.ctx.handle)
        return self

    @classmethod
    def from_host(cls, ctx, nodes, primitive_indices):
        """A device Bvh2 from its parts: Bvh2Node[] (48-byte layout) and primitive_indices (bvh2/mod.rs:31-85)."""
        nodes = np.ascontiguousarray(nodes, dtype=BVH2_NODE)
        prims = np.ascontiguousarray(primitive_indices, dtype=np.uint32)
        out = C.c_void_p()
        check(lib().obvhs_bvh2_create_from_host(ctx.handle, nodes.ctypes.data_as(C.c_void_p), len(nodes),
                                                prims.ctypes.data_as(C.c_void_p), len(prims), C.byref(out)),
              ctx.handle)
        return cls(ctx, out)

    def set_node_aabbs(self, node_ids, aabbs):
        """Bvh2Node::set_aabb (bvh2/node.rs:116-125) for each (id, box); nothing above is refit."""
        di, n = _as_device(self.ctx, node_ids if isinstance(node_ids, DeviceArray)
                           else np.asarray(node_ids, dtype=np.uint32), np.uint32)
        da, m = _as_device(self.ctx, aabbs, AABB)
        if n != m:
            raise ValueError("one box per node id")
        check(lib().obvhs_bvh2_set_node_aabbs(self.ctx.handle, self.handle, di.ptr, da.ptr, n), self.ctx.handle)
        return self

    def primitives_to_nodes(self):
        """Bvh2::compute_primitives_to_nodes (bvh2/mod.rs:647-663) as a device u32 array (INVALID where none)."""
        out = DeviceArray(self.ctx, np.uint32, self._info()[1])
        check(lib().obvhs_bvh2_primitives_to_nodes(self.ctx.handle, self.handle, out.ptr), self.ctx.handle)
        return out

    def close(self):
        if self.handle and self.owned:
            lib().obvhs_bvh2_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PlocBuilder:
    """ploc/mod.rs:34-159.  Keeps the context's scratch allocations between builds like the crate's builder."""

    def __init__(self, ctx):
        self.ctx = ctx

    def build(self, search_distance, aabbs, indices=None, sort_precision=SortPrecision.U64,
              search_depth_threshold=0):
        return self.build_with_bvh(None, search_distance, aabbs, indices, sort_precision, search_depth_threshold)

    def build_with_bvh(self, bvh, search_distance, aabbs, indices=None, sort_precision=SortPrecision.U64,
                       search_depth_threshold=0):
        sd = PlocSearchDistance.for_build(search_distance)
        is_tris = (isinstance(aabbs, DeviceArray) and aabbs.dtype == TRIANGLE) or \
            (isinstance(aabbs, np.ndarray) and aabbs.dtype == TRIANGLE)
        dt = TRIANGLE if is_tris else AABB
        d, n = _as_device(self.ctx, aabbs, dt)
        di = None
        if indices is not None:
            di, ni = _as_device(self.ctx, np.asarray(indices, dtype=np.uint32) if not isinstance(indices, DeviceArray)
                                else indices, np.uint32)
            if ni != n:
                raise ValueError("indices must have one entry per primitive")
        out = C.c_void_p()
        fn = lib().obvhs_ploc_build_triangles if is_tris else lib().obvhs_ploc_build_aabbs
        reuse = bvh.handle if bvh is not None else None
        sp = SortPrecision(sort_precision)
        check(fn(self.ctx.handle, d.ptr, di.ptr if di is not None else None, n, int(sd), int(sp),
                 int(search_depth_threshold), reuse, C.byref(out)), self.ctx.handle)
        if bvh is None:
            bvh = Bvh2(self.ctx, out)
        bvh.sort_precision = sp
        return bvh

    def full_rebuild(self, bvh, search_distance, sort_precision=SortPrecision.U64, search_depth_threshold=0):
        """ploc/rebuild.rs:53-79: rebuild `bvh` from its current leaves."""
        sd = PlocSearchDistance.for_build(search_distance)
        check(lib().obvhs_bvh2_full_rebuild(self.ctx.handle, bvh.handle, int(sd), int(SortPrecision(sort_precision)),
                                            int(search_depth_threshold)), self.ctx.handle)
        return bvh

    def partial_rebuild(self, bvh, should_remove, search_distance, sort_precision=SortPrecision.U64,
                        search_depth_threshold=0):
        """ploc/rebuild.rs:98-134.  `should_remove` is one flag per node (a device u8 array such as
        compute_rebuild_path_flags returns, or a host bool/u8 array): the closure |i| flags[i]."""
        sd = PlocSearchDistance.for_build(search_distance)
        sp = SortPrecision(sort_precision)
        n = bvh.node_count
        if isinstance(should_remove, DeviceArray):
            flags = should_remove
            if flags.count < n:
                raise ValueError("one flag per node")
        else:
            host = np.ascontiguousarray(np.asarray(should_remove).astype(np.uint8))
            if host.size != n:
                raise ValueError("one flag per node")
            flags = DeviceArray.from_host(self.ctx, host if n else np.zeros(1, np.uint8))
        check(lib().obvhs_bvh2_partial_rebuild(self.ctx.handle, bvh.handle, flags.ptr, int(sd), int(sp),
                                               int(search_depth_threshold)), self.ctx.handle)
        return bvh


def compute_rebuild_path_flags(bvh, leaves):
    """ploc/rebuild.rs:12-42: a device u8 array, 1 for each given leaf and its ancestors."""
    ctx = bvh.ctx
    di, n = _as_device(ctx, leaves if isinstance(leaves, DeviceArray) else np.asarray(leaves, dtype=np.uint32),
                       np.uint32)
    flags = DeviceArray(ctx, np.uint8, bvh.node_count)
    check(lib().obvhs_memset_device(ctx.handle, flags.ptr, 0, flags.nbytes), ctx.handle)
    check(lib().obvhs_bvh2_rebuild_path_flags(ctx.handle, bvh.handle, di.ptr if n else None, n, flags.ptr),
          ctx.handle)
    return flags


class CwBvh:
    """Handle of a device-resident CwBvh (cwbvh/mod.rs:43-55)."""

    def __init__(self, ctx, handle, stats=None):
        self.ctx = ctx
        self.handle = handle
        self.stats = stats

    def _info(self):
        nc, pc, bb = C.c_uint32(), C.c_uint32(), native.Aabb()
        check(lib().obvhs_cwbvh_info(self.handle, C.byref(nc), C.byref(pc), C.byref(bb)))
        return nc.value, pc.value, bb

    @property
    def node_count(self):
        return self._info()[0]

    @property
    def total_aabb(self):
        bb = self._info()[2]
        return np.array(bb.min[:3], dtype=np.float32), np.array(bb.max[:3], dtype=np.float32)

    @property
    def nodes(self):
        n = self._info()[0]
        out = np.zeros(n, dtype=CWBVH_NODE)
        check(lib().obvhs_cwbvh_copy_nodes(self.ctx.handle, self.handle, out.ctypes.data_as(C.c_void_p)),
              self.ctx.handle)
        return out

    @property
    def primitive_indices(self):
        n = self._info()[1]
        out = np.zeros(n, dtype=np.uint32)
        check(lib().obvhs_cwbvh_copy_primitive_indices(self.ctx.handle, self.handle,
                                                       out.ctypes.data_as(C.c_void_p)), self.ctx.handle)
        return out

    def reorder_triangles(self, tris):
        """bvh_tris[i] = tris[primitive_indices[i]] on the device (examples/obj_cwbvh.rs:63-67)."""
        d, n = _as_device(self.ctx, tris, TRIANGLE)
        out = DeviceArray(self.ctx, TRIANGLE, self._info()[1])
        check(lib().obvhs_cwbvh_reorder_triangles(self.ctx.handle, self.handle, d.ptr, out.ptr), self.ctx.handle)
        return out

    def ray_traverse(self, rays, bvh_tris, counted=False):
        """CwBvh::ray_traverse for every ray (cwbvh/mod.rs:169-198).  bvh_tris in primitive_indices order.
        Returns RayHit[n] (primitive_id indexes primitive_indices); with counted=True also per-ray node and
        triangle test counts."""
        dr, n = _as_device(self.ctx, rays, RAY)
        dt, _ = _as_device(self.ctx, bvh_tris, TRIANGLE)
        hits = DeviceArray(self.ctx, RAY_HIT, n)
        if counted:
            nv = DeviceArray(self.ctx, np.uint32, n)
            tv = DeviceArray(self.ctx, np.uint32, n)
            check(lib().obvhs_cwbvh_ray_traverse_counted(self.ctx.handle, self.handle, dt.ptr, dr.ptr, n, hits.ptr,
                                                         nv.ptr, tv.ptr), self.ctx.handle)
            return hits.to_host(), nv.to_host(), tv.to_host()
        check(lib().obvhs_cwbvh_ray_traverse(self.ctx.handle, self.handle, dt.ptr, dr.ptr, n, hits.ptr),
              self.ctx.handle)
        return hits.to_host()

    def ray_traverse_miss(self, rays, bvh_tris):
        """CwBvh::ray_traverse_miss for every ray (cwbvh/mod.rs:201-225): True where the ray hit nothing."""
        dr, n = _as_device(self.ctx, rays, RAY)
        dt, _ = _as_device(self.ctx, bvh_tris, TRIANGLE)
        miss = DeviceArray(self.ctx, np.uint8, n)
        check(lib().obvhs_cwbvh_ray_traverse_miss(self.ctx.handle, self.handle, dt.ptr, dr.ptr, n, miss.ptr),
              self.ctx.handle)
        return miss.to_host().astype(bool)

    def ray_traverse_watertight(self, rays, bvh_tris, miss=False):
        """The same traversal with the watertight triangle test (an extra mode beside Triangle::intersect; no
        counterpart in the crate): RayHit[n], or with miss=True the occlusion flags (True where nothing was hit)."""
        dr, n = _as_device(self.ctx, rays, RAY)
        dt, _ = _as_device(self.ctx, bvh_tris, TRIANGLE)
        if miss:
            out = DeviceArray(self.ctx, np.uint8, n)
            check(lib().obvhs_cwbvh_ray_traverse_miss_watertight(self.ctx.handle, self.handle, dt.ptr, dr.ptr, n,
                                                                 out.ptr), self.ctx.handle)
            return out.to_host().astype(bool)
        hits = DeviceArray(self.ctx, RAY_HIT, n)
        check(lib().obvhs_cwbvh_ray_traverse_watertight(self.ctx.handle, self.handle, dt.ptr, dr.ptr, n, hits.ptr),
              self.ctx.handle)
        return hits.to_host()

    def ray_traverse_rt(self, rays, bvh_rt_tris, miss=False):
        """ray_traverse / ray_traverse_miss with RtTriangle primitives (rt_triangle.rs:200-239, the crate's
        demoscene layout), bvh_rt_tris in primitive_indices order (rt_triangles()); the same hits as ray_traverse."""
        dr, n = _as_device(self.ctx, rays, RAY)
        dt, _ = _as_device(self.ctx, bvh_rt_tris, RT_TRIANGLE)
        if miss:
            out = DeviceArray(self.ctx, np.uint8, n)
            check(lib().obvhs_cwbvh_ray_traverse_miss_rt(self.ctx.handle, self.handle, dt.ptr, dr.ptr, n, out.ptr),
                  self.ctx.handle)
            return out.to_host().astype(bool)
        hits = DeviceArray(self.ctx, RAY_HIT, n)
        check(lib().obvhs_cwbvh_ray_traverse_rt(self.ctx.handle, self.handle, dt.ptr, dr.ptr, n, hits.ptr),
              self.ctx.handle)
        return hits.to_host()

    @property
    def bvh2(self):
        """The BVH2 this CwBvh was built from by build_cwbvh_from_tris (borrowed; None for conversions)."""
        out = C.c_void_p()
        check(lib().obvhs_cwbvh_bvh2(self.handle, C.byref(out)))
        return Bvh2(self.ctx, out, owned=False, keepalive=self) if out.value else None

    def close(self):
        if self.handle:
            lib().obvhs_cwbvh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bvh2_to_cwbvh(bvh2, max_prims_per_leaf, order_children=True, include_exact_node_aabbs=False):
    """cwbvh/bvh2_to_cwbvh.rs:490-510."""
    out = C.c_void_p()
    check(lib().obvhs_bvh2_to_cwbvh(bvh2.ctx.handle, bvh2.handle, int(max_prims_per_leaf), int(bool(order_children)),
                                    int(bool(include_exact_node_aabbs)), None, C.byref(out)), bvh2.ctx.handle)
    return CwBvh(bvh2.ctx, out)


def build_cwbvh_from_tris(ctx, triangles, config, reuse=None, split_max_total=0, ratio_sequence=None):
    """cwbvh/builder.rs:20-85.  split_max_total (0 = no limit) caps the pre-split leaf count; ratio_sequence is the
    reinsertion ratio sequence (None = the crate's default)."""
    d, n = _as_device(ctx, triangles, TRIANGLE)
    out = C.c_void_p()
    stats = BuildStats()
    params = config.to_c()
    seq = None if ratio_sequence is None else np.ascontiguousarray(ratio_sequence, dtype=np.float32)
    ext = native.BuildExt(int(split_max_total), None if seq is None else seq.ctypes.data_as(C.c_void_p),
                          0 if seq is None else len(seq))
    check(lib().obvhs_build_cwbvh_from_tris_ex(ctx.handle, d.ptr, n, C.byref(params), C.byref(ext),
                                               reuse.handle if reuse is not None else None, C.byref(out),
                                               C.byref(stats)), ctx.handle)
    if reuse is not None:
        reuse.stats = stats
        return reuse
    return CwBvh(ctx, out, stats)


class CwBvhBatch:
    """Many CwBvhs built together (one per mesh), each identical to build_cwbvh_from_tris on that mesh alone."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.handle = handle

    def info(self):
        m, nodes, prims, passes = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().obvhs_batch_info(self.handle, C.byref(m), C.byref(nodes), C.byref(prims), C.byref(passes)))
        return dict(mesh_count=m.value, total_nodes=nodes.value, total_prims=prims.value, ploc_passes=passes.value)

    def mesh_info(self, i):
        vals = [C.c_uint32() for _ in range(4)]
        check(lib().obvhs_batch_mesh_info(self.handle, int(i), *[C.byref(v) for v in vals]))
        return tuple(v.value for v in vals)  # node_offset, node_count, prim_offset, prim_count

    def all_nodes(self):
        out = np.zeros(self.info()["total_nodes"], dtype=CWBVH_NODE)
        check(lib().obvhs_batch_copy_nodes(self.ctx.handle, self.handle, out.ctypes.data_as(C.c_void_p)),
              self.ctx.handle)
        return out

    def all_primitive_indices(self):
        out = np.zeros(self.info()["total_prims"], dtype=np.uint32)
        check(lib().obvhs_batch_copy_primitive_indices(self.ctx.handle, self.handle,
                                                       out.ctypes.data_as(C.c_void_p)), self.ctx.handle)
        return out

    def reorder_triangles(self, tris):
        d, n = _as_device(self.ctx, tris, TRIANGLE)
        out = DeviceArray(self.ctx, TRIANGLE, n)
        check(lib().obvhs_batch_reorder_triangles(self.ctx.handle, self.handle, d.ptr, out.ptr), self.ctx.handle)
        return out

    def ray_traverse(self, rays, ray_mesh, bvh_tris):
        dr, n = _as_device(self.ctx, rays, RAY)
        dm, _ = _as_device(self.ctx, ray_mesh, np.uint32)
        dt, _ = _as_device(self.ctx, bvh_tris, TRIANGLE)
        hits = DeviceArray(self.ctx, RAY_HIT, n)
        check(lib().obvhs_batch_ray_traverse(self.ctx.handle, self.handle, dt.ptr, dr.ptr, dm.ptr, n, hits.ptr),
              self.ctx.handle)
        return hits.to_host()

    def ray_traverse_miss(self, rays, ray_mesh, bvh_tris):
        dr, n = _as_device(self.ctx, rays, RAY)
        dm, _ = _as_device(self.ctx, ray_mesh, np.uint32)
        dt, _ = _as_device(self.ctx, bvh_tris, TRIANGLE)
        miss = DeviceArray(self.ctx, np.uint8, n)
        check(lib().obvhs_batch_ray_traverse_miss(self.ctx.handle, self.handle, dt.ptr, dr.ptr, dm.ptr, n,
                                                  miss.ptr), self.ctx.handle)
        return miss.to_host().astype(bool)

    def build_tlas(self, config):
        """A CwBvh over the meshes' bounds (the TLAS of ray_traverse_tlas_blas, cwbvh/mod.rs:341-491)."""
        out = C.c_void_p()
        p = config.to_c()
        check(lib().obvhs_batch_build_tlas(self.ctx.handle, self.handle, C.byref(p), C.byref(out)), self.ctx.handle)
        return CwBvh(self.ctx, out)

    def tlas_ray_traverse(self, tlas, rays, bvh_tris):
        """CwBvh::ray_traverse_tlas_blas with this batch's meshes as BLASes: RayHit[n] with geometry_id = mesh."""
        dr, n = _as_device(self.ctx, rays, RAY)
        dt, _ = _as_device(self.ctx, bvh_tris, TRIANGLE)
        hits = DeviceArray(self.ctx, RAY_HIT, n)
        check(lib().obvhs_batch_tlas_ray_traverse(self.ctx.handle, tlas.handle, self.handle, dt.ptr, dr.ptr, n,
                                                  hits.ptr), self.ctx.handle)
        return hits.to_host()

    def close(self):
        if self.handle:
            lib().obvhs_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_cwbvh_batch(ctx, triangles, offsets, config, reuse=None):
    """Builds one CwBvh per mesh; mesh i is triangles[offsets[i]:offsets[i+1]] (concatenated Triangle array)."""
    d, n = _as_device(ctx, triangles, TRIANGLE)
    offs = np.ascontiguousarray(offsets, dtype=np.uint32)
    if offs[0] != 0 or offs[-1] != n:
        raise ValueError("offsets must start at 0 and end at the triangle count")
    out = C.c_void_p()
    params = config.to_c()
    check(lib().obvhs_batch_build_cwbvh(ctx.handle, d.ptr, offs.ctypes.data_as(C.c_void_p), len(offs) - 1,
                                        C.byref(params), reuse.handle if reuse is not None else None,
                                        C.byref(out)), ctx.handle)
    if reuse is not None:
        return reuse
    return CwBvhBatch(ctx, out)


REQ_TRACE_CLOSEST, REQ_TRACE_MISS, REQ_REFIT, REQ_REBUILD = 0, 1, 2, 3


def _info_dict(inf):
    return dict(status=native.STATUS.get(inf.status, inf.status), kind=inf.kind, scene=inf.scene,
                version=inf.version, batch=inf.batch, batch_requests=inf.batch_requests, submit_ns=inf.submit_ns,
                start_ns=inf.start_ns, done_ns=inf.done_ns)


class Server:
    """Resident scenes, coalesced ray requests and overlapped refit / rebuild updates (the C server).

    Requests name device arrays (DeviceArray of a client Context); they must stay alive until waited for."""

    def __init__(self, device=0, hbm_budget_bytes=0, max_batch_rays=0, max_inflight=0, update_streams=0):
        cfg = native.ServerConfig(int(hbm_budget_bytes), int(max_batch_rays), int(max_inflight), int(update_streams), 0)
        h = C.c_void_p()
        check(lib().obvhs_server_create(int(device), C.byref(cfg), C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().obvhs_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_scene(self, d_tris, params):
        """Builds a resident scene from a device Triangle array (copied) with BvhBuildParams."""
        sid = C.c_uint32()
        p = params.to_c()
        check(lib().obvhs_server_add_scene(self.handle, d_tris.ptr, d_tris.count, C.byref(p), C.byref(sid)))
        return sid.value

    def submit_trace(self, scene, d_rays, d_out, miss=False, n=None):
        rid = C.c_uint64()
        n = d_rays.count if n is None else n
        check(lib().obvhs_server_submit_trace(self.handle, int(scene), REQ_TRACE_MISS if miss else REQ_TRACE_CLOSEST,
                                              d_rays.ptr, n, d_out.ptr, C.byref(rid)))
        return rid.value

    def submit_update(self, scene, d_tris, rebuild=False):
        rid = C.c_uint64()
        check(lib().obvhs_server_submit_update(self.handle, int(scene), REQ_REBUILD if rebuild else REQ_REFIT,
                                               d_tris.ptr, d_tris.count, C.byref(rid)))
        return rid.value

    def wait(self, rid):
        inf = native.RequestInfo()
        lib().obvhs_server_wait(self.handle, int(rid), C.byref(inf))
        return _info_dict(inf)

    def replay(self, requests):
        """requests: list of dicts {arrival_ns, kind, scene, rays, out, tris} (device arrays); submitted open-loop
        at their arrival offsets, then all waited for.  Returns one info dict per request."""
        arr = (native.ServerRequest * len(requests))()
        for i, q in enumerate(requests):
            kind = q["kind"]
            arr[i].arrival_ns = int(q.get("arrival_ns", 0))
            arr[i].kind = kind
            arr[i].scene = int(q["scene"])
            if kind in (REQ_TRACE_CLOSEST, REQ_TRACE_MISS):
                arr[i].rays, arr[i].out, arr[i].n = q["rays"].ptr, q["out"].ptr, q.get("n", q["rays"].count)
            else:
                arr[i].tris, arr[i].n = q["tris"].ptr, q["tris"].count
        infos = (native.RequestInfo * len(requests))()
        lib().obvhs_server_replay(self.handle, arr, len(requests), infos)
        return [_info_dict(x) for x in infos]

    def scene_info(self, scene):
        v, p, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().obvhs_server_scene_info(self.handle, int(scene), C.byref(v), C.byref(p), C.byref(c)))
        return dict(version=v.value, prim_count=p.value, cwbvh_nodes=c.value)

    def stats(self):
        st = native.ServerStats()
        check(lib().obvhs_server_get_stats(self.handle, C.byref(st)))
        return dict(launches=st.launches, rays=st.rays, segments=st.segments, resident_bytes=st.resident_bytes,
                    scenes=st.scenes)


def rays_new(ctx, origins, directions, tmin=0.0, tmax=float("inf")):
    """Ray::new (ray.rs:34-52) on the device: a device Ray array from (n, 3) or (n, 4) origins and directions
    (host arrays are uploaded as 16-byte vec3a rows)."""
    o = np.zeros((len(origins), 4), np.float32)
    d = np.zeros((len(directions), 4), np.float32)
    o[:, :3] = np.asarray(origins, np.float32)[:, :3]
    d[:, :3] = np.asarray(directions, np.float32)[:, :3]
    if len(o) != len(d):
        raise ValueError("origins and directions differ in length")
    do, dd = ctx.upload(o), ctx.upload(d)
    out = DeviceArray(ctx, RAY, len(o))
    check(lib().obvhs_rays_new(ctx.handle, do.ptr, dd.ptr, len(o), float(tmin), float(tmax), out.ptr), ctx.handle)
    return out


def ray_hits_none(ctx, n):
    """RayHit::none() (ray.rs:76-83) for n device hit records."""
    out = DeviceArray(ctx, RAY_HIT, n)
    check(lib().obvhs_ray_hits_none(ctx.handle, out.ptr, n), ctx.handle)
    return out


def rt_triangles(ctx, triangles):
    """RtTriangle::from(&Triangle) for every triangle (rt_triangle.rs:162-181): a device RtTriangle array."""
    d, n = _as_device(ctx, triangles, TRIANGLE)
    out = DeviceArray(ctx, RT_TRIANGLE, n)
    check(lib().obvhs_rt_triangles_from_triangles(ctx.handle, d.ptr, n, out.ptr), ctx.handle)
    return out


def split_aabbs_preset(ctx, triangles, max_total=0, with_stats=False):
    """The pre-split leaf set of build_cwbvh_from_tris (cwbvh/builder.rs:26-55): (Aabb array, u32 indices), plus
    the builder's average and largest half area when with_stats."""
    d, n = _as_device(ctx, triangles, TRIANGLE)
    cap = max(1, 2 * n)
    while True:
        da, di = DeviceArray(ctx, AABB, cap), DeviceArray(ctx, np.uint32, cap)
        cnt, avg, big = C.c_uint32(), C.c_float(), C.c_float()
        st = lib().obvhs_split_aabbs_preset(ctx.handle, d.ptr, n, int(max_total), da.ptr, di.ptr, cap, C.byref(cnt),
                                            C.byref(avg), C.byref(big))
        if st == native.STATUS_CODE["OBVHS_ERR_CAPACITY"] and cnt.value > cap:
            cap = cnt.value
            continue
        check(st, ctx.handle)
        out = (da.to_host()[:cnt.value].copy(), di.to_host()[:cnt.value].copy())
        return out + (avg.value, big.value) if with_stats else out


def log2_threshold_table():
    out = np.zeros(256, dtype=np.uint32)
    check(lib().obvhs_log2_threshold_table(out.ctypes.data_as(C.c_void_p)))
    return out


def size_class(n):
    """Size class of n primitives or keys: 0 below 4,096, 1 below 65,536, 2 below 1,048,576, else 3."""
    return int(lib().obvhs_size_class(int(n)))


def tune_sort(ctx, n):
    """Times the radix sort of n keys with each tile size; stores and returns the fastest (keys per thread)."""
    out = C.c_uint32()
    check(lib().obvhs_tune_sort(ctx.handle, int(n), C.byref(out)), ctx.handle)
    return out.value


def tune_traversal(ctx, cwbvh, rays, bvh_tris, any_hit=False):
    """Times traversal of the rays with each LDS stack depth, one lane per ray and chunked; stores the fastest
    for the BVH's class and returns its depth (the chunk choice is in ctx.tune_table()["trav_chunk"])."""
    dr, n = _as_device(ctx, rays, RAY)
    dt, _ = _as_device(ctx, bvh_tris, TRIANGLE)
    out = C.c_uint32()
    check(lib().obvhs_tune_traversal(ctx.handle, cwbvh.handle, dt.ptr, dr.ptr, n, int(bool(any_hit)),
                                     C.byref(out)), ctx.handle)
    return out.value


def tune_ploc_radius(ctx, triangles, max_prims_per_leaf=3, tolerance=0.01, baseline=PlocSearchDistance.Medium):
    """Builds the triangles with every PlocSearchDistance and picks the smallest whose CWBVH cost is within
    `tolerance` of the cost at `baseline` (Auto/0: of the lowest cost); stores it for the size class (Auto builds
    use it).  Returns (distance, costs of the six distances)."""
    d, n = _as_device(ctx, triangles, TRIANGLE)
    out = C.c_uint32()
    costs = np.zeros(6, np.float32)
    check(lib().obvhs_tune_ploc_radius(ctx.handle, d.ptr, n, int(max_prims_per_leaf), int(baseline),
                                       float(tolerance), C.byref(out), costs.ctypes.data_as(C.c_void_p)),
          ctx.handle)
    return PlocSearchDistance(out.value), costs


__all__ = ["Server", "CwBvhBatch", "build_cwbvh_batch", "split_aabbs_preset", "PlocSearchDistance", "SortPrecision",
           "BvhBuildParams", "Context", "DeviceArray", "Bvh2", "PlocBuilder", "compute_rebuild_path_flags",
           "CwBvh", "bvh2_to_cwbvh", "build_cwbvh_from_tris", "log2_threshold_table", "ObvhsError", "size_class",
           "tune_sort", "tune_traversal", "tune_ploc_radius"]
