"""ctypes binding of libobvhs_hip.so (the C ABI declared in include/obvhs_hip.h).

The library is loaded from the package directory only; when it is missing the import of this module fails loudly --
there is no CPU fallback anywhere in the product path."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libobvhs_hip.so")

STATUS = {
    0: "OBVHS_OK", 1: "OBVHS_ERR_INVALID_ARG", 2: "OBVHS_ERR_NAN_INPUT", 3: "OBVHS_ERR_PLOC_SELF_MERGE",
    4: "OBVHS_ERR_CW_DISTRIBUTE", 5: "OBVHS_ERR_CW_NONUNIT_LEAF", 6: "OBVHS_ERR_CW_ORDER", 7: "OBVHS_ERR_CW_PRIMCOUNT",
    8: "OBVHS_ERR_CAPACITY", 9: "OBVHS_ERR_DEPTH", 10: "OBVHS_ERR_OOM", 11: "OBVHS_ERR_DEVICE",
    12: "OBVHS_ERR_NO_DEVICE",
}
STATUS_CODE = {v: k for k, v in STATUS.items()}


class ObvhsError(RuntimeError):
    def __init__(self, status, message=""):
        self.status = status
        self.name = STATUS.get(status, f"status {status}")
        super().__init__(f"{self.name}: {message}" if message else self.name)


class BuildParams(C.Structure):
    """obvhs_build_params -- BvhBuildParams (lib.rs:207-230)."""
    _fields_ = [("pre_split", C.c_uint32), ("ploc_search_distance", C.c_uint32),
                ("search_depth_threshold", C.c_uint32), ("reinsertion_batch_ratio", C.c_float),
                ("post_collapse_reinsertion_batch_ratio_multiplier", C.c_float), ("sort_precision", C.c_uint32),
                ("max_prims_per_leaf", C.c_uint32), ("collapse_traversal_cost", C.c_float)]


class BuildStats(C.Structure):
    _fields_ = [("prim_count", C.c_uint32), ("ploc_iterations", C.c_uint32), ("bvh2_node_count", C.c_uint32),
                ("cwbvh_node_count", C.c_uint32), ("ploc_cluster_sum", C.c_uint64)]


class BuildExt(C.Structure):
    """obvhs_build_ext: the split budget and the reinsertion ratio sequence."""
    _fields_ = [("split_max_total", C.c_uint64), ("ratio_sequence", C.c_void_p), ("ratio_sequence_len", C.c_uint32)]


class TuneTable(C.Structure):
    """obvhs_tune_table: per size class PLOC distance for "auto" builds, sort tile, traversal LDS stack depth,
    chunked traversal launches."""
    _fields_ = [("ploc_radius", C.c_uint32 * 4), ("sort_items", C.c_uint32 * 4), ("trav_lds_depth", C.c_uint32 * 4),
                ("trav_chunk", C.c_uint32 * 4)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_batch_rays", C.c_uint32), ("max_inflight", C.c_uint32),
                ("update_streams", C.c_uint32), ("pad", C.c_uint32)]


class RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("kind", C.c_uint32), ("scene", C.c_uint32), ("version", C.c_uint32),
                ("batch", C.c_uint64), ("batch_requests", C.c_uint32), ("pad", C.c_uint32),
                ("submit_ns", C.c_int64), ("start_ns", C.c_int64), ("done_ns", C.c_int64)]


class ServerRequest(C.Structure):
    _fields_ = [("arrival_ns", C.c_uint64), ("kind", C.c_uint32), ("scene", C.c_uint32), ("rays", C.c_void_p),
                ("out", C.c_void_p), ("tris", C.c_void_p), ("n", C.c_uint32), ("pad", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("rays", C.c_uint64), ("segments", C.c_uint64),
                ("resident_bytes", C.c_uint64), ("scenes", C.c_uint32), ("pad", C.c_uint32)]


class Aabb(C.Structure):
    _fields_ = [("min", C.c_float * 4), ("max", C.c_float * 4)]


P = C.c_void_p
U32 = C.c_uint32
U64 = C.c_uint64
PU32 = C.POINTER(C.c_uint32)
INT = C.c_int

# name -> (restype, argtypes)
SIGNATURES = {
    "obvhs_status_string": (C.c_char_p, [INT]),
    "obvhs_context_create": (INT, [INT, P, C.POINTER(P)]),
    "obvhs_context_destroy": (INT, [P]),
    "obvhs_context_synchronize": (INT, [P]),
    "obvhs_context_set_trace": (INT, [P, INT]),
    "obvhs_size_class": (U32, [C.c_uint64]),
    "obvhs_tune_get": (INT, [P, P]),
    "obvhs_tune_set": (INT, [P, P]),
    "obvhs_tune_sort": (INT, [P, U32, P]),
    "obvhs_tune_traversal": (INT, [P, P, P, P, U32, INT, P]),
    "obvhs_tune_ploc_radius": (INT, [P, P, U32, U32, U32, C.c_float, P, P]),
    "obvhs_last_error_message": (C.c_char_p, [P]),
    "obvhs_device_malloc": (INT, [P, U64, C.POINTER(P)]),
    "obvhs_device_free": (INT, [P, P]),
    "obvhs_memcpy_htod": (INT, [P, P, P, U64]),
    "obvhs_memcpy_dtoh": (INT, [P, P, P, U64]),
    "obvhs_memset_device": (INT, [P, P, INT, U64]),
    "obvhs_memcpy_dtod": (INT, [P, P, P, U64]),
    "obvhs_timer_start": (INT, [P]),
    "obvhs_timer_stop": (INT, [P, C.POINTER(C.c_float)]),
    "obvhs_ploc_build_aabbs": (INT, [P, P, P, U32, U32, U32, U32, P, C.POINTER(P)]),
    "obvhs_ploc_build_triangles": (INT, [P, P, P, U32, U32, U32, U32, P, C.POINTER(P)]),
    "obvhs_bvh2_info": (INT, [P, PU32, PU32, PU32]),
    "obvhs_bvh2_copy_nodes": (INT, [P, P, P]),
    "obvhs_bvh2_copy_nodes_small": (INT, [P, P, P]),
    "obvhs_bvh2_copy_primitive_indices": (INT, [P, P, P]),
    "obvhs_bvh2_copy_sort_trace": (INT, [P, P, P, P]),
    "obvhs_bvh2_destroy": (INT, [P]),
    "obvhs_bvh2_to_cwbvh": (INT, [P, P, U32, INT, INT, P, C.POINTER(P)]),
    "obvhs_build_cwbvh_from_tris": (INT, [P, P, U32, C.POINTER(BuildParams), P, C.POINTER(P),
                                          C.POINTER(BuildStats)]),
    "obvhs_cwbvh_info": (INT, [P, PU32, PU32, C.POINTER(Aabb)]),
    "obvhs_cwbvh_copy_nodes": (INT, [P, P, P]),
    "obvhs_cwbvh_copy_primitive_indices": (INT, [P, P, P]),
    "obvhs_cwbvh_device_arrays": (INT, [P, C.POINTER(P), C.POINTER(P)]),
    "obvhs_cwbvh_reorder_triangles": (INT, [P, P, P, P]),
    "obvhs_cwbvh_destroy": (INT, [P]),
    "obvhs_cwbvh_ray_traverse": (INT, [P, P, P, P, U32, P]),
    "obvhs_cwbvh_ray_traverse_miss": (INT, [P, P, P, P, U32, P]),
    "obvhs_cwbvh_ray_traverse_counted": (INT, [P, P, P, P, U32, P, P, P]),
    "obvhs_cwbvh_ray_traverse_watertight": (INT, [P, P, P, P, U32, P]),
    "obvhs_cwbvh_ray_traverse_miss_watertight": (INT, [P, P, P, P, U32, P]),
    "obvhs_rays_new": (INT, [P, P, P, U32, C.c_float, C.c_float, P]),
    "obvhs_ray_hits_none": (INT, [P, P, U32]),
    "obvhs_rt_triangles_from_triangles": (INT, [P, P, U32, P]),
    "obvhs_cwbvh_ray_traverse_rt": (INT, [P, P, P, P, U32, P]),
    "obvhs_cwbvh_ray_traverse_miss_rt": (INT, [P, P, P, P, U32, P]),
    "obvhs_cwbvh_ray_traverse_noalloc": (INT, [P, P, P, P, U32, P]),
    "obvhs_log2_threshold_table": (INT, [P]),
    "obvhs_batch_build_cwbvh": (INT, [P, P, P, U32, C.POINTER(BuildParams), P, C.POINTER(P)]),
    "obvhs_batch_info": (INT, [P, PU32, PU32, PU32, PU32]),
    "obvhs_batch_mesh_info": (INT, [P, U32, PU32, PU32, PU32, PU32]),
    "obvhs_batch_copy_nodes": (INT, [P, P, P]),
    "obvhs_batch_copy_primitive_indices": (INT, [P, P, P]),
    "obvhs_batch_reorder_triangles": (INT, [P, P, P, P]),
    "obvhs_batch_ray_traverse": (INT, [P, P, P, P, P, U32, P]),
    "obvhs_batch_ray_traverse_miss": (INT, [P, P, P, P, P, U32, P]),
    "obvhs_batch_destroy": (INT, [P]),
    "obvhs_batch_build_tlas": (INT, [P, P, C.POINTER(BuildParams), C.POINTER(P)]),
    "obvhs_batch_tlas_ray_traverse": (INT, [P, P, P, P, P, U32, P]),
    "obvhs_build_cwbvh_from_tris_ex": (INT, [P, P, U32, C.POINTER(BuildParams), C.POINTER(BuildExt), P,
                                             C.POINTER(P), C.POINTER(BuildStats)]),
    "obvhs_cwbvh_bvh2": (INT, [P, C.POINTER(P)]),
    "obvhs_bvh2_reinsertion": (INT, [P, P, C.c_float, P, U32]),
    "obvhs_bvh2_refit_all": (INT, [P, P]),
    "obvhs_bvh2_refit_from_triangles": (INT, [P, P, P]),
    "obvhs_bvh2_create_from_host": (INT, [P, P, U32, P, U32, P]),
    "obvhs_bvh2_set_node_aabbs": (INT, [P, P, P, P, U32]),
    "obvhs_bvh2_primitives_to_nodes": (INT, [P, P, P]),
    "obvhs_bvh2_rebuild_path_flags": (INT, [P, P, P, U32, P]),
    "obvhs_bvh2_partial_rebuild": (INT, [P, P, P, U32, U32, U32]),
    "obvhs_bvh2_full_rebuild": (INT, [P, P, U32, U32, U32]),
    "obvhs_split_aabbs_preset": (INT, [P, P, U32, U64, P, P, U32, PU32, C.POINTER(C.c_float),
                                       C.POINTER(C.c_float)]),
    "obvhs_server_create": (INT, [INT, C.POINTER(ServerConfig), C.POINTER(P)]),
    "obvhs_server_destroy": (INT, [P]),
    "obvhs_server_add_scene": (INT, [P, P, U32, C.POINTER(BuildParams), PU32]),
    "obvhs_server_submit_trace": (INT, [P, U32, U32, P, U32, P, C.POINTER(U64)]),
    "obvhs_server_submit_update": (INT, [P, U32, U32, P, U32, C.POINTER(U64)]),
    "obvhs_server_wait": (INT, [P, U64, C.POINTER(RequestInfo)]),
    "obvhs_server_replay": (INT, [P, C.POINTER(ServerRequest), U32, C.POINTER(RequestInfo)]),
    "obvhs_server_scene_info": (INT, [P, U32, PU32, PU32, PU32]),
    "obvhs_server_get_stats": (INT, [P, C.POINTER(ServerStats)]),
}

_lib = None


def lib():
    """Loads the library (once).  Raises ImportError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m obvhs_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, ctx=None):
    if status != 0:
        msg = ""
        if ctx is not None:
            m = lib().obvhs_last_error_message(ctx)
            msg = m.decode() if m else ""
        raise ObvhsError(status, msg)
