"""ctypes loader of the CPU oracle (oracle/liboracle.so) -- the checker the GPU path is compared against.

Only the test suite, __graft_entry__.smoke() and the CPU-baseline leg of bench.py load it."""
import ctypes as C
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "oracle", "liboracle.so")

_lib = None
P = C.c_void_p


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            import importlib.util
            spec = importlib.util.spec_from_file_location("build_oracle", os.path.join(REPO, "oracle",
                                                                                       "build_oracle.py"))
            m = importlib.util.module_from_spec(spec)
            spec.loader.exec_module(m)
            m.build()
        L = C.CDLL(LIB)
        L.orc_sizeof.restype = C.c_int
        L.orc_sizeof.argtypes = [C.c_int]
        L.orc_tri_aabbs.argtypes = [P, C.c_uint64, P]
        L.orc_morton_encode_u64_unorm.restype = C.c_uint64
        L.orc_morton_encode_u64_unorm.argtypes = [C.c_double] * 3
        L.orc_split_by_3_u64.restype = C.c_uint64
        L.orc_split_by_3_u64.argtypes = [C.c_uint32]
        L.orc_f64_as_u32.restype = C.c_uint32
        L.orc_f64_as_u32.argtypes = [C.c_double]
        L.orc_half_area.restype = C.c_float
        L.orc_half_area.argtypes = [P]
        L.orc_ploc_build.restype = C.c_int
        L.orc_ploc_build.argtypes = [P, P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, P, P, P, P, P, C.c_uint32,
                                     P, P, P, C.c_uint32, P]
        L.orc_split_by_3_u128.argtypes = [C.c_uint64, P]
        L.orc_morton_encode_u128_unorm.argtypes = [C.c_double, C.c_double, C.c_double, P]
        L.orc_bvh2_to_cwbvh.restype = C.c_int
        L.orc_bvh2_to_cwbvh.argtypes = [P, C.c_uint64, P, C.c_uint64, C.c_uint32, C.c_int, P, C.c_uint64, P,
                                        C.c_uint64, P, P, P]
        L.orc_cwbvh_traverse.argtypes = [P, C.c_uint64, P, P, C.c_uint64, C.c_int, P, P, P, P, C.c_int]
        L.orc_brute_force.argtypes = [P, C.c_uint64, P, C.c_uint64, P]
        L.orc_brute_force_wt.argtypes = [P, C.c_uint64, P, C.c_uint64, P]
        L.orc_tri_intersect_wt.restype = C.c_float
        L.orc_tri_intersect_wt.argtypes = [P, P]
        L.orc_tri_intersect.restype = C.c_float
        L.orc_tri_intersect.argtypes = [P, P]
        L.orc_intersect_node.restype = C.c_uint32
        L.orc_intersect_node.argtypes = [P, P, C.c_uint32]
        L.orc_ray_octant_inv4.restype = C.c_uint32
        L.orc_ray_octant_inv4.argtypes = [P]
        L.orc_bvh2_validate.restype = C.c_int
        L.orc_bvh2_validate.argtypes = [P, C.c_uint64, P, C.c_uint64, P, C.c_int, P]
        L.orc_cwbvh_validate.restype = C.c_int
        L.orc_cwbvh_validate.argtypes = [P, C.c_uint64, P, C.c_uint64, P, P]
        L.orc_now.restype = C.c_double
        L.orc_log2f.restype = C.c_float
        L.orc_log2f.argtypes = [C.c_float]
        L.orc_set_mutation.argtypes = [C.c_int]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_presplit.restype = C.c_int
        L.orc_presplit.argtypes = [P, C.c_uint64, C.c_uint64, P, P, C.c_uint64, P, P, P]
        L.orc_split_triangle.argtypes = [P, C.c_uint32, C.c_float, P, P]
        L.orc_reinsertion.argtypes = [P, C.c_uint64, C.c_float, P, C.c_uint64]
        L.orc_find_reinsertion.argtypes = [P, C.c_uint64, C.c_uint32, P, P, P]
        L.orc_reinsert_node.argtypes = [P, C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_refit_all.argtypes = [P, C.c_uint64]
        L.orc_tlas_traverse.argtypes = [P, C.c_uint64, P, P, P, P, P, P, C.c_uint32, P, C.c_uint64, P]
        L.orc_rebuild_path_flags.restype = C.c_int
        L.orc_rebuild_path_flags.argtypes = [P, C.c_uint64, P, C.c_uint64, P]
        L.orc_partial_rebuild.restype = C.c_int
        L.orc_partial_rebuild.argtypes = [P, C.c_uint64, P, C.c_uint32, C.c_uint32, P, C.c_uint32]
        L.orc_full_rebuild.restype = C.c_int
        L.orc_full_rebuild.argtypes = [P, C.c_uint64, C.c_uint32, C.c_uint32, P, P, C.c_uint32]
        L.orc_cw_cost_table.restype = C.c_int
        L.orc_cw_cost_table.argtypes = [P, C.c_uint64, C.c_uint32, P, P, P, P]
        L.orc_primitives_to_nodes.restype = C.c_int
        L.orc_primitives_to_nodes.argtypes = [P, C.c_uint64, P, C.c_uint64, P]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code):
        self.code = code
        super().__init__(f"oracle error code {code}")


def split_by_3_u128(a):
    """split_by_3_u128 (morton.rs:66-75) as a Python int."""
    out = np.zeros(2, np.uint64)
    lib().orc_split_by_3_u128(int(a), _ptr(out))
    return int(out[0]) | int(out[1]) << 64


def morton_encode_u128_unorm(x, y, z):
    """morton_encode_u128_unorm (morton.rs:85-88) as a Python int."""
    out = np.zeros(2, np.uint64)
    lib().orc_morton_encode_u128_unorm(float(x), float(y), float(z), _ptr(out))
    return int(out[0]) | int(out[1]) << 64


def tri_aabbs(tris):
    from obvhs_amd.types import AABB
    out = np.zeros(len(tris), dtype=AABB)
    lib().orc_tri_aabbs(_ptr(tris), len(tris), _ptr(out))
    return out


def ploc_build(aabbs, search_distance=14, search_depth_threshold=0, indices=None, basic_search=False,
               sort_precision=64):
    """Returns dict(nodes (Bvh2Node[2n-1]), prim_indices, codes, sorted_index, iter_counts, depth, max_offset);
    codes are u64[n] for SortPrecision::U64 and u64[n, 2] (low, high words) for U128."""
    from obvhs_amd.types import BVH2_NODE
    n = len(aabbs)
    aabbs = np.ascontiguousarray(aabbs)
    idx = np.arange(n, dtype=np.uint32) if indices is None else np.ascontiguousarray(indices, dtype=np.uint32)
    nodes = np.zeros(max(2 * n - 1, 0), dtype=BVH2_NODE)
    prims = np.zeros(n, dtype=np.uint32)
    codes = np.zeros(n, dtype=np.uint64)
    codes_hi = np.zeros(n, dtype=np.uint64)
    order = np.zeros(n, dtype=np.uint32)
    cap = 1 << 20
    iters = np.zeros(cap, dtype=np.uint32)
    ni, depth, maxoff = C.c_uint32(), C.c_uint32(), C.c_uint32()
    rc = lib().orc_ploc_build(_ptr(aabbs), _ptr(idx), n, search_distance, search_depth_threshold,
                              int(basic_search), _ptr(nodes), _ptr(prims), _ptr(codes), _ptr(order), _ptr(iters),
                              cap, C.byref(ni), C.byref(depth), C.byref(maxoff), sort_precision, _ptr(codes_hi))
    if rc != 0:
        raise OracleError(rc)
    if sort_precision == 128:
        codes = np.stack([codes, codes_hi], axis=1)
    return dict(nodes=nodes, prim_indices=prims, codes=codes, sorted_index=order,
                iter_counts=iters[:min(ni.value, cap)].copy(), depth=depth.value, max_offset=maxoff.value)


def bvh2_to_cwbvh(nodes, prim_indices, max_prims_per_leaf=3, order_children=True):
    from obvhs_amd.types import CWBVH_NODE
    nn = len(nodes)
    np_ = len(prim_indices)
    cap = max(nn, 1)
    out = np.zeros(cap, dtype=CWBVH_NODE)
    prims = np.zeros(max(np_, 1), dtype=np.uint32)
    cw2b = np.zeros(cap, dtype=np.uint32)
    nnodes, nprims = C.c_uint64(), C.c_uint64()
    rc = lib().orc_bvh2_to_cwbvh(_ptr(np.ascontiguousarray(nodes)), nn, _ptr(np.ascontiguousarray(prim_indices)),
                                 np_, max_prims_per_leaf, int(order_children), _ptr(out), cap, _ptr(prims),
                                 len(prims), C.byref(nnodes), C.byref(nprims), _ptr(cw2b))
    if rc != 0:
        raise OracleError(rc)
    return out[:nnodes.value].copy(), prims[:nprims.value].copy(), cw2b[:nnodes.value].copy()


def cw_cost_table(nodes, max_prims_per_leaf=3):
    """The conversion's cost table (bvh2_to_cwbvh.rs:220-344): cost, kind (0 LEAF, 1 INTERNAL, 2 DISTRIBUTE),
    distribute_left, distribute_right, each of shape (len(nodes), 7)."""
    nn = len(nodes)
    cost = np.zeros(nn * 7, dtype=np.float32)
    kind, dl, dr = (np.zeros(nn * 7, dtype=np.uint8) for _ in range(3))
    rc = lib().orc_cw_cost_table(_ptr(np.ascontiguousarray(nodes)), nn, max_prims_per_leaf, _ptr(cost), _ptr(kind),
                                 _ptr(dl), _ptr(dr))
    if rc != 0:
        raise OracleError(rc)
    return tuple(a.reshape(nn, 7) for a in (cost, kind, dl, dr))


def traverse(cw_nodes, bvh_tris, rays, mode=0, threads=1, counts=False):
    from obvhs_amd.types import RAY_HIT
    n = len(rays)
    hits = np.zeros(n, dtype=RAY_HIT) if (mode & 1) == 0 else None  # modes 2, 3: the watertight test
    miss = np.zeros(n, dtype=np.uint8) if (mode & 1) == 1 else None
    nv = np.zeros(n, dtype=np.uint32) if counts else None
    tv = np.zeros(n, dtype=np.uint32) if counts else None
    lib().orc_cwbvh_traverse(_ptr(np.ascontiguousarray(cw_nodes)), len(cw_nodes),
                             _ptr(np.ascontiguousarray(bvh_tris)), _ptr(np.ascontiguousarray(rays)), n, mode,
                             _ptr(hits), _ptr(miss), _ptr(nv), _ptr(tv), threads)
    res = hits if (mode & 1) == 0 else miss.astype(bool)
    if counts:
        return res, nv, tv
    return res


def brute_force_wt(tris, rays):
    """Closest hit over all triangles with the watertight test (first index on equal t)."""
    from obvhs_amd.types import RAY_HIT
    out = np.zeros(len(rays), dtype=RAY_HIT)
    lib().orc_brute_force_wt(_ptr(np.ascontiguousarray(tris)), len(tris), _ptr(np.ascontiguousarray(rays)),
                             len(rays), _ptr(out))
    return out


def tri_intersect_wt(tri, ray):
    return lib().orc_tri_intersect_wt(_ptr(np.ascontiguousarray(tri)), _ptr(np.ascontiguousarray(ray)))


def brute_force(tris, rays):
    from obvhs_amd.types import RAY_HIT
    out = np.zeros(len(rays), dtype=RAY_HIT)
    lib().orc_brute_force(_ptr(np.ascontiguousarray(tris)), len(tris), _ptr(np.ascontiguousarray(rays)), len(rays),
                          _ptr(out))
    return out


def bvh2_validate(nodes, prim_indices, prim_aabbs, tight_fit=True, ordered=True, splits=False):
    md = C.c_uint32()
    flags = int(bool(tight_fit)) | (0 if ordered else 2) | (4 if splits else 0)
    rc = lib().orc_bvh2_validate(_ptr(np.ascontiguousarray(nodes)), len(nodes),
                                 _ptr(np.ascontiguousarray(prim_indices)), len(prim_indices),
                                 _ptr(np.ascontiguousarray(prim_aabbs)), flags, C.byref(md))
    return rc, md.value


def cwbvh_validate(cw_nodes, prim_indices, prim_aabbs):
    md = C.c_uint32()
    rc = lib().orc_cwbvh_validate(_ptr(np.ascontiguousarray(cw_nodes)), len(cw_nodes),
                                  _ptr(np.ascontiguousarray(prim_indices)), len(prim_indices),
                                  _ptr(np.ascontiguousarray(prim_aabbs)), C.byref(md))
    return rc, md.value


def presplit(tris, max_total=0):
    """Leaf boxes and primitive indices after the pre-splits of build_cwbvh_from_tris (cwbvh/builder.rs:26-55)."""
    from obvhs_amd.types import AABB
    n = len(tris)
    cap = max(1, n * 8 if max_total == 0 else max_total)
    aabbs = np.zeros(cap, dtype=AABB)
    idx = np.zeros(cap, dtype=np.uint32)
    cnt, avg, big = C.c_uint64(), C.c_float(), C.c_float()
    rc = lib().orc_presplit(_ptr(np.ascontiguousarray(tris)), n, max_total, _ptr(aabbs), _ptr(idx), cap,
                            C.byref(cnt), C.byref(avg), C.byref(big))
    if rc != 0:
        raise OracleError(rc)
    return aabbs[:cnt.value].copy(), idx[:cnt.value].copy(), avg.value, big.value


def split_triangle(tri, dim, pos):
    from obvhs_amd.types import AABB
    left, right = np.zeros(1, dtype=AABB), np.zeros(1, dtype=AABB)
    lib().orc_split_triangle(_ptr(np.ascontiguousarray(tri)), dim, pos, _ptr(left), _ptr(right))
    return left[0], right[0]


def reinsertion(nodes, ratio, seq=None):
    """ReinsertionOptimizer::default().run(bvh, ratio, seq) on a copy of the nodes (reinsertion.rs:40-61)."""
    out = np.ascontiguousarray(nodes).copy()
    s = None if seq is None else np.ascontiguousarray(seq, dtype=np.float32)
    lib().orc_reinsertion(_ptr(out), len(out), ratio, _ptr(s), 0 if s is None else len(s))
    return out


def find_reinsertion(nodes, node_id):
    f, t, a = C.c_uint32(), C.c_uint32(), C.c_float()
    lib().orc_find_reinsertion(_ptr(np.ascontiguousarray(nodes)), len(nodes), node_id, C.byref(f), C.byref(t),
                               C.byref(a))
    return f.value, t.value, a.value


def reinsert_node(nodes, frm, to):
    out = np.ascontiguousarray(nodes).copy()
    lib().orc_reinsert_node(_ptr(out), len(out), frm, to)
    return out


def refit_all(nodes):
    out = np.ascontiguousarray(nodes).copy()
    lib().orc_refit_all(_ptr(out), len(out))
    return out


def rebuild_path_flags(nodes, leaves):
    """compute_rebuild_path_flags: one byte per node."""
    nodes = np.ascontiguousarray(nodes)
    leaves = np.ascontiguousarray(leaves, dtype=np.uint32)
    flags = np.zeros(len(nodes), np.uint8)
    rc = lib().orc_rebuild_path_flags(_ptr(nodes), len(nodes), _ptr(leaves), len(leaves), _ptr(flags))
    return rc, flags


def partial_rebuild(nodes, flags, search_distance=1, search_depth_threshold=0, sort_precision=64):
    """PlocBuilder::partial_rebuild on a copy; returns (rc, nodes, pass count)."""
    out = np.ascontiguousarray(nodes).copy()
    flags = np.ascontiguousarray(flags, dtype=np.uint8)
    assert len(flags) == len(out)
    depth = C.c_uint32()
    rc = lib().orc_partial_rebuild(_ptr(out), len(out), _ptr(flags), search_distance, search_depth_threshold,
                                   C.byref(depth), sort_precision)
    return rc, out, depth.value


def full_rebuild(nodes, search_distance=1, search_depth_threshold=0, sort_precision=64):
    """PlocBuilder::full_rebuild on a copy; returns (rc, nodes, pass count)."""
    out = np.ascontiguousarray(nodes).copy()
    nn, depth = C.c_uint64(), C.c_uint32()
    rc = lib().orc_full_rebuild(_ptr(out), len(out), search_distance, search_depth_threshold, C.byref(nn),
                                C.byref(depth), sort_precision)
    return rc, out[:nn.value] if rc == 0 else out, depth.value


def primitives_to_nodes(nodes, prim_indices):
    prim_indices = np.ascontiguousarray(prim_indices, dtype=np.uint32)
    out = np.zeros(len(prim_indices), np.uint32)
    rc = lib().orc_primitives_to_nodes(_ptr(np.ascontiguousarray(nodes)), len(nodes), _ptr(prim_indices),
                                       len(prim_indices), _ptr(out))
    return rc, out


def build_cwbvh(tris, search_distance=14, search_depth_threshold=0, max_prims_per_leaf=3, pre_split=False,
                reinsertion_batch_ratio=0.0, ratio_sequence=None, split_max_total=0, sort_precision=64):
    """Oracle build_cwbvh_from_tris (cwbvh/builder.rs:20-85): [pre-splits] -> PLOC -> [reinsertion] ->
    bvh2_to_cwbvh(.., clamp(max_prims, 1, 3), true, false).  Returns (bvh2 dict, cw nodes, cw primitive indices);
    the dict's "nodes" are the BVH2 after reinsertion."""
    if pre_split and len(tris):
        aabbs, idx, _, _ = presplit(tris, split_max_total)
    else:
        aabbs, idx = tri_aabbs(tris), None
    b = ploc_build(aabbs, search_distance, search_depth_threshold, indices=idx, sort_precision=sort_precision)
    if reinsertion_batch_ratio > 0.0 and len(b["nodes"]):
        b["nodes"] = reinsertion(b["nodes"], reinsertion_batch_ratio, ratio_sequence)
    cw, prims, _ = bvh2_to_cwbvh(b["nodes"], b["prim_indices"], max(1, min(3, max_prims_per_leaf)), True)
    return b, cw, prims


def tlas_traverse(tlas_nodes, tlas_prims, blas_list, rays):
    """CwBvh::ray_traverse_tlas_blas (cwbvh/mod.rs:341-491) for every ray.  blas_list: [(cw_nodes, bvh_tris)]."""
    from obvhs_amd.types import CWBVH_NODE, RAY_HIT, TRIANGLE
    nodes = np.concatenate([b[0] for b in blas_list]) if blas_list else np.zeros(0, dtype=CWBVH_NODE)
    tris = np.concatenate([b[1] for b in blas_list]) if blas_list else np.zeros(0, dtype=TRIANGLE)
    ncount = np.array([len(b[0]) for b in blas_list], dtype=np.uint32)
    tcount = np.array([len(b[1]) for b in blas_list], dtype=np.uint32)
    noff = np.concatenate([[0], np.cumsum(ncount)[:-1]]).astype(np.uint32)
    toff = np.concatenate([[0], np.cumsum(tcount)[:-1]]).astype(np.uint32)
    out = np.zeros(len(rays), dtype=RAY_HIT)
    lib().orc_tlas_traverse(_ptr(np.ascontiguousarray(tlas_nodes)), len(tlas_nodes),
                            _ptr(np.ascontiguousarray(tlas_prims, dtype=np.uint32)), _ptr(nodes), _ptr(noff),
                            _ptr(ncount), _ptr(tris), _ptr(toff), len(blas_list),
                            _ptr(np.ascontiguousarray(rays)), len(rays), _ptr(out))
    return out


class threads:
    """Context manager: the oracle's builds run multithreaded (the crate's `parallel` sections) inside it."""

    def __init__(self, n):
        self.n = n

    def __enter__(self):
        lib().orc_set_threads(self.n)
        return self

    def __exit__(self, *exc):
        lib().orc_set_threads(1)
