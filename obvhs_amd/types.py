"""numpy views of the obvhs plain-data types (byte-identical to the crate's repr(C) layouts).

aabb.rs:11-16 (Aabb, 32 B), triangle.rs:8-13 (Triangle, 48 B), rt_triangle.rs:147-160 (RtTriangle, 64 B),
ray.rs:15-30 (Ray, 64 B), ray.rs:63-70 (RayHit, 16 B),
bvh2/node.rs:38-66 (Bvh2Node, 48 B), bvh2/node.rs:12-36 (small Bvh2Node, 32 B), cwbvh/node.rs:12-54 (CwBvhNode, 80 B).
"""
import numpy as np

INVALID_ID = 0xFFFFFFFF  # ray.rs:72

VEC3A = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("w", "<f4")])
AABB = np.dtype([("min", "<f4", (4,)), ("max", "<f4", (4,))])
TRIANGLE = np.dtype([("v0", "<f4", (4,)), ("v1", "<f4", (4,)), ("v2", "<f4", (4,))])
RT_TRIANGLE = np.dtype([("v0", "<f4", (4,)), ("e1", "<f4", (4,)), ("e2", "<f4", (4,)), ("ng", "<f4", (4,))])
RAY = np.dtype([("origin", "<f4", (4,)), ("direction", "<f4", (4,)), ("inv_direction", "<f4", (4,)),
                ("tmin", "<f4"), ("tmax", "<f4"), ("pad", "<f4", (2,))])
RAY_HIT = np.dtype([("primitive_id", "<u4"), ("geometry_id", "<u4"), ("instance_id", "<u4"), ("t", "<f4")])
BVH2_NODE = np.dtype([("min", "<f4", (4,)), ("max", "<f4", (4,)), ("prim_count", "<u4"), ("first_index", "<u4"),
                      ("meta1", "<u4"), ("meta2", "<u4")])
BVH2_NODE_SMALL = np.dtype([("min", "<f4", (3,)), ("prim_count", "<u4"), ("max", "<f4", (3,)),
                            ("first_index", "<u4")])
CWBVH_NODE = np.dtype([("p", "<f4", (3,)), ("e", "u1", (3,)), ("imask", "u1"), ("child_base_idx", "<u4"),
                       ("primitive_base_idx", "<u4"), ("child_meta", "u1", (8,)),
                       ("child_min_x", "u1", (8,)), ("child_max_x", "u1", (8,)),
                       ("child_min_y", "u1", (8,)), ("child_max_y", "u1", (8,)),
                       ("child_min_z", "u1", (8,)), ("child_max_z", "u1", (8,))])

assert AABB.itemsize == 32 and TRIANGLE.itemsize == 48 and RAY.itemsize == 64 and RAY_HIT.itemsize == 16
assert RT_TRIANGLE.itemsize == 64
assert BVH2_NODE.itemsize == 48 and BVH2_NODE_SMALL.itemsize == 32 and CWBVH_NODE.itemsize == 80

F32_EPS = np.float32(1.1920929e-07)  # f32::EPSILON


def triangles_from_vertices(v0, v1, v2):
    """Packs (n,3) float32 vertex arrays into Triangle[n] (the Vec3A padding lane is zero)."""
    v0, v1, v2 = (np.asarray(v, dtype=np.float32).reshape(-1, 3) for v in (v0, v1, v2))
    out = np.zeros(len(v0), dtype=TRIANGLE)
    out["v0"][:, :3] = v0
    out["v1"][:, :3] = v1
    out["v2"][:, :3] = v2
    return out


def aabbs_from_minmax(lo, hi):
    lo, hi = (np.asarray(v, dtype=np.float32).reshape(-1, 3) for v in (lo, hi))
    out = np.zeros(len(lo), dtype=AABB)
    out["min"][:, :3] = lo
    out["max"][:, :3] = hi
    return out


def safe_inverse(x):
    """ray.rs:6-12: |x| <= EPSILON -> signum(x) / EPSILON (signum(+-0) = +-1), else 1/x, in f32."""
    x = np.asarray(x, dtype=np.float32)
    with np.errstate(divide="ignore"):
        sig = np.where(np.signbit(x), np.float32(-1.0), np.float32(1.0)).astype(np.float32)
        small = np.abs(x) <= F32_EPS
        inv = np.where(small, sig / F32_EPS, np.float32(1.0) / np.where(small, np.float32(1.0), x))
    return inv.astype(np.float32)


def make_rays(origins, directions, tmin=0.0, tmax=np.inf):
    """Ray::new (ray.rs:34-52) for arrays of origins/directions."""
    o = np.asarray(origins, dtype=np.float32).reshape(-1, 3)
    d = np.asarray(directions, dtype=np.float32).reshape(-1, 3)
    n = max(len(o), len(d))
    out = np.zeros(n, dtype=RAY)
    out["origin"][:, :3] = o
    out["direction"][:, :3] = d
    out["inv_direction"][:, :3] = safe_inverse(d)
    out["tmin"] = np.float32(tmin)
    out["tmax"] = np.float32(tmax)
    return out


def ray_hit_none(n):
    """RayHit::none() (ray.rs:76-83)."""
    h = np.zeros(n, dtype=RAY_HIT)
    h["primitive_id"] = INVALID_ID
    h["geometry_id"] = INVALID_ID
    h["instance_id"] = INVALID_ID
    h["t"] = np.inf
    return h
