"""Meshes and integer-hash sampling (test_util.rs), vectorised with numpy.

The hash functions are exact u32 arithmetic; the meshes follow the crate's f32 operation order (glam normalize is
v * (1 / sqrt(dot(v, v))), dot = (x*x' + y*y') + z*z')."""
import numpy as np

from .types import triangles_from_vertices

_U = np.uint32


def uhash(x):
    """test_util.rs:9-17 (lowbias32)."""
    x = np.asarray(x, dtype=np.uint32)
    with np.errstate(over="ignore"):
        x = x ^ (x >> _U(16))
        x = (x * _U(0x7FEB352D)).astype(np.uint32)
        x = x ^ (x >> _U(15))
        x = (x * _U(0x846CA68B)).astype(np.uint32)
        x = x ^ (x >> _U(16))
    return x.astype(np.uint32)


def uhash2(a, b):
    """test_util.rs:33-35"""
    a = np.asarray(a, dtype=np.uint32)
    b = np.asarray(b, dtype=np.uint32)
    with np.errstate(over="ignore"):
        return uhash((a * _U(1597334673)).astype(np.uint32) ^ (b * _U(3812015801)).astype(np.uint32))


def unormf(n):
    """test_util.rs:38-40: n as f32 * (1 / 4294967296.0f32)."""
    return (np.asarray(n, dtype=np.uint32).astype(np.float32) * np.float32(1.0 / 4294967296.0)).astype(np.float32)


def hash_noise(x, y, frame):
    """test_util.rs:43-46: unormf(uhash2(x, (y << 11) + frame))."""
    y = np.asarray(y, dtype=np.uint32)
    with np.errstate(over="ignore"):
        b = ((y << _U(11)) + np.asarray(frame, dtype=np.uint32)).astype(np.uint32)
    return unormf(uhash2(x, b))


def _normalize(v):
    v = np.asarray(v, dtype=np.float32)
    d = (v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2]
    r = (np.float32(1.0) / np.sqrt(d)).astype(np.float32)
    return (v * r[..., None]).astype(np.float32)


def _vec(a, b, c):
    return np.array([a, b, c], dtype=np.float32)


# test_util.rs:184-197: cube with side length 2 centred at the origin.
CUBE = triangles_from_vertices(
    np.array([[-1, 1, -1], [1, 1, 1], [-1, 1, 1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
              [-1, 1, -1], [1, 1, 1], [-1, 1, 1], [1, -1, -1], [1, 1, -1], [-1, 1, -1]], dtype=np.float32),
    np.array([[1, 1, 1], [-1, -1, 1], [-1, -1, -1], [-1, -1, 1], [1, -1, 1], [1, -1, -1],
              [-1, 1, 1], [-1, 1, 1], [-1, 1, -1], [1, -1, 1], [1, 1, 1], [1, 1, -1]], dtype=np.float32),
    np.array([[1, 1, -1], [1, -1, 1], [-1, -1, 1], [-1, -1, -1], [1, -1, -1], [-1, -1, -1],
              [1, 1, 1], [-1, -1, 1], [-1, -1, -1], [-1, -1, 1], [1, -1, 1], [1, -1, -1]], dtype=np.float32))

# test_util.rs:200-203: plane with side length 2 centred at the origin (normal +Y).
PLANE = triangles_from_vertices(np.array([[1, 0, 1], [1, 0, 1]], dtype=np.float32),
                                np.array([[-1, 0, -1], [1, 0, -1]], dtype=np.float32),
                                np.array([[-1, 0, 1], [-1, 0, -1]], dtype=np.float32))


def icosphere(subdivisions):
    """test_util.rs:205-252 (unit radius: the base vertices are normalised)."""
    phi = (np.float32(1.0) + np.sqrt(np.float32(5.0))) / np.float32(2.0)
    a, b, c, d, e = np.float32(1.0), np.float32(-1.0), np.float32(0.0), phi, -phi
    p = np.array([_vec(b, d, c), _vec(a, d, c), _vec(b, e, c), _vec(a, e, c), _vec(c, b, d), _vec(c, a, d),
                  _vec(c, b, e), _vec(c, a, e), _vec(d, c, b), _vec(d, c, a), _vec(e, c, b), _vec(e, c, a)])
    p = _normalize(p)
    idx = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                    [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                    [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    t0, t1, t2 = p[idx[:, 0]], p[idx[:, 1]], p[idx[:, 2]]
    half = np.float32(0.5)
    for _ in range(subdivisions):
        m01 = _normalize((t0 + t1) * half)
        m12 = _normalize((t1 + t2) * half)
        m20 = _normalize((t2 + t0) * half)
        # new_tris per old triangle: (v0,m01,m20), (v1,m12,m01), (v2,m20,m12), (m01,m12,m20)
        n0 = np.stack([t0, t1, t2, m01], axis=1)
        n1 = np.stack([m01, m12, m20, m12], axis=1)
        n2 = np.stack([m20, m01, m12, m20], axis=1)
        t0, t1, t2 = (x.reshape(-1, 3) for x in (n0, n1, n2))
    return triangles_from_vertices(t0, t1, t2)


def height_to_triangles(height_map, x_resolution, z_resolution):
    """test_util.rs:255-287.  height_map(x_index_array, z_index_array) -> f32 array."""
    xs, zs = np.meshgrid(np.arange(x_resolution), np.arange(z_resolution))
    xs, zs = xs.ravel(), zs.ravel()
    xr, zr = np.float32(x_resolution), np.float32(z_resolution)
    two, one = np.float32(2.0), np.float32(1.0)
    fx = (xs.astype(np.float32) / xr) * two - one
    fz = (zs.astype(np.float32) / zr) * two - one
    fx2 = ((xs + 1).astype(np.float32) / xr) * two - one
    fz2 = ((zs + 1).astype(np.float32) / zr) * two - one
    h = lambda x, z: np.asarray(height_map(x, z), dtype=np.float32) * np.ones_like(fx)
    v00 = np.stack([fx, h(xs, zs), fz], axis=1)
    v10 = np.stack([fx2, h(xs + 1, zs), fz], axis=1)
    v01 = np.stack([fx, h(xs, zs + 1), fz2], axis=1)
    v11 = np.stack([fx2, h(xs + 1, zs + 1), fz2], axis=1)
    a0 = np.stack([v00, v10], axis=1).reshape(-1, 3)
    a1 = np.stack([v01, v01], axis=1).reshape(-1, 3)
    a2 = np.stack([v10, v11], axis=1).reshape(-1, 3)
    return triangles_from_vertices(a0, a1, a2)


def cosine_sample_hemisphere(u0, u1):
    """test_util.rs:63-71 (f32)."""
    u0 = np.asarray(u0, dtype=np.float32)
    u1 = np.asarray(u1, dtype=np.float32)
    r = np.sqrt(u0)
    theta = u1 * np.float32(2.0 * np.pi)
    z = np.sqrt(np.maximum(np.float32(0.0), np.float32(1.0) - u0))
    return np.stack([r * np.cos(theta), r * np.sin(theta), z], axis=-1).astype(np.float32)


def uniform_sample_sphere(u0, u1):
    """test_util.rs:74-79 (f32)."""
    u0 = np.asarray(u0, dtype=np.float32)
    u1 = np.asarray(u1, dtype=np.float32)
    z = np.float32(1.0) - np.float32(2.0) * u0
    r = np.sqrt(np.maximum(np.float32(0.0), np.float32(1.0) - z * z))
    theta = u1 * np.float32(2.0 * np.pi)
    return np.stack([r * np.cos(theta), r * np.sin(theta), z], axis=-1).astype(np.float32)


def build_orthonormal_basis(n):
    """test_util.rs:50-60: columns (tangent, bitangent, n) for unit normals n[..., 3] (f32)."""
    n = np.asarray(n, dtype=np.float32)
    sign = np.where(np.signbit(n[..., 2]), np.float32(-1.0), np.float32(1.0)).astype(np.float32)
    a = np.float32(-1.0) / (sign + n[..., 2])
    b = n[..., 0] * n[..., 1] * a
    t = np.stack([np.float32(1.0) + sign * n[..., 0] * n[..., 0] * a, sign * b, -sign * n[..., 0]], axis=-1)
    bt = np.stack([b, sign + n[..., 1] * n[..., 1] * a, -n[..., 1]], axis=-1)
    return t.astype(np.float32), bt.astype(np.float32), n
