"""Seeded synthetic workloads (triangle scenes and ray batches) for the benchmark and the parity tests.

All randomness comes from the crate's integer hash (test_util.rs:9-46) so inputs are identical on every machine.

* soup(n)            -- n small triangles: centroids U[0,1]^3, vertex offsets U[-0.005, 0.005]^3.
* pinhole_rays(...)  -- primary rays of a pinhole camera (pixel centres, vertical field of view).
* heightfield_scene  -- a g x g grid over [0,1]^2 (two triangles per cell), height = sum of 8 seeded sines with
                        frequencies in [1,32], phases in [0,2pi), amplitudes proportional to 1/frequency, plus seeded
                        12-triangle boxes resting on the surface.
* bounce_rays        -- one cosine-weighted diffuse bounce per primary hit (geometric normal, seeded).
* thin_triangles     -- long thin triangles (aspect ratio >= 50) scattered over the unit square, to exercise
                        spatial pre-splits.
* large_workload     -- a 2896 x 2896 heightfield plus 500,000 thin triangles, incoherent closest-hit rays (origins
                        uniform on a sphere around the scene, uniform directions) and shadow rays from seeded
                        surface points toward a seeded point light (tmax = distance to the light).
* serving_scenes     -- resident scenes for the serving workload: sizes log-uniform in [100K, 4M] triangles, even
                        indices heightfields (grid sized to the count), odd indices soups.
* jitter_triangles   -- seeded per-vertex jitter of +-frac of the mesh's AABB extent (refit / rebuild requests).
* icosphere_batch    -- many independent icosphere meshes (counts log-uniform, rounded down to 20*4^s), vertex
                        radius jitter +-10% keyed on the vertex position (shared vertices move together), and a
                        ray batch per mesh from a radius-2 sphere toward uniform points in the mesh AABB.
"""
import numpy as np

from .test_util import (CUBE, build_orthonormal_basis, cosine_sample_hemisphere, hash_noise, icosphere, uhash,
                        uhash2, unormf, uniform_sample_sphere)
from .types import RAY, make_rays, safe_inverse, triangles_from_vertices


def soup(n, seed=0):
    """Triangle soup of n triangles (65,536 for the CPU-sized plain workload)."""
    i = np.arange(n, dtype=np.uint32)
    c = np.stack([hash_noise(i, a, seed) for a in range(3)], axis=1)
    verts = []
    for k in range(3):
        off = np.stack([hash_noise(i, 3 + 3 * k + a, seed) for a in range(3)], axis=1)
        verts.append(c + (off - np.float32(0.5)) * np.float32(0.01))
    return triangles_from_vertices(*verts)


def signed_zero_soup(n, seed=0):
    """n triangles with vertices on a 1/4 lattice of [-1, 1]^3 (many coordinates exactly zero), each zero
    coordinate given a seeded sign: boxes that are equal in value but differ in the sign of a zero, the inputs on
    which refit_from_fast's early exit (bvh2/mod.rs:722-751) decides which bits a box keeps."""
    i = np.arange(n, dtype=np.uint32)
    q = np.float32(0.25)
    c = np.stack([np.floor(hash_noise(i, a, seed) * np.float32(9.0)) - np.float32(4.0) for a in range(3)], axis=1)
    verts = []
    for k in range(3):
        off = np.stack([np.floor(hash_noise(i, 3 + 3 * k + a, seed) * np.float32(3.0)) - np.float32(1.0)
                        for a in range(3)], axis=1)
        v = ((c + off) * q).astype(np.float32)
        neg = np.stack([hash_noise(i, 12 + 3 * k + a, seed) < np.float32(0.5) for a in range(3)], axis=1)
        v = np.where((v == 0) & neg, np.float32(-0.0), np.where(v == 0, np.float32(0.0), v)).astype(np.float32)
        verts.append(v)
    return triangles_from_vertices(*verts)


def slab_edge_rays(tris, planes, n, seed=0):
    """n rays whose CWBVH node-test slab terms (cwbvh/simd.rs:17-100) meet special values, built field by field (the
    crate's Ray fields are public, so a caller's inverse direction need not be Ray::new's safe_inverse):
      * directions axis-aligned, in a coordinate plane with a +0 or -0 component, with a component of 1e-30 or
        1e-38, or general;
      * on an axis whose direction component is zero or tiny, the origin coordinate is, most of the time, that of a
        node plane from `planes` (CWBVH node origins p): (p - o) * inv is then 0 * inf = NaN where inv is infinite;
      * inverse directions 1/d (infinite on a zero component) or safe_inverse(d), a few with a NaN component;
      * origins up to 3e38 out on one axis (slab terms overflow to infinity);
      * tmax infinite, 1e30, NaN, 0, -1 or half the scene's extent; tmin 0 or 1e-3."""
    rng = np.random.default_rng(seed)
    v = np.concatenate([tris["v0"][:, :3], tris["v1"][:, :3], tris["v2"][:, :3]]).astype(np.float64)
    lo, hi = v.min(axis=0), v.max(axis=0)
    o = lo + (hi - lo) * (rng.random((n, 3)) * 1.4 - 0.2)
    d = rng.normal(size=(n, 3))
    kind = rng.integers(0, 6, n)
    ax = rng.integers(0, 3, n)
    rows = np.arange(n)
    m = kind == 0  # axis-aligned
    d[m] = 0.0
    d[m, ax[m]] = rng.choice([-1.0, 1.0], int(m.sum()))
    m = kind == 1  # in a coordinate plane; the zero's sign drawn
    d[m, ax[m]] = np.where(rng.random(int(m.sum())) < 0.5, -0.0, 0.0)
    m = kind == 2  # one tiny component
    d[m, ax[m]] = rng.choice([1e-30, -1e-30, 1e-38, -1e-38], int(m.sum()))
    m = kind == 3  # far out on one axis, aimed back at the scene
    far = rng.choice([-3e38, 3e38, -1e30, 1e30], int(m.sum()))
    o[m, ax[m]] = far
    d[m, ax[m]] = -np.sign(far) * rng.uniform(1e-3, 1.0, int(m.sum()))
    d = d.astype(np.float32)
    p = np.asarray(planes, dtype=np.float32).reshape(-1, 3)[rng.integers(0, len(planes), n)]
    on = (np.abs(d) < np.float32(1e-20)) & (rng.random((n, 3)) < 0.7)
    o = np.where(on, p, o.astype(np.float32)).astype(np.float32)
    with np.errstate(divide="ignore", over="ignore"):
        inv = np.where(rng.random((n, 1)) < 0.6, np.float32(1.0) / d, safe_inverse(d)).astype(np.float32)
    m = rng.random(n) < 0.02
    inv[rows[m], ax[m]] = np.float32(np.nan)
    out = np.zeros(n, dtype=RAY)
    out["origin"][:, :3] = o
    out["direction"][:, :3] = d
    out["inv_direction"][:, :3] = inv
    out["tmin"] = np.where(rng.random(n) < 0.8, np.float32(0.0), np.float32(1e-3))
    tmax_choices = np.array([np.inf, 1e30, np.nan, 0.0, -1.0, 0.5 * float((hi - lo).max())], dtype=np.float32)
    out["tmax"] = tmax_choices[rng.choice(6, n, p=[0.55, 0.2, 0.08, 0.05, 0.05, 0.07])]
    return out


def look_at_basis(eye, target, up=(0.0, 1.0, 0.0)):
    eye = np.asarray(eye, dtype=np.float64)
    fwd = np.asarray(target, dtype=np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(up, dtype=np.float64))
    if np.linalg.norm(right) < 1e-12:
        right = np.cross(fwd, np.array([1.0, 0.0, 0.0]))
    right /= np.linalg.norm(right)
    upv = np.cross(right, fwd)
    return fwd, right, upv


def pinhole_rays(width, height, eye, forward=(0.0, 0.0, 1.0), up=(0.0, 1.0, 0.0), fov_deg=60.0, tmax=1e30,
                 target=None):
    """Rays through pixel centres; vertical field of view fov_deg.  Directions are normalised in f64 and stored
    in f32 together with Ray::new's safe inverse (ray.rs:34-52)."""
    eye = np.asarray(eye, dtype=np.float64)
    if target is None:
        target = eye + np.asarray(forward, dtype=np.float64)
    fwd, right, upv = look_at_basis(eye, target, up)
    tan_half = np.tan(np.radians(fov_deg) * 0.5)
    aspect = width / height
    ys, xs = np.meshgrid(np.arange(height), np.arange(width), indexing="ij")
    sx = ((xs + 0.5) / width * 2.0 - 1.0) * tan_half * aspect
    sy = (1.0 - (ys + 0.5) / height * 2.0) * tan_half
    d = fwd[None, None, :] + sx[..., None] * right[None, None, :] + sy[..., None] * upv[None, None, :]
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    d = d.reshape(-1, 3).astype(np.float32)
    o = np.broadcast_to(eye.astype(np.float32), d.shape)
    return make_rays(o, d, 0.0, tmax)


def plain_workload(n=65536, seed=0, width=256, height=256):
    """65,536-triangle soup traced by 256x256 pinhole rays from (0.5,0.5,-1.5) looking +z, 60 deg, tmax 1e30."""
    tris = soup(n, seed)
    rays = pinhole_rays(width, height, eye=(0.5, 0.5, -1.5), forward=(0.0, 0.0, 1.0), fov_deg=60.0, tmax=1e30)
    return tris, rays


def _sines(seed, count=8):
    k = np.arange(count, dtype=np.uint32)
    f = 1.0 + 31.0 * hash_noise(k, 100, seed).astype(np.float64)
    phase = 2.0 * np.pi * hash_noise(k, 101, seed).astype(np.float64)
    ang = 2.0 * np.pi * hash_noise(k, 102, seed).astype(np.float64)
    amp = 0.05 / f
    return f, phase, ang, amp


def _height(x, z, sines):
    f, phase, ang, amp = sines
    x = np.asarray(x, dtype=np.float64)[..., None]
    z = np.asarray(z, dtype=np.float64)[..., None]
    return np.sum(amp * np.sin(2.0 * np.pi * f * (x * np.cos(ang) + z * np.sin(ang)) + phase), axis=-1)


def heightfield_scene(grid=724, boxes=20000, seed=0):
    """Heightfield of grid x grid cells (2 triangles each) over [0,1]^2 in x/z with y up, plus seeded boxes."""
    sines = _sines(seed)
    g = np.arange(grid + 1, dtype=np.float64) / grid
    X, Z = np.meshgrid(g, g)  # [z, x]
    H = _height(X, Z, sines).astype(np.float32)
    Xf, Zf = X.astype(np.float32), Z.astype(np.float32)
    v = lambda zi, xi: np.stack([Xf[zi, xi], H[zi, xi], Zf[zi, xi]], axis=-1)
    zi, xi = np.meshgrid(np.arange(grid), np.arange(grid), indexing="ij")
    zi, xi = zi.ravel(), xi.ravel()
    v00, v10, v01, v11 = v(zi, xi), v(zi, xi + 1), v(zi + 1, xi), v(zi + 1, xi + 1)
    t_a = np.stack([v00, v10], axis=1).reshape(-1, 3)
    t_b = np.stack([v01, v01], axis=1).reshape(-1, 3)
    t_c = np.stack([v10, v11], axis=1).reshape(-1, 3)
    terrain = triangles_from_vertices(t_a, t_b, t_c)
    if boxes == 0:
        return terrain
    b = np.arange(boxes, dtype=np.uint32)
    cx = hash_noise(b, 200, seed).astype(np.float64)
    cz = hash_noise(b, 201, seed).astype(np.float64)
    s = 0.001 + 0.003 * hash_noise(b, 202, seed).astype(np.float64)
    cy = _height(cx, cz, sines) + s
    center = np.stack([cx, cy, cz], axis=1)
    cube = [CUBE["v0"][:, :3].astype(np.float64), CUBE["v1"][:, :3].astype(np.float64),
            CUBE["v2"][:, :3].astype(np.float64)]
    bv = [(center[:, None, :] + s[:, None, None] * c[None, :, :]).reshape(-1, 3).astype(np.float32) for c in cube]
    box_tris = triangles_from_vertices(*bv)
    return np.concatenate([terrain, box_tris])


def heightfield_workload(grid=724, boxes=20000, width=1920, height=1080, seed=0):
    """Heightfield scene with 1920x1080 primary rays from above the terrain."""
    tris = heightfield_scene(grid, boxes, seed)
    rays = pinhole_rays(width, height, eye=(0.5, 0.45, -0.35), target=(0.5, 0.0, 0.55), fov_deg=60.0, tmax=1e30)
    return tris, rays


def triangle_normals(tris):
    v0 = tris["v0"][:, :3].astype(np.float64)
    v1 = tris["v1"][:, :3].astype(np.float64)
    v2 = tris["v2"][:, :3].astype(np.float64)
    n = np.cross(v1 - v0, v2 - v0)
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    return n / np.where(ln > 0, ln, 1.0)


def bounce_rays(rays, hits, bvh_tris, seed=0, eps=1e-4, tmax=1e30):
    """One cosine-weighted bounce per hit (hits[i].primitive_id indexes bvh_tris).  Misses produce no ray."""
    hit = np.isfinite(hits["t"]) & (hits["primitive_id"] != 0xFFFFFFFF)
    idx = np.nonzero(hit)[0]
    o = rays["origin"][idx, :3].astype(np.float64)
    d = rays["direction"][idx, :3].astype(np.float64)
    t = hits["t"][idx].astype(np.float64)
    p = o + d * t[:, None]
    n = triangle_normals(bvh_tris[hits["primitive_id"][idx]])
    flip = np.sum(n * d, axis=1) > 0
    n[flip] = -n[flip]
    u0 = hash_noise(idx.astype(np.uint32), 300, seed)
    u1 = hash_noise(idx.astype(np.uint32), 301, seed)
    local = cosine_sample_hemisphere(u0, u1).astype(np.float64)
    tb, bt, nn = build_orthonormal_basis(n.astype(np.float32))
    wd = local[:, 0:1] * tb + local[:, 1:2] * bt + local[:, 2:3] * nn
    wd /= np.linalg.norm(wd, axis=1, keepdims=True)
    origin = p + n * eps
    return make_rays(origin.astype(np.float32), wd.astype(np.float32), 0.0, tmax)


def empty_rays(n):
    return np.zeros(n, dtype=RAY)


def thin_triangles(n, seed=0, min_len=0.05, max_len=0.5):
    """n long thin triangles over [0,1]^2 (heights in [0, 0.3]): length in [min_len, max_len], random direction in
    3D, base width length / 50 .. length / 200 (aspect ratio >= 50)."""
    i = np.arange(n, dtype=np.uint32)
    f = np.uint32(seed) + np.uint32(7000)
    u = [hash_noise(i, a, f).astype(np.float64) for a in range(8)]
    start = np.stack([u[0], u[1] * 0.3, u[2]], axis=1)
    length = min_len + (max_len - min_len) * u[3]
    d = uniform_sample_sphere(u[4].astype(np.float32), u[5].astype(np.float32)).astype(np.float64)
    end = start + d * length[:, None]
    # a unit vector perpendicular to d for the base
    helper = np.where(np.abs(d[:, :1]) < 0.9, np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 1.0, 0.0]]))
    perp = np.cross(d, helper)
    perp /= np.linalg.norm(perp, axis=1, keepdims=True)
    width = length / (50.0 + 150.0 * u[6])
    v0 = start - perp * (width / 2)[:, None]
    v1 = start + perp * (width / 2)[:, None]
    return triangles_from_vertices(v0.astype(np.float32), v1.astype(np.float32), end.astype(np.float32))


def large_workload(grid=2896, thin=500000, n_rays=1 << 22, n_shadow=1 << 22, seed=0):
    """Returns (tris, closest-hit rays, shadow rays, light position)."""
    tris = np.concatenate([heightfield_scene(grid, 0, seed), thin_triangles(thin, seed)])
    f = np.uint32(seed) + np.uint32(9000)
    i = np.arange(n_rays, dtype=np.uint32)
    center = np.array([0.5, 0.1, 0.5], dtype=np.float32)
    o = center + uniform_sample_sphere(hash_noise(i, 0, f), hash_noise(i, 1, f)) * np.float32(0.8)
    d = uniform_sample_sphere(hash_noise(i, 2, f), hash_noise(i, 3, f))
    rays = make_rays(o, d, 0.0, 1e30)
    # shadow rays: seeded points on seeded triangles, nudged along the geometric normal toward the light
    light = np.array([0.2 + 0.6 * float(hash_noise(np.uint32(0), 10, f)), 1.5,
                      0.2 + 0.6 * float(hash_noise(np.uint32(0), 11, f))])
    k = np.arange(n_shadow, dtype=np.uint32)
    t = (uhash2(k, f + np.uint32(7)) % np.uint32(len(tris))).astype(np.int64)
    a = hash_noise(k, 4, f).astype(np.float64)
    b = hash_noise(k, 5, f).astype(np.float64)
    flip = a + b > 1.0
    a[flip], b[flip] = 1.0 - a[flip], 1.0 - b[flip]
    v0 = tris["v0"][t, :3].astype(np.float64)
    p = v0 + a[:, None] * (tris["v1"][t, :3] - v0) + b[:, None] * (tris["v2"][t, :3] - v0)
    nrm = triangle_normals(tris[t])
    to_light = light[None, :] - p
    nrm[np.sum(nrm * to_light, axis=1) < 0] *= -1.0
    p = p + nrm * 1e-4
    to_light = light[None, :] - p
    dist = np.linalg.norm(to_light, axis=1)
    shadow = make_rays(p.astype(np.float32), (to_light / dist[:, None]).astype(np.float32), 0.0, 0.0)
    shadow["tmax"] = (dist * (1.0 - 1e-4)).astype(np.float32)
    return tris, rays, shadow, light


def jitter_triangles(tris, seed=0, frac=0.01):
    """Every vertex moved by U[-frac, frac] of the mesh's AABB extent per axis (seeded)."""
    out = tris.copy()
    if len(tris) == 0:
        return out
    lo = np.minimum(np.minimum(tris["v0"][:, :3].min(axis=0), tris["v1"][:, :3].min(axis=0)),
                    tris["v2"][:, :3].min(axis=0)).astype(np.float64)
    hi = np.maximum(np.maximum(tris["v0"][:, :3].max(axis=0), tris["v1"][:, :3].max(axis=0)),
                    tris["v2"][:, :3].max(axis=0)).astype(np.float64)
    ext = hi - lo
    i = np.arange(len(tris), dtype=np.uint32)
    f = np.uint32(seed) + np.uint32(11000)
    for k, name in enumerate(("v0", "v1", "v2")):
        for a in range(3):
            u = hash_noise(i, 3 * k + a, f).astype(np.float64) * 2.0 - 1.0
            out[name][:, a] = (tris[name][:, a].astype(np.float64) + u * frac * ext[a]).astype(np.float32)
    return out


def serving_scenes(count=64, seed=0, lo=100_000, hi=4_000_000):
    """[(triangles, 'heightfield' | 'soup')]: log-uniform sizes, alternating heightfields and soups (seeded)."""
    k = np.arange(count, dtype=np.uint32)
    u = hash_noise(k, 20, np.uint32(seed) + np.uint32(12000)).astype(np.float64)
    sizes = (lo * np.power(hi / lo, u)).astype(np.int64)
    out = []
    for i, n in enumerate(sizes):
        if i % 2 == 0:
            grid = max(8, int(round(np.sqrt(n / 2.0))))
            out.append((heightfield_scene(grid, 0, seed=seed * 1000 + i), "heightfield"))
        else:
            out.append((soup(int(n), seed=seed * 1000 + i), "soup"))
    return out


ICO_MAX_SUBDIV = 5


def _icosphere_base(s, _cache={}):
    if s not in _cache:
        t = icosphere(s)
        _cache[s] = np.stack([t["v0"][:, :3], t["v1"][:, :3], t["v2"][:, :3]], axis=1).copy()  # (20*4^s, 3, 3)
    return _cache[s]


def icosphere_batch_counts(mesh_count, seed=0, min_tris=64, max_tris=65536):
    """Subdivision level per mesh: a log-uniform count in [min_tris, max_tris] rounded down to 20*4^s (s >= 0)."""
    i = np.arange(mesh_count, dtype=np.uint32)
    u = unormf(uhash2(i, np.uint32(0x9E3779B9) ^ np.uint32(seed))).astype(np.float64)
    c = np.float64(min_tris) * np.power(np.float64(max_tris) / min_tris, u)
    s = np.floor(np.log(c / 20.0) / np.log(4.0)).astype(np.int64)
    return np.clip(s, 0, ICO_MAX_SUBDIV)


def icosphere_batch(mesh_count=4096, seed=0, rays_per_mesh=1024, min_tris=64, max_tris=65536, levels=None):
    """Returns (tris, offsets, rays, ray_mesh): concatenated Triangle array, mesh_count+1 prefix offsets, rays with
    rays_per_mesh rays per mesh (grouped by mesh), and each ray's mesh id.  `levels` overrides the subdivision
    level of each mesh."""
    levels = icosphere_batch_counts(mesh_count, seed, min_tris, max_tris) if levels is None else np.asarray(levels)
    counts = (20 * np.power(4, levels)).astype(np.int64)
    offsets = np.zeros(mesh_count + 1, dtype=np.int64)
    np.cumsum(counts, out=offsets[1:])
    verts = np.empty((int(offsets[-1]), 3, 3), dtype=np.float32)
    mesh_ids = np.arange(mesh_count, dtype=np.uint32)
    for lvl in np.unique(levels):
        sel = np.nonzero(levels == lvl)[0]
        base = _icosphere_base(int(lvl))
        bits = base.view(np.uint32)
        key = uhash2(uhash2(bits[..., 0], bits[..., 1]), bits[..., 2])  # (T, 3): one key per vertex position
        for chunk in np.array_split(sel, max(1, len(sel) * base.shape[0] // (1 << 22) + 1)):
            h = uhash2(key[None, :, :], (mesh_ids[chunk] ^ np.uint32(seed))[:, None, None])
            r = np.float32(1.0) + (unormf(h) * np.float32(2.0) - np.float32(1.0)) * np.float32(0.1)
            v = base[None] * r[..., None]
            for j, m in enumerate(chunk):
                verts[offsets[m]:offsets[m + 1]] = v[j]
    tris = triangles_from_vertices(verts[:, 0], verts[:, 1], verts[:, 2])
    # rays: origins uniform on the radius-2 sphere, directions toward a uniform point of the mesh AABB
    lo = np.empty((mesh_count, 3), dtype=np.float32)
    hi = np.empty((mesh_count, 3), dtype=np.float32)
    for m in range(mesh_count):
        v = verts[offsets[m]:offsets[m + 1]].reshape(-1, 3)
        lo[m], hi[m] = v.min(axis=0), v.max(axis=0)
    n = mesh_count * rays_per_mesh
    k = np.arange(n, dtype=np.uint32)
    ray_mesh = (k // np.uint32(rays_per_mesh)).astype(np.uint32)
    frame = np.uint32(seed) + np.uint32(1000)
    o = uniform_sample_sphere(hash_noise(k, 0, frame), hash_noise(k, 1, frame)) * np.float32(2.0)
    u = np.stack([hash_noise(k, 2 + a, frame) for a in range(3)], axis=1)
    target = lo[ray_mesh] + (hi[ray_mesh] - lo[ray_mesh]) * u
    d = (target - o).astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = make_rays(o, d.astype(np.float32), 0.0, 1e30)
    return tris, offsets.astype(np.uint32), rays, ray_mesh
