// CWBVH closest-hit and occlusion traversal on gfx950.
//
// One lane per ray runs the reference's traversal state machine exactly (traverse!, cwbvh/traverse_macro.rs:59-126;
// CwBvh::ray_traverse cwbvh/mod.rs:169-198 and ray_traverse_miss 201-225), so node visit order, primitive test
// order and the resulting hits match the CPU path bit for bit:
//   * node groups and primitive groups are consumed highest bit first (firstbithigh, cwbvh/mod.rs:996-998);
//   * the 32-entry group stack keeps StackStack semantics (faststack.rs:299-310): a push past slot 31 overwrites
//     slot 31, a pop of the empty stack ends traversal;
//   * the node test gives the SSE path's results (cwbvh/simd.rs:17-100: mul then add, SSE min/max, tmin clamped to
//     1e-4, inclusive compare) and the primitive test is Triangle::intersect (triangle.rs:35-76).
// A step of a lane's walk tests at most TRI_PER_STEP pending triangles and then, once none is left, the next node;
// an empty node group is replaced by the stack's top at the end of the step.  Each ray's sequence of tests, pushes
// and pops is the crate's; only how lanes share a wave's steps changes.
// Stack storage: the bottom LDS entries of each lane's stack live in LDS (one 64-lane column per wave, entry-major;
// 4, 8 or 16 entries, the depth the context's tuning table holds for the BVH's size class, 8 by default), deeper
// entries in a per-lane column of an HBM spill array, so LDS does not limit residency.  A launch has one lane per
// ray (up to 16M lanes; beyond that the lanes stride over the rays), or, chunked, rows of rays dealt over the
// resident waves; each lane owns one spill column.  Nodes are fetched as five 16-byte loads, triangles (in
// primitive_indices order) as three.
#include <algorithm>
#include <mutex>

#include "common.hpp"
#include "cwdp.hpp"
#include "internal.hpp"

#ifndef TRAV_LDS_STACK
#define TRAV_LDS_STACK 8
#endif

namespace obvhs {

namespace {

constexpr int TRAV_THREADS = 64;
constexpr int STACK_SIZE = 32;  // TRAVERSAL_STACK_SIZE, cwbvh/mod.rs:60
constexpr int LDS_STACK = TRAV_LDS_STACK;  // entries kept in LDS; STACK_SIZE - LDS_STACK spill to HBM
constexpr float NODE_EPSILON = 0.0001f;  // cwbvh/node.rs:82
#ifndef OBVHS_TRI_PER_STEP
#define OBVHS_TRI_PER_STEP 2
#endif
// Triangle tests a lane makes per step of the walk; a lane with more pending keeps them for the next steps and
// tests its next node only once they are done, so the order of every ray's tests is the crate's.
constexpr int TRI_PER_STEP = OBVHS_TRI_PER_STEP;
constexpr uint64_t TRAV_MAX_LANES = 1u << 24;  // spill columns (224 B each) for at most 16M lanes: 3.8 GB

struct RayData {
    float ox, oy, oz;
    float dx, dy, dz;
    float ix, iy, iz;
    float tmin, tmax;
};

__device__ __forceinline__ uint32_t firstbithigh(uint32_t v) { return 31u - (uint32_t)__clz((int)v); }

// cwbvh/mod.rs:1001-1010
__device__ __forceinline__ uint32_t octant_inv4(float dx, float dy, float dz) {
    return (dx < 0.0f ? 0u : 0x04040404u) | (dy < 0.0f ? 0u : 0x02020202u) | (dz < 0.0f ? 0u : 0x01010101u);
}

__device__ __forceinline__ float byte_f(uint32_t w, int b) { return (float)((w >> (8 * b)) & 0xffu); }

// CwBvhNode::intersect_ray_simd (cwbvh/simd.rs:17-100) with get_child_and_index_bits (cwbvh/node.rs:207-231).
// Node words: q0 = {p.x, p.y, p.z, e|imask}, q1 = {child_base, prim_base, meta lo, meta hi},
// q2 = {min_x lo/hi, max_x lo/hi}, q3 = {min_y, max_y}, q4 = {min_z, max_z}.
// ROBUST (the watertight mode only): each quantised child box widened by one quantum per side and the slab interval
// by a relative 2^-20, so rounding in the quantisation and in the slab arithmetic cannot cull a child the ray
// touches (the crate's test is exact only in exact arithmetic).
// Same hit mask as the SSE code with fewer instructions:
//   * the crate swaps each child's converted near/far planes by the ray's direction signs; swapping the packed plane
//     words once per node selects the same bytes;
//   * SSE max(a, max(b, c)) is NaN exactly when c is, and otherwise the largest non-NaN of a, b, c, which IEEE
//     max3 (v_max3_f32) also gives; only the z slab (c) needs a NaN check, after which the clamp to
//     [1e-4, ray tmax] follows: tmin = 1e-4 when the z entry is NaN, tmax = ray tmax when the z exit is NaN (min
//     likewise).  A NaN ray tmax fails every SSE compare; -inf stands in for it.  Results can differ from the SSE
//     chain only in the sign of a zero, which no compare sees.
template <bool ROBUST = false>
__device__ __forceinline__ uint32_t intersect_node(const uint4& q0, const uint4& q1, const uint4& q2,
                                                   const uint4& q3, const uint4& q4, const RayData& r,
                                                   float ray_tmax, uint32_t oct_inv4) {
    const float ex = __uint_as_float((q0.w & 0xffu) << 23);
    const float ey = __uint_as_float(((q0.w >> 8) & 0xffu) << 23);
    const float ez = __uint_as_float(((q0.w >> 16) & 0xffu) << 23);
    const float adx = ex * r.ix, ady = ey * r.iy, adz = ez * r.iz;
    const float aox = (__uint_as_float(q0.x) - r.ox) * r.ix;
    const float aoy = (__uint_as_float(q0.y) - r.oy) * r.iy;
    const float aoz = (__uint_as_float(q0.z) - r.oz) * r.iz;
    const bool rdx = r.dx < 0.0f, rdy = r.dy < 0.0f, rdz = r.dz < 0.0f;

    uint64_t oct_inv8 = (uint64_t)oct_inv4;
    oct_inv8 |= oct_inv8 << 32;
    const uint64_t meta8 = (uint64_t)q1.z | ((uint64_t)q1.w << 32);
    const uint64_t is_inner8 = (meta8 & (meta8 << 1)) & 0x1010101010101010ull;
    const uint64_t inner_mask8 = (is_inner8 >> 4) * 0xffull;
    const uint64_t bit_index8 = (meta8 ^ (oct_inv8 & inner_mask8)) & 0x1f1f1f1f1f1f1f1full;
    const uint64_t child_bits8 = (meta8 >> 5) & 0x0707070707070707ull;

    // near (entry) and far (exit) plane words per axis and half
    const uint32_t lx0 = rdx ? q2.z : q2.x, hx0 = rdx ? q2.x : q2.z, lx1 = rdx ? q2.w : q2.y, hx1 = rdx ? q2.y : q2.w;
    const uint32_t ly0 = rdy ? q3.z : q3.x, hy0 = rdy ? q3.x : q3.z, ly1 = rdy ? q3.w : q3.y, hy1 = rdy ? q3.y : q3.w;
    const uint32_t lz0 = rdz ? q4.z : q4.x, hz0 = rdz ? q4.x : q4.z, lz1 = rdz ? q4.w : q4.y, hz1 = rdz ? q4.y : q4.w;
    // ROBUST: the near plane moves one quantum outward (down the axis for a positive direction), the far plane up
    const float plx = ROBUST ? (rdx ? 1.0f : -1.0f) : 0.0f;
    const float ply = ROBUST ? (rdy ? 1.0f : -1.0f) : 0.0f;
    const float plz = ROBUST ? (rdz ? 1.0f : -1.0f) : 0.0f;
    const float rt = ray_tmax == ray_tmax ? ray_tmax : -INFINITY;

    uint32_t hit_mask = 0u;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int h = c >> 2, b = c & 3;
        float xmin = byte_f(h ? lx1 : lx0, b), xmax = byte_f(h ? hx1 : hx0, b);
        float ymin = byte_f(h ? ly1 : ly0, b), ymax = byte_f(h ? hy1 : hy0, b);
        float zmin = byte_f(h ? lz1 : lz0, b), zmax = byte_f(h ? hz1 : hz0, b);
        if (ROBUST) {
            xmin += plx; xmax -= plx;
            ymin += ply; ymax -= ply;
            zmin += plz; zmax -= plz;
        }
        const float tminx = xmin * adx + aox, tmaxx = xmax * adx + aox;
        const float tminy = ymin * ady + aoy, tmaxy = ymax * ady + aoy;
        const float tminz = zmin * adz + aoz, tmaxz = zmax * adz + aoz;
        float tmin = fmaxf(tminx, fmaxf(tminy, tminz));
        float tmax = fminf(tmaxx, fminf(tmaxy, tmaxz));
        if (ROBUST) {
            const float k = 9.5367431640625e-07f;  // 2^-20
            tmin = tmin - fabsf(tmin) * k;
            tmax = tmax + fabsf(tmax) * k;
        }
        tmin = tminz == tminz ? fmaxf(tmin, NODE_EPSILON) : NODE_EPSILON;
        tmax = tmaxz == tmaxz ? fminf(tmax, rt) : rt;
        if (tmin <= tmax) {
            const uint32_t cb = (uint32_t)(child_bits8 >> (c * 8)) & 0xffu;
            const uint32_t bi = (uint32_t)(bit_index8 >> (c * 8)) & 0xffu;
            hit_mask |= cb << bi;
        }
    }
    return hit_mask;
}

// RtTriangle::intersect (rt_triangle.rs:200-239) on v0, e1 = v0 - v1, e2 = v2 - v0, ng = e1 x e2; glam
// dot = (x*x'+y*y')+z*z', cross unfused.  Triangle::intersect (triangle.rs:35-76) is the same arithmetic after
// computing e1, e2 and ng from the vertices (tri_intersect below), so both give the same bits.
__device__ __forceinline__ float tri_intersect_rt(const float4& v0, float e1x, float e1y, float e1z, float e2x,
                                                  float e2y, float e2z, float nx, float ny, float nz,
                                                  const RayData& r, float tmin, float tmax) {
    const float cx = v0.x - r.ox, cy = v0.y - r.oy, cz = v0.z - r.oz;
    const float rx = r.dy * cz - r.dz * cy, ry = r.dz * cx - r.dx * cz, rz = r.dx * cy - r.dy * cx;
    const float inv_det = 1.0f / ((nx * r.dx + ny * r.dy) + nz * r.dz);
    const float u = ((rx * e2x + ry * e2y) + rz * e2z) * inv_det;
    const float v = ((rx * e1x + ry * e1y) + rz * e1z) * inv_det;
    const float w = 1.0f - u - v;
    const uint32_t hit = __float_as_uint(u) | __float_as_uint(v) | __float_as_uint(w);
    if (inv_det != 0.0f && (hit & 0x80000000u) == 0u) {
        const float t = ((nx * cx + ny * cy) + nz * cz) * inv_det;
        if (t >= tmin && t <= tmax) return t;
    }
    return INFINITY;
}

// The RtTriangle fields of a Triangle (RtTriangle::new, rt_triangle.rs:170-180).
__device__ __forceinline__ void rt_fields(const float4& v0, const float4& v1, const float4& v2, float& e1x,
                                          float& e1y, float& e1z, float& e2x, float& e2y, float& e2z, float& nx,
                                          float& ny, float& nz) {
    e1x = v0.x - v1.x, e1y = v0.y - v1.y, e1z = v0.z - v1.z;
    e2x = v2.x - v0.x, e2y = v2.y - v0.y, e2z = v2.z - v0.z;
    nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
}

// Triangle::intersect (triangle.rs:35-76).
__device__ __forceinline__ float tri_intersect(const float4& v0, const float4& v1, const float4& v2,
                                               const RayData& r, float tmin, float tmax) {
    float e1x, e1y, e1z, e2x, e2y, e2z, nx, ny, nz;
    rt_fields(v0, v1, v2, e1x, e1y, e1z, e2x, e2y, e2z, nx, ny, nz);
    return tri_intersect_rt(v0, e1x, e1y, e1z, e2x, e2y, e2z, nx, ny, nz, r, tmin, tmax);
}

// ---- watertight mode (no counterpart in the reference: an extra primitive test beside Triangle::intersect) ----
// Woop, Benthin, Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1) 2013: the ray's dominant axis kz, the other
// two axes kx, ky (swapped when d[kz] < 0, which keeps the winding) and the shear constants are computed once per
// ray; per triangle the vertices are translated to the ray origin, sheared onto the plane z = 0 and tested with 2D
// edge functions U, V, W, recomputed in f64 when one is exactly 0 (f32 products are exact in f64), so a ray through
// an edge or a vertex shared by several triangles hits at least one of them.  No culling; t = T / det in
// [tmin, tmax] inclusive, as Triangle::intersect's range test.
struct WtRay {
    uint32_t kx, ky, kz;
    float sx, sy, sz;
};

__device__ __forceinline__ float sel3(float x, float y, float z, uint32_t k) { return k == 0u ? x : (k == 1u ? y : z); }

__device__ __forceinline__ WtRay wt_ray(const RayData& r) {
    const float ax = fabsf(r.dx), ay = fabsf(r.dy), az = fabsf(r.dz);
    WtRay w;
    w.kz = (ax >= ay && ax >= az) ? 0u : (ay >= az ? 1u : 2u);
    w.kx = w.kz == 2u ? 0u : w.kz + 1u;
    w.ky = w.kx == 2u ? 0u : w.kx + 1u;
    const float dz = sel3(r.dx, r.dy, r.dz, w.kz);
    if (dz < 0.0f) {
        const uint32_t t = w.kx;
        w.kx = w.ky;
        w.ky = t;
    }
    w.sx = sel3(r.dx, r.dy, r.dz, w.kx) / dz;
    w.sy = sel3(r.dx, r.dy, r.dz, w.ky) / dz;
    w.sz = 1.0f / dz;
    return w;
}

__device__ __forceinline__ float tri_intersect_wt(const float4& v0, const float4& v1, const float4& v2,
                                                  const RayData& r, const WtRay& w, float tmin, float tmax) {
    const float a0 = v0.x - r.ox, a1 = v0.y - r.oy, a2 = v0.z - r.oz;
    const float b0 = v1.x - r.ox, b1 = v1.y - r.oy, b2 = v1.z - r.oz;
    const float c0 = v2.x - r.ox, c1 = v2.y - r.oy, c2 = v2.z - r.oz;
    const float akz = sel3(a0, a1, a2, w.kz), bkz = sel3(b0, b1, b2, w.kz), ckz = sel3(c0, c1, c2, w.kz);
    const float ax = sel3(a0, a1, a2, w.kx) - w.sx * akz, ay = sel3(a0, a1, a2, w.ky) - w.sy * akz;
    const float bx = sel3(b0, b1, b2, w.kx) - w.sx * bkz, by = sel3(b0, b1, b2, w.ky) - w.sy * bkz;
    const float cx = sel3(c0, c1, c2, w.kx) - w.sx * ckz, cy = sel3(c0, c1, c2, w.ky) - w.sy * ckz;
    float u = cx * by - cy * bx, v = ax * cy - ay * cx, ww = bx * ay - by * ax;
    if (u == 0.0f || v == 0.0f || ww == 0.0f) {
        u = (float)((double)cx * (double)by - (double)cy * (double)bx);
        v = (float)((double)ax * (double)cy - (double)ay * (double)cx);
        ww = (float)((double)bx * (double)ay - (double)by * (double)ax);
    }
    if ((u < 0.0f || v < 0.0f || ww < 0.0f) && (u > 0.0f || v > 0.0f || ww > 0.0f)) return INFINITY;
    const float det = (u + v) + ww;
    if (det == 0.0f) return INFINITY;
    const float az = w.sz * akz, bz = w.sz * bkz, cz = w.sz * ckz;
    float tt = (u * az + v * bz) + ww * cz;
    float adet = det;
    if (det < 0.0f) {
        tt = -tt;
        adet = -det;
    }
    const float t = tt / adet;
    return (t >= tmin && t <= tmax) ? t : INFINITY;
}

// MODE of the traversal kernels: bit 0 = occlusion (ray_traverse_miss) instead of closest hit, bit 1 = the
// watertight primitive test instead of Triangle::intersect, bit 2 = the primitives are RtTriangles (four 16-byte
// rows v0, e1, e2, ng, rt_triangle.rs:147-160) tested with RtTriangle::intersect (no watertight form).
template <int MODE>
__device__ __forceinline__ float prim_test(const float4* __restrict__ tris, uint32_t pid, const RayData& r,
                                           const WtRay& w, float tmin, float tmax) {
    if (MODE & 4) {
        const float4 v0 = tris[4 * (size_t)pid], e1 = tris[4 * (size_t)pid + 1], e2 = tris[4 * (size_t)pid + 2],
                     ng = tris[4 * (size_t)pid + 3];
        return tri_intersect_rt(v0, e1.x, e1.y, e1.z, e2.x, e2.y, e2.z, ng.x, ng.y, ng.z, r, tmin, tmax);
    }
    const float4 v0 = tris[3 * (size_t)pid], v1 = tris[3 * (size_t)pid + 1], v2 = tris[3 * (size_t)pid + 2];
    if (MODE & 2) return tri_intersect_wt(v0, v1, v2, r, w, tmin, tmax);
    return tri_intersect(v0, v1, v2, r, tmin, tmax);
}

// Where a launch's rays come from: one BVH (SRC_PLAIN), a mesh of a batched build per ray (SRC_MESHES), or request
// segments that each name their own BVH, ray array and output array (SRC_SEGMENTS, the serving path).
constexpr int SRC_PLAIN = 0, SRC_MESHES = 1, SRC_SEGMENTS = 2;

template <int MODE, bool COUNT, int SRC = SRC_PLAIN, int LDS = LDS_STACK>
__global__ void __launch_bounds__(TRAV_THREADS)
    k_traverse(const uint4* __restrict__ nodes_in, uint32_t node_count_in, uint32_t prim_count_in,
               const float4* __restrict__ tris_in, const float4* __restrict__ rays_in, uint32_t n,
               obvhs_ray_hit* __restrict__ hits_in, uint8_t* __restrict__ miss_in, uint32_t* __restrict__ node_visits,
               uint32_t* __restrict__ tri_tests, const MeshDesc* __restrict__ meshes,
               const uint32_t* __restrict__ ray_mesh, const TraceSegment* __restrict__ segs,
               uint2* __restrict__ spill) {
    __shared__ uint2 stack[LDS][TRAV_THREADS];
    const uint32_t lane = threadIdx.x;
    const uint32_t lanes = gridDim.x * TRAV_THREADS;
    const uint32_t gid = blockIdx.x * TRAV_THREADS + lane;
    uint2* my_spill = spill + gid;  // entry k >= LDS at my_spill[(k - LDS) * lanes]
    // Consecutive blocks take consecutive 64-ray spans.  Measured alternatives on the 1.29M-triangle heightfield
    // workload (with a grid of the resident capacity striding over the rays): giving each XCD a contiguous eighth of
    // the rays (so coherent rays share one L2) was about 25% slower; waves fetching further 64-ray spans from a
    // counter (one device-scope atomic per span) made the primary launch 1.8x slower; a copy of the nodes at a
    // 128-byte stride (one cache line per node) was 7% slower.
    for (uint32_t i = gid; i < n; i += lanes) {
    const uint4* nodes = nodes_in;
    const float4* tris = tris_in;
    const float4* rays = rays_in;
    obvhs_ray_hit* hits = hits_in;
    uint8_t* miss_out = miss_in;
    uint32_t node_count = node_count_in, prim_count = prim_count_in;
    size_t ri = i;  // index into rays and outputs
    if (SRC == SRC_MESHES) {  // batched: the ray's mesh selects its CWBVH and triangle slices; node_count = mesh count
        const uint32_t mi = ray_mesh[i];
        if (mi < node_count) {
            const MeshDesc md = meshes[mi];
            nodes += (size_t)md.cw_base * 5;
            tris += (size_t)md.prim_base * 3;
            node_count = md.cw_count;
            prim_count = md.prim_count;
        } else {  // out-of-range mesh id: the ray sees an empty BVH (a miss)
            node_count = 0u;
            prim_count = 0u;
        }
    }
    if (SRC == SRC_SEGMENTS) {  // node_count = segment count; segments are sorted by `begin`, the first begins at 0
        uint32_t lo = 0, hi = node_count;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (segs[mid].begin <= i) lo = mid;
            else hi = mid;
        }
        const TraceSegment sg = segs[lo];
        nodes = reinterpret_cast<const uint4*>(sg.nodes);
        tris = reinterpret_cast<const float4*>(sg.tris);
        rays = reinterpret_cast<const float4*>(sg.rays);
        hits = reinterpret_cast<obvhs_ray_hit*>(sg.out);
        miss_out = reinterpret_cast<uint8_t*>(sg.out);
        node_count = sg.node_count;
        prim_count = sg.prim_count;
        ri = i - sg.begin;
    }
    const float4 ra = rays[4 * ri], rb = rays[4 * ri + 1], rc = rays[4 * ri + 2], rd = rays[4 * ri + 3];
    const RayData r{ra.x, ra.y, ra.z, rb.x, rb.y, rb.z, rc.x, rc.y, rc.z, rd.x, rd.y};
    const uint32_t oct = octant_inv4(r.dx, r.dy, r.dz);
    const WtRay wr = (MODE & 2) ? wt_ray(r) : WtRay{};
    float tmax = r.tmax;
    uint32_t hit_id = INVALID32;
    float hit_t = INFINITY;
    bool is_miss = true;
    uint32_t nv = 0, tv = 0;
    int sp = 0;
    uint2 cur = node_count == 0u ? make_uint2(0u, 0u) : make_uint2(0u, 0x80000000u);
    uint2 prim = make_uint2(0u, 0u);
    for (;;) {
        // The next node (the node group's highest remaining child) does not depend on the triangle tests below, so
        // its five loads are issued first and their latency overlaps the triangles' (heightfield: 4.00 -> 4.08
        // Grays/s; 72 VGPRs, 7 waves per SIMD).
        const bool node_next = (cur.y & 0xff000000u) != 0u && __popc(prim.y) <= TRI_PER_STEP;
        uint32_t child = 0u;
        uint4 q0, q1, q2, q3, q4;
        if (node_next) {
            const uint32_t slot_index = (firstbithigh(cur.y) - 24u) ^ (oct & 0xffu);
            child = cur.x + (uint32_t)__popc(cur.y & ~(0xffffffffu << slot_index));
            if (child < node_count) {
                const uint4* np = nodes + (size_t)child * 5;
                q0 = np[0];
                q1 = np[1];
                q2 = np[2];
                q3 = np[3];
                q4 = np[4];
            }
        }
        bool stop = false;
        for (int k = 0; k < TRI_PER_STEP && prim.y != 0u; ++k) {
            const uint32_t local = firstbithigh(prim.y);
            prim.y &= ~(1u << local);
            const uint32_t pid = prim.x + local;
            if (pid >= prim_count) continue;  // malformed input only: never index past the triangle buffer
            if (COUNT) tv++;
            if ((MODE & 1) == 0) {
                const float t = prim_test<MODE>(tris, pid, r, wr, r.tmin, tmax);
                if (t < tmax) {
                    hit_id = pid;
                    hit_t = t;
                    tmax = t;
                }
            } else {
                const float t = prim_test<MODE>(tris, pid, r, wr, r.tmin, r.tmax);
                if (t < r.tmax) {
                    is_miss = false;
                    stop = true;
                    break;
                }
            }
        }
        if (stop) break;
        if (node_next) {  // the group's triangles are all tested (at most TRI_PER_STEP were pending)
            cur.y &= ~(1u << firstbithigh(cur.y));
            if ((cur.y & 0xff000000u) != 0u) {
                if (sp < LDS) stack[sp][lane] = cur;
                else my_spill[(size_t)(sp - LDS) * lanes] = cur;
                sp = min(sp + 1, STACK_SIZE - 1);
            }
            if (child >= node_count) break;  // malformed input only: never index past the node buffer
            if (COUNT) nv++;
            const uint32_t hm = intersect_node<(MODE & 2) != 0>(q0, q1, q2, q3, q4, r, tmax, oct);
            const uint32_t imask = (q0.w >> 24) & 0xffu;
            cur = make_uint2(q1.x, (hm & 0xff000000u) | imask);
            prim = make_uint2(q1.y, hm & 0x00ffffffu);
        }
        // No node group left: pop the next one now, before this node's triangles are tested in the next step (the
        // stack does not depend on them), so that step also tests a node and its five loads go out first.  The
        // crate pops once both groups are empty; the stack's pushes and pops keep their order, so the visits do.
        if ((cur.y & 0xff000000u) == 0u) {
            if (sp == 0) {
                if (prim.y == 0u) break;
            } else {
                sp -= 1;
                cur = sp < LDS ? stack[sp][lane] : my_spill[(size_t)(sp - LDS) * lanes];
            }
        }
    }
    if ((MODE & 1) == 0) {
        obvhs_ray_hit h;
        h.primitive_id = hit_id;
        h.geometry_id = INVALID32;
        h.instance_id = INVALID32;
        h.t = hit_t;
        reinterpret_cast<uint4*>(hits)[ri] = make_uint4(h.primitive_id, h.geometry_id, h.instance_id,
                                                        __float_as_uint(h.t));
    } else {
        miss_out[ri] = is_miss ? 1 : 0;
    }
    if (COUNT) {
        node_visits[i] = nv;
        tri_tests[i] = tv;
    }
    }  // rays of this lane
}

// k_traverse for one BVH when a launch holds several rays per resident lane: each 64-lane workgroup (one wave)
// owns per_wave / 64 rows of 64 rays, rows blockIdx.x, blockIdx.x + W, blockIdx.x + 2W, ... (W workgroups, so the
// waves sweep the rays together, as a one-ray-per-lane grid does), and a lane whose ray ends takes the wave's next
// ray at the next step (a wave-uniform count of the rays handed out, no atomics) instead of idling until the wave's
// longest ray ends.  The steps of every ray are k_traverse's, so the hits are identical.
// SRC_MESHES: as in k_traverse, each ray's mesh (ray_mesh) selects its CWBVH and triangle slices; node_count_in is
// then the mesh count.
template <int MODE, int LDS, int SRC = SRC_PLAIN>
__global__ void __launch_bounds__(TRAV_THREADS)
    k_traverse_chunk(const uint4* __restrict__ nodes_in, uint32_t node_count_in, uint32_t prim_count_in,
                     const float4* __restrict__ tris_in, const float4* __restrict__ rays, uint32_t n,
                     obvhs_ray_hit* __restrict__ hits, uint8_t* __restrict__ miss_out, uint2* __restrict__ spill,
                     uint32_t per_wave, const MeshDesc* __restrict__ meshes, const uint32_t* __restrict__ ray_mesh) {
    __shared__ uint2 stack[LDS][TRAV_THREADS];
    const uint32_t lane = threadIdx.x;
    const uint32_t lanes = gridDim.x * TRAV_THREADS;
    uint2* my_spill = spill + blockIdx.x * TRAV_THREADS + lane;  // entry k >= LDS at my_spill[(k - LDS) * lanes]
    uint32_t handed = 0u;  // rays of the chunk handed out so far (the same in every lane)
    bool live = false;
    uint32_t i = 0u;
    RayData r{};
    WtRay wr{};
    uint32_t oct = 0u, hit_id = INVALID32;
    float tmax = 0.0f, hit_t = INFINITY;
    bool is_miss = true;
    int sp = 0;
    uint2 cur = make_uint2(0u, 0u), prim = make_uint2(0u, 0u);
    const uint4* nodes = nodes_in;
    const float4* tris = tris_in;
    uint32_t node_count = node_count_in, prim_count = prim_count_in;
    for (;;) {
        const uint64_t want = __ballot(!live);
        if (want != 0ull) {
            if (!live) {
                // the h-th ray handed out in this wave: row h / 64 of the wave's rows blockIdx.x, blockIdx.x + W, ...
                const uint32_t h = handed + (uint32_t)__popcll(want & lanemask_lt());
                const uint64_t k = ((uint64_t)(h / TRAV_THREADS) * gridDim.x + blockIdx.x) * TRAV_THREADS +
                                   h % TRAV_THREADS;
                if (h < per_wave && k < n) {
                    i = (uint32_t)k;
                    if (SRC == SRC_MESHES) {  // the ray's mesh; an out-of-range id sees an empty BVH (a miss)
                        const uint32_t mi = ray_mesh[i];
                        nodes = nodes_in;
                        tris = tris_in;
                        node_count = 0u;
                        prim_count = 0u;
                        if (mi < node_count_in) {
                            const MeshDesc md = meshes[mi];
                            nodes += (size_t)md.cw_base * 5;
                            tris += (size_t)md.prim_base * 3;
                            node_count = md.cw_count;
                            prim_count = md.prim_count;
                        }
                    }
                    const float4 ra = rays[4 * (size_t)i], rb = rays[4 * (size_t)i + 1], rc = rays[4 * (size_t)i + 2],
                                 rd = rays[4 * (size_t)i + 3];
                    r = RayData{ra.x, ra.y, ra.z, rb.x, rb.y, rb.z, rc.x, rc.y, rc.z, rd.x, rd.y};
                    oct = octant_inv4(r.dx, r.dy, r.dz);
                    if (MODE & 2) wr = wt_ray(r);
                    tmax = r.tmax;
                    hit_id = INVALID32;
                    hit_t = INFINITY;
                    is_miss = true;
                    sp = 0;
                    cur = node_count == 0u ? make_uint2(0u, 0u) : make_uint2(0u, 0x80000000u);
                    prim = make_uint2(0u, 0u);
                    live = true;
                }
            }
            handed = min(handed + (uint32_t)__popcll(want), per_wave);
        }
        if (__ballot(live) == 0ull) break;  // the chunk is done (wave-uniform)
        if (!live) continue;
        // one step of k_traverse's loop
        bool done = false;
        const bool node_next = (cur.y & 0xff000000u) != 0u && __popc(prim.y) <= TRI_PER_STEP;
        uint32_t child = 0u;
        uint4 q0, q1, q2, q3, q4;
        if (node_next) {
            const uint32_t slot_index = (firstbithigh(cur.y) - 24u) ^ (oct & 0xffu);
            child = cur.x + (uint32_t)__popc(cur.y & ~(0xffffffffu << slot_index));
            if (child < node_count) {
                const uint4* np = nodes + (size_t)child * 5;
                q0 = np[0];
                q1 = np[1];
                q2 = np[2];
                q3 = np[3];
                q4 = np[4];
            }
        }
        for (int k = 0; k < TRI_PER_STEP && prim.y != 0u; ++k) {
            const uint32_t local = firstbithigh(prim.y);
            prim.y &= ~(1u << local);
            const uint32_t pid = prim.x + local;
            if (pid >= prim_count) continue;  // malformed input only: never index past the triangle buffer
            if ((MODE & 1) == 0) {
                const float t = prim_test<MODE>(tris, pid, r, wr, r.tmin, tmax);
                if (t < tmax) {
                    hit_id = pid;
                    hit_t = t;
                    tmax = t;
                }
            } else {
                const float t = prim_test<MODE>(tris, pid, r, wr, r.tmin, r.tmax);
                if (t < r.tmax) {
                    is_miss = false;
                    done = true;
                    break;
                }
            }
        }
        if (!done) {
            if (node_next) {  // the group's triangles are all tested (at most TRI_PER_STEP were pending)
                cur.y &= ~(1u << firstbithigh(cur.y));
                if ((cur.y & 0xff000000u) != 0u) {
                    if (sp < LDS) stack[sp][lane] = cur;
                    else my_spill[(size_t)(sp - LDS) * lanes] = cur;
                    sp = min(sp + 1, STACK_SIZE - 1);
                }
                if (child >= node_count) {
                    done = true;  // malformed input only: never index past the node buffer
                } else {
                    const uint32_t hm = intersect_node<(MODE & 2) != 0>(q0, q1, q2, q3, q4, r, tmax, oct);
                    const uint32_t imask = (q0.w >> 24) & 0xffu;
                    cur = make_uint2(q1.x, (hm & 0xff000000u) | imask);
                    prim = make_uint2(q1.y, hm & 0x00ffffffu);
                }
            }
            if (!done && (cur.y & 0xff000000u) == 0u) {  // the early pop of k_traverse
                if (sp == 0) {
                    done = prim.y == 0u;
                } else {
                    sp -= 1;
                    cur = sp < LDS ? stack[sp][lane] : my_spill[(size_t)(sp - LDS) * lanes];
                }
            }
        }
        if (done) {
            if ((MODE & 1) == 0)
                reinterpret_cast<uint4*>(hits)[i] = make_uint4(hit_id, INVALID32, INVALID32, __float_as_uint(hit_t));
            else
                miss_out[i] = is_miss ? 1 : 0;
            live = false;
        }
    }
}

// CwBvh::ray_traverse_tlas_blas (cwbvh/mod.rs:341-491): one stack for both levels, TLAS primitives name BLASes
// (meshes of a batched build) through the TLAS's primitive_indices, the closure is Triangle::intersect on the
// mesh's triangles in its primitive order.  Restated as written, including that a TLAS primitive group popped after
// a BLAS finishes is treated as a node group and dropped (mod.rs:396-401, 469).  Hits report the BLAS-local
// primitive slot and the BLAS index (RayHit::geometry_id).
__global__ void __launch_bounds__(TRAV_THREADS)
    k_traverse_tlas(const uint4* __restrict__ tlas, uint32_t tlas_count, const uint32_t* __restrict__ tlas_prims,
                    uint32_t tlas_prim_count, const uint4* __restrict__ blas_nodes, const MeshDesc* __restrict__ meshes,
                    uint32_t mesh_count, const float4* __restrict__ tris, const float4* __restrict__ rays, uint32_t n,
                    obvhs_ray_hit* __restrict__ hits, uint2* __restrict__ spill) {
    __shared__ uint2 stack[LDS_STACK][TRAV_THREADS];
    const uint32_t lane = threadIdx.x;
    const uint32_t lanes = gridDim.x * TRAV_THREADS;
    const uint32_t gid = blockIdx.x * TRAV_THREADS + lane;
    uint2* my_spill = spill + gid;
    for (uint32_t i = gid; i < n; i += lanes) {
        const float4 ra = rays[4 * (size_t)i], rb = rays[4 * (size_t)i + 1], rc = rays[4 * (size_t)i + 2],
                     rd = rays[4 * (size_t)i + 3];
        RayData r{ra.x, ra.y, ra.z, rb.x, rb.y, rb.z, rc.x, rc.y, rc.z, rd.x, rd.y};
        const uint32_t oct = octant_inv4(r.dx, r.dy, r.dz);
        uint32_t hit_id = INVALID32, hit_geom = INVALID32;
        int sp = 0;  // StackStack::index
        uint32_t tlas_stack_size = INVALID32, mesh = INVALID32;
        const uint4* nodes = tlas;
        uint32_t node_count = tlas_count;
        const float4* mtris = tris;
        uint32_t mprims = 0;
        uint2 cur = tlas_count == 0u ? make_uint2(0u, 0u) : make_uint2(0u, 0x80000000u);
        for (;;) {
            uint2 prim = make_uint2(0u, 0u);
            if ((cur.y & 0xff000000u) != 0u) {
                const uint32_t hits_imask = cur.y;
                const uint32_t off = firstbithigh(hits_imask);
                const uint32_t base = cur.x;
                cur.y &= ~(1u << off);
                if ((cur.y & 0xff000000u) != 0u) {
                    if (sp < LDS_STACK) stack[sp][lane] = cur;
                    else my_spill[(size_t)(sp - LDS_STACK) * lanes] = cur;
                    sp = min(sp + 1, STACK_SIZE - 1);
                }
                const uint32_t slot_index = (off - 24u) ^ (oct & 0xffu);
                const uint32_t rel = (uint32_t)__popc(hits_imask & ~(0xffffffffu << slot_index));
                const uint32_t child = base + rel;
                if (child >= node_count) break;  // malformed input only
                const uint4* np = nodes + (size_t)child * 5;
                const uint4 q0 = np[0], q1 = np[1], q2 = np[2], q3 = np[3], q4 = np[4];
                const uint32_t hm = intersect_node(q0, q1, q2, q3, q4, r, r.tmax, oct);
                cur = make_uint2(q1.x, (hm & 0xff000000u) | ((q0.w >> 24) & 0xffu));
                prim = make_uint2(q1.y, hm & 0x00ffffffu);
            } else {
                cur = make_uint2(0u, 0u);
            }
            while (prim.y != 0u) {
                const uint32_t local = firstbithigh(prim.y);
                prim.y &= ~(1u << local);
                const uint32_t global = prim.x + local;
                if (tlas_stack_size == INVALID32) {  // in the TLAS: the primitive is a BLAS
                    if (prim.y != 0u) {
                        if (sp < LDS_STACK) stack[sp][lane] = prim;
                        else my_spill[(size_t)(sp - LDS_STACK) * lanes] = prim;
                        sp = min(sp + 1, STACK_SIZE - 1);
                    }
                    if ((cur.y & 0xff000000u) != 0u) {
                        if (sp < LDS_STACK) stack[sp][lane] = cur;
                        else my_spill[(size_t)(sp - LDS_STACK) * lanes] = cur;
                        sp = min(sp + 1, STACK_SIZE - 1);
                    }
                    tlas_stack_size = (uint32_t)sp;
                    const uint32_t b = global < tlas_prim_count ? tlas_prims[global] : INVALID32;
                    if (b < mesh_count) {
                        const MeshDesc md = meshes[b];
                        nodes = blas_nodes + (size_t)md.cw_base * 5;
                        node_count = md.cw_count;
                        mtris = tris + (size_t)md.prim_base * 3;
                        mprims = md.prim_count;
                        mesh = b;
                        cur = node_count == 0u ? make_uint2(0u, 0u) : make_uint2(0u, 0x80000000u);
                    } else {  // malformed input only
                        cur = make_uint2(0u, 0u);
                    }
                    break;
                }
                if (global < mprims) {  // in a BLAS: Triangle::intersect with the ray's current tmax
                    const float4 v0 = mtris[3 * (size_t)global], v1 = mtris[3 * (size_t)global + 1],
                                 v2 = mtris[3 * (size_t)global + 2];
                    const float t = tri_intersect(v0, v1, v2, r, r.tmin, r.tmax);
                    if (t < r.tmax) {
                        hit_id = global;
                        hit_geom = mesh;
                        r.tmax = t;
                    }
                }
            }
            if ((cur.y & 0xff000000u) == 0u) {
                if (sp == 0) break;
                if ((uint32_t)sp == tlas_stack_size) {  // the BLAS is done: back to the TLAS
                    tlas_stack_size = INVALID32;
                    mesh = INVALID32;
                    nodes = tlas;
                    node_count = tlas_count;
                }
                sp -= 1;
                cur = sp < LDS_STACK ? stack[sp][lane] : my_spill[(size_t)(sp - LDS_STACK) * lanes];
            }
        }
        reinterpret_cast<uint4*>(hits)[i] = make_uint4(hit_id, hit_geom, INVALID32,
                                                       __float_as_uint(hit_id != INVALID32 ? r.tmax : INFINITY));
    }
}

// RtTriangle::from(&Triangle) (rt_triangle.rs:162-181) for every triangle: v0, e1 = v0 - v1, e2 = v2 - v0,
// ng = e1 x e2, each a Vec3A row (w = 0).
__global__ void __launch_bounds__(256) k_rt_triangles(const float4* __restrict__ tris, uint32_t n,
                                                      float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 v0 = tris[3 * (size_t)i], v1 = tris[3 * (size_t)i + 1], v2 = tris[3 * (size_t)i + 2];
    float e1x, e1y, e1z, e2x, e2y, e2z, nx, ny, nz;
    rt_fields(v0, v1, v2, e1x, e1y, e1z, e2x, e2y, e2z, nx, ny, nz);
    out[4 * (size_t)i] = make_float4(v0.x, v0.y, v0.z, 0.0f);
    out[4 * (size_t)i + 1] = make_float4(e1x, e1y, e1z, 0.0f);
    out[4 * (size_t)i + 2] = make_float4(e2x, e2y, e2z, 0.0f);
    out[4 * (size_t)i + 3] = make_float4(nx, ny, nz, 0.0f);
}

}  // namespace

// safe_inverse (ray.rs:6-12); signum(+-0) = +-1, signum(NaN) = NaN (and |NaN| <= EPSILON is false).
__device__ __forceinline__ float safe_inverse(float x) {
    constexpr float EPS = 1.1920928955078125e-07f;  // f32::EPSILON = 2^-23
    return fabsf(x) <= EPS ? copysignf(1.0f, x) / EPS : 1.0f / x;
}

// Ray::new (ray.rs:34-52), one thread per ray.
__global__ void __launch_bounds__(256) k_rays_new(const float4* __restrict__ o, const float4* __restrict__ d,
                                                  uint32_t n, float tmin, float tmax, float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 a = o[i], b = d[i];
    out[4 * (size_t)i] = make_float4(a.x, a.y, a.z, 0.0f);
    out[4 * (size_t)i + 1] = make_float4(b.x, b.y, b.z, 0.0f);
    out[4 * (size_t)i + 2] = make_float4(safe_inverse(b.x), safe_inverse(b.y), safe_inverse(b.z), 0.0f);
    out[4 * (size_t)i + 3] = make_float4(tmin, tmax, 0.0f, 0.0f);
}

// RayHit::none (ray.rs:76-83).
__global__ void __launch_bounds__(256) k_hits_none(uint4* __restrict__ hits, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) hits[i] = make_uint4(INVALID32, INVALID32, INVALID32, __float_as_uint(INFINITY));
}

hipError_t launch_rays_new(const obvhs_vec3a* origins, const obvhs_vec3a* directions, uint32_t n, float tmin,
                           float tmax, obvhs_ray* out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    k_rays_new<<<(n + 255) / 256, 256, 0, stream>>>(reinterpret_cast<const float4*>(origins),
                                                    reinterpret_cast<const float4*>(directions), n, tmin, tmax,
                                                    reinterpret_cast<float4*>(out));
    return hipGetLastError();
}

hipError_t launch_hits_none(obvhs_ray_hit* hits, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    k_hits_none<<<(n + 255) / 256, 256, 0, stream>>>(reinterpret_cast<uint4*>(hits), n);
    return hipGetLastError();
}

hipError_t launch_rt_triangles(const obvhs_triangle* tris, uint32_t n, obvhs_rt_triangle* out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    k_rt_triangles<<<(n + 255) / 256, 256, 0, stream>>>(reinterpret_cast<const float4*>(tris), n,
                                                        reinterpret_cast<float4*>(out));
    return hipGetLastError();
}

// Launches k_traverse<MODE, COUNT, SRC, depth> with the LDS stack depth the launch was tuned for (4, 8 or 16).
template <int MODE, bool COUNT, int SRC, class... A>
static void launch_depth(uint32_t depth, uint32_t blocks, hipStream_t stream, A... args) {
    if (depth == 4u)
        k_traverse<MODE, COUNT, SRC, 4><<<blocks, TRAV_THREADS, 0, stream>>>(args...);
    else if (depth == 16u)
        k_traverse<MODE, COUNT, SRC, 16><<<blocks, TRAV_THREADS, 0, stream>>>(args...);
    else
        k_traverse<MODE, COUNT, SRC, 8><<<blocks, TRAV_THREADS, 0, stream>>>(args...);
}

// Chunked launches: a grid of the waves the device keeps resident for the kernel instance (its occupancy from its
// register and LDS use, times the compute units) that the rows of 64 rays are dealt over, rows w, w + W, ... to wave
// w, so no wave waits for a slot while others run (with a grid of 32 waves per CU and 7 resident, the last 4 per CU
// started as the first finished and ran alone at the end).  At most CHUNK_MAX_ROWS_PER_WAVE rows per wave; larger
// launches have more waves than slots.
constexpr uint32_t CHUNK_MAX_ROWS_PER_WAVE = 16;
struct ChunkPlan {
    uint32_t blocks = 0, per_wave = 0;  // per_wave = rays (64 per row) a wave's lanes take in turn
};

template <class K>
static uint32_t resident_waves_per_cu(K kernel) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, TRAV_THREADS, 0) != hipSuccess || nb <= 0) {
        (void)hipGetLastError();
        return 0u;
    }
    return (uint32_t)nb;
}

// The plan for k_traverse_chunk<MODE, LDS, SRC> over n rays; blocks = 0 when the launch should not be chunked (fewer
// than two rows per resident wave, or more waves than spill columns).
template <int MODE, int LDS, int SRC>
static ChunkPlan chunk_plan(uint32_t n, const TravScratch& ts) {
    static const uint32_t per_cu = resident_waves_per_cu(k_traverse_chunk<MODE, LDS, SRC>);
    ChunkPlan p;
    if (!ts.chunked || per_cu == 0u || ts.resident == 0u) return p;
    const uint64_t cus = std::max<uint32_t>(1u, ts.resident / (32u * TRAV_THREADS));
    const uint64_t waves = cus * per_cu;
    const uint64_t rows = ((uint64_t)n + TRAV_THREADS - 1) / TRAV_THREADS;
    if (rows < 2u * waves) return p;
    // every resident slot takes a wave (rows dealt one at a time: the waves' row counts differ by at most one)
    const uint64_t per = std::min<uint64_t>((rows + waves - 1) / waves, CHUNK_MAX_ROWS_PER_WAVE);
    const uint64_t blocks = std::max<uint64_t>(waves, (rows + per - 1) / per);
    if (blocks * TRAV_THREADS > ts.lanes) return p;
    p.blocks = (uint32_t)blocks;
    p.per_wave = (uint32_t)(per * TRAV_THREADS);
    return p;
}

// k_traverse_chunk<MODE, depth, SRC> with the launch's LDS stack depth, when chunk_plan chunks the launch; false
// (nothing launched) otherwise.
template <int MODE, int SRC, int LDS>
static bool launch_chunk_depth(uint32_t n, const TravScratch& ts, hipStream_t stream, const uint4* nd,
                               uint32_t node_count, uint32_t prim_count, const float4* tr, const float4* ry,
                               obvhs_ray_hit* hits, uint8_t* miss, uint2* sp, const MeshDesc* meshes,
                               const uint32_t* ray_mesh) {
    const ChunkPlan p = chunk_plan<MODE, LDS, SRC>(n, ts);
    if (!p.blocks) return false;
    k_traverse_chunk<MODE, LDS, SRC><<<p.blocks, TRAV_THREADS, 0, stream>>>(
        nd, node_count, prim_count, tr, ry, n, hits, miss, sp, p.per_wave, meshes, ray_mesh);
    return true;
}

template <int MODE, int SRC>
static bool launch_chunk(uint32_t n, const TravScratch& ts, hipStream_t stream, const uint4* nd, uint32_t node_count,
                         uint32_t prim_count, const float4* tr, const float4* ry, obvhs_ray_hit* hits, uint8_t* miss,
                         uint2* sp, const MeshDesc* meshes, const uint32_t* ray_mesh) {
    if (ts.lds_depth == 4u)
        return launch_chunk_depth<MODE, SRC, 4>(n, ts, stream, nd, node_count, prim_count, tr, ry, hits, miss, sp,
                                                meshes, ray_mesh);
    if (ts.lds_depth == 16u)
        return launch_chunk_depth<MODE, SRC, 16>(n, ts, stream, nd, node_count, prim_count, tr, ry, hits, miss, sp,
                                                 meshes, ray_mesh);
    return launch_chunk_depth<MODE, SRC, 8>(n, ts, stream, nd, node_count, prim_count, tr, ry, hits, miss, sp, meshes,
                                            ray_mesh);
}

// Grid: one lane per ray, at most the spill buffer's lane count (lanes then stride over the rays).
static uint32_t trav_blocks(uint32_t n, const TravScratch& ts) {
    const uint32_t want = (n + TRAV_THREADS - 1) / TRAV_THREADS;
    const uint32_t b = std::min<uint32_t>(want, ts.lanes / TRAV_THREADS);
    return b == 0u ? 1u : b;
}

hipError_t launch_traverse(const CwNode* nodes, uint32_t node_count, uint32_t prim_count, const obvhs_triangle* tris,
                           const obvhs_ray* rays, uint32_t n, int mode, obvhs_ray_hit* hits, uint8_t* miss,
                           uint32_t* node_visits, uint32_t* tri_tests, const TravScratch& ts, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t blocks = trav_blocks(n, ts);
    const uint4* nd = reinterpret_cast<const uint4*>(nodes);
    const float4* tr = reinterpret_cast<const float4*>(tris);
    const float4* ry = reinterpret_cast<const float4*>(rays);
    uint2* sp = static_cast<uint2*>(ts.spill);
    const MeshDesc* no_meshes = nullptr;
    const uint32_t* no_map = nullptr;
    const TraceSegment* no_segs = nullptr;
    // Chunked launches (when the tuning table asks for them: incoherent rays) need at least two rows per resident
    // wave; each lane then takes that many rays in turn, refilled in the wave.
    if (mode > 5 || (node_visits && mode != 0)) return hipErrorInvalidValue;
    if (!node_visits) {
        bool chunked = false;
#define OBVHS_CHUNK(M) \
    chunked = launch_chunk<M, SRC_PLAIN>(n, ts, stream, nd, node_count, prim_count, tr, ry, hits, miss, sp, no_meshes, no_map)
        if (mode == 0) OBVHS_CHUNK(0);
        else if (mode == 1) OBVHS_CHUNK(1);
        else if (mode == 2) OBVHS_CHUNK(2);
        else if (mode == 3) OBVHS_CHUNK(3);
        else if (mode == 4) OBVHS_CHUNK(4);
        else OBVHS_CHUNK(5);
#undef OBVHS_CHUNK
        if (chunked) return hipGetLastError();
    }
#define OBVHS_LANES(M, C)                                                                                       \
    launch_depth<M, C, SRC_PLAIN>(ts.lds_depth, blocks, stream, nd, node_count, prim_count, tr, ry, n, hits, miss, \
                                  node_visits, tri_tests, no_meshes, no_map, no_segs, sp)
    if (node_visits) OBVHS_LANES(0, true);
    else if (mode == 0) OBVHS_LANES(0, false);
    else if (mode == 1) OBVHS_LANES(1, false);
    else if (mode == 2) OBVHS_LANES(2, false);
    else if (mode == 3) OBVHS_LANES(3, false);
    else if (mode == 4) OBVHS_LANES(4, false);
    else OBVHS_LANES(5, false);
#undef OBVHS_LANES
    return hipGetLastError();
}

hipError_t launch_traverse_meshes(const CwNode* nodes, const MeshDesc* meshes, uint32_t mesh_count,
                                  const uint32_t* ray_mesh, const obvhs_triangle* tris, const obvhs_ray* rays,
                                  uint32_t n, int mode, obvhs_ray_hit* hits, uint8_t* miss, const TravScratch& ts,
                                  hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t blocks = trav_blocks(n, ts);
    const uint4* nd = reinterpret_cast<const uint4*>(nodes);
    const float4* tr = reinterpret_cast<const float4*>(tris);
    const float4* ry = reinterpret_cast<const float4*>(rays);
    uint2* sp = static_cast<uint2*>(ts.spill);
    uint32_t* no_counts = nullptr;
    const TraceSegment* no_segs = nullptr;
    const bool chunked =
        mode == 0 ? launch_chunk<0, SRC_MESHES>(n, ts, stream, nd, mesh_count, 0u, tr, ry, hits, miss, sp, meshes,
                                                ray_mesh)
                  : launch_chunk<1, SRC_MESHES>(n, ts, stream, nd, mesh_count, 0u, tr, ry, hits, miss, sp, meshes,
                                                ray_mesh);
    if (chunked) return hipGetLastError();
    if (mode == 0)
        launch_depth<0, false, SRC_MESHES>(ts.lds_depth, blocks, stream, nd, mesh_count, 0u, tr, ry, n, hits, miss,
                                           no_counts, no_counts, meshes, ray_mesh, no_segs, sp);
    else
        launch_depth<1, false, SRC_MESHES>(ts.lds_depth, blocks, stream, nd, mesh_count, 0u, tr, ry, n, hits, miss,
                                           no_counts, no_counts, meshes, ray_mesh, no_segs, sp);
    return hipGetLastError();
}

hipError_t launch_traverse_segments(const TraceSegment* segs, uint32_t nseg, uint32_t n, int mode,
                                    const TravScratch& ts, hipStream_t stream) {
    if (n == 0 || nseg == 0) return hipSuccess;
    const uint32_t blocks = trav_blocks(n, ts);
    uint2* sp = static_cast<uint2*>(ts.spill);
    const uint4* no_nodes = nullptr;
    const float4* no_f4 = nullptr;
    obvhs_ray_hit* no_hits = nullptr;
    uint8_t* no_miss = nullptr;
    uint32_t* no_counts = nullptr;
    const MeshDesc* no_meshes = nullptr;
    const uint32_t* no_map = nullptr;
    if (mode == 0)
        launch_depth<0, false, SRC_SEGMENTS>(ts.lds_depth, blocks, stream, no_nodes, nseg, 0u, no_f4, no_f4, n,
                                             no_hits, no_miss, no_counts, no_counts, no_meshes, no_map, segs, sp);
    else
        launch_depth<1, false, SRC_SEGMENTS>(ts.lds_depth, blocks, stream, no_nodes, nseg, 0u, no_f4, no_f4, n,
                                             no_hits, no_miss, no_counts, no_counts, no_meshes, no_map, segs, sp);
    return hipGetLastError();
}

// Spill columns for a launch of n rays: one lane per ray up to TRAV_MAX_LANES lanes, at least the device's resident
// capacity (compute units x 32 waves).  One ray per lane lets the dispatcher keep the CUs full as waves finish;
// measured on the 1.29M-triangle heightfield, 2.07M primary rays took 0.300 ms against 0.364 ms on a persistent
// grid of the resident capacity walking the rays with a stride (incoherent bounce rays: no change).  Launches of
// more rays stride over TRAV_MAX_LANES lanes.  The spill costs capacity, not bandwidth: only stacks deeper than
// the LDS part touch it.
obvhs_status traverse_scratch(obvhs_context* ctx, TravScratch& ts, uint64_t n) {
    if (ctx->trav_lanes == 0) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess)
            return set_hip_error(ctx, hipGetLastError(), "device properties");
        ctx->trav_lanes = (uint32_t)std::max(1, prop.multiProcessorCount) * 32u * TRAV_THREADS;
    }
    const uint64_t want = (std::min<uint64_t>(n, TRAV_MAX_LANES) + TRAV_THREADS - 1) / TRAV_THREADS * TRAV_THREADS;
    const uint32_t lanes = (uint32_t)std::max<uint64_t>(ctx->trav_lanes, want);
    // columns for the shallowest LDS depth a launch may use (4): every depth's spill fits
    const size_t bytes = (size_t)lanes * (STACK_SIZE - 4) * sizeof(uint2);
    const hipError_t e = ctx->scratch.trav_spill.ensure(bytes);
    if (e != hipSuccess) return set_hip_error(ctx, e, "alloc traversal spill");
    ts.spill = ctx->scratch.trav_spill.ptr;
    ts.lanes = lanes;
    ts.resident = ctx->trav_lanes;
    return OBVHS_OK;
}

}  // namespace obvhs

// ---- closest-hit traversal that allocates nothing (obvhs_cwbvh_ray_traverse_noalloc) ----
// Its spill columns live in this array, which the library's code object holds on every device it is loaded on until
// the process ends: NOALLOC_LANES columns of STACK_SIZE - NOALLOC_MIN_LDS entries (8 B each) = 96 MiB.  Host code
// never passes `scratch` itself to a kernel (on the host the name is only a stand-in for the symbol): obvhs_context_create
// looks up each device's address of it once with hipGetSymbolAddress and every launch takes that address.
namespace {
constexpr uint32_t NOALLOC_LANES = 1u << 19;  // 256 CUs x 32 waves x 64 lanes: the MI355X's resident lanes
constexpr uint32_t NOALLOC_MIN_LDS = 8;       // LDS stack entries at least (the rest of the 32 spill)
constexpr uint64_t NOALLOC_MAX_RAYS = 1ull << 31;  // lanes stride over the rays; the ray index stays below 2^32
constexpr size_t NOALLOC_ENTRIES = (size_t)NOALLOC_LANES * (obvhs::STACK_SIZE - NOALLOC_MIN_LDS);
constexpr int NOALLOC_MAX_DEVICES = 64;
}  // namespace

static __device__ __align__(256) uint2 scratch[NOALLOC_ENTRIES];

namespace obvhs {

namespace {
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

// Called by obvhs_context_create with the context's device current: the device's address of `scratch`, looked up
// once per device and kept for every later context on it.
obvhs_status noalloc_scratch_init(obvhs_context* ctx) {
    if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES) return OBVHS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[ctx->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return set_hip_error(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol,
                                                         "static traversal scratch address");
        g_noalloc_ptr[ctx->device] = p;
    }
    ctx->noalloc_spill = g_noalloc_ptr[ctx->device];
    if (ctx->trav_lanes == 0) {
        hipDeviceProp_t prop;
        const hipError_t e = hipGetDeviceProperties(&prop, ctx->device);
        if (e != hipSuccess) return set_hip_error(ctx, e, "device properties");
        ctx->trav_lanes = (uint32_t)std::max(1, prop.multiProcessorCount) * 32u * TRAV_THREADS;
    }
    return OBVHS_OK;
}

}  // namespace obvhs

extern "C" obvhs_status obvhs_cwbvh_ray_traverse_noalloc(obvhs_context* ctx, const obvhs_cwbvh* bvh,
                                                         const obvhs_triangle* d_bvh_tris, const obvhs_ray* d_rays,
                                                         uint32_t n, obvhs_ray_hit* d_hits) {
    using namespace obvhs;
    if (!ctx || !bvh || (n && (!d_rays || !d_hits))) return OBVHS_ERR_INVALID_ARG;
    if (bvh->node_count && !d_bvh_tris) return OBVHS_ERR_INVALID_ARG;
    if ((uint64_t)n > NOALLOC_MAX_RAYS) {
        ctx->last_error = "obvhs_cwbvh_ray_traverse_noalloc: at most 2^31 rays per call";
        return OBVHS_ERR_CAPACITY;
    }
    if (!ctx->noalloc_spill) return OBVHS_ERR_DEVICE;
    if (n == 0) return OBVHS_OK;
    TravScratch ts;
    ts.spill = ctx->noalloc_spill;
    ts.lanes = NOALLOC_LANES;
    ts.resident = std::min<uint32_t>(ctx->trav_lanes, NOALLOC_LANES);
    ts.lds_depth = std::max<uint32_t>(lds_depth_for(ctx, bvh->prim_count), NOALLOC_MIN_LDS);
    ts.chunked = chunk_for(ctx, bvh->prim_count);
    const hipError_t e = launch_traverse(bvh->nodes.as<CwNode>(), bvh->node_count, bvh->prim_count, d_bvh_tris,
                                         d_rays, n, 0, d_hits, nullptr, nullptr, nullptr, ts, ctx->stream);
    if (e != hipSuccess) return set_hip_error(ctx, e, "traverse noalloc");
    return OBVHS_OK;
}

namespace obvhs {

hipError_t launch_traverse_tlas(const CwNode* tlas, uint32_t tlas_count, const uint32_t* tlas_prims,
                                uint32_t tlas_prim_count, const CwNode* blas_nodes, const MeshDesc* meshes,
                                uint32_t mesh_count, const obvhs_triangle* tris, const obvhs_ray* rays, uint32_t n,
                                obvhs_ray_hit* hits, const TravScratch& ts, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint32_t blocks = trav_blocks(n, ts);
    k_traverse_tlas<<<blocks, TRAV_THREADS, 0, stream>>>(
        reinterpret_cast<const uint4*>(tlas), tlas_count, tlas_prims, tlas_prim_count,
        reinterpret_cast<const uint4*>(blas_nodes), meshes, mesh_count, reinterpret_cast<const float4*>(tris),
        reinterpret_cast<const float4*>(rays), n, hits, static_cast<uint2*>(ts.spill));
    return hipGetLastError();
}

}  // namespace obvhs
