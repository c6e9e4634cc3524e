// Shared device types and arithmetic helpers for the obvhs MI355X (gfx950) kernels.
//
// Byte layouts match the obvhs crate (aabb.rs:11-16, triangle.rs:8-13, ray.rs:15-30, ray.rs:63-70,
// bvh2/node.rs:38-66 and 12-36, cwbvh/node.rs:12-54).  Arithmetic reproduces the crate's x86 SSE semantics:
// min(a,b) = a<b ? a : b and max(a,b) = a>b ? a : b (glam Vec3A on SSE), no FMA contraction (the library is built
// with -ffp-contract=off), IEEE division and denormals preserved.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/obvhs_hip.h"

namespace obvhs {

constexpr uint32_t INVALID32 = 0xffffffffu;

// Device-side error bits, or-ed into a per-context error word and turned into an obvhs_status on the host.
enum : uint32_t {
    DERR_NAN_INPUT = 1u << 0,
    DERR_SELF_MERGE = 1u << 1,
    DERR_CW_DISTRIBUTE = 1u << 2,
    DERR_CW_NONUNIT_LEAF = 1u << 3,
    DERR_CW_ORDER = 1u << 4,
    DERR_CW_PRIMCOUNT = 1u << 5,
    DERR_CW_DEPTH = 1u << 6,
    DERR_CAPACITY = 1u << 7,
};

// PLOC working cluster / compact BVH2 node: the obvhs `small_bvh2_node` layout (bvh2/node.rs:12-36), 32 bytes.
struct __align__(16) Node32 {
    float minx, miny, minz;
    uint32_t prim_count;
    float maxx, maxy, maxz;
    uint32_t first_index;
};
static_assert(sizeof(Node32) == 32, "Node32 must be 32 bytes");

struct __align__(16) CwNode {
    float p[3];
    uint8_t e[3];
    uint8_t imask;
    uint32_t child_base_idx;
    uint32_t primitive_base_idx;
    uint8_t child_meta[8];
    uint8_t child_min_x[8], child_max_x[8], child_min_y[8], child_max_y[8], child_min_z[8], child_max_z[8];
};
static_assert(sizeof(CwNode) == 80, "CwBvhNode must be 80 bytes");

__device__ __forceinline__ float smin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float smax(float a, float b) { return a > b ? a : b; }

struct Box {
    float lx, ly, lz, hx, hy, hz;
};

__device__ __forceinline__ Box box_union(const Box& a, const Box& b) {
    return Box{smin(a.lx, b.lx), smin(a.ly, b.ly), smin(a.lz, b.lz),
               smax(a.hx, b.hx), smax(a.hy, b.hy), smax(a.hz, b.hz)};
}

// aabb.rs:151-154 -- (dx + dy) * dz + dx * dy, unfused.
__device__ __forceinline__ float half_area(const Box& a) {
    float dx = a.hx - a.lx, dy = a.hy - a.ly, dz = a.hz - a.lz;
    return (dx + dy) * dz + dx * dy;
}

// half_area(box_union(a, b)) for boxes without NaN: the IEEE minimum / maximum instructions (one each) give the
// union's components in value (only the sign of a zero may differ), so the area compares equal; for neighbour
// searches, whose costs only meet compares.
__device__ __forceinline__ float union_area_nonan(const Box& a, const Box& b) {
    const Box u{__builtin_elementwise_minimum(a.lx, b.lx), __builtin_elementwise_minimum(a.ly, b.ly),
                __builtin_elementwise_minimum(a.lz, b.lz), __builtin_elementwise_maximum(a.hx, b.hx),
                __builtin_elementwise_maximum(a.hy, b.hy), __builtin_elementwise_maximum(a.hz, b.hz)};
    return half_area(u);
}

__device__ __forceinline__ bool box_has_nan_any(const Box& b) {
    return b.lx != b.lx || b.ly != b.ly || b.lz != b.lz || b.hx != b.hx || b.hy != b.hy || b.hz != b.hz;
}

__device__ __forceinline__ Box node_box(const Node32& n) { return Box{n.minx, n.miny, n.minz, n.maxx, n.maxy, n.maxz}; }

__device__ __forceinline__ Node32 load_node(const Node32* p) {
    const float4* q = reinterpret_cast<const float4*>(p);
    float4 a = q[0], b = q[1];
    Node32 n;
    n.minx = a.x;
    n.miny = a.y;
    n.minz = a.z;
    n.prim_count = __float_as_uint(a.w);
    n.maxx = b.x;
    n.maxy = b.y;
    n.maxz = b.z;
    n.first_index = __float_as_uint(b.w);
    return n;
}

__device__ __forceinline__ void store_node(Node32* p, const Node32& n) {
    float4* q = reinterpret_cast<float4*>(p);
    q[0] = make_float4(n.minx, n.miny, n.minz, __uint_as_float(n.prim_count));
    q[1] = make_float4(n.maxx, n.maxy, n.maxz, __uint_as_float(n.first_index));
}

__device__ __forceinline__ Node32 make_node(const Box& b, uint32_t prim_count, uint32_t first_index) {
    return Node32{b.lx, b.ly, b.lz, prim_count, b.hx, b.hy, b.hz, first_index};
}

// Order-preserving map of f32 onto u32 (for atomic min/max of the scene bounds).
__device__ __forceinline__ uint32_t f32_order_key(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ __forceinline__ float f32_from_order_key(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    union {
        uint32_t u;
        float f;
    } c;
    c.u = u;
    return c.f;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

}  // namespace obvhs
