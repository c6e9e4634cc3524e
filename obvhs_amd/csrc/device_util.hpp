// Small device helpers shared by the PLOC kernels of the single and batched builds.
#pragma once

#include "common.hpp"

namespace obvhs {

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = (int)__lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = (uint32_t)__shfl_up((int)v, off);
        if (lane >= off) v += t;
    }
    return v;
}

__device__ __forceinline__ bool box_has_nan(const Box& b) {
    return isnan(b.lx) || isnan(b.ly) || isnan(b.lz) || isnan(b.hx) || isnan(b.hy) || isnan(b.hz);
}

// morton.rs:35-43
__device__ __forceinline__ uint64_t split_by_3_u64(uint32_t a) {
    uint64_t x = (uint64_t)a & 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}
// Rust `f64 as u32` (saturating, NaN -> 0).
__device__ __forceinline__ uint32_t f64_as_u32(double v) {
    if (!(v >= 1.0)) return 0u;
    if (v >= 4294967296.0) return 0xffffffffu;
    return (uint32_t)v;
}

// Rust `f64 as u64` (saturating, NaN -> 0).
__device__ __forceinline__ uint64_t f64_as_u64(double v) {
    if (!(v >= 1.0)) return 0u;
    if (v >= 18446744073709551616.0) return ~0ull;
    return (uint64_t)v;
}

// 126-bit Morton code of the unit-cube point (px, py, pz) (morton_encode_u128_unorm, morton.rs:63-89) as its low
// and high 64-bit words.  Each 42-bit coordinate is x = xh * 2^21 + xl; spreading puts bit i of x at 3i, so xl
// fills bits [0, 63) and xh bits [63, 126): the code is code64(xl, yl, zl) + code64(xh, yh, zh) * 2^63.
__device__ __forceinline__ void morton128_of_point(double px, double py, double pz, uint64_t& lo, uint64_t& hi) {
    const double s = 4398046511104.0;  // 2^42
    const uint64_t m = 0x3ffffffffffull;
    const uint64_t x = f64_as_u64(px * s) & m, y = f64_as_u64(py * s) & m, z = f64_as_u64(pz * s) & m;
    // split_by_3_u64 keeps the low 21 bits of its argument
    const uint64_t a = split_by_3_u64((uint32_t)x) | split_by_3_u64((uint32_t)y) << 1 |
                       split_by_3_u64((uint32_t)z) << 2;
    const uint64_t b = split_by_3_u64((uint32_t)(x >> 21)) | split_by_3_u64((uint32_t)(y >> 21)) << 1 |
                       split_by_3_u64((uint32_t)(z >> 21)) << 2;
    lo = a | b << 63;
    hi = b >> 1;
}

// Morton code of a leaf box under the scene normalisation (ploc/mod.rs:287-288, 782-785; morton.rs:55-58).
__device__ __forceinline__ uint64_t morton_of_box(const Box& b, double sx, double sy, double sz, double ox, double oy,
                                                  double oz) {
    const float cx = (b.hx + b.lx) * 0.5f, cy = (b.hy + b.ly) * 0.5f, cz = (b.hz + b.lz) * 0.5f;
    const double px = (double)cx * sx + ox, py = (double)cy * sy + oy, pz = (double)cz * sz + oz;
    const double s = 2097152.0;
    return split_by_3_u64(f64_as_u32(px * s)) | split_by_3_u64(f64_as_u32(py * s)) << 1 |
           split_by_3_u64(f64_as_u32(pz * s)) << 2;
}

}  // namespace obvhs
