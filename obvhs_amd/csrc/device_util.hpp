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
