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

// Exclusive scans over a 1024-thread workgroup of two u32 values at once; ta / tb receive the totals.  se / sm: 16
// words of LDS each (reused: the function ends with a barrier).
__device__ __forceinline__ void block_exscan2_1024(uint32_t a, uint32_t b, uint32_t* se, uint32_t* sm, uint32_t& ea,
                                                   uint32_t& eb, uint32_t& ta, uint32_t& tb) {
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t ia = wave_inclusive_scan(a), ib = wave_inclusive_scan(b);
    if (lane == 63u) {
        se[wave] = ia;
        sm[wave] = ib;
    }
    __syncthreads();
    uint32_t pa = 0u, pb = 0u;
    ta = 0u;
    tb = 0u;
#pragma unroll
    for (uint32_t w = 0; w < 16; ++w) {
        const uint32_t xa = se[w], xb = sm[w];
        if (w < wave) {
            pa += xa;
            pb += xb;
        }
        ta += xa;
        tb += xb;
    }
    ea = pa + ia - a;
    eb = pb + ib - b;
    __syncthreads();
}

// Two-level scan of per-tile (emit, merge) counts, SCAN_CHUNK counts (4 per thread) per 1,024-thread workgroup.
// scan_chunk_total: the chunk's totals into chunk_tot[blockIdx.x].  scan_chunk_apply: every count becomes its
// exclusive prefix over all tiles (the chunks before this one from chunk_tot); returns the totals of all chunks.
constexpr uint32_t SCAN_CHUNK = 4096;

__device__ __forceinline__ void scan_chunk_total(const uint2* __restrict__ counts, uint32_t n,
                                                 uint2* __restrict__ chunk_tot) {
    __shared__ uint32_t se[16], sm[16];
    const uint32_t b0 = blockIdx.x * SCAN_CHUNK + threadIdx.x * 4u;
    uint32_t E = 0u, M = 0u;
#pragma unroll
    for (uint32_t k = 0; k < 4u; ++k)
        if (b0 + k < n) {
            const uint2 v = counts[b0 + k];
            E += v.x;
            M += v.y;
        }
    uint32_t ee, em, te, tm;
    block_exscan2_1024(E, M, se, sm, ee, em, te, tm);
    if (threadIdx.x == 0) chunk_tot[blockIdx.x] = make_uint2(te, tm);
}

__device__ __forceinline__ uint2 scan_chunk_apply(uint2* __restrict__ counts, uint32_t n,
                                                  const uint2* __restrict__ chunk_tot, uint32_t nchunks) {
    __shared__ uint32_t se[16], sm[16];
    uint32_t pe = 0u, pm = 0u, ae = 0u, am = 0u;
    for (uint32_t c = threadIdx.x; c < nchunks; c += 1024u) {
        const uint2 v = chunk_tot[c];
        if (c < blockIdx.x) {
            pe += v.x;
            pm += v.y;
        }
        ae += v.x;
        am += v.y;
    }
    uint32_t x0, x1, before_e, before_m, all_e, all_m;
    block_exscan2_1024(pe, pm, se, sm, x0, x1, before_e, before_m);
    block_exscan2_1024(ae, am, se, sm, x0, x1, all_e, all_m);
    const uint32_t b0 = blockIdx.x * SCAN_CHUNK + threadIdx.x * 4u;
    uint2 v[4];
    uint32_t E = 0u, M = 0u;
#pragma unroll
    for (uint32_t k = 0; k < 4u; ++k) {
        v[k] = b0 + k < n ? counts[b0 + k] : make_uint2(0u, 0u);
        E += v[k].x;
        M += v[k].y;
    }
    uint32_t ee, em, te, tm;
    block_exscan2_1024(E, M, se, sm, ee, em, te, tm);
    ee += before_e;
    em += before_m;
#pragma unroll
    for (uint32_t k = 0; k < 4u; ++k)
        if (b0 + k < n) {
            counts[b0 + k] = make_uint2(ee, em);
            ee += v[k].x;
            em += v[k].y;
        }
    return make_uint2(all_e, all_m);
}

}  // namespace obvhs
