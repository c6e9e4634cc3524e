// Stable LSD radix sort of (u64 Morton code, u32 leaf index) pairs for gfx950.
//
// Replaces the Morton sort of ploc/mod.rs:811-827.  The reference uses unstable sorts whose tie order is not
// reproducible; this sort is stable, so ties between equal codes keep ascending leaf index (the project rule the
// oracle follows too).
//
// Per 8-bit digit pass (reduce-then-scan, three launches):
//   upsweep   : per-tile digit histogram in LDS -> counts[digit][tile] (no global atomics);
//   scan      : one workgroup per digit, exclusive scan of counts[digit][*] within the digit, and the digit's
//               total;
//   downsweep : the digit bases (an exclusive scan of the 256 totals, in every tile), then wave-level multisplit
//               ranking (eight 64-lane ballots per item) gives each key its stable rank inside the tile; the tile's
//               keys are reordered in LDS and written out coalesced per digit run, then its values through the same
//               LDS (half the LDS of staging both at once, so twice the tiles per CU).
// Tiles are 256 threads x ITEMS keys (ITEMS = 4, 8 or 16, the sort tile the context's tuning table picks per size
// class; 8 by default); each wave owns a contiguous 64 x ITEMS key segment.  Every tile size gives the same order.
#include "common.hpp"
#include "device_util.hpp"
#include "internal.hpp"

namespace obvhs {

namespace {

constexpr int SORT_THREADS = 256;
constexpr int SORT_MIN_ITEMS = 4;  // the scratch is sized for the smallest tile (the most tiles)
constexpr int SORT_WAVES = SORT_THREADS / 64;

template <int ITEMS>
__global__ void __launch_bounds__(SORT_THREADS)
    k_radix_upsweep(const uint64_t* __restrict__ keys, uint32_t n, uint32_t shift, uint32_t num_tiles,
                    uint32_t* __restrict__ counts) {
    constexpr int SORT_TILE = SORT_THREADS * ITEMS;
    __shared__ uint32_t hist[256];
    const uint32_t tid = threadIdx.x;
    hist[tid] = 0u;
    __syncthreads();
    const uint32_t base = blockIdx.x * SORT_TILE;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        uint32_t idx = base + j * SORT_THREADS + tid;
        if (idx < n) {
            uint32_t d = (uint32_t)(keys[idx] >> shift) & 0xffu;
            atomicAdd(&hist[d], 1u);
        }
    }
    __syncthreads();
    counts[tid * num_tiles + blockIdx.x] = hist[tid];
}

// Exclusive scans over the 256 threads of two u32 values at once (wave scans, then the four wave totals through
// LDS); ta and tb receive the totals.  tmp: 8 words of LDS.
__device__ __forceinline__ void block_exclusive_scan2_256(uint32_t a, uint32_t b, uint32_t* tmp, uint32_t& ea,
                                                          uint32_t& eb, uint32_t& ta, uint32_t& tb) {
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t ia = wave_inclusive_scan(a), ib = wave_inclusive_scan(b);
    if (lane == 63u) {
        tmp[wave] = ia;
        tmp[4 + wave] = ib;
    }
    __syncthreads();
    uint32_t pa = 0u, pb = 0u;
    ta = 0u;
    tb = 0u;
#pragma unroll
    for (uint32_t w = 0; w < SORT_WAVES; ++w) {
        const uint32_t xa = tmp[w], xb = tmp[4 + w];
        if (w < wave) {
            pa += xa;
            pb += xb;
        }
        ta += xa;
        tb += xb;
    }
    ea = pa + ia - a;
    eb = pb + ib - b;
    __syncthreads();  // tmp is reused by the caller's next scan
}

// One workgroup per digit: exclusive scan of the digit's per-tile counts within the digit, and the digit's total.
__global__ void __launch_bounds__(SORT_THREADS)
    k_radix_scan(uint32_t* __restrict__ counts, uint32_t* __restrict__ totals, uint32_t num_tiles) {
    __shared__ uint32_t tmp[8];
    const uint32_t d = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    uint32_t carry = 0u;
    uint32_t* row = counts + (size_t)d * num_tiles;
    for (uint32_t start = 0; start < num_tiles; start += 2 * SORT_THREADS) {
        const uint32_t i = start + 2 * tid;
        const uint32_t v0 = i < num_tiles ? row[i] : 0u, v1 = i + 1 < num_tiles ? row[i + 1] : 0u;
        uint32_t ex, dummy_e, chunk_total, dummy_t;
        block_exclusive_scan2_256(v0 + v1, 0u, tmp, ex, dummy_e, chunk_total, dummy_t);
        if (i < num_tiles) row[i] = carry + ex;
        if (i + 1 < num_tiles) row[i + 1] = carry + ex + v0;
        carry += chunk_total;
    }
    if (tid == 0) totals[d] = carry;
}

template <int ITEMS>
__global__ void __launch_bounds__(SORT_THREADS)
    k_radix_downsweep(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                      uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n, uint32_t shift,
                      const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ totals,
                      uint32_t num_tiles) {
    constexpr int SORT_TILE = SORT_THREADS * ITEMS;
    __shared__ uint32_t whist[SORT_WAVES][256];
    __shared__ uint32_t digit_start[256];
    __shared__ uint32_t gbase[256];
    __shared__ uint32_t scan_tmp[8];
    __shared__ uint64_t stage[SORT_TILE];  // the tile's keys, then its values

    const uint32_t tid = threadIdx.x;
    const uint32_t wave = tid >> 6;
    const uint32_t lane = tid & 63u;
    const uint32_t tile = blockIdx.x;
    const uint32_t tile_base = tile * SORT_TILE;
    const uint32_t tile_count = min((uint32_t)SORT_TILE, n - tile_base);

#pragma unroll
    for (int w = 0; w < SORT_WAVES; ++w) whist[w][tid] = 0u;
    const uint32_t dtotal = totals[tid], toff = offsets[(size_t)tid * num_tiles + tile];
    __syncthreads();

    const uint32_t seg = tile_base + wave * (64 * ITEMS);
    uint64_t k[ITEMS];
    uint32_t v[ITEMS];
    uint32_t rank[ITEMS];
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        uint32_t idx = seg + j * 64 + lane;
        bool valid = idx < n;
        k[j] = valid ? kin[idx] : 0ull;
        v[j] = valid ? (vin ? vin[idx] : idx) : 0u;
        uint32_t d = (uint32_t)(k[j] >> shift) & 0xffu;
        uint64_t peers = ballot64(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool bit = (d >> b) & 1u;
            uint64_t bm = ballot64(bit);
            peers &= bit ? bm : ~bm;
        }
        uint32_t below = (uint32_t)__popcll(peers & lt);
        uint32_t base = 0;
        if (valid) base = whist[wave][d];
        __builtin_amdgcn_wave_barrier();
        if (valid && below == 0) whist[wave][d] = base + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rank[j] = base + below;
    }
    __syncthreads();
    // Exclusive prefix across waves for each digit, then across digits for the tile.
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < SORT_WAVES; ++w) {
        uint32_t c = whist[w][tid];
        whist[w][tid] = run;
        run += c;
    }
    uint32_t start, dbase, tile_total, n_total;
    block_exclusive_scan2_256(run, dtotal, scan_tmp, start, dbase, tile_total, n_total);
    digit_start[tid] = start;
    gbase[tid] = dbase + toff;
    __syncthreads();
    uint32_t pos[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t d = (uint32_t)(k[j] >> shift) & 0xffu;
        pos[j] = digit_start[d] + whist[wave][d] + rank[j];
        if (seg + j * 64 + lane < n) stage[pos[j]] = k[j];
    }
    __syncthreads();
    uint32_t dst[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t s = tid + (uint32_t)j * SORT_THREADS;
        if (s < tile_count) {
            const uint64_t key = stage[s];
            const uint32_t d = (uint32_t)(key >> shift) & 0xffu;
            dst[j] = gbase[d] + (s - digit_start[d]);
            kout[dst[j]] = key;
        }
    }
    __syncthreads();
    uint32_t* sv = reinterpret_cast<uint32_t*>(stage);
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
        if (seg + j * 64 + lane < n) sv[pos[j]] = v[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t s = tid + (uint32_t)j * SORT_THREADS;
        if (s < tile_count) vout[dst[j]] = sv[s];
    }
}

}  // namespace

size_t radix_sort_scratch_bytes(uint32_t n) {
    constexpr uint32_t MIN_TILE = SORT_THREADS * SORT_MIN_ITEMS;
    uint32_t tiles = (n + MIN_TILE - 1) / MIN_TILE;
    size_t b = 0;
    b += (size_t)n * 8;                     // alternate keys
    b += (size_t)n * 4;                     // alternate values
    b += (size_t)256 * tiles * 4 + 256;     // counts
    b += 256 * 4 + 256;                     // digit totals of the pass
    return b + 1024;
}

// Sorts keys[0..n) ascending (stable) with values.  On entry values may be null (identity 0..n).  `bits` is the
// number of low key bits that can be non-zero (63 for Morton64).  Result lands in keys/values (even pass count)
// or is copied back.  scratch must hold radix_sort_scratch_bytes(n).  items: keys per thread of a tile (4, 8, 16).
template <int ITEMS>
static hipError_t radix_sort_impl(uint64_t* keys, uint32_t* values, bool values_identity, uint32_t n, int bits,
                                  void* scratch, hipStream_t stream) {
    constexpr int SORT_TILE = SORT_THREADS * ITEMS;
    uint32_t tiles = (n + SORT_TILE - 1) / SORT_TILE;
    char* p = static_cast<char*>(scratch);
    auto take = [&](size_t bytes) {
        char* r = p;
        p += (bytes + 255) & ~size_t(255);
        return r;
    };
    uint64_t* kalt = reinterpret_cast<uint64_t*>(take((size_t)n * 8));
    uint32_t* valt = reinterpret_cast<uint32_t*>(take((size_t)n * 4));
    uint32_t* counts = reinterpret_cast<uint32_t*>(take((size_t)256 * tiles * 4));
    uint32_t* totals = reinterpret_cast<uint32_t*>(take(256 * 4));
    int passes = (bits + 7) / 8;
    uint64_t* kin = keys;
    uint32_t* vin = values;
    uint64_t* kout = kalt;
    uint32_t* vout = valt;
    for (int pass = 0; pass < passes; ++pass) {
        uint32_t shift = (uint32_t)(pass * 8);
        k_radix_upsweep<ITEMS><<<tiles, SORT_THREADS, 0, stream>>>(kin, n, shift, tiles, counts);
        k_radix_scan<<<256, SORT_THREADS, 0, stream>>>(counts, totals, tiles);
        const uint32_t* vsrc = (pass == 0 && values_identity) ? nullptr : vin;
        k_radix_downsweep<ITEMS><<<tiles, SORT_THREADS, 0, stream>>>(kin, vsrc, kout, vout, n, shift, counts,
                                                                     totals, tiles);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (kin != keys) {
        e = hipMemcpyAsync(keys, kin, (size_t)n * 8, hipMemcpyDeviceToDevice, stream);
        if (e != hipSuccess) return e;
        e = hipMemcpyAsync(values, vin, (size_t)n * 4, hipMemcpyDeviceToDevice, stream);
    }
    return e;
}

hipError_t radix_sort_pairs(uint64_t* keys, uint32_t* values, bool values_identity, uint32_t n, int bits,
                            void* scratch, hipStream_t stream, int items) {
    if (n == 0) return hipSuccess;
    switch (items) {
        case 4: return radix_sort_impl<4>(keys, values, values_identity, n, bits, scratch, stream);
        case 16: return radix_sort_impl<16>(keys, values, values_identity, n, bits, scratch, stream);
        default: return radix_sort_impl<8>(keys, values, values_identity, n, bits, scratch, stream);
    }
}

}  // namespace obvhs
