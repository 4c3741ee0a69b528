// sjb_device.hpp — device-side primitives shared by every gfx950 kernel.
//
// Checksum formulas are the ones documented in include/sjb.h.  Reductions
// are written for 64-lane wavefronts (CDNA4): a wave reduces with
// __shfl_xor over 6 steps, then one LDS slot per wave finishes the block.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sjb.h"

namespace sjb {

constexpr uint64_t kSeedA = 0x243F6A8885A308D3ULL;
constexpr uint64_t kSeedB = 0x13198A2E03707344ULL;
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;
constexpr uint32_t kHashMul = 0x9E3779B1u;  // Fibonacci hashing: bijective on u32
constexpr int kWave = 64;

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ uint64_t hash_a(uint32_t k, uint32_t v) {
    return mix64(((uint64_t(k) << 32) | v) ^ kSeedA);
}
__host__ __device__ __forceinline__ uint64_t hash_b(uint32_t k, uint32_t v) {
    return mix64(((uint64_t(k) << 32) | v) ^ kSeedB);
}
__host__ __device__ __forceinline__ uint64_t pair_hash_from_a(uint64_t a, uint32_t sv) {
    return mix64(a + kGolden * ((uint64_t(sv) << 1) | 1ULL));
}
__host__ __device__ __forceinline__ uint32_t key_hash(uint32_t k) { return k * kHashMul; }

// Running (count, C1, C2) triple; every field is a sum mod 2^64, so any
// reduction order gives the same bits.
struct Acc {
    uint64_t count, c1, c2;
    __device__ __forceinline__ void zero() { count = c1 = c2 = 0; }
    __device__ __forceinline__ void add(const Acc &o) {
        count += o.count;
        c1 += o.c1;
        c2 += o.c2;
    }
};

__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Block-wide sum of an Acc; result valid in thread 0.  `scratch` holds
// 3 * (blockDim/64) u64.  Ends with a barrier so scratch can be reused.
__device__ __forceinline__ Acc block_sum(Acc a, uint64_t *scratch) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    a.count = wave_sum(a.count);
    a.c1 = wave_sum(a.c1);
    a.c2 = wave_sum(a.c2);
    if (lane == 0) {
        scratch[wid] = a.count;
        scratch[nw + wid] = a.c1;
        scratch[2 * nw + wid] = a.c2;
    }
    __syncthreads();
    Acc r;
    r.zero();
    if (threadIdx.x == 0) {
        for (int w = 0; w < nw; w++) {
            r.count += scratch[w];
            r.c1 += scratch[nw + w];
            r.c2 += scratch[2 * nw + w];
        }
    }
    __syncthreads();
    return r;
}

// Exclusive prefix of a per-lane u32 over the wave via ballot-free shuffles.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t *total) {
    const int lane = threadIdx.x & (kWave - 1);
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        uint32_t y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    *total = __shfl(x, kWave - 1, kWave);
    return x - v;
}

// Block-exclusive scan of one u32 per thread; returns the exclusive prefix,
// *total the block sum.  `ws` holds blockDim/64 u32.  Every thread of the
// block must call it (two barriers).
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t *ws, uint32_t *total) {
    uint32_t wt;
    const uint32_t ex = wave_excl_scan(v, &wt);
    const int wid = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) ws[wid] = wt;
    __syncthreads();
    uint32_t pre = 0, all = 0;
    for (int w = 0; w < int(blockDim.x / kWave); w++) {
        if (w < wid) pre += ws[w];
        all += ws[w];
    }
    __syncthreads();
    *total = all;
    return pre + ex;
}

// Lane rank among the active lanes whose predicate is set (mbcnt on the
// ballot mask), used for wave-aggregated output reservations.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

}  // namespace sjb
