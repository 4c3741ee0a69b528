// sjb_device.hpp — device-side primitives shared by every gfx950 kernel.
//
// Checksum formulas are the ones documented in include/sjb.h.  Reductions
// are written for 64-lane wavefronts (CDNA4).  Where the whole wave is
// active (next to a barrier) a wave reduces and scans with DPP lane moves
// and readlane — VALU instructions of a few cycles' latency — instead of
// __shfl, which goes through the LDS unit (ds_bpermute) at LDS latency per
// step; then one LDS slot per wave finishes the block.
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

// Sum over the active lanes of the wave (any exec mask).
__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// DPP move of lane values within 16-lane rows (VALU).  Lanes whose source
// lies outside the row, and rows outside ROW_MASK, get 0.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return uint32_t(__builtin_amdgcn_update_dpp(0, int(v), CTRL, ROW_MASK, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
    return uint64_t(dpp_u32<CTRL>(uint32_t(v))) | (uint64_t(dpp_u32<CTRL>(uint32_t(v >> 32))) << 32);
}
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int lane) {
    return uint64_t(uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v)), lane))) |
           (uint64_t(uint32_t(__builtin_amdgcn_readlane(int(uint32_t(v >> 32)), lane))) << 32);
}

// Sum over the 64 lanes, returned in every lane.  Every lane must be active:
// DPP and readlane read an inactive lane's register as it stands.
__device__ __forceinline__ uint64_t wave_sum_full(uint64_t v) {
    v += dpp_u64<0xB1>(v);   // quad_perm [1,0,3,2]: + lane ^ 1
    v += dpp_u64<0x4E>(v);   // quad_perm [2,3,0,1]: + lane ^ 2 -> quad sums
    v += dpp_u64<0x141>(v);  // row_half_mirror: + the other quad of the 8 lanes
    v += dpp_u64<0x140>(v);  // row_mirror: + the other 8 lanes -> row sums
    return readlane_u64(v, 0) + readlane_u64(v, 16) + readlane_u64(v, 32) + readlane_u64(v, 48);
}

// Minimum and maximum over the 64 lanes, returned in every lane.  Every lane
// must be active (as for wave_sum_full); the four lane moves have a source
// in every lane.
__device__ __forceinline__ void wave_minmax_full(uint32_t &mn, uint32_t &mx) {
    mn = min(mn, dpp_u32<0xB1>(mn));
    mx = max(mx, dpp_u32<0xB1>(mx));
    mn = min(mn, dpp_u32<0x4E>(mn));
    mx = max(mx, dpp_u32<0x4E>(mx));
    mn = min(mn, dpp_u32<0x141>(mn));
    mx = max(mx, dpp_u32<0x141>(mx));
    mn = min(mn, dpp_u32<0x140>(mn));
    mx = max(mx, dpp_u32<0x140>(mx));
    const uint32_t n0 = min(uint32_t(__builtin_amdgcn_readlane(int(mn), 0)), uint32_t(__builtin_amdgcn_readlane(int(mn), 16)));
    const uint32_t n1 = min(uint32_t(__builtin_amdgcn_readlane(int(mn), 32)), uint32_t(__builtin_amdgcn_readlane(int(mn), 48)));
    const uint32_t x0 = max(uint32_t(__builtin_amdgcn_readlane(int(mx), 0)), uint32_t(__builtin_amdgcn_readlane(int(mx), 16)));
    const uint32_t x1 = max(uint32_t(__builtin_amdgcn_readlane(int(mx), 32)), uint32_t(__builtin_amdgcn_readlane(int(mx), 48)));
    mn = min(n0, n1);
    mx = max(x0, x1);
}

// Exclusive prefix over the 64 lanes; *total = the wave's sum.  Every lane
// must be active (as for wave_sum_full).
__device__ __forceinline__ uint32_t wave_excl_scan_full(uint32_t v, uint32_t *total) {
    uint32_t x = v;
    x += dpp_u32<0x111>(x);       // row_shr:1
    x += dpp_u32<0x112>(x);       // row_shr:2
    x += dpp_u32<0x114>(x);       // row_shr:4
    x += dpp_u32<0x118>(x);       // row_shr:8 -> inclusive scan of each 16-lane row
    x += dpp_u32<0x142, 0xA>(x);  // row_bcast:15: rows 1 and 3 add the previous row's last lane
    x += dpp_u32<0x143, 0xC>(x);  // row_bcast:31: rows 2 and 3 add lane 31 (rows 0-1)
    *total = uint32_t(__builtin_amdgcn_readlane(int(x), kWave - 1));
    return x - v;
}

// Block-wide sum of an Acc; result valid in thread 0.  `scratch` holds
// 3 * (blockDim/64) u64.  Ends with a barrier so scratch can be reused.
__device__ __forceinline__ Acc block_sum(Acc a, uint64_t *scratch) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    a.count = wave_sum_full(a.count);
    a.c1 = wave_sum_full(a.c1);
    a.c2 = wave_sum_full(a.c2);
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
    const uint32_t ex = wave_excl_scan_full(v, &wt);
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
