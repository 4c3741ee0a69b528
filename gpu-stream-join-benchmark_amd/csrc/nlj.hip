// nlj.hip — nested-loop join kernels (lazy and eager NLJ).
//
// Replaces nestedLoopCountKernel / nestedLoopJoinKernel
// (src/sjb/windowing/lazy/executor/GPU/GPU_NLJ.cu:28-64), the eager
// countKernel (src/sjb/windowing/eager/executor/GPU/GPU_NLJ.cu:40-54) and the
// bitmap kernels getJoinResultBitmap / getJoinResultBitmapEager
// (GPU_NLJ.cu:66-87, eager/executor/GPU/GPU_NLJ_BM.cu:55-76).
//
// Count kernel: each thread keeps 4 right (probe) keys in registers; the
// left keys of the workgroup's range are compacted once into a 4-B key array
// and come in through scalar loads (16 keys per instruction), shared by the
// whole wave, straight into the compare's scalar operand.  Each compare is
// one VALU instruction, v_cmpx_ne_u32, which takes the lanes holding an
// equal key out of EXEC; the lanes gone after a block of 64 compares are the
// lanes with a match, so a tile with no match costs nothing else.  Only
// waves with such a lane rescan the tile (32 keys per load batch, a cheap
// min-xor test per key) and add count, C1 and C2 per matching pair.  The NLJ
// is VALU-bound (n_L*n_R compares), not HBM-bound.  Workgroups add their
// sums to the window result with 64-bit atomics (k_nlj_keys zeroes it).
//
// Bitmap method: each row of a wave holds 64 consecutive right keys; for a
// left key, __ballot(lk == rk) over a row is exactly the 64-bit match word
// of that left tuple (the reference builds 32-bit words with one atomicAdd
// each and decodes them on the host).  The count kernel's v_cmpx prefilter
// finds the tiles with a match; only those are rescanned with ballots, and
// non-zero words are appended as (l, w, mask) triples and decoded on the
// device into ResultTuple rows.
#include <algorithm>

#include "sjb_device.hpp"
#include "sjb_internal.hpp"

namespace sjb {

constexpr int kNljThreads = 256;
#ifndef SJB_NLJ_PER
#define SJB_NLJ_PER 4
#endif
// 128-key tiles and 16 workgroups per CU (two rounds of the 8 that fit, so
// early finishers take more work) measured best at 2^16 x 2^16 per window:
// 32.1 T compares/s against 29.2 T with 256-key tiles and 8 per CU.
#ifndef SJB_NLJ_LTILE
#define SJB_NLJ_LTILE 128
#endif
#ifndef SJB_NLJ_WG_PER_CU
#define SJB_NLJ_WG_PER_CU 16
#endif
constexpr int kNljPer = SJB_NLJ_PER;  // right keys per thread
constexpr int kNljRTile = kNljThreads * kNljPer;  // right tuples per workgroup
constexpr int kNljLTile = SJB_NLJ_LTILE;          // left keys per tile (exact-pass granularity)
constexpr int kNljMaxBlocks = 65536;

constexpr int kNljKeySlack = 64;  // keys readable past the end of the key array

// Calls f(e, i) for every left key e < tn of the tile and right key i of
// this lane that are equal.  Keys come in 32 at a time through scalar loads
// (reads up to 31 keys past tn stay inside the key array's slack); each key
// costs a min-xor over the lane's keys and one bit of a per-lane hit mask,
// and only the set bits are revisited (a vector reload of that key).
template <class F>
__device__ __forceinline__ void nlj_tile_hits(const uint32_t *__restrict__ tk, int tn,
                                              const uint32_t (&sk)[kNljPer], const bool (&valid)[kNljPer],
                                              F &&f) {
    for (int e = 0; e < tn; e += 32) {
        uint32_t hm = 0;
#pragma unroll
        for (int c = 0; c < 32; c++) {
            const uint32_t lk = tk[e + c];
            uint32_t x = lk ^ sk[0];
#pragma unroll
            for (int i = 1; i < kNljPer; i++) x = min(x, lk ^ sk[i]);
            hm |= uint32_t(x == 0) << c;
        }
        if (tn - e < 32) hm &= (1u << (tn - e)) - 1u;
        while (hm) {
            const int c = __builtin_ctz(hm);
            hm &= hm - 1u;
            const uint32_t lk = tk[e + c];
#pragma unroll
            for (int i = 0; i < kNljPer; i++)
                if (valid[i] && lk == sk[i]) f(e + c, i);
        }
    }
}

// Lanes of the wave holding a right key equal to one of 16 left keys (kc,
// scalar registers).  Each compare is one v_cmpx_ne_u32, which drops from
// EXEC every lane whose key equals the left key, so 16 x kNljPer compares
// cost as many VALU instructions and nothing else: the lanes gone from EXEC
// at the end are the lanes with a match.  EXEC is saved and restored inside
// the one asm statement, so no compiler-generated instruction runs under the
// reduced mask.  v_cmpx also writes VCC (clobbered).  s_nop 4 gives the EXEC
// write of the last VALU compare five wait states before the scalar read of
// EXEC.
#define SJB_CMPX4(k, a, b, c, d)                        \
    "v_cmpx_ne_u32_e32 vcc, %[k" #k "], %[r" #a "]\n\t" \
    "v_cmpx_ne_u32_e32 vcc, %[k" #k "], %[r" #b "]\n\t" \
    "v_cmpx_ne_u32_e32 vcc, %[k" #k "], %[r" #c "]\n\t" \
    "v_cmpx_ne_u32_e32 vcc, %[k" #k "], %[r" #d "]\n\t"
#if SJB_NLJ_PER == 4
#define SJB_CMPX(k) SJB_CMPX4(k, 0, 1, 2, 3)
#define SJB_CMPX_RIGHT [r0] "v"(sk[0]), [r1] "v"(sk[1]), [r2] "v"(sk[2]), [r3] "v"(sk[3])
#elif SJB_NLJ_PER == 8
#define SJB_CMPX(k) SJB_CMPX4(k, 0, 1, 2, 3) SJB_CMPX4(k, 4, 5, 6, 7)
#define SJB_CMPX_RIGHT                                                                            \
    [r0] "v"(sk[0]), [r1] "v"(sk[1]), [r2] "v"(sk[2]), [r3] "v"(sk[3]), [r4] "v"(sk[4]), \
        [r5] "v"(sk[5]), [r6] "v"(sk[6]), [r7] "v"(sk[7])
#else
#error "SJB_NLJ_PER must be 4 or 8"
#endif
__device__ __forceinline__ uint64_t nlj_match16(const uint32_t (&kc)[16], const uint32_t (&sk)[kNljPer]) {
    uint64_t saved, gone;
    asm volatile(
        "s_mov_b64 %[sv], exec\n\t"
        SJB_CMPX(0) SJB_CMPX(1) SJB_CMPX(2) SJB_CMPX(3) SJB_CMPX(4) SJB_CMPX(5) SJB_CMPX(6) SJB_CMPX(7)
        SJB_CMPX(8) SJB_CMPX(9) SJB_CMPX(10) SJB_CMPX(11) SJB_CMPX(12) SJB_CMPX(13) SJB_CMPX(14) SJB_CMPX(15)
        "s_nop 4\n\t"
        "s_andn2_b64 %[gone], %[sv], exec\n\t"
        "s_mov_b64 exec, %[sv]"
        : [sv] "=&s"(saved), [gone] "=&s"(gone)
        : [k0] "s"(kc[0]), [k1] "s"(kc[1]), [k2] "s"(kc[2]), [k3] "s"(kc[3]), [k4] "s"(kc[4]),
          [k5] "s"(kc[5]), [k6] "s"(kc[6]), [k7] "s"(kc[7]), [k8] "s"(kc[8]), [k9] "s"(kc[9]),
          [k10] "s"(kc[10]), [k11] "s"(kc[11]), [k12] "s"(kc[12]), [k13] "s"(kc[13]), [k14] "s"(kc[14]),
          [k15] "s"(kc[15]), SJB_CMPX_RIGHT
        : "vcc", "scc");
    return gone;
}
#undef SJB_CMPX
#undef SJB_CMPX4
#undef SJB_CMPX_RIGHT

template <bool C2, bool WRITE>
__global__ __launch_bounds__(kNljThreads) void k_nlj(
    const uint32_t *__restrict__ lkeys, const sjb_tuple_t *__restrict__ L, uint64_t nL,
    const sjb_tuple_t *__restrict__ R, uint64_t nR, uint64_t l_split, sjb_window_result_t *__restrict__ res,
    sjb_result_tuple_t *__restrict__ out, uint64_t out_cap, int wrap, unsigned long long *__restrict__ cursor) {
    __shared__ uint64_t s_red[3 * kNljThreads / kWave];
    __shared__ uint32_t s_ws[kNljThreads / kWave];
    __shared__ unsigned long long s_base;

    const uint64_t rbase = uint64_t(blockIdx.x) * kNljRTile;
    uint32_t sk[kNljPer], sv[kNljPer];
    bool valid[kNljPer];
#pragma unroll
    for (int i = 0; i < kNljPer; i++) {
        const uint64_t r = rbase + threadIdx.x + uint64_t(i) * kNljThreads;
        valid[i] = r < nR;
        const uint2 kv = valid[i] ? *reinterpret_cast<const uint2 *>(&R[r]) : make_uint2(0, 0);
        sk[i] = kv.x;
        sv[i] = kv.y;
    }
    Acc acc;
    acc.zero();
    const uint64_t lbeg = uint64_t(blockIdx.y) * l_split;
    const uint64_t lend = min(nL, lbeg + l_split);

    for (uint64_t t0 = lbeg; t0 < lend; t0 += kNljLTile) {
        const int tn = int(min<uint64_t>(kNljLTile, lend - t0));
        // Left keys are the same for every lane: they come in through scalar
        // loads (16 per instruction) straight into the compare's scalar
        // operand (nlj_match16).
        const uint32_t *__restrict__ tk = lkeys + t0;
        const int t16 = tn & ~31;
        // software-pipelined key loads: the next 16 keys are requested
        // before the current 16 are compared (the reads past t16 stay in
        // the key array's slack and are not compared)
        uint32_t kc[16], kn[16];
        uint64_t hit = 0;  // lanes with a match in this tile (wave-uniform)
#pragma unroll
        for (int c = 0; c < 16; c++) kc[c] = tk[c];
        for (int e = 0; e < t16; e += 32) {  // t16 is a multiple of 32
            __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): kc has landed
#pragma unroll
            for (int c = 0; c < 16; c++) kn[c] = tk[e + 16 + c];
            hit |= nlj_match16(kc, sk);
            __builtin_amdgcn_s_waitcnt(0xc07f);  // kn has landed
#pragma unroll
            for (int c = 0; c < 16; c++) kc[c] = tk[e + 32 + c];
            hit |= nlj_match16(kn, sk);
        }
        uint32_t m = 0xFFFFFFFFu;
        for (int e = t16; e < tn; e++) {
            const uint32_t lk = tk[e];
#pragma unroll
            for (int i = 0; i < kNljPer; i++) m = min(m, lk ^ sk[i]);
        }
        // Exact pass only for lanes with an equal key.
        uint32_t mine = 0;
        if (((hit >> __lane_id()) & 1) || m == 0) {
            nlj_tile_hits(tk, tn, sk, valid, [&](int e, int i) {
                mine++;
                if (!WRITE) {
                    const uint64_t a = hash_a(sk[i], L[t0 + e].val);
                    acc.count++;
                    acc.c1 += a * hash_b(sk[i], sv[i]);
                    if (C2) acc.c2 += pair_hash_from_a(a, sv[i]);
                }
            });
        }
        if (WRITE) {
            uint32_t tot;
            const uint32_t ex = block_excl_scan_u32(mine, s_ws, &tot);
            if (threadIdx.x == 0) s_base = tot ? atomicAdd(cursor, (unsigned long long)tot) : 0ull;
            __syncthreads();
            if (mine) {
                uint64_t pos = s_base + ex;
                nlj_tile_hits(tk, tn, sk, valid, [&](int e, int i) {
                    const sjb_tuple_t lt = L[t0 + e];
                    const sjb_tuple_t st = R[rbase + threadIdx.x + uint64_t(i) * kNljThreads];
                    const uint64_t a = hash_a(lt.key, lt.val);
                    acc.count++;
                    acc.c1 += a * hash_b(st.key, st.val);
                    if (C2) acc.c2 += pair_hash_from_a(a, st.val);
                    uint64_t q = pos++;
                    if (wrap) q %= out_cap;
                    if (wrap || q < out_cap) {
                        sjb_result_tuple_t row;
                        row.lkey = lt.key;
                        row.rkey = st.key;
                        row.lval = lt.val;
                        row.rval = st.val;
                        row.lts = lt.ts;
                        row.rts = st.ts;
                        out[q] = row;
                    }
                });
            }
            __syncthreads();  // s_base is read before the next tile's reservation
        }
    }
    const Acc r = block_sum(acc, s_red);
    // the window result was zeroed by k_nlj_keys; only non-zero sums are added
    if (threadIdx.x == 0) {
        auto *d = reinterpret_cast<unsigned long long *>(res);
        if (r.count) atomicAdd(d, (unsigned long long)r.count);
        if (r.c1) atomicAdd(d + 1, (unsigned long long)r.c1);
        if (r.c2) atomicAdd(d + 2, (unsigned long long)r.c2);
    }
}

// Keys of n tuples, compacted (4 B each) for the scalar loads of k_nlj and
// k_nlj_bitmap.  It also zeroes the window result that k_nlj's workgroups
// add into (res non-null).
__global__ void k_nlj_keys(const sjb_tuple_t *__restrict__ L, uint64_t n, uint32_t *__restrict__ keys,
                           sjb_window_result_t *__restrict__ res) {
    if (res && blockIdx.x == 0 && threadIdx.x < sizeof(sjb_window_result_t) / 8)
        reinterpret_cast<uint64_t *>(res)[threadIdx.x] = 0;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
        keys[i] = L[i].key;
}

// Grid of k_nlj / k_nlj_bitmap (x: right tiles of kNljRTile tuples, y: left
// ranges of l_split keys, a multiple of kNljLTile) and the compacted left
// keys of the window (k_nlj_keys, which zeroes *res when res is non-null).
static int nlj_prepare(sjb_ctx *c, const sjb_tuple_t *L, uint64_t nL, uint64_t nR, sjb_window_result_t *res,
                       dim3 *grid, uint64_t *l_split_out) {
    const uint64_t gx = (nR + kNljRTile - 1) / kNljRTile;
    // Split the left side so the grid fills the chip (SJB_NLJ_WG_PER_CU
    // workgroups per CU).
    uint64_t want = uint64_t(c->n_cu) * SJB_NLJ_WG_PER_CU;
    uint64_t gy = (want + gx - 1) / gx;
    uint64_t max_gy = (nL + kNljLTile - 1) / kNljLTile;
    if (gy > max_gy) gy = max_gy;
    if (gy < 1) gy = 1;
    if (gx * gy > kNljMaxBlocks) gy = kNljMaxBlocks / gx;
    if (gy < 1 || gx > 65535) return SJB_E_CAPACITY;
    uint64_t l_split = (nL + gy - 1) / gy;
    l_split = (l_split + kNljLTile - 1) / kNljLTile * kNljLTile;
    gy = (nL + l_split - 1) / l_split;
    *grid = dim3((unsigned)gx, (unsigned)gy);
    *l_split_out = l_split;
    if (c->nlj_keys_cap < nL) {  // grow-only key buffer, owned by the ctx
        const uint64_t want_keys = std::max<uint64_t>(nL, c->cap);
        void *p = c->d_nlj_keys;
        int rc = dev_regrow(c, &p, c->d_nlj_keys ? (c->nlj_keys_cap + kNljKeySlack) * sizeof(uint32_t) : 0,
                            (want_keys + kNljKeySlack) * sizeof(uint32_t));
        c->d_nlj_keys = static_cast<uint32_t *>(p);
        c->nlj_keys_cap = c->d_nlj_keys ? want_keys : 0;
        if (rc) return rc;
    }
    k_nlj_keys<<<unsigned(std::min<uint64_t>((nL + 255) / 256, 4096)), 256, 0, c->stream>>>(L, nL, c->d_nlj_keys,
                                                                                          res);
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

int launch_nlj(sjb_ctx *c, const sjb_tuple_t *L, uint64_t nL, const sjb_tuple_t *R, uint64_t nR,
               uint32_t flags, int write, sjb_result_tuple_t *out, uint64_t out_cap, int wrap,
               sjb_window_result_t *d_out) {
    dim3 grid;
    uint64_t l_split;
    int rc = nlj_prepare(c, L, nL, nR, d_out, &grid, &l_split);
    if (rc) return rc;
    const bool c2 = (flags & SJB_F_C2) != 0;
    auto *cur = reinterpret_cast<unsigned long long *>(c->d_cursor);
    const uint32_t *lkeys = c->d_nlj_keys;
    if (write) {
        if (c2)
            k_nlj<true, true><<<grid, kNljThreads, 0, c->stream>>>(lkeys, L, nL, R, nR, l_split, d_out,
                                                                   out, out_cap, wrap, cur);
        else
            k_nlj<false, true><<<grid, kNljThreads, 0, c->stream>>>(lkeys, L, nL, R, nR, l_split, d_out,
                                                                    out, out_cap, wrap, cur);
    } else {
        if (c2)
            k_nlj<true, false><<<grid, kNljThreads, 0, c->stream>>>(lkeys, L, nL, R, nR, l_split, d_out,
                                                                    out, out_cap, wrap, cur);
        else
            k_nlj<false, false><<<grid, kNljThreads, 0, c->stream>>>(lkeys, L, nL, R, nR, l_split,
                                                                     d_out, out, out_cap, wrap, cur);
    }
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

// ---------------------------------------------------------------------------
// Bitmap method
// ---------------------------------------------------------------------------
// Masks of the Bitmap method, on the count kernel's layout: each lane holds
// kNljPer right keys (row i of a wave = one word of 64 consecutive right
// tuples), left keys come in by scalar loads, and nlj_match16's v_cmpx
// compares find the lanes with an equal key.  Only a wave whose tile has
// such a lane rescans that tile, one ballot per (left key, word): the ballot
// is the word's match mask for that left tuple, and non-zero masks are
// appended as (left index, word, mask) triples.
__global__ __launch_bounds__(kNljThreads) void k_nlj_bitmap(
    const uint32_t *__restrict__ lkeys, uint64_t nL, const sjb_tuple_t *__restrict__ R, uint64_t nR,
    uint64_t l_split, uint64_t *__restrict__ masks, uint64_t mask_cap,
    unsigned long long *__restrict__ n_masks) {
    const uint64_t rbase = uint64_t(blockIdx.x) * kNljRTile;
    const uint64_t wbase = rbase + (threadIdx.x & ~uint32_t(kWave - 1));  // lane 0's right index, row 0
    const bool leader = __lane_id() == 0;
    uint32_t sk[kNljPer];
    bool valid[kNljPer];
#pragma unroll
    for (int i = 0; i < kNljPer; i++) {
        const uint64_t r = rbase + threadIdx.x + uint64_t(i) * kNljThreads;
        valid[i] = r < nR;
        sk[i] = valid[i] ? R[r].key : 0u;
    }
    const uint64_t lbeg = uint64_t(blockIdx.y) * l_split;
    const uint64_t lend = min(nL, lbeg + l_split);
    for (uint64_t t0 = lbeg; t0 < lend; t0 += kNljLTile) {
        const int tn = int(min<uint64_t>(kNljLTile, lend - t0));
        const uint32_t *__restrict__ tk = lkeys + t0;
        const int t16 = tn & ~31;
        // the count kernel's prefilter (reads past t16 stay in the slack)
        uint32_t kc[16], kn[16];
        uint64_t hit = 0;
#pragma unroll
        for (int c = 0; c < 16; c++) kc[c] = tk[c];
        for (int e = 0; e < t16; e += 32) {
            __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): kc has landed
#pragma unroll
            for (int c = 0; c < 16; c++) kn[c] = tk[e + 16 + c];
            hit |= nlj_match16(kc, sk);
            __builtin_amdgcn_s_waitcnt(0xc07f);  // kn has landed
#pragma unroll
            for (int c = 0; c < 16; c++) kc[c] = tk[e + 32 + c];
            hit |= nlj_match16(kn, sk);
        }
        uint32_t m = 0xFFFFFFFFu;
        for (int e = t16; e < tn; e++) {
            const uint32_t lk = tk[e];
#pragma unroll
            for (int i = 0; i < kNljPer; i++) m = min(m, lk ^ sk[i]);
        }
        if ((hit | __ballot(m == 0)) == 0) continue;  // wave-uniform
        for (int e = 0; e < tn; e++) {
            const uint32_t lk = tk[e];
#pragma unroll
            for (int i = 0; i < kNljPer; i++) {
                const uint64_t mask = __ballot(valid[i] && lk == sk[i]);
                if (mask != 0 && leader) {
                    const unsigned long long slot = atomicAdd(n_masks, 1ull);
                    if (slot < mask_cap) {
                        masks[3 * slot] = t0 + uint64_t(e);
                        masks[3 * slot + 1] = (wbase + uint64_t(i) * kNljThreads) / kWave;
                        masks[3 * slot + 2] = mask;
                    }
                }
            }
        }
    }
}

// One thread per mask triple: reserve popcount(mask) rows per block, write
// them, and accumulate count/C1/C2.
__global__ __launch_bounds__(256) void k_bitmap_decode(
    const uint64_t *__restrict__ masks, const unsigned long long *__restrict__ n_masks,
    const sjb_tuple_t *__restrict__ L, const sjb_tuple_t *__restrict__ R,
    sjb_result_tuple_t *__restrict__ out, uint64_t out_cap, unsigned long long *__restrict__ cursor,
    uint64_t *__restrict__ partials) {
    __shared__ uint64_t s_red[3 * 256 / kWave];
    __shared__ uint32_t s_ws[256 / kWave];
    __shared__ unsigned long long s_base;
    const uint64_t n = *n_masks;
    Acc acc;
    acc.zero();
    for (uint64_t i0 = uint64_t(blockIdx.x) * 256; i0 < n; i0 += uint64_t(gridDim.x) * 256) {
        const uint64_t i = i0 + threadIdx.x;
        uint64_t mask = 0, l = 0, w = 0;
        if (i < n) {
            l = masks[3 * i];
            w = masks[3 * i + 1];
            mask = masks[3 * i + 2];
        }
        uint32_t tot;
        const uint32_t ex = block_excl_scan_u32(uint32_t(__popcll(mask)), s_ws, &tot);
        if (threadIdx.x == 0) s_base = tot ? atomicAdd(cursor, (unsigned long long)tot) : 0ull;
        __syncthreads();
        uint64_t pos = s_base + ex;
        if (mask) {
            const sjb_tuple_t lt = L[l];
            const uint64_t a = hash_a(lt.key, lt.val);
            while (mask) {
                const int bit = __ffsll((long long)mask) - 1;
                mask &= mask - 1;
                const sjb_tuple_t st = R[w * kWave + bit];
                acc.count++;
                acc.c1 += a * hash_b(st.key, st.val);
                acc.c2 += pair_hash_from_a(a, st.val);
                if (pos < out_cap) {
                    sjb_result_tuple_t row;
                    row.lkey = lt.key;
                    row.rkey = st.key;
                    row.lval = lt.val;
                    row.rval = st.val;
                    row.lts = lt.ts;
                    row.rts = st.ts;
                    out[pos] = row;
                }
                pos++;
            }
        }
        __syncthreads();
    }
    const Acc r = block_sum(acc, s_red);
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x] = r.count;
        partials[3 * blockIdx.x + 1] = r.c1;
        partials[3 * blockIdx.x + 2] = r.c2;
    }
}

int launch_nlj_bitmap(sjb_ctx *c, const sjb_tuple_t *L, uint64_t nL, const sjb_tuple_t *R,
                      uint64_t nR, uint32_t flags, sjb_result_tuple_t *out, uint64_t out_cap,
                      sjb_window_result_t *d_out, uint64_t *n_masks_host) {
    (void)flags;
    dim3 grid;
    uint64_t l_split;
    int rc0 = nlj_prepare(c, L, nL, nR, nullptr, &grid, &l_split);
    if (rc0) return rc0;
    auto *cnt = reinterpret_cast<unsigned long long *>(c->d_cursor) + 1;
    auto *cur = reinterpret_cast<unsigned long long *>(c->d_cursor);
    // The number of (left tuple, 64-key word) masks with a match is a
    // property of the window, not of the launch: when the first pass finds
    // more than the buffer holds (it stores none past the end, but counts
    // them all), the buffer grows to that count and the pass runs again.
    unsigned long long nm = 0;
    for (int pass = 0;; pass++) {
        int rc = launch_fill_u64(c->stream, c->d_cursor, 0, 2);
        if (rc) return rc;
        k_nlj_bitmap<<<grid, kNljThreads, 0, c->stream>>>(c->d_nlj_keys, nL, R, nR, l_split, c->d_masks,
                                                          c->mask_cap, cnt);
        SJB_HIP_TRY(hipGetLastError());
        SJB_HIP_TRY(hipMemcpyAsync(&nm, cnt, sizeof(nm), hipMemcpyDeviceToHost, c->stream));
        SJB_HIP_TRY(hipStreamSynchronize(c->stream));
        *n_masks_host = nm;
        if (nm <= c->mask_cap) break;
        if (pass > 0) return SJB_E_CAPACITY;  // cannot happen: the count is deterministic
        const uint64_t want = std::max<uint64_t>(nm, c->mask_cap + c->mask_cap / 2);
        void *p = c->d_masks;
        rc = dev_regrow(c, &p, 3 * c->mask_cap * sizeof(uint64_t), 3 * want * sizeof(uint64_t));
        c->d_masks = static_cast<uint64_t *>(p);
        c->mask_cap = c->d_masks ? want : 0;
        if (rc) return rc;
    }
    const int blocks = 512;
    k_bitmap_decode<<<blocks, 256, 0, c->stream>>>(c->d_masks, cnt, L, R, out, out_cap, cur,
                                                   c->d_partials);
    SJB_HIP_TRY(hipGetLastError());
    return launch_reduce(c, blocks, d_out, 0);
}

}  // namespace sjb
