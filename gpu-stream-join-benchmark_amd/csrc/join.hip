// join.hip — per-partition join kernels of the lazy HJ and SMJ operators.
//
// After partition.hip, partition p of the window is left records
// d_part[lb, le) and right records d_part[rb, re).  One workgroup joins one
// partition entirely in LDS:
//
//   HJ  (replaces buildHashTable + probeCount/probeWrite/probeJoin,
//        src/sjb/windowing/lazy/executor/GPU/GPU_HJ.cu:22-133): the build
//        tile is inserted into an LDS chained table (u32 heads, u16 links,
//        bucket = hash bits below the partition bits); every probe record
//        walks its chain and compares keys.
//   SMJ (replaces smjSortKernel + smjMergeKernel{Count,Write,_ES},
//        GPU_SMJ.cu:19-113): both tiles are bitonic-sorted in LDS as
//        (key << 32 | payload); each probe element co-ranks its key in the
//        sorted build tile (lower/upper bound) — the merge step — and the
//        equal-key run [lo, hi) gives count hi-lo and, in grouped mode,
//        C1 = (prefixA[hi] - prefixA[lo]) * B(s) in closed form
//        (the reference's count closed form is GPU_SMJ.cu:34-44).
//
// Build partitions larger than one LDS tile are processed tile by tile
// (each probe record meets every tile), so any partition size is correct.
//
// Materialising variants (payload = tuple index) reserve output slots per
// round with one global atomic per workgroup (block scan of the lanes'
// match counts), gather both tuples and write ResultTuple rows; they replace
// the per-match/per-thread atomics of probeWrite/probeJoin
// (GPU_HJ.cu:87,123).  `wrap` writes pos % cap (the Atomic method's ring,
// GPU_HJ.cu:92); otherwise rows past cap are dropped and counted.
#include "sjb_device.hpp"
#include "sjb_internal.hpp"

namespace sjb {

constexpr int kHjLogNB = 12;
constexpr int kHjNB = 1 << kHjLogNB;
constexpr uint32_t kNil32 = 0xFFFFFFFFu;
constexpr uint16_t kNil16 = 0xFFFFu;

struct PartRange {
    uint64_t lb, le, rb, re;
};

__device__ __forceinline__ PartRange part_range(const uint32_t *__restrict__ hist, int logF,
                                                int ncL, int ncR, uint64_t nL, uint64_t nR, int p) {
    const int F = 1 << logF;
    PartRange r;
    r.lb = hist[uint64_t(p) * ncL];
    r.le = p + 1 < F ? hist[uint64_t(p + 1) * ncL] : nL;
    const uint32_t *hr = hist + uint64_t(F) * ncL;
    r.rb = hr[uint64_t(p) * ncR];
    r.re = p + 1 < F ? hr[uint64_t(p + 1) * ncR] : nL + nR;
    return r;
}

__device__ __forceinline__ void write_row(sjb_result_tuple_t *__restrict__ out, uint64_t pos,
                                          uint64_t cap, int wrap, const sjb_tuple_t &l,
                                          const sjb_tuple_t &r) {
    if (wrap)
        pos %= cap;
    else if (pos >= cap)
        return;
    sjb_result_tuple_t row;
    row.lkey = l.key;
    row.rkey = r.key;
    row.lval = l.val;
    row.rval = r.val;
    row.lts = l.ts;
    row.rts = r.ts;
    out[pos] = row;
}

// ---------------------------------------------------------------------------
// HJ
// ---------------------------------------------------------------------------
template <bool C2, bool WRITE>
__global__ __launch_bounds__(kJoinThreads) void k_hj_join(
    const Rec *__restrict__ part, const uint32_t *__restrict__ hist, int logF, int ncL, int ncR,
    uint64_t nL, uint64_t nR, int bshift, const sjb_tuple_t *__restrict__ L,
    const sjb_tuple_t *__restrict__ R, uint64_t *__restrict__ partials,
    sjb_result_tuple_t *__restrict__ out, uint64_t out_cap, int wrap,
    unsigned long long *__restrict__ cursor) {
    __shared__ uint32_t s_key[kHjCap];
    __shared__ uint32_t s_pay[kHjCap];
    __shared__ uint16_t s_next[kHjCap];
    __shared__ uint32_t s_head[kHjNB];
    __shared__ uint64_t s_red[3 * kJoinThreads / kWave];
    __shared__ uint32_t s_ws[kJoinThreads / kWave];
    __shared__ unsigned long long s_base;

    const int p = blockIdx.x;
    const PartRange pr = part_range(hist, logF, ncL, ncR, nL, nR, p);
    Acc acc;
    acc.zero();

    for (uint64_t t0 = pr.lb; t0 < pr.le; t0 += kHjCap) {
        const int tn = int(min<uint64_t>(kHjCap, pr.le - t0));
        for (int i = threadIdx.x; i < kHjNB; i += kJoinThreads) s_head[i] = kNil32;
        __syncthreads();
        for (int i = threadIdx.x; i < tn; i += kJoinThreads) {
            const Rec r = part[t0 + i];
            s_key[i] = r.x;
            s_pay[i] = r.y;
            const uint32_t b = (key_hash(r.x) << bshift) >> (32 - kHjLogNB);
            s_next[i] = uint16_t(atomicExch(&s_head[b], uint32_t(i)));
        }
        __syncthreads();
        if (!WRITE) {
            for (uint64_t j = pr.rb + threadIdx.x; j < pr.re; j += kJoinThreads) {
                const Rec s = part[j];
                const uint32_t b = (key_hash(s.x) << bshift) >> (32 - kHjLogNB);
                for (uint32_t e = s_head[b]; e < kNil16; e = s_next[e]) {
                    if (s_key[e] == s.x) {
                        const uint64_t a = hash_a(s.x, s_pay[e]);
                        acc.count++;
                        acc.c1 += a * hash_b(s.x, s.y);
                        if (C2) acc.c2 += pair_hash_from_a(a, s.y);
                    }
                }
            }
        } else {
            // Rounds of one probe record per thread: count, reserve, write.
            for (uint64_t j0 = pr.rb; j0 < pr.re; j0 += kJoinThreads) {
                const uint64_t j = j0 + threadIdx.x;
                Rec s = make_uint2(0, 0);
                uint32_t b = 0, m = 0;
                if (j < pr.re) {
                    s = part[j];
                    b = (key_hash(s.x) << bshift) >> (32 - kHjLogNB);
                    for (uint32_t e = s_head[b]; e < kNil16; e = s_next[e]) m += s_key[e] == s.x;
                }
                uint32_t tot;
                uint32_t ex = block_excl_scan_u32(m, s_ws, &tot);
                if (threadIdx.x == 0) s_base = tot ? atomicAdd(cursor, (unsigned long long)tot) : 0ull;
                __syncthreads();
                if (m) {
                    uint64_t pos = s_base + ex;
                    const sjb_tuple_t st = R[s.y];
                    for (uint32_t e = s_head[b]; e < kNil16; e = s_next[e]) {
                        if (s_key[e] != s.x) continue;
                        const sjb_tuple_t lt = L[s_pay[e]];
                        const uint64_t a = hash_a(lt.key, lt.val);
                        acc.count++;
                        acc.c1 += a * hash_b(st.key, st.val);
                        if (C2) acc.c2 += pair_hash_from_a(a, st.val);
                        write_row(out, pos++, out_cap, wrap, lt, st);
                    }
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }
    const Acc r = block_sum(acc, s_red);
    if (threadIdx.x == 0) {
        partials[3 * p] = r.count;
        partials[3 * p + 1] = r.c1;
        partials[3 * p + 2] = r.c2;
    }
}

// ---------------------------------------------------------------------------
// SMJ
// ---------------------------------------------------------------------------
__device__ __forceinline__ void bitonic_sort(uint64_t *s, int n) {
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t a = s[i], b = s[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        s[i] = b;
                        s[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Loads n records into s as (key << 32 | payload), pads to pow2 with ~0.
// Padding entries sort last; they equal a real (0xFFFFFFFF, 0xFFFFFFFF)
// record bit for bit, so the first n sorted entries are exactly the tile.
__device__ __forceinline__ int load_sorted(uint64_t *s, const Rec *__restrict__ src, int n) {
    int np = 2;
    while (np < n) np <<= 1;
    for (int i = threadIdx.x; i < np; i += blockDim.x) {
        uint64_t v = ~0ull;
        if (i < n) {
            const Rec r = src[i];
            v = (uint64_t(r.x) << 32) | r.y;
        }
        s[i] = v;
    }
    __syncthreads();
    bitonic_sort(s, np);
    return np;
}

__device__ __forceinline__ int lower_bound_key(const uint64_t *s, int n, uint32_t k) {
    const uint64_t x = uint64_t(k) << 32;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s[mid] < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}
__device__ __forceinline__ int upper_bound_key(const uint64_t *s, int lo, int n, uint32_t k) {
    int hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (uint32_t(s[mid] >> 32) <= k)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// GROUPED: count + C1 via run closed forms (C2 not produced).
template <bool C2, bool WRITE, bool GROUPED>
__global__ __launch_bounds__(kJoinThreads) void k_smj_join(
    const Rec *__restrict__ part, const uint32_t *__restrict__ hist, int logF, int ncL, int ncR,
    uint64_t nL, uint64_t nR, const sjb_tuple_t *__restrict__ L, const sjb_tuple_t *__restrict__ R,
    uint64_t *__restrict__ partials, sjb_result_tuple_t *__restrict__ out, uint64_t out_cap,
    int wrap, unsigned long long *__restrict__ cursor) {
    __shared__ uint64_t s_b[kSmjCap];
    __shared__ uint64_t s_p[kSmjCap];
    __shared__ uint64_t s_pa[GROUPED ? kSmjCap + 1 : 1];
    __shared__ uint64_t s_red[3 * kJoinThreads / kWave];
    __shared__ uint32_t s_ws[kJoinThreads / kWave];
    __shared__ unsigned long long s_base;

    const int p = blockIdx.x;
    const PartRange pr = part_range(hist, logF, ncL, ncR, nL, nR, p);
    Acc acc;
    acc.zero();

    for (uint64_t t0 = pr.lb; t0 < pr.le; t0 += kSmjCap) {
        const int tn = int(min<uint64_t>(kSmjCap, pr.le - t0));
        load_sorted(s_b, part + t0, tn);
        if (GROUPED) {
            // prefix sums of A over the sorted build tile (serial per chunk of
            // 8 entries, then a block scan of chunk sums)
            constexpr int kPer = kSmjCap / kJoinThreads;
            uint64_t loc[kPer];
            uint64_t sum = 0;
            for (int q = 0; q < kPer; q++) {
                const int i = threadIdx.x * kPer + q;
                uint64_t a = 0;
                if (i < tn) a = hash_a(uint32_t(s_b[i] >> 32), uint32_t(s_b[i]));
                loc[q] = a;
                sum += a;
            }
            // exclusive scan of u64 sums across the block through LDS s_pa
            s_pa[threadIdx.x] = sum;
            __syncthreads();
            if (threadIdx.x == 0) {
                uint64_t run = 0;
                for (int t = 0; t < kJoinThreads; t++) {
                    const uint64_t v = s_pa[t];
                    s_pa[t] = run;
                    run += v;
                }
            }
            __syncthreads();
            uint64_t run = s_pa[threadIdx.x];
            __syncthreads();
            for (int q = 0; q < kPer; q++) {
                const int i = threadIdx.x * kPer + q;
                s_pa[i] = run;
                run += loc[q];
            }
            if (threadIdx.x == kJoinThreads - 1) s_pa[kSmjCap] = run;
            __syncthreads();
        }
        for (uint64_t u0 = pr.rb; u0 < pr.re; u0 += kSmjCap) {
            const int un = int(min<uint64_t>(kSmjCap, pr.re - u0));
            load_sorted(s_p, part + u0, un);
            if (!WRITE) {
                for (int j = threadIdx.x; j < un; j += kJoinThreads) {
                    const uint32_t k = uint32_t(s_p[j] >> 32), sv = uint32_t(s_p[j]);
                    const int lo = lower_bound_key(s_b, tn, k);
                    const int hi = upper_bound_key(s_b, lo, tn, k);
                    if (hi == lo) continue;
                    if (GROUPED) {
                        acc.count += uint64_t(hi - lo);
                        acc.c1 += (s_pa[hi] - s_pa[lo]) * hash_b(k, sv);
                    } else {
                        for (int e = lo; e < hi; e++) {
                            const uint64_t a = hash_a(k, uint32_t(s_b[e]));
                            acc.count++;
                            acc.c1 += a * hash_b(k, sv);
                            if (C2) acc.c2 += pair_hash_from_a(a, sv);
                        }
                    }
                }
            } else {
                for (int j0 = 0; j0 < un; j0 += kJoinThreads) {
                    const int j = j0 + threadIdx.x;
                    int lo = 0, hi = 0;
                    uint32_t k = 0, si = 0;
                    if (j < un) {
                        k = uint32_t(s_p[j] >> 32);
                        si = uint32_t(s_p[j]);
                        lo = lower_bound_key(s_b, tn, k);
                        hi = upper_bound_key(s_b, lo, tn, k);
                    }
                    uint32_t tot;
                    const uint32_t ex = block_excl_scan_u32(uint32_t(hi - lo), s_ws, &tot);
                    if (threadIdx.x == 0)
                        s_base = tot ? atomicAdd(cursor, (unsigned long long)tot) : 0ull;
                    __syncthreads();
                    if (hi > lo) {
                        uint64_t pos = s_base + ex;
                        const sjb_tuple_t st = R[si];
                        for (int e = lo; e < hi; e++) {
                            const sjb_tuple_t lt = L[uint32_t(s_b[e])];
                            const uint64_t a = hash_a(lt.key, lt.val);
                            acc.count++;
                            acc.c1 += a * hash_b(st.key, st.val);
                            if (C2) acc.c2 += pair_hash_from_a(a, st.val);
                            write_row(out, pos++, out_cap, wrap, lt, st);
                        }
                    }
                    __syncthreads();
                }
            }
            __syncthreads();
        }
    }
    const Acc r = block_sum(acc, s_red);
    if (threadIdx.x == 0) {
        partials[3 * p] = r.count;
        partials[3 * p + 1] = r.c1;
        partials[3 * p + 2] = r.c2;
    }
}

// ---------------------------------------------------------------------------
// partial reduction -> one sjb_window_result_t
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_reduce(const uint64_t *__restrict__ partials, int n,
                                                sjb_window_result_t *__restrict__ out) {
    __shared__ uint64_t s_red[3 * 256 / kWave];
    Acc a;
    a.zero();
    for (int i = threadIdx.x; i < n; i += 256) {
        a.count += partials[3 * i];
        a.c1 += partials[3 * i + 1];
        a.c2 += partials[3 * i + 2];
    }
    const Acc r = block_sum(a, s_red);
    if (threadIdx.x == 0) {
        out->count = r.count;
        out->checksum1 = r.c1;
        out->checksum2 = r.c2;
        out->pairs = 0;
    }
}

int launch_reduce(sjb_ctx *c, int n_partials, sjb_window_result_t *d_out) {
    k_reduce<<<1, 256, 0, c->stream>>>(c->d_partials, n_partials, d_out);
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

int launch_hj_join(sjb_ctx *c, const WindowPlan &p, const sjb_tuple_t *L, const sjb_tuple_t *R,
                   sjb_result_tuple_t *out, uint64_t out_cap, int wrap) {
    const int F = 1 << p.logF;
    const bool c2 = (p.flags & SJB_F_C2) != 0;
    const int bshift = p.kind == PART_HASH ? p.logF : 0;
    auto *cur = reinterpret_cast<unsigned long long *>(c->d_cursor);
    if (p.payload == PAY_IDX) {
        if (c2)
            k_hj_join<true, true><<<F, kJoinThreads, 0, c->stream>>>(
                c->d_part, c->d_hist, p.logF, p.ncL, p.ncR, p.nL, p.nR, bshift, L, R, c->d_partials,
                out, out_cap, wrap, cur);
        else
            k_hj_join<false, true><<<F, kJoinThreads, 0, c->stream>>>(
                c->d_part, c->d_hist, p.logF, p.ncL, p.ncR, p.nL, p.nR, bshift, L, R, c->d_partials,
                out, out_cap, wrap, cur);
    } else {
        if (c2)
            k_hj_join<true, false><<<F, kJoinThreads, 0, c->stream>>>(
                c->d_part, c->d_hist, p.logF, p.ncL, p.ncR, p.nL, p.nR, bshift, L, R, c->d_partials,
                out, out_cap, wrap, cur);
        else
            k_hj_join<false, false><<<F, kJoinThreads, 0, c->stream>>>(
                c->d_part, c->d_hist, p.logF, p.ncL, p.ncR, p.nL, p.nR, bshift, L, R, c->d_partials,
                out, out_cap, wrap, cur);
    }
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

int launch_smj_join(sjb_ctx *c, const WindowPlan &p, const sjb_tuple_t *L, const sjb_tuple_t *R,
                    sjb_result_tuple_t *out, uint64_t out_cap, int wrap) {
    const int F = 1 << p.logF;
    const bool c2 = (p.flags & SJB_F_C2) != 0;
    const bool grouped = (p.flags & SJB_F_GROUPED) != 0 && p.payload == PAY_VAL;
    auto *cur = reinterpret_cast<unsigned long long *>(c->d_cursor);
#define SJB_SMJ(C2_, W_, G_)                                                                      \
    k_smj_join<C2_, W_, G_><<<F, kJoinThreads, 0, c->stream>>>(c->d_part, c->d_hist, p.logF,     \
                                                               p.ncL, p.ncR, p.nL, p.nR, L, R,   \
                                                               c->d_partials, out, out_cap, wrap, \
                                                               cur)
    if (p.payload == PAY_IDX) {
        if (c2)
            SJB_SMJ(true, true, false);
        else
            SJB_SMJ(false, true, false);
    } else if (grouped) {
        SJB_SMJ(false, false, true);
    } else if (c2) {
        SJB_SMJ(true, false, false);
    } else {
        SJB_SMJ(false, false, false);
    }
#undef SJB_SMJ
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

}  // namespace sjb
