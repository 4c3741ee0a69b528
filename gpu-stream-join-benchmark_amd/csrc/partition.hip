// partition.hip — bounded-fanout partitioning of both window sides.
//
// Replaces the reference's D-bucket counting sort (computeHistogram +
// host prefix sum + buildHashTable, src/sjb/windowing/lazy/executor/GPU/
// GPU_HJ.cu:9-39 and GPU_HJ_NoOutput.cu:53-76; smjHistogramKernel +
// thrust::exclusive_scan + smjSortKernel, GPU_SMJ.cu:9-32) with a fanout
// F = 2^logF that is independent of the key range, in ONE pass over the
// input (k_partition below): each 16384-tuple chunk is sorted by partition
// in LDS and written back contiguously as 8-byte (key, payload) records,
// together with its F+1 local partition offsets.  No global histogram, no
// scan, no host round trip.  Both sides go through one launch: blocks
// [0, ncL) take left chunks and [ncL, ncL+ncR) right chunks; left records
// land in d_part[0, nL), right records in d_part[nL, nL+nR).
#include <algorithm>

#include "sjb_device.hpp"
#include "sjb_internal.hpp"

namespace sjb {

// Partition function of one key.  HASH: top logF bits of the Fibonacci hash
// (bijective, so equal keys and only equal keys collide before truncation).
// RANGE: (key - kmin) >> shift, i.e. MSD radix on the key digits above the
// window's minimum.  LEARNED: floor(F * cdf(key)) from the RMI.
// Learned partition of key k into F parts (RMI of RmiParams; fminf/fmaxf
// also map a NaN prediction to 0).
__device__ __forceinline__ uint32_t rmi_partition(const RmiParams *__restrict__ m, uint32_t k, uint32_t F) {
    const float x = k < m->kmin ? 0.f : float(k - m->kmin) * m->scale;
    float r;
    if (m->root_kind == ROOT_LINEAR) {
        r = fmaf(m->root_a, x, m->root_b);
    } else {
        const float u = fminf(fmaxf(x, 0.f), 1.f) * float(kRootTable);
        const int i = min(int(u), kRootTable - 1);
        r = fmaf(m->table[i + 1] - m->table[i], u - float(i), m->table[i]);
    }
    r = fminf(fmaxf(r * float(kRmiLeaves), 0.f), float(kRmiLeaves - 1));
    const int leaf = int(r);
    float c = fmaf(m->leaf_a[leaf], x, m->leaf_b[leaf]);
    c = fminf(fmaxf(c * float(F), 0.f), float(F - 1));
    return uint32_t(c);
}

// Heavy-key set of a learned model in LDS (the RmiParams lookup table:
// 64-slot open addressing, slot -> key, index in RmiParams::heavy, -1 =
// empty, plus a 4096-bit filter on 12 hash bits, so most keys, not heavy,
// are rejected with one LDS read instead of a probe sequence).
// Each wave also has a queue of its heavy hits (key, val, heavy index):
// hits are hashed and summed when 64 are queued, with every lane busy.
struct HeavyQ {
    uint32_t key[kWave];
    uint32_t val[kWave];
    int h[kWave];
};
struct HeavySet {
    uint32_t key[64];
    int idx[64];
    uint32_t filter[128];
    uint32_t cnt[kMaxHeavy];
    uint64_t sum[kMaxHeavy];
    HeavyQ q[kPartThreads / kWave];
};

#ifndef SJB_HEAVY_REG
#define SJB_HEAVY_REG 1
#endif
constexpr int kHeavyReg = SJB_HEAVY_REG;  // most frequent heavy keys summed in registers
constexpr int kHeavyRegArr = kHeavyReg > 0 ? kHeavyReg : 1;

// Hashes and sums the n (<= 64) queued heavy hits of the calling wave, one
// per lane: the kHeavyReg most frequent keys into the lane's registers, the
// others with LDS atomics.  The queue is free again afterwards.
__device__ __forceinline__ void heavy_drain(HeavySet &hs, const HeavyQ &q, uint32_t n, bool left,
                                            uint32_t (&hcnt)[kHeavyRegArr], uint64_t (&hsum)[kHeavyRegArr]) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if (lane < n) {
        const uint32_t k = q.key[lane], v = q.val[lane];
        const int h = q.h[lane];
        const uint64_t x = left ? hash_a(k, v) : hash_b(k, v);
        if (h < kHeavyReg) {
#pragma unroll
            for (int i = 0; i < kHeavyReg; i++) {
                hcnt[i] += h == i;
                hsum[i] += h == i ? x : 0ull;
            }
        } else {
            atomicAdd(&hs.cnt[h], 1u);
            atomicAdd(reinterpret_cast<unsigned long long *>(&hs.sum[h]), (unsigned long long)x);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

// The model's precomputed lookup table, one word (or two) per thread: the
// global loads are issued first (heavy_fetch), the chunk's own loads next,
// and the LDS copy (heavy_store, which also clears the counters) waits
// only for the table; the caller puts a barrier after heavy_store.
struct HeavyFetch {
    uint32_t key, filter;
    int idx;
};
__device__ __forceinline__ HeavyFetch heavy_fetch(const RmiParams *__restrict__ m) {
    const uint32_t t = threadIdx.x;
    HeavyFetch f{0u, 0u, -1};
    if (t < 64) {
        f.key = m->hslot_key[t];
        f.idx = m->hslot_idx[t];
    }
    if (t < 128) f.filter = m->hfilter[t];
    return f;
}
__device__ __forceinline__ void heavy_store(HeavySet &h, const HeavyFetch &f) {
    const uint32_t t = threadIdx.x;
    if (t < 64) {
        h.key[t] = f.key;
        h.idx[t] = f.idx;
    }
    if (t < 128) h.filter[t] = f.filter;
    if (t < kMaxHeavy) {
        h.cnt[t] = 0;
        h.sum[t] = 0;
    }
}

__device__ __forceinline__ int heavy_find(const HeavySet &h, uint32_t k) {
    const uint32_t x = k * kHashMul;
    if (!((h.filter[x >> 25] >> ((x >> 20) & 31)) & 1u)) return -1;
    uint32_t s = x >> 26;
    for (;;) {
        const int i = h.idx[s];
        if (i < 0) return -1;
        if (h.key[s] == k) return i;
        s = (s + 1) & 63;
    }
}

template <int KIND>
struct PartFn {
    uint32_t logF, fmask, kmin, shift;
    const RmiParams *rmi;
    // kmin / shift: range partitioning parameters (ignored by HASH/LEARNED)
    __device__ __forceinline__ void init(int logF_, uint32_t kmin_, uint32_t shift_,
                                         const RmiParams *rmi_) {
        logF = logF_;
        fmask = (1u << logF_) - 1u;
        rmi = rmi_;
        if (KIND == PART_RANGE) {
            kmin = kmin_;
            shift = shift_;
        } else {
            kmin = 0;
            shift = 32 - logF_;
        }
    }
    __device__ __forceinline__ uint32_t operator()(uint32_t k) const {
        if (KIND == PART_HASH) return logF ? key_hash(k) >> shift : 0u;
        if (KIND == PART_RANGE) {
            // keys outside [kmin, ...] clamp to the end partitions: the
            // function stays a function of the key, so equal keys meet
            if (k < kmin) return 0u;
            const uint32_t p = (k - kmin) >> shift;
            return p > fmask ? fmask : p;
        }
        // LEARNED
        return rmi_partition(rmi, k, fmask + 1);
    }
};

// ---------------------------------------------------------------------------
// min / max key over both sides (range partitioning parameters)
// ---------------------------------------------------------------------------
// Key bounds for range partitioning from an evenly spaced sample of at most
// kMinmaxSample keys of the window (every key when the window is that
// small).  Keys outside the sampled bounds are clamped to the first / last
// partition by the range function, which stays a function of the key, so
// equal keys still meet; the sample only sets the balance.  The sample is
// spread over many workgroups (one key per thread: the loads of one
// workgroup alone would queue behind each other), each writing its partial
// bounds; the range partition pass folds them (range_params).
constexpr int kMinmaxThreads = 1024;
constexpr uint32_t kMinmaxSample = 1u << 16;
__global__ __launch_bounds__(kMinmaxThreads) void k_minmax_partial(const sjb_tuple_t *__restrict__ L, uint64_t nL,
                                                                   const sjb_tuple_t *__restrict__ R, uint32_t ns,
                                                                   uint64_t step, uint32_t frac,
                                                                   uint32_t *__restrict__ partial) {
    __shared__ uint32_t smin[kMinmaxThreads / kWave], smax[kMinmaxThreads / kWave];
    uint32_t mn = 0xFFFFFFFFu, mx = 0;
    for (uint32_t q = blockIdx.x * kMinmaxThreads + threadIdx.x; q < ns; q += gridDim.x * kMinmaxThreads) {
        // sample q sits at floor(q * n / ns) = q * step + floor(q * frac / 2^32)
        const uint64_t i = uint64_t(q) * step + ((uint64_t(q) * frac) >> 32);
        const uint32_t k = i < nL ? L[i].key : R[i - nL].key;
        mn = min(mn, k);
        mx = max(mx, k);
    }
    wave_minmax_full(mn, mx);
    const int wid = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        smin[wid] = mn;
        smax[wid] = mx;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w = 1; w < kMinmaxThreads / kWave; w++) {
        mn = min(mn, smin[w]);
        mx = max(mx, smax[w]);
    }
    partial[2 * blockIdx.x] = mn;
    partial[2 * blockIdx.x + 1] = mx;
}

// Range-partition parameters (kmin, shift) from the nblocks (min, max)
// pairs of k_minmax_partial: every partition workgroup folds them itself —
// its first wave, DPP min/max — so no reduction kernel runs between the
// sample and the partition pass.  Every thread of the block (one barrier).
__device__ __forceinline__ void range_params(const uint32_t *__restrict__ partial, int nblocks, int logF,
                                             uint32_t *kmin, uint32_t *shift) {
    __shared__ uint32_t s_rp[2];
    if (threadIdx.x < kWave) {  // the whole first wave
        uint32_t mn = 0xFFFFFFFFu, mx = 0;
        for (int b = threadIdx.x; b < nblocks; b += kWave) {
            mn = min(mn, partial[2 * b]);
            mx = max(mx, partial[2 * b + 1]);
        }
        wave_minmax_full(mn, mx);
        if (threadIdx.x == 0) {
            if (mn > mx) mn = mx = 0;
            const uint32_t span = mx - mn;
            const int bits = span ? 32 - __clz(span) : 0;
            s_rp[0] = mn;
            s_rp[1] = bits > logF ? uint32_t(bits - logF) : 0u;
        }
    }
    __syncthreads();
    *kmin = s_rp[0];
    *shift = s_rp[1];
}

// Samples the window's keys (k_minmax_partial) into c->d_scratch; returns
// the number of (min, max) pairs the range partition pass folds.
int launch_minmax(sjb_ctx *c, hipStream_t s, const sjb_tuple_t *L, uint64_t nL, const sjb_tuple_t *R,
                  uint64_t nR, int *blocks_out) {
    const uint64_t n = nL + nR;
    const uint32_t ns = uint32_t(std::min<uint64_t>(n, kMinmaxSample));
    const int blocks = std::max(1, int((ns + kMinmaxThreads - 1) / kMinmaxThreads));
    const uint64_t step = ns ? n / ns : 0;
    const uint32_t frac = ns ? uint32_t(((n % ns) << 32) / ns) : 0u;
    // The scan-block buffer doubles as scratch for the partial bounds.
    k_minmax_partial<<<blocks, kMinmaxThreads, 0, s>>>(L, nL, R, ns, step, frac, c->d_scratch);
    SJB_HIP_TRY(hipGetLastError());
    *blocks_out = blocks;
    return SJB_OK;
}

// ---------------------------------------------------------------------------
// single-pass chunk-major partitioning
// ---------------------------------------------------------------------------
// One workgroup of kPartThreads threads owns one chunk of up to CHUNK
// tuples of one side.  It loads its tuples (CHUNK/kPartThreads per thread,
// coalesced), ranks them in an LDS histogram over F partitions, scans the
// histogram, places the 8-byte records partition-sorted into an LDS staging
// buffer and writes the whole chunk back contiguously — every global write
// is a full, coalesced line.  The chunk's local partition offsets (F+1
// entries) go to `o`.  The join gathers partition p as the run
// [o_c[p], o_c[p+1]) of every chunk c.
//
// PEEL (learned model, grouped mode): records whose key is in the model's
// heavy set are neither ranked nor written back (o[F] ends partition F-1)
// but aggregated: per heavy key the count and the sum of A (left side) or
// B (right side) are added to heavy_acc[4*h + (left ? 0 : 2) + {0,1}].
// Hits are queued per wave and hashed 64 at a time (heavy_drain); heavy
// keys are ordered by sample frequency and the kHeavyReg most frequent are
// summed in per-thread registers, reduced once per wave (under skew the
// top key is a large share of the hits, and LDS atomics on one address
// serialise); the others go to LDS atomics.  Measured on Zipf 1.2: one
// register key is best — each further one costs a select per hit and two
// wave reductions per chunk, more than the atomics it saves.
template <int KIND, int PAY, int CHUNK, bool PEEL>
__device__ __forceinline__ void partition_chunk(const sjb_tuple_t *__restrict__ in, uint64_t base,
                                                int len, int logF, const PartFn<KIND> &part,
                                                uint32_t *__restrict__ o, Rec *__restrict__ out,
                                                uint32_t *lds, HeavySet *hs, const RmiParams *rmi,
                                                bool left, uint64_t *__restrict__ heavy_acc) {
    const int F = 1 << logF;
    const int NB = F + (PEEL ? 1 : 0);                     // buckets
    uint32_t *hist = lds;                                  // NB + 1 counters
    uint32_t *ws = lds + F + 2;                            // wave sums
    Rec *stage = reinterpret_cast<Rec *>(lds + ((F + 2 + kPartThreads / kWave + 1) & ~1));

    for (int q = threadIdx.x; q <= NB; q += kPartThreads) hist[q] = 0;
    HeavyFetch hf{0u, 0u, -1};
    if (PEEL) hf = heavy_fetch(rmi);
    constexpr int kPer = CHUNK / kPartThreads;
    Rec rec[kPer];
    uint32_t pid[kPer], rank[kPer];
    uint32_t hcnt[kHeavyRegArr];
    uint64_t hsum[kHeavyRegArr];
#pragma unroll
    for (int i = 0; i < kHeavyReg; i++) {
        hcnt[i] = 0;
        hsum[i] = 0;
    }
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int e = j * kPartThreads + threadIdx.x;
        if (e < len) {
            const uint2 kv = *reinterpret_cast<const uint2 *>(&in[base + e]);
            rec[j] = make_uint2(kv.x, PAY == PAY_VAL ? kv.y : uint32_t(base + e));
        }
    }
    if (PEEL) heavy_store(*hs, hf);
    __syncthreads();  // zeroed histogram (and the heavy-key table) visible
    uint32_t qn = 0;  // queued heavy hits of this wave (wave-uniform)
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int e = j * kPartThreads + threadIdx.x;
        if (PEEL) {
            const int h = e < len ? heavy_find(*hs, rec[j].x) : -1;
            const bool hv = h >= 0;
            const uint64_t mask = __ballot(hv);
            if (mask) {  // wave-uniform
                HeavyQ &q = hs->q[threadIdx.x / kWave];
                const uint32_t c = uint32_t(__popcll(mask));
                if (qn + c > uint32_t(kWave)) {
                    heavy_drain(*hs, q, qn, left, hcnt, hsum);
                    qn = 0;
                }
                if (hv) {
                    const uint32_t pos = qn + lanes_below(mask);
                    q.key[pos] = rec[j].x;
                    q.val[pos] = rec[j].y;
                    q.h[pos] = h;
                }
                qn += c;
            }
            if (e < len) {
                pid[j] = hv ? uint32_t(F) : part(rec[j].x);
                if (!hv) rank[j] = atomicAdd(&hist[pid[j]], 1u);
            }
        } else if (e < len) {
            pid[j] = part(rec[j].x);
            rank[j] = atomicAdd(&hist[pid[j]], 1u);
        }
    }
    if (PEEL) {
        heavy_drain(*hs, hs->q[threadIdx.x / kWave], qn, left, hcnt, hsum);
#pragma unroll
        for (int i = 0; i < kHeavyReg; i++) {
            const uint64_t c = wave_sum_full(uint64_t(hcnt[i]));
            const uint64_t v = wave_sum_full(hsum[i]);
            if ((threadIdx.x & (kWave - 1)) == 0 && c) {
                atomicAdd(&hs->cnt[i], uint32_t(c));
                atomicAdd(reinterpret_cast<unsigned long long *>(&hs->sum[i]), (unsigned long long)v);
            }
        }
    }
    __syncthreads();
    // exclusive scan of hist[0..NB) in place; hist[NB] = records ranked
    {
        constexpr int kMaxPer = ((1 << kMaxLogF) + 1 + kPartThreads - 1) / kPartThreads;
        uint32_t v[kMaxPer];
        uint32_t sum = 0;
        const int per = (NB + kPartThreads - 1) / kPartThreads;
        for (int q = 0; q < per && q < kMaxPer; q++) {
            const int idx = threadIdx.x * per + q;
            v[q] = idx < NB ? hist[idx] : 0;
            sum += v[q];
        }
        uint32_t tot;
        uint32_t run = block_excl_scan_u32(sum, ws, &tot);
        for (int q = 0; q < per && q < kMaxPer; q++) {
            const int idx = threadIdx.x * per + q;
            if (idx < NB) hist[idx] = run;
            run += v[q];
        }
        if (threadIdx.x == 0) hist[NB] = tot;
    }
    __syncthreads();
    for (int q = threadIdx.x; q <= F; q += kPartThreads) o[q] = hist[q];
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const int e = j * kPartThreads + threadIdx.x;
        if (e < len && (!PEEL || pid[j] < uint32_t(F))) stage[hist[pid[j]] + rank[j]] = rec[j];
    }
    __syncthreads();
    const int keep = PEEL ? int(hist[F]) : len;
    for (int e = threadIdx.x; e < keep; e += kPartThreads) out[e] = stage[e];
    if (PEEL && threadIdx.x < kMaxHeavy && hs->cnt[threadIdx.x]) {
        uint64_t *a = heavy_acc + 4 * threadIdx.x + (left ? 0 : 2);
        atomicAdd(reinterpret_cast<unsigned long long *>(a), (unsigned long long)hs->cnt[threadIdx.x]);
        atomicAdd(reinterpret_cast<unsigned long long *>(a + 1), (unsigned long long)hs->sum[threadIdx.x]);
    }
}

// One window: blocks [0, ncL) take left chunks of CH tuples and
// [ncL, ncL+ncR) right chunks; chunk b's offsets row is offs[b * (F+1)].
// Block 0 zeroes the window result the join adds into (direct non-null).
template <int KIND, int PAY, bool PEEL, int CH>
__global__ __launch_bounds__(kPartThreads) void k_partition(
    const sjb_tuple_t *__restrict__ L, uint64_t nL, int ncL, const sjb_tuple_t *__restrict__ R,
    uint64_t nR, int logF, const uint32_t *__restrict__ range_partial, int range_blocks,
    const RmiParams *__restrict__ rmi, uint32_t *__restrict__ offs, Rec *__restrict__ part_out,
    uint64_t *__restrict__ heavy_acc, sjb_window_result_t *__restrict__ direct) {
    extern __shared__ uint32_t lds[];
    __shared__ HeavySet hs[1];
    if (direct && blockIdx.x == 0 && threadIdx.x < sizeof(sjb_window_result_t) / 8)
        reinterpret_cast<uint64_t *>(direct)[threadIdx.x] = 0;
    uint32_t kmin = 0, shift = 0;
    if constexpr (KIND == PART_RANGE) range_params(range_partial, range_blocks, logF, &kmin, &shift);
    PartFn<KIND> part;
    part.init(logF, kmin, shift, rmi);
    const bool left = int(blockIdx.x) < ncL;
    const int chunk = left ? blockIdx.x : blockIdx.x - ncL;
    const uint64_t n = left ? nL : nR;
    const uint64_t base = uint64_t(chunk) * CH;
    const int len = int(min<uint64_t>(CH, n - base));
    partition_chunk<KIND, PAY, CH, PEEL>(left ? L : R, base, len, logF, part,
                                             offs + uint64_t(blockIdx.x) * ((1 << logF) + 1),
                                             part_out + (left ? 0 : nL) + base, lds, hs, rmi, left,
                                             heavy_acc);
}

// ---------------------------------------------------------------------------
// persistent, software-pipelined variant of k_partition (no heavy peeling)
// ---------------------------------------------------------------------------
// The staging buffer holds a whole chunk (128 KiB), so only one partition
// workgroup fits on a CU and, one chunk per workgroup, a CU would sit idle on
// HBM while it ranks, scans and stages.  Here each workgroup walks the chunks
// b = blockIdx.x, blockIdx.x + gridDim.x, ... and issues the loads of chunk
// b + gridDim.x as soon as chunk b's keys are ranked: they are in flight
// while chunk b is scanned, staged and written back.  Same records, offsets
// and layout as k_partition.
constexpr int kPartPer = kChunk / kPartThreads;  // tuples per thread per chunk

struct ChunkGeo {
    const sjb_tuple_t *in;
    uint64_t base;  // first tuple of the chunk within its side
    uint64_t obase; // first record of the chunk in the partition output
    int len;
};

__device__ __forceinline__ ChunkGeo chunk_geo(int b, const sjb_tuple_t *L, uint64_t nL, int ncL,
                                              const sjb_tuple_t *R, uint64_t nR) {
    ChunkGeo g;
    const bool left = b < ncL;
    const int c = left ? b : b - ncL;
    const uint64_t n = left ? nL : nR;
    g.in = left ? L : R;
    g.base = uint64_t(c) * kChunk;
    g.obase = (left ? 0 : nL) + g.base;
    g.len = int(min<uint64_t>(kChunk, n - g.base));
    return g;
}

// Loads are clamped to the chunk's last tuple instead of guarded, so they
// are issued back to back without branches; lanes past len are ignored.
// Addresses are a uniform base plus a 32-bit lane offset (one VGPR each).
// Loads records [J0, J1) of every thread.
template <int PAY, int J0 = 0, int J1 = kPartPer>
__device__ __forceinline__ void chunk_load(const ChunkGeo &g, Rec (&rec)[kPartPer]) {
    const char *__restrict__ in = reinterpret_cast<const char *>(g.in + g.base);
    const uint32_t last = uint32_t(g.len - 1);
#pragma unroll
    for (int j = J0; j < J1; j++) {
        const uint32_t e = min(uint32_t(j * kPartThreads) + threadIdx.x, last);
        const uint2 kv = *reinterpret_cast<const uint2 *>(in + e * uint32_t(sizeof(sjb_tuple_t)));
        rec[j] = make_uint2(kv.x, PAY == PAY_VAL ? kv.y : uint32_t(g.base) + e);
    }
}

// rank[j/2] holds the ranks of records j (low half) and j+1 (high half)
// within their partitions (< CHUNK = 2^14); the partition is recomputed
// when the record is staged.  The partition function runs on every
// (clamped, valid) record outside the len guard, so every record register
// is consumed unconditionally here.
template <int KIND>
__device__ __forceinline__ void chunk_rank(const PartFn<KIND> &part, const Rec (&rec)[kPartPer], int len,
                                           uint32_t *hist, uint32_t (&rank)[kPartPer / 2]) {
#pragma unroll
    for (int j = 0; j < kPartPer; j++) {
        const int e = j * kPartThreads + threadIdx.x;
        const uint32_t p = part(rec[j].x);
        uint32_t r = 0;
        if (e < len) r = atomicAdd(&hist[p], 1u);
        if (j & 1)
            rank[j / 2] |= r << 16;
        else
            rank[j / 2] = r;
    }
}

template <int KIND>
__device__ __forceinline__ void chunk_stage(const PartFn<KIND> &part, const Rec (&rec)[kPartPer],
                                            const uint32_t (&rank)[kPartPer / 2], int len,
                                            const uint32_t *hist, Rec *stage) {
#pragma unroll
    for (int j = 0; j < kPartPer; j++) {
        const int e = j * kPartThreads + threadIdx.x;
        const uint32_t r = (j & 1) ? rank[j / 2] >> 16 : rank[j / 2] & 0xFFFFu;
        if (e < len) stage[hist[part(rec[j].x)] + r] = rec[j];
    }
}

// Exclusive scan of hist[0, F) in place, hist[F] = total; every thread of
// the block calls it; ends with a barrier.
__device__ __forceinline__ void hist_scan(uint32_t *hist, uint32_t *ws, int F) {
    constexpr int kMaxPer = ((1 << kMaxLogF) + kPartThreads - 1) / kPartThreads;
    uint32_t v[kMaxPer];
    uint32_t sum = 0;
    const int per = (F + kPartThreads - 1) / kPartThreads;
    for (int q = 0; q < per && q < kMaxPer; q++) {
        const int idx = threadIdx.x * per + q;
        v[q] = idx < F ? hist[idx] : 0;
        sum += v[q];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan_u32(sum, ws, &tot);
    for (int q = 0; q < per && q < kMaxPer; q++) {
        const int idx = threadIdx.x * per + q;
        if (idx < F) hist[idx] = run;
        run += v[q];
    }
    if (threadIdx.x == 0) hist[F] = tot;
    __syncthreads();
}

static_assert(kChunk <= (1 << 16), "chunk_rank packs ranks into 16 bits");

// One chunk b (records in `cur`, geometry g): rank, scan, stage, write the
// offsets row and the records.  The next chunk's loads (gn, into `nxt`),
// when there is one, are issued in two halves: the first right after the
// ranking, the second once `cur` is staged (its registers are then free),
// so at most one and a half chunks occupy registers.  Starts and ends with
// the LDS free.
template <int KIND, int PAY>
__device__ __forceinline__ void chunk_step(const PartFn<KIND> &part, int F, uint32_t *hist, uint32_t *ws,
                                           Rec *stage, const Rec (&cur)[kPartPer], const ChunkGeo &g, int b,
                                           Rec (&nxt)[kPartPer], const ChunkGeo &gn, bool has_next,
                                           uint32_t *__restrict__ offs, Rec *__restrict__ part_out) {
    for (int q = threadIdx.x; q <= F; q += kPartThreads) hist[q] = 0;
    __syncthreads();
    uint32_t rank[kPartPer / 2];
    chunk_rank<KIND>(part, cur, g.len, hist, rank);
    if (has_next) chunk_load<PAY, 0, kPartPer / 2>(gn, nxt);
    __syncthreads();
    hist_scan(hist, ws, F);
    chunk_stage<KIND>(part, cur, rank, g.len, hist, stage);
    if (has_next) chunk_load<PAY, kPartPer / 2, kPartPer>(gn, nxt);
    // unsigned 32-bit lane offsets from uniform bases: the stores address
    // through a scalar base, with no per-lane 64-bit pointers kept live
    // The lane index is re-read here (the empty asm hides its value) so the
    // lane addresses below are formed per step rather than hoisted out of
    // the chunk loop and kept live (or spilled) across it.
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    uint32_t *o = offs + uint32_t(b) * uint32_t(F + 1);  // < 2^23 words
    for (uint32_t q = tid; q <= uint32_t(F); q += kPartThreads) o[q] = hist[q];
    __syncthreads();
    Rec *out = part_out + g.obase;
    for (uint32_t e = tid; e < uint32_t(g.len); e += kPartThreads) out[e] = stage[e];
    __syncthreads();
}

// Two register sets alternate (no copies between them), so waiting for one
// chunk's loads never waits for the previous chunk's stores.
template <int KIND, int PAY>
__global__ __launch_bounds__(kPartThreads) void k_partition_pipe(
    const sjb_tuple_t *__restrict__ L, uint64_t nL, int ncL, const sjb_tuple_t *__restrict__ R,
    uint64_t nR, int ncR, int logF, const uint32_t *__restrict__ range_partial, int range_blocks,
    const RmiParams *__restrict__ rmi, uint32_t *__restrict__ offs, Rec *__restrict__ part_out,
    sjb_window_result_t *__restrict__ direct) {
    extern __shared__ uint32_t lds[];
    const int F = 1 << logF;
    // the join kernel accumulates into `direct`; it starts from zero here
    if (direct && blockIdx.x == 0 && threadIdx.x < 4)
        reinterpret_cast<uint64_t *>(direct)[threadIdx.x] = 0;
    uint32_t *hist = lds;       // F + 1 counters
    uint32_t *ws = lds + F + 2; // wave sums
    Rec *stage = reinterpret_cast<Rec *>(lds + ((F + 2 + kPartThreads / kWave + 1) & ~1));
    const int nch = ncL + ncR;
    const int G = int(gridDim.x);
    int b = blockIdx.x;
    if (b >= nch) return;  // the whole workgroup
    ChunkGeo ga = chunk_geo(b, L, nL, ncL, R, nR);
    Rec ra[kPartPer], rb[kPartPer];
    chunk_load<PAY>(ga, ra);
    // range parameters folded while the first chunk's loads are in flight
    uint32_t kmin = 0, shift = 0;
    if constexpr (KIND == PART_RANGE) range_params(range_partial, range_blocks, logF, &kmin, &shift);
    PartFn<KIND> part;
    part.init(logF, kmin, shift, rmi);
    for (;;) {
        const int b1 = b + G;
        const bool h1 = b1 < nch;
        const ChunkGeo gb = h1 ? chunk_geo(b1, L, nL, ncL, R, nR) : ga;
        chunk_step<KIND, PAY>(part, F, hist, ws, stage, ra, ga, b, rb, gb, h1, offs, part_out);
        if (!h1) break;
        const int b2 = b1 + G;
        const bool h2 = b2 < nch;
        const ChunkGeo gc = h2 ? chunk_geo(b2, L, nL, ncL, R, nR) : gb;
        chunk_step<KIND, PAY>(part, F, hist, ws, stage, rb, gb, b1, ra, gc, h2, offs, part_out);
        if (!h2) break;
        b = b2;
        ga = gc;
    }
}

// Range parameters of a batched window from a fixed strided sample of up
// to 256 keys per side.  Every chunk workgroup of the window draws the same
// sample, so all of them use the same partition function; keys outside the
// sampled range clamp to the end partitions (exact, only the balance
// depends on the sample).
__device__ __forceinline__ void sample_range(const sjb_tuple_t *__restrict__ L, uint32_t nL,
                                             const sjb_tuple_t *__restrict__ R, uint32_t nR,
                                             int logF, uint32_t *s_mm, uint32_t *kmin,
                                             uint32_t *shift) {
    constexpr int kS = 256;
    if (threadIdx.x < 2) s_mm[threadIdx.x] = threadIdx.x ? 0u : 0xFFFFFFFFu;
    __syncthreads();
    uint32_t mn = 0xFFFFFFFFu, mx = 0;
    const int t = threadIdx.x;
    if (t < 2 * kS) {
        const bool left = t < kS;
        const uint32_t n = left ? nL : nR;
        const uint32_t j = left ? t : t - kS;
        const uint32_t ns = n < uint32_t(kS) ? n : uint32_t(kS);
        if (j < ns) {
            const uint32_t idx = uint32_t((uint64_t(j) * n) / ns);
            const uint32_t k = (left ? L : R)[idx].key;
            mn = k;
            mx = k;
        }
    }
    wave_minmax_full(mn, mx);
    if ((threadIdx.x & (kWave - 1)) == 0 && t < 2 * kS) {
        atomicMin(&s_mm[0], mn);
        atomicMax(&s_mm[1], mx);
    }
    __syncthreads();
    mn = s_mm[0];
    mx = s_mm[1];
    if (mn > mx) mn = mx = 0;
    const uint32_t span = mx - mn;
    const int bits = span ? 32 - __clz(span) : 0;
    *kmin = mn;
    *shift = bits > logF ? uint32_t(bits - logF) : 0u;
}

// Batched windows: block b is partition-pass item b of the batch (one chunk
// of one side of one window).
template <int KIND, int CHUNK, bool PEEL>
__global__ __launch_bounds__(kPartThreads) void k_partition_batch(BatchArgs a) {
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_mm[2];
    __shared__ HeavySet hs[1];
    const uint32_t item = blockIdx.x;
    const BatchWin bw = a.win[a.chunk_win[item]];
    const uint32_t local = item - bw.chunk_base;
    const bool left = local < bw.ncL;
    const uint32_t chunk = left ? local : local - bw.ncL;
    const int logF = int(bw.logF);
    uint32_t kmin = 0, shift = 0;
    if (KIND == PART_RANGE)
        sample_range(a.L + bw.lo, bw.nL, a.R + bw.ro, bw.nR, logF, s_mm, &kmin, &shift);
    PartFn<KIND> part;
    part.init(logF, kmin, shift, a.rmi);
    const uint32_t n = left ? bw.nL : bw.nR;
    const uint64_t base = uint64_t(chunk) * CHUNK;
    const int len = int(min<uint64_t>(CHUNK, n - base));
    partition_chunk<KIND, PAY_VAL, CHUNK, PEEL>(
        left ? a.L + bw.lo : a.R + bw.ro, base, len, logF, part,
        a.offs + bw.offs_base + uint64_t(local) * ((1u << logF) + 1),
        a.part + (left ? bw.lpart : bw.rpart) + base, lds, hs, a.rmi, left,
        PEEL ? a.heavy + uint64_t(bw.id) * kMaxHeavy * 4 : nullptr);
}

// Adds the peeled heavy keys' pairs of every window of a batch to its
// result: count += cntL * cntR, C1 += sumA * sumB per heavy key (C1
// factorises per key).  One thread per window, after the join kernel.
__global__ void k_heavy_finalize(sjb_window_result_t *__restrict__ res, uint32_t n,
                                 const uint64_t *__restrict__ heavy, uint32_t n_heavy) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    const uint64_t *h = heavy + uint64_t(w) * kMaxHeavy * 4;
    uint64_t cnt = 0, c1 = 0;
    for (uint32_t i = 0; i < n_heavy; i++) {
        cnt += h[4 * i] * h[4 * i + 2];
        c1 += h[4 * i + 1] * h[4 * i + 3];
    }
    res[w].count += cnt;
    res[w].checksum1 += c1;
}

// Zeroes the per-window results and the work-queue heads of a batch (a
// kernel, so a captured graph holds no memset node).
__global__ void k_batch_clear(sjb_window_result_t *__restrict__ res, uint32_t n,
                              uint32_t *__restrict__ qhead, uint64_t *__restrict__ heavy) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        res[i].count = 0;
        res[i].checksum1 = 0;
        res[i].checksum2 = 0;
        res[i].pairs = 0;
        if (heavy)
            for (int j = 0; j < kMaxHeavy * 4; j++) heavy[uint64_t(i) * kMaxHeavy * 4 + j] = 0;
    }
    if (i < kXcds * 16) qhead[i] = 0;
}

// log2 fanout for a window whose build side has n_build tuples: the
// expected partition size stays below what one LDS tile holds (HJ: ~0.92 of
// the CSR table, SMJ: half a sort tile, grouped HJ: 3/8 of the aggregation
// table), so uniform keys never spill into a second tile.  `bias` shifts
// the choice (autotuner); the result is clamped to [0, kMaxLogF].
int batch_log_fanout(int alg, uint32_t flags, uint64_t n_build, int bias) {
    uint64_t target = (kHjCap * 23) / 25;
    if (alg == SJB_ALG_SMJ)
        target = kSmjCap / 2;
    else if (flags & SJB_F_GROUPED)
        target = 768;
    int logF = 0;
    while (logF < kMaxLogF && (n_build >> logF) > target) logF++;
    logF += bias;
    return std::min(std::max(logF, 0), kMaxLogF);
}

// Fanout of a lone window: the tile rule above, raised to a floor while
// partitions keep >= 64 build records — a small window, or one with few
// distinct keys and many matches, would otherwise run its join on a few
// dozen workgroups.  Windows in kSmallChunk or kTinyChunk chunks (small): ~1024 build
// records per partition and at least 2^7 partitions, at most 2^9 (measured
// best at 2^14..2^19 tuples per side: 2^7 up to 2^16, 2^8 at 2^18, 2^9 at
// 2^19); windows in kChunk chunks: 2^9 partitions (two join workgroups per
// CU).
int choose_logF(int alg, uint32_t flags, uint64_t n_build, int requested, bool small) {
    if (requested > 0) return std::min(std::max(requested, 1), kMaxLogF);
    const int logF = std::max(batch_log_fanout(alg, flags, n_build, 0), 1);
    int floor_logF = 9;
    if (small) {
        floor_logF = 7;
        while (floor_logF < 9 && (n_build >> (floor_logF + 1)) >= 1024) floor_logF++;
    }
    while (floor_logF > logF && (n_build >> floor_logF) < 64) floor_logF--;
    return std::max(logF, floor_logF);
}

size_t partition_lds_bytes_chunk(int logF, int chunk) {
    const size_t words = ((size_t(1) << logF) + 2 + kPartThreads / kWave + 1) & ~size_t(1);
    return words * 4 + size_t(chunk) * sizeof(Rec);
}
size_t partition_lds_bytes(int logF) { return partition_lds_bytes_chunk(logF, kChunk); }

template <int KIND, bool PEEL, int CH>
static int partition_chunks(const WindowPlan &p, const sjb_tuple_t *L, const sjb_tuple_t *R, sjb_ctx *c) {
    const size_t lds = partition_lds_bytes_chunk(p.logF, CH);
    const int nblk = p.ncL + p.ncR;
    if (p.payload == PAY_VAL)
        k_partition<KIND, PAY_VAL, PEEL, CH><<<nblk, kPartThreads, lds, p.ps>>>(
            L, p.nL, p.ncL, R, p.nR, p.logF, c->d_scratch, p.range_blocks, c->d_rmi, p.offs, p.part, p.heavy,
            p.direct);
    else
        k_partition<KIND, PAY_IDX, false, CH><<<nblk, kPartThreads, lds, p.ps>>>(
            L, p.nL, p.ncL, R, p.nR, p.logF, c->d_scratch, p.range_blocks, c->d_rmi, p.offs, p.part, p.heavy,
            p.direct);
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

template <int KIND, bool PEEL>
static int partition_kind(sjb_ctx *c, const WindowPlan &p, const sjb_tuple_t *L,
                          const sjb_tuple_t *R) {
    if (p.chunk == kSmallChunk) return partition_chunks<KIND, PEEL, kSmallChunk>(p, L, R, c);
    if (p.chunk == kTinyChunk) return partition_chunks<KIND, PEEL, kTinyChunk>(p, L, R, c);
    if (p.chunk != kChunk) return SJB_E_INVALID;
    const size_t lds = partition_lds_bytes(p.logF);
    const int nblk = p.ncL + p.ncR;
#if SJB_PART_PIPE
    if constexpr (!PEEL && KIND != PART_LEARNED) {
        if (p.payload == PAY_VAL) {
            const int grid = std::min(nblk, std::max(c->n_cu, 1));
            k_partition_pipe<KIND, PAY_VAL><<<grid, kPartThreads, lds, p.ps>>>(
                L, p.nL, p.ncL, R, p.nR, p.ncR, p.logF, c->d_scratch, p.range_blocks, c->d_rmi, p.offs, p.part,
                p.direct);
            SJB_HIP_TRY(hipGetLastError());
            return SJB_OK;
        }
    }
#endif
    (void)lds;
    (void)nblk;
    return partition_chunks<KIND, PEEL, kChunk>(p, L, R, c);
}

int launch_partition(sjb_ctx *c, const WindowPlan &p, const sjb_tuple_t *L, const sjb_tuple_t *R) {
    if (p.ncL > kMaxChunks || p.ncR > kMaxChunks) return SJB_E_CAPACITY;
    if (uint64_t((1u << p.logF) + 1) * uint64_t(p.ncL + p.ncR) > c->offs_cap) return SJB_E_CAPACITY;
    switch (p.kind) {
        case PART_HASH:
            return p.peel ? partition_kind<PART_HASH, true>(c, p, L, R)
                          : partition_kind<PART_HASH, false>(c, p, L, R);
        case PART_RANGE: return partition_kind<PART_RANGE, false>(c, p, L, R);
        case PART_LEARNED:
            return p.peel ? partition_kind<PART_LEARNED, true>(c, p, L, R)
                          : partition_kind<PART_LEARNED, false>(c, p, L, R);
    }
    return SJB_E_INVALID;
}

template <int KIND, int CHUNK, bool PEEL>
static int partition_batch_kind(hipStream_t s, const BatchArgs &a, uint32_t n_chunks, size_t dyn) {
    k_partition_batch<KIND, CHUNK, PEEL><<<n_chunks, kPartThreads, dyn, s>>>(a);
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

int launch_partition_batch(hipStream_t s, int kind, bool peel, const BatchArgs &a, uint32_t n_chunks,
                           size_t dyn) {
    if (n_chunks == 0) return SJB_OK;
    const bool big = a.chunk == kChunk;
    if (!big && a.chunk != 4096) return SJB_E_INVALID;
    switch (kind) {
        case PART_HASH:
            if (peel)
                return big ? partition_batch_kind<PART_HASH, kChunk, true>(s, a, n_chunks, dyn)
                           : partition_batch_kind<PART_HASH, 4096, true>(s, a, n_chunks, dyn);
            return big ? partition_batch_kind<PART_HASH, kChunk, false>(s, a, n_chunks, dyn)
                       : partition_batch_kind<PART_HASH, 4096, false>(s, a, n_chunks, dyn);
        case PART_RANGE:
            if (peel)
                return big ? partition_batch_kind<PART_RANGE, kChunk, true>(s, a, n_chunks, dyn)
                           : partition_batch_kind<PART_RANGE, 4096, true>(s, a, n_chunks, dyn);
            return big ? partition_batch_kind<PART_RANGE, kChunk, false>(s, a, n_chunks, dyn)
                       : partition_batch_kind<PART_RANGE, 4096, false>(s, a, n_chunks, dyn);
        case PART_LEARNED:
            if (peel)
                return big ? partition_batch_kind<PART_LEARNED, kChunk, true>(s, a, n_chunks, dyn)
                           : partition_batch_kind<PART_LEARNED, 4096, true>(s, a, n_chunks, dyn);
            return big ? partition_batch_kind<PART_LEARNED, kChunk, false>(s, a, n_chunks, dyn)
                       : partition_batch_kind<PART_LEARNED, 4096, false>(s, a, n_chunks, dyn);
    }
    return SJB_E_INVALID;
}

// Partition sizes of one window: sizes[p] = left records of partition p,
// sizes[F + p] = right records (sums over the chunks' runs).
__global__ void k_part_sizes(const uint32_t *__restrict__ offs, int logF, int ncL, int ncR,
                             uint32_t *__restrict__ sizes) {
    const int F = 1 << logF;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= F) return;
    uint32_t a = 0, b = 0;
    for (int q = 0; q < ncL + ncR; q++) {
        const uint32_t *o = offs + uint64_t(q) * (F + 1);
        const uint32_t len = o[p + 1] - o[p];
        if (q < ncL)
            a += len;
        else
            b += len;
    }
    sizes[p] = a;
    sizes[F + p] = b;
}

int launch_part_sizes(hipStream_t s, const uint32_t *offs, int logF, int ncL, int ncR, uint32_t *sizes) {
    const int F = 1 << logF;
    k_part_sizes<<<(F + 255) / 256, 256, 0, s>>>(offs, logF, ncL, ncR, sizes);
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

int launch_heavy_finalize(hipStream_t s, sjb_window_result_t *res, uint32_t n, const uint64_t *heavy,
                          uint32_t n_heavy) {
    k_heavy_finalize<<<(n + 255) / 256, 256, 0, s>>>(res, n, heavy, n_heavy);
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

int launch_batch_clear(hipStream_t s, sjb_window_result_t *res, uint32_t n, uint32_t *qhead,
                       uint64_t *heavy) {
    const uint32_t m = std::max<uint32_t>(n, kXcds * 16);
    k_batch_clear<<<(m + 255) / 256, 256, 0, s>>>(res, n, qhead, heavy);
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

// Raises the dynamic-LDS limit of the partition kernels (the staging buffer
// is up to 128 KiB); called once per context.
int partition_setup(sjb_ctx *c) {
    const int need = int(partition_lds_bytes(kMaxLogF));
    (void)c;
    const void *fns[] = {
        (const void *)k_partition<PART_HASH, PAY_VAL, false, kChunk>,
        (const void *)k_partition<PART_HASH, PAY_IDX, false, kChunk>,
        (const void *)k_partition<PART_HASH, PAY_VAL, true, kChunk>,
        (const void *)k_partition_batch<PART_HASH, kChunk, true>,
        (const void *)k_partition_batch<PART_HASH, 4096, true>,
        (const void *)k_partition<PART_RANGE, PAY_VAL, false, kChunk>,
        (const void *)k_partition<PART_RANGE, PAY_IDX, false, kChunk>,
        (const void *)k_partition<PART_LEARNED, PAY_VAL, false, kChunk>,
        (const void *)k_partition<PART_LEARNED, PAY_IDX, false, kChunk>,
        (const void *)k_partition<PART_LEARNED, PAY_VAL, true, kChunk>,
        (const void *)k_partition_batch<PART_HASH, kChunk, false>,
        (const void *)k_partition_batch<PART_RANGE, kChunk, false>,
        (const void *)k_partition_batch<PART_RANGE, kChunk, true>,
        (const void *)k_partition_batch<PART_RANGE, 4096, true>,
        (const void *)k_partition_batch<PART_LEARNED, kChunk, false>,
        (const void *)k_partition_batch<PART_LEARNED, kChunk, true>,
        (const void *)k_partition_batch<PART_HASH, 4096, false>,
        (const void *)k_partition_batch<PART_RANGE, 4096, false>,
        (const void *)k_partition_batch<PART_LEARNED, 4096, false>,
        (const void *)k_partition_batch<PART_LEARNED, 4096, true>,
        (const void *)k_partition<PART_HASH, PAY_VAL, false, kSmallChunk>,
        (const void *)k_partition<PART_HASH, PAY_IDX, false, kSmallChunk>,
        (const void *)k_partition<PART_RANGE, PAY_VAL, false, kSmallChunk>,
        (const void *)k_partition<PART_RANGE, PAY_IDX, false, kSmallChunk>,
        (const void *)k_partition_pipe<PART_HASH, PAY_VAL>,
        (const void *)k_partition_pipe<PART_RANGE, PAY_VAL>,
    };
    for (const void *f : fns)
        (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, need);
    (void)hipGetLastError();
    return SJB_OK;
}

// ---------------------------------------------------------------------------
// fills (used instead of hipMemsetAsync so a captured graph never holds a
// memset node)
// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u64(uint64_t *p, uint64_t v, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

int launch_fill_u32(hipStream_t s, uint32_t *p, uint32_t v, uint64_t n) {
    if (n == 0) return SJB_OK;
    int blocks = int(std::min<uint64_t>((n + 255) / 256, 4096));
    k_fill_u32<<<blocks, 256, 0, s>>>(p, v, n);
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}
int launch_fill_u64(hipStream_t s, uint64_t *p, uint64_t v, uint64_t n) {
    if (n == 0) return SJB_OK;
    int blocks = int(std::min<uint64_t>((n + 255) / 256, 4096));
    k_fill_u64<<<blocks, 256, 0, s>>>(p, v, n);
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

}  // namespace sjb
