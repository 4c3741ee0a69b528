// join.hip — per-partition join kernels of the lazy HJ and SMJ operators.
//
// After k_partition (partition.hip), partition p of a window is the union
// over chunks c of the runs [offs[c][p], offs[c][p+1]) of chunk c's
// partition-sorted records.  One workgroup joins one partition in LDS:
//
//   * run table: the workgroup reads the run bounds of every chunk of both
//     sides (L2-resident: the offsets array is a few MiB) and scans the run
//     lengths, so record j of the partition is found by a binary search
//     over the run prefix;
//   * records are gathered in batches of kBatch loads per thread, issued
//     before any of them is used, so one memory latency covers a batch;
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
// batch with one global atomic per workgroup (block scan of the lanes'
// match counts), gather both tuples and write ResultTuple rows; they replace
// the per-match/per-thread atomics of probeWrite/probeJoin
// (GPU_HJ.cu:87,123).  `wrap` writes pos % cap (the Atomic method's ring,
// GPU_HJ.cu:92); otherwise rows past cap are dropped and counted.
#include "sjb_device.hpp"
#include "sjb_internal.hpp"

namespace sjb {

constexpr int kHjNB = 1 << kHjLogNB;
constexpr int kBatch = 8;
constexpr int kHjPerT = (kHjCap + kJoinThreads - 1) / kJoinThreads;  // records per thread per tile
#ifndef SJB_HJ_PROBE
#define SJB_HJ_PROBE 2048
#endif
constexpr int kHjProbe = SJB_HJ_PROBE;                  // probe records staged in LDS per pass
#ifndef SJB_BATCH_HJ_WAVES
#define SJB_BATCH_HJ_WAVES 6
#endif
// workgroups per CU launched for the resident count join (three fit)
#ifndef SJB_JOIN_PIPE_WG_PER_CU
#define SJB_JOIN_PIPE_WG_PER_CU 3
#endif
#ifndef SJB_BATCH_SMJ_WAVES
#define SJB_BATCH_SMJ_WAVES 6
#endif
constexpr int kHjPerP = kHjProbe / kJoinThreads;        // staged probe records per thread

// Record index of element j (< pref[nc]) by binary search over the run
// prefix (used only past the run-id table).
__device__ __forceinline__ uint32_t run_index(const uint32_t *pref, const uint32_t *base, int nc,
                                              uint32_t j) {
    int lo = 0, hi = nc;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pref[mid] <= j)
            lo = mid;
        else
            hi = mid;
    }
    return base[lo] + j;
}

// One partition of a window as seen from its join workgroup: for each side
// the exclusive prefix pref[c] of the run lengths, base[c] = (start of chunk
// c's run in d_part) - pref[c] (mod 2^32), so element j of run c is record
// base[c] + j, and a run-id table (element j -> run) for the first `idcap`
// elements: a record index costs two LDS reads.
struct PartView {
    uint32_t *base[2];
    uint32_t *pref[2];
    uint16_t *runid[2];
    int nc[2];
    uint32_t n[2];
    uint32_t idcap;
    __device__ __forceinline__ uint32_t index(int side, uint32_t j) const {
        if (j < idcap) {
            const uint32_t c = runid[side][j];
            return base[side][c] + j;
        }
        return run_index(pref[side], base[side], nc[side], j);
    }
    // index() for j known to be below idcap.
    __device__ __forceinline__ uint32_t index_tab(int side, uint32_t j) const {
        return base[side][runid[side][j]] + j;
    }
};

// Dynamic LDS bytes of a PartView: base/pref for both sides plus two
// run-id tables of idcap entries.
__host__ __device__ constexpr size_t part_view_bytes(int ncL, int ncR, int idcap) {
    return size_t(4) * (size_t(ncL) + ncR + (ncL + 1) + (ncR + 1)) + size_t(2) * 2 * idcap + 16;
}

// Builds the PartView of partition p in `dyn` (dynamic LDS); ends with a
// barrier.  `offs` is the window's first chunk-offsets row; left chunk c
// starts at record lbase + c*chunk of d_part, right chunk c at
// rbase + c*chunk.
__device__ __forceinline__ PartView load_view(uint32_t *dyn, const uint32_t *__restrict__ offs,
                                              int logF, int p, int ncL, int ncR, uint32_t lbase,
                                              uint32_t rbase, uint32_t chunk, uint32_t idcap,
                                              uint32_t *ws) {
    PartView v;
    v.nc[0] = ncL;
    v.nc[1] = ncR;
    v.base[0] = dyn;
    v.base[1] = dyn + ncL;
    v.pref[0] = dyn + ncL + ncR;
    v.pref[1] = v.pref[0] + ncL + 1;
    v.runid[0] = reinterpret_cast<uint16_t *>(v.pref[1] + ncR + 1);
    v.runid[1] = v.runid[0] + idcap;
    v.idcap = idcap;
    const int F1 = (1 << logF) + 1;
    // Run bounds of both sides: slot q < ncL is left chunk q, slot ncL + c is
    // right chunk c; one pass of global reads and one scan when they fit.
    const int nq = ncL + ncR;
    uint32_t carry = 0;
    for (int q0 = 0; q0 < nq; q0 += blockDim.x) {
        const int q = q0 + threadIdx.x;
        uint32_t len = 0;
        if (q < nq) {
            const bool right = q >= ncL;
            const int c = right ? q - ncL : q;
            const uint32_t *o = offs + uint64_t(q) * F1;
            const uint32_t a = o[p], b = o[p + 1];
            v.base[right][c] = (right ? rbase : lbase) + uint32_t(c) * chunk + a;  // run start for now
            len = b - a;
        }
        uint32_t tot;
        const uint32_t ex = block_excl_scan_u32(len, ws, &tot);
        if (q < nq) {
            const bool right = q >= ncL;
            // right prefixes are rebased below once the left total is known
            if (right)
                v.pref[1][q - ncL] = carry + ex;
            else
                v.pref[0][q] = carry + ex;
        }
        carry += tot;
    }
    __syncthreads();
    const uint32_t nleft = ncL ? (ncR ? v.pref[1][0] : carry) : 0;
    __syncthreads();  // every thread has read pref[1][0] before it is rebased
    for (int c = threadIdx.x; c < ncR; c += blockDim.x) v.pref[1][c] -= nleft;
    if (threadIdx.x == 0) {
        v.pref[0][ncL] = nleft;
        v.pref[1][ncR] = carry - nleft;
    }
    v.n[0] = nleft;
    v.n[1] = carry - nleft;
    __syncthreads();
    // run-id fill: one thread per run of either side (slot q as above), so
    // both sides fill at once; the run start becomes base = start - pref
    for (int q = threadIdx.x; q < nq; q += blockDim.x) {
        const int side = q >= ncL;
        const int c = side ? q - ncL : q;
        const uint32_t a = v.pref[side][c];
        v.base[side][c] -= a;
        const uint32_t b = min(v.pref[side][c + 1], idcap);
        for (uint32_t j = a; j < b; j++) v.runid[side][j] = uint16_t(c);
    }
    __syncthreads();
    return v;
}

// Thread q's run bounds (a, b) of partition p for load_view_small: slot q <
// ncL is left chunk q, slot ncL + c right chunk c; zero past the slots.
__device__ __forceinline__ void view_bounds(const uint32_t *__restrict__ offs, int logF, int p, int nq,
                                            uint32_t &a, uint32_t &b) {
    a = b = 0;
    const int q = threadIdx.x;
    if (q < nq) {
        const uint32_t *o = offs + uint64_t(q) * ((1 << logF) + 1);
        a = o[p];
        b = o[p + 1];
    }
}

// The view of load_view_small from bounds loaded by view_bounds.  Ends with
// a barrier.
__device__ __forceinline__ PartView view_small_from(uint32_t *dyn, uint32_t a, uint32_t b, int ncL, int ncR,
                                                    uint32_t lbase, uint32_t rbase, uint32_t chunk,
                                                    uint32_t idcap, uint32_t *ws) {
    PartView v;
    v.nc[0] = ncL;
    v.nc[1] = ncR;
    v.base[0] = dyn;
    v.base[1] = dyn + ncL;
    v.pref[0] = dyn + ncL + ncR;
    v.pref[1] = v.pref[0] + ncL + 1;
    v.runid[0] = reinterpret_cast<uint16_t *>(v.pref[1] + ncR + 1);
    v.runid[1] = v.runid[0] + idcap;
    v.idcap = idcap;
    const int nq = ncL + ncR;
    const int q = threadIdx.x;
    const bool right = q >= ncL;
    const int c = right ? q - ncL : q;
    uint32_t wt;
    const uint32_t ex = wave_excl_scan_full(b - a, &wt);
    const int wid = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) ws[wid] = wt;
    __syncthreads();
    uint32_t pre = 0, nleft = 0, tot = 0;
    const int lw = ncL / kWave;
    for (int w = 0; w < int(blockDim.x / kWave); w++) {
        const uint32_t x = ws[w];
        if (w < wid) pre += x;
        if (w < lw) nleft += x;
        tot += x;
    }
    if (q < nq) {
        const uint32_t pf = pre + ex - (right ? nleft : 0u);
        v.base[right][c] = (right ? rbase : lbase) + uint32_t(c) * chunk + a - pf;
        v.pref[right][c] = pf;
    }
    if (threadIdx.x == 0) {
        v.pref[0][ncL] = nleft;
        v.pref[1][ncR] = tot - nleft;
    }
    v.n[0] = nleft;
    v.n[1] = tot - nleft;
    __syncthreads();
    if (q < nq) {
        const uint32_t lo = v.pref[right][c];
        const uint32_t hi = min(v.pref[right][c + 1], idcap);
        for (uint32_t j = lo; j < hi; j++) v.runid[right][j] = uint16_t(c);
    }
    __syncthreads();
    return v;
}

// load_view when ncL + ncR <= blockDim and ncL is a whole number of waves
// (every window of up to 2^22 tuples per side with 64-aligned chunk counts,
// e.g. 2^22 = 256 chunks): one slot per thread, one scan, the left total
// from the wave sums, three barriers instead of six.  Ends with a barrier.
__device__ __forceinline__ PartView load_view_small(uint32_t *dyn, const uint32_t *__restrict__ offs, int logF,
                                                    int p, int ncL, int ncR, uint32_t lbase, uint32_t rbase,
                                                    uint32_t chunk, uint32_t idcap, uint32_t *ws) {
    uint32_t a, b;
    view_bounds(offs, logF, p, ncL + ncR, a, b);
    return view_small_from(dyn, a, b, ncL, ncR, lbase, rbase, chunk, idcap, ws);
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
// Partition handled by workgroup `bid`.  Workgroups are dealt round-robin
// over the 8 XCDs of the MI355X; giving each XCD a contiguous range of
// partitions makes the runs of neighbouring partitions (which share cache
// lines inside every chunk) and their offset words hit the same L2.
__device__ __forceinline__ int xcd_partition(int bid, int F) {
    constexpr int kXcd = 8;
    if (F % kXcd) return bid;
    return (bid % kXcd) * (F / kXcd) + bid / kXcd;
}

// LDS table of one build tile in CSR form: records sorted by bucket
// (bucket = hash bits below the partition bits), off[b] .. off[b+1] is
// bucket b.  Built by an LDS counting sort; a probe reads two bounds and
// scans contiguous keys (independent loads, no pointer chasing).
struct HjTable {
    uint2 kv[kHjCap];  // (key, payload): one 8-byte LDS read per entry
    uint32_t off[kHjNB + 1];
};

__device__ __forceinline__ uint32_t hj_bucket(uint32_t k, int bshift) {
    return (key_hash(k) << bshift) >> (32 - kHjLogNB);
}

// Builds the CSR table from up to kHjPerT register-resident records per
// thread (element i = k * blockDim + tid < tn).  Starts and ends with a
// barrier; off[] must be zero on entry.
__device__ __forceinline__ void hj_build(HjTable &t, const Rec (&r)[kHjPerT], uint32_t tn, int bshift,
                                         uint32_t *ws) {
    uint32_t slot[kHjPerT];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kHjPerT; k++) {
        const uint32_t i = k * kJoinThreads + threadIdx.x;
        if (i < tn) {
            const uint32_t b = hj_bucket(r[k].x, bshift);
            slot[k] = (b << 16) | atomicAdd(&t.off[b], 1u);
        }
    }
    __syncthreads();
    // exclusive scan of the kHjNB bucket counts
    constexpr int kPerB = (kHjNB + kJoinThreads - 1) / kJoinThreads;
    uint32_t v[kPerB];
    uint32_t sum = 0;
#pragma unroll
    for (int q = 0; q < kPerB; q++) {
        const int idx = threadIdx.x * kPerB + q;
        v[q] = idx < kHjNB ? t.off[idx] : 0;
        sum += v[q];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan_u32(sum, ws, &tot);
#pragma unroll
    for (int q = 0; q < kPerB; q++) {
        const int idx = threadIdx.x * kPerB + q;
        if (idx < kHjNB) t.off[idx] = run;
        run += v[q];
    }
    if (threadIdx.x == 0) t.off[kHjNB] = tot;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kHjPerT; k++) {
        const uint32_t i = k * kJoinThreads + threadIdx.x;
        if (i < tn) {
            const uint32_t pos = t.off[slot[k] >> 16] + (slot[k] & 0xFFFFu);
            t.kv[pos] = make_uint2(r[k].x, r[k].y);
        }
    }
    __syncthreads();
}

// Gathers elements [base, base+n) of one side of the partition, kBatch
// loads per thread in flight, and hands each record to f(i, rec).
template <class Fn>
__device__ __forceinline__ void gather(const Rec *__restrict__ part, const PartView &v, int side,
                                       uint32_t base, uint32_t n, Fn &&f) {
    for (uint32_t i0 = 0; i0 < n; i0 += kJoinThreads * kBatch) {
        Rec r[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            const uint32_t i = i0 + k * kJoinThreads + threadIdx.x;
            if (i < n) r[k] = part[v.index(side, base + i)];
        }
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            const uint32_t i = i0 + k * kJoinThreads + threadIdx.x;
            if (i < n) f(i, r[k]);
        }
    }
}

template <bool C2>
__device__ __forceinline__ void pair_acc(uint32_t k, uint32_t rv, uint32_t sv, Acc &acc) {
    const uint64_t a = hash_a(k, rv);
    acc.c1 += a * hash_b(k, sv);
    if (C2) acc.c2 += pair_hash_from_a(a, sv);
}

// Matches of one wave waiting for their checksum math: key, probe val and
// build val.  Queuing them and running the 64-bit math up to 64 matches at
// a time keeps the lanes busy, where running it in place costs a full wave
// pass for each probe round in which any lane matched.
struct PairQ {
    uint32_t key[kWave];
    uint32_t sv[kWave];
    uint32_t rv[kWave];
};

// Lanes [0, n) (n <= 64) each take one queued match; the queue is free
// again afterwards.
template <bool C2>
__device__ __forceinline__ void pairq_drain(const PairQ &q, uint32_t n, Acc &acc) {
    const uint32_t lane = threadIdx.x & (kWave - 1);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    if (lane < n) pair_acc<C2>(q.key[lane], q.rv[lane], q.sv[lane], acc);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

// Count / C1 / C2 of up to kHjPerP register-resident probe records per
// thread (element j = k * blockDim + tid < un) against the current tile
// (payload = val).  The bucket scan compares keys only; the first two
// matches of every probe go to the wave's queue (later matches, rare under
// uniform keys, are summed in place).  ALLQ (a batched-join variant the
// autotuner picks for skewed windows): every match goes through the queue
// instead — fewer idle lanes under hot keys, ~5% more work under uniform
// keys.
template <bool C2, bool ALLQ>
__device__ __forceinline__ void hj_probe_regs(const HjTable &t, const Rec (&rp)[kHjPerP], uint32_t un,
                                              int bshift, Acc &acc, PairQ &q) {
    uint32_t qn = 0;  // queued matches (wave-uniform)
    if constexpr (ALLQ) {
        // every match through the queue: per round each lane advances its
        // bucket scan to its next match and the wave queues the round's
        // matches, so a hot key's long match list runs on full waves
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < kHjPerP; k++) {
            const uint32_t j = k * kJoinThreads + threadIdx.x;
            const Rec s = rp[k];
            uint32_t e = 0, end = 0;
            if (j < un) {
                const uint32_t b = hj_bucket(s.x, bshift);
                e = t.off[b];
                end = t.off[b + 1];
            }
            for (;;) {
                while (e < end && t.kv[e].x != s.x) e++;
                const bool has = e < end;
                const uint64_t mask = __ballot(has);
                if (mask == 0) break;  // wave-uniform
                const uint32_t c = uint32_t(__popcll(mask));
                if (qn + c > uint32_t(kWave)) {
                    pairq_drain<C2>(q, qn, acc);
                    qn = 0;
                }
                if (has) {
                    const uint32_t pos = qn + lanes_below(mask);
                    q.key[pos] = s.x;
                    q.sv[pos] = s.y;
                    q.rv[pos] = t.kv[e].y;
                    cnt++;
                    e++;
                }
                qn += c;
            }
        }
        acc.count += cnt;
        pairq_drain<C2>(q, qn, acc);
        return;
    }
#pragma unroll
    for (int k = 0; k < kHjPerP; k++) {
        const uint32_t j = k * kJoinThreads + threadIdx.x;
        const Rec s = rp[k];
        uint32_t m = 0, rv0 = 0, rv1 = 0, lo = 0, hi = 0;
        if (j < un) {
            const uint32_t b = hj_bucket(s.x, bshift);
            lo = t.off[b];
            hi = t.off[b + 1];
            // branch-free scan: selects keep the first two matches' payloads
            for (uint32_t e = lo; e < hi; e++) {
                const uint2 kv = t.kv[e];
                const bool eq = kv.x == s.x;
                rv1 = eq && m == 1 ? kv.y : rv1;
                rv0 = eq && m == 0 ? kv.y : rv0;
                m += eq;
            }
        }
        acc.count += m;
        // third and later matches (rare under uniform keys): the lanes that
        // have them scan their bucket again and sum those pairs in place
        if (__ballot(m > 2) != 0) {  // wave-uniform
            if (m > 2) {
                uint32_t seen = 0;
                for (uint32_t e = lo; e < hi; e++) {
                    const uint2 kv = t.kv[e];
                    if (kv.x == s.x) {
                        if (seen >= 2) pair_acc<C2>(s.x, kv.y, s.y, acc);
                        seen++;
                    }
                }
            }
        }
        for (uint32_t r = 0; r < 2; r++) {
            const bool has = m > r;
            const uint64_t mask = __ballot(has);
            if (mask == 0) break;  // wave-uniform
            const uint32_t c = uint32_t(__popcll(mask));
            if (qn + c > uint32_t(kWave)) {
                pairq_drain<C2>(q, qn, acc);
                qn = 0;
            }
            if (has) {
                const uint32_t pos = qn + lanes_below(mask);
                q.key[pos] = s.x;
                q.sv[pos] = s.y;
                q.rv[pos] = r == 0 ? rv0 : rv1;
            }
            qn += c;
        }
    }
    pairq_drain<C2>(q, qn, acc);
}

// Probes the un records staged in LDS against the current tile and writes
// pair rows (payload = tuple index): per round one block scan of the lanes'
// match counts and one global reservation, then rows are written.
template <bool C2>
__device__ __forceinline__ void hj_probe_write(const HjTable &t, const Rec *sp, uint32_t un, int bshift,
                                               Acc &acc, const sjb_tuple_t *__restrict__ L,
                                               const sjb_tuple_t *__restrict__ R,
                                               sjb_result_tuple_t *__restrict__ out, uint64_t cap, int wrap,
                                               unsigned long long *__restrict__ cursor, uint32_t *ws,
                                               unsigned long long *s_base) {
    for (uint32_t j0 = 0; j0 < un; j0 += kJoinThreads) {
        const uint32_t j = j0 + threadIdx.x;
        Rec s = make_uint2(0, 0);
        uint32_t m = 0;
        if (j < un) {
            s = sp[j];
            const uint32_t b = hj_bucket(s.x, bshift);
            for (uint32_t e = t.off[b]; e < t.off[b + 1]; e++) m += t.kv[e].x == s.x;
        }
        uint32_t tot;
        const uint32_t ex = block_excl_scan_u32(m, ws, &tot);
        if (threadIdx.x == 0) *s_base = tot ? atomicAdd(cursor, (unsigned long long)tot) : 0ull;
        __syncthreads();
        if (m) {
            uint64_t pos = *s_base + ex;
            const sjb_tuple_t st = R[s.y];
            const uint32_t b = hj_bucket(s.x, bshift);
            for (uint32_t e = t.off[b]; e < t.off[b + 1]; e++) {
                if (t.kv[e].x != s.x) continue;
                const sjb_tuple_t lt = L[t.kv[e].y];
                const uint64_t a = hash_a(lt.key, lt.val);
                acc.count++;
                acc.c1 += a * hash_b(st.key, st.val);
                if (C2) acc.c2 += pair_hash_from_a(a, st.val);
                write_row(out, pos++, cap, wrap, lt, st);
            }
        }
        __syncthreads();
    }
}

// Arguments every per-partition join body needs.
struct JoinIO {
    const Rec *part;
    const sjb_tuple_t *L, *R;             // window inputs (pair output only)
    sjb_result_tuple_t *out;
    uint64_t out_cap;
    int wrap;
    unsigned long long *cursor;
};

// LDS of the HJ body: the CSR build table and, when pair rows are written,
// the probe staging buffer; otherwise probe records stay in registers and
// each wave has a match queue.
template <bool WRITE>
struct HjLds;
template <>
struct HjLds<true> {
    HjTable t;
    Rec probe[kHjProbe];
};
template <>
struct HjLds<false> {
    HjTable t;
    PairQ q[kJoinThreads / kWave];
};

// Joins one partition (HJ).  The first build tile and the first kHjProbe
// probe records are gathered with all loads in flight at once (build
// records go straight into the LDS table; probe records stay in registers,
// or, when pair rows are written, go to an LDS staging buffer); larger
// partitions continue tile by tile.  The caller has built `v` (load_view
// ends with a barrier).
template <bool C2, bool WRITE, bool ALLQ>
__device__ __forceinline__ void hj_join_part(const JoinIO &io, const PartView &v, int bshift,
                                             HjLds<WRITE> &m, uint32_t *s_ws, unsigned long long *s_base,
                                             Acc &acc) {
    const uint32_t nb = v.n[0], np = v.n[1];
    if (!nb || !np) return;
    HjTable &t = m.t;
    const Rec *__restrict__ part = io.part;
    for (int i = threadIdx.x; i <= kHjNB; i += kJoinThreads) t.off[i] = 0;
    // The first tile and probe pass lie inside the run-id table: an HJ view
    // holds kHjCap entries per side (idcap_of, join_idcap).
    static_assert(kHjProbe <= kHjCap, "first probe pass must fit the run-id table");
    const uint32_t tn0 = min<uint32_t>(nb, kHjCap), pn0 = min<uint32_t>(np, kHjProbe);
    Rec rb[kHjPerT];
    Rec rp[kHjPerP];
#pragma unroll
    for (int k = 0; k < kHjPerT; k++) {
        const uint32_t i = k * kJoinThreads + threadIdx.x;
        if (i < tn0) rb[k] = part[v.index_tab(0, i)];
    }
#pragma unroll
    for (int k = 0; k < kHjPerP; k++) {
        const uint32_t j = k * kJoinThreads + threadIdx.x;
        rp[k] = make_uint2(0, 0);
        if (j < pn0) rp[k] = part[v.index_tab(1, j)];
    }
    if constexpr (WRITE) {
#pragma unroll
        for (int k = 0; k < kHjPerP; k++) {
            const uint32_t j = k * kJoinThreads + threadIdx.x;
            if (j < pn0) m.probe[j] = rp[k];
        }
    }
    uint32_t resident = 0;  // first element of the resident probe chunk
    for (uint32_t t0 = 0; t0 < nb; t0 += kHjCap) {
        const uint32_t tn = min<uint32_t>(kHjCap, nb - t0);
        if (t0 > 0) {
            __syncthreads();
            for (int i = threadIdx.x; i <= kHjNB; i += kJoinThreads) t.off[i] = 0;
#pragma unroll
            for (int k = 0; k < kHjPerT; k++) {
                const uint32_t i = k * kJoinThreads + threadIdx.x;
                if (i < tn) rb[k] = part[v.index(0, t0 + i)];
            }
        }
        hj_build(t, rb, tn, bshift, s_ws);
        for (uint32_t u0 = 0; u0 < np; u0 += kHjProbe) {
            const uint32_t un = min<uint32_t>(kHjProbe, np - u0);
            if constexpr (WRITE) {
                if (u0 != resident) {
                    __syncthreads();
                    gather(part, v, 1, u0, un, [&](uint32_t i, Rec r) { m.probe[i] = r; });
                    resident = u0;
                    __syncthreads();
                }
                hj_probe_write<C2>(t, m.probe, un, bshift, acc, io.L, io.R, io.out, io.out_cap, io.wrap,
                                   io.cursor, s_ws, s_base);
            } else {
                if (u0 != resident) {
#pragma unroll
                    for (int k = 0; k < kHjPerP; k++) {
                        const uint32_t j = k * kJoinThreads + threadIdx.x;
                        if (j < un) rp[k] = part[v.index(1, u0 + j)];
                    }
                    resident = u0;
                }
                hj_probe_regs<C2, ALLQ>(t, rp, un, bshift, acc, m.q[threadIdx.x / kWave]);
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Aggregated (group-by-key) build: grouped mode, count + C1 only
// ---------------------------------------------------------------------------
// C1 factorises per key: sum over pairs of A(r)*B(s) = sum over probe
// records s of B(s) * (sum of A over the build records with s's key).  The
// build side of a partition is aggregated into an LDS open-addressing table
// of distinct keys: count, the val of the key's first record, and the sum
// of A over its other records — A of the first record is only computed if a
// probe matches, so unique build keys cost no hash at build time.  A probe
// record costs one lookup whatever the key's multiplicity, so the work is
// O(n) at any skew.  The table is filled in rounds of at most kJoinThreads
// records; before a round starts with more than kAggHigh distinct keys in
// the table, every probe record is looked up against it and it is cleared,
// so the table never exceeds 3/4 load.  Key 0xFFFFFFFF (the empty marker)
// has its own slot.
constexpr int kAggLog = kJoinThreads >= 1024 ? 12 : 11;
constexpr int kAggSlots = 1 << kAggLog;
constexpr uint32_t kAggHigh = kAggSlots / 2;
constexpr uint32_t kEmptyKey = 0xFFFFFFFFu;
constexpr int kAggIdCap = 1536;  // run-id table entries per side of a grouped HJ view
static_assert(kAggHigh + kJoinThreads <= (kAggSlots * 3) / 4, "aggregation table load");

struct AggLds {
    uint32_t key[kAggSlots];
    uint32_t cnt[kAggSlots];
    uint32_t fval[kAggSlots];   // val of the record that made the count 0 -> n
    uint64_t suma[kAggSlots];   // sum of A over the other records
    uint64_t ff_suma;           // key 0xFFFFFFFF
    uint32_t ff_cnt, ff_fval;
    uint32_t distinct;
};

__device__ __forceinline__ uint32_t agg_slot(uint32_t k, int bshift) {
    return (key_hash(k) << bshift) >> (32 - kAggLog);
}

// Every thread of the block; the caller puts a barrier after it.
__device__ __forceinline__ void agg_clear(AggLds &t) {
    for (int i = threadIdx.x; i < kAggSlots; i += kJoinThreads) {
        t.key[i] = kEmptyKey;
        t.cnt[i] = 0;
        t.suma[i] = 0;
    }
    if (threadIdx.x == 0) {
        t.ff_suma = 0;
        t.ff_cnt = 0;
        t.distinct = 0;
    }
}

// Slot of key k (inserted if absent); kAggSlots for key 0xFFFFFFFF.
__device__ __forceinline__ uint32_t agg_find_insert(AggLds &t, uint32_t k, int bshift) {
    if (k == kEmptyKey) return kAggSlots;
    uint32_t s = agg_slot(k, bshift);
    for (;;) {
        uint32_t cur = __hip_atomic_load(&t.key[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == kEmptyKey) {
            cur = atomicCAS(&t.key[s], kEmptyKey, k);
            if (cur == kEmptyKey) {
                atomicAdd(&t.distinct, 1u);
                return s;
            }
        }
        if (cur == k) return s;
        s = (s + 1) & (kAggSlots - 1);
    }
}

// Adds `n` records of key k to its slot: `v` is one of them (the one that
// becomes the slot's first record if the count was 0) and `rest` the sum of
// A over the other n-1.
__device__ __forceinline__ void agg_add(AggLds &t, uint32_t k, uint32_t n, uint32_t v, uint64_t rest,
                                        int bshift) {
    const uint32_t s = agg_find_insert(t, k, bshift);
    uint32_t *cnt = s == kAggSlots ? &t.ff_cnt : &t.cnt[s];
    uint64_t *sum = s == kAggSlots ? &t.ff_suma : &t.suma[s];
    const uint32_t old = atomicAdd(cnt, n);
    if (old == 0)
        (s == kAggSlots ? t.ff_fval : t.fval[s]) = v;
    else
        rest += hash_a(k, v);
    if (rest) atomicAdd(reinterpret_cast<unsigned long long *>(sum), (unsigned long long)rest);
}

// Inserts one record per lane (`valid` lanes); called by every lane of the
// wave.  When several lanes of the wave share the first valid lane's key
// (the common case under heavy skew, where one key dominates a partition),
// they are combined into one update.
__device__ __forceinline__ void agg_insert(AggLds &t, bool valid, Rec r, int bshift) {
    const uint64_t vmask = __ballot(valid);
    if (vmask == 0) return;
    const int lane = threadIdx.x & (kWave - 1);
    const int lead = __ffsll((unsigned long long)vmask) - 1;
    const uint32_t k0 = uint32_t(__builtin_amdgcn_readlane(int(r.x), lead));
    const bool same = valid && r.x == k0;
    const uint64_t smask = __ballot(same);
    if (__popcll(smask) > 1) {  // wave-uniform
        const uint64_t sa = wave_sum_full(same && lane != lead ? hash_a(r.x, r.y) : 0ull);
        if (lane == lead) agg_add(t, k0, uint32_t(__popcll(smask)), r.y, sa, bshift);
        if (valid && !same) agg_add(t, r.x, 1u, r.y, 0ull, bshift);
    } else if (valid) {
        agg_add(t, r.x, 1u, r.y, 0ull, bshift);
    }
}

__device__ __forceinline__ void agg_probe(const AggLds &t, Rec s, int bshift, Acc &acc) {
    uint32_t cnt = 0, fv = 0;
    uint64_t sa = 0;
    if (s.x == kEmptyKey) {
        cnt = t.ff_cnt;
        fv = t.ff_fval;
        sa = t.ff_suma;
    } else {
        uint32_t sl = agg_slot(s.x, bshift);
        for (;;) {
            const uint32_t cur = t.key[sl];
            if (cur == s.x) {
                cnt = t.cnt[sl];
                fv = t.fval[sl];
                sa = t.suma[sl];
                break;
            }
            if (cur == kEmptyKey) break;
            sl = (sl + 1) & (kAggSlots - 1);
        }
    }
    if (cnt) {
        acc.count += cnt;
        acc.c1 += (sa + hash_a(s.x, fv)) * hash_b(s.x, s.y);
    }
}

__device__ __forceinline__ void agg_join_part(const JoinIO &io, const PartView &v, int bshift,
                                              AggLds &t, Acc &acc) {
    const uint32_t nb = v.n[0], np = v.n[1];
    if (!nb || !np) return;
    const Rec *__restrict__ part = io.part;
    agg_clear(t);
    __syncthreads();
    // probe every probe record against the table (read-only meanwhile):
    // kBatch register-resident records per thread per pass
    auto probe_all = [&]() {
        for (uint32_t u0 = 0; u0 < np; u0 += kJoinThreads * kBatch) {
            Rec r[kBatch];
#pragma unroll
            for (int k = 0; k < kBatch; k++) {
                const uint32_t j = u0 + k * kJoinThreads + threadIdx.x;
                if (j < np) r[k] = part[v.index(1, j)];
            }
#pragma unroll
            for (int k = 0; k < kBatch; k++) {
                const uint32_t j = u0 + k * kJoinThreads + threadIdx.x;
                if (j < np) agg_probe(t, r[k], bshift, acc);
            }
        }
    };
    for (uint32_t i0 = 0; i0 < nb; i0 += kJoinThreads * kBatch) {
        Rec r[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            const uint32_t i = i0 + k * kJoinThreads + threadIdx.x;
            r[k] = i < nb ? part[v.index(0, i)] : make_uint2(0u, 0u);
        }
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            if (i0 + k * kJoinThreads >= nb) break;  // uniform
            __syncthreads();
            const bool full = t.distinct > kAggHigh;
            __syncthreads();
            if (full) {
                probe_all();
                __syncthreads();
                agg_clear(t);
                __syncthreads();
            }
            const uint32_t i = i0 + k * kJoinThreads + threadIdx.x;
            agg_insert(t, i < nb, r[k], bshift);
        }
    }
    __syncthreads();
    probe_all();
}

// ---------------------------------------------------------------------------
// SMJ
// ---------------------------------------------------------------------------
// Value of lane (lane ^ m), m a power of two below 64 known at compile time
// after unrolling; every lane of the wave active.  m = 1, 2: DPP quad_perm;
// m = 4, 8: a row_shl:m move into the banks (4-lane groups of a 16-lane
// row) whose partner lies above and a row_shr:m move into the others —
// VALU instructions, no LDS round trip; m = 16: ds_swizzle (xor mask within
// 32 lanes, no address computation); m = 32: ds_bpermute.
__device__ __forceinline__ uint32_t lane_xor(uint32_t v, int m) {
    switch (m) {
        case 1: return dpp_u32<0xB1>(v);  // quad_perm [1,0,3,2]
        case 2: return dpp_u32<0x4E>(v);  // quad_perm [2,3,0,1]
        case 4: {
            const int up = __builtin_amdgcn_update_dpp(0, int(v), 0x104, 0xF, 0x5, false);  // row_shl:4, banks 0, 2
            return uint32_t(__builtin_amdgcn_update_dpp(up, int(v), 0x114, 0xF, 0xA, false));  // row_shr:4, banks 1, 3
        }
        case 8: {
            const int up = __builtin_amdgcn_update_dpp(0, int(v), 0x108, 0xF, 0x3, false);  // row_shl:8, banks 0, 1
            return uint32_t(__builtin_amdgcn_update_dpp(up, int(v), 0x118, 0xF, 0xC, false));  // row_shr:8, banks 2, 3
        }
        case 16: return uint32_t(__builtin_amdgcn_ds_swizzle(int(v), 0x401F));  // and 0x1F, xor 0x10
        default: return uint32_t(__shfl_xor(int(v), m));
    }
}
__device__ __forceinline__ uint64_t shfl_xor_v(uint64_t v, int m) {
    const uint32_t lo = lane_xor(uint32_t(v), m);
    const uint32_t hi = lane_xor(uint32_t(v >> 32), m);
    return (uint64_t(hi) << 32) | lo;
}
__device__ __forceinline__ uint32_t shfl_xor_v(uint32_t v, int m) { return lane_xor(v, m); }

static_assert(kSmjCap == 4 * kJoinThreads, "SMJ sort tiles hold up to 4 elements per thread");

// Bitonic sorting network over N = kJoinThreads * E values held in
// registers, ascending: thread t holds elements E*t .. E*t+E-1 in x.
// Strides below E are compare-exchanges inside the thread, strides up to
// 64*E exchange with lane t ^ (j/E) of the same wave (cross-lane
// shuffles), and only the larger strides (6 of the log2(N)(log2(N)+1)/2
// stages) go through LDS, one barrier each, alternating between buf0 and
// buf1 (lane-contiguous layout, no bank conflicts; each holds N values).
// Ends with a barrier after the last LDS stage's reads.
template <int E, class T>
__device__ __forceinline__ void bitonic_network(T (&x)[E], T *buf0, T *buf1) {
    constexpr int N = kJoinThreads * E;
    constexpr int LOGN = E == 1 ? 9 : E == 2 ? 10 : 11;
    static_assert(N == (1 << LOGN), "tile of kJoinThreads * E elements");
    // opaque copy of the lane index: the network's shuffle and LDS addresses
    // are computed here, not shared with (and held live across) other code
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    int flip = 0;
#pragma unroll
    for (int lk = 1; lk <= LOGN; lk++) {
        const int k = 1 << lk;
#pragma unroll
        for (int lj = lk - 1; lj >= 0; lj--) {
            const int j = 1 << lj;
            if (j < E) {
#pragma unroll
                for (int r = 0; r < E; r++) {
                    if (r & j) continue;
                    const bool up = ((E * t + r) & k) == 0;
                    const T a = x[r], b = x[r | j];
                    const bool sw = (a > b) == up;
                    x[r] = sw ? b : a;
                    x[r | j] = sw ? a : b;
                }
            } else {
                const int tj = j / E;  // partner thread distance
                const bool keep_min = ((t & tj) == 0) == (((E * t) & k) == 0);
                T y[E];
                if (tj < kWave) {
#pragma unroll
                    for (int r = 0; r < E; r++) y[r] = shfl_xor_v(x[r], tj);
                } else {
                    T *buf = flip ? buf1 : buf0;
                    flip ^= 1;
#pragma unroll
                    for (int r = 0; r < E; r++) buf[r * kJoinThreads + t] = x[r];
                    __syncthreads();
#pragma unroll
                    for (int r = 0; r < E; r++) y[r] = buf[r * kJoinThreads + (t ^ tj)];
                }
#pragma unroll
                for (int r = 0; r < E; r++) {
                    const bool lt = x[r] < y[r];
                    x[r] = (lt == keep_min) ? x[r] : y[r];
                }
            }
        }
    }
    __syncthreads();
}

// Sorts the 64-bit (key << 32 | payload) tile s[0, N) in place.  When the
// tile's keys span less than 2^20 (kmin: smallest key) the network runs on
// 32-bit values ((key - kmin) << 11 | position), half the shuffles and
// compares of the 64-bit values, and the records are permuted once at the
// end; ~0 marks padding (a real value is below 2^31).  Otherwise the
// 64-bit values themselves are sorted.  tmp: scratch of N 64-bit values.
// Expects s written and a barrier passed; ends with s sorted and a barrier.
template <int E>
__device__ __forceinline__ void sort_tile(uint64_t *s, uint64_t *tmp, int n, bool narrow, uint32_t kmin) {
    constexpr int N = kJoinThreads * E;
    const int t = threadIdx.x;
    if (narrow) {
        uint32_t x[E];
#pragma unroll
        for (int r = 0; r < E; r++) {
            const int i = E * t + r;
            x[r] = i < n ? ((uint32_t(s[i] >> 32) - kmin) << 11) | uint32_t(i) : ~0u;
        }
        uint32_t *b32 = reinterpret_cast<uint32_t *>(tmp);
        bitonic_network<E>(x, b32, b32 + N);
        uint64_t v[E];
#pragma unroll
        for (int r = 0; r < E; r++) v[r] = x[r] == ~0u ? ~0ull : s[x[r] & 2047u];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < E; r++) s[E * t + r] = v[r];
    } else {
        uint64_t x[E];
#pragma unroll
        for (int r = 0; r < E; r++) x[r] = s[E * t + r];
        // the last LDS stage reads buf1 = s only after every thread's
        // reads above (its barrier follows them)
        bitonic_network<E>(x, tmp, s);
#pragma unroll
        for (int r = 0; r < E; r++) s[E * t + r] = x[r];
    }
    __syncthreads();
}

// Key span handled by the direct counting sort (one u32 counter per key
// value, in the scratch of the sort tile).
constexpr uint32_t kSmjDirect = 4096;
// Run-id table entries per side of a non-grouped SMJ view (records past them
// are located by binary search over the runs): 1536 keeps the workgroup,
// with the key directory, at three per CU; range partitions hold ~1024.
constexpr int kSmjIdCap = 1536;
static_assert(kSmjDirect * sizeof(uint32_t) <= kSmjCap * sizeof(uint64_t), "counters fit the sort scratch");
static_assert(kSmjDirect % kJoinThreads == 0, "counters split evenly over the threads");

// Sorts the n records of s (n <= kSmjCap) by key when every key lies in
// [kmin, kmin + kSmjDirect): a counting sort with one counter per key value
// — one LDS atomic per record (its rank among equal keys), a block scan of
// the counters, one scattered write per record — instead of a sorting
// network.  Equal keys come out in arbitrary order (no consumer depends on
// it).  cnt: kSmjDirect counters, zeroed before the caller's barrier; dir
// receives the tile's key directory: dir[d] is the first position of key
// kmin + d, dir[kSmjDirect] = n, so key kmin + d occupies [dir[d],
// dir[d+1]).  ws: kJoinThreads / kWave words.  Ends with s sorted and a
// barrier.
__device__ __forceinline__ void counting_sort_tile(uint64_t *s, uint32_t *cnt, uint16_t *dir, uint32_t *ws, int n,
                                                   uint32_t kmin) {
    constexpr int E = kSmjCap / kJoinThreads;
    constexpr int P = int(kSmjDirect) / kJoinThreads;
    static_assert(kSmjDirect <= (1u << 16) && kSmjCap <= (1 << 16), "key offset and rank share a word");
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));  // addresses computed here, not hoisted and held live
    // only (key offset << 16 | rank) stays live across the scan; the
    // records are read again for the scatter
    uint32_t kr[E];
#pragma unroll
    for (int r = 0; r < E; r++) {
        const int i = E * t + r;
        const uint32_t d = uint32_t(s[i] >> 32) - kmin;
        kr[r] = i < n ? (d << 16) | atomicAdd(&cnt[d], 1u) : 0u;
    }
    __syncthreads();
    uint32_t sum = 0;
#pragma unroll
    for (int q = 0; q < P; q++) sum += cnt[P * t + q];
    uint32_t tot;
    uint32_t run = block_excl_scan_u32(sum, ws, &tot);
#pragma unroll
    for (int q = 0; q < P; q++) {
        const uint32_t c = cnt[P * t + q];
        dir[P * t + q] = uint16_t(run);
        run += c;
    }
    if (t == 0) dir[kSmjDirect] = uint16_t(n);
    __syncthreads();
    uint64_t x[E];
#pragma unroll
    for (int r = 0; r < E; r++) x[r] = s[E * t + r];
    __syncthreads();  // every record read before the in-place scatter
#pragma unroll
    for (int r = 0; r < E; r++) {
        const int i = E * t + r;
        if (i < n) s[dir[kr[r] >> 16] + (kr[r] & 0xFFFFu)] = x[r];
    }
    __syncthreads();
}

// Elements of the sort tile holding n records: 512, 1024 or 2048.
__device__ __forceinline__ int sort_tile_size(int n) {
    return n <= kJoinThreads ? kJoinThreads : n <= 2 * kJoinThreads ? 2 * kJoinThreads : 4 * kJoinThreads;
}

// Gathers elements [t0, t0+n) of one side into s as (key << 32 | payload).
// SORT: pads to the sort tile size with ~0 and sorts (tmp: scratch of
// kSmjCap 64-bit words; s_mm: 3 * kJoinThreads / kWave words of LDS for the
// tile's key bounds and scan scratch) — by a counting sort when the keys
// span less than kSmjDirect values, else by a sorting network; padding
// entries sort last and equal a real (0xFFFFFFFF, 0xFFFFFFFF) record bit
// for bit, so the first n sorted entries are exactly the tile.  Ends with a
// barrier.
// How a build tile was sorted: direct = counting sort, whose key directory
// (relative to kmin) then answers probes.
struct TileSort {
    bool direct;
    uint32_t kmin;
};

template <bool SORT, bool DIRECT = false>
__device__ __forceinline__ TileSort gather_tile(uint64_t *s, uint64_t *tmp, uint32_t *s_mm, uint16_t *dir,
                                                const Rec *__restrict__ part, const PartView &v, int side,
                                                uint32_t t0, int n) {
    const int np2 = SORT ? sort_tile_size(n) : n;
    uint32_t mn = 0xFFFFFFFFu, mx = 0;
    for (int i0 = 0; i0 < np2; i0 += kJoinThreads * kBatch) {
        Rec r[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            const int i = i0 + k * kJoinThreads + threadIdx.x;
            if (i < n) r[k] = part[v.index(side, t0 + i)];
        }
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            const int i = i0 + k * kJoinThreads + threadIdx.x;
            if (i < np2) s[i] = i < n ? ((uint64_t(r[k].x) << 32) | r[k].y) : ~0ull;
            if (SORT && i < n) {
                mn = min(mn, r[k].x);
                mx = max(mx, r[k].x);
            }
        }
    }
    if (SORT) {
        // counters of the direct counting sort (used if the keys' span allows)
        if (DIRECT) {
            uint32_t *cnt = reinterpret_cast<uint32_t *>(tmp);
#pragma unroll
            for (int q = 0; q < int(kSmjDirect) / kJoinThreads; q++)
                cnt[q * kJoinThreads + threadIdx.x] = 0;
        }
        wave_minmax_full(mn, mx);
        constexpr int kW = kJoinThreads / kWave;
        if ((threadIdx.x & (kWave - 1)) == 0) {
            s_mm[threadIdx.x / kWave] = mn;
            s_mm[kW + threadIdx.x / kWave] = mx;
        }
    }
    __syncthreads();
    if (SORT) {
        constexpr int kW = kJoinThreads / kWave;
#pragma unroll
        for (int w = 0; w < kW; w++) {
            mn = min(mn, s_mm[w]);
            mx = max(mx, s_mm[kW + w]);
        }
        // block-uniform values (the compiler cannot see that on its own):
        // uniform branches keep the sort variants' registers apart
        mn = __builtin_amdgcn_readfirstlane(mn);
        const bool narrow = __builtin_amdgcn_readfirstlane(mx - mn) < (1u << 20);  // n >= 1: mn <= mx
        const int nu = __builtin_amdgcn_readfirstlane(n);
        const int np2u = __builtin_amdgcn_readfirstlane(np2);
        if (DIRECT && __builtin_amdgcn_readfirstlane(mx - mn) < kSmjDirect) {
            counting_sort_tile(s, reinterpret_cast<uint32_t *>(tmp), dir, s_mm + 2 * kW, nu, mn);
            return TileSort{true, mn};
        }
        if (np2u == kJoinThreads)
            sort_tile<1>(s, tmp, nu, narrow, mn);
        else if (np2u == 2 * kJoinThreads)
            sort_tile<2>(s, tmp, nu, narrow, mn);
        else
            sort_tile<4>(s, tmp, nu, narrow, mn);
    }
    return TileSort{false, 0u};
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

template <bool GROUPED>
struct SmjLds {
    uint64_t b[kSmjCap];                          // sorted build tile
    uint64_t p[kSmjCap];                          // probe tile (unsorted)
    uint32_t mm[3 * kJoinThreads / kWave];        // per-wave key bounds of a build tile + scan scratch
    uint16_t dir[kSmjDirect + 2];                 // counting-sorted tile: first position of each key
    uint64_t pa[GROUPED ? kSmjCap + 1 : 1];       // grouped: prefix sums of A over the sorted build tile
};

// Count / C1 / C2 of the un probe records in m.p against a build tile that
// the counting sort ordered (m.dir: first rank of each key offset from
// kmin).  Every lookup is done first, kept as (first rank | count << 16) per
// probe slot; after a barrier the key directory holds the waves' match
// queues, and the first two matches of every probe go through the queue so
// the checksum math runs on full waves (later matches, rare under uniform
// keys, are summed in place).  The directory is gone afterwards: for the
// last probe tile of a build tile.
template <bool C2>
__device__ __forceinline__ void smj_probe_queued(SmjLds<false> &m, int un, uint32_t kmin, Acc &acc) {
    constexpr int kSlots = kSmjCap / kJoinThreads;
    static_assert(sizeof(m.dir) >= (kJoinThreads / kWave) * sizeof(PairQ), "one queue per wave in the directory");
    static_assert(offsetof(SmjLds<false>, dir) % alignof(PairQ) == 0, "queues aligned in the directory");
    uint32_t lc[kSlots];
#pragma unroll
    for (int k = 0; k < kSlots; k++) {
        const int j = k * kJoinThreads + int(threadIdx.x);
        lc[k] = 0;
        if (j < un) {
            const uint32_t d = uint32_t(m.p[j] >> 32) - kmin;
            if (d < kSmjDirect) {
                const uint32_t lo = m.dir[d], hi = m.dir[d + 1];
                lc[k] = lo | ((hi - lo) << 16);
            }
        }
    }
    __syncthreads();  // every lookup done: the directory becomes the queues
    PairQ &q = reinterpret_cast<PairQ *>(m.dir)[threadIdx.x / kWave];
    uint32_t qn = 0;  // queued matches (wave-uniform)
#pragma unroll
    for (int k = 0; k < kSlots; k++) {
        const int j = k * kJoinThreads + int(threadIdx.x);
        const uint32_t lo = lc[k] & 0xFFFFu, cnt = lc[k] >> 16;
        uint32_t key = 0, sv = 0;
        if (cnt) {
            const uint64_t r = m.p[j];
            key = uint32_t(r >> 32);
            sv = uint32_t(r);
        }
        acc.count += cnt;
        for (uint32_t e = 2; e < cnt; e++) pair_acc<C2>(key, uint32_t(m.b[lo + e]), sv, acc);
        for (uint32_t e = 0; e < 2; e++) {
            const bool has = cnt > e;
            const uint64_t mask = __ballot(has);
            if (mask == 0) break;  // wave-uniform
            const uint32_t c = uint32_t(__popcll(mask));
            if (qn + c > uint32_t(kWave)) {
                pairq_drain<C2>(q, qn, acc);
                qn = 0;
            }
            if (has) {
                const uint32_t pos = qn + lanes_below(mask);
                q.key[pos] = key;
                q.sv[pos] = sv;
                q.rv[pos] = uint32_t(m.b[lo + e]);
            }
            qn += c;
        }
    }
    pairq_drain<C2>(q, qn, acc);
}

// Joins one partition (SMJ): each build tile is bitonic-sorted in LDS and
// every probe element is merged into it by co-ranking its key (lower /
// upper bound: the one-sided merge path), so the probe tile itself needs
// no sort.  GROUPED: count + C1 from prefix sums of A over the sorted tile;
// a grouped partition whose build side exceeds one tile (heavy keys) goes
// to the aggregated build instead, which is O(n) at any skew.
template <bool C2, bool WRITE, bool GROUPED, bool PER_WINDOW, bool EXPAND>
__device__ __forceinline__ void smj_join_part(const JoinIO &io, const PartView &v, SmjLds<GROUPED> &m,
                                              uint32_t *s_ws, unsigned long long *s_base, Acc &acc) {
    const uint32_t nb = v.n[0], np = v.n[1];
    const Rec *__restrict__ part = io.part;
    uint64_t *s_b = m.b, *s_p = m.p, *s_pa = m.pa;
    for (uint32_t t0 = 0; np && t0 < nb; t0 += kSmjCap) {
        const int tn = int(min<uint32_t>(kSmjCap, nb - t0));
        const TileSort ts = gather_tile<true, PER_WINDOW>(s_b, s_p, m.mm, m.dir, part, v, 0, t0, tn);
        if constexpr (GROUPED) {
            // prefix sums of A over the sorted build tile: each thread sums
            // kPer consecutive entries, then a block scan of the sums
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
            s_pa[threadIdx.x] = sum;
            __syncthreads();
            if (threadIdx.x == 0) {
                uint64_t run = 0;
                for (int t = 0; t < kJoinThreads; t++) {
                    const uint64_t x = s_pa[t];
                    s_pa[t] = run;
                    run += x;
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
        for (uint32_t u0 = 0; u0 < np; u0 += kSmjCap) {
            const int un = int(min<uint32_t>(kSmjCap, np - u0));
            gather_tile<false>(s_p, nullptr, nullptr, nullptr, part, v, 1, u0, un);
            if (!WRITE && !GROUPED && EXPAND) {
                // a wave in which some probe has more than two matches
                // expands its match lists cooperatively — lane g of each
                // round takes match g of the wave's concatenated lists — so
                // a hot key's list runs on full waves instead of one lane
                const uint32_t lane = threadIdx.x & (kWave - 1);
                for (int j0 = 0; j0 < int(un); j0 += kJoinThreads) {
                    const int j = j0 + int(threadIdx.x);
                    uint32_t k = 0, sv = 0;
                    int lo = 0, hi = 0;
                    if (j < int(un)) {
                        k = uint32_t(s_p[j] >> 32);
                        sv = uint32_t(s_p[j]);
                        lo = lower_bound_key(s_b, tn, k);
                        hi = upper_bound_key(s_b, lo, tn, k);
                    }
                    const uint32_t cnt = uint32_t(hi - lo);
                    if (__ballot(cnt > 2) == 0) {  // wave-uniform
                        for (int e = lo; e < hi; e++) {
                            const uint64_t a = hash_a(k, uint32_t(s_b[e]));
                            acc.count++;
                            acc.c1 += a * hash_b(k, sv);
                            if (C2) acc.c2 += pair_hash_from_a(a, sv);
                        }
                        continue;
                    }
                    uint32_t tot;
                    const uint32_t ex = wave_excl_scan(cnt, &tot);
                    for (uint32_t g0 = 0; g0 < tot; g0 += kWave) {
                        const uint32_t g = g0 + lane;
                        // owner: the last lane whose list starts at or before g
                        int olo = 0, ohi = kWave;
#pragma unroll
                        for (int st = 0; st < 6; st++) {
                            const int mid = (olo + ohi) >> 1;
                            if (uint32_t(__shfl(int(ex), mid)) <= g)
                                olo = mid;
                            else
                                ohi = mid;
                        }
                        const uint32_t e = uint32_t(__shfl(lo, olo)) + (g - uint32_t(__shfl(int(ex), olo)));
                        const uint32_t svo = uint32_t(__shfl(int(sv), olo));
                        if (g < tot) {
                            const uint64_t b = s_b[e];
                            const uint32_t ko = uint32_t(b >> 32);  // equals the probe's key
                            const uint64_t a = hash_a(ko, uint32_t(b));
                            acc.count++;
                            acc.c1 += a * hash_b(ko, svo);
                            if (C2) acc.c2 += pair_hash_from_a(a, svo);
                        }
                    }
                }
            } else if (!WRITE) {
                bool queued = false;
                if constexpr (PER_WINDOW && !GROUPED) {
                    if (ts.direct && u0 + uint32_t(kSmjCap) >= np) {  // block-uniform
                        smj_probe_queued<C2>(m, un, ts.kmin, acc);
                        queued = true;
                    }
                }
                for (int j = threadIdx.x; !queued && j < un; j += kJoinThreads) {
                    const uint32_t k = uint32_t(s_p[j] >> 32), sv = uint32_t(s_p[j]);
                    int lo, hi;
                    if (PER_WINDOW && ts.direct) {  // the counting sort's key directory
                        const uint32_t d = k - ts.kmin;
                        if (d >= kSmjDirect) continue;
                        lo = m.dir[d];
                        hi = m.dir[d + 1];
                    } else {
                        lo = lower_bound_key(s_b, tn, k);
                        hi = upper_bound_key(s_b, lo, tn, k);
                    }
                    if (hi == lo) continue;
                    if constexpr (GROUPED) {
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
                        *s_base = tot ? atomicAdd(io.cursor, (unsigned long long)tot) : 0ull;
                    __syncthreads();
                    if (hi > lo) {
                        uint64_t pos = *s_base + ex;
                        const sjb_tuple_t st = io.R[si];
                        for (int e = lo; e < hi; e++) {
                            const sjb_tuple_t lt = io.L[uint32_t(s_b[e])];
                            const uint64_t a = hash_a(lt.key, lt.val);
                            acc.count++;
                            acc.c1 += a * hash_b(st.key, st.val);
                            if (C2) acc.c2 += pair_hash_from_a(a, st.val);
                            write_row(io.out, pos++, io.out_cap, io.wrap, lt, st);
                        }
                    }
                    __syncthreads();
                }
            }
            __syncthreads();
        }
    }
}

// LDS bytes of each join body.
template <int ALG, bool GROUPED, bool WRITE>
struct JoinLdsBytes {
    static constexpr size_t hj = GROUPED ? sizeof(AggLds) : sizeof(HjLds<WRITE>);
    static constexpr size_t smj = GROUPED ? (sizeof(SmjLds<true>) > sizeof(AggLds) ? sizeof(SmjLds<true>) : sizeof(AggLds))
                                          : sizeof(SmjLds<false>);
    static constexpr size_t value = ALG == SJB_ALG_SMJ ? smj : hj;
};

// Joins partition view v with the body of (ALG, C2, WRITE, GROUPED).
// PER_WINDOW: the per-window join (k_join) — SMJ build tiles with narrow key
// spans use the counting sort (the persistent batched join keeps only the
// sorting networks, whose registers it can hold at three workgroups per CU);
// EXPAND (batched join only) runs long match lists on full waves: the HJ
// probe queues every match, the SMJ probe expands them cooperatively.
template <int ALG, bool C2, bool WRITE, bool GROUPED, bool PER_WINDOW, bool EXPAND = false>
__device__ __forceinline__ void join_part(const JoinIO &io, const PartView &v, int bshift, uint8_t *mem,
                                          uint32_t *s_ws, unsigned long long *s_base, Acc &acc) {
    if constexpr (ALG == SJB_ALG_HJ) {
        if constexpr (GROUPED)
            agg_join_part(io, v, bshift, *reinterpret_cast<AggLds *>(mem), acc);
        else
            hj_join_part<C2, WRITE, EXPAND>(io, v, bshift, *reinterpret_cast<HjLds<WRITE> *>(mem), s_ws,
                                                 s_base, acc);
    } else {
        if (GROUPED && v.n[0] > uint32_t(kSmjCap))
            agg_join_part(io, v, bshift, *reinterpret_cast<AggLds *>(mem), acc);
        else
            smj_join_part<C2, WRITE, GROUPED, PER_WINDOW, EXPAND>(io, v, *reinterpret_cast<SmjLds<GROUPED> *>(mem),
                                                                  s_ws, s_base,
                                                      acc);
    }
}

// Run-id table entries per side of a view (records past them are located
// by binary search): one build tile; grouped HJ keeps fewer so its
// workgroup fits three to a CU.
template <int ALG, bool GROUPED>
__host__ __device__ constexpr int idcap_of() {
    return ALG == SJB_ALG_SMJ ? (GROUPED ? kSmjCap : kSmjIdCap) : (GROUPED ? kAggIdCap : kHjCap);
}

// ---------------------------------------------------------------------------
// one window: one workgroup per partition
// ---------------------------------------------------------------------------
// Block sum of a workgroup's (count, C1, C2) with a single barrier (nothing
// runs after it): to partials[3p ...], or added to `direct` when it is set.
__device__ __forceinline__ void join_tail(Acc acc, uint64_t *s_red, uint64_t *__restrict__ partials, int p,
                                          sjb_window_result_t *__restrict__ direct) {
    acc.count = wave_sum_full(acc.count);
    acc.c1 = wave_sum_full(acc.c1);
    acc.c2 = wave_sum_full(acc.c2);
    constexpr int kW = kJoinThreads / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        const int w = threadIdx.x / kWave;
        s_red[w] = acc.count;
        s_red[kW + w] = acc.c1;
        s_red[2 * kW + w] = acc.c2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        uint64_t t = 0;
#pragma unroll
        for (int w = 0; w < kW; w++) t += s_red[threadIdx.x * kW + w];
        // sums mod 2^64: the atomics give the same bits in any order
        if (!direct)
            partials[3 * p + threadIdx.x] = t;
        else if (t)
            atomicAdd(reinterpret_cast<unsigned long long *>(direct) + threadIdx.x, (unsigned long long)t);
    }
}

// The workgroup's (count, C1, C2) goes to partials[3p ...] and is summed
// by k_reduce, or, when `direct` is set (the partition pass zeroed it), is
// added to the window result with three 64-bit atomics.
template <int ALG, bool C2, bool WRITE, bool GROUPED>
// Count bodies whose LDS allows three workgroups per CU are held to 80
// VGPRs (six waves per SIMD) to match.
__global__ __launch_bounds__(kJoinThreads)
__attribute__((amdgpu_waves_per_eu(WRITE || (ALG == SJB_ALG_SMJ && GROUPED) ? 1 : 6))) void k_join(
    JoinIO io, const uint32_t *__restrict__ offs,
                                                       int logF, int ncL, int ncR, uint32_t nL, int chunk,
                                                       int bshift, uint64_t *__restrict__ partials,
                                                       sjb_window_result_t *__restrict__ direct) {
    __shared__ __align__(16) uint8_t s_mem[JoinLdsBytes<ALG, GROUPED, WRITE>::value];
    __shared__ uint32_t s_ws[kJoinThreads / kWave];
    __shared__ uint64_t s_red[3 * kJoinThreads / kWave];
    __shared__ unsigned long long s_base;
    extern __shared__ uint32_t dyn[];

    const int p = xcd_partition(blockIdx.x, 1 << logF);
    const PartView v =
        ncL + ncR <= kJoinThreads && ncL % kWave == 0
            ? load_view_small(dyn, offs, logF, p, ncL, ncR, 0u, nL, uint32_t(chunk), idcap_of<ALG, GROUPED>(), s_ws)
            : load_view(dyn, offs, logF, p, ncL, ncR, 0u, nL, uint32_t(chunk), idcap_of<ALG, GROUPED>(), s_ws);
    Acc acc;
    acc.zero();
    join_part<ALG, C2, WRITE, GROUPED, true>(io, v, bshift, s_mem, s_ws, &s_base, acc);
    join_tail(acc, s_red, partials, p, direct);
}

// Count-path HJ / SMJ join of one window on a resident grid (the host
// launches it when the partition pass zeroed `direct` and load_view_small
// applies):
// workgroup b joins partitions xcd_partition(b), xcd_partition(b + G), ...
// (G a multiple of the 8 XCDs, so all of them on b's XCD), and loads the
// next partition's run bounds while the current one's records are gathered,
// so that global round trip leaves the per-partition chain.  Its
// (count, C1, C2) over all of them is added to the window result once.
template <int ALG, bool C2>
__global__ __launch_bounds__(kJoinThreads) __attribute__((amdgpu_waves_per_eu(6))) void k_join_pipe(
    JoinIO io, const uint32_t *__restrict__ offs, int logF, int ncL, int ncR, uint32_t nL, int chunk, int bshift,
    sjb_window_result_t *__restrict__ direct) {
    __shared__ __align__(16) uint8_t s_mem[JoinLdsBytes<ALG, false, false>::value];
    __shared__ uint32_t s_ws[kJoinThreads / kWave];
    __shared__ uint64_t s_red[3 * kJoinThreads / kWave];
    __shared__ unsigned long long s_base;
    extern __shared__ uint32_t dyn[];

    const int F = 1 << logF, G = int(gridDim.x), nq = ncL + ncR;
    int it = blockIdx.x;
    uint32_t a, b;
    view_bounds(offs, logF, xcd_partition(it, F), nq, a, b);
    Acc acc;
    acc.zero();
    for (; it < F; it += G) {
        __syncthreads();  // the previous partition's LDS is no longer read
        const PartView v =
            view_small_from(dyn, a, b, ncL, ncR, 0u, nL, uint32_t(chunk), idcap_of<ALG, false>(), s_ws);
        if (it + G < F) view_bounds(offs, logF, xcd_partition(it + G, F), nq, a, b);
        join_part<ALG, C2, false, false, true>(io, v, bshift, s_mem, s_ws, &s_base, acc);
    }
    join_tail(acc, s_red, nullptr, 0, direct);
}

// ---------------------------------------------------------------------------
// batched windows: one persistent launch over (window, partition) items
// ---------------------------------------------------------------------------
// The host deals the windows over kXcds queues (largest first, each to the
// least loaded queue) and orders every queue by decreasing window size.
// Workgroup b starts on queue b % kXcds — the XCD the hardware dispatched it
// to, so neighbouring partitions of a window (which share cache lines of
// every chunk's runs) are joined under one L2 — takes `grab` items per
// atomic step and, once its queue is empty, drains the others in turn.
// Every workgroup visits every queue once and exits when all are empty, so
// the grid always drains.  Each item adds its (count, C1, C2) to its
// window's result with three 64-bit atomics (sums mod 2^64: any order gives
// the same bits).
// HJ bodies fit three workgroups per CU in LDS; the register budget is
// capped to match (6 waves per SIMD -> at most 80 VGPRs).
template <int ALG, bool C2, bool GROUPED, bool EXPAND = false>
__global__ __attribute__((amdgpu_flat_work_group_size(1, kJoinThreads),
                          amdgpu_waves_per_eu(ALG == SJB_ALG_HJ ? SJB_BATCH_HJ_WAVES
                                              : GROUPED        ? 4
                                                               : SJB_BATCH_SMJ_WAVES))) void
k_join_batch(BatchArgs a) {
    __shared__ __align__(16) uint8_t s_mem[JoinLdsBytes<ALG, GROUPED, false>::value];
    __shared__ uint32_t s_ws[kJoinThreads / kWave];
    __shared__ uint64_t s_red[3 * kJoinThreads / kWave];
    __shared__ unsigned long long s_base;
    __shared__ uint32_t s_item;
    extern __shared__ uint32_t dyn[];

    JoinIO io;
    io.part = a.part;
    io.L = a.L;
    io.R = a.R;
    io.out = nullptr;
    io.out_cap = 0;
    io.wrap = 0;
    io.cursor = nullptr;
    const int home = blockIdx.x % kXcds;
    for (int qi = 0; qi < kXcds; qi++) {
        const int x = (home + qi) % kXcds;
        const uint32_t qlen = a.qbeg[x + 1] - a.qbeg[x];
        for (;;) {
            __syncthreads();
            if (threadIdx.x == 0) s_item = qlen ? atomicAdd(&a.qhead[x * 16], a.grab) : qlen;
            __syncthreads();
            // uniform: readfirstlane keeps the item and its window
            // descriptor in scalar registers
            const uint32_t i0 = __builtin_amdgcn_readfirstlane(s_item);
            if (i0 >= qlen) break;
            const uint32_t i1 = min(qlen, i0 + a.grab);
            for (uint32_t i = i0; i < i1; i++) {
                const uint32_t item = a.qbeg[x] + i;
                const BatchWin bw = a.win[a.item_win[item]];
                const int p = int(item - bw.item_base);
                const int logF = int(bw.logF);
                __syncthreads();  // the previous item's LDS is no longer read
                const PartView v = load_view(dyn, a.offs + bw.offs_base, logF, p, int(bw.ncL),
                                             int(bw.ncR), bw.lpart, bw.rpart, a.chunk, a.idcap, s_ws);
                Acc acc;
                acc.zero();
                const int bshift = ALG == SJB_ALG_HJ ? logF : 0;
                join_part<ALG, C2, false, GROUPED, false, EXPAND>(io, v, bshift, s_mem, s_ws, &s_base, acc);
                const Acc r = block_sum(acc, s_red);
                if (threadIdx.x == 0 && (r.count | r.c1 | r.c2)) {
                    sjb_window_result_t *o = a.res + bw.id;
                    atomicAdd(reinterpret_cast<unsigned long long *>(&o->count), (unsigned long long)r.count);
                    atomicAdd(reinterpret_cast<unsigned long long *>(&o->checksum1), (unsigned long long)r.c1);
                    atomicAdd(reinterpret_cast<unsigned long long *>(&o->checksum2), (unsigned long long)r.c2);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// partial reduction -> one sjb_window_result_t
// ---------------------------------------------------------------------------
// n_heavy > 0: also adds the peeled heavy keys' pairs (count cntL * cntR,
// C1 sumA * sumB per key; 4 u64 per key in heavy[]).
__global__ __launch_bounds__(1024) void k_reduce(const uint64_t *__restrict__ partials, int n,
                                                 const uint64_t *__restrict__ heavy, uint32_t n_heavy,
                                                 sjb_window_result_t *__restrict__ out) {
    __shared__ uint64_t s_red[3 * 1024 / kWave];
    Acc a;
    a.zero();
    constexpr int kU = 8;
    for (int i0 = 0; i0 < n; i0 += 1024 * kU) {
        uint64_t v[kU][3];
#pragma unroll
        for (int u = 0; u < kU; u++) {
            const int i = i0 + u * 1024 + threadIdx.x;
            v[u][0] = v[u][1] = v[u][2] = 0;
            if (i < n) {
                v[u][0] = partials[3 * i];
                v[u][1] = partials[3 * i + 1];
                v[u][2] = partials[3 * i + 2];
            }
        }
#pragma unroll
        for (int u = 0; u < kU; u++) {
            a.count += v[u][0];
            a.c1 += v[u][1];
            a.c2 += v[u][2];
        }
    }
    if (threadIdx.x < n_heavy) {
        const uint64_t *h = heavy + 4 * threadIdx.x;
        a.count += h[0] * h[2];
        a.c1 += h[1] * h[3];
    }
    const Acc r = block_sum(a, s_red);
    if (threadIdx.x == 0) {
        out->count = r.count;
        out->checksum1 = r.c1;
        out->checksum2 = r.c2;
        out->pairs = 0;
    }
}

int launch_reduce(sjb_ctx *c, int n_partials, sjb_window_result_t *d_out, uint32_t n_heavy,
                  const uint64_t *heavy) {
    k_reduce<<<1, 1024, 0, c->stream>>>(c->d_partials, n_partials, heavy ? heavy : c->d_heavy, n_heavy, d_out);
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

int join_idcap(int alg, bool grouped) {
    return alg == SJB_ALG_SMJ ? (grouped ? kSmjCap : kSmjIdCap) : (grouped ? kAggIdCap : kHjCap);
}

size_t join_view_bytes(int alg, bool grouped, int ncL, int ncR) {
    return part_view_bytes(ncL, ncR, join_idcap(alg, grouped));
}

template <int ALG>
static int launch_join_alg(sjb_ctx *c, const WindowPlan &p, const sjb_tuple_t *L, const sjb_tuple_t *R,
                           sjb_result_tuple_t *out, uint64_t out_cap, int wrap) {
    const int F = 1 << p.logF;
    const bool c2 = (p.flags & SJB_F_C2) != 0;
    const bool grouped = (p.flags & SJB_F_GROUPED) != 0 && p.payload == PAY_VAL;
    const int bshift = p.kind == PART_HASH ? p.logF : 0;
    const size_t dynb = part_view_bytes(p.ncL, p.ncR, grouped ? idcap_of<ALG, true>() : idcap_of<ALG, false>());
    JoinIO io;
    io.part = p.part;
    io.L = L;
    io.R = R;
    io.out = out;
    io.out_cap = out_cap;
    io.wrap = wrap;
    io.cursor = reinterpret_cast<unsigned long long *>(c->d_cursor);
    if (p.payload == PAY_VAL && !grouped && p.direct && p.ncL + p.ncR <= kJoinThreads && p.ncL % kWave == 0) {
        // resident grid: three workgroups per CU, a multiple of the 8 XCDs
        const int slots = c->n_cu > 0 ? SJB_JOIN_PIPE_WG_PER_CU * c->n_cu : F;
        const int G = F <= slots ? F : std::max(slots & ~7, 8);
        if (c2)
            k_join_pipe<ALG, true><<<G, kJoinThreads, dynb, c->stream>>>(io, p.offs, p.logF, p.ncL, p.ncR,
                                                                         uint32_t(p.nL), p.chunk, bshift, p.direct);
        else
            k_join_pipe<ALG, false><<<G, kJoinThreads, dynb, c->stream>>>(io, p.offs, p.logF, p.ncL, p.ncR,
                                                                          uint32_t(p.nL), p.chunk, bshift, p.direct);
        SJB_HIP_TRY(hipGetLastError());
        return SJB_OK;
    }
#define SJB_J(C2_, W_, G_)                                                                        \
    k_join<ALG, C2_, W_, G_><<<F, kJoinThreads, dynb, c->stream>>>(io, p.offs, p.logF, p.ncL,      \
                                                                 p.ncR, uint32_t(p.nL), p.chunk, bshift, \
                                                                 c->d_partials, p.direct)
    if (p.payload == PAY_IDX) {
        if (c2)
            SJB_J(true, true, false);
        else
            SJB_J(false, true, false);
    } else if (grouped) {
        SJB_J(false, false, true);
    } else if (c2) {
        SJB_J(true, false, false);
    } else {
        SJB_J(false, false, false);
    }
#undef SJB_J
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

int launch_hj_join(sjb_ctx *c, const WindowPlan &p, const sjb_tuple_t *L, const sjb_tuple_t *R,
                   sjb_result_tuple_t *out, uint64_t out_cap, int wrap) {
    return launch_join_alg<SJB_ALG_HJ>(c, p, L, R, out, out_cap, wrap);
}

int launch_smj_join(sjb_ctx *c, const WindowPlan &p, const sjb_tuple_t *L, const sjb_tuple_t *R,
                    sjb_result_tuple_t *out, uint64_t out_cap, int wrap) {
    return launch_join_alg<SJB_ALG_SMJ>(c, p, L, R, out, out_cap, wrap);
}

int launch_join_batch(hipStream_t s, int alg, uint32_t flags, const BatchArgs &a, int blocks,
                      size_t dyn) {
    const bool c2 = (flags & SJB_F_C2) != 0;
    const bool grouped = (flags & SJB_F_GROUPED) != 0;
#define SJB_JB(A_, C2_, G_) k_join_batch<A_, C2_, G_><<<blocks, kJoinThreads, dyn, s>>>(a)
    if (alg == SJB_ALG_HJ) {
        if (grouped)
            SJB_JB(SJB_ALG_HJ, false, true);
        else if (a.expand)
            c2 ? k_join_batch<SJB_ALG_HJ, true, false, true><<<blocks, kJoinThreads, dyn, s>>>(a)
               : k_join_batch<SJB_ALG_HJ, false, false, true><<<blocks, kJoinThreads, dyn, s>>>(a);
        else if (c2)
            SJB_JB(SJB_ALG_HJ, true, false);
        else
            SJB_JB(SJB_ALG_HJ, false, false);
    } else if (alg == SJB_ALG_SMJ) {
        if (grouped)
            SJB_JB(SJB_ALG_SMJ, false, true);
        else if (c2 && a.expand)
            k_join_batch<SJB_ALG_SMJ, true, false, true><<<blocks, kJoinThreads, dyn, s>>>(a);
        else if (c2)
            SJB_JB(SJB_ALG_SMJ, true, false);
        else
            SJB_JB(SJB_ALG_SMJ, false, false);
    } else {
        return SJB_E_UNSUPPORTED;
    }
#undef SJB_JB
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

}  // namespace sjb
