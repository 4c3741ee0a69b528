// slide.hip — eager sliding-window join with device-resident state and a
// native serving loop (count + C1 of the current window after every event).
//
// The reference's eager operators join each incoming batch against the
// opposite side's window state and clear the state when a tumbling window
// closes (include/sjb/windowing/eager/CountBased.cuh:36-72,
// src/sjb/windowing/eager/executor/GPU/GPU_HJ.cu:78-120).  Here the window
// slides: after every event (one left batch and one right batch) it holds
// the last W tuples of each side, and the result of that window is
// maintained incrementally and exactly.  Count and C1 are sums over pairs,
// so with the state of each side kept as a group-by-key table
// (key -> number of tuples, sum of A for the left side or of B for the
// right side) the join of a batch against a side costs one lookup per
// batch tuple:  count += cnt_other(k),  C1 += hash_own(t) * sum_other(k).
// One event is four ordered steps (each a kernel; "vs" = probe against):
//
//   1. left tuples leaving the window:   result -= L_out vs R_old;  L -= L_out
//   2. right tuples leaving the window:  result -= R_out vs L_keep; R -= R_out
//   3. new left tuples:                  result += L_in vs R_keep;  L += L_in
//   4. new right tuples:                 result += R_in vs L_new;   R += R_in
//
// which is J(L_new, R_new) = J(L_old, R_old) - J(L_out, R_old)
//   - J(L_keep, R_out) + J(L_in, R_keep) + J(L_new, R_in).
// Tuples of each side also live in a device ring of W slots (the expiring
// ones are read back from it).  Removing a key leaves a slot with count 0;
// tables have SJB_SLIDE_SLOTS * W slots (8W: a rebuild, ~1 ms at W = 2^22,
// comes rarely enough to stay out of the 99th latency percentile — slots
// used since a rebuild grow with the distinct keys seen, ever more slowly
// under skew; with 4W, rebuilds every ~200 events put p99 at 1.2 ms), and
// when slots in use could pass half of them the table is rebuilt from the
// ring (a rebuilt table holds at most W keys).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

#include "sjb_device.hpp"
#include "sjb_internal.hpp"

namespace sjb {

struct SlideSlot {
    uint32_t key;  // 0xFFFFFFFF = empty (key 0xFFFFFFFF itself uses slot `cap`)
    uint32_t cnt;
    uint64_t sum;
};
static_assert(sizeof(SlideSlot) == 16, "slot");
constexpr uint32_t kSlideEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t slide_home(uint32_t k, int log_cap) {
    return uint32_t((uint64_t(k) * 0x9E3779B97F4A7C15ULL) >> (64 - log_cap));
}

// A whole slot in one 16-byte load (key, count and sum arrive together).
__device__ __forceinline__ SlideSlot slide_load(const SlideSlot *p) {
    const uint4 v = *reinterpret_cast<const uint4 *>(p);
    SlideSlot r;
    r.key = v.x;
    r.cnt = v.y;
    r.sum = uint64_t(v.z) | (uint64_t(v.w) << 32);
    return r;
}

// Count and sum of key k in table t (0 if absent), probing linearly from
// its home slot s, whose content `first` is already loaded.
__device__ __forceinline__ void slide_lookup_from(const SlideSlot *__restrict__ t, int log_cap, uint32_t k,
                                                  uint32_t s, SlideSlot first, uint32_t *cnt, uint64_t *sum) {
    const uint32_t cap = 1u << log_cap;
    SlideSlot cur = first;
    for (;;) {
        if (cur.key == k) {
            *cnt = cur.cnt;
            *sum = cur.sum;
            return;
        }
        if (cur.key == kSlideEmpty) {
            *cnt = 0;
            *sum = 0;
            return;
        }
        s = (s + 1) & (cap - 1);
        cur = slide_load(&t[s]);
    }
}

// Slot of key k in table t, inserted if absent (*fresh set when this call
// inserted it), probing linearly from its home slot s, whose key `first`
// is already loaded.
__device__ __forceinline__ uint32_t slide_find_insert_from(SlideSlot *__restrict__ t, int log_cap, uint32_t k,
                                                           uint32_t s, uint32_t first, bool *fresh) {
    const uint32_t cap = 1u << log_cap;
    uint32_t cur = first;
    for (;;) {
        if (cur == kSlideEmpty) {
            cur = atomicCAS(&t[s].key, kSlideEmpty, k);
            if (cur == kSlideEmpty) {
                *fresh = true;
                return s;
            }
        }
        if (cur == k) return s;
        s = (s + 1) & (cap - 1);
        cur = __hip_atomic_load(&t[s].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// One step for n tuples of side SIDE (0 left, 1 right) read from
// src[(base + i) % src_mod], one tuple per thread: each workgroup first
// groups its tuples by key in LDS (count and sum of the own hash per key),
// then per distinct key probes `other` (if PROBE: count += c * cnt_other,
// C1 += sum_own * sum_other, times `sign`) and adds `sign` x (c, sum_own) to
// the key's entry in `own` — one global lookup and one update per distinct
// key of the group instead of per tuple, so hot keys cost no atomic
// contention.  Tuples are also copied into ring[(dst + i) % W] (if ring).
// Sign is +1 or -1 (two's complement).
// Tuples (threads) per step workgroup: larger groups combine more updates
// of hot keys, smaller ones spread a batch over more CUs; 512 measured
// best for batches of 2^16 Zipf(1.0) tuples (1024: -10%, 256: -6%).
#ifndef SJB_SLIDE_SLOTS
#define SJB_SLIDE_SLOTS 8  // table slots per window tuple (a power of two)
#endif
static_assert(SJB_SLIDE_SLOTS >= 2 && SJB_SLIDE_SLOTS <= 8 && (SJB_SLIDE_SLOTS & (SJB_SLIDE_SLOTS - 1)) == 0,
              "slot indices stay 32-bit for windows up to 2^28");
#ifndef SJB_SLIDE_THREADS
#define SJB_SLIDE_THREADS 512
#endif
constexpr int kSlideThreads = SJB_SLIDE_THREADS;
constexpr int kSlideLds = 2 * kSlideThreads;  // LDS group slots, load <= 1/2
constexpr int kSlideLdsLog = kSlideThreads == 1024 ? 11 : kSlideThreads == 512 ? 10 : 9;
static_assert((1 << kSlideLdsLog) == kSlideLds, "slide LDS slots");
template <int SIDE, bool PROBE>
__global__ __launch_bounds__(kSlideThreads) void k_slide_step(
    const sjb_tuple_t *__restrict__ src, uint64_t base, uint64_t src_mod, uint64_t n,
    SlideSlot *__restrict__ own, const SlideSlot *__restrict__ other, int log_cap, uint64_t sign,
    uint32_t *__restrict__ used, sjb_tuple_t *__restrict__ ring, uint64_t dst, uint64_t W,
    uint64_t *__restrict__ acc) {
    __shared__ uint32_t s_key[kSlideLds + 1];  // slot kSlideLds: key 0xFFFFFFFF
    __shared__ uint32_t s_cnt[kSlideLds + 1];
    __shared__ uint64_t s_sum[kSlideLds + 1];
    __shared__ uint64_t s_red[2 * kSlideThreads / kWave];
    __shared__ uint16_t s_occ[kSlideLds + 1];  // occupied group slots, compacted
    __shared__ uint32_t s_ws[kSlideThreads / kWave];
    for (int j = threadIdx.x; j <= kSlideLds; j += kSlideThreads) {
        s_key[j] = kSlideEmpty;
        s_cnt[j] = 0;
        s_sum[j] = 0;
    }
    __syncthreads();
    const uint64_t i = uint64_t(blockIdx.x) * kSlideThreads + threadIdx.x;
    if (i < n) {
        const sjb_tuple_t t = src[(base + i) % src_mod];
        if (ring) ring[(dst + i) % W] = t;
        const uint64_t h = SIDE == 0 ? hash_a(t.key, t.val) : hash_b(t.key, t.val);
        uint32_t sl = kSlideLds;
        if (t.key != kSlideEmpty) {
            sl = (t.key * kHashMul) >> (32 - kSlideLdsLog);
            for (;;) {
                uint32_t cur = __hip_atomic_load(&s_key[sl], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (cur == kSlideEmpty) cur = atomicCAS(&s_key[sl], kSlideEmpty, t.key);
                if (cur == kSlideEmpty || cur == t.key) break;
                sl = (sl + 1) & (kSlideLds - 1);
            }
        }
        atomicAdd(&s_cnt[sl], 1u);
        atomicAdd(reinterpret_cast<unsigned long long *>(&s_sum[sl]), (unsigned long long)h);
    }
    __syncthreads();
    // compact the occupied slots, so each thread takes at most one distinct
    // key when the group has at most kSlideThreads of them (its global
    // lookup and update are one dependent chain, not two or three)
    {
        constexpr int kPerT = (kSlideLds + kSlideThreads) / kSlideThreads;  // slots per thread (incl. the extra)
        uint32_t occ = 0;
#pragma unroll
        for (int r = 0; r < kPerT; r++) {
            const int j = threadIdx.x * kPerT + r;
            occ += (j <= kSlideLds && s_cnt[j] != 0) ? 1u : 0u;
        }
        uint32_t tot;
        uint32_t pos = block_excl_scan_u32(occ, s_ws, &tot);
#pragma unroll
        for (int r = 0; r < kPerT; r++) {
            const int j = threadIdx.x * kPerT + r;
            if (j <= kSlideLds && s_cnt[j] != 0) s_occ[pos++] = uint16_t(j);
        }
        if (threadIdx.x == 0) s_ws[0] = tot;
    }
    __syncthreads();
    const uint32_t n_occ = s_ws[0];
    uint64_t cnt = 0, c1 = 0;
    uint32_t fresh_n = 0;
    for (uint32_t t = threadIdx.x; t < n_occ; t += kSlideThreads) {
        const int j = s_occ[t];
        const uint32_t cb = s_cnt[j];
        const uint32_t k = j == kSlideLds ? kSlideEmpty : s_key[j];
        const uint64_t sb = s_sum[j];
        uint32_t g;
        if (k == kSlideEmpty) {  // key 0xFFFFFFFF: the extra slot `cap` of both tables
            g = 1u << log_cap;
            if (PROBE) {
                cnt += uint64_t(cb) * other[g].cnt;
                c1 += sb * other[g].sum;
            }
        } else {
            // the first probe of both tables (same home slot) in flight at once
            const uint32_t s0 = slide_home(k, log_cap);
            SlideSlot fo{};
            if (PROBE) fo = slide_load(&other[s0]);
            const uint32_t fk = __hip_atomic_load(&own[s0].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (PROBE) {
                uint32_t oc;
                uint64_t os;
                slide_lookup_from(other, log_cap, k, s0, fo, &oc, &os);
                cnt += uint64_t(cb) * oc;
                c1 += sb * os;
            }
            bool fresh = false;
            g = slide_find_insert_from(own, log_cap, k, s0, fk, &fresh);
            fresh_n += fresh;
        }
        atomicAdd(&own[g].cnt, uint32_t(uint64_t(cb) * sign));
        atomicAdd(reinterpret_cast<unsigned long long *>(&own[g].sum), (unsigned long long)(sb * sign));
    }
    // new keys counted once per wave instead of one global atomic each
    const uint64_t fw = wave_sum_full(uint64_t(fresh_n));
    if ((threadIdx.x & (kWave - 1)) == 0 && fw) atomicAdd(used, uint32_t(fw));
    if (PROBE) {
        cnt = wave_sum_full(cnt * sign);
        c1 = wave_sum_full(c1 * sign);
        if ((threadIdx.x & (kWave - 1)) == 0) {
            s_red[threadIdx.x / kWave] = cnt;
            s_red[kSlideThreads / kWave + threadIdx.x / kWave] = c1;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t x = 0, y = 0;
            for (int w = 0; w < kSlideThreads / kWave; w++) {
                x += s_red[w];
                y += s_red[kSlideThreads / kWave + w];
            }
            if (x | y) {
                atomicAdd(reinterpret_cast<unsigned long long *>(acc), (unsigned long long)x);
                atomicAdd(reinterpret_cast<unsigned long long *>(acc + 1), (unsigned long long)y);
            }
        }
    }
}

__global__ void k_slide_clear(SlideSlot *__restrict__ t, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
        t[i].key = kSlideEmpty;
        t[i].cnt = 0;
        t[i].sum = 0;
    }
}

// Snapshot of the running result and the tables' used slots for the host.
__global__ void k_slide_snapshot(const uint64_t *__restrict__ acc, const uint32_t *__restrict__ used,
                                 uint64_t *__restrict__ out) {
    if (threadIdx.x == 0) {
        out[0] = acc[0];
        out[1] = acc[1];
        out[2] = 0;
        out[3] = (uint64_t(used[0]) << 32) | used[1];
    }
}

}  // namespace sjb

using namespace sjb;

struct sjb_slide {
    sjb_ctx *c = nullptr;
    uint64_t W = 0, max_batch = 0;
    int log_cap = 0;
    sjb_tuple_t *d_ring[2] = {};
    SlideSlot *d_table[2] = {};
    uint32_t *d_used = nullptr;         // [2]
    uint64_t *d_acc = nullptr;          // [2]: count, C1
    uint64_t *d_snap = nullptr;         // snapshots, 4 u64 per event slot
    sjb_tuple_t *d_in[2][2] = {};       // staging [parity][side]
    hipStream_t copy = nullptr;
    hipEvent_t copied[2] = {}, consumed[2] = {};
    uint64_t n_total[2] = {0, 0};       // tuples pushed per side
    // Table fill tracking: slots in use are at most snap_used (from the
    // latest completed snapshot) plus the tuples inserted since then
    // (added - snap_added); `epoch` counts rebuilds so that a snapshot
    // taken before a rebuild is not used after it.
    uint64_t added[2] = {0, 0}, snap_used[2] = {0, 0}, snap_added[2] = {0, 0}, epoch[2] = {0, 0};
    uint64_t events = 0, rebuilds = 0;
    uint64_t bytes = 0;
    std::vector<void *> allocs;
};

namespace sjb {

static void slide_free(sjb_slide *s) {
    if (!s) return;
    for (hipEvent_t e : {s->copied[0], s->copied[1], s->consumed[0], s->consumed[1]})
        if (e) (void)hipEventDestroy(e);
    if (s->copy) (void)hipStreamDestroy(s->copy);
    for (void *p : s->allocs) (void)hipFree(p);
    if (s->c) s->c->ext_bytes -= s->bytes;
    delete s;
}

static int slide_alloc(sjb_slide *s, void **p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return SJB_E_NOMEM;
    }
    s->allocs.push_back(*p);
    return SJB_OK;
}

static int launch_step(sjb_slide *s, int side, bool probe, const sjb_tuple_t *src, uint64_t base,
                       uint64_t src_mod, uint64_t n, int64_t sign, bool to_ring) {
    if (n == 0) return SJB_OK;
    hipStream_t st = s->c->stream;
    const uint64_t sg = sign > 0 ? 1ull : ~0ull;
    SlideSlot *own = s->d_table[side];
    const SlideSlot *other = s->d_table[1 - side];
    sjb_tuple_t *ring = to_ring ? s->d_ring[side] : nullptr;
    const uint64_t dst = s->n_total[side] % s->W;
    const unsigned blocks = unsigned((n + kSlideThreads - 1) / kSlideThreads);
#define SJB_STEP(S_, P_)                                                                                   \
    k_slide_step<S_, P_><<<blocks, kSlideThreads, 0, st>>>(src, base, src_mod, n, own, other, s->log_cap, \
                                                           sg, s->d_used + side, ring, dst, s->W, s->d_acc)
    if (side == 0) {
        if (probe) SJB_STEP(0, true); else SJB_STEP(0, false);
    } else {
        if (probe) SJB_STEP(1, true); else SJB_STEP(1, false);
    }
#undef SJB_STEP
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

// Rebuilds one side's table from the tuples in its ring.
static int slide_rebuild(sjb_slide *s, int side) {
    hipStream_t st = s->c->stream;
    const uint64_t slots = (1ull << s->log_cap) + 1;
    k_slide_clear<<<1024, 256, 0, st>>>(s->d_table[side], slots);
    SJB_HIP_TRY(hipGetLastError());
    int rc = launch_fill_u32(st, s->d_used + side, 0u, 1);
    if (rc) return rc;
    const uint64_t live = std::min(s->n_total[side], s->W);
    const uint64_t first = s->n_total[side] - live;
    rc = launch_step(s, side, false, s->d_ring[side], first % s->W, s->W, live, +1, false);
    if (rc) return rc;
    s->snap_used[side] = live;  // upper bound until the next snapshot
    s->snap_added[side] = s->added[side];
    s->epoch[side]++;
    s->rebuilds++;
    return SJB_OK;
}

// Takes the used-slot counts of a completed snapshot (packed as in
// k_slide_snapshot) that was enqueued when added[] was `added_at` and the
// rebuild epochs were `epoch_at`.
static void slide_note_snapshot(sjb_slide *s, uint64_t packed, const uint64_t *added_at, const uint64_t *epoch_at) {
    for (int side = 0; side < 2; side++) {
        if (epoch_at[side] != s->epoch[side] || added_at[side] < s->snap_added[side]) continue;
        s->snap_used[side] = side == 0 ? packed >> 32 : packed & 0xFFFFFFFFu;
        s->snap_added[side] = added_at[side];
    }
}

// Enqueues one event whose batches are already in device memory (L, R);
// the snapshot of the result lands in snap (4 u64, device-accessible).
static int slide_event(sjb_slide *s, const sjb_tuple_t *L, uint64_t nl, const sjb_tuple_t *R, uint64_t nr,
                       uint64_t *snap) {
    // rebuild a table when its used slots (last snapshot + this event's
    // insertions) could pass half of the table
    const uint64_t half = (1ull << s->log_cap) / 2;
    for (int side = 0; side < 2; side++) {
        const uint64_t add = side == 0 ? nl : nr;
        if (s->snap_used[side] + (s->added[side] - s->snap_added[side]) + add > half) {
            int rc = slide_rebuild(s, side);
            if (rc) return rc;
        }
        s->added[side] += add;
    }
    // the window keeps the last W tuples of each side: the oldest ones
    // leave when the current window plus the batch exceeds W
    const uint64_t W = s->W;
    const uint64_t curL = std::min(s->n_total[0], W), curR = std::min(s->n_total[1], W);
    const uint64_t outL = curL + nl > W ? curL + nl - W : 0;
    const uint64_t outR = curR + nr > W ? curR + nr - W : 0;
    const uint64_t oldL = (s->n_total[0] - curL) % W;
    const uint64_t oldR = (s->n_total[1] - curR) % W;
    int rc;
    if ((rc = launch_step(s, 0, true, s->d_ring[0], oldL, W, outL, -1, false))) return rc;
    if ((rc = launch_step(s, 1, true, s->d_ring[1], oldR, W, outR, -1, false))) return rc;
    if ((rc = launch_step(s, 0, true, L, 0, ~0ull, nl, +1, true))) return rc;
    s->n_total[0] += nl;
    if ((rc = launch_step(s, 1, true, R, 0, ~0ull, nr, +1, true))) return rc;
    s->n_total[1] += nr;
    k_slide_snapshot<<<1, 64, 0, s->c->stream>>>(s->d_acc, s->d_used, snap);
    SJB_HIP_TRY(hipGetLastError());
    s->events++;
    return SJB_OK;
}

// Copies one event's host batches into staging buffer `parity` on the copy
// stream (after the previous user of that buffer finished) and makes the
// ctx stream wait for the copy.
static int slide_stage(sjb_slide *s, int parity, const sjb_tuple_t *L, uint64_t nl, const sjb_tuple_t *R,
                       uint64_t nr) {
    SJB_HIP_TRY(hipStreamWaitEvent(s->copy, s->consumed[parity], 0));
    if (nl)
        SJB_HIP_TRY(hipMemcpyAsync(s->d_in[parity][0], L, nl * sizeof(sjb_tuple_t), hipMemcpyHostToDevice, s->copy));
    if (nr)
        SJB_HIP_TRY(hipMemcpyAsync(s->d_in[parity][1], R, nr * sizeof(sjb_tuple_t), hipMemcpyHostToDevice, s->copy));
    SJB_HIP_TRY(hipEventRecord(s->copied[parity], s->copy));
    SJB_HIP_TRY(hipStreamWaitEvent(s->c->stream, s->copied[parity], 0));
    return SJB_OK;
}

static int slide_push_enqueue(sjb_slide *s, const sjb_tuple_t *L, uint64_t nl, const sjb_tuple_t *R, uint64_t nr,
                              uint64_t *snap) {
    const int parity = int(s->events & 1);
    int rc = slide_stage(s, parity, L, nl, R, nr);
    if (rc) return rc;
    rc = slide_event(s, s->d_in[parity][0], nl, s->d_in[parity][1], nr, snap);
    if (rc) return rc;
    SJB_HIP_TRY(hipEventRecord(s->consumed[parity], s->c->stream));
    return SJB_OK;
}

}  // namespace sjb

extern "C" {

int sjb_slide_create(sjb_ctx_t *c, uint64_t window, uint64_t max_batch, sjb_slide_t **out) {
    if (!c || !out || window == 0 || max_batch == 0 || max_batch > window || window > (1ull << 28))
        return SJB_E_INVALID;
    *out = nullptr;
    SJB_HIP_TRY(hipSetDevice(c->device));
    int log_cap = 10;
    while ((1ull << log_cap) < uint64_t(SJB_SLIDE_SLOTS) * window) log_cap++;
    const uint64_t slots = (1ull << log_cap) + 1;
    const uint64_t bytes = 2 * (window * sizeof(sjb_tuple_t) + slots * sizeof(SlideSlot)) +
                           4 * max_batch * sizeof(sjb_tuple_t) + 4096;
    if (c->opts.hbm_budget_bytes && c->bytes + c->ext_bytes + bytes > c->opts.hbm_budget_bytes)
        return SJB_E_BUDGET;
    sjb_slide *s = new sjb_slide();
    s->c = c;
    s->W = window;
    s->max_batch = max_batch;
    s->log_cap = log_cap;
    int rc = SJB_OK;
    for (int side = 0; side < 2 && !rc; side++) {
        if ((rc = slide_alloc(s, (void **)&s->d_ring[side], window * sizeof(sjb_tuple_t)))) break;
        rc = slide_alloc(s, (void **)&s->d_table[side], slots * sizeof(SlideSlot));
    }
    for (int p = 0; p < 2 && !rc; p++)
        for (int side = 0; side < 2 && !rc; side++)
            rc = slide_alloc(s, (void **)&s->d_in[p][side], max_batch * sizeof(sjb_tuple_t));
    if (!rc) rc = slide_alloc(s, (void **)&s->d_used, 64);
    if (!rc) rc = slide_alloc(s, (void **)&s->d_acc, 64);
    if (!rc) rc = slide_alloc(s, (void **)&s->d_snap, 64);
    if (!rc && hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) != hipSuccess) rc = SJB_E_HIP;
    for (int p = 0; p < 2 && !rc; p++)
        if (hipEventCreateWithFlags(&s->copied[p], hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&s->consumed[p], hipEventDisableTiming) != hipSuccess)
            rc = SJB_E_HIP;
    if (rc) {
        s->c = nullptr;
        slide_free(s);
        return rc;
    }
    s->bytes = bytes;
    c->ext_bytes += bytes;
    for (int p = 0; p < 2; p++) (void)hipEventRecord(s->consumed[p], c->stream);
    if ((rc = sjb_slide_reset(s))) {
        slide_free(s);
        return rc;
    }
    *out = s;
    return SJB_OK;
}

int sjb_slide_reset(sjb_slide_t *s) {
    if (!s) return SJB_E_INVALID;
    SJB_HIP_TRY(hipSetDevice(s->c->device));
    hipStream_t st = s->c->stream;
    const uint64_t slots = (1ull << s->log_cap) + 1;
    for (int side = 0; side < 2; side++) k_slide_clear<<<1024, 256, 0, st>>>(s->d_table[side], slots);
    SJB_HIP_TRY(hipGetLastError());
    int rc = launch_fill_u64(st, s->d_acc, 0, 8);
    if (!rc) rc = launch_fill_u32(st, s->d_used, 0u, 16);
    if (rc) return rc;
    SJB_HIP_TRY(hipStreamSynchronize(st));
    for (int side = 0; side < 2; side++) {
        s->n_total[side] = 0;
        s->added[side] = s->snap_used[side] = s->snap_added[side] = 0;
        s->epoch[side]++;
    }
    s->events = 0;
    return SJB_OK;
}

int sjb_slide_push(sjb_slide_t *s, const sjb_tuple_t *left, uint64_t n_left, const sjb_tuple_t *right,
                   uint64_t n_right, sjb_window_result_t *result) {
    if (!s || !result || n_left > s->max_batch || n_right > s->max_batch) return SJB_E_INVALID;
    if ((n_left && !left) || (n_right && !right)) return SJB_E_INVALID;
    SJB_HIP_TRY(hipSetDevice(s->c->device));
    int rc = slide_push_enqueue(s, left, n_left, right, n_right, s->d_snap);
    if (rc) return rc;
    uint64_t snap[4];
    SJB_HIP_TRY(hipMemcpyAsync(snap, s->d_snap, sizeof(snap), hipMemcpyDeviceToHost, s->c->stream));
    SJB_HIP_TRY(hipStreamSynchronize(s->c->stream));
    slide_note_snapshot(s, snap[3], s->added, s->epoch);
    result->count = snap[0];
    result->checksum1 = snap[1];
    result->checksum2 = 0;
    result->pairs = 0;
    return SJB_OK;
}

// Serving loop: event i (left batch h_left[left_off[i], left_off[i+1]),
// right batch likewise) arrives arrival_s[i] seconds after the call starts;
// it is enqueued as soon as it has arrived and the previous event has been
// enqueued.  latency_s[i] = time its result became host-visible minus its
// arrival (completion observed by a second host thread).  Host batches
// should be pinned for asynchronous copies.
int sjb_slide_serve(sjb_slide_t *s, const sjb_tuple_t *h_left, const uint64_t *left_off,
                    const sjb_tuple_t *h_right, const uint64_t *right_off, uint32_t n_events,
                    const double *arrival_s, double *latency_s, sjb_window_result_t *results,
                    double *elapsed_s) {
    if (!s || !left_off || !right_off || !arrival_s || !latency_s || !results) return SJB_E_INVALID;
    for (uint32_t i = 0; i < n_events; i++) {
        if (left_off[i + 1] < left_off[i] || right_off[i + 1] < right_off[i]) return SJB_E_INVALID;
        if (left_off[i + 1] - left_off[i] > s->max_batch || right_off[i + 1] - right_off[i] > s->max_batch)
            return SJB_E_INVALID;
    }
    if (n_events == 0) return SJB_OK;
    SJB_HIP_TRY(hipSetDevice(s->c->device));
    void *h_snap = nullptr;
    if (hipHostMalloc(&h_snap, size_t(n_events) * 32, hipHostMallocDefault) != hipSuccess) return SJB_E_NOMEM;
    uint64_t *snaps = static_cast<uint64_t *>(h_snap);
    std::vector<hipEvent_t> done(n_events, nullptr);
    int rc = SJB_OK;
    for (uint32_t i = 0; i < n_events && !rc; i++)
        if (hipEventCreateWithFlags(&done[i], hipEventDisableTiming) != hipSuccess) rc = SJB_E_HIP;
    using clk = std::chrono::steady_clock;
    std::vector<double> done_at(n_events, 0.0);
    std::vector<uint64_t> added_at(2 * size_t(n_events)), epoch_at(2 * size_t(n_events));
    std::atomic<uint32_t> enqueued{0};
    std::atomic<bool> failed{false};
    const auto t0 = clk::now();
    auto secs = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };
    std::thread waiter([&] {
        (void)hipSetDevice(s->c->device);
        for (uint32_t i = 0; i < n_events; i++) {
            while (enqueued.load(std::memory_order_acquire) <= i) {
                if (failed.load()) return;
                std::this_thread::yield();
            }
            if (hipEventSynchronize(done[i]) != hipSuccess) {
                failed.store(true);
                return;
            }
            done_at[i] = secs();
        }
    });
    for (uint32_t i = 0; i < n_events && !rc; i++) {
        while (secs() < arrival_s[i]) {
        }
        rc = slide_push_enqueue(s, h_left + left_off[i], left_off[i + 1] - left_off[i], h_right + right_off[i],
                                right_off[i + 1] - right_off[i], s->d_snap + 4 * (i & 1));
        if (!rc && hipMemcpyAsync(snaps + 4 * size_t(i), s->d_snap + 4 * (i & 1), 32, hipMemcpyDeviceToHost,
                                  s->c->stream) != hipSuccess)
            rc = SJB_E_HIP;
        if (!rc && hipEventRecord(done[i], s->c->stream) != hipSuccess) rc = SJB_E_HIP;
        if (!rc) {
            for (int side = 0; side < 2; side++) {
                added_at[2 * size_t(i) + side] = s->added[side];
                epoch_at[2 * size_t(i) + side] = s->epoch[side];
            }
            // used slots of the previous event if it is done (one event of
            // lag): keeps the rebuild decisions tight
            if (i > 0 && hipEventQuery(done[i - 1]) == hipSuccess)
                slide_note_snapshot(s, snaps[4 * size_t(i - 1) + 3], &added_at[2 * size_t(i - 1)],
                                    &epoch_at[2 * size_t(i - 1)]);
            enqueued.store(i + 1, std::memory_order_release);
        }
    }
    if (rc) failed.store(true);
    waiter.join();
    if (!rc && failed.load()) rc = SJB_E_HIP;
    if (hipStreamSynchronize(s->c->stream) != hipSuccess && !rc) rc = SJB_E_HIP;
    if (elapsed_s) *elapsed_s = secs();
    if (!rc) {
        for (uint32_t i = 0; i < n_events; i++) {
            latency_s[i] = done_at[i] - arrival_s[i];
            results[i].count = snaps[4 * size_t(i)];
            results[i].checksum1 = snaps[4 * size_t(i) + 1];
            results[i].checksum2 = 0;
            results[i].pairs = 0;
        }
        slide_note_snapshot(s, snaps[4 * size_t(n_events - 1) + 3], &added_at[2 * size_t(n_events - 1)],
                            &epoch_at[2 * size_t(n_events - 1)]);
    }
    for (hipEvent_t e : done)
        if (e) (void)hipEventDestroy(e);
    (void)hipHostFree(h_snap);
    return rc;
}

int sjb_slide_info(sjb_slide_t *s, sjb_slide_info_t *info) {
    if (!s || !info) return SJB_E_INVALID;
    info->window = s->W;
    info->max_batch = s->max_batch;
    info->table_slots = 1ull << s->log_cap;
    info->device_bytes = s->bytes;
    info->events = s->events;
    info->rebuilds = s->rebuilds;
    info->tuples_left = s->n_total[0];
    info->tuples_right = s->n_total[1];
    return SJB_OK;
}

int sjb_slide_destroy(sjb_slide_t *s) {
    if (!s) return SJB_OK;
    (void)hipSetDevice(s->c->device);
    (void)hipStreamSynchronize(s->c->stream);
    (void)hipStreamSynchronize(s->copy);
    slide_free(s);
    return SJB_OK;
}

}  // extern "C"
