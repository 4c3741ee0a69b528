// pk.hip — persistent-kernel eager hash join.
//
// The reference's design (src/sjb/windowing/eager/executor/GPU/GPU_HJ_PK.cu:
// 10-249, not compiled there) keeps 1024-thread build and probe kernels
// alive, spinning on per-thread signal bytes in host-mapped memory, with the
// hash tables in pinned host memory, so every access crosses PCIe.
//
// Here a grid of G = kPkBlocksPerCu workgroups per CU stays resident; the
// window state is the device-resident chained table of eager.hip.  Posting a
// batch costs the host no runtime call:
//   host  : writes the batch's descriptor (operation, side, size, and the
//           device-side address of its tuples: the caller's page-locked
//           buffer, read in place, or a page-locked copy of a pageable one)
//           into a ring in coherent page-locked host memory, then the
//           doorbell `tail` next to it, and waits for the batch's result;
//   lead  : workgroup 0 polls the doorbell across the bus and forwards the
//           descriptor into device memory, then publishes one word,
//           `posted` = (sequence + 1) << 20 | slices;
//   grid  : every workgroup polls `posted` (relaxed loads, sleeps between
//           them); slice s (kPkSlice tuples) of the batch belongs to
//           workgroup s mod G, which probes the opposite side's chains and
//           appends the slice to its own side (different sides, so a batch's
//           probes and appends may interleave).  The host posts a batch only
//           after the previous one's result, so no workgroup is ever more
//           than one batch behind: a workgroup that finds a later batch
//           posted knows its own batch is complete and moves on;
//   finish: each workgroup with slices adds its partial (count, C1, C2) to
//           the slot, releases its writes at device scope and counts its
//           slices in; the one completing the last slice publishes the
//           result to coherent host memory, advances the side's size and
//           clears the slot;
//   exit  : the lead leaves after `idle_limit` polls without a post, or on
//           the stop word, recording the batch it was waiting for; every
//           other workgroup leaves when it reaches that batch.  A workgroup
//           never leaves with slices of a posted batch pending, so every
//           wave finishes; the host relaunches the grid on the next post.
#include <string.h>

#include <atomic>
#include <chrono>
#include <thread>

#include "sjb_device.hpp"
#include "sjb_internal.hpp"

namespace sjb {

constexpr int kPkThreads = 256;
#ifndef SJB_PK_ILP
#define SJB_PK_ILP 1
#endif
#ifndef SJB_PK_BLOCKS_PER_CU
#define SJB_PK_BLOCKS_PER_CU 1
#endif
constexpr int kPkIlp = SJB_PK_ILP;                  // tuples per thread per slice, chains walked together
constexpr uint32_t kPkSlice = kPkThreads * kPkIlp;  // tuples per claimed slice
constexpr int kPkBlocksPerCu = SJB_PK_BLOCKS_PER_CU;
constexpr uint32_t kPkClearSlice = kPkThreads * 16;  // head words cleared per CLEAR slice
constexpr uint32_t kPkMaxSlices = (1u << 20) - 1;    // slices per batch (20-bit field of `posted`)
constexpr int kPkQ = 8;                              // descriptor / batch-slot ring
constexpr uint32_t kPkNil = 0xFFFFFFFFu;
constexpr uint64_t kPkRunning = ~0ull;               // exit_at while the grid runs

enum PkOp : uint32_t { PK_BATCH = 1, PK_CLEAR = 2 };

struct PkDesc {
    uint32_t op;
    uint32_t side;           // 0 = left, 1 = right
    uint64_t n;
    const sjb_tuple_t *src;  // the batch's tuples (page-locked host memory)
    uint64_t seq;
};

struct PkHost {              // coherent page-locked host memory, written by the host
    uint64_t tail;           // descriptors [0, tail) posted
    uint64_t stop;           // non-zero: the grid leaves
    uint64_t pad[6];
    PkDesc desc[kPkQ];
};

// Per ring slot (device memory): slices completed, partial sums and the
// forwarded descriptor.
struct PkSlot {
    uint64_t done;
    uint64_t acc[3];
    uint32_t op, side;
    uint64_t n;
    const sjb_tuple_t *src;
    uint64_t pad;
};

struct PkCtl {               // device memory
    uint64_t posted;         // (batches forwarded) << 20 | slices of the last one
    uint64_t exit_at;        // the lead left waiting for this batch (kPkRunning: it runs)
    uint64_t n[2];           // state sizes (advanced by each batch's finisher)
    uint64_t pad[4];
    PkSlot slot[kPkQ];
};

struct PkRes {               // coherent page-locked host memory, written by the device
    uint64_t done;           // descriptors [0, done) are complete
    uint64_t count[kPkQ], c1[kPkQ], c2[kPkQ];
};

__host__ __device__ __forceinline__ uint32_t pk_slices(uint32_t op, uint64_t n, uint64_t nb) {
    const uint64_t units = op == PK_CLEAR ? (nb + kPkClearSlice - 1) / kPkClearSlice
                                          : (n + kPkSlice - 1) / kPkSlice;
    return uint32_t(units > 0 ? units : 1);  // an empty batch still has one (empty) slice to finish it
}

// Probes slice s of the batch against the opposite side and appends it to
// its own side; kPkIlp tuples per thread, their chains walked in lockstep.
__device__ __forceinline__ void pk_slice(uint32_t side, uint64_t n, const sjb_tuple_t *batch, uint32_t s,
                                         uint64_t base, sjb_tuple_t *tup0, sjb_tuple_t *tup1, uint32_t *next0,
                                         uint32_t *next1, uint32_t *head0, uint32_t *head1, uint32_t nb_mask,
                                         int c2, Acc &acc) {
    const int own = int(side), other = 1 - own;
    const sjb_tuple_t *otup = other ? tup1 : tup0;
    const uint32_t *onext = other ? next1 : next0;
    const uint32_t *ohead = other ? head1 : head0;
    sjb_tuple_t t[kPkIlp];
    uint32_t g[kPkIlp];
#pragma unroll
    for (int j = 0; j < kPkIlp; j++) {
        const uint64_t i = uint64_t(s) * kPkSlice + uint64_t(j) * kPkThreads + threadIdx.x;
        if (i < n) t[j] = batch[i];
    }
#pragma unroll
    for (int j = 0; j < kPkIlp; j++) {
        const uint64_t i = uint64_t(s) * kPkSlice + uint64_t(j) * kPkThreads + threadIdx.x;
        g[j] = i < n ? ohead[key_hash(t[j].key) & nb_mask] : kPkNil;
    }
    for (;;) {
        bool any = false;
        uint2 o[kPkIlp];
        uint32_t nx[kPkIlp];
#pragma unroll
        for (int j = 0; j < kPkIlp; j++) {
            if (g[j] != kPkNil) {
                o[j] = *reinterpret_cast<const uint2 *>(&otup[g[j]]);
                nx[j] = onext[g[j]];
            }
        }
#pragma unroll
        for (int j = 0; j < kPkIlp; j++) {
            if (g[j] == kPkNil) continue;
            if (o[j].x == t[j].key) {
                const uint32_t lv = own == 0 ? t[j].val : o[j].y, rv = own == 0 ? o[j].y : t[j].val;
                const uint64_t a = hash_a(t[j].key, lv);
                acc.count++;
                acc.c1 += a * hash_b(t[j].key, rv);
                if (c2) acc.c2 += pair_hash_from_a(a, rv);
            }
            g[j] = nx[j];
            any |= g[j] != kPkNil;
        }
        if (!any) break;
    }
    sjb_tuple_t *tt = own ? tup1 : tup0;
    uint32_t *tn = own ? next1 : next0;
    uint32_t *th = own ? head1 : head0;
#pragma unroll
    for (int j = 0; j < kPkIlp; j++) {
        const uint64_t i = uint64_t(s) * kPkSlice + uint64_t(j) * kPkThreads + threadIdx.x;
        if (i < n) {
            const uint32_t gg = uint32_t(base + i);
            tt[gg] = t[j];
            tn[gg] = atomicExch(&th[key_hash(t[j].key) & nb_mask], gg);
        }
    }
}

// The lead's wait for post `seq`: forwards it, or records that the grid
// leaves.  Returns true to work on `seq`.
__device__ __forceinline__ bool pk_lead_wait(PkCtl *ctl, const PkHost *host, uint64_t seq, uint64_t nb,
                                             uint64_t idle_limit) {
    uint64_t polls = 0;
    for (;;) {
        if (__hip_atomic_load(&host->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) > seq) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "one-as");  // the descriptor written before the doorbell
            const int q = int(seq % kPkQ);
            const PkDesc *hd = &host->desc[q];
            const uint32_t op = __hip_atomic_load(&hd->op, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            const uint32_t side = __hip_atomic_load(&hd->side, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            const uint64_t n = __hip_atomic_load(&hd->n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            const sjb_tuple_t *src = reinterpret_cast<const sjb_tuple_t *>(__hip_atomic_load(
                reinterpret_cast<const uint64_t *>(&hd->src), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM));
            PkSlot *sl = &ctl->slot[q];
            sl->op = op;
            sl->side = side;
            sl->n = n;
            sl->src = src;
            const uint64_t total = pk_slices(op, n, nb);
            __hip_atomic_store(&ctl->posted, ((seq + 1) << 20) | total, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            return true;
        }
        if (__hip_atomic_load(&host->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) || ++polls > idle_limit) {
            __hip_atomic_store(&ctl->exit_at, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(4);
    }
}

__global__ __launch_bounds__(kPkThreads) void k_pk_eager_hj(
    PkCtl *ctl, const PkHost *host, sjb_tuple_t *tup0, sjb_tuple_t *tup1, uint32_t *next0, uint32_t *next1,
    uint32_t *head0, uint32_t *head1, uint32_t nb_mask, uint64_t nb, PkRes *res, uint64_t seq,
    uint64_t idle_limit, int c2) {
    __shared__ uint32_t s_go;     // 0 = leave, 1 = batch `seq` posted, 2 = already complete
    __shared__ uint32_t s_total;  // slices of batch `seq`
    __shared__ uint32_t s_last;   // this workgroup completed the batch
    __shared__ uint64_t s_red[3 * kPkThreads / kWave];
    const bool lead = blockIdx.x == 0;
    const uint32_t G = gridDim.x;
    for (;;) {
        if (threadIdx.x == 0) {
            uint32_t go = 1;
            if (lead) go = pk_lead_wait(ctl, host, seq, nb, idle_limit) ? 1 : 0;
            while (go) {
                const uint64_t w = __hip_atomic_load(&ctl->posted, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t p = w >> 20;
                if (p == seq + 1) {
                    s_total = uint32_t(w & kPkMaxSlices);
                    break;
                }
                if (p > seq + 1) {  // a later batch is posted: batch `seq` is complete
                    go = 2;
                    break;
                }
                if (__hip_atomic_load(&ctl->exit_at, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= seq) {
                    go = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(8);
            }
            s_go = go;
        }
        __syncthreads();
        const uint32_t go = s_go;
        if (!go) break;
        const uint32_t total = s_total;
        if (go == 1 && blockIdx.x < total) {
            // this workgroup's slices: see the forwarded descriptor, the
            // batch in host memory and the previous batches' appends
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "one-as");
            const int q = int(seq % kPkQ);
            PkSlot *sl = &ctl->slot[q];
            const uint32_t op = sl->op, side = sl->side;
            const uint64_t n = sl->n;
            const sjb_tuple_t *src = sl->src;
            const uint64_t base =
                op == PK_BATCH ? __hip_atomic_load(&ctl->n[side], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
            Acc acc;
            acc.zero();
            uint32_t mine = 0;
            for (uint32_t s = blockIdx.x; s < total; s += G, mine++) {
                if (op == PK_CLEAR) {
                    for (uint64_t i = uint64_t(s) * kPkClearSlice + threadIdx.x;
                         i < min<uint64_t>(nb, uint64_t(s + 1) * kPkClearSlice); i += kPkThreads) {
                        head0[i] = kPkNil;
                        head1[i] = kPkNil;
                    }
                } else {
                    pk_slice(side, n, src, s, base, tup0, tup1, next0, next1, head0, head1, nb_mask, c2, acc);
                }
            }
            const Acc r = block_sum(acc, s_red);  // ends with a barrier: every wave's stores issued
            if (threadIdx.x == 0) {
                __hip_atomic_fetch_add(&sl->acc[0], r.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(&sl->acc[1], r.c1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(&sl->acc[2], r.c2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                // the appends and sums visible device-wide before the count
                const uint64_t old =
                    __hip_atomic_fetch_add(&sl->done, uint64_t(mine), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                s_last = old + mine == uint64_t(total);
            }
            __syncthreads();
            if (s_last && threadIdx.x == 0) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                const uint64_t cnt = __hip_atomic_load(&sl->acc[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t v1 = __hip_atomic_load(&sl->acc[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint64_t v2 = __hip_atomic_load(&sl->acc[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (op == PK_CLEAR) {
                    ctl->n[0] = 0;
                    ctl->n[1] = 0;
                } else {
                    ctl->n[side] = base + n;
                }
                // the slot is free for batch seq + kPkQ
                sl->done = 0;
                sl->acc[0] = sl->acc[1] = sl->acc[2] = 0;
                __hip_atomic_store(&res->count[q], cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(&res->c1[q], v1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(&res->c2[q], v2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(&res->done, seq + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
        __syncthreads();
        seq++;
    }
}

static int pk_launch(sjb_ctx *c, uint64_t seq) {
    // the grid runs until its lead leaves again
    *c->h_pk_word = kPkRunning;
    SJB_HIP_TRY(hipMemcpyAsync(&static_cast<PkCtl *>(c->d_pk_ctl)->exit_at, c->h_pk_word, sizeof(uint64_t),
                               hipMemcpyHostToDevice, c->pk_stream));
    const unsigned grid = unsigned(std::max(c->n_cu, 1) * kPkBlocksPerCu);
    k_pk_eager_hj<<<grid, kPkThreads, 0, c->pk_stream>>>(
        static_cast<PkCtl *>(c->d_pk_ctl), static_cast<const PkHost *>(c->h_pk_host), c->d_eager_tuples[0],
        c->d_eager_tuples[1], c->d_eager_next[0], c->d_eager_next[1], c->d_eager_head[0], c->d_eager_head[1],
        uint32_t(c->eager_nb - 1), c->eager_nb, static_cast<PkRes *>(c->h_pk_res), seq, c->pk_idle_limit,
        (c->opts.flags & SJB_F_C2) ? 1 : 0);
    SJB_HIP_TRY(hipGetLastError());
    c->pk_running = true;
    c->pk_launches++;
    return SJB_OK;
}

static int pk_setup(sjb_ctx *c) {
    if (c->pk_ready) return SJB_OK;
    int rc = eager_init(c);
    if (rc) return rc;
    c->pk_slot_cap = c->opts.batch_size;
    if (pk_slices(PK_BATCH, c->pk_slot_cap, c->eager_nb) > kPkMaxSlices ||
        pk_slices(PK_CLEAR, 0, c->eager_nb) > kPkMaxSlices)
        return SJB_E_CAPACITY;
    SJB_HIP_TRY(hipStreamCreateWithFlags(&c->pk_stream, hipStreamNonBlocking));
    SJB_HIP_TRY(hipStreamCreateWithFlags(&c->pk_copy, hipStreamNonBlocking));
    if ((rc = dev_alloc(c, &c->d_pk_ctl, sizeof(PkCtl))) ||
        (rc = host_alloc(c, (void **)&c->h_pk_slots, kPkQ * c->pk_slot_cap * sizeof(sjb_tuple_t))) ||
        (rc = host_alloc(c, (void **)&c->h_pk_word, sizeof(uint64_t))))
        return rc;
    // doorbell, descriptors and completion words: coherent memory both
    // sides read and write while the grid runs
    for (void **slot : {&c->h_pk_host, &c->h_pk_res}) {
        void *p = nullptr;
        const size_t bytes = slot == &c->h_pk_host ? sizeof(PkHost) : sizeof(PkRes);
        if (hipHostMalloc(&p, bytes, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess) {
            (void)hipGetLastError();
            return SJB_E_NOMEM;
        }
        c->host_allocs.push_back(p);
        memset(p, 0, bytes);
        *slot = p;
    }
    SJB_HIP_TRY(hipMemsetAsync(c->d_pk_ctl, 0, sizeof(PkCtl), c->pk_copy));
    SJB_HIP_TRY(hipStreamSynchronize(c->pk_copy));
    if (c->pk_idle_limit == 0) {
        const char *e = getenv("SJB_PK_IDLE_POLLS");
        c->pk_idle_limit = e ? strtoull(e, nullptr, 10) : 2000000ull;
    }
    c->pk_ready = true;
    return SJB_OK;
}

// Posts descriptor `seq` and waits for its completion, relaunching the grid
// when it has left.
static int pk_post(sjb_ctx *c, uint32_t op, uint32_t side, const sjb_tuple_t *batch, uint64_t n,
                   uint64_t *count, uint64_t *c1, uint64_t *c2) {
    const uint64_t seq = c->pk_seq;
    const int q = int(seq % kPkQ);
    const sjb_tuple_t *src = nullptr;
    if (n) {
        // page-locked batches are read in place; pageable ones through a
        // page-locked copy; the device reads through the runtime's
        // device-side address of that memory
        hipPointerAttribute_t a;
        const bool pinned = hipPointerGetAttributes(&a, batch) == hipSuccess && a.devicePointer != nullptr &&
                            (a.type == hipMemoryTypeHost || a.type == hipMemoryTypeDevice);
        if (!pinned) {
            (void)hipGetLastError();
            sjb_tuple_t *dst = c->h_pk_slots + uint64_t(q) * c->pk_slot_cap;
            memcpy(dst, batch, n * sizeof(sjb_tuple_t));
            SJB_HIP_TRY(hipPointerGetAttributes(&a, dst));
        }
        if (!a.devicePointer) return SJB_E_INVALID;
        // devicePointer may be the allocation's base (then hostPointer is
        // too) or the queried address itself (offset 0)
        const void *hbase = a.type == hipMemoryTypeHost ? a.hostPointer : a.devicePointer;
        const uint64_t off = pinned && hbase ? uint64_t(reinterpret_cast<const char *>(batch) -
                                                        static_cast<const char *>(hbase))
                                             : 0;
        src = reinterpret_cast<const sjb_tuple_t *>(static_cast<const char *>(a.devicePointer) + off);
    }
    volatile PkHost *host = static_cast<volatile PkHost *>(c->h_pk_host);
    host->desc[q].op = op;
    host->desc[q].side = side;
    host->desc[q].n = n;
    host->desc[q].src = src;
    host->desc[q].seq = seq;
    std::atomic_thread_fence(std::memory_order_release);
    host->tail = seq + 1;
    c->pk_seq = seq + 1;
    volatile PkRes *res = static_cast<volatile PkRes *>(c->h_pk_res);
    const auto t0 = std::chrono::steady_clock::now();
    auto last_check = t0;
    for (uint64_t spins = 0;; spins++) {
        if (res->done > seq) break;
        // an idle grid leaves by itself; relaunch it for this descriptor
        // (checked now and then: a stream query costs more than a poll)
        const auto now = std::chrono::steady_clock::now();
        if (!c->pk_running || now - last_check > std::chrono::microseconds(200)) {
            last_check = now;
            if (c->pk_running && hipStreamQuery(c->pk_stream) == hipSuccess) c->pk_running = false;
            if (!c->pk_running) {
                if (res->done > seq) break;
                int rc = pk_launch(c, seq);
                if (rc) return rc;
            }
        }
        if (now - t0 > std::chrono::seconds(20)) return SJB_E_TIMEOUT;
        if (spins > 64) std::this_thread::yield();
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    if (count) *count = res->count[q];
    if (c1) *c1 = res->c1[q];
    if (c2) *c2 = res->c2[q];
    return SJB_OK;
}

}  // namespace sjb

using namespace sjb;

extern "C" {

int sjb_pk_execute(sjb_ctx_t *c, const sjb_tuple_t *batch, uint64_t n, int is_left, sjb_stats_t *stats) {
    if (!c || !stats || (n && !batch)) return SJB_E_INVALID;
    memset(stats, 0, sizeof(*stats));
    SJB_HIP_TRY(hipSetDevice(c->device));
    int rc = pk_setup(c);
    if (rc) return rc;
    const int own = is_left ? 0 : 1;
    if (n > c->pk_slot_cap || c->pk_n[own] + n > c->eager_cap) return SJB_E_CAPACITY;
    const auto t0 = std::chrono::steady_clock::now();
    rc = pk_post(c, PK_BATCH, uint32_t(own), batch, n, &stats->result_count, &stats->checksum1,
                 &stats->checksum2);
    if (rc) return rc;
    c->pk_n[own] += n;
    stats->c2_valid = (c->opts.flags & SJB_F_C2) ? 1 : 0;
    if (!stats->c2_valid) stats->checksum2 = 0;
    stats->device_ns = uint64_t(
        std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count());
    return SJB_OK;
}

int sjb_pk_clear(sjb_ctx_t *c) {
    if (!c) return SJB_E_INVALID;
    if (!c->pk_ready) return SJB_OK;
    SJB_HIP_TRY(hipSetDevice(c->device));
    int rc = pk_post(c, PK_CLEAR, 0, nullptr, 0, nullptr, nullptr, nullptr);
    if (rc) return rc;
    c->pk_n[0] = c->pk_n[1] = 0;
    return SJB_OK;
}

int sjb_pk_stop(sjb_ctx_t *c) {
    if (!c) return SJB_E_INVALID;
    if (!c->pk_ready) return SJB_OK;
    SJB_HIP_TRY(hipSetDevice(c->device));
    volatile PkHost *host = static_cast<volatile PkHost *>(c->h_pk_host);
    host->stop = 1;
    SJB_HIP_TRY(hipStreamSynchronize(c->pk_stream));
    c->pk_running = false;
    host->stop = 0;
    return SJB_OK;
}

uint64_t sjb_pk_launches(sjb_ctx_t *c) { return c ? c->pk_launches : 0; }

}  // extern "C"
