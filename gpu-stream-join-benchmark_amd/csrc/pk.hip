// pk.hip — persistent-kernel eager hash join.
//
// The reference's design reference (src/sjb/windowing/eager/executor/GPU/
// GPU_HJ_PK.cu:10-249, not compiled there) keeps 1024-thread build and probe
// kernels alive, spinning on per-thread signal bytes in host-mapped memory,
// with the hash tables in pinned host memory, so every access crosses PCIe.
//
// Here ONE resident workgroup (1024 threads) serves a device-memory work
// queue:
//   host  : copies the batch into a device slot and the descriptor into the
//           descriptor ring on a copy stream, then bumps `tail` with
//           hipStreamWriteValue64 on the same stream (ordered after the
//           copies);
//   kernel: thread 0 polls `tail` with system-scope acquire loads
//           (s_sleep between polls), the workgroup probes the opposite
//           side's device-resident state and appends the batch to its own
//           side (the same state layout as eager.hip), then thread 0 writes
//           the result into coherent pinned host memory and publishes
//           `done` with a system-scope release store;
//   exit  : a stop word, or `idle_limit` polls without work (the host then
//           relaunches on the next batch), so every wave always finishes.
#include <string.h>

#include <atomic>
#include <chrono>
#include <thread>

#include "sjb_device.hpp"
#include "sjb_internal.hpp"

namespace sjb {

constexpr int kPkThreads = 1024;
#ifndef SJB_PK_ILP
#define SJB_PK_ILP 6
#endif
constexpr int kPkIlp = SJB_PK_ILP;  // batch tuples per thread whose memory chains are in flight together
constexpr int kPkQ = 8;  // descriptor / batch-slot ring
constexpr uint32_t kPkNil = 0xFFFFFFFFu;

enum PkOp : uint32_t { PK_BATCH = 1, PK_CLEAR = 2 };

struct PkDesc {
    uint32_t op;
    uint32_t side;  // 0 = left, 1 = right
    uint64_t n;
    uint64_t seq;
    uint64_t pad;
};

struct PkCtl {            // device memory
    uint64_t tail;        // descriptors [0, tail) have been posted
    uint64_t stop;        // non-zero: exit
    uint64_t n[2];        // state sizes (kernel-maintained)
};

struct PkRes {            // coherent pinned host memory
    uint64_t done;        // descriptors [0, done) are complete
    uint64_t exited;      // seq at which the kernel last exited (+1)
    uint64_t count[kPkQ], c1[kPkQ], c2[kPkQ];
};

__global__ __launch_bounds__(kPkThreads) void k_pk_eager_hj(
    PkCtl *ctl, const PkDesc *descs, const sjb_tuple_t *slots, uint64_t slot_cap, sjb_tuple_t *tup0,
    sjb_tuple_t *tup1, uint32_t *next0, uint32_t *next1, uint32_t *head0, uint32_t *head1, uint32_t nb_mask,
    uint64_t nb, PkRes *res, uint64_t seq, uint64_t idle_limit, int c2) {
    __shared__ uint64_t s_cmd;  // 0 = work on `seq`, 1 = exit
    __shared__ uint64_t s_n[2]; // state sizes (mirrored to ctl->n for a relaunch)
    __shared__ uint64_t s_red[3 * kPkThreads / kWave];
    if (threadIdx.x == 0) {
        s_n[0] = ctl->n[0];
        s_n[1] = ctl->n[1];
    }
    __syncthreads();
    for (;;) {
        if (threadIdx.x == 0) {
            uint64_t polls = 0, cmd = 1;
            for (;;) {
                const uint64_t tail = __hip_atomic_load(&ctl->tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
                if (tail > seq) {
                    cmd = 0;
                    break;
                }
                if (__hip_atomic_load(&ctl->stop, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM)) break;
                if (++polls > idle_limit) break;
                __builtin_amdgcn_s_sleep(8);
            }
            s_cmd = cmd;
        }
        __syncthreads();
        const uint64_t cmd = s_cmd;
        __threadfence_system();
        if (cmd) break;
        const PkDesc d = descs[seq % kPkQ];
        Acc acc;
        acc.zero();
        if (d.op == PK_CLEAR) {
            for (uint64_t i = threadIdx.x; i < nb; i += kPkThreads) {
                head0[i] = kPkNil;
                head1[i] = kPkNil;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                s_n[0] = s_n[1] = 0;
                ctl->n[0] = ctl->n[1] = 0;
            }
        } else {
            const sjb_tuple_t *batch = slots + (seq % kPkQ) * slot_cap;
            const int own = int(d.side), other = 1 - own;
            const sjb_tuple_t *otup = other ? tup1 : tup0;
            const uint32_t *onext = other ? next1 : next0;
            const uint32_t *ohead = other ? head1 : head0;
            // probe the opposite side: kPkIlp tuples per thread at a time,
            // their chains walked in lockstep (all loads of a step in flight);
            // key and value are all the checksums need (8-byte loads)
            for (uint64_t i0 = 0; i0 < d.n; i0 += uint64_t(kPkThreads) * kPkIlp) {
                uint2 t[kPkIlp];
                uint32_t g[kPkIlp];
#pragma unroll
                for (int j = 0; j < kPkIlp; j++) {
                    const uint64_t i = i0 + uint64_t(j) * kPkThreads + threadIdx.x;
                    t[j] = i < d.n ? *reinterpret_cast<const uint2 *>(&batch[i]) : make_uint2(0, 0);
                }
#pragma unroll
                for (int j = 0; j < kPkIlp; j++) {
                    const uint64_t i = i0 + uint64_t(j) * kPkThreads + threadIdx.x;
                    g[j] = i < d.n ? ohead[key_hash(t[j].x) & nb_mask] : kPkNil;
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
                        if (o[j].x == t[j].x) {
                            const uint32_t lv = own == 0 ? t[j].y : o[j].y, rv = own == 0 ? o[j].y : t[j].y;
                            const uint64_t a = hash_a(t[j].x, lv);
                            acc.count++;
                            acc.c1 += a * hash_b(t[j].x, rv);
                            if (c2) acc.c2 += pair_hash_from_a(a, rv);
                        }
                        g[j] = nx[j];
                        any |= g[j] != kPkNil;
                    }
                    if (!any) break;
                }
            }
            // append to the own side (kPkIlp tuples per thread at a time)
            sjb_tuple_t *tt = own ? tup1 : tup0;
            uint32_t *tn = own ? next1 : next0;
            uint32_t *th = own ? head1 : head0;
            const uint64_t base = s_n[own];
            for (uint64_t i0 = 0; i0 < d.n; i0 += uint64_t(kPkThreads) * kPkIlp) {
                sjb_tuple_t t[kPkIlp];
#pragma unroll
                for (int j = 0; j < kPkIlp; j++) {
                    const uint64_t i = i0 + uint64_t(j) * kPkThreads + threadIdx.x;
                    if (i < d.n) t[j] = batch[i];
                }
#pragma unroll
                for (int j = 0; j < kPkIlp; j++) {
                    const uint64_t i = i0 + uint64_t(j) * kPkThreads + threadIdx.x;
                    if (i < d.n) {
                        const uint32_t gg = uint32_t(base + i);
                        tt[gg] = t[j];
                        tn[gg] = atomicExch(&th[key_hash(t[j].key) & nb_mask], gg);
                    }
                }
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                s_n[own] = base + d.n;
                ctl->n[own] = base + d.n;
            }
        }
        const Acc r = block_sum(acc, s_red);
        if (threadIdx.x == 0) {
            const int q = int(seq % kPkQ);
            __hip_atomic_store(&res->count[q], r.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&res->c1[q], r.c1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&res->c2[q], r.c2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(&res->done, seq + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        __threadfence();
        __syncthreads();
        seq++;
    }
    if (threadIdx.x == 0) __hip_atomic_store(&res->exited, seq + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

static int pk_launch(sjb_ctx *c) {
    k_pk_eager_hj<<<1, kPkThreads, 0, c->pk_stream>>>(
        static_cast<PkCtl *>(c->d_pk_ctl), static_cast<const PkDesc *>(c->d_pk_descs), c->d_pk_slots,
        c->pk_slot_cap, c->d_eager_tuples[0], c->d_eager_tuples[1], c->d_eager_next[0], c->d_eager_next[1],
        c->d_eager_head[0], c->d_eager_head[1], uint32_t(c->eager_nb - 1), c->eager_nb,
        static_cast<PkRes *>(c->h_pk_res), c->pk_seq, c->pk_idle_limit, (c->opts.flags & SJB_F_C2) ? 1 : 0);
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
    SJB_HIP_TRY(hipStreamCreateWithFlags(&c->pk_stream, hipStreamNonBlocking));
    SJB_HIP_TRY(hipStreamCreateWithFlags(&c->pk_copy, hipStreamNonBlocking));
    if ((rc = dev_alloc(c, &c->d_pk_ctl, sizeof(PkCtl))) ||
        (rc = dev_alloc(c, &c->d_pk_descs, kPkQ * sizeof(PkDesc))) ||
        (rc = dev_alloc(c, (void **)&c->d_pk_slots, kPkQ * c->pk_slot_cap * sizeof(sjb_tuple_t))) ||
        (rc = host_alloc(c, &c->h_pk_stage, kPkQ * sizeof(PkDesc))))
        return rc;
    // completion words: coherent pinned memory the kernel stores into
    void *res = nullptr;
    if (hipHostMalloc(&res, sizeof(PkRes), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess) {
        (void)hipGetLastError();
        return SJB_E_NOMEM;
    }
    c->host_allocs.push_back(res);
    c->h_pk_res = res;
    memset(res, 0, sizeof(PkRes));
    SJB_HIP_TRY(hipMemsetAsync(c->d_pk_ctl, 0, sizeof(PkCtl), c->pk_copy));
    SJB_HIP_TRY(hipStreamSynchronize(c->pk_copy));
    if (c->pk_idle_limit == 0) {
        const char *e = getenv("SJB_PK_IDLE_POLLS");
        c->pk_idle_limit = e ? strtoull(e, nullptr, 10) : 2000000ull;
    }
    c->pk_ready = true;
    return SJB_OK;
}

// Posts descriptor `seq` (and its batch) and waits for its completion.
static int pk_post(sjb_ctx *c, uint32_t op, uint32_t side, const sjb_tuple_t *batch, uint64_t n,
                   uint64_t *count, uint64_t *c1, uint64_t *c2) {
    const uint64_t seq = c->pk_seq;
    const int q = int(seq % kPkQ);
    PkDesc *stage = static_cast<PkDesc *>(c->h_pk_stage) + q;
    stage->op = op;
    stage->side = side;
    stage->n = n;
    stage->seq = seq;
    if (n) SJB_HIP_TRY(hipMemcpyAsync(c->d_pk_slots + uint64_t(q) * c->pk_slot_cap, batch, n * sizeof(sjb_tuple_t),
                                      hipMemcpyHostToDevice, c->pk_copy));
    SJB_HIP_TRY(hipMemcpyAsync(static_cast<PkDesc *>(c->d_pk_descs) + q, stage, sizeof(PkDesc),
                               hipMemcpyHostToDevice, c->pk_copy));
    SJB_HIP_TRY(hipStreamWriteValue64(c->pk_copy, &static_cast<PkCtl *>(c->d_pk_ctl)->tail, seq + 1, 0));
    SJB_HIP_TRY(hipStreamSynchronize(c->pk_copy));
    c->pk_seq = seq + 1;
    volatile PkRes *res = static_cast<volatile PkRes *>(c->h_pk_res);
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        if (res->done > seq) break;
        // an idle kernel exits by itself; relaunch it for this descriptor
        if (c->pk_running && hipStreamQuery(c->pk_stream) == hipSuccess) {
            c->pk_running = false;
            if (res->done > seq) break;
            const uint64_t save = c->pk_seq;
            c->pk_seq = seq;
            int rc = pk_launch(c);
            c->pk_seq = save;
            if (rc) return rc;
        } else if (!c->pk_running) {
            const uint64_t save = c->pk_seq;
            c->pk_seq = seq;
            int rc = pk_launch(c);
            c->pk_seq = save;
            if (rc) return rc;
        }
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(20)) return SJB_E_TIMEOUT;
        std::this_thread::yield();
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
    SJB_HIP_TRY(hipStreamWriteValue64(c->pk_copy, &static_cast<PkCtl *>(c->d_pk_ctl)->stop, 1, 0));
    SJB_HIP_TRY(hipStreamSynchronize(c->pk_copy));
    SJB_HIP_TRY(hipStreamSynchronize(c->pk_stream));
    c->pk_running = false;
    SJB_HIP_TRY(hipStreamWriteValue64(c->pk_copy, &static_cast<PkCtl *>(c->d_pk_ctl)->stop, 0, 0));
    SJB_HIP_TRY(hipStreamSynchronize(c->pk_copy));
    return SJB_OK;
}

uint64_t sjb_pk_launches(sjb_ctx_t *c) { return c ? c->pk_launches : 0; }

}  // extern "C"
