// eager.hip — per-batch (eager) symmetric joins with device-resident state.
//
// Replaces the eager GPU executors GPU_HJ / GPU_HJ_Atomic
// (src/sjb/windowing/eager/executor/GPU/GPU_HJ.cu:33-120,
// GPU_HJ_Atomic.cu:34-116) and GPU_NLJ / GPU_NLJ_BM (GPU_NLJ.cu:10-73,
// GPU_NLJ_BM.cu:10-137).
//
// The reference reserves numDistinctKeys x windowSize 16-byte slots per side
// (GPU_HJ.cu:15-16: 16 GB at D=1000, W=2^20) and has its kernels read the
// batch from pinned host memory over PCIe.  Here each side's window state is
//   tuples[W]   append-only copy of the side's tuples (device),
//   next[W]     chain link of each tuple,
//   head[NB]    chain head per bucket, NB = pow2 >= W,
// i.e. O(W) memory for any key range.  A batch is copied H2D once, probed
// against the opposite side (walk the chain of its bucket, compare keys),
// then appended to its own side (one atomicExch per tuple on head[]).
// The count is exact per batch (the reference shares one device counter
// across concurrently running batches and zeroes it from thread 1 while
// others add, GPU_HJ.cu:58-60).
#include <string.h>

#include <algorithm>

#include "sjb_device.hpp"
#include "sjb_internal.hpp"

namespace sjb {

constexpr uint32_t kEagerNil = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t eager_bucket(uint32_t k, uint32_t nb_mask) {
    return key_hash(k) & nb_mask;
}

// Appends the batch to its own side.  Workgroup 0 also sums the probe's
// n_partials per-block (count, C1, C2) partials into *res (the batch's
// result), so no separate reduction launch follows the probe.
__global__ __launch_bounds__(256) void k_eager_build(const sjb_tuple_t *__restrict__ batch, uint64_t n,
                                                     uint64_t base, sjb_tuple_t *__restrict__ tuples,
                                                     uint32_t *__restrict__ next,
                                                     uint32_t *__restrict__ head, uint32_t nb_mask,
                                                     const uint64_t *__restrict__ partials, uint32_t n_partials,
                                                     sjb_window_result_t *__restrict__ res) {
    __shared__ uint64_t s_red[3 * 256 / kWave];
    if (blockIdx.x == 0 && n_partials) {  // uniform per workgroup
        Acc a;
        a.zero();
        for (uint32_t b = threadIdx.x; b < n_partials; b += blockDim.x) {
            a.count += partials[3 * b];
            a.c1 += partials[3 * b + 1];
            a.c2 += partials[3 * b + 2];
        }
        const Acc r = block_sum(a, s_red);
        if (threadIdx.x == 0) {
            res->count = r.count;
            res->checksum1 = r.c1;
            res->checksum2 = r.c2;
            res->pairs = 0;
        }
    }
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const sjb_tuple_t t = batch[i];
    const uint32_t g = uint32_t(base + i);
    tuples[g] = t;
    next[g] = atomicExch(&head[eager_bucket(t.key, nb_mask)], g);
}

// is_left: the batch is the left (R) side, so pairs are (batch, state).
template <bool C2, bool WRITE>
__global__ __launch_bounds__(256) void k_eager_probe(
    const sjb_tuple_t *__restrict__ batch, uint64_t n, const sjb_tuple_t *__restrict__ tuples,
    const uint32_t *__restrict__ next, const uint32_t *__restrict__ head, uint32_t nb_mask,
    int is_left, uint64_t *__restrict__ partials, sjb_result_tuple_t *__restrict__ out,
    uint64_t out_cap, int wrap, unsigned long long *__restrict__ cursor) {
    __shared__ uint64_t s_red[3 * 256 / kWave];
    __shared__ uint32_t s_ws[256 / kWave];
    __shared__ unsigned long long s_base;
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    Acc acc;
    acc.zero();
    sjb_tuple_t t{};
    uint32_t h = kEagerNil;
    if (i < n) {
        t = batch[i];
        h = head[eager_bucket(t.key, nb_mask)];
    }
    uint32_t m = 0;
    for (uint32_t g = h; g != kEagerNil; g = next[g]) {
        const sjb_tuple_t o = tuples[g];
        if (o.key != t.key) continue;
        const sjb_tuple_t &l = is_left ? t : o;
        const sjb_tuple_t &r = is_left ? o : t;
        const uint64_t a = hash_a(l.key, l.val);
        acc.count++;
        acc.c1 += a * hash_b(r.key, r.val);
        if (C2) acc.c2 += pair_hash_from_a(a, r.val);
        m++;
    }
    if (WRITE) {
        uint32_t tot;
        const uint32_t ex = block_excl_scan_u32(m, s_ws, &tot);
        if (threadIdx.x == 0) s_base = tot ? atomicAdd(cursor, (unsigned long long)tot) : 0ull;
        __syncthreads();
        uint64_t pos = s_base + ex;
        if (m) {
            for (uint32_t g = h; g != kEagerNil; g = next[g]) {
                const sjb_tuple_t o = tuples[g];
                if (o.key != t.key) continue;
                const sjb_tuple_t &l = is_left ? t : o;
                const sjb_tuple_t &r = is_left ? o : t;
                uint64_t q = pos++;
                if (wrap) q %= out_cap;
                if (wrap || q < out_cap) {
                    sjb_result_tuple_t row;
                    row.lkey = l.key;
                    row.rkey = r.key;
                    row.lval = l.val;
                    row.rval = r.val;
                    row.lts = l.ts;
                    row.rts = r.ts;
                    out[q] = row;
                }
            }
        }
    }
    const Acc r = block_sum(acc, s_red);
    if (threadIdx.x == 0) {
        partials[3 * blockIdx.x] = r.count;
        partials[3 * blockIdx.x + 1] = r.c1;
        partials[3 * blockIdx.x + 2] = r.c2;
    }
}

int eager_init(sjb_ctx *c) {
    if (c->eager_ready) return SJB_OK;
    const uint64_t W = c->opts.max_window_tuples;
    if (W >= kEagerNil) return SJB_E_CAPACITY;
    uint64_t nb = 1024;
    while (nb < W) nb <<= 1;
    const uint64_t need = 2 * (W * (sizeof(sjb_tuple_t) + 4) + nb * 4) + W * sizeof(sjb_tuple_t);
    if (c->opts.hbm_budget_bytes && c->bytes + need > c->opts.hbm_budget_bytes) return SJB_E_BUDGET;
    int rc;
    for (int s = 0; s < 2; s++) {
        if ((rc = dev_alloc(c, (void **)&c->d_eager_tuples[s], W * sizeof(sjb_tuple_t))) ||
            (rc = dev_alloc(c, (void **)&c->d_eager_next[s], W * sizeof(uint32_t))) ||
            (rc = dev_alloc(c, (void **)&c->d_eager_head[s], nb * sizeof(uint32_t))))
            return rc;
    }
    c->eager_cap = W;
    c->eager_nb = nb;
    c->batch_cap = W;
    if ((rc = dev_alloc(c, (void **)&c->d_batch, W * sizeof(sjb_tuple_t)))) return rc;
    c->eager_ready = true;
    return sjb_eager_clear(c);
}

}  // namespace sjb

using namespace sjb;

extern "C" {

int sjb_eager_clear(sjb_ctx_t *c) {
    if (!c) return SJB_E_INVALID;
    if (!c->eager_ready) return SJB_OK;
    SJB_HIP_TRY(hipSetDevice(c->device));
    for (int s = 0; s < 2; s++) {
        int rc = launch_fill_u32(c->stream, c->d_eager_head[s], kEagerNil, c->eager_nb);
        if (rc) return rc;
        c->eager_n[s] = 0;
    }
    SJB_HIP_TRY(hipStreamSynchronize(c->stream));
    return SJB_OK;
}

int sjb_eager_end_of_stream(sjb_ctx_t *c) {
    if (!c) return SJB_E_INVALID;
    SJB_HIP_TRY(hipStreamSynchronize(c->stream));
    return SJB_OK;
}

int sjb_eager_execute(sjb_ctx_t *c, int algorithm, int method, uint32_t flags,
                      const sjb_tuple_t *batch, uint64_t n, int is_left, sjb_result_tuple_t *out,
                      uint64_t out_cap, sjb_stats_t *stats) {
    if (!c || !stats || (n && !batch)) return SJB_E_INVALID;
    if (algorithm != SJB_ALG_HJ && algorithm != SJB_ALG_NLJ) return SJB_E_UNSUPPORTED;
    if (method != SJB_OUT_NOOUTPUT && method != SJB_OUT_ATOMIC && method != SJB_OUT_BITMAP &&
        method != SJB_OUT_COUNTKERNEL)
        return SJB_E_UNSUPPORTED;
    if (method == SJB_OUT_BITMAP && algorithm != SJB_ALG_NLJ) return SJB_E_UNSUPPORTED;
    if (method != SJB_OUT_NOOUTPUT && out_cap > 0 && !out) return SJB_E_INVALID;
    if (method == SJB_OUT_ATOMIC && out_cap == 0) return SJB_E_INVALID;
    memset(stats, 0, sizeof(*stats));
    if (c->pk_ready) return SJB_E_STATE;  // the window state belongs to the persistent kernel
    SJB_HIP_TRY(hipSetDevice(c->device));
    int rc = eager_init(c);
    if (rc) return rc;
    flags |= c->opts.flags;
    const int own = is_left ? 0 : 1, other = 1 - own;
    if (c->eager_n[own] + n > c->eager_cap) return SJB_E_CAPACITY;
    if (n == 0) return SJB_OK;
    const bool timing = (flags & SJB_F_TIMING) != 0;
    SJB_HIP_TRY(hipMemcpyAsync(c->d_batch, batch, n * sizeof(sjb_tuple_t), hipMemcpyHostToDevice,
                               c->stream));
    if (timing) SJB_HIP_TRY(hipEventRecord(c->ev0, c->stream));
    const bool write = method != SJB_OUT_NOOUTPUT;
    const bool c2 = write || (flags & SJB_F_C2);
    OutBuf ob;
    if (write) {
        rc = out_prepare(c, out, out_cap, &ob);
        if (rc) return rc;
        out = ob.dev;
    }
    const int wrap = method == SJB_OUT_ATOMIC ? 1 : 0;
    auto *cur = reinterpret_cast<unsigned long long *>(c->d_cursor);
    if (write) {
        rc = launch_fill_u64(c->stream, c->d_cursor, 0, 2);
        if (rc) return rc;
    }
    const uint64_t n_other = c->eager_n[other];
    sjb_window_result_t res{};
    uint32_t probe_partials = 0;
    if (write && !wrap && n_other) {
        // Count first so an undersized output buffer is rejected before the
        // state changes (the caller can grow the buffer and call again).
        if (algorithm == SJB_ALG_HJ) {
            const unsigned blocks = unsigned((n + 255) / 256);
            k_eager_probe<false, false><<<blocks, 256, 0, c->stream>>>(
                c->d_batch, n, c->d_eager_tuples[other], c->d_eager_next[other], c->d_eager_head[other],
                uint32_t(c->eager_nb - 1), is_left, c->d_partials, nullptr, 0, 0, cur);
            SJB_HIP_TRY(hipGetLastError());
            rc = launch_reduce(c, int(blocks), c->d_result, 0);
        } else {
            const sjb_tuple_t *L0 = is_left ? c->d_batch : c->d_eager_tuples[other];
            const sjb_tuple_t *R0 = is_left ? c->d_eager_tuples[other] : c->d_batch;
            rc = launch_nlj(c, L0, is_left ? n : n_other, R0, is_left ? n_other : n, 0, 0, nullptr, 0, 0,
                            c->d_result);
        }
        if (rc) return rc;
        SJB_HIP_TRY(hipMemcpyAsync(c->h_result, c->d_result, sizeof(sjb_window_result_t),
                                   hipMemcpyDeviceToHost, c->stream));
        SJB_HIP_TRY(hipStreamSynchronize(c->stream));
        if (c->h_result->count > out_cap) {
            stats->result_count = c->h_result->count;
            return SJB_E_CAPACITY;
        }
        rc = launch_fill_u64(c->stream, c->d_cursor, 0, 2);
        if (rc) return rc;
    }
    if (n_other == 0) {
        rc = launch_fill_u64(c->stream, reinterpret_cast<uint64_t *>(c->d_result), 0, 4);
        if (rc) return rc;
    } else if (algorithm == SJB_ALG_HJ) {
        const unsigned blocks = unsigned((n + 255) / 256);
        const uint32_t mask = uint32_t(c->eager_nb - 1);
#define SJB_EP(C2_, W_)                                                                           \
    k_eager_probe<C2_, W_><<<blocks, 256, 0, c->stream>>>(                                        \
        c->d_batch, n, c->d_eager_tuples[other], c->d_eager_next[other], c->d_eager_head[other], \
        mask, is_left, c->d_partials, out, out_cap, wrap, cur)
        if (write)
            SJB_EP(true, true);
        else if (c2)
            SJB_EP(true, false);
        else
            SJB_EP(false, false);
#undef SJB_EP
        SJB_HIP_TRY(hipGetLastError());
        probe_partials = blocks;  // summed by the append kernel
    } else {
        // NLJ: incoming batch x the other side's resident window
        // (eager countKernel / getJoinResultBitmapEager).  The reference
        // misses the latest opposing batch (GPU_NLJ.cu:20,28); here the state
        // holds every appended tuple.
        const sjb_tuple_t *L = is_left ? c->d_batch : c->d_eager_tuples[other];
        const sjb_tuple_t *R = is_left ? c->d_eager_tuples[other] : c->d_batch;
        const uint64_t nL = is_left ? n : n_other, nR = is_left ? n_other : n;
        if (method == SJB_OUT_BITMAP) {
            uint64_t nm = 0;
            rc = launch_nlj_bitmap(c, L, nL, R, nR, flags, out, out_cap, c->d_result, &nm);
        } else {
            rc = launch_nlj(c, L, nL, R, nR, c2 ? (flags | SJB_F_C2) : flags, write ? 1 : 0, out,
                            out_cap, wrap, c->d_result);
        }
        if (rc) return rc;
    }
    // append the batch to its own side (and, after an HJ probe, sum its
    // partials into the result)
    if (algorithm == SJB_ALG_HJ) {
        k_eager_build<<<unsigned((n + 255) / 256), 256, 0, c->stream>>>(
            c->d_batch, n, c->eager_n[own], c->d_eager_tuples[own], c->d_eager_next[own],
            c->d_eager_head[own], uint32_t(c->eager_nb - 1), c->d_partials, probe_partials, c->d_result);
        SJB_HIP_TRY(hipGetLastError());
    } else {
        SJB_HIP_TRY(hipMemcpyAsync(c->d_eager_tuples[own] + c->eager_n[own], c->d_batch,
                                   n * sizeof(sjb_tuple_t), hipMemcpyDeviceToDevice, c->stream));
    }
    if (timing) SJB_HIP_TRY(hipEventRecord(c->ev1, c->stream));
    c->eager_n[own] += n;
    SJB_HIP_TRY(hipMemcpyAsync(c->h_result, c->d_result, sizeof(sjb_window_result_t),
                               hipMemcpyDeviceToHost, c->stream));
    SJB_HIP_TRY(hipStreamSynchronize(c->stream));
    res = *c->h_result;
    stats->result_count = res.count;
    stats->checksum1 = res.checksum1;
    stats->checksum2 = c2 ? res.checksum2 : 0;
    stats->c2_valid = c2 ? 1 : 0;
    if (write) {
        stats->pairs_written = std::min(res.count, out_cap);
        rc = out_finish(c, ob, stats->pairs_written);
        if (rc) return rc;
        if (!wrap && res.count > out_cap) rc = SJB_E_CAPACITY;
    }
    if (timing) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) {
            stats->device_ns = uint64_t(ms * 1.0e6f);
            stats->probe_time_us = uint64_t(ms * 1000.f);
        }
    }
    return rc;
}

}  // extern "C"
