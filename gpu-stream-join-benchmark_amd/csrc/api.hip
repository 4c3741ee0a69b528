// api.hip — the C ABI (include/sjb.h): contexts and lazy window joins.
//
// A context owns one HIP stream and every device buffer a window of
// max_window_tuples per side needs, allocated once at sjb_open (the
// reference allocates per executor in the constructor, e.g.
// src/sjb/windowing/lazy/executor/GPU/GPU_HJ.cu:266-287, and per window in
// its CountKernel/ES executors).  A window is enqueued with no host sync
// between its kernels (the reference performs 11 cudaStreamSynchronize
// calls and a host prefix sum per HJ window,
// src/sjb/windowing/lazy/executor/GPU/GPU_HJ_NoOutput.cu:43-98).
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cstring>
#include <chrono>

#include "sjb_device.hpp"
#include "sjb_internal.hpp"

namespace sjb {

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    if (getenv("SJB_VERBOSE"))
        fprintf(stderr, "sjb: %s failed at %s:%d: %s\n", what, file, line, hipGetErrorString(e));
    if (e == hipErrorOutOfMemory) return SJB_E_NOMEM;
    return SJB_E_HIP;
}

int dev_alloc(sjb_ctx *c, void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return e == hipErrorOutOfMemory ? SJB_E_NOMEM : SJB_E_HIP;
    }
    c->dev_allocs.push_back(*p);
    c->bytes += bytes;
    return SJB_OK;
}

// Frees a tracked device buffer of `bytes` bytes (as passed to dev_alloc)
// and gives its bytes back to the ctx's account; *p becomes null.
void dev_free(sjb_ctx *c, void **p, size_t bytes) {
    if (!*p) return;
    if (bytes == 0) bytes = 256;
    auto &v = c->dev_allocs;
    v.erase(std::remove(v.begin(), v.end(), *p), v.end());
    (void)hipFree(*p);
    c->bytes -= bytes;
    *p = nullptr;
}

// Replaces a tracked device buffer by one of `bytes` bytes (contents not
// kept), checking the ctx's HBM budget with the old buffer given back.
int dev_regrow(sjb_ctx *c, void **p, size_t old_bytes, size_t bytes) {
    const uint64_t budget = c->opts.hbm_budget_bytes;
    const uint64_t held = c->bytes + c->ext_bytes - (*p ? (old_bytes ? old_bytes : 256) : 0);
    if (budget && held + bytes > budget) return SJB_E_BUDGET;
    dev_free(c, p, old_bytes);
    return dev_alloc(c, p, bytes);
}

int host_alloc(sjb_ctx *c, void **p, size_t bytes) {
    *p = nullptr;
    hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return SJB_E_NOMEM;
    }
    c->host_allocs.push_back(*p);
    return SJB_OK;
}

bool device_accessible(const void *p) {
    if (!p) return false;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.devicePointer != nullptr;
}

int out_prepare(sjb_ctx *c, sjb_result_tuple_t *out, uint64_t cap, OutBuf *ob) {
    ob->cap = cap;
    ob->host = nullptr;
    ob->dev = out;
    if (cap == 0 || out == nullptr || device_accessible(out)) return SJB_OK;
    if (c->out_stage_cap < cap) {
        // geometric growth: callers that size their buffer to each window's
        // count would otherwise reallocate on every larger window
        const uint64_t want = std::max<uint64_t>(cap, c->out_stage_cap + c->out_stage_cap / 2);
        void *p = c->d_out_stage;
        int rc = dev_regrow(c, &p, c->out_stage_cap * sizeof(sjb_result_tuple_t),
                            want * sizeof(sjb_result_tuple_t));
        c->d_out_stage = static_cast<sjb_result_tuple_t *>(p);
        c->out_stage_cap = c->d_out_stage ? want : 0;
        if (rc) return rc;
    }
    ob->dev = c->d_out_stage;
    ob->host = out;
    return SJB_OK;
}

int out_finish(sjb_ctx *c, const OutBuf &ob, uint64_t rows) {
    if (!ob.host || rows == 0) return SJB_OK;
    SJB_HIP_TRY(hipMemcpyAsync(ob.host, ob.dev, std::min(rows, ob.cap) * sizeof(sjb_result_tuple_t),
                               hipMemcpyDeviceToHost, c->stream));
    SJB_HIP_TRY(hipStreamSynchronize(c->stream));
    return SJB_OK;
}

static bool valid_alg(int alg) { return alg == SJB_ALG_NLJ || alg == SJB_ALG_HJ || alg == SJB_ALG_SMJ; }
static bool valid_method(int m) { return m >= SJB_OUT_NOOUTPUT && m <= SJB_OUT_ATOMIC; }

// Plans one HJ/SMJ window (fanout, partition function) and enqueues its
// partition pass (plus the key-range sample or the heavy-key reset it
// needs).  Both sides non-empty.
int window_partition(sjb_ctx *c, int alg, uint32_t flags, const sjb_tuple_t *L, uint64_t nL,
                     const sjb_tuple_t *R, uint64_t nR, int write, WindowPlan *out, int buf,
                     sjb_window_result_t *direct) {
    WindowPlan &p = *out;
    p.ps = buf < 0 ? c->stream : c->ov_stream;
    p.part = buf < 0 ? c->d_part : c->ov_part[buf];
    p.offs = buf < 0 ? c->d_offs : c->ov_offs[buf];
    p.heavy = buf < 0 ? c->d_heavy : c->ov_heavy[buf];
    p.alg = alg;
    p.flags = flags;
    p.nL = nL;
    p.nR = nR;
    const bool small = c->small_chunks && nL + nR <= kSmallWindow;
    p.logF = choose_logF(alg, flags, nL, int(c->opts.fanout_log2), small);
    const bool learned = (flags & SJB_F_LEARNED) && c->rmi_valid;
    p.kind = part_kind(alg, learned);
    p.payload = write ? PAY_IDX : PAY_VAL;
    p.peel = learned && (flags & SJB_F_GROUPED) && !write && c->rmi_nheavy > 0;
    p.chunk = !small ? kChunk : nL + nR <= kTinyWindow ? kTinyChunk : kSmallChunk;
    p.range_blocks = 0;
    p.ncL = int((nL + p.chunk - 1) / p.chunk);
    p.ncR = int((nR + p.chunk - 1) / p.chunk);
    if (p.ncL > kMaxChunks || p.ncR > kMaxChunks) return SJB_E_CAPACITY;
    // count paths sum in the join (the partition pass zeroes the result)
    p.direct = direct && !write && !p.peel && p.kind != PART_LEARNED ? direct : nullptr;
    int rc;
    if (p.kind == PART_RANGE) {
        rc = launch_minmax(c, p.ps, L, nL, R, nR, &p.range_blocks);
        if (rc) return rc;
    }
    if (p.peel) {
        rc = launch_fill_u64(p.ps, p.heavy, 0, 4 * kMaxHeavy);
        if (rc) return rc;
    }
    return launch_partition(c, p, L, R);
}

// Second buffer set, stream and events of overlapped async windows
// (allocated on first use).
static int overlap_setup(sjb_ctx *c) {
    if (c->ov_stream) return SJB_OK;
    int rc;
    if ((rc = dev_alloc(c, (void **)&c->ov_part[1], 2 * c->cap * sizeof(Rec))) ||
        (rc = dev_alloc(c, (void **)&c->ov_offs[1], c->offs_cap * sizeof(uint32_t))) ||
        (rc = dev_alloc(c, (void **)&c->ov_heavy[1], 4 * kMaxHeavy * sizeof(uint64_t))))
        return rc;
    c->ov_part[0] = c->d_part;
    c->ov_offs[0] = c->d_offs;
    c->ov_heavy[0] = c->d_heavy;
    for (int b = 0; b < 2; b++) {
        SJB_HIP_TRY(hipEventCreateWithFlags(&c->ov_part_done[b], hipEventDisableTiming));
        SJB_HIP_TRY(hipEventCreateWithFlags(&c->ov_join_done[b], hipEventDisableTiming));
        SJB_HIP_TRY(hipEventRecord(c->ov_join_done[b], c->stream));
    }
    SJB_HIP_TRY(hipStreamCreateWithFlags(&c->ov_stream, hipStreamNonBlocking));
    return SJB_OK;
}

// Enqueues one window (count/checksums or pair output) on the ctx stream;
// the result lands in d_res.  write: 0 = count, 1 = materialise.
int window_enqueue(sjb_ctx *c, int alg, uint32_t flags, const sjb_tuple_t *L, uint64_t nL,
                   const sjb_tuple_t *R, uint64_t nR, int write, sjb_result_tuple_t *out,
                   uint64_t out_cap, int wrap, sjb_window_result_t *d_res, int *logF_used,
                   const hipEvent_t *ev, int buf) {
    int rc;
    if (write) {
        rc = launch_fill_u64(c->stream, c->d_cursor, 0, 2);
        if (rc) return rc;
    }
    if (nL == 0 || nR == 0) {
        *logF_used = 0;
        return launch_fill_u64(c->stream, reinterpret_cast<uint64_t *>(d_res), 0,
                               sizeof(sjb_window_result_t) / 8);
    }
    if (alg == SJB_ALG_NLJ) {
        *logF_used = 0;
        if (ev) SJB_HIP_TRY(hipEventRecord(ev[0], c->stream));
        if (ev) SJB_HIP_TRY(hipEventRecord(ev[1], c->stream));
        rc = launch_nlj(c, L, nL, R, nR, flags, write, out, out_cap, wrap, d_res);
        if (rc) return rc;
        if (ev) SJB_HIP_TRY(hipEventRecord(ev[3], c->stream));
        return SJB_OK;
    }
    // overlapped: the partition pass waits for the last join that read this
    // buffer set, and the join waits for the partition pass
    hipStream_t ps = buf < 0 ? c->stream : c->ov_stream;
    if (buf >= 0) SJB_HIP_TRY(hipStreamWaitEvent(ps, c->ov_join_done[buf], 0));
    if (ev) SJB_HIP_TRY(hipEventRecord(ev[0], ps));
    WindowPlan p;
    rc = window_partition(c, alg, flags, L, nL, R, nR, write, &p, buf, d_res);
    if (rc) return rc;
    *logF_used = p.logF;
    if (ev) SJB_HIP_TRY(hipEventRecord(ev[1], ps));
    if (buf >= 0) {
        SJB_HIP_TRY(hipEventRecord(c->ov_part_done[buf], ps));
        SJB_HIP_TRY(hipStreamWaitEvent(c->stream, c->ov_part_done[buf], 0));
    }
    // ev[2] (join start) differs from ev[1] only when the partition pass ran
    // on the second stream; otherwise ev[1] stands for it (one event less)
    if (ev && buf >= 0) SJB_HIP_TRY(hipEventRecord(ev[2], c->stream));
    rc = alg == SJB_ALG_HJ ? launch_hj_join(c, p, L, R, out, out_cap, wrap)
                           : launch_smj_join(c, p, L, R, out, out_cap, wrap);
    if (rc) return rc;
    if (!p.direct) {
        rc = launch_reduce(c, 1 << p.logF, d_res, p.peel ? c->rmi_nheavy : 0, p.heavy);
        if (rc) return rc;
    }
    if (ev) SJB_HIP_TRY(hipEventRecord(ev[3], c->stream));
    // buffer set 0 is also the ctx's own: its next overlapped partition pass
    // waits for this join
    if (c->ov_stream) SJB_HIP_TRY(hipEventRecord(c->ov_join_done[buf < 0 ? 0 : buf], c->stream));
    return SJB_OK;
}

static int fetch_result(sjb_ctx *c, sjb_window_result_t *host) {
    SJB_HIP_TRY(hipMemcpyAsync(c->h_result, c->d_result, sizeof(sjb_window_result_t),
                               hipMemcpyDeviceToHost, c->stream));
    SJB_HIP_TRY(hipStreamSynchronize(c->stream));
    *host = *c->h_result;
    return SJB_OK;
}

static void fill_times(sjb_ctx *c, int alg, bool timing, sjb_stats_t *st) {
    if (!timing) return;
    float a = 0.f, b = 0.f;
    if (hipEventElapsedTime(&a, c->ev[0], c->ev[1]) != hipSuccess) a = 0.f;
    if (hipEventElapsedTime(&b, c->ev[2], c->ev[3]) != hipSuccess) b = 0.f;
    const uint64_t ua = uint64_t(a * 1000.f), ub = uint64_t(b * 1000.f);
    st->device_ns += uint64_t((a + b) * 1.0e6f);
    if (alg == SJB_ALG_SMJ) {
        st->sort_time_us += ua;
        st->merge_time_us += ub;
    } else {
        st->build_time_us += ua;
        st->probe_time_us += ub;
    }
}

// Core of sjb_lazy_join for device-resident inputs.
static int lazy_join_device(sjb_ctx *c, int alg, int method, uint32_t flags, const sjb_tuple_t *L,
                            uint64_t nL, const sjb_tuple_t *R, uint64_t nR, sjb_result_tuple_t *out,
                            uint64_t out_cap, sjb_stats_t *st) {
    flags |= c->opts.flags;
    const bool timing = (flags & SJB_F_TIMING) != 0;
    sjb_window_result_t res{};
    OutBuf ob;
    if (method != SJB_OUT_NOOUTPUT) {
        if (out_cap > 0 && out == nullptr) return SJB_E_INVALID;
        int r0 = out_prepare(c, out, out_cap, &ob);
        if (r0) return r0;
        out = ob.dev;
    }
    int logF = 0, rc;
    const bool c2_enum = (flags & SJB_F_C2) != 0 && !(flags & SJB_F_GROUPED);
    if (method == SJB_OUT_BITMAP) {
        if (alg != SJB_ALG_NLJ) return SJB_E_UNSUPPORTED;
        uint64_t nm = 0;
        rc = launch_fill_u64(c->stream, c->d_cursor, 0, 2);
        if (rc) return rc;
        if (nL && nR) {
            if (timing) SJB_HIP_TRY(hipEventRecord(c->ev0, c->stream));
            if (timing) SJB_HIP_TRY(hipEventRecord(c->ev1, c->stream));
            rc = launch_nlj_bitmap(c, L, nL, R, nR, flags, out, out_cap, c->d_result, &nm);
            if (rc) return rc;
            if (timing) SJB_HIP_TRY(hipEventRecord(c->ev2, c->stream));
            rc = fetch_result(c, &res);
            if (rc) return rc;
        }
        st->result_count = res.count;
        st->checksum1 = res.checksum1;
        st->checksum2 = res.checksum2;
        st->c2_valid = 1;
        st->pairs_written = std::min(res.count, out_cap);
        fill_times(c, alg, timing && nL && nR, st);
        rc = out_finish(c, ob, st->pairs_written);
        if (rc) return rc;
        return res.count > out_cap ? SJB_E_CAPACITY : SJB_OK;
    }
    if (method == SJB_OUT_NOOUTPUT) {
        rc = window_enqueue(c, alg, flags, L, nL, R, nR, 0, nullptr, 0, 0, c->d_result, &logF,
                            timing ? c->ev : nullptr);
        if (rc) return rc;
        rc = fetch_result(c, &res);
        if (rc) return rc;
        st->c2_valid = c2_enum ? 1 : 0;
    } else {
        if (out_cap > 0 && out == nullptr) return SJB_E_INVALID;
        if (method == SJB_OUT_COUNTKERNEL) {
            // pass 1: exact count (GPU_HJ_CountKernel.cu:83-101 counts first)
            rc = window_enqueue(c, alg, flags & ~SJB_F_C2, L, nL, R, nR, 0, nullptr, 0, 0,
                                c->d_result, &logF, nullptr);
            if (rc) return rc;
            rc = fetch_result(c, &res);
            if (rc) return rc;
            if (res.count > out_cap) {
                st->result_count = res.count;
                st->checksum1 = res.checksum1;
                st->fanout_log2 = uint32_t(logF);
                return SJB_E_CAPACITY;
            }
        }
        if (method == SJB_OUT_ATOMIC && out_cap == 0) return SJB_E_INVALID;
        const int wrap = method == SJB_OUT_ATOMIC ? 1 : 0;
        rc = window_enqueue(c, alg, flags | SJB_F_C2, L, nL, R, nR, 1, out, out_cap, wrap,
                            c->d_result, &logF, timing ? c->ev : nullptr);
        if (rc) return rc;
        rc = fetch_result(c, &res);
        if (rc) return rc;
        st->c2_valid = 1;
        st->pairs_written = std::min(res.count, out_cap);
        rc = out_finish(c, ob, st->pairs_written);
        if (rc) return rc;
        if (method != SJB_OUT_ATOMIC && res.count > out_cap) rc = SJB_E_CAPACITY;
    }
    st->result_count = res.count;
    st->checksum1 = res.checksum1;
    st->checksum2 = st->c2_valid ? res.checksum2 : 0;
    st->fanout_log2 = uint32_t(logF);
    fill_times(c, alg, timing && nL && nR, st);
    return rc;
}

}  // namespace sjb

using namespace sjb;

extern "C" {

int sjb_abi_version(void) { return SJB_ABI_VERSION; }

const char *sjb_strerror(int code) {
    switch (code) {
        case SJB_OK: return "ok";
        case SJB_E_INVALID: return "invalid argument";
        case SJB_E_CAPACITY: return "capacity exceeded";
        case SJB_E_NOMEM: return "out of memory";
        case SJB_E_HIP: return "HIP runtime error";
        case SJB_E_UNSUPPORTED: return "unsupported combination";
        case SJB_E_NODEVICE: return "no usable device";
        case SJB_E_BUDGET: return "HBM budget exceeded";
        case SJB_E_TIMEOUT: return "bounded wait expired";
        case SJB_E_STATE: return "invalid state";
        case SJB_E_MISMATCH: return "launch shapes gave different results";
    }
    return "unknown error";
}

int sjb_open(int device, const sjb_options_t *opts, sjb_ctx_t **out) {
    if (!out) return SJB_E_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return SJB_E_NODEVICE;
    }
    if (device < 0 || device >= ndev) return SJB_E_INVALID;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return SJB_E_NODEVICE;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return SJB_E_NODEVICE;
    SJB_HIP_TRY(hipSetDevice(device));

    sjb_ctx *c = new sjb_ctx();
    c->device = device;
    if (opts) c->opts = *opts;
    if (c->opts.max_window_tuples == 0) c->opts.max_window_tuples = 1ull << 22;
    if (c->opts.num_distinct_keys == 0) c->opts.num_distinct_keys = 1000;
    if (c->opts.batch_size == 0) c->opts.batch_size = 8192;
    c->n_cu = prop.multiProcessorCount;
    c->cap = c->opts.max_window_tuples;
    int rc = SJB_OK;
    auto fail = [&](int r) {
        sjb_close(c);
        return r;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(SJB_E_HIP);
    if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->ev2) != hipSuccess || hipEventCreate(&c->ev3) != hipSuccess)
        return fail(SJB_E_HIP);
    c->ev[0] = c->ev0;
    c->ev[1] = c->ev1;
    c->ev[2] = c->ev1;  // one stream: partition end = join start
    c->ev[3] = c->ev2;

    const uint64_t cap = c->cap;
    if (c->opts.hbm_budget_bytes) {
        const uint64_t need = cap * (2 * sizeof(sjb_tuple_t) + 2 * sizeof(Rec));
        if (need > c->opts.hbm_budget_bytes) return fail(SJB_E_BUDGET);
    }
    const uint64_t nchunks = std::min<uint64_t>((cap + kChunk - 1) / kChunk, kMaxChunks);
    // offsets rows: two sides of kChunk chunks, or a small window's
    // kSmallChunk chunks (both sides together at most kSmallWindow tuples)
    // or kTinyChunk chunks (at most kTinyWindow tuples)
    const uint64_t small_rows = std::max<uint64_t>(std::min<uint64_t>(2 * cap, kSmallWindow) / kSmallChunk,
                                                   std::min<uint64_t>(2 * cap, kTinyWindow) / kTinyChunk) + 2;
    c->offs_cap = ((uint64_t(1) << kMaxLogF) + 1) * std::max<uint64_t>(2 * std::max<uint64_t>(nchunks, 1), small_rows);
    const uint64_t n_partials = std::max<uint64_t>(65536, 1ull << kMaxLogF);
    c->mask_cap = std::max<uint64_t>(cap, 1ull << 20);
    if ((rc = dev_alloc(c, (void **)&c->d_stage[0], cap * sizeof(sjb_tuple_t))) ||
        (rc = dev_alloc(c, (void **)&c->d_stage[1], cap * sizeof(sjb_tuple_t))) ||
        (rc = dev_alloc(c, (void **)&c->d_part, 2 * cap * sizeof(Rec))) ||
        (rc = dev_alloc(c, (void **)&c->d_offs, c->offs_cap * sizeof(uint32_t))) ||
        (rc = dev_alloc(c, (void **)&c->d_scratch, 4096 * sizeof(uint32_t))) ||
        (rc = dev_alloc(c, (void **)&c->d_partials, 3 * n_partials * sizeof(uint64_t))) ||
        (rc = dev_alloc(c, (void **)&c->d_result, 64 * sizeof(sjb_window_result_t))) ||
        (rc = dev_alloc(c, (void **)&c->d_rmi, sizeof(RmiParams))) ||
        (rc = dev_alloc(c, (void **)&c->d_heavy, 4 * kMaxHeavy * sizeof(uint64_t))) ||
        (rc = dev_alloc(c, (void **)&c->d_cursor, 4 * sizeof(uint64_t))) ||
        (rc = dev_alloc(c, (void **)&c->d_masks, 3 * c->mask_cap * sizeof(uint64_t))) ||
        (rc = host_alloc(c, (void **)&c->h_result, 64 * sizeof(sjb_window_result_t))))
        return fail(rc);
    if (launch_fill_u64(c->stream, c->d_cursor, 0, 4) != SJB_OK) return fail(SJB_E_HIP);
    partition_setup(c);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(SJB_E_HIP);
    *out = c;
    return SJB_OK;
}

int sjb_close(sjb_ctx_t *c) {
    if (!c) return SJB_E_INVALID;
    (void)hipSetDevice(c->device);
    if (c->pk_ready) (void)sjb_pk_stop(c);
    if (c->pk_stream) (void)hipStreamDestroy(c->pk_stream);
    if (c->pk_copy) (void)hipStreamDestroy(c->pk_copy);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    sjb::batch_plan_free(c->scratch_plan);
    c->scratch_plan = nullptr;
    for (void *p : c->dev_allocs) (void)hipFree(p);
    for (void *p : c->host_allocs) (void)hipHostFree(p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev2) (void)hipEventDestroy(c->ev2);
    if (c->ev3) (void)hipEventDestroy(c->ev3);
    for (hipEvent_t e : c->tring) (void)hipEventDestroy(e);
    if (c->ov_stream) {
        (void)hipStreamSynchronize(c->ov_stream);
        (void)hipStreamDestroy(c->ov_stream);
    }
    for (int b = 0; b < 2; b++) {
        if (c->ov_part_done[b]) (void)hipEventDestroy(c->ov_part_done[b]);
        if (c->ov_join_done[b]) (void)hipEventDestroy(c->ov_join_done[b]);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SJB_OK;
}

void *sjb_stream(sjb_ctx_t *c) { return c ? (void *)c->stream : nullptr; }

int sjb_sync(sjb_ctx_t *c) {
    if (!c) return SJB_E_INVALID;
    if (c->ov_stream) SJB_HIP_TRY(hipStreamSynchronize(c->ov_stream));
    SJB_HIP_TRY(hipStreamSynchronize(c->stream));
    return SJB_OK;
}

uint64_t sjb_device_bytes(sjb_ctx_t *c) { return c ? c->bytes : 0; }

int sjb_host_alloc(uint64_t bytes, void **p) {
    if (!p) return SJB_E_INVALID;
    *p = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return SJB_E_NODEVICE;
    }
    if (hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return SJB_E_NOMEM;
    }
    return SJB_OK;
}

int sjb_host_free(void *p) {
    if (!p) return SJB_OK;
    SJB_HIP_TRY(hipHostFree(p));
    return SJB_OK;
}

int sjb_lazy_join(sjb_ctx_t *c, int algorithm, int method, uint32_t flags, const sjb_tuple_t *d_left,
                  uint64_t n_left, const sjb_tuple_t *d_right, uint64_t n_right,
                  sjb_result_tuple_t *out, uint64_t out_cap, sjb_stats_t *stats) {
    if (!c || !stats || !valid_alg(algorithm) || !valid_method(method)) return SJB_E_INVALID;
    if ((n_left && !d_left) || (n_right && !d_right)) return SJB_E_INVALID;
    memset(stats, 0, sizeof(*stats));
    if (n_left > c->cap || n_right > c->cap) return SJB_E_CAPACITY;
    SJB_HIP_TRY(hipSetDevice(c->device));
    if ((n_left && !device_accessible(d_left)) || (n_right && !device_accessible(d_right)))
        return SJB_E_INVALID;
    return lazy_join_device(c, algorithm, method, flags, d_left, n_left, d_right, n_right, out,
                            out_cap, stats);
}

// Gathers ring slots into the ctx staging buffers.  Adjacent slots are
// merged into one copy (the reference copies one contiguous run from
// min(indexes) and is wrong when slots are not adjacent,
// src/sjb/windowing/lazy/executor/GPU/GPU_HJ.cu:235-241).
static int gather(sjb_ctx *c, sjb_tuple_t *dst, const sjb_tuple_t *ring, const uint64_t *idx,
                  uint64_t n_idx, uint64_t B) {
    uint64_t w = 0;
    for (uint64_t i = 0; i < n_idx;) {
        uint64_t j = i + 1;
        while (j < n_idx && idx[j] == idx[j - 1] + B) j++;
        const uint64_t cnt = (j - i) * B;
        SJB_HIP_TRY(hipMemcpyAsync(dst + w, ring + idx[i], cnt * sizeof(sjb_tuple_t),
                                   hipMemcpyHostToDevice, c->stream));
        w += cnt;
        i = j;
    }
    return SJB_OK;
}

int sjb_lazy_execute(sjb_ctx_t *c, int algorithm, int method, uint32_t flags,
                     const sjb_tuple_t *left_ring, const uint64_t *left_idx, uint64_t n_left_idx,
                     const sjb_tuple_t *right_ring, const uint64_t *right_idx, uint64_t n_right_idx,
                     uint64_t batch_size, sjb_result_tuple_t *out, uint64_t out_cap,
                     sjb_stats_t *stats) {
    if (!c || !stats || !valid_alg(algorithm) || !valid_method(method) || batch_size == 0)
        return SJB_E_INVALID;
    if ((n_left_idx && (!left_ring || !left_idx)) || (n_right_idx && (!right_ring || !right_idx)))
        return SJB_E_INVALID;
    memset(stats, 0, sizeof(*stats));
    const uint64_t nL = n_left_idx * batch_size, nR = n_right_idx * batch_size;
    if (nL > c->cap || nR > c->cap) return SJB_E_CAPACITY;
    SJB_HIP_TRY(hipSetDevice(c->device));
    int rc = gather(c, c->d_stage[0], left_ring, left_idx, n_left_idx, batch_size);
    if (rc) return rc;
    rc = gather(c, c->d_stage[1], right_ring, right_idx, n_right_idx, batch_size);
    if (rc) return rc;
    return lazy_join_device(c, algorithm, method, flags, c->d_stage[0], nL, c->d_stage[1], nR, out,
                            out_cap, stats);
}

int sjb_partition_report(sjb_ctx_t *c, int algorithm, uint32_t flags, const sjb_tuple_t *d_left,
                         uint64_t n_left, const sjb_tuple_t *d_right, uint64_t n_right,
                         sjb_partition_report_t *rep) {
    if (!c || !rep || (algorithm != SJB_ALG_HJ && algorithm != SJB_ALG_SMJ)) return SJB_E_INVALID;
    if (n_left == 0 || n_right == 0 || n_left > c->cap || n_right > c->cap) return SJB_E_INVALID;
    if (!device_accessible(d_left) || !device_accessible(d_right)) return SJB_E_INVALID;
    SJB_HIP_TRY(hipSetDevice(c->device));
    flags |= c->opts.flags;
    WindowPlan p;
    int rc = window_partition(c, algorithm, flags, d_left, n_left, d_right, n_right, 0, &p);
    if (rc) return rc;
    const int F = 1 << p.logF;
    uint32_t *d_sizes = reinterpret_cast<uint32_t *>(c->d_partials);  // 2F u32
    rc = launch_part_sizes(c->stream, c->d_offs, p.logF, p.ncL, p.ncR, d_sizes);
    if (rc) return rc;
    std::vector<uint32_t> sz(2 * size_t(F));
    uint64_t heavy[4 * kMaxHeavy] = {};
    SJB_HIP_TRY(hipMemcpyAsync(sz.data(), d_sizes, sz.size() * 4, hipMemcpyDeviceToHost, c->stream));
    if (p.peel)
        SJB_HIP_TRY(hipMemcpyAsync(heavy, c->d_heavy, sizeof(heavy), hipMemcpyDeviceToHost, c->stream));
    SJB_HIP_TRY(hipStreamSynchronize(c->stream));
    const uint64_t cap = algorithm == SJB_ALG_SMJ ? uint64_t(kSmjCap) : uint64_t(kHjCap);
    uint64_t tot = 0, over = 0, mxb = 0, mxp = 0;
    for (int i = 0; i < F; i++) {
        tot += sz[i];
        mxb = std::max<uint64_t>(mxb, sz[i]);
        mxp = std::max<uint64_t>(mxp, sz[F + i]);
        if (sz[i] > cap) over += sz[i];
    }
    std::memset(rep, 0, sizeof(*rep));
    rep->fanout_log2 = uint32_t(p.logF);
    rep->kind = uint32_t(p.kind);
    rep->max_build = mxb;
    rep->max_probe = mxp;
    rep->mean_build = double(tot) / F;
    rep->imbalance = tot ? double(mxb) / rep->mean_build : 1.0;
    rep->fixup_frac = tot ? double(over) / double(tot) : 0.0;
    for (int h = 0; h < kMaxHeavy; h++) {
        rep->heavy_left += heavy[4 * h];
        rep->heavy_right += heavy[4 * h + 2];
    }
    return SJB_OK;
}

int sjb_lazy_join_async(sjb_ctx_t *c, int algorithm, uint32_t flags, const sjb_tuple_t *d_left,
                        uint64_t n_left, const sjb_tuple_t *d_right, uint64_t n_right,
                        sjb_window_result_t *d_result) {
    if (!c || !d_result || !valid_alg(algorithm)) return SJB_E_INVALID;
    if ((n_left && !d_left) || (n_right && !d_right)) return SJB_E_INVALID;
    if (n_left > c->cap || n_right > c->cap) return SJB_E_CAPACITY;
    if ((n_left && !device_accessible(d_left)) || (n_right && !device_accessible(d_right)) ||
        !device_accessible(d_result))
        return SJB_E_INVALID;
    int logF;
    flags |= c->opts.flags;
    const hipEvent_t *ev = nullptr;
    int buf = -1;
    if ((flags & SJB_F_OVERLAP) && algorithm != SJB_ALG_NLJ) {
        SJB_HIP_TRY(hipSetDevice(c->device));
        int rc = overlap_setup(c);
        if (rc) return rc;
        buf = int(c->ov_next++ & 1);
    }
    if ((flags & SJB_F_TIMING) && !c->tring.empty()) {
        const uint64_t slot = c->tring_next % (c->tring.size() / 4);
        ev = &c->tring[4 * slot];
        c->tring_split[slot] = buf >= 0;
        c->tring_next++;
    }
    return window_enqueue(c, algorithm, flags, d_left, n_left, d_right, n_right, 0, nullptr, 0, 0,
                          d_result, &logF, ev, buf);
}

int sjb_timing_ring(sjb_ctx_t *c, uint32_t n) {
    if (!c || n > (1u << 20)) return SJB_E_INVALID;
    SJB_HIP_TRY(hipSetDevice(c->device));
    SJB_HIP_TRY(hipStreamSynchronize(c->stream));
    for (hipEvent_t e : c->tring) (void)hipEventDestroy(e);
    c->tring.clear();
    c->tring_split.assign(n, 0);
    c->tring_next = 0;
    // Timestamp-only events: no system-scope fence (cache writeback and
    // invalidate) when they are recorded, so timing a window costs it
    // almost nothing; they are never used to synchronise with the host.
    for (uint32_t i = 0; i < 4 * n; i++) {
        hipEvent_t e;
        SJB_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        c->tring.push_back(e);
    }
    return SJB_OK;
}

int sjb_timing_read(sjb_ctx_t *c, float *part_ms, float *join_ms, uint32_t n, uint32_t *n_read) {
    if (!c || !part_ms || !join_ms || !n_read) return SJB_E_INVALID;
    SJB_HIP_TRY(hipSetDevice(c->device));
    SJB_HIP_TRY(hipStreamSynchronize(c->stream));
    const uint64_t slots = c->tring.size() / 4;
    const uint64_t have = std::min<uint64_t>({uint64_t(n), slots, c->tring_next});
    for (uint64_t i = 0; i < have; i++) {
        const uint64_t w = c->tring_next - have + i;  // oldest first
        const hipEvent_t *e = &c->tring[4 * (w % slots)];
        SJB_HIP_TRY(hipEventElapsedTime(&part_ms[i], e[0], e[1]));
        SJB_HIP_TRY(hipEventElapsedTime(&join_ms[i], c->tring_split[w % slots] ? e[2] : e[1], e[3]));
    }
    *n_read = uint32_t(have);
    return SJB_OK;
}

int sjb_lazy_join_batch(sjb_ctx_t *c, int algorithm, uint32_t flags, const sjb_tuple_t *d_left,
                        const uint64_t *left_off, const sjb_tuple_t *d_right,
                        const uint64_t *right_off, uint32_t n_windows, sjb_window_result_t *results) {
    if (!c || !left_off || !right_off || !results || !valid_alg(algorithm))
        return SJB_E_INVALID;
    return batch_join(c, algorithm, flags | c->opts.flags, d_left, left_off, d_right, right_off,
                      n_windows, results);
}

uint64_t sjb_mix64(uint64_t z) { return mix64(z); }
uint64_t sjb_hash_a(uint32_t k, uint32_t v) { return hash_a(k, v); }
uint64_t sjb_hash_b(uint32_t k, uint32_t v) { return hash_b(k, v); }
uint64_t sjb_pair_hash(uint32_t k, uint32_t rv, uint32_t sv) {
    return pair_hash_from_a(hash_a(k, rv), sv);
}

}  // extern "C"
