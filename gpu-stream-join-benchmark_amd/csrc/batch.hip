// batch.hip — many uneven windows per launch (count/checksums).
//
// The reference joins one window per executor call
// (include/sjb/windowing/lazy/CountBased.cuh:135-138): every window pays its
// own launches and host synchronisations.  Here a batch of windows is one
// sequence of three kernels on the ctx stream:
//
//   k_batch_clear      zero the per-window results and the queue heads
//   k_partition_batch  one workgroup per (window, side, chunk) item
//   k_join_batch       persistent workgroups draining per-XCD queues of
//                      (window, partition) items
//
// A plan (sjb_batch_plan_create) fixes the windows, buffers and launch
// shapes once, uploads the descriptors and captures the three launches in a
// hipGraph; sjb_batch_plan_launch replays it.  The launch shape (fanout
// bias, partition chunk, resident join workgroups, queue grab) comes from
// the ctx's tuning, which sjb_batch_autotune measures.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "sjb_device.hpp"
#include "sjb_internal.hpp"

struct sjb_batch_plan {
    sjb_ctx *c = nullptr;
    int alg = 0;
    uint32_t flags = 0;
    int kind = 0;
    sjb::BatchTune tune;
    uint32_t n_win = 0;
    sjb_window_result_t *d_res = nullptr;
    // device buffers (grow-only, owned by the plan)
    sjb::BatchWin *d_win = nullptr;
    uint32_t *d_chunk_win = nullptr, *d_item_win = nullptr;
    sjb::Rec *d_part = nullptr;
    uint32_t *d_offs = nullptr;
    uint32_t *d_qhead = nullptr;
    uint64_t *d_heavy = nullptr;      // learned + grouped: per-window heavy-key sums
    size_t cap_heavy = 0;
    bool peel = false;
    bool drift_fallback = false;      // learned CDF judged stale on this batch, radix ranges used
    bool drift_refit = false;         // ... and refitted to this batch's key sample instead
    bool peel_sampled = false;        // SJB_F_PEEL: heavy keys of this batch's sample, no model
    sjb::RmiParams rmi_refit{};       // the refitted model (host copy)
    sjb::RmiParams *d_rmi = nullptr;  // ... and on the device
    size_t cap_rmi = 0;
    uint32_t n_heavy = 0;
    size_t cap_win = 0, cap_chunks = 0, cap_items = 0, cap_part = 0, cap_offs = 0;
    uint64_t dev_bytes = 0;
    // pinned staging of the descriptors
    void *h_stage = nullptr;
    size_t h_stage_cap = 0;
    sjb::BatchArgs args{};
    uint32_t n_chunks = 0, n_items = 0;
    int blocks = 0;
    size_t dyn_part = 0, dyn_join = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace sjb {

template <class T>
static int grow(sjb_batch_plan *p, T **ptr, size_t *cap, size_t need) {
    if (need <= *cap && *ptr) return SJB_OK;
    size_t n = std::max<size_t>(need, 1);
    if (*ptr) {
        (void)hipFree(*ptr);
        p->dev_bytes -= *cap * sizeof(T);
        p->c->ext_bytes -= *cap * sizeof(T);
        *ptr = nullptr;
        *cap = 0;
    }
    const uint64_t budget = p->c->opts.hbm_budget_bytes;
    if (budget && p->c->bytes + p->c->ext_bytes + n * sizeof(T) > budget) return SJB_E_BUDGET;
    if (hipMalloc(reinterpret_cast<void **>(ptr), n * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return SJB_E_NOMEM;
    }
    *cap = n;
    p->dev_bytes += n * sizeof(T);
    p->c->ext_bytes += n * sizeof(T);
    return SJB_OK;
}

void batch_plan_free(sjb_batch_plan *p) {
    if (!p) return;
    if (p->c) p->c->ext_bytes -= p->dev_bytes;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    void *dev[] = {p->d_win, p->d_chunk_win, p->d_item_win, p->d_part, p->d_offs, p->d_qhead, p->d_heavy,
                   p->d_rmi};
    for (void *d : dev)
        if (d) (void)hipFree(d);
    if (p->h_stage) (void)hipHostFree(p->h_stage);
    delete p;
}

// Host planning: per-window fanout and chunk counts, the deal of windows
// over the kXcds queues (largest first, each to the least loaded queue),
// descriptor order (queue by queue, decreasing size), item maps, buffer
// sizes; uploads descriptors and maps on the ctx stream.
static int plan_build(sjb_batch_plan *p, const sjb_tuple_t *L, const uint64_t *loff,
                      const sjb_tuple_t *R, const uint64_t *roff) {
    sjb_ctx *c = p->c;
    const uint32_t n = p->n_win;
    const uint32_t chunk = 1u << p->tune.chunk_log2;
    const uint64_t max_side = uint64_t(chunk) * kMaxChunks;
    for (uint32_t w = 0; w < n; w++) {
        if (loff[w + 1] < loff[w] || roff[w + 1] < roff[w]) return SJB_E_INVALID;
        if (loff[w + 1] - loff[w] > max_side || roff[w + 1] - roff[w] > max_side) return SJB_E_CAPACITY;
    }
    const uint64_t nl_tot = loff[n] - loff[0], nr_tot = roff[n] - roff[0];
    if (nl_tot + nr_tot >= (1ull << 32)) return SJB_E_CAPACITY;
    // The learned CDF judges itself on this batch's keys (an evenly spaced
    // sample of the left stream).  Stale after a drift, it is refitted to
    // that sample (root = the sample's CDF, leaves = chords of it; a
    // plan-private copy), or — with SJB_F_NO_REFIT, or when the refit does
    // not beat radix — the batch uses radix ranges.  Peeled heavy keys stay.
    p->drift_fallback = false;
    p->drift_refit = false;
    p->peel_sampled = false;
    if (p->kind != PART_LEARNED && (p->flags & SJB_F_PEEL) && (p->flags & SJB_F_GROUPED) && nl_tot > 0) {
        // model-free peeling: heavy keys of this batch's own key sample,
        // radix ranges (SMJ) or hash partitions (HJ) for the rest
        uint32_t nh = 0;
        int rc = sampled_heavy(c, L + loff[0], nl_tot, &p->rmi_refit, &nh);
        if (rc) return rc;
        if (nh > 0) {
            if ((rc = grow(p, &p->d_rmi, &p->cap_rmi, 1))) return rc;
            SJB_HIP_TRY(hipMemcpyAsync(p->d_rmi, &p->rmi_refit, sizeof(RmiParams), hipMemcpyHostToDevice,
                                       c->stream));
            p->peel = true;
            p->peel_sampled = true;
            p->n_heavy = nh;
        }
    }
    if (p->kind == PART_LEARNED && nl_tot > 0) {
        std::vector<uint32_t> keys;
        int rc = rmi_sample_keys(c, L + loff[0], nl_tot, keys);
        if (rc) return rc;
        const int verdict = rmi_judge(c, keys, p->peel, !(p->flags & SJB_F_NO_REFIT), &p->rmi_refit);
        if (verdict == kDriftRefit) {
            if ((rc = grow(p, &p->d_rmi, &p->cap_rmi, 1))) return rc;
            SJB_HIP_TRY(hipMemcpyAsync(p->d_rmi, &p->rmi_refit, sizeof(RmiParams), hipMemcpyHostToDevice,
                                       c->stream));
            p->drift_refit = true;
        } else if (verdict == kDriftRanges) {
            p->kind = PART_RANGE;
            p->drift_fallback = true;
        }
    }

    struct W {
        uint32_t nL, nR, ncL, ncR, logF;
        uint64_t cost;
    };
    std::vector<W> ws(n);
    for (uint32_t w = 0; w < n; w++) {
        W &x = ws[w];
        x.nL = uint32_t(loff[w + 1] - loff[w]);
        x.nR = uint32_t(roff[w + 1] - roff[w]);
        const bool empty = x.nL == 0 || x.nR == 0;
        x.logF = empty ? 0 : uint32_t(batch_log_fanout(p->alg, p->flags, x.nL, p->tune.fanout_bias));
        x.ncL = empty ? 0 : (x.nL + chunk - 1) / chunk;
        x.ncR = empty ? 0 : (x.nR + chunk - 1) / chunk;
        // join cost ~ tuples read plus a per-partition overhead
        x.cost = empty ? 0 : uint64_t(x.nL) + x.nR + (uint64_t(256) << x.logF);
    }
    std::vector<uint32_t> by_cost(n);
    std::iota(by_cost.begin(), by_cost.end(), 0u);
    std::stable_sort(by_cost.begin(), by_cost.end(),
                     [&](uint32_t a, uint32_t b) { return ws[a].cost > ws[b].cost; });
    std::vector<std::vector<uint32_t>> bins(kXcds);
    uint64_t load[kXcds] = {};
    for (uint32_t w : by_cost) {
        int best = 0;
        for (int x = 1; x < kXcds; x++)
            if (load[x] < load[best]) best = x;
        bins[best].push_back(w);
        load[best] += ws[w].cost;
    }

    std::vector<BatchWin> desc;
    desc.reserve(n);
    uint64_t offs_words = 0;
    uint32_t n_chunks = 0, n_items = 0;
    int max_ncL = 0, max_ncR = 0, max_logF = 0;
    for (int x = 0; x < kXcds; x++) {
        p->args.qbeg[x] = n_items;
        for (uint32_t w : bins[x]) {
            const W &s = ws[w];
            BatchWin d{};
            d.lo = loff[w];
            d.ro = roff[w];
            d.offs_base = offs_words;
            d.lpart = uint32_t(loff[w] - loff[0]);
            d.rpart = uint32_t(nl_tot + roff[w] - roff[0]);
            d.nL = s.nL;
            d.nR = s.nR;
            d.ncL = s.ncL;
            d.ncR = s.ncR;
            d.chunk_base = n_chunks;
            d.item_base = n_items;
            d.id = w;
            d.logF = s.logF;
            desc.push_back(d);
            const bool empty = s.ncL == 0;
            offs_words += uint64_t(s.ncL + s.ncR) * ((1u << s.logF) + 1);
            n_chunks += s.ncL + s.ncR;
            n_items += empty ? 0 : (1u << s.logF);
            max_ncL = std::max<int>(max_ncL, int(s.ncL));
            max_ncR = std::max<int>(max_ncR, int(s.ncR));
            max_logF = std::max<int>(max_logF, int(s.logF));
        }
    }
    p->args.qbeg[kXcds] = n_items;
    p->n_chunks = n_chunks;
    p->n_items = n_items;

    std::vector<uint32_t> chunk_win(n_chunks), item_win(n_items);
    for (uint32_t i = 0; i < uint32_t(desc.size()); i++) {
        const BatchWin &d = desc[i];
        for (uint32_t k = 0; k < d.ncL + d.ncR; k++) chunk_win[d.chunk_base + k] = i;
        if (d.ncL)
            for (uint32_t k = 0; k < (1u << d.logF); k++) item_win[d.item_base + k] = i;
    }

    int rc;
    if ((rc = grow(p, &p->d_win, &p->cap_win, std::max<size_t>(n, 1))) ||
        (rc = grow(p, &p->d_chunk_win, &p->cap_chunks, std::max<size_t>(n_chunks, 1))) ||
        (rc = grow(p, &p->d_item_win, &p->cap_items, std::max<size_t>(n_items, 1))) ||
        (rc = grow(p, &p->d_part, &p->cap_part, std::max<size_t>(nl_tot + nr_tot, 1))) ||
        (rc = grow(p, &p->d_offs, &p->cap_offs, std::max<size_t>(offs_words, 1))))
        return rc;
    if (!p->d_qhead) {
        size_t cap = 0;
        if ((rc = grow(p, &p->d_qhead, &cap, kXcds * 16))) return rc;
    }
    if (p->peel && (rc = grow(p, &p->d_heavy, &p->cap_heavy, size_t(n) * kMaxHeavy * 4))) return rc;

    // stage descriptors + maps in pinned memory, one H2D copy each
    const size_t bytes_win = size_t(n) * sizeof(BatchWin);
    const size_t bytes_cw = size_t(n_chunks) * 4, bytes_iw = size_t(n_items) * 4;
    const size_t need = bytes_win + bytes_cw + bytes_iw + 64;
    SJB_HIP_TRY(hipStreamSynchronize(c->stream));  // staging may still feed a copy
    if (need > p->h_stage_cap) {
        if (p->h_stage) (void)hipHostFree(p->h_stage);
        p->h_stage = nullptr;
        p->h_stage_cap = 0;
        if (hipHostMalloc(&p->h_stage, need, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            p->h_stage = nullptr;
            return SJB_E_NOMEM;
        }
        p->h_stage_cap = need;
    }
    uint8_t *h = static_cast<uint8_t *>(p->h_stage);
    if (bytes_win) std::memcpy(h, desc.data(), bytes_win);
    if (bytes_cw) std::memcpy(h + bytes_win, chunk_win.data(), bytes_cw);
    if (bytes_iw) std::memcpy(h + bytes_win + bytes_cw, item_win.data(), bytes_iw);
    if (bytes_win)
        SJB_HIP_TRY(hipMemcpyAsync(p->d_win, h, bytes_win, hipMemcpyHostToDevice, c->stream));
    if (bytes_cw)
        SJB_HIP_TRY(hipMemcpyAsync(p->d_chunk_win, h + bytes_win, bytes_cw, hipMemcpyHostToDevice,
                                   c->stream));
    if (bytes_iw)
        SJB_HIP_TRY(hipMemcpyAsync(p->d_item_win, h + bytes_win + bytes_cw, bytes_iw,
                                   hipMemcpyHostToDevice, c->stream));

    BatchArgs &a = p->args;
    a.L = L;
    a.R = R;
    a.win = p->d_win;
    a.chunk_win = p->d_chunk_win;
    a.item_win = p->d_item_win;
    a.part = p->d_part;
    a.offs = p->d_offs;
    a.res = p->d_res;
    a.qhead = p->d_qhead;
    a.chunk = chunk;
    a.grab = uint32_t(std::max(p->tune.grab, 1));
    a.expand = uint32_t(p->tune.expand);
    const bool grouped = (p->flags & SJB_F_GROUPED) != 0;
    a.idcap = uint32_t(join_idcap(p->alg, grouped));
    a.rmi = (p->drift_refit || p->peel_sampled) ? p->d_rmi : c->d_rmi;
    a.heavy = p->peel ? p->d_heavy : nullptr;
    p->dyn_part = partition_lds_bytes_chunk(max_logF, int(chunk));
    p->dyn_join = join_view_bytes(p->alg, grouped, max_ncL, max_ncR);
    p->blocks = int(std::min<uint64_t>(uint64_t(c->n_cu) * std::max(p->tune.blocks_per_cu, 1),
                                       std::max<uint32_t>(n_items, 1)));
    return SJB_OK;
}

static int plan_enqueue(sjb_batch_plan *p, hipStream_t s) {
    int rc = launch_batch_clear(s, p->d_res, p->n_win, p->d_qhead, p->peel ? p->d_heavy : nullptr);
    if (rc) return rc;
    rc = launch_partition_batch(s, p->kind, p->peel, p->args, p->n_chunks, p->dyn_part);
    if (rc) return rc;
    if (p->n_items) {
        rc = launch_join_batch(s, p->alg, p->flags, p->args, p->blocks, p->dyn_join);
        if (rc) return rc;
    }
    if (p->peel) return launch_heavy_finalize(s, p->d_res, p->n_win, p->d_heavy, p->n_heavy);
    return SJB_OK;
}

static int plan_capture(sjb_batch_plan *p) {
    hipStream_t s = p->c->stream;
    SJB_HIP_TRY(hipStreamSynchronize(s));
    SJB_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = plan_enqueue(p, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    SJB_HIP_TRY(e);
    p->graph = g;
    SJB_HIP_TRY(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
    return SJB_OK;
}

static int plan_init(sjb_ctx *c, sjb_batch_plan *p, int alg, uint32_t flags, uint32_t n,
                     sjb_window_result_t *d_res) {
    if (alg != SJB_ALG_HJ && alg != SJB_ALG_SMJ) return SJB_E_UNSUPPORTED;
    flags |= c->opts.flags;
    p->c = c;
    p->alg = alg;
    p->flags = flags;
    const bool learned = (flags & SJB_F_LEARNED) && c->rmi_valid;
    p->kind = part_kind(alg, learned);
    p->peel = learned && (flags & SJB_F_GROUPED) && c->rmi_nheavy > 0;
    p->n_heavy = p->peel ? c->rmi_nheavy : 0;
    p->tune = c->tune;
    p->n_win = n;
    p->d_res = d_res;
    return SJB_OK;
}

static int check_batch_args(sjb_ctx *c, const sjb_tuple_t *L, const uint64_t *loff,
                            const sjb_tuple_t *R, const uint64_t *roff, uint32_t n) {
    if (!c || !loff || !roff) return SJB_E_INVALID;
    if (n == 0) return SJB_OK;
    if (loff[n] < loff[0] || roff[n] < roff[0]) return SJB_E_INVALID;
    if ((loff[n] > loff[0] && !device_accessible(L)) || (roff[n] > roff[0] && !device_accessible(R)))
        return SJB_E_INVALID;
    return SJB_OK;
}

static int ensure_results(sjb_ctx *c, uint32_t n) {
    if (c->batch_res_cap >= n) return SJB_OK;
    void *p = nullptr;
    int rc = dev_alloc(c, &p, size_t(n) * sizeof(sjb_window_result_t));
    if (rc) return rc;
    c->d_batch_res = static_cast<sjb_window_result_t *>(p);
    c->batch_res_cap = n;
    return SJB_OK;
}

// Direct (uncaptured) batched join into c->d_batch_res with the ctx's
// scratch plan and a given tuning.
static int batch_run(sjb_ctx *c, int alg, uint32_t flags, const sjb_tuple_t *L, const uint64_t *loff,
                     const sjb_tuple_t *R, const uint64_t *roff, uint32_t n, const BatchTune &t) {
    int rc = ensure_results(c, n);
    if (rc) return rc;
    if (!c->scratch_plan) c->scratch_plan = new sjb_batch_plan();
    sjb_batch_plan *p = c->scratch_plan;
    if ((rc = plan_init(c, p, alg, flags, n, c->d_batch_res))) return rc;
    p->tune = t;
    if ((rc = plan_build(p, L, loff, R, roff))) return rc;
    return plan_enqueue(p, c->stream);
}

int batch_join(sjb_ctx *c, int alg, uint32_t flags, const sjb_tuple_t *L, const uint64_t *loff,
               const sjb_tuple_t *R, const uint64_t *roff, uint32_t n_windows,
               sjb_window_result_t *results) {
    int rc = check_batch_args(c, L, loff, R, roff, n_windows);
    if (rc || n_windows == 0) return rc;
    SJB_HIP_TRY(hipSetDevice(c->device));
    if (alg == SJB_ALG_NLJ) {
        // NLJ has no partitions: windows are enqueued back to back
        if ((rc = ensure_results(c, n_windows))) return rc;
        for (uint32_t w = 0; w < n_windows; w++) {
            if (loff[w + 1] < loff[w] || roff[w + 1] < roff[w]) return SJB_E_INVALID;
            if (loff[w + 1] - loff[w] > c->cap || roff[w + 1] - roff[w] > c->cap) return SJB_E_CAPACITY;
        }
        for (uint32_t w = 0; w < n_windows; w++) {
            int logF;
            rc = window_enqueue(c, alg, flags | c->opts.flags, L + loff[w], loff[w + 1] - loff[w],
                                R + roff[w], roff[w + 1] - roff[w], 0, nullptr, 0, 0,
                                c->d_batch_res + w, &logF, nullptr);
            if (rc) return rc;
        }
    } else {
        rc = batch_run(c, alg, flags, L, loff, R, roff, n_windows, c->tune);
        if (rc) return rc;
    }
    SJB_HIP_TRY(hipMemcpyAsync(results, c->d_batch_res, size_t(n_windows) * sizeof(sjb_window_result_t),
                               hipMemcpyDeviceToHost, c->stream));
    SJB_HIP_TRY(hipStreamSynchronize(c->stream));
    return SJB_OK;
}

}  // namespace sjb

using namespace sjb;

extern "C" {

int sjb_batch_plan_create(sjb_ctx_t *c, int alg, uint32_t flags, const sjb_tuple_t *d_left,
                          const uint64_t *left_off, const sjb_tuple_t *d_right,
                          const uint64_t *right_off, uint32_t n_windows,
                          sjb_window_result_t *d_results, sjb_batch_plan_t **out) {
    if (!out) return SJB_E_INVALID;
    *out = nullptr;
    int rc = check_batch_args(c, d_left, left_off, d_right, right_off, n_windows);
    if (rc) return rc;
    if (n_windows == 0 || !d_results || !device_accessible(d_results)) return SJB_E_INVALID;
    SJB_HIP_TRY(hipSetDevice(c->device));
    sjb_batch_plan *p = new sjb_batch_plan();
    if ((rc = plan_init(c, p, alg, flags, n_windows, d_results)) ||
        (rc = plan_build(p, d_left, left_off, d_right, right_off)) || (rc = plan_capture(p))) {
        (void)hipStreamSynchronize(c->stream);
        batch_plan_free(p);
        return rc;
    }
    SJB_HIP_TRY(hipStreamSynchronize(c->stream));
    *out = p;
    return SJB_OK;
}

int sjb_batch_plan_launch(sjb_batch_plan_t *p) {
    if (!p || !p->exec) return SJB_E_INVALID;
    SJB_HIP_TRY(hipSetDevice(p->c->device));
    SJB_HIP_TRY(hipGraphLaunch(p->exec, p->c->stream));
    return SJB_OK;
}

int sjb_batch_plan_info(sjb_batch_plan_t *p, sjb_batch_info_t *info) {
    if (!p || !info) return SJB_E_INVALID;
    info->n_windows = p->n_win;
    info->n_chunks = p->n_chunks;
    info->n_items = p->n_items;
    info->blocks = uint32_t(p->blocks);
    info->device_bytes = p->dev_bytes;
    info->part_kind = uint32_t(p->kind);
    info->drift_fallback = p->drift_fallback ? 1u : 0u;
    info->drift_refit = p->drift_refit ? 1u : 0u;
    info->n_heavy = p->peel ? p->n_heavy : 0u;
    info->tune.fanout_bias = p->tune.fanout_bias;
    info->tune.chunk_log2 = uint32_t(p->tune.chunk_log2);
    info->tune.blocks_per_cu = uint32_t(p->tune.blocks_per_cu);
    info->tune.grab = uint32_t(p->tune.grab);
    info->tune.expand = uint32_t(p->tune.expand);
    info->tune.ms = 0.0;
    return SJB_OK;
}

int sjb_batch_plan_destroy(sjb_batch_plan_t *p) {
    if (!p) return SJB_OK;
    (void)hipSetDevice(p->c->device);
    (void)hipStreamSynchronize(p->c->stream);
    batch_plan_free(p);
    return SJB_OK;
}

int sjb_batch_get_tune(sjb_ctx_t *c, sjb_batch_tune_t *t) {
    if (!c || !t) return SJB_E_INVALID;
    t->fanout_bias = c->tune.fanout_bias;
    t->chunk_log2 = uint32_t(c->tune.chunk_log2);
    t->blocks_per_cu = uint32_t(c->tune.blocks_per_cu);
    t->grab = uint32_t(c->tune.grab);
    t->expand = uint32_t(c->tune.expand);
    t->reserved = 0;
    t->ms = 0.0;
    return SJB_OK;
}

int sjb_batch_set_tune(sjb_ctx_t *c, const sjb_batch_tune_t *t) {
    if (!c || !t) return SJB_E_INVALID;
    if (t->fanout_bias < -4 || t->fanout_bias > 4) return SJB_E_INVALID;
    if (t->chunk_log2 != 12 && t->chunk_log2 != 14) return SJB_E_INVALID;
    if (t->blocks_per_cu < 1 || t->blocks_per_cu > 3 || t->grab < 1 || t->grab > 64) return SJB_E_INVALID;
    if (t->expand > 1) return SJB_E_INVALID;
    c->tune.fanout_bias = t->fanout_bias;
    c->tune.chunk_log2 = int(t->chunk_log2);
    c->tune.blocks_per_cu = int(t->blocks_per_cu);
    c->tune.grab = int(t->grab);
    c->tune.expand = int(t->expand);
    return SJB_OK;
}

// Measures every candidate launch shape on the given batch (one untimed
// run, then `reps` timed runs, median), checks that every candidate gives
// bit-identical per-window results, keeps the fastest in the ctx and
// reports it.  Candidates: fanout bias {-1,0,+1} x chunk {2^12,2^14} x
// join workgroups per CU {1,2,3} x queue grab {1,4} (x long match lists on
// full waves {off, on} for non-grouped HJ and SMJ with C2).
int sjb_batch_autotune(sjb_ctx_t *c, int alg, uint32_t flags, const sjb_tuple_t *d_left,
                       const uint64_t *left_off, const sjb_tuple_t *d_right,
                       const uint64_t *right_off, uint32_t n_windows, uint32_t reps,
                       sjb_batch_tune_t *best, uint32_t *n_tried) {
    int rc = check_batch_args(c, d_left, left_off, d_right, right_off, n_windows);
    if (rc) return rc;
    if (n_windows == 0 || !best || (alg != SJB_ALG_HJ && alg != SJB_ALG_SMJ)) return SJB_E_INVALID;
    SJB_HIP_TRY(hipSetDevice(c->device));
    reps = std::max<uint32_t>(reps, 1);
    std::vector<sjb_window_result_t> ref(n_windows), got(n_windows);
    BatchTune best_t = c->tune;
    double best_ms = 1e300;
    uint32_t tried = 0;
    // non-grouped HJ and SMJ with C2 also try running long match lists on
    // full waves (it pays under skewed keys, costs a little without)
    const uint32_t fl = flags | c->opts.flags;
    const int n_expand =
        !(fl & SJB_F_GROUPED) && (alg == SJB_ALG_HJ || (fl & SJB_F_C2)) ? 2 : 1;
    for (int bias = -1; bias <= 1; bias++)
        for (int cl : {12, 14})
            for (int bpc : {1, 2, 3})
                for (int grab : {1, 4})
                for (int ex = 0; ex < n_expand; ex++) {
                    BatchTune t;
                    t.fanout_bias = bias;
                    t.chunk_log2 = cl;
                    t.blocks_per_cu = bpc;
                    t.grab = grab;
                    t.expand = ex;
                    rc = batch_run(c, alg, flags, d_left, left_off, d_right, right_off, n_windows, t);
                    if (rc == SJB_E_CAPACITY) continue;  // window too large for this chunk
                    if (rc) return rc;
                    std::vector<float> ms(reps);
                    for (uint32_t r = 0; r < reps; r++) {
                        SJB_HIP_TRY(hipEventRecord(c->ev0, c->stream));
                        rc = plan_enqueue(c->scratch_plan, c->stream);
                        if (rc) return rc;
                        SJB_HIP_TRY(hipEventRecord(c->ev1, c->stream));
                        SJB_HIP_TRY(hipEventSynchronize(c->ev1));
                        SJB_HIP_TRY(hipEventElapsedTime(&ms[r], c->ev0, c->ev1));
                    }
                    SJB_HIP_TRY(hipMemcpyAsync(got.data(), c->d_batch_res,
                                               size_t(n_windows) * sizeof(sjb_window_result_t),
                                               hipMemcpyDeviceToHost, c->stream));
                    SJB_HIP_TRY(hipStreamSynchronize(c->stream));
                    if (tried == 0)
                        ref = got;
                    else if (std::memcmp(ref.data(), got.data(), n_windows * sizeof(sjb_window_result_t)))
                        return SJB_E_MISMATCH;
                    tried++;
                    std::sort(ms.begin(), ms.end());
                    const double med = ms[reps / 2];
                    if (med < best_ms) {
                        best_ms = med;
                        best_t = t;
                    }
                }
    if (tried == 0) return SJB_E_CAPACITY;
    c->tune = best_t;
    best->fanout_bias = best_t.fanout_bias;
    best->chunk_log2 = uint32_t(best_t.chunk_log2);
    best->blocks_per_cu = uint32_t(best_t.blocks_per_cu);
    best->grab = uint32_t(best_t.grab);
    best->expand = uint32_t(best_t.expand);
    best->reserved = 0;
    best->ms = best_ms;
    if (n_tried) *n_tried = tried;
    return SJB_OK;
}

}  // extern "C"
