// sjb_internal.hpp — host-side internals shared by the translation units
// of libsjb_hip.so: the context object, launch helpers and the kernel
// launcher declarations.  Nothing here crosses the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "sjb.h"

#define SJB_HIP_TRY(expr)                                   \
    do {                                                    \
        hipError_t e_ = (expr);                             \
        if (e_ != hipSuccess) return sjb::hip_fail(e_, #expr, __FILE__, __LINE__); \
    } while (0)

namespace sjb {

int hip_fail(hipError_t e, const char *what, const char *file, int line);

// Partition record: key plus a 32-bit payload (the tuple's val for
// count/checksum work, or its index in the window for pair output).
typedef uint2 Rec;

// Tuning constants (overridable at build time with -DSJB_*).
#ifndef SJB_CHUNK
#define SJB_CHUNK 16384
#endif
#ifndef SJB_PART_THREADS
#define SJB_PART_THREADS 1024
#endif
#ifndef SJB_PART_PIPE
#define SJB_PART_PIPE 1
#endif
#ifndef SJB_JOIN_THREADS
#define SJB_JOIN_THREADS 512
#endif
#ifndef SJB_HJ_CAP
#define SJB_HJ_CAP 2560
#endif
#ifndef SJB_HJ_LOGNB
#define SJB_HJ_LOGNB 11
#endif
#ifndef SJB_SMJ_CAP
#define SJB_SMJ_CAP 2048
#endif
#ifndef SJB_SMALL_CHUNK
#define SJB_SMALL_CHUNK 2048
#endif
#ifndef SJB_SMALL_WINDOW
#define SJB_SMALL_WINDOW (1 << 20)
#endif
#ifndef SJB_TINY_CHUNK
#define SJB_TINY_CHUNK 1024
#endif
#ifndef SJB_TINY_WINDOW
#define SJB_TINY_WINDOW (1 << 17)
#endif
constexpr int kChunk = SJB_CHUNK;              // tuples per partition-pass workgroup
// Windows of at most kSmallWindow tuples (both sides) are partitioned in
// kSmallChunk-tuple chunks, one workgroup each, and those of at most
// kTinyWindow in kTinyChunk chunks: in kChunk chunks a 2^17-tuple window
// would keep 8 of the chip's workgroups busy.  (Smaller chunks mean longer
// partition views in the join, so larger windows keep larger chunks.)
constexpr int kSmallChunk = SJB_SMALL_CHUNK;
constexpr uint64_t kSmallWindow = SJB_SMALL_WINDOW;
constexpr int kTinyChunk = SJB_TINY_CHUNK;
constexpr uint64_t kTinyWindow = SJB_TINY_WINDOW;
constexpr int kPartThreads = SJB_PART_THREADS; // threads of the partition pass
constexpr int kMaxLogF = 12;                   // fanout bound: 4096 partitions
constexpr int kMaxChunks = 1024;               // chunks per side
constexpr int kJoinThreads = SJB_JOIN_THREADS; // threads of the per-partition join
constexpr int kHjCap = SJB_HJ_CAP;             // build entries per LDS tile (HJ)
constexpr int kHjLogNB = SJB_HJ_LOGNB;         // log2 LDS hash buckets (HJ)
constexpr int kSmjCap = SJB_SMJ_CAP;           // elements per LDS sort tile (SMJ)

enum PartKind : int { PART_HASH = 0, PART_RANGE = 1, PART_LEARNED = 2 };

// Partition function of a window.  HJ hashes (equi-depth over distinct
// keys already; a learned model only contributes its heavy keys, which
// grouped joins peel); SMJ uses MSD radix on the key bits, or with a
// learned model equi-depth CDF ranges.
inline int part_kind(int alg, bool learned) {
    if (alg != SJB_ALG_SMJ) return PART_HASH;
    return learned ? PART_LEARNED : PART_RANGE;
}
enum Payload : int { PAY_VAL = 0, PAY_IDX = 1 };

// Learned CDF model (two-stage RMI).  A key maps to x = (key - kmin) *
// scale in [0, 1) (keys below kmin to 0); the root predicts a leaf from x
// (linear, or the trained 1-32-32-1 MLP sampled at kRootTable+1 points and
// interpolated), the leaf's linear model gives cdf(x), and the partition is
// floor(F * cdf).  Keys listed in heavy[] are aggregated apart in grouped
// mode (peeled) instead of partitioned.
constexpr int kRmiLeaves = 1024;
constexpr int kRootTable = 1024;
constexpr int kMlpH = 32;
constexpr int kMlpParams = kMlpH + kMlpH + kMlpH * kMlpH + kMlpH + kMlpH + 1;  // 1153
constexpr int kMaxHeavy = 32;  // heavy-key slots
enum RootKind : uint32_t { ROOT_LINEAR = 0, ROOT_TABLE = 1 };
struct RmiParams {
    uint32_t kmin;
    float scale;
    uint32_t root_kind;
    uint32_t n_heavy;
    float root_a, root_b;           // linear root: leaf = (root_a * x + root_b) * kRmiLeaves
    float table[kRootTable + 1];    // sampled MLP root at x = i / kRootTable
    float leaf_a[kRmiLeaves];       // cdf = leaf_a * x + leaf_b
    float leaf_b[kRmiLeaves];
    uint32_t heavy[kMaxHeavy];
    // heavy-key lookup structure, built on the host (heavy_table_build) and
    // copied into LDS by every partition workgroup: 64-slot open addressing
    // on the top 6 bits of key * kHashMul (slot -> key, heavy index, -1 =
    // empty) and a 4096-bit filter on the top 12 bits
    uint32_t hslot_key[64];
    int32_t hslot_idx[64];
    uint32_t hfilter[128];
};

// One window of a batched launch (device descriptor, built on the host).
// Records of the window's left side are partitioned into
// d_part[lpart, lpart + nL), right side into d_part[rpart, rpart + nR).
struct BatchWin {
    uint64_t lo, ro;         // first tuple of the window in the left / right stream
    uint64_t offs_base;      // first word of the window's chunk-offsets table
    uint32_t lpart, rpart;   // first partitioned record of each side
    uint32_t nL, nR;
    uint32_t ncL, ncR;       // partition-pass chunks per side
    uint32_t chunk_base;     // first partition-pass item of the window
    uint32_t item_base;      // first join item (= partition) of the window
    uint32_t id;             // caller's window index (result slot)
    uint32_t logF;
};

// Arguments of the batched partition and join kernels (by value).
constexpr int kXcds = 8;  // MI355X: 8 XCDs, workgroups dealt round-robin
struct BatchArgs {
    const sjb_tuple_t *L, *R;
    const BatchWin *win;
    const uint32_t *chunk_win;   // partition-pass item -> window
    const uint32_t *item_win;    // join item -> window
    Rec *part;
    uint32_t *offs;
    sjb_window_result_t *res;    // indexed by BatchWin::id
    uint32_t *qhead;             // per-XCD queue heads, 64 B apart
    uint32_t qbeg[kXcds + 1];    // join items of XCD x: [qbeg[x], qbeg[x+1])
    uint32_t chunk;              // partition-pass chunk (tuples)
    uint32_t grab;               // join items taken per queue access
    uint32_t expand;             // non-grouped: long match lists on full waves
    uint32_t idcap;              // run-id table entries of a PartView
    const RmiParams *rmi;
    uint64_t *heavy;             // learned + grouped: 4 u64 per (window, heavy key)
};

// Launch configuration of the batched path, chosen by sjb_batch_autotune.
struct BatchTune {
    int fanout_bias = 0;       // added to the per-window log2 fanout
    int chunk_log2 = 14;       // partition-pass chunk: 2^12 or 2^14 tuples
    int blocks_per_cu = 2;     // resident join workgroups per CU
    int grab = 1;              // join items per queue access
    int expand = 0;            // non-grouped: long match lists on full waves
};

struct WindowPlan {
    int alg;
    int logF;
    int kind;      // PartKind
    int payload;   // Payload
    uint32_t flags;
    uint64_t nL, nR;
    int chunk;     // tuples per partition chunk: kChunk, kSmallChunk or kTinyChunk
    int range_blocks;  // range kind: (min, max) pairs of the key sample in ctx d_scratch
    int ncL, ncR;  // chunks per side
    bool peel;     // learned + grouped: heavy keys aggregated apart
    // where the partition pass runs and writes (the ctx's own buffers and
    // stream, or the second set of an overlapped window)
    hipStream_t ps;
    Rec *part;
    uint32_t *offs;
    uint64_t *heavy;
    // non-null: the partition pass zeroes this result and every join
    // workgroup adds its (count, C1, C2) to it with 64-bit atomics, so no
    // reduction kernel runs (count paths without heavy-key peeling)
    sjb_window_result_t *direct;
};

}  // namespace sjb

struct sjb_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    sjb_options_t opts{};
    int n_cu = 0;
    bool small_chunks = true;         // small windows in kSmallChunk chunks (offs_cap sized for them)
    uint64_t bytes = 0;
    uint64_t ext_bytes = 0;           // device bytes of live batch plans and sliding windows
    std::vector<void *> dev_allocs;
    std::vector<void *> host_allocs;

    // lazy window buffers
    uint64_t cap = 0;                 // per-side tuple capacity
    sjb_tuple_t *d_stage[2] = {};     // H2D gather target of sjb_lazy_execute
    sjb::Rec *d_part = nullptr;       // partitioned records, left then right
    uint32_t *d_offs = nullptr;       // (F+1) local partition offsets per chunk
    uint64_t offs_cap = 0;
    uint32_t *d_scratch = nullptr;    // small per-call scratch (min/max partials)
    uint32_t *d_nlj_keys = nullptr;   // NLJ: compacted left keys (grow-only)
    uint64_t nlj_keys_cap = 0;
    uint64_t *d_partials = nullptr;   // 4 u64 per partition
    sjb_window_result_t *d_result = nullptr;
    sjb_window_result_t *h_result = nullptr;  // pinned
    sjb::RmiParams *d_rmi = nullptr;
    bool rmi_valid = false;
    uint32_t rmi_nheavy = 0;          // host copy of d_rmi->n_heavy
    std::vector<uint8_t> rmi_host;    // host copy of *d_rmi (a RmiParams) after a fit
    // learned-model fit buffers (rmi.hip), allocated on the first fit
    uint32_t *d_fit_keys = nullptr;   // sorted key sample
    float *d_fit_x = nullptr, *d_fit_t = nullptr;
    float *d_mlp = nullptr;           // params, grads, Adam m, Adam v (4 x kMlpParams)
    double *d_fit_acc = nullptr;      // loss accumulator + leaf statistics
    float *d_loss_hist = nullptr;
    uint32_t loss_hist_cap = 0;
    uint64_t *d_cursor = nullptr;     // pair-output cursor
    uint64_t *d_heavy = nullptr;      // heavy-key aggregates
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
    hipEvent_t ev[4] = {};             // per-phase timing of one call (partition start / end, join start / end)
    std::vector<hipEvent_t> tring;    // timing ring: 4 events per async window
    std::vector<uint8_t> tring_split; // per slot: join start is its own event (ev[2])
    uint64_t tring_next = 0;
    // overlapped async windows (SJB_F_OVERLAP): the partition pass of a
    // window runs on ov_stream into buffer set (window index & 1) while the
    // previous window's join runs on the ctx stream; set 0 is the ctx's own
    // d_part / d_offs / d_heavy
    hipStream_t ov_stream = nullptr;
    sjb::Rec *ov_part[2] = {};
    uint32_t *ov_offs[2] = {};
    uint64_t *ov_heavy[2] = {};
    hipEvent_t ov_part_done[2] = {}, ov_join_done[2] = {};
    uint64_t ov_next = 0;

    // device staging of pair output when the caller's buffer is pageable
    sjb_result_tuple_t *d_out_stage = nullptr;
    uint64_t out_stage_cap = 0;

    // batched windows
    sjb_window_result_t *d_batch_res = nullptr;
    uint64_t batch_res_cap = 0;
    sjb::BatchTune tune;
    struct sjb_batch_plan *scratch_plan = nullptr;  // reused by sjb_lazy_join_batch

    // NLJ bitmap path
    uint64_t *d_masks = nullptr;      // (left idx, word idx, mask) triples
    uint64_t mask_cap = 0;

    // eager state
    bool eager_ready = false;
    uint64_t eager_cap = 0;           // tuples per side
    uint64_t eager_nb = 0;            // buckets per side (pow2)
    sjb_tuple_t *d_eager_tuples[2] = {};
    uint32_t *d_eager_next[2] = {};
    uint32_t *d_eager_head[2] = {};
    uint64_t eager_n[2] = {0, 0};
    sjb_tuple_t *d_batch = nullptr;   // staged incoming batch
    uint64_t batch_cap = 0;

    // persistent-kernel eager executor (pk.hip)
    hipStream_t pk_stream = nullptr;   // runs the persistent kernel
    hipStream_t pk_copy = nullptr;     // control-word resets before a relaunch
    void *d_pk_ctl = nullptr;          // forwarded posts, completion, claims (device)
    void *h_pk_host = nullptr;         // doorbell + descriptor ring (coherent pinned host)
    sjb_tuple_t *h_pk_slots = nullptr; // pinned copies of pageable batches (ring)
    void *h_pk_res = nullptr;          // results + completion (coherent pinned host)
    uint64_t *h_pk_word = nullptr;     // pinned source of control-word resets
    uint64_t pk_seq = 0;               // next sequence number
    uint64_t pk_slot_cap = 0;          // tuples per batch slot
    uint64_t pk_n[2] = {0, 0};         // host view of the state sizes
    uint64_t pk_idle_limit = 0;        // polls before an idle kernel exits
    uint64_t pk_launches = 0;
    bool pk_ready = false;
    bool pk_running = false;
};

namespace sjb {

// allocation helpers (tracked; freed by sjb_close)
int dev_alloc(sjb_ctx *c, void **p, size_t bytes);
int host_alloc(sjb_ctx *c, void **p, size_t bytes);
void dev_free(sjb_ctx *c, void **p, size_t bytes);
int dev_regrow(sjb_ctx *c, void **p, size_t old_bytes, size_t bytes);

// ---- launchers (partition.hip) ----
int choose_logF(int alg, uint32_t flags, uint64_t n_build, int requested, bool small);
int launch_minmax(sjb_ctx *c, hipStream_t s, const sjb_tuple_t *L, uint64_t nL, const sjb_tuple_t *R,
                  uint64_t nR, int *blocks_out);
int launch_partition(sjb_ctx *c, const WindowPlan &p, const sjb_tuple_t *L, const sjb_tuple_t *R);
int partition_setup(sjb_ctx *c);
int launch_fill_u32(hipStream_t s, uint32_t *p, uint32_t v, uint64_t n);
int launch_fill_u64(hipStream_t s, uint64_t *p, uint64_t v, uint64_t n);

// ---- launchers (join.hip) ----
int launch_hj_join(sjb_ctx *c, const WindowPlan &p, const sjb_tuple_t *L, const sjb_tuple_t *R,
                   sjb_result_tuple_t *out, uint64_t out_cap, int wrap);
int launch_smj_join(sjb_ctx *c, const WindowPlan &p, const sjb_tuple_t *L, const sjb_tuple_t *R,
                    sjb_result_tuple_t *out, uint64_t out_cap, int wrap);
int launch_reduce(sjb_ctx *c, int n_partials, sjb_window_result_t *d_out, uint32_t n_heavy,
                  const uint64_t *heavy = nullptr);

// ---- launchers (nlj.hip) ----
int launch_nlj(sjb_ctx *c, const sjb_tuple_t *L, uint64_t nL, const sjb_tuple_t *R, uint64_t nR,
               uint32_t flags, int write, sjb_result_tuple_t *out, uint64_t out_cap, int wrap,
               sjb_window_result_t *d_out);
int launch_nlj_bitmap(sjb_ctx *c, const sjb_tuple_t *L, uint64_t nL, const sjb_tuple_t *R,
                      uint64_t nR, uint32_t flags, sjb_result_tuple_t *out, uint64_t out_cap,
                      sjb_window_result_t *d_out, uint64_t *n_masks);

int launch_part_sizes(hipStream_t s, const uint32_t *offs, int logF, int ncL, int ncR, uint32_t *sizes);

// ---- api.hip ----
int window_partition(sjb_ctx *c, int alg, uint32_t flags, const sjb_tuple_t *L, uint64_t nL,
                     const sjb_tuple_t *R, uint64_t nR, int write, WindowPlan *p, int buf = -1,
                     sjb_window_result_t *direct = nullptr);  // direct: see WindowPlan::direct
int window_enqueue(sjb_ctx *c, int alg, uint32_t flags, const sjb_tuple_t *L, uint64_t nL,
                   const sjb_tuple_t *R, uint64_t nR, int write, sjb_result_tuple_t *out,
                   uint64_t out_cap, int wrap, sjb_window_result_t *d_res, int *logF_used,
                   const hipEvent_t *ev, int buf = -1);  // ev: 4 events (partition start / end, join
                                                         // start / end) or null; buf: overlapped buffer set

// ---- batch.hip ----
int batch_join(sjb_ctx *c, int alg, uint32_t flags, const sjb_tuple_t *L, const uint64_t *loff,
               const sjb_tuple_t *R, const uint64_t *roff, uint32_t n_windows,
               sjb_window_result_t *results);
void batch_plan_free(struct sjb_batch_plan *p);
int batch_log_fanout(int alg, uint32_t flags, uint64_t n_build, int bias);
// batched kernels (partition.hip / join.hip)
int launch_partition_batch(hipStream_t s, int kind, bool peel, const BatchArgs &a, uint32_t n_chunks,
                           size_t dyn_bytes);
int launch_heavy_finalize(hipStream_t s, sjb_window_result_t *res, uint32_t n, const uint64_t *heavy,
                          uint32_t n_heavy);
int launch_join_batch(hipStream_t s, int alg, uint32_t flags, const BatchArgs &a, int blocks,
                      size_t dyn_bytes);
int launch_batch_clear(hipStream_t s, sjb_window_result_t *res, uint32_t n, uint32_t *qhead,
                       uint64_t *heavy);
size_t partition_lds_bytes_chunk(int logF, int chunk);
size_t join_view_bytes(int alg, bool grouped, int ncL, int ncR);
// verdicts of rmi_judge
constexpr int kDriftNone = 0, kDriftRefit = 1, kDriftRanges = 2;
int rmi_judge(const sjb_ctx *c, const std::vector<uint32_t> &keys, bool peel, bool allow_refit,
              RmiParams *out);
int rmi_sample_keys(sjb_ctx *c, const sjb_tuple_t *T, uint64_t n, std::vector<uint32_t> &keys);
int sampled_heavy(sjb_ctx *c, const sjb_tuple_t *T, uint64_t n, RmiParams *out, uint32_t *n_heavy);
int join_idcap(int alg, bool grouped);

// ---- pointer checks / output staging (api.hip) ----
// True when the device can dereference p (device memory, or pinned /
// registered host memory).
bool device_accessible(const void *p);
struct OutBuf {
    sjb_result_tuple_t *dev = nullptr;   // what kernels write
    sjb_result_tuple_t *host = nullptr;  // caller's pageable buffer (copy back) or null
    uint64_t cap = 0;
};
int out_prepare(sjb_ctx *c, sjb_result_tuple_t *out, uint64_t cap, OutBuf *ob);
int out_finish(sjb_ctx *c, const OutBuf &ob, uint64_t rows);

// ---- eager.hip ----
int eager_init(sjb_ctx *c);

}  // namespace sjb
