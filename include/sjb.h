/*
 * sjb.h — C ABI of the MI355X (gfx950) windowed stream-join operators.
 *
 * This is the drop-in boundary for the reference's join-operator layer
 * (TU-Berlin-DIMA/gpu-stream-join-benchmark).  Every entry point takes plain
 * pointers and sizes, returns an int status (0 = ok, negative = SJB_E_*),
 * never exits the process and never throws across the ABI.  The C++ adapters
 * in gpu-stream-join-benchmark_amd/host/ wrap these calls in the reference's
 * class interfaces (BaseLazyExecutor / BaseEagerExecutor / Sink / factory).
 *
 * Sides: "left" = R = build side, "right" = S = probe side, as in the
 * reference (src/sjb/windowing/lazy/executor/GPU/GPU_HJ.cu:122-128).
 *
 * Checksums (the reference defines none; see sjb_mix64 below):
 *   A(k,v)  = mix64(((u64)k << 32 | v) ^ SJB_SEED_A)
 *   B(k,v)  = mix64(((u64)k << 32 | v) ^ SJB_SEED_B)
 *   C1      = sum over matched pairs (r,s) of A(k, r.val) * B(k, s.val)  mod 2^64
 *   H2      = mix64(A(k, r.val) + SJB_GOLDEN * ((u64)s.val << 1 | 1))
 *   C2      = sum over matched pairs of H2                              mod 2^64
 * C1 factorises per key group, so it is computable in O(n) at any skew;
 * C2 needs every pair and is only produced when SJB_F_C2 is requested.
 */
#ifndef SJB_H
#define SJB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SJB_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
#define SJB_OK 0
#define SJB_E_INVALID (-1)     /* bad argument / unknown enum value            */
#define SJB_E_CAPACITY (-2)    /* output or window larger than the buffer/ctx  */
#define SJB_E_NOMEM (-3)       /* device or host allocation failed             */
#define SJB_E_HIP (-4)         /* a HIP runtime call failed                    */
#define SJB_E_UNSUPPORTED (-5) /* combination not offered (mirrors the factory) */
#define SJB_E_NODEVICE (-6)    /* no usable gfx950 device                      */
#define SJB_E_BUDGET (-7)      /* HBM budget would be exceeded                 */
#define SJB_E_TIMEOUT (-8)     /* a bounded wait (persistent kernel) expired   */
#define SJB_E_STATE (-9)       /* call not valid in the current ctx state      */
#define SJB_E_MISMATCH (-10)   /* autotuner: two launch shapes disagreed       */

/* ---- enums: same values as include/sjb/windowing/BaseWindowing.cuh:9-40 - */
/* ALGORITHM_TYPE */
#define SJB_ALG_EMPTY 0
#define SJB_ALG_NLJ 1
#define SJB_ALG_HJ 2
#define SJB_ALG_SMJ 3
/* RESULT_WRITING_METHOD */
#define SJB_OUT_NOOUTPUT 0
#define SJB_OUT_COUNTKERNEL 1
#define SJB_OUT_ESTIMATEDSELECTIVITY 2
#define SJB_OUT_BITMAP 3
#define SJB_OUT_ATOMIC 4
/* PROGRESSIVENESS_TYPE */
#define SJB_EAGER 0
#define SJB_LAZY 1
/* DEVICE_TYPE */
#define SJB_DEV_CPU 0
#define SJB_DEV_GPU 1

/* ---- per-call flags ---------------------------------------------------- */
#define SJB_F_C2 0x1u          /* also compute C2 (enumerates every pair)      */
#define SJB_F_GROUPED 0x2u     /* key-group aggregation (count + C1 only);     */
                               /* O(n) at any skew, C2 not produced            */
#define SJB_F_LEARNED 0x4u     /* partition with the ctx's learned CDF model   */
#define SJB_F_TIMING 0x8u      /* record per-phase device time (hipEvents)     */
#define SJB_F_OVERLAP 0x10u    /* sjb_lazy_join_async: the window's partition  */
                               /* pass may run on a second stream, overlapping */
                               /* the previous window's join; the inputs must  */
                               /* be ready when the call is made               */
#define SJB_F_NO_REFIT 0x20u   /* batched + SJB_F_LEARNED: a model found stale */
                               /* on the batch's keys is not refitted to them; */
                               /* the batch uses radix ranges instead          */
#define SJB_F_PEEL 0x40u       /* batched + SJB_F_GROUPED, no SJB_F_LEARNED:   */
                               /* the most frequent keys of a sample of the    */
                               /* batch's left stream (>= 3 of 16384 samples,  */
                               /* at most 32) are aggregated apart; the rest is */
                               /* hash (HJ) / radix-range (SMJ) partitioned    */
                               /* as without it — the model-free baseline of   */
                               /* the learned partitioner                      */

/* ---- data layout ------------------------------------------------------- */
/* Identical to Tuple (include/sjb/utils/Tuple.hpp:6-11): 16 B, AoS.        */
typedef struct sjb_tuple {
    uint32_t key;
    uint32_t val;
    uint64_t ts;
} sjb_tuple_t;

/* Identical to ResultTuple (include/sjb/utils/ResultTuple.hpp:5-13): 32 B. */
typedef struct sjb_result_tuple {
    uint32_t lkey;
    uint32_t rkey;
    uint32_t lval;
    uint32_t rval;
    uint64_t lts;
    uint64_t rts;
} sjb_result_tuple_t;

/* Per-window result: ExecutionStatistic (include/sjb/utils/ExecutionStatistic.hpp:6-20)
 * plus the checksums and device timings this build adds.                    */
typedef struct sjb_stats {
    uint64_t result_count;  /* ExecutionStatistic::resultCount                */
    uint64_t build_time_us; /* ExecutionStatistic::buildTime  (host clock)    */
    uint64_t probe_time_us; /* ExecutionStatistic::probeTime                  */
    uint64_t sort_time_us;  /* ExecutionStatistic::sortTime                   */
    uint64_t merge_time_us; /* ExecutionStatistic::mergeTime                  */
    uint64_t checksum1;     /* C1                                              */
    uint64_t checksum2;     /* C2 (valid when c2_valid)                        */
    uint64_t pairs_written; /* result tuples stored into the caller's buffer  */
    uint64_t device_ns;     /* device time of the call (SJB_F_TIMING)         */
    uint32_t c2_valid;
    uint32_t fanout_log2;   /* partition fanout actually used                 */
} sjb_stats_t;

/* Compact per-window result used by the batched and async paths.           */
typedef struct sjb_window_result {
    uint64_t count;
    uint64_t checksum1;
    uint64_t checksum2;
    uint64_t pairs; /* pairs written (materialising methods) */
} sjb_window_result_t;

/* Context options (the reference's executor constructor arguments:
 * numDistinctKeys, batchSize, maxTupleInWindow — e.g.
 * src/sjb/windowing/lazy/executor/GPU/GPU_HJ.cu:266-287).                  */
typedef struct sjb_options {
    uint64_t max_window_tuples; /* per-side capacity of one lazy window        */
    uint64_t num_distinct_keys; /* D: bucket count of the reference CPU/eager  */
                                /* joins and the ES size estimate (1/D)       */
    uint64_t batch_size;        /* B: tuples per batch / ring slot             */
    uint64_t hbm_budget_bytes;  /* 0 = no budget                               */
    uint32_t fanout_log2;       /* 0 = automatic                               */
    uint32_t flags;             /* default SJB_F_* applied to every call       */
} sjb_options_t;

typedef struct sjb_ctx sjb_ctx_t;

/* ---- library / context ------------------------------------------------- */
int sjb_abi_version(void);
const char *sjb_strerror(int code);

/* Opens a context on `device`: one HIP stream plus every device buffer a
 * window of max_window_tuples per side needs (no per-window hipMalloc).
 * Replaces the executor constructors, e.g. GPU_HJ::GPU_HJ
 * (src/sjb/windowing/lazy/executor/GPU/GPU_HJ.cu:266-287).                  */
int sjb_open(int device, const sjb_options_t *opts, sjb_ctx_t **out);
int sjb_close(sjb_ctx_t *ctx);
/* The ctx's HIP stream (hipStream_t) so callers can order or time work.   */
void *sjb_stream(sjb_ctx_t *ctx);
int sjb_sync(sjb_ctx_t *ctx);
/* Device bytes currently held by the ctx.                                  */
uint64_t sjb_device_bytes(sjb_ctx_t *ctx);

/* Page-locked host memory the device copies from (and writes to)
 * asynchronously; SJB_E_NODEVICE without a device.  Replaces the
 * cudaMallocHost calls of the reference's source streams, ring buffers and
 * sink (src/experiment/Runner.cpp:23,29,
 * include/sjb/windowing/lazy/CountBased.cuh:35-36, src/sjb/sink/Sink.cpp:8-9). */
int sjb_host_alloc(uint64_t bytes, void **p);
int sjb_host_free(void *p);

/* ---- lazy (tumbling) window join --------------------------------------- */
/* One window, inputs already resident in device memory.  Replaces
 * BaseLazyExecutor::execute (include/sjb/windowing/lazy/executor/BaseLazyExecutor.hpp:26-30)
 * for the GPU executors GPU_{HJ,SMJ,NLJ}_{NoOutput,CountKernel,
 * EstimatedSelectivity,Atomic,Bitmap} (src/sjb/windowing/lazy/executor/GPU/).
 * `out` (host or device pointer, capacity `out_cap` result tuples) is used
 * by the materialising methods; if the join has more matches than out_cap,
 * CountKernel/EstimatedSelectivity return SJB_E_CAPACITY with
 * stats->result_count = required size, and Atomic wraps (pos % out_cap) as
 * the reference sink ring does (src/sjb/windowing/lazy/executor/GPU/GPU_HJ.cu:92).
 * Synchronous: returns when stats are host-visible.                         */
int sjb_lazy_join(sjb_ctx_t *ctx, int algorithm, int method, uint32_t flags,
                  const sjb_tuple_t *d_left, uint64_t n_left,
                  const sjb_tuple_t *d_right, uint64_t n_right,
                  sjb_result_tuple_t *out, uint64_t out_cap, sjb_stats_t *stats);

/* Same, but gathers the window from host ring-buffer slots: `left_idx` holds
 * n_left_idx tuple offsets of batch_size-tuple slots in left_ring — exactly
 * the arguments of BaseLazyExecutor::execute.  The H2D copy is part of the
 * call, as in the reference (src/sjb/windowing/lazy/executor/GPU/GPU_HJ_NoOutput.cu:37-41). */
int sjb_lazy_execute(sjb_ctx_t *ctx, int algorithm, int method, uint32_t flags,
                     const sjb_tuple_t *left_ring, const uint64_t *left_idx, uint64_t n_left_idx,
                     const sjb_tuple_t *right_ring, const uint64_t *right_idx, uint64_t n_right_idx,
                     uint64_t batch_size, sjb_result_tuple_t *out, uint64_t out_cap,
                     sjb_stats_t *stats);

/* Asynchronous count/checksum window (NoOutput): enqueues on the ctx stream
 * and writes the result into device memory `d_result` (one
 * sjb_window_result_t).  Used for back-to-back windows with no host sync. */
int sjb_lazy_join_async(sjb_ctx_t *ctx, int algorithm, uint32_t flags,
                        const sjb_tuple_t *d_left, uint64_t n_left,
                        const sjb_tuple_t *d_right, uint64_t n_right,
                        sjb_window_result_t *d_result);

/* Per-window device timing of sjb_lazy_join_async calls made with
 * SJB_F_TIMING: sjb_timing_ring allocates a ring of n windows x 4 HIP events
 * (partition start / end, join start / end, recorded on the stream of each
 * phase around the partition kernels and the join + reduce kernels);
 * sjb_timing_read returns the last min(n, recorded) windows' partition and
 * join milliseconds, oldest first (synchronises the ctx).  With
 * SJB_F_OVERLAP the phases of consecutive windows may overlap in time.     */
int sjb_timing_ring(sjb_ctx_t *ctx, uint32_t n);
int sjb_timing_read(sjb_ctx_t *ctx, float *part_ms, float *join_ms, uint32_t n, uint32_t *n_read);

/* Many uneven windows in one launch sequence (count/checksum only).  Window
 * w is d_left[left_off[w], left_off[w+1]) x d_right[right_off[w], right_off[w+1]);
 * offsets are host arrays of n_windows+1 entries.  Results (host array) are
 * identical to n_windows calls of sjb_lazy_join.                          */
int sjb_lazy_join_batch(sjb_ctx_t *ctx, int algorithm, uint32_t flags,
                        const sjb_tuple_t *d_left, const uint64_t *left_off,
                        const sjb_tuple_t *d_right, const uint64_t *right_off,
                        uint32_t n_windows, sjb_window_result_t *results);

/* ---- batched windows: plans, hipGraph replay, autotuning --------------- */
/* A plan fixes a batch layout (the windows of sjb_lazy_join_batch over fixed
 * device streams d_left / d_right), owns the partition buffers, uploads the
 * window descriptors once and captures the batch's launches (clear,
 * partition, persistent join) in a hipGraph.  sjb_batch_plan_launch replays
 * the graph on the ctx stream (asynchronous); per-window results land in
 * the caller's device array d_results[n_windows] (HJ / SMJ; count, C1 and,
 * with SJB_F_C2, C2).  This replaces the reference's one-executor-call-per-
 * window loop (include/sjb/windowing/lazy/CountBased.cuh:103-166) for a
 * known window schedule; it has no single reference counterpart.           */
typedef struct sjb_batch_plan sjb_batch_plan_t;

/* Launch shape of the batched path (the autotuner's search space).        */
typedef struct sjb_batch_tune {
    int32_t fanout_bias;    /* added to every window's log2 partition fanout */
    uint32_t chunk_log2;    /* partition-pass chunk: 12 or 14                */
    uint32_t blocks_per_cu; /* resident persistent join workgroups per CU    */
    uint32_t grab;          /* join work items taken per queue access        */
    uint32_t expand;        /* non-grouped joins, 0 or 1: 1 = long match     */
                            /* lists run on full waves (HJ: every match via  */
                            /* the wave's queue; SMJ with C2: cooperative    */
                            /* expansion) — pays under skewed keys           */
    uint32_t reserved;      /* 0                                             */
    double ms;              /* autotune: median batch time of this shape     */
} sjb_batch_tune_t;

typedef struct sjb_batch_info {
    uint32_t n_windows;
    uint32_t n_chunks;      /* partition-pass work items                     */
    uint32_t n_items;       /* join work items (window partitions)           */
    uint32_t blocks;        /* persistent join workgroups                    */
    uint64_t device_bytes;  /* device memory held by the plan                */
    sjb_batch_tune_t tune;
    uint32_t part_kind;     /* partitioning used: 0 hash, 1 radix ranges,    */
                            /* 2 learned CDF                                 */
    uint32_t drift_fallback;/* 1: SJB_F_LEARNED asked, but the model was     */
                            /* stale on a sample of this batch's keys and no */
                            /* refit was made or helped: more imbalanced     */
                            /* than radix ranges, so ranges are used         */
    uint32_t drift_refit;   /* 1: the model was stale on that sample, so the */
                            /* plan refitted a private copy of it to the     */
                            /* sample (root = the sample's CDF as a table,   */
                            /* leaves = chords of it; heavy keys kept) and   */
                            /* partitions by it                              */
    uint32_t n_heavy;       /* keys aggregated apart (learned or SJB_F_PEEL) */
} sjb_batch_info_t;

int sjb_batch_plan_create(sjb_ctx_t *ctx, int algorithm, uint32_t flags,
                          const sjb_tuple_t *d_left, const uint64_t *left_off,
                          const sjb_tuple_t *d_right, const uint64_t *right_off,
                          uint32_t n_windows, sjb_window_result_t *d_results,
                          sjb_batch_plan_t **plan);
int sjb_batch_plan_launch(sjb_batch_plan_t *plan);
int sjb_batch_plan_info(sjb_batch_plan_t *plan, sjb_batch_info_t *info);
int sjb_batch_plan_destroy(sjb_batch_plan_t *plan);

/* Tries every launch shape of the search space on the given batch, checks
 * that all of them give bit-identical per-window results (SJB_E_MISMATCH
 * otherwise), stores the fastest in the ctx (used by later batched calls
 * and plans) and returns it in *best.                                       */
int sjb_batch_autotune(sjb_ctx_t *ctx, int algorithm, uint32_t flags,
                       const sjb_tuple_t *d_left, const uint64_t *left_off,
                       const sjb_tuple_t *d_right, const uint64_t *right_off,
                       uint32_t n_windows, uint32_t reps, sjb_batch_tune_t *best,
                       uint32_t *n_tried);
int sjb_batch_get_tune(sjb_ctx_t *ctx, sjb_batch_tune_t *tune);
int sjb_batch_set_tune(sjb_ctx_t *ctx, const sjb_batch_tune_t *tune);

/* ---- learned partitioner ------------------------------------------------ */
/* A two-stage RMI over the key CDF (root: linear or a 1-32-32-1 ReLU MLP
 * trained on the device by Adam; 1024 linear leaves), fitted on a sample of
 * earlier windows' tuples (d_train, device memory).  SMJ joins called with
 * SJB_F_LEARNED partition by floor(F * cdf(key)) (equi-depth key ranges)
 * instead of MSD radix; HJ keeps hash partitions, which are equi-depth over
 * distinct keys already.  In grouped mode (both algorithms) the model's
 * heavy keys (up to 32 most frequent sample keys occurring at least
 * heavy_min times) are aggregated apart instead of partitioned.  The fit compares the learned partition of the
 * sample with MSD radix on the key bits; when the learned one is more than
 * 1.25x as imbalanced the model is not used (report->fallback = 1) and
 * SJB_F_LEARNED joins take the plain radix / hash path.  Batch plans judge
 * the model again on a sample of their own keys and, when it has gone stale
 * (a drift since the fit), partition by a copy refitted to that sample, or
 * by radix ranges (sjb_batch_info_t.drift_refit / drift_fallback;
 * SJB_F_NO_REFIT).  Results never
 * depend on the model: only partition balance does.  No reference
 * counterpart (the reference partitions by key % D only).                 */
#define SJB_RMI_LINEAR 0
#define SJB_RMI_MLP 1
typedef struct sjb_rmi_config {
    uint32_t root;       /* SJB_RMI_LINEAR or SJB_RMI_MLP                   */
    uint32_t steps;      /* MLP: full-batch Adam steps                      */
    float lr;            /* MLP: peak learning rate (cosine decay to 10%)   */
    uint32_t seed;       /* MLP: initial weights                            */
    uint32_t heavy_min;  /* min sample count of a heavy key (0 = none)      */
    uint32_t pad;
} sjb_rmi_config_t;
typedef struct sjb_rmi_report {
    double loss_init;         /* MLP: mean squared CDF error before training */
    double loss_final;        /* ... and after the last step                 */
    double fit_ms;            /* device time of the whole fit                */
    double imbalance_learned; /* max/mean partition size of the sample      */
    double imbalance_radix;   /* same for MSD radix on the key bits          */
    uint32_t n_heavy, leaves_used, n_sample, fallback;
} sjb_rmi_report_t;
int sjb_rmi_fit(sjb_ctx_t *ctx, const sjb_tuple_t *d_train, uint64_t n_train,
                const sjb_rmi_config_t *cfg, sjb_rmi_report_t *report);
/* Forgets the model (SJB_F_LEARNED then takes the radix / hash path).      */
int sjb_rmi_clear(sjb_ctx_t *ctx);
/* One forward/backward pass of the MLP root on host arrays: params
 * (1153 floats: W1[32], b1[32], W2[32][32] row = output unit, b2[32],
 * w3[32], b3), samples x[n], targets t[n] (n <= 65536); writes the gradient
 * of mean((y - t)^2) and the loss.  Exposed for gradient checks.          */
int sjb_mlp_grad(sjb_ctx_t *ctx, const float *params, const float *x, const float *t, uint32_t n,
                 float *grad, double *loss);

/* Partition statistics of one window (the learned path's judge): runs the
 * partition pass of sjb_lazy_join(algorithm, flags) and reports per-side
 * partition sizes.  fixup_frac = share of build tuples in partitions larger
 * than one LDS tile (they take the multi-tile path).                       */
typedef struct sjb_partition_report {
    uint32_t fanout_log2;
    uint32_t kind;            /* 0 hash, 1 range (radix), 2 learned          */
    uint64_t max_build, max_probe;
    double mean_build;
    double imbalance;         /* max_build / mean_build                      */
    double fixup_frac;
    uint64_t heavy_left, heavy_right; /* tuples aggregated as heavy keys     */
} sjb_partition_report_t;
int sjb_partition_report(sjb_ctx_t *ctx, int algorithm, uint32_t flags,
                         const sjb_tuple_t *d_left, uint64_t n_left,
                         const sjb_tuple_t *d_right, uint64_t n_right,
                         sjb_partition_report_t *report);

/* ---- eager sliding-window join (device-resident state, serving loop) --- */
/* A sliding count window over two streams: after every event (one left and
 * one right batch, either may be empty) the window holds the last `window`
 * tuples of each side, and the event's result is count and C1 of that
 * window's join.  State stays on the device (per side: a ring of `window`
 * tuples and a group-by-key table of count and sum of A / B); an event costs
 * O(batch) lookups whatever the window size (expired and new tuples are
 * probed against the other side, then applied to their own).  Generalises
 * the reference's eager executors (BaseEagerExecutor::execute per batch,
 * include/sjb/windowing/eager/executor/BaseEagerExecutor.hpp:16-20; state
 * cleared per tumbling window by include/sjb/windowing/eager/CountBased.cuh:36-72)
 * to a sliding window.  Admission: sjb_slide_create returns SJB_E_BUDGET
 * when the state would exceed the ctx's hbm_budget_bytes together with the
 * ctx, its batch plans and its other sliding windows.                      */
typedef struct sjb_slide sjb_slide_t;
typedef struct sjb_slide_info {
    uint64_t window, max_batch, table_slots, device_bytes;
    uint64_t events, rebuilds, tuples_left, tuples_right;
} sjb_slide_info_t;
int sjb_slide_create(sjb_ctx_t *ctx, uint64_t window, uint64_t max_batch, sjb_slide_t **slide);
int sjb_slide_reset(sjb_slide_t *slide);
/* One event from host batches (at most max_batch tuples per side); returns
 * when the window's result is host-visible.                                 */
int sjb_slide_push(sjb_slide_t *slide, const sjb_tuple_t *left, uint64_t n_left,
                   const sjb_tuple_t *right, uint64_t n_right, sjb_window_result_t *result);
/* Serving loop: event i = h_left[left_off[i], left_off[i+1]) and
 * h_right[right_off[i], right_off[i+1]) arrives arrival_s[i] seconds after
 * the call starts and is enqueued once it has arrived; latency_s[i] = time
 * its result became host-visible minus its arrival.  Host batches should be
 * pinned (hipHostMalloc) for asynchronous copies.                           */
int sjb_slide_serve(sjb_slide_t *slide, const sjb_tuple_t *h_left, const uint64_t *left_off,
                    const sjb_tuple_t *h_right, const uint64_t *right_off, uint32_t n_events,
                    const double *arrival_s, double *latency_s, sjb_window_result_t *results,
                    double *elapsed_s);
int sjb_slide_info(sjb_slide_t *slide, sjb_slide_info_t *info);
int sjb_slide_destroy(sjb_slide_t *slide);

/* ---- eager (per-batch) join -------------------------------------------- */
/* Joins one incoming batch against the opposite side's resident window state
 * and then appends it to its own side's state.  Replaces
 * BaseEagerExecutor::execute/clearStates/onEndOfStream
 * (include/sjb/windowing/eager/executor/BaseEagerExecutor.hpp:16-20) for the
 * executors GPU_HJ, GPU_HJ_Atomic, GPU_NLJ, GPU_NLJ_BM
 * (src/sjb/windowing/eager/executor/GPU/).  `batch` is a host pointer.     */
int sjb_eager_execute(sjb_ctx_t *ctx, int algorithm, int method, uint32_t flags,
                      const sjb_tuple_t *batch, uint64_t n, int is_left,
                      sjb_result_tuple_t *out, uint64_t out_cap, sjb_stats_t *stats);
int sjb_eager_clear(sjb_ctx_t *ctx);
int sjb_eager_end_of_stream(sjb_ctx_t *ctx);

/* ---- persistent-kernel eager join -------------------------------------- */
/* The eager hash join of sjb_eager_execute (count/checksums) served by one
 * resident workgroup that polls a device-memory work queue, so a batch costs
 * one H2D copy and a doorbell write instead of kernel launches.  Replaces the
 * design of GPU_HJ_PK (src/sjb/windowing/eager/executor/GPU/GPU_HJ_PK.cu:10-249,
 * include/sjb/windowing/eager/executor/GPU/GPU_HJ_PK.cuh:19-40; not compiled
 * in the reference).  Batches hold at most opts.batch_size tuples.  The
 * kernel exits by itself when idle (environment SJB_PK_IDLE_POLLS polls) and
 * is relaunched on the next batch; sjb_pk_stop ends it, sjb_close stops it. */
int sjb_pk_execute(sjb_ctx_t *ctx, const sjb_tuple_t *batch, uint64_t n, int is_left, sjb_stats_t *stats);
int sjb_pk_clear(sjb_ctx_t *ctx);
int sjb_pk_stop(sjb_ctx_t *ctx);
/* Number of persistent-kernel launches so far (relaunches after idle exits). */
uint64_t sjb_pk_launches(sjb_ctx_t *ctx);

/* ---- allocation-free window join ---------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): one HJ or SMJ window like sjb_lazy_join_async
 * (count + C1, with SJB_F_C2 also C2, or SJB_F_GROUPED), without a context,
 * enqueued on the caller's `stream` (a hipStream_t of the current device;
 * null = that device's null stream); the result lands in device memory
 * d_result when the stream reaches it.  It calls no hipMalloc / hipFree and
 * clears its counters with kernels (no memset nodes).  Scratch memory is one
 * static device array of about 19 MB per GPU, held until the process ends,
 * so each side holds at most SJB_NOALLOC_MAX_TUPLES tuples (SJB_E_CAPACITY
 * beyond); other flags and NLJ return SJB_E_UNSUPPORTED.  Every call on a
 * GPU shares that GPU's array: calls that overlap in time on one GPU race,
 * so order them (one stream, or events).                                  */
#define SJB_NOALLOC_MAX_TUPLES (1u << 20)
int sjb_lazy_join_noalloc(int algorithm, uint32_t flags, const sjb_tuple_t *d_left, uint64_t n_left,
                          const sjb_tuple_t *d_right, uint64_t n_right, sjb_window_result_t *d_result,
                          void *stream);

/* ---- data generation (host, deterministic) ----------------------------- */
/* Fills n tuples: key distribution by `dist` (0 uniform over [0,key_range),
 * 1 Zipf(theta) over [0,key_range) with ranks scrambled by a fixed
 * bijection), val = independent uniform u32, ts = ts_base + index.        */
#define SJB_DIST_UNIFORM 0
#define SJB_DIST_ZIPF 1
int sjb_generate(sjb_tuple_t *out, uint64_t n, int dist, uint64_t key_range, double theta,
                 uint64_t seed, uint64_t ts_base);
/* Drifting mixture over u32 keys for window `window`: with probability
 * zipf_frac a Zipf(theta) key over [0, 2^32) (ranks scrambled as above),
 * otherwise a draw of one of 8 Gaussian components: weight of component j
 * proportional to 1 + (5j mod 4), mean (0.06 + 0.12 j) * 2^32 shifted by
 * 0.01 * 2^32 every 64 windows (mod 2^32), sigma (0.004 + 0.002 (j mod 4)) *
 * 2^32.  val and ts as sjb_generate.                                       */
int sjb_generate_mix(sjb_tuple_t *out, uint64_t n, uint64_t window, double zipf_frac, double theta,
                     uint64_t seed, uint64_t ts_base);

/* ---- checksum primitives (host reference of the device formulas) ------ */
uint64_t sjb_mix64(uint64_t z);
uint64_t sjb_hash_a(uint32_t key, uint32_t val);
uint64_t sjb_hash_b(uint32_t key, uint32_t val);
uint64_t sjb_pair_hash(uint32_t key, uint32_t rval, uint32_t sval);

#ifdef __cplusplus
}
#endif

#endif /* SJB_H */
