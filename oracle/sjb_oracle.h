/*
 * sjb_oracle.h — CPU oracle for the windowed stream-join hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or the timed CPU baseline) — never as the product path.
 *
 * It restates, in plain C, the reference's CPU join semantics:
 *   lazy CPU_HJ   src/sjb/windowing/lazy/executor/CPU/CPU_HJ.cpp:30-136
 *   lazy CPU_SMJ  src/sjb/windowing/lazy/executor/CPU/CPU_SMJ.cpp:23-139
 *   lazy CPU_NLJ  src/sjb/windowing/lazy/executor/CPU/CPU_NLJ.cpp:11-32
 *   eager CPU_HJ  src/sjb/windowing/eager/executor/CPU/CPU_HJ.cpp:35-82
 * and extends each with the checksums C1/C2 and the pair list (the
 * reference joins only count).  Defects listed in the reference's own code
 * (raw-key histogram indexing, hash-vs-key compare in CPU_HJ.cpp:113,
 * key copied into val in CPU_SMJ.cpp:48) are not reproduced.
 */
#ifndef SJB_ORACLE_H
#define SJB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_tuple {
    uint32_t key;
    uint32_t val;
    uint64_t ts;
} or_tuple;

typedef struct or_pair { /* layout of ResultTuple */
    uint32_t lkey, rkey, lval, rval;
    uint64_t lts, rts;
} or_pair;

typedef struct or_result {
    uint64_t count;
    uint64_t c1;
    uint64_t c2;
    uint64_t pairs; /* pairs stored (<= cap) */
} or_result;

uint64_t or_mix64(uint64_t z);
uint64_t or_hash_a(uint32_t k, uint32_t v);
uint64_t or_hash_b(uint32_t k, uint32_t v);
uint64_t or_pair_hash(uint32_t k, uint32_t rv, uint32_t sv);

/* Lazy joins of one window.  want_c2: enumerate pairs for C2.  pairs may be
 * NULL (cap 0).  Return 0, or -1 on allocation failure.                    */
int or_lazy_hj(const or_tuple *R, uint64_t nR, const or_tuple *S, uint64_t nS, uint64_t D,
               int want_c2, or_result *res, or_pair *pairs, uint64_t cap);
int or_lazy_smj(const or_tuple *R, uint64_t nR, const or_tuple *S, uint64_t nS,
                int want_c2, or_result *res, or_pair *pairs, uint64_t cap);
int or_lazy_nlj(const or_tuple *R, uint64_t nR, const or_tuple *S, uint64_t nS,
                or_result *res, or_pair *pairs, uint64_t cap);

/* Batched: n_windows windows given by offset arrays (n_windows+1 entries). */
int or_lazy_batch(int algorithm, const or_tuple *R, const uint64_t *roff, const or_tuple *S,
                  const uint64_t *soff, uint32_t n_windows, uint64_t D, int want_c2,
                  or_result *res);

/* Eager symmetric hash join (build own side, probe the other). */
typedef struct or_eager or_eager;
or_eager *or_eager_new(uint64_t D, uint64_t max_window);
void or_eager_free(or_eager *e);
void or_eager_clear(or_eager *e);
int or_eager_execute(or_eager *e, const or_tuple *batch, uint64_t n, int is_left, int want_c2,
                     or_result *res, or_pair *pairs, uint64_t cap);

/* Sliding count window (size W, slide `slide`) over two streams of equal
 * length n: window t covers tuples [t*slide + slide - W, t*slide + slide)
 * clipped at 0, for t = 0 .. n/slide - 1.  Brute force (sort+merge) per
 * window: count and C1.                                                    */
int or_sliding(const or_tuple *R, const or_tuple *S, uint64_t n, uint64_t W, uint64_t slide,
               or_result *res);

/* Sliding count window with batches of varying size: after event i the
 * window of each side is its last W tuples, R[max(0, endR[i]-W), endR[i])
 * and likewise for S.  Count and C1 per event (closed form per key).      */
int or_sliding_var(const or_tuple *R, const uint64_t *endR, const or_tuple *S, const uint64_t *endS,
                   uint64_t n_events, uint64_t W, or_result *res);

/* Timed CPU baseline: lazy hash join over consecutive tumbling windows of W
 * tuples with `threads` probe threads (INTRA decomposition of
 * executable/runExperiment.cpp:163-166).  Returns elapsed seconds.        */
double or_bench_lazy_hj(const or_tuple *R, const or_tuple *S, uint64_t n, uint64_t W, uint64_t D,
                        int threads, or_result *total);
/* Timed sliding-window baseline: after event i (left stream R up to
 * endR[i], right stream S up to endS[i]) the window is the last W tuples of
 * each side; res[i] = count and C1 of its join, maintained incrementally
 * with one hash map per side (single thread).  Returns elapsed seconds.   */
double or_bench_slide(const or_tuple *R, const uint64_t *endR, const or_tuple *S, const uint64_t *endS,
                      uint64_t n_events, uint64_t W, or_result *res);
/* The same windows with the INTER decomposition (runExperiment.cpp:167-170):
 * `threads` joiners with one probe thread each take whole windows
 * round-robin.  Returns elapsed seconds (-1 on allocation failure).       */
double or_bench_lazy_hj_inter(const or_tuple *R, const or_tuple *S, uint64_t n, uint64_t W, uint64_t D,
                              int threads, or_result *total);

/* Learned partitioner root, float64: the 1-32-32-1 ReLU MLP
 *   y = w3 . relu(W2 relu(W1 x + b1) + b2) + b3,  loss = mean (y - t)^2
 * with parameters P[1153] = W1[32], b1[32], W2[32][32] (row i = output unit
 * i), b2[32], w3[32], b3.  Writes d loss / d P into grad and returns the
 * loss (analytic back-propagation; relu'(0) = 0).  No reference
 * counterpart: pinned by central finite differences in tests.           */
double or_mlp_grad(const double *P, const double *x, const double *t, uint32_t n, double *grad);
/* loss only (for finite differences) */
double or_mlp_loss(const double *P, const double *x, const double *t, uint32_t n);

#ifdef __cplusplus
}
#endif
#endif
