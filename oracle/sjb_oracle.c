/*
 * sjb_oracle.c — CPU oracle (test infrastructure only; see sjb_oracle.h).
 *
 * Every function cites the reference code whose semantics it restates.
 * The checksum formulas are the ones documented in include/sjb.h; they are
 * written out here independently of the product code.
 */
#define _GNU_SOURCE
#include "sjb_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ---------------------------------------------------------------------- */
/* Checksum primitives.  mix64 is the SplitMix64 output function (Steele,  */
/* Lea, Flood, "Fast splittable pseudorandom number generators", 2014).    */
/* ---------------------------------------------------------------------- */
#define OR_SEED_A 0x243F6A8885A308D3ULL
#define OR_SEED_B 0x13198A2E03707344ULL
#define OR_GOLDEN 0x9E3779B97F4A7C15ULL

uint64_t or_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t or_hash_a(uint32_t k, uint32_t v) { return or_mix64((((uint64_t)k << 32) | v) ^ OR_SEED_A); }
uint64_t or_hash_b(uint32_t k, uint32_t v) { return or_mix64((((uint64_t)k << 32) | v) ^ OR_SEED_B); }
uint64_t or_pair_hash(uint32_t k, uint32_t rv, uint32_t sv) {
    return or_mix64(or_hash_a(k, rv) + OR_GOLDEN * ((((uint64_t)sv) << 1) | 1ULL));
}

static inline void emit(or_result *res, or_pair *pairs, uint64_t cap, const or_tuple *r,
                        const or_tuple *s, int want_c2) {
    res->c1 += or_hash_a(r->key, r->val) * or_hash_b(s->key, s->val);
    if (want_c2) res->c2 += or_pair_hash(r->key, r->val, s->val);
    if (pairs && res->count < cap) {
        or_pair *p = &pairs[res->count];
        p->lkey = r->key;
        p->rkey = s->key;
        p->lval = r->val;
        p->rval = s->val;
        p->lts = r->ts;
        p->rts = s->ts;
        res->pairs = res->count + 1;
    }
    res->count++;
}

/* ---------------------------------------------------------------------- */
/* Lazy hash join — CPU_HJ::execute / computePrefixSum / build / probe     */
/* (src/sjb/windowing/lazy/executor/CPU/CPU_HJ.cpp:30-136):                */
/*   occupation[h]++ over the build side, exclusive prefix sum over D      */
/*   buckets (:121-136), counting-sort scatter store[prefix[h]+inserted[h]++] */
/*   (:84-99), probe scans bucket h comparing keys (:101-119).             */
/* h = key % D for both the histogram and the scatter (the reference      */
/* indexes the histogram by raw key, :126, and compares the hash to the   */
/* stored key, :113; both are defects when key >= D and are not kept).    */
/* ---------------------------------------------------------------------- */
int or_lazy_hj(const or_tuple *R, uint64_t nR, const or_tuple *S, uint64_t nS, uint64_t D,
               int want_c2, or_result *res, or_pair *pairs, uint64_t cap) {
    memset(res, 0, sizeof(*res));
    if (D == 0) return -1;
    uint64_t *occ = calloc(D, sizeof(uint64_t));
    uint64_t *prefix = calloc(D, sizeof(uint64_t));
    uint64_t *ins = calloc(D, sizeof(uint64_t));
    or_tuple *store = malloc((nR ? nR : 1) * sizeof(or_tuple));
    if (!occ || !prefix || !ins || !store) {
        free(occ); free(prefix); free(ins); free(store);
        return -1;
    }
    for (uint64_t i = 0; i < nR; i++) occ[R[i].key % D]++;
    uint64_t run = 0;
    for (uint64_t h = 0; h < D; h++) {
        prefix[h] = run;
        run += occ[h];
    }
    for (uint64_t i = 0; i < nR; i++) {
        uint64_t h = R[i].key % D;
        store[prefix[h] + ins[h]++] = R[i];
    }
    for (uint64_t i = 0; i < nS; i++) {
        uint64_t h = S[i].key % D;
        for (uint64_t j = prefix[h]; j < prefix[h] + occ[h]; j++)
            if (store[j].key == S[i].key) emit(res, pairs, cap, &store[j], &S[i], want_c2);
    }
    free(occ); free(prefix); free(ins); free(store);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Lazy sort-merge join — CPU_SMJ::execute                                */
/* (src/sjb/windowing/lazy/executor/CPU/CPU_SMJ.cpp:23-139): copy both    */
/* sides to local buffers, sort each by key (compareEventTuple = key <,   */
/* include/sjb/windowing/lazy/executor/CPU/CPU_SMJ.hpp:27-29), then the   */
/* cartesian product of every equal-key run (:99-125).  The reference     */
/* walks runs through D-sized prefix arrays; this walks them with two     */
/* pointers, which needs no bound on the key range.  Count and C1 use the */
/* per-run closed forms |L_k|*|R_k| and (sum A)(sum B); pairs and C2 are  */
/* enumerated only when asked for.                                         */
/* ---------------------------------------------------------------------- */
static int cmp_key_val(const void *a, const void *b) {
    const or_tuple *x = a, *y = b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    if (x->val != y->val) return x->val < y->val ? -1 : 1;
    if (x->ts != y->ts) return x->ts < y->ts ? -1 : 1;
    return 0;
}

int or_lazy_smj(const or_tuple *R, uint64_t nR, const or_tuple *S, uint64_t nS, int want_c2,
                or_result *res, or_pair *pairs, uint64_t cap) {
    memset(res, 0, sizeof(*res));
    or_tuple *L = malloc((nR ? nR : 1) * sizeof(or_tuple));
    or_tuple *Rt = malloc((nS ? nS : 1) * sizeof(or_tuple));
    if (!L || !Rt) {
        free(L); free(Rt);
        return -1;
    }
    memcpy(L, R, nR * sizeof(or_tuple));
    memcpy(Rt, S, nS * sizeof(or_tuple));
    qsort(L, nR, sizeof(or_tuple), cmp_key_val);
    qsort(Rt, nS, sizeof(or_tuple), cmp_key_val);
    uint64_t i = 0, j = 0;
    while (i < nR && j < nS) {
        if (L[i].key < Rt[j].key) { i++; continue; }
        if (L[i].key > Rt[j].key) { j++; continue; }
        uint32_t k = L[i].key;
        uint64_t i1 = i, j1 = j;
        while (i1 < nR && L[i1].key == k) i1++;
        while (j1 < nS && Rt[j1].key == k) j1++;
        if (want_c2 || pairs) {
            for (uint64_t a = i; a < i1; a++)
                for (uint64_t b = j; b < j1; b++) emit(res, pairs, cap, &L[a], &Rt[b], want_c2);
        } else {
            uint64_t sa = 0, sb = 0;
            for (uint64_t a = i; a < i1; a++) sa += or_hash_a(k, L[a].val);
            for (uint64_t b = j; b < j1; b++) sb += or_hash_b(k, Rt[b].val);
            res->count += (i1 - i) * (j1 - j);
            res->c1 += sa * sb;
        }
        i = i1;
        j = j1;
    }
    free(L); free(Rt);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Lazy nested-loop join — CPU_NLJ::execute                                */
/* (src/sjb/windowing/lazy/executor/CPU/CPU_NLJ.cpp:11-32): every left     */
/* tuple against every right tuple, counter += (l.key == r.key).          */
/* ---------------------------------------------------------------------- */
int or_lazy_nlj(const or_tuple *R, uint64_t nR, const or_tuple *S, uint64_t nS, or_result *res,
                or_pair *pairs, uint64_t cap) {
    memset(res, 0, sizeof(*res));
    for (uint64_t i = 0; i < nR; i++)
        for (uint64_t j = 0; j < nS; j++)
            if (R[i].key == S[j].key) emit(res, pairs, cap, &R[i], &S[j], 1);
    return 0;
}

int or_lazy_batch(int algorithm, const or_tuple *R, const uint64_t *roff, const or_tuple *S,
                  const uint64_t *soff, uint32_t n_windows, uint64_t D, int want_c2,
                  or_result *res) {
    for (uint32_t w = 0; w < n_windows; w++) {
        const or_tuple *r = R + roff[w];
        const or_tuple *s = S + soff[w];
        uint64_t nr = roff[w + 1] - roff[w], ns = soff[w + 1] - soff[w];
        int rc;
        if (algorithm == 1)
            rc = or_lazy_nlj(r, nr, s, ns, &res[w], NULL, 0);
        else if (algorithm == 2)
            rc = or_lazy_hj(r, nr, s, ns, D, want_c2, &res[w], NULL, 0);
        else
            rc = or_lazy_smj(r, nr, s, ns, want_c2, &res[w], NULL, 0);
        if (rc) return rc;
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Eager symmetric hash join — Eager CPU_HJ                                */
/* (src/sjb/windowing/eager/executor/CPU/CPU_HJ.cpp:35-82): execute()      */
/* builds the batch into its own side's bucketed table and probes the      */
/* other side's table (:35-48); clearStates() resets both occupations     */
/* (:25-31).  The reference reserves D x maxTupleInWindow slots per side   */
/* (:9-10); here each bucket is a singly linked list over an append-only   */
/* array, which holds the same multiset in O(window) memory.               */
/* ---------------------------------------------------------------------- */
typedef struct or_side {
    or_tuple *tuples;
    uint64_t *next;
    uint64_t *head;
    uint64_t n;
} or_side;

struct or_eager {
    uint64_t D, cap;
    or_side side[2];
};

or_eager *or_eager_new(uint64_t D, uint64_t max_window) {
    or_eager *e = calloc(1, sizeof(or_eager));
    if (!e || D == 0) { free(e); return NULL; }
    e->D = D;
    e->cap = max_window;
    for (int s = 0; s < 2; s++) {
        e->side[s].tuples = malloc((max_window ? max_window : 1) * sizeof(or_tuple));
        e->side[s].next = malloc((max_window ? max_window : 1) * sizeof(uint64_t));
        e->side[s].head = malloc(D * sizeof(uint64_t));
        if (!e->side[s].tuples || !e->side[s].next || !e->side[s].head) {
            or_eager_free(e);
            return NULL;
        }
    }
    or_eager_clear(e);
    return e;
}

void or_eager_free(or_eager *e) {
    if (!e) return;
    for (int s = 0; s < 2; s++) {
        free(e->side[s].tuples);
        free(e->side[s].next);
        free(e->side[s].head);
    }
    free(e);
}

void or_eager_clear(or_eager *e) {
    for (int s = 0; s < 2; s++) {
        e->side[s].n = 0;
        memset(e->side[s].head, 0xFF, e->D * sizeof(uint64_t));
    }
}

int or_eager_execute(or_eager *e, const or_tuple *batch, uint64_t n, int is_left, int want_c2,
                     or_result *res, or_pair *pairs, uint64_t cap) {
    memset(res, 0, sizeof(*res));
    or_side *own = &e->side[is_left ? 0 : 1];
    or_side *other = &e->side[is_left ? 1 : 0];
    if (own->n + n > e->cap) return -2;
    /* build own side (:38/:41) */
    for (uint64_t i = 0; i < n; i++) {
        uint64_t h = batch[i].key % e->D;
        uint64_t g = own->n++;
        own->tuples[g] = batch[i];
        own->next[g] = own->head[h];
        own->head[h] = g;
    }
    /* probe the other side (:39/:42) */
    for (uint64_t i = 0; i < n; i++) {
        uint64_t h = batch[i].key % e->D;
        for (uint64_t g = other->head[h]; g != UINT64_MAX; g = other->next[g]) {
            if (other->tuples[g].key != batch[i].key) continue;
            if (is_left)
                emit(res, pairs, cap, &batch[i], &other->tuples[g], want_c2);
            else
                emit(res, pairs, cap, &other->tuples[g], &batch[i], want_c2);
        }
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Sliding window brute force (no reference counterpart; the reference has */
/* tumbling windows only).  Each window is joined from scratch with the   */
/* sort-merge closed form.                                                 */
/* ---------------------------------------------------------------------- */
int or_sliding(const or_tuple *R, const or_tuple *S, uint64_t n, uint64_t W, uint64_t slide,
               or_result *res) {
    if (slide == 0) return -1;
    uint64_t nw = n / slide;
    for (uint64_t t = 0; t < nw; t++) {
        uint64_t end = (t + 1) * slide;
        uint64_t begin = end > W ? end - W : 0;
        int rc = or_lazy_smj(R + begin, end - begin, S + begin, end - begin, 0, &res[t], NULL, 0);
        if (rc) return rc;
    }
    return 0;
}

int or_sliding_var(const or_tuple *R, const uint64_t *endR, const or_tuple *S, const uint64_t *endS,
                   uint64_t n_events, uint64_t W, or_result *res) {
    for (uint64_t i = 0; i < n_events; i++) {
        const uint64_t bR = endR[i] > W ? endR[i] - W : 0, bS = endS[i] > W ? endS[i] - W : 0;
        int rc = or_lazy_smj(R + bR, endR[i] - bR, S + bS, endS[i] - bS, 0, &res[i], NULL, 0);
        if (rc) return rc;
    }
    return 0;
}

/* ---------------------------------------------------------------------- */
/* Timed CPU baseline: the lazy CPU_HJ over tumbling windows of W tuples   */
/* per side with the INTRA decomposition (1 joiner, `threads` probe        */
/* threads: executable/runExperiment.cpp:163-166; the probe is split over  */
/* threads like the parallel_for_each of CPU_HJ.cpp:53-65).                */
/* ---------------------------------------------------------------------- */
typedef struct probe_arg {
    const or_tuple *S;
    uint64_t begin, end, D;
    const uint64_t *occ, *prefix;
    const or_tuple *store;
    or_result r;
} probe_arg;

static void *probe_worker(void *p) {
    probe_arg *a = p;
    memset(&a->r, 0, sizeof(a->r));
    for (uint64_t i = a->begin; i < a->end; i++) {
        uint64_t h = a->S[i].key % a->D;
        for (uint64_t j = a->prefix[h]; j < a->prefix[h] + a->occ[h]; j++)
            if (a->store[j].key == a->S[i].key) {
                a->r.count++;
                a->r.c1 += or_hash_a(a->store[j].key, a->store[j].val) *
                           or_hash_b(a->S[i].key, a->S[i].val);
                a->r.c2 += or_pair_hash(a->S[i].key, a->store[j].val, a->S[i].val);
            }
    }
    return NULL;
}

double or_bench_lazy_hj(const or_tuple *R, const or_tuple *S, uint64_t n, uint64_t W, uint64_t D,
                        int threads, or_result *total) {
    memset(total, 0, sizeof(*total));
    if (threads < 1) threads = 1;
    if (threads > 1024) threads = 1024;
    uint64_t *occ = malloc(D * sizeof(uint64_t));
    uint64_t *prefix = malloc(D * sizeof(uint64_t));
    uint64_t *ins = malloc(D * sizeof(uint64_t));
    or_tuple *store = malloc(W * sizeof(or_tuple));
    pthread_t *tid = malloc((size_t)threads * sizeof(pthread_t));
    probe_arg *args = malloc((size_t)threads * sizeof(probe_arg));
    if (!occ || !prefix || !ins || !store || !tid || !args) {
        free(occ); free(prefix); free(ins); free(store); free(tid); free(args);
        return -1.0;
    }
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (uint64_t w = 0; w + W <= n; w += W) {
        const or_tuple *r = R + w, *s = S + w;
        memset(occ, 0, D * sizeof(uint64_t));
        memset(ins, 0, D * sizeof(uint64_t));
        for (uint64_t i = 0; i < W; i++) occ[r[i].key % D]++;
        uint64_t run = 0;
        for (uint64_t h = 0; h < D; h++) {
            prefix[h] = run;
            run += occ[h];
        }
        for (uint64_t i = 0; i < W; i++) {
            uint64_t h = r[i].key % D;
            store[prefix[h] + ins[h]++] = r[i];
        }
        for (int t = 0; t < threads; t++) {
            args[t].S = s;
            args[t].begin = W * (uint64_t)t / threads;
            args[t].end = W * (uint64_t)(t + 1) / threads;
            args[t].D = D;
            args[t].occ = occ;
            args[t].prefix = prefix;
            args[t].store = store;
            if (threads > 1)
                pthread_create(&tid[t], NULL, probe_worker, &args[t]);
            else
                probe_worker(&args[t]);
        }
        for (int t = 0; t < threads; t++) {
            if (threads > 1) pthread_join(tid[t], NULL);
            total->count += args[t].r.count;
            total->c1 += args[t].r.c1;
            total->c2 += args[t].r.c2;
        }
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(occ); free(prefix); free(ins); free(store);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}


/* ---- learned partitioner root (float64) ---------------------------------- */
#define OR_H 32
#define OR_W1 0
#define OR_B1 32
#define OR_W2 64
#define OR_B2 (64 + OR_H * OR_H)
#define OR_W3 (OR_B2 + OR_H)
#define OR_B3 (OR_W3 + OR_H)

static double mlp_forward(const double *P, double x, double *h1, double *h2) {
    double y = P[OR_B3];
    for (int j = 0; j < OR_H; j++) {
        const double z = P[OR_W1 + j] * x + P[OR_B1 + j];
        h1[j] = z > 0 ? z : 0;
    }
    for (int i = 0; i < OR_H; i++) {
        double z = P[OR_B2 + i];
        for (int j = 0; j < OR_H; j++) z += P[OR_W2 + i * OR_H + j] * h1[j];
        h2[i] = z > 0 ? z : 0;
        y += P[OR_W3 + i] * h2[i];
    }
    return y;
}

double or_mlp_loss(const double *P, const double *x, const double *t, uint32_t n) {
    double h1[OR_H], h2[OR_H], loss = 0;
    for (uint32_t s = 0; s < n; s++) {
        const double e = mlp_forward(P, x[s], h1, h2) - t[s];
        loss += e * e;
    }
    return loss / n;
}

double or_mlp_grad(const double *P, const double *x, const double *t, uint32_t n, double *grad) {
    double h1[OR_H], h2[OR_H], dz2[OR_H], loss = 0;
    memset(grad, 0, sizeof(double) * (OR_B3 + 1));
    for (uint32_t s = 0; s < n; s++) {
        const double e = mlp_forward(P, x[s], h1, h2) - t[s];
        loss += e * e;
        const double dy = 2.0 * e / n;
        grad[OR_B3] += dy;
        for (int i = 0; i < OR_H; i++) {
            grad[OR_W3 + i] += dy * h2[i];
            dz2[i] = h2[i] > 0 ? dy * P[OR_W3 + i] : 0;
            grad[OR_B2 + i] += dz2[i];
            for (int j = 0; j < OR_H; j++) grad[OR_W2 + i * OR_H + j] += dz2[i] * h1[j];
        }
        for (int j = 0; j < OR_H; j++) {
            if (h1[j] <= 0) continue;
            double dh = 0;
            for (int i = 0; i < OR_H; i++) dh += dz2[i] * P[OR_W2 + i * OR_H + j];
            grad[OR_W1 + j] += dh * x[s];
            grad[OR_B1 + j] += dh;
        }
    }
    return loss / n;
}

/* ---------------------------------------------------------------------- */
/* Timed CPU baseline, INTER decomposition: `threads` joiners, each with   */
/* its own single-threaded CPU_HJ (executable/runExperiment.cpp:167-170:   */
/* INTER = joinerThreads executors, one probe thread each), take whole     */
/* tumbling windows round-robin (window w to joiner w mod threads).        */
/* ---------------------------------------------------------------------- */
typedef struct joiner_arg {
    const or_tuple *R, *S;
    uint64_t n, W, D;
    int id, threads;
    or_result r;
    int fail;
} joiner_arg;

static void *joiner_worker(void *p) {
    joiner_arg *a = p;
    memset(&a->r, 0, sizeof(a->r));
    const uint64_t D = a->D, W = a->W;
    uint64_t *occ = malloc(D * sizeof(uint64_t));
    uint64_t *prefix = malloc(D * sizeof(uint64_t));
    uint64_t *ins = malloc(D * sizeof(uint64_t));
    or_tuple *store = malloc((W ? W : 1) * sizeof(or_tuple));
    if (!occ || !prefix || !ins || !store) {
        a->fail = 1;
        free(occ); free(prefix); free(ins); free(store);
        return NULL;
    }
    uint64_t w = 0;
    for (uint64_t o = 0; o + W <= a->n; o += W, w++) {
        if ((int)(w % (uint64_t)a->threads) != a->id) continue;
        const or_tuple *r = a->R + o, *s = a->S + o;
        memset(occ, 0, D * sizeof(uint64_t));
        memset(ins, 0, D * sizeof(uint64_t));
        for (uint64_t i = 0; i < W; i++) occ[r[i].key % D]++;
        uint64_t run = 0;
        for (uint64_t h = 0; h < D; h++) {
            prefix[h] = run;
            run += occ[h];
        }
        for (uint64_t i = 0; i < W; i++) {
            uint64_t h = r[i].key % D;
            store[prefix[h] + ins[h]++] = r[i];
        }
        for (uint64_t i = 0; i < W; i++) {
            uint64_t h = s[i].key % D;
            for (uint64_t j = prefix[h]; j < prefix[h] + occ[h]; j++)
                if (store[j].key == s[i].key) {
                    a->r.count++;
                    a->r.c1 += or_hash_a(store[j].key, store[j].val) * or_hash_b(s[i].key, s[i].val);
                    a->r.c2 += or_pair_hash(s[i].key, store[j].val, s[i].val);
                }
        }
    }
    free(occ); free(prefix); free(ins); free(store);
    return NULL;
}

double or_bench_lazy_hj_inter(const or_tuple *R, const or_tuple *S, uint64_t n, uint64_t W, uint64_t D,
                              int threads, or_result *total) {
    memset(total, 0, sizeof(*total));
    if (threads < 1) threads = 1;
    if (threads > 1024) threads = 1024;
    if (D == 0 || W == 0) return -1.0;
    pthread_t *tid = malloc((size_t)threads * sizeof(pthread_t));
    joiner_arg *args = calloc((size_t)threads, sizeof(joiner_arg));
    if (!tid || !args) {
        free(tid); free(args);
        return -1.0;
    }
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int t = 0; t < threads; t++) {
        args[t].R = R;
        args[t].S = S;
        args[t].n = n;
        args[t].W = W;
        args[t].D = D;
        args[t].id = t;
        args[t].threads = threads;
        pthread_create(&tid[t], NULL, joiner_worker, &args[t]);
    }
    int fail = 0;
    for (int t = 0; t < threads; t++) {
        pthread_join(tid[t], NULL);
        fail |= args[t].fail;
        total->count += args[t].r.count;
        total->c1 += args[t].r.c1;
        total->c2 += args[t].r.c2;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(tid); free(args);
    if (fail) return -1.0;
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ---------------------------------------------------------------------- */
/* Timed CPU baseline of the sliding-window join (no reference             */
/* counterpart: the reference windows are tumbling).  The same exact       */
/* incremental method as the device path, single-threaded: per side a     */
/* hash map key -> (tuples in the window, sum of A or B); after every      */
/* event the window is the last W tuples of each side and                  */
/*   J(L',R') = J(L,R) - J(Lout,R) - J(Lkeep,Rout) + J(Lin,Rkeep) + J(L',Rin) */
/* for count and C1.  res[i] is the window result after event i.           */
/* ---------------------------------------------------------------------- */
typedef struct slide_slot {
    uint32_t key;
    uint32_t used;
    uint64_t cnt;
    uint64_t sum;
} slide_slot;

static slide_slot *slide_find(slide_slot *t, uint64_t mask, uint32_t key) {
    uint64_t h = (or_mix64(key) >> 7) & mask;
    while (t[h].used && t[h].key != key) h = (h + 1) & mask;
    if (!t[h].used) {
        t[h].used = 1;
        t[h].key = key;
    }
    return &t[h];
}

double or_bench_slide(const or_tuple *R, const uint64_t *endR, const or_tuple *S, const uint64_t *endS,
                      uint64_t n_events, uint64_t W, or_result *res) {
    const uint64_t nR = n_events ? endR[n_events - 1] : 0, nS = n_events ? endS[n_events - 1] : 0;
    uint64_t cap = 1024;
    while (cap < 2 * (nR > nS ? nR : nS)) cap <<= 1;
    slide_slot *tl = calloc(cap, sizeof(slide_slot)), *tr = calloc(cap, sizeof(slide_slot));
    if (!tl || !tr) {
        free(tl); free(tr);
        return -1.0;
    }
    const uint64_t mask = cap - 1;
    uint64_t count = 0, c1 = 0, pl = 0, pr = 0;  /* tuples of each side seen so far */
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (uint64_t e = 0; e < n_events; e++) {
        const uint64_t el = endR[e], er = endS[e];
        const uint64_t lo_old = pl > W ? pl - W : 0, lo_new = el > W ? el - W : 0;
        const uint64_t ro_old = pr > W ? pr - W : 0, ro_new = er > W ? er - W : 0;
        for (uint64_t i = lo_old; i < lo_new && i < pl; i++) { /* left tuples leaving */
            slide_slot *o = slide_find(tr, mask, R[i].key);
            const uint64_t a = or_hash_a(R[i].key, R[i].val);
            count -= o->cnt;
            c1 -= a * o->sum;
            slide_slot *m = slide_find(tl, mask, R[i].key);
            m->cnt--;
            m->sum -= a;
        }
        for (uint64_t i = ro_old; i < ro_new && i < pr; i++) { /* right tuples leaving */
            slide_slot *o = slide_find(tl, mask, S[i].key);
            const uint64_t b = or_hash_b(S[i].key, S[i].val);
            count -= o->cnt;
            c1 -= o->sum * b;
            slide_slot *m = slide_find(tr, mask, S[i].key);
            m->cnt--;
            m->sum -= b;
        }
        for (uint64_t i = pl > lo_new ? pl : lo_new; i < el; i++) { /* left tuples arriving */
            slide_slot *o = slide_find(tr, mask, R[i].key);
            const uint64_t a = or_hash_a(R[i].key, R[i].val);
            count += o->cnt;
            c1 += a * o->sum;
            slide_slot *m = slide_find(tl, mask, R[i].key);
            m->cnt++;
            m->sum += a;
        }
        for (uint64_t i = pr > ro_new ? pr : ro_new; i < er; i++) { /* right tuples arriving */
            slide_slot *o = slide_find(tl, mask, S[i].key);
            const uint64_t b = or_hash_b(S[i].key, S[i].val);
            count += o->cnt;
            c1 += o->sum * b;
            slide_slot *m = slide_find(tr, mask, S[i].key);
            m->cnt++;
            m->sum += b;
        }
        pl = el;
        pr = er;
        res[e].count = count;
        res[e].c1 = c1;
        res[e].c2 = 0;
        res[e].pairs = 0;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(tl); free(tr);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
