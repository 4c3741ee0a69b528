// Lazy executors.  GPU_* adapt BaseLazyExecutor::execute to the C ABI
// (sjb_lazy_execute gathers the ring slots and joins on the device).
// CPU_* restate the reference CPU joins (src/sjb/windowing/lazy/executor/
// CPU/CPU_{HJ,SMJ,NLJ}.cpp) with std::thread instead of TBB and add C1.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <thread>

#include "sjb/utils/Error.hpp"
#include "sjb/windowing/lazy/executor/Executors.hpp"

namespace Windowing {
namespace Lazy {
namespace Executor {

GPU_Lazy::GPU_Lazy(int algorithm, int method, uint64_t numDistinctKeys, uint64_t batchSize,
                   uint64_t maxTupleInWindow, double estimatedSelectivity, uint32_t flags)
    : algorithm(algorithm), method(method), numDistinctKeys(numDistinctKeys), batchSize(batchSize),
      maxTupleInWindow(maxTupleInWindow), estimatedSelectivity(estimatedSelectivity), flags(flags) {
    sjb_options_t o{};
    o.max_window_tuples = maxTupleInWindow;
    o.num_distinct_keys = numDistinctKeys;
    o.batch_size = batchSize;
    sjbCheck(sjb_open(0, &o, &ctx), "sjb_open");
}

GPU_Lazy::~GPU_Lazy() {
    if (ctx) sjb_close(ctx);
}

ExecutionStatistic GPU_Lazy::execute(Tuple *leftRingBuffer, Tuple *rightRingBuffer,
                                     std::vector<uint64_t> leftIndexesToJoin,
                                     std::vector<uint64_t> rightIndexesToJoin, Sink &sink) {
    ExecutionStatistic es;
    sjb_stats_t st{};
    const uint64_t nL = leftIndexesToJoin.size() * batchSize, nR = rightIndexesToJoin.size() * batchSize;
    auto call = [&](int m, ResultTuple *out, uint64_t cap) {
        return sjb_lazy_execute(ctx, algorithm, m, flags | SJB_F_TIMING, as_abi(leftRingBuffer),
                                leftIndexesToJoin.data(), leftIndexesToJoin.size(), as_abi(rightRingBuffer),
                                rightIndexesToJoin.data(), rightIndexesToJoin.size(), batchSize,
                                reinterpret_cast<sjb_result_tuple_t *>(out), cap, &st);
    };
    if (method == SJB_OUT_NOOUTPUT) {
        sjbCheck(call(SJB_OUT_NOOUTPUT, nullptr, 0), "sjb_lazy_execute");
        es.resultCount = st.result_count;
    } else if (method == SJB_OUT_ATOMIC) {
        // pairs wrap around the sink ring; the count goes to counterGPU
        // (the reference's Atomic executors add to the pinned counterGPU)
        sjbCheck(call(SJB_OUT_ATOMIC, sink.sinkBuffer, sink.getSinkBufferSize()), "sjb_lazy_execute");
        __atomic_fetch_add(sink.counterGPU, (unsigned long long)st.result_count, __ATOMIC_RELAXED);
    } else {
        // CountKernel / EstimatedSelectivity / Bitmap: write every pair, then
        // hand them to the sink in sink-sized chunks, like the reference
        // (GPU_HJ_CountKernel.cu:117-127).  ES starts with the 1.125 x
        // (1/D) x nL x nR estimate (GPU_HJ_EstimatedSelectivity.cu:109); a
        // too-small buffer is regrown to the exact count and rerun.
        uint64_t cap = scratch.size();
        if (method == SJB_OUT_ESTIMATEDSELECTIVITY && cap == 0)
            cap = uint64_t(1.125 * std::ceil(estimatedSelectivity * double(nL) * double(nR))) + 1;
        if (scratch.size() < cap) scratch.resize(cap);
        int rc = call(method, scratch.data(), scratch.size());
        if (rc == SJB_E_CAPACITY) {
            scratch.resize(st.result_count + 1);
            rc = call(method, scratch.data(), scratch.size());
        }
        sjbCheck(rc, "sjb_lazy_execute");
        uint64_t written = 0;
        while (written < st.result_count) {
            const uint64_t n = std::min(sink.getSinkBufferSize(), st.result_count - written);
            sink.incrementCounterAndStore(scratch.data() + written, n);
            written += n;
        }
    }
    if (algorithm == SJB_ALG_SMJ) {
        es.sortTime = st.sort_time_us;
        es.mergeTime = st.merge_time_us;
    } else {
        es.buildTime = st.build_time_us;
        es.probeTime = st.probe_time_us;
    }
    es.checksum1 = st.checksum1;
    es.checksum2 = st.checksum2;
    es.checksum2Valid = st.c2_valid != 0;
    sink.addToStatistics(es);
    return es;
}

static std::vector<Tuple> gatherWindow(Tuple *ring, const std::vector<uint64_t> &idx, uint64_t B) {
    std::vector<Tuple> w(idx.size() * B);
    for (size_t i = 0; i < idx.size(); i++) std::copy(ring + idx[i], ring + idx[i] + B, w.begin() + i * B);
    return w;
}

// ---- CPU_HJ (CPU_HJ.cpp:30-136): counting-sort build on key % D, probe
// threads scan the bucket of each right tuple.
CPU_HJ::CPU_HJ(uint64_t distinctKeys, uint64_t batchSize, uint64_t nProbeThreads, uint64_t maxTupleInWindow)
    : distinctKeys(distinctKeys ? distinctKeys : 1), batchSize(batchSize),
      nProbeThreads(nProbeThreads ? nProbeThreads : 1), maxTupleInWindow(maxTupleInWindow),
      htStore(maxTupleInWindow) {}

ExecutionStatistic CPU_HJ::execute(Tuple *leftRingBuffer, Tuple *rightRingBuffer,
                                   std::vector<uint64_t> leftIndexesToJoin,
                                   std::vector<uint64_t> rightIndexesToJoin, Sink &sink) {
    const uint64_t D = distinctKeys;
    std::vector<uint64_t> occ(D, 0), prefix(D, 0), ins(D, 0);
    const auto t0 = std::chrono::high_resolution_clock::now();
    const std::vector<Tuple> L = gatherWindow(leftRingBuffer, leftIndexesToJoin, batchSize);
    const std::vector<Tuple> R = gatherWindow(rightRingBuffer, rightIndexesToJoin, batchSize);
    if (htStore.size() < L.size()) htStore.resize(L.size());
    for (const Tuple &t : L) occ[t.key % D]++;
    uint64_t run = 0;
    for (uint64_t h = 0; h < D; h++) {
        prefix[h] = run;
        run += occ[h];
    }
    for (const Tuple &t : L) {
        const uint64_t h = t.key % D;
        htStore[prefix[h] + ins[h]++] = t;
    }
    const auto t1 = std::chrono::high_resolution_clock::now();
    const unsigned T = unsigned(std::min<uint64_t>(nProbeThreads, 256));
    std::vector<uint64_t> cnt(T, 0), c1(T, 0);
    auto work = [&](unsigned t) {
        const uint64_t b = R.size() * t / T, e = R.size() * (t + 1) / T;
        for (uint64_t i = b; i < e; i++) {
            const uint64_t h = R[i].key % D;
            for (uint64_t j = prefix[h]; j < prefix[h] + occ[h]; j++)
                if (htStore[j].key == R[i].key) {
                    cnt[t]++;
                    c1[t] += sjb_hash_a(htStore[j].key, htStore[j].val) * sjb_hash_b(R[i].key, R[i].val);
                }
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < T; t++) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    const auto t2 = std::chrono::high_resolution_clock::now();
    ExecutionStatistic es;
    for (unsigned t = 0; t < T; t++) {
        es.resultCount += cnt[t];
        es.checksum1 += c1[t];
    }
    es.buildTime = uint64_t(std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count());
    es.probeTime = uint64_t(std::chrono::duration_cast<std::chrono::microseconds>(t2 - t1).count());
    sink.addToStatistics(es);
    return es;
}

// ---- CPU_SMJ (CPU_SMJ.cpp:23-139): sort both sides by key, then the
// cartesian product of every equal-key run (closed form per run).
CPU_SMJ::CPU_SMJ(uint64_t numDistinctKey, uint64_t batchSize, uint64_t nSortThreads)
    : numDistinctKey(numDistinctKey), batchSize(batchSize), nSortThreads(nSortThreads) {}

static void parallelSort(std::vector<Tuple> &v, unsigned T) {
    auto cmp = [](const Tuple &a, const Tuple &b) { return a.key < b.key; };
    if (T <= 1 || v.size() < 4096) {
        std::sort(v.begin(), v.end(), cmp);
        return;
    }
    std::vector<std::thread> th;
    std::vector<size_t> cut(T + 1);
    for (unsigned t = 0; t <= T; t++) cut[t] = v.size() * t / T;
    for (unsigned t = 0; t < T; t++)
        th.emplace_back([&, t] { std::sort(v.begin() + cut[t], v.begin() + cut[t + 1], cmp); });
    for (auto &x : th) x.join();
    for (unsigned w = 1; w < T; w *= 2)
        for (unsigned t = 0; t + w < T; t += 2 * w)
            std::inplace_merge(v.begin() + cut[t], v.begin() + cut[t + w], v.begin() + cut[std::min(t + 2 * w, T)], cmp);
}

ExecutionStatistic CPU_SMJ::execute(Tuple *leftRingBuffer, Tuple *rightRingBuffer,
                                    std::vector<uint64_t> leftIndexesToJoin,
                                    std::vector<uint64_t> rightIndexesToJoin, Sink &sink) {
    const auto t0 = std::chrono::high_resolution_clock::now();
    std::vector<Tuple> L = gatherWindow(leftRingBuffer, leftIndexesToJoin, batchSize);
    std::vector<Tuple> R = gatherWindow(rightRingBuffer, rightIndexesToJoin, batchSize);
    const unsigned T = unsigned(std::max<uint64_t>(1, std::min<uint64_t>(nSortThreads, 64)));
    parallelSort(L, T);
    parallelSort(R, T);
    const auto t1 = std::chrono::high_resolution_clock::now();
    ExecutionStatistic es;
    size_t i = 0, j = 0;
    while (i < L.size() && j < R.size()) {
        if (L[i].key < R[j].key) { i++; continue; }
        if (L[i].key > R[j].key) { j++; continue; }
        const uint32_t k = L[i].key;
        size_t i1 = i, j1 = j;
        uint64_t sa = 0, sb = 0;
        while (i1 < L.size() && L[i1].key == k) sa += sjb_hash_a(k, L[i1++].val);
        while (j1 < R.size() && R[j1].key == k) sb += sjb_hash_b(k, R[j1++].val);
        es.resultCount += uint64_t(i1 - i) * uint64_t(j1 - j);
        es.checksum1 += sa * sb;
        i = i1;
        j = j1;
    }
    const auto t2 = std::chrono::high_resolution_clock::now();
    es.sortTime = uint64_t(std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count());
    es.mergeTime = uint64_t(std::chrono::duration_cast<std::chrono::microseconds>(t2 - t1).count());
    sink.addToStatistics(es);
    return es;
}

// ---- CPU_NLJ (CPU_NLJ.cpp:11-32): every left tuple against every right tuple.
CPU_NLJ::CPU_NLJ(uint64_t batchSize) : batchSize(batchSize) {}

ExecutionStatistic CPU_NLJ::execute(Tuple *leftRingBuffer, Tuple *rightRingBuffer,
                                    std::vector<uint64_t> leftIndexesToJoin,
                                    std::vector<uint64_t> rightIndexesToJoin, Sink &sink) {
    ExecutionStatistic es;
    for (uint64_t i : leftIndexesToJoin)
        for (uint64_t a = i; a < i + batchSize; a++)
            for (uint64_t k : rightIndexesToJoin)
                for (uint64_t b = k; b < k + batchSize; b++)
                    if (leftRingBuffer[a].key == rightRingBuffer[b].key) {
                        es.resultCount++;
                        es.checksum1 += sjb_hash_a(leftRingBuffer[a].key, leftRingBuffer[a].val) *
                                        sjb_hash_b(rightRingBuffer[b].key, rightRingBuffer[b].val);
                    }
    sink.addToStatistics(es);
    return es;
}

}  // namespace Executor
}  // namespace Lazy
}  // namespace Windowing
