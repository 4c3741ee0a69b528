// Eager executors: GPU_* adapt BaseEagerExecutor to sjb_eager_*; CPU_HJ
// restates src/sjb/windowing/eager/executor/CPU/CPU_HJ.cpp:35-82 (build the
// batch into its own side, probe the other side) with bucket chains in
// O(window) memory instead of D x maxTupleInWindow slots.
#include <algorithm>

#include "sjb/utils/Error.hpp"
#include "sjb/windowing/eager/executor/Executors.hpp"

namespace Windowing {
namespace Eager {
namespace Executor {

// The eager execute returns only the batch's match count (the reference's
// signature); its checksums reach the sink as an ExecutionStatistic.
static void addChecksums(Sink &sink, uint64_t c1, uint64_t c2, bool c2Valid) {
    ExecutionStatistic es;
    es.checksum1 = c1;
    es.checksum2 = c2;
    es.checksum2Valid = c2Valid;
    sink.addToStatistics(es);
}

GPU_Eager::GPU_Eager(int algorithm, int method, uint64_t numDistinctKeys, uint64_t windowSize,
                     uint64_t batchSize, uint32_t flags)
    : algorithm(algorithm), method(method), numDistinctKeys(numDistinctKeys), windowSize(windowSize),
      batchSize(batchSize), flags(flags) {
    sjb_options_t o{};
    o.max_window_tuples = windowSize;
    o.num_distinct_keys = numDistinctKeys;
    o.batch_size = batchSize;
    sjbCheck(sjb_open(0, &o, &ctx), "sjb_open");
}

GPU_Eager::~GPU_Eager() {
    if (ctx) sjb_close(ctx);
}

uint64_t GPU_Eager::execute(Tuple *tupleBuffer, bool isLeftSide, uint64_t, Sink &sink) {
    sjb_stats_t st{};
    if (method == SJB_OUT_NOOUTPUT) {
        sjbCheck(sjb_eager_execute(ctx, algorithm, method, flags, as_abi(tupleBuffer), batchSize,
                                   isLeftSide ? 1 : 0, nullptr, 0, &st),
                 "sjb_eager_execute");
        addChecksums(sink, st.checksum1, st.checksum2, st.c2_valid != 0);
        return st.result_count;
    }
    if (method == SJB_OUT_ATOMIC) {
        // pairs go round the sink ring; the count goes to counterGPU
        sjbCheck(sjb_eager_execute(ctx, algorithm, method, flags, as_abi(tupleBuffer), batchSize,
                                   isLeftSide ? 1 : 0,
                                   reinterpret_cast<sjb_result_tuple_t *>(sink.sinkBuffer),
                                   sink.getSinkBufferSize(), &st),
                 "sjb_eager_execute");
        __atomic_fetch_add(sink.counterGPU, (unsigned long long)st.result_count, __ATOMIC_RELAXED);
        addChecksums(sink, st.checksum1, st.checksum2, true);
        return 0;
    }
    // Bitmap: decoded pairs are handed to the sink (the reference decodes
    // the bitmap on the host and stores every pair, GPU_NLJ_BM.cu:100-130)
    if (scratch.empty()) scratch.resize(batchSize * 4 + 1);
    int rc = sjb_eager_execute(ctx, algorithm, method, flags, as_abi(tupleBuffer), batchSize, isLeftSide ? 1 : 0,
                               reinterpret_cast<sjb_result_tuple_t *>(scratch.data()), scratch.size(), &st);
    if (rc == SJB_E_CAPACITY) {
        // rejected before the state changed: grow to the exact count, retry
        scratch.resize(st.result_count + 1);
        rc = sjb_eager_execute(ctx, algorithm, method, flags, as_abi(tupleBuffer), batchSize, isLeftSide ? 1 : 0,
                               reinterpret_cast<sjb_result_tuple_t *>(scratch.data()), scratch.size(), &st);
    }
    sjbCheck(rc, "sjb_eager_execute");
    uint64_t written = 0;
    while (written < st.result_count) {
        const uint64_t n = std::min(sink.getSinkBufferSize(), st.result_count - written);
        sink.incrementCounterAndStore(scratch.data() + written, n);
        written += n;
    }
    addChecksums(sink, st.checksum1, st.checksum2, true);
    return 0;
}

void GPU_Eager::clearStates() { sjbCheck(sjb_eager_clear(ctx), "sjb_eager_clear"); }
void GPU_Eager::onEndOfStream() { sjbCheck(sjb_eager_end_of_stream(ctx), "sjb_eager_end_of_stream"); }

GPU_HJ_PK::GPU_HJ_PK(uint64_t numDistinctKeys, uint64_t windowSize, uint64_t batchSize) : batchSize(batchSize) {
    sjb_options_t o{};
    o.max_window_tuples = windowSize;
    o.num_distinct_keys = numDistinctKeys;
    o.batch_size = batchSize;
    o.flags = SJB_F_C2;
    sjbCheck(sjb_open(0, &o, &ctx), "sjb_open");
}

GPU_HJ_PK::~GPU_HJ_PK() {
    if (ctx) sjb_close(ctx);
}

uint64_t GPU_HJ_PK::execute(Tuple *tupleBuffer, bool isLeftSide, uint64_t, Sink &sink) {
    sjb_stats_t st{};
    sjbCheck(sjb_pk_execute(ctx, as_abi(tupleBuffer), batchSize, isLeftSide ? 1 : 0, &st), "sjb_pk_execute");
    addChecksums(sink, st.checksum1, st.checksum2, true);
    return st.result_count;
}

void GPU_HJ_PK::clearStates() { sjbCheck(sjb_pk_clear(ctx), "sjb_pk_clear"); }
void GPU_HJ_PK::onEndOfStream() { sjbCheck(sjb_pk_stop(ctx), "sjb_pk_stop"); }

CPU_HJ::CPU_HJ(uint64_t maxTupleInWindow, uint64_t numDistinctKeys, uint64_t batchSize)
    : maxTupleInWindow(maxTupleInWindow), numDistinctKeys(numDistinctKeys ? numDistinctKeys : 1),
      batchSize(batchSize) {
    for (auto &s : side) {
        s.tuples.reserve(maxTupleInWindow);
        s.next.reserve(maxTupleInWindow);
        s.head.assign(this->numDistinctKeys, UINT64_MAX);
    }
}

void CPU_HJ::clearStates() {
    for (auto &s : side) {
        s.tuples.clear();
        s.next.clear();
        std::fill(s.head.begin(), s.head.end(), UINT64_MAX);
    }
}

uint64_t CPU_HJ::execute(Tuple *tupleBuffer, bool isLeftSide, uint64_t, Sink &sink) {
    Side &own = side[isLeftSide ? 0 : 1];
    Side &other = side[isLeftSide ? 1 : 0];
    for (uint64_t i = 0; i < batchSize; i++) {
        const uint64_t h = tupleBuffer[i].key % numDistinctKeys;
        own.next.push_back(own.head[h]);
        own.head[h] = own.tuples.size();
        own.tuples.push_back(tupleBuffer[i]);
    }
    uint64_t count = 0, c1 = 0;
    for (uint64_t i = 0; i < batchSize; i++) {
        const Tuple &t = tupleBuffer[i];
        for (uint64_t g = other.head[t.key % numDistinctKeys]; g != UINT64_MAX; g = other.next[g]) {
            const Tuple &o = other.tuples[g];
            if (o.key != t.key) continue;
            count++;
            const Tuple &l = isLeftSide ? t : o;
            const Tuple &r = isLeftSide ? o : t;
            c1 += sjb_hash_a(l.key, l.val) * sjb_hash_b(r.key, r.val);
        }
    }
    addChecksums(sink, c1, 0, false);
    return count;
}

}  // namespace Executor
}  // namespace Eager
}  // namespace Windowing
