// host_tests — CPU-only checks of the host layer (run by tests/test_host.py).
// Exit status 0 when every check passes.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "experiment/Runner.hpp"
#include "sjb/sink/Sink.h"
#include "sjb/source/Source.hpp"
#include "sjb/utils/DataGenerator.hpp"
#include "sjb/utils/Error.hpp"
#include "sjb/windowing/WindowingFactory.hpp"
#include "sjb/windowing/eager/CountBased.hpp"
#include "sjb/windowing/eager/executor/Executors.hpp"
#include "sjb/windowing/lazy/CountBased.hpp"
#include "sjb/windowing/lazy/executor/Executors.hpp"
#include "yaml_lite.hpp"

static int failures = 0, checks = 0;
#define CHECK(cond)                                                             \
    do {                                                                        \
        checks++;                                                               \
        if (!(cond)) {                                                          \
            failures++;                                                         \
            fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);     \
        }                                                                       \
    } while (0)

static std::vector<Tuple> rnd(uint64_t n, uint32_t range, uint32_t seed) {
    std::mt19937_64 g(seed);
    std::vector<Tuple> v(n);
    for (uint64_t i = 0; i < n; i++) v[i] = Tuple{uint32_t(g() % range), uint32_t(g()), i};
    return v;
}

// brute-force count and C1 of L x R
static std::pair<uint64_t, uint64_t> brute(const Tuple *L, uint64_t nl, const Tuple *R, uint64_t nr) {
    uint64_t c = 0, c1 = 0;
    for (uint64_t i = 0; i < nl; i++)
        for (uint64_t j = 0; j < nr; j++)
            if (L[i].key == R[j].key) {
                c++;
                c1 += sjb_hash_a(L[i].key, L[i].val) * sjb_hash_b(R[j].key, R[j].val);
            }
    return {c, c1};
}

static void test_yaml() {
    auto d = yaml_lite::parse(
        "algTypes: [ \"HJ\", 'SMJ' ] # comment\n"
        "windowSizesMB: [ 1,2, 4 ]\n"
        "measuredMetric: \"Throughput\"  # \"Throughput\" | \"x\"\n"
        "repeat: 5\n\n# full-line comment\n");
    CHECK(d["algTypes"].size() == 2 && d["algTypes"][0] == "HJ" && d["algTypes"][1] == "SMJ");
    CHECK(d["windowSizesMB"].size() == 3 && d["windowSizesMB"][2] == "4");
    CHECK(d["measuredMetric"].size() == 1 && d["measuredMetric"][0] == "Throughput");
    CHECK(d["repeat"][0] == "5");
    bool threw = false;
    try {
        yaml_lite::parse("novalue\n");
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    CHECK(threw);
}

static void test_cpu_executors_agree() {
    using namespace Windowing::Lazy::Executor;
    const uint64_t B = 64, k = 5;
    auto L = rnd(B * k, 50, 1), R = rnd(B * k, 50, 2);
    std::vector<uint64_t> idx;
    for (uint64_t i = 0; i < k; i++) idx.push_back(i * B);
    Sink s1, s2, s3;
    auto ref = brute(L.data(), L.size(), R.data(), R.size());
    CPU_HJ hj(64, B, 3, B * k);
    CPU_SMJ smj(64, B, 2);
    CPU_NLJ nlj(B);
    auto a = hj.execute(L.data(), R.data(), idx, idx, s1);
    auto b = smj.execute(L.data(), R.data(), idx, idx, s2);
    auto c = nlj.execute(L.data(), R.data(), idx, idx, s3);
    CHECK(a.resultCount == ref.first && b.resultCount == ref.first && c.resultCount == ref.first);
    CHECK(a.checksum1 == ref.second && b.checksum1 == ref.second && c.checksum1 == ref.second);
}

static void test_lazy_windows_canonical() {
    using namespace Windowing::Lazy::Executor;
    const uint64_t B = 32, W = 4 * B, n = 5 * W;
    auto L = rnd(n, 40, 3), R = rnd(n, 40, 4);
    Sink sink;
    std::vector<std::shared_ptr<BaseLazyExecutor>> js{std::make_shared<CPU_HJ>(1000, B, 2, W)};
    auto w = std::make_shared<Windowing::Lazy::CountBased>(sink, 2 * W, 8, W, B, js, "lazy_cpu_hj_nm");
    Source src(n, B);
    src.start(L.data(), R.data(), 1, w);
    uint64_t expect = 0;
    for (uint64_t o = 0; o < n; o += W) expect += brute(L.data() + o, W, R.data() + o, W).first;
    CHECK(sink.getMatchCount() == expect);
    CHECK(sink.getTotalExecutedWindows().load() == n / W);
    CHECK(sink.getPossibleMatches().load() == (n / W) * W * W);
}

// Windows must not depend on how concurrent source threads interleave:
// 32 source threads, several pooled executors running windows at once and a
// ring of only two windows per side, five runs; every run must give the
// canonical per-window totals (count and C1 over L[wW,(w+1)W) x R[wW,(w+1)W)).
static void test_lazy_windows_concurrent_sources() {
    using namespace Windowing::Lazy::Executor;
    const uint64_t B = 16, W = 8 * B, nw = 40, n = nw * W;
    auto L = rnd(n, 64, 13), R = rnd(n, 64, 14);
    uint64_t expect = 0, expect_c1 = 0;
    for (uint64_t o = 0; o < n; o += W) {
        auto b = brute(L.data() + o, W, R.data() + o, W);
        expect += b.first;
        expect_c1 += b.second;
    }
    // the same tuples split into batches in push order would give another
    // total: the check below can tell canonical windows from arrival order
    uint64_t shifted = 0;
    for (uint64_t o = B; o + W <= n; o += W) shifted += brute(L.data() + o, W, R.data() + o - B, W).first;
    CHECK(shifted != expect);
    for (int run = 0; run < 5; run++) {
        Sink sink;
        std::vector<std::shared_ptr<BaseLazyExecutor>> js;
        for (int j = 0; j < 3; j++) js.push_back(std::make_shared<CPU_HJ>(1000, B, 1, W));
        auto w = std::make_shared<Windowing::Lazy::CountBased>(sink, 2 * W, 1u << 30, W, B, js, "lazy_cpu_hj_nm");
        Source(n, B).start(L.data(), R.data(), 32, w);
        CHECK(sink.getMatchCount() == expect);
        CHECK(sink.getChecksum1() == expect_c1);
        CHECK(sink.getTotalExecutedWindows().load() == nw);
    }
}

// Eager windows with concurrent sources: batches are joined in stream order
// (L0,R0,L1,R1,...), so the totals equal the single-threaded run.
static void test_eager_concurrent_sources() {
    using namespace Windowing::Eager;
    const uint64_t B = 16, W = 4 * B, n = 24 * W;
    auto L = rnd(n, 30, 15), R = rnd(n, 30, 16);
    uint64_t expect = 0;
    for (uint64_t o = 0; o < n; o += W) expect += brute(L.data() + o, W, R.data() + o, W).first;
    for (int run = 0; run < 3; run++) {
        Sink sink;
        auto ex = std::make_shared<Executor::CPU_HJ>(W, 1000, B);
        auto w = std::make_shared<CountBased>(sink, W, B, ex, "eager_cpu_hj_nm", CANONICAL);
        Source(n, B).start(L.data(), R.data(), 16, w);
        CHECK(sink.getMatchCount() == expect);
    }
}

// Reference callers use the Sink's public members directly
// (src/experiment/Runner.cpp:44 resets counterGPU[0]; the Atomic executors
// add to it and write into sinkBuffer).
static void test_sink_reference_members() {
    Sink sink(8);
    sink.counterGPU[0] = 0;
    sink.counterGPU[0] += 5;
    CHECK(sink.getMatchCount(true) == 5 && sink.getMatchCount(false) == 0);
    ResultTuple *ring = sink.sinkBuffer;
    CHECK(ring != nullptr && sink.getSinkBufferSize() == 8 && sink.sinkBufferSize == 8);
    ring[7].lkey = 42;
    CHECK(sink.sinkBuffer[7].lkey == 42);
    ResultTuple rows[3] = {};
    rows[0].lkey = 1;
    rows[2].rVal = 9;
    sink.incrementCounterAndStore(rows, 3);
    CHECK(sink.getMatchCount() == 3 && sink.sinkBuffer[0].lkey == 1 && sink.sinkBuffer[2].rVal == 9);
    sink.incrementCounterAndStore(11, 12, 13, 14, 15, 16);  // position 3 of the ring
    CHECK(sink.getMatchCount() == 4 && sink.sinkBuffer[3].lkey == 11 && sink.sinkBuffer[3].rTs == 16);
    ExecutionStatistic es;
    es.checksum1 = 7;
    es.checksum2 = 100;
    es.checksum2Valid = false;
    sink.addToStatistics(es);
    es.checksum2Valid = true;
    sink.addToStatistics(es);
    CHECK(sink.getChecksum1() == 14 && sink.getChecksum2() == 100);
}

static void test_eager_close_rules() {
    using namespace Windowing::Eager;
    const uint64_t B = 16, W = 3 * B, n = 4 * W;
    auto L = rnd(n, 30, 5), R = rnd(n, 30, 6);
    // canonical: window w = L[wW,(w+1)W) x R[wW,(w+1)W)
    {
        Sink sink;
        auto ex = std::make_shared<Executor::CPU_HJ>(W, 1000, B);
        auto w = std::make_shared<CountBased>(sink, W, B, ex, "eager_cpu_hj_nm", CANONICAL);
        Source(n, B).start(L.data(), R.data(), 1, w);
        uint64_t expect = 0;
        for (uint64_t o = 0; o < n; o += W) expect += brute(L.data() + o, W, R.data() + o, W).first;
        CHECK(sink.getMatchCount() == expect);
    }
    // reference rule (IW/eager/CountBased.cuh:39-45): a batch first clears
    // the state if EITHER side already holds >= W tuples.  Replay the rule
    // to assign every batch to a window, then brute-force each window.
    {
        Sink sink;
        auto ex = std::make_shared<Executor::CPU_HJ>(W, 1000, B);
        auto w = std::make_shared<CountBased>(sink, W, B, ex, "eager_cpu_hj_nm", REFERENCE);
        Source(n, B).start(L.data(), R.data(), 1, w);
        std::vector<std::vector<uint64_t>> wl(1), wr(1);  // batch ids per window
        uint64_t tl = 0, tr = 0;
        for (uint64_t b = 0; b < n / B; b++) {
            for (int side = 0; side < 2; side++) {
                if (tl >= W || tr >= W) {
                    wl.emplace_back();
                    wr.emplace_back();
                    tl = tr = 0;
                }
                (side == 0 ? wl : wr).back().push_back(b);
                (side == 0 ? tl : tr) += B;
            }
        }
        uint64_t expect = 0;
        for (size_t x = 0; x < wl.size(); x++)
            for (uint64_t bl : wl[x])
                for (uint64_t br : wr[x]) expect += brute(L.data() + bl * B, B, R.data() + br * B, B).first;
        CHECK(sink.getMatchCount() == expect);
        CHECK(wl.size() > n / W);  // the rule opens more, shorter windows
    }
}

static void test_factory() {
    using namespace Windowing;
    Sink sink;
    AlgorithmConfig a{HJ, CPU, LAZY, NoOutput};
    DataConfig d{1 << 12, 0.0f, DataGenerator::STEADY, 64, 100, 1000};
    QueryConfig q{256};
    EngineConfig e{1024, 8, 2, 1, 256};
    CHECK(WindowingFactory::create(a, d, q, e, sink)->getAlgorithmName() == "lazy_cpu_hj_nm");
    a.underlyingAlgorithm = SMJ;
    CHECK(WindowingFactory::create(a, d, q, e, sink)->getAlgorithmName() == "lazy_cpu_smj");
    a.underlyingAlgorithm = NLJ;
    CHECK(WindowingFactory::create(a, d, q, e, sink)->getAlgorithmName() == "lazy_cpu_nlj_nm");
    a.progressiveness = EAGER;
    a.underlyingAlgorithm = HJ;
    CHECK(WindowingFactory::create(a, d, q, e, sink)->getAlgorithmName() == "eager_cpu_hj_nm");
    a.underlyingAlgorithm = SMJ;
    bool threw = false;
    try {
        WindowingFactory::create(a, d, q, e, sink);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    CHECK(threw);
    // GPU executors need a device: without one the C ABI reports NODEVICE
    if (!getenv("SJB_HAVE_GPU")) {
        a = AlgorithmConfig{HJ, GPU, LAZY, NoOutput};
        int code = 0;
        try {
            WindowingFactory::create(a, d, q, e, sink);
        } catch (const SjbError &err) {
            code = err.code;
        }
        CHECK(code == SJB_E_NODEVICE);
    }
}

static void test_misc() {
    Sink s;
    CHECK(s.getAverageBuildTime() == 0);  // no windows: no division by zero
    std::vector<Tuple> t(5000);
    DataGenerator::generate(t.data(), t.size(), 0.0f, DataGenerator::STEADY, 100, 7, 1000);
    CHECK(t[0].ts == 0 && t[4999].ts == 999 && t[5].ts == 1);
    bool sorted = true;
    for (size_t i = 1; i < t.size(); i++) sorted &= t[i].ts >= t[i - 1].ts && t[i].key < 100;
    CHECK(sorted);
    DataGenerator::generate(t.data(), t.size(), 0.9f, DataGenerator::BURST, 100, 7, 1000);
    sorted = true;
    for (size_t i = 1; i < t.size(); i++) sorted &= t[i].ts >= t[i - 1].ts && t[i].ts < 1000;
    CHECK(sorted);
    CHECK(Runner::computeMedian({5, 1, 3}) == 3 && Runner::computeMedian({4, 1, 3, 2}) == 2.5);
    CHECK(Runner::csvHeader(false) ==
          "algorithm_name,write_result,ingestion_rate,batch_size,zipf_theta,distinct_key,window_size,"
          "probe_threads,joiner_threads,selectivity,time\n");
}

int main() {
    test_yaml();
    test_cpu_executors_agree();
    test_lazy_windows_canonical();
    test_lazy_windows_concurrent_sources();
    test_eager_concurrent_sources();
    test_sink_reference_members();
    test_eager_close_rules();
    test_factory();
    test_misc();
    printf("host_tests: %d checks, %d failures\n", checks, failures);
    return failures ? 1 : 0;
}
