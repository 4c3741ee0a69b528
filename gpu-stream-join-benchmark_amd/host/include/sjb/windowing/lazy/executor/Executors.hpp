// Lazy executors with the reference's class names and constructor
// arguments (factory: src/sjb/windowing/WindowingFactory.cpp:61-219).
//
// GPU executors are adapters over the C ABI (one sjb_ctx each, i.e. one
// HIP stream and one set of device buffers, like the reference's one CUDA
// stream set per executor).  CPU executors restate the reference CPU joins
// in portable C++ (std::thread instead of TBB).
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

#include "sjb.h"
#include "sjb/sink/Sink.h"
#include "sjb/windowing/lazy/executor/BaseLazyExecutor.hpp"

namespace Windowing {
namespace Lazy {
namespace Executor {

// Common GPU adapter: algorithm x result-writing method.
class GPU_Lazy : public BaseLazyExecutor {
public:
    GPU_Lazy(int algorithm, int method, uint64_t numDistinctKeys, uint64_t batchSize,
             uint64_t maxTupleInWindow, double estimatedSelectivity = 0.0, uint32_t flags = SJB_F_C2);
    ~GPU_Lazy() override;
    ExecutionStatistic execute(Tuple *leftRingBuffer, Tuple *rightRingBuffer,
                               std::vector<uint64_t> leftIndexesToJoin,
                               std::vector<uint64_t> rightIndexesToJoin, Sink &sink) override;
    sjb_ctx_t *context() { return ctx; }

protected:
    int algorithm, method;
    uint64_t numDistinctKeys, batchSize, maxTupleInWindow;
    double estimatedSelectivity;
    uint32_t flags;
    sjb_ctx_t *ctx = nullptr;
    std::vector<ResultTuple> scratch;  // exact-size buffer when the sink ring is too small
};

// NoOutput executors report the match count (and C1): grouped mode gives
// both in O(n) whatever the keys' multiplicity (the reference's SMJ count is
// the same closed form, sum over keys of L_k * R_k, GPU_SMJ.cu:34-44); the
// other methods enumerate every pair and add C2.
#define SJB_GPU_LAZY_3(NAME, ALG, METHOD)                                                   \
    class NAME : public GPU_Lazy {                                                          \
    public:                                                                                 \
        NAME(uint64_t numDistinctKeys, uint64_t batchSize, uint64_t maxTupleInWindow)       \
            : GPU_Lazy(ALG, METHOD, numDistinctKeys, batchSize, maxTupleInWindow, 0.0,      \
                       (METHOD) == SJB_OUT_NOOUTPUT ? SJB_F_GROUPED : SJB_F_C2) {}          \
    };

// HJ (src/sjb/windowing/lazy/executor/GPU/GPU_HJ_*.cu)
SJB_GPU_LAZY_3(GPU_HJ_NoOutput, SJB_ALG_HJ, SJB_OUT_NOOUTPUT)
SJB_GPU_LAZY_3(GPU_HJ_CountKernel, SJB_ALG_HJ, SJB_OUT_COUNTKERNEL)
SJB_GPU_LAZY_3(GPU_HJ_Atomic, SJB_ALG_HJ, SJB_OUT_ATOMIC)
class GPU_HJ_EstimatedSelectivity : public GPU_Lazy {
public:
    GPU_HJ_EstimatedSelectivity(double estimatedSelectivity, uint64_t numDistinctKeys, uint64_t batchSize,
                                uint64_t maxTupleInWindow)
        : GPU_Lazy(SJB_ALG_HJ, SJB_OUT_ESTIMATEDSELECTIVITY, numDistinctKeys, batchSize, maxTupleInWindow,
                   estimatedSelectivity) {}
};
// SMJ (src/sjb/windowing/lazy/executor/GPU/GPU_SMJ_*.cu)
SJB_GPU_LAZY_3(GPU_SMJ_NoOutput, SJB_ALG_SMJ, SJB_OUT_NOOUTPUT)
SJB_GPU_LAZY_3(GPU_SMJ_CountKernel, SJB_ALG_SMJ, SJB_OUT_COUNTKERNEL)
SJB_GPU_LAZY_3(GPU_SMJ_Atomic, SJB_ALG_SMJ, SJB_OUT_ATOMIC)
class GPU_SMJ_EstimatedSelectivity : public GPU_Lazy {
public:
    GPU_SMJ_EstimatedSelectivity(uint64_t numDistinctKeys, uint64_t batchSize, uint64_t maxTupleInWindow,
                                 double estimatedSelectivity)
        : GPU_Lazy(SJB_ALG_SMJ, SJB_OUT_ESTIMATEDSELECTIVITY, numDistinctKeys, batchSize, maxTupleInWindow,
                   estimatedSelectivity) {}
};
#undef SJB_GPU_LAZY_3

// NLJ (src/sjb/windowing/lazy/executor/GPU/GPU_NLJ_*.cu)
#define SJB_GPU_NLJ(NAME, METHOD)                                                        \
    class NAME : public GPU_Lazy {                                                       \
    public:                                                                              \
        NAME(uint64_t batchSize, uint64_t maxTupleInWindow)                              \
            : GPU_Lazy(SJB_ALG_NLJ, METHOD, 1000, batchSize, maxTupleInWindow) {}         \
    };
SJB_GPU_NLJ(GPU_NLJ_NoOutput, SJB_OUT_NOOUTPUT)
SJB_GPU_NLJ(GPU_NLJ_CountKernel, SJB_OUT_COUNTKERNEL)
SJB_GPU_NLJ(GPU_NLJ_Bitmap, SJB_OUT_BITMAP)
#undef SJB_GPU_NLJ
class GPU_NLJ_EstimatedSelectivity : public GPU_Lazy {
public:
    GPU_NLJ_EstimatedSelectivity(uint64_t batchSize, uint64_t maxTupleInWindow, double estimatedSelectivity)
        : GPU_Lazy(SJB_ALG_NLJ, SJB_OUT_ESTIMATEDSELECTIVITY, 1000, batchSize, maxTupleInWindow,
                   estimatedSelectivity) {}
};

// CPU (src/sjb/windowing/lazy/executor/CPU/CPU_{HJ,SMJ,NLJ}.cpp)
class CPU_HJ : public BaseLazyExecutor {
public:
    CPU_HJ(uint64_t distinctKeys, uint64_t batchSize, uint64_t nProbeThreads, uint64_t maxTupleInWindow);
    ExecutionStatistic execute(Tuple *leftRingBuffer, Tuple *rightRingBuffer,
                               std::vector<uint64_t> leftIndexesToJoin,
                               std::vector<uint64_t> rightIndexesToJoin, Sink &sink) override;

private:
    uint64_t distinctKeys, batchSize, nProbeThreads, maxTupleInWindow;
    std::vector<Tuple> htStore;
};

class CPU_SMJ : public BaseLazyExecutor {
public:
    CPU_SMJ(uint64_t numDistinctKey, uint64_t batchSize, uint64_t nSortThreads);
    ExecutionStatistic execute(Tuple *leftRingBuffer, Tuple *rightRingBuffer,
                               std::vector<uint64_t> leftIndexesToJoin,
                               std::vector<uint64_t> rightIndexesToJoin, Sink &sink) override;

private:
    uint64_t numDistinctKey, batchSize, nSortThreads;
};

class CPU_NLJ : public BaseLazyExecutor {
public:
    explicit CPU_NLJ(uint64_t batchSize);
    ExecutionStatistic execute(Tuple *leftRingBuffer, Tuple *rightRingBuffer,
                               std::vector<uint64_t> leftIndexesToJoin,
                               std::vector<uint64_t> rightIndexesToJoin, Sink &sink) override;

private:
    uint64_t batchSize;
};

}  // namespace Executor
}  // namespace Lazy
}  // namespace Windowing
