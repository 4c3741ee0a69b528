// Eager executors with the reference's class names and constructor
// arguments (factory: src/sjb/windowing/WindowingFactory.cpp:221-296).
#pragma once
#include <cstdint>
#include <vector>

#include "sjb.h"
#include "sjb/windowing/eager/executor/BaseEagerExecutor.hpp"

namespace Windowing {
namespace Eager {
namespace Executor {

class GPU_Eager : public BaseEagerExecutor {
public:
    GPU_Eager(int algorithm, int method, uint64_t numDistinctKeys, uint64_t windowSize, uint64_t batchSize,
              uint32_t flags = SJB_F_C2);
    ~GPU_Eager() override;
    uint64_t execute(Tuple *tupleBuffer, bool isLeftSide, uint64_t windowTupleCount, Sink &sink) override;
    void clearStates() override;
    void onEndOfStream() override;
    sjb_ctx_t *context() { return ctx; }

protected:
    int algorithm, method;
    uint64_t numDistinctKeys, windowSize, batchSize;
    uint32_t flags;
    sjb_ctx_t *ctx = nullptr;
    std::vector<ResultTuple> scratch;
};

// src/sjb/windowing/eager/executor/GPU/GPU_HJ.cu, GPU_HJ_Atomic.cu
class GPU_HJ : public GPU_Eager {
public:
    GPU_HJ(uint64_t numDistinctKeys, uint64_t windowSize, uint64_t batchSize)
        : GPU_Eager(SJB_ALG_HJ, SJB_OUT_NOOUTPUT, numDistinctKeys, windowSize, batchSize) {}
};
class GPU_HJ_Atomic : public GPU_Eager {
public:
    GPU_HJ_Atomic(uint64_t numDistinctKeys, uint64_t windowSize, uint64_t batchSize)
        : GPU_Eager(SJB_ALG_HJ, SJB_OUT_ATOMIC, numDistinctKeys, windowSize, batchSize) {}
};
// src/sjb/windowing/eager/executor/GPU/GPU_NLJ.cu, GPU_NLJ_BM.cu
class GPU_NLJ : public GPU_Eager {
public:
    GPU_NLJ(uint64_t batchSize, uint64_t windowSize)
        : GPU_Eager(SJB_ALG_NLJ, SJB_OUT_NOOUTPUT, 1000, windowSize, batchSize) {}
};
class GPU_NLJ_BM : public GPU_Eager {
public:
    GPU_NLJ_BM(uint64_t batchSize, uint64_t windowSize)
        : GPU_Eager(SJB_ALG_NLJ, SJB_OUT_BITMAP, 1000, windowSize, batchSize) {}
};

// Persistent-kernel eager hash join (the design of
// src/sjb/windowing/eager/executor/GPU/GPU_HJ_PK.cu, GPU_HJ_PK.cuh:19-40):
// batches go through sjb_pk_execute to one resident workgroup.
class GPU_HJ_PK : public BaseEagerExecutor {
public:
    GPU_HJ_PK(uint64_t numDistinctKeys, uint64_t windowSize, uint64_t batchSize);
    ~GPU_HJ_PK() override;
    uint64_t execute(Tuple *tupleBuffer, bool isLeftSide, uint64_t windowTupleCount, Sink &sink) override;
    void clearStates() override;
    void onEndOfStream() override;

private:
    uint64_t batchSize;
    sjb_ctx_t *ctx = nullptr;
};

// src/sjb/windowing/eager/executor/CPU/CPU_HJ.cpp: build own side, probe the other.
class CPU_HJ : public BaseEagerExecutor {
public:
    CPU_HJ(uint64_t maxTupleInWindow, uint64_t numDistinctKeys, uint64_t batchSize);
    uint64_t execute(Tuple *tupleBuffer, bool isLeftSide, uint64_t windowTupleCount, Sink &sink) override;
    void clearStates() override;
    void onEndOfStream() override {}

private:
    struct Side {
        std::vector<Tuple> tuples;
        std::vector<uint64_t> next, head;
    };
    Side side[2];
    uint64_t maxTupleInWindow, numDistinctKeys, batchSize;
};

}  // namespace Executor
}  // namespace Eager
}  // namespace Windowing
