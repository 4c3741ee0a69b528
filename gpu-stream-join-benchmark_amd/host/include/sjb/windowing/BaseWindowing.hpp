// Windowing enums and the abstract windowing; values and names of
// include/sjb/windowing/BaseWindowing.cuh:9-64.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <utility>

#include "sjb/sink/Sink.h"
#include "sjb/utils/Tuple.hpp"

namespace Windowing {
enum ALGORITHM_TYPE : uint8_t { EMPTY = 0, NLJ = 1, HJ = 2, SMJ = 3, SABER = 4, HELLS = 5 };
enum PROGRESSIVENESS_TYPE : uint8_t { EAGER = 0, LAZY = 1 };
enum RESULT_WRITING_METHOD : uint8_t { NoOutput = 0, CountKernel = 1, EstimatedSelectivity = 2, Bitmap = 3, Atomic = 4 };
enum DEVICE_TYPE : uint8_t { CPU = 0, GPU = 1 };
enum PARALLELIZATION_STRATEGY : uint8_t { INTER = 0, INTRA = 1, HYBRID = 2 };

class BaseWindowing {
public:
    BaseWindowing(Sink &sink, std::string algorithmName) : sink(sink), algorithmName(std::move(algorithmName)) {}
    virtual ~BaseWindowing() = default;
    virtual void onIncomingLeft(Tuple *tupleBuffer) = 0;
    virtual void onIncomingRight(Tuple *tupleBuffer) = 0;
    // The same, for batch `batchIndex` of its stream (0, 1, 2, ... in
    // stream order).  With the index a windowing can place the batch by its
    // position in the stream rather than by arrival order, so windows do not
    // depend on how concurrent source threads interleave (the reference
    // windows by arrival order, include/sjb/windowing/lazy/CountBased.cuh:112-125).
    // Callers that do not know the index use the calls above.
    virtual void onIncomingLeft(Tuple *tupleBuffer, uint64_t batchIndex) {
        (void)batchIndex;
        onIncomingLeft(tupleBuffer);
    }
    virtual void onIncomingRight(Tuple *tupleBuffer, uint64_t batchIndex) {
        (void)batchIndex;
        onIncomingRight(tupleBuffer);
    }
    virtual void onEndOfStream() = 0;
    const std::string &getAlgorithmName() const { return algorithmName; }

protected:
    Sink &sink;
    std::string algorithmName;
    bool endOfStream = false;
};
}  // namespace Windowing
