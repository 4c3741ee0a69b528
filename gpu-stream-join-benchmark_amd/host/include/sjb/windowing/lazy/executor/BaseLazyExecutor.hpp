// Lazy (per-window) operator interface: the signature of
// include/sjb/windowing/lazy/executor/BaseLazyExecutor.hpp:26-30.
#pragma once
#include <cstdint>
#include <vector>

#include "sjb/utils/ExecutionStatistic.hpp"
#include "sjb/utils/Tuple.hpp"

class Sink;

namespace Windowing {
namespace Lazy {
namespace Executor {
class BaseLazyExecutor {
public:
    explicit BaseLazyExecutor() {}
    virtual ~BaseLazyExecutor() = default;
    /**
     * Joins one window.  leftIndexesToJoin / rightIndexesToJoin are tuple
     * offsets of batch-sized slots in the ring buffers.
     */
    virtual ExecutionStatistic execute(Tuple *leftRingBuffer, Tuple *rightRingBuffer,
                                       std::vector<uint64_t> leftIndexesToJoin,
                                       std::vector<uint64_t> rightIndexesToJoin, Sink &sink) = 0;
};
}  // namespace Executor
}  // namespace Lazy
}  // namespace Windowing
