// Lazy tumbling count window (include/sjb/windowing/lazy/CountBased.cuh:19-196):
// incoming batches are copied into free ring-buffer slots; once both sides
// hold windowSize/batchSize executable slots, one window's slots are popped
// and joined by a pooled executor, then returned.  Thread-safe (sources may
// push from several threads); the ring-full condition throws instead of
// exit(1) (CountBased.cuh:58-61).
#pragma once
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "sjb/windowing/BaseWindowing.hpp"
#include "sjb/windowing/lazy/executor/BaseLazyExecutor.hpp"

namespace Windowing {
namespace Lazy {
class CountBased : public BaseWindowing {
public:
    CountBased(Sink &sink, uint64_t ringBufferSize, uint64_t ringBufferMaxTries, uint64_t windowSize,
               uint64_t batchSize, std::vector<std::shared_ptr<Executor::BaseLazyExecutor>> joiners,
               std::string algorithmName);
    void onIncomingLeft(Tuple *tupleBuffer) override;
    void onIncomingRight(Tuple *tupleBuffer) override;
    void onEndOfStream() override {}
    void execute();

    std::vector<Tuple> leftRingBuffer, rightRingBuffer;

private:
    uint64_t popFree(std::deque<uint64_t> &q);
    void onIncoming(Tuple *tupleBuffer, bool left);

    std::mutex mtx;
    std::deque<uint64_t> availableLeft, availableRight, executableLeft, executableRight;
    std::mutex joinerMtx;
    std::condition_variable joinerCv;
    std::deque<std::shared_ptr<Executor::BaseLazyExecutor>> joiners;
    uint64_t ringBufferSize, ringBufferMaxTries, windowSize, batchSize;
};
}  // namespace Lazy
}  // namespace Windowing
