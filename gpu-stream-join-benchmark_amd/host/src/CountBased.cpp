// Lazy and eager count-based windowing.
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <thread>

#include "sjb/windowing/eager/CountBased.hpp"
#include "sjb/windowing/lazy/CountBased.hpp"

namespace Windowing {
namespace Lazy {

CountBased::CountBased(Sink &sink, uint64_t ringBufferSize, uint64_t ringBufferMaxTries, uint64_t windowSize,
                       uint64_t batchSize, std::vector<std::shared_ptr<Executor::BaseLazyExecutor>> js,
                       std::string algorithmName)
    : BaseWindowing(sink, std::move(algorithmName)), ringBufferSize(ringBufferSize),
      ringBufferMaxTries(ringBufferMaxTries), windowSize(windowSize), batchSize(batchSize) {
    if (batchSize == 0 || windowSize < batchSize) throw std::invalid_argument("window must hold >= 1 batch");
    // The ring must hold at least one window per side plus a batch in flight.
    if (this->ringBufferSize < windowSize + batchSize) this->ringBufferSize = windowSize + batchSize;
    leftRingBuffer.resize(this->ringBufferSize);
    rightRingBuffer.resize(this->ringBufferSize);
    for (uint64_t i = 0; i + batchSize <= this->ringBufferSize; i += batchSize) {
        availableLeft.push_back(i);
        availableRight.push_back(i);
    }
    for (auto &j : js) joiners.push_back(j);
}

uint64_t CountBased::popFree(std::deque<uint64_t> &q) {
    for (uint64_t trial = 0;; trial++) {
        {
            std::lock_guard<std::mutex> g(mtx);
            if (!q.empty()) {
                const uint64_t v = q.front();
                q.pop_front();
                return v;
            }
        }
        if (trial > ringBufferMaxTries) throw std::runtime_error("Ring buffer is full");
        std::this_thread::sleep_for(std::chrono::microseconds(1));
    }
}

void CountBased::onIncoming(Tuple *tupleBuffer, bool left) {
    const uint64_t slot = popFree(left ? availableLeft : availableRight);
    std::memcpy((left ? leftRingBuffer : rightRingBuffer).data() + slot, tupleBuffer, batchSize * sizeof(Tuple));
    {
        std::lock_guard<std::mutex> g(mtx);
        (left ? executableLeft : executableRight).push_back(slot);
    }
    execute();
}

void CountBased::onIncomingLeft(Tuple *tupleBuffer) { onIncoming(tupleBuffer, true); }
void CountBased::onIncomingRight(Tuple *tupleBuffer) { onIncoming(tupleBuffer, false); }

// CountBased.cuh:103-166
void CountBased::execute() {
    const uint64_t k = windowSize / batchSize;
    std::vector<uint64_t> li, ri;
    {
        std::lock_guard<std::mutex> g(mtx);
        if (executableLeft.size() >= k && executableRight.size() >= k) {
            for (uint64_t i = 0; i < k; i++) {
                li.push_back(executableLeft.front());
                executableLeft.pop_front();
                ri.push_back(executableRight.front());
                executableRight.pop_front();
            }
        }
    }
    if (li.empty()) return;
    std::shared_ptr<Executor::BaseLazyExecutor> joiner;
    {
        std::unique_lock<std::mutex> g(joinerMtx);
        joinerCv.wait(g, [&] { return !joiners.empty(); });
        joiner = joiners.front();
        joiners.pop_front();
    }
    ExecutionStatistic es;
    try {
        es = joiner->execute(leftRingBuffer.data(), rightRingBuffer.data(), li, ri, sink);
    } catch (...) {
        std::lock_guard<std::mutex> g(joinerMtx);
        joiners.push_back(joiner);
        joinerCv.notify_one();
        throw;
    }
    sink.addToPossibleMatches(batchSize * li.size() * batchSize * ri.size());
    sink.addToCounter(es.resultCount, true);
    sink.addToBuildTime(es.buildTime);
    sink.addToProbeTime(es.probeTime);
    sink.addToSortTime(es.sortTime);
    sink.addToMergeTime(es.mergeTime);
    sink.addToExecutedWindows(1);
    {
        std::lock_guard<std::mutex> g(mtx);
        for (uint64_t i = 0; i < k; i++) {
            availableLeft.push_back(li[i]);
            availableRight.push_back(ri[i]);
        }
    }
    {
        std::lock_guard<std::mutex> g(joinerMtx);
        joiners.push_back(joiner);
    }
    joinerCv.notify_one();
}

}  // namespace Lazy

namespace Eager {

CountBased::CountBased(Sink &sink, uint64_t windowSize, uint64_t batchSize,
                       std::shared_ptr<Executor::BaseEagerExecutor> joiner, std::string algorithmName,
                       CLOSE_RULE rule)
    : BaseWindowing(sink, std::move(algorithmName)), windowSize(windowSize), batchSize(batchSize),
      joiner(std::move(joiner)), rule(rule) {}

// CountBased.cuh:36-72
void CountBased::onIncoming(Tuple *tupleBuffer, bool left) {
    std::lock_guard<std::mutex> g(mtx);
    bool closed = false;
    const uint64_t own = left ? totalIncomingLeft : totalIncomingRight;
    const bool close = rule == REFERENCE
                           ? (totalIncomingLeft >= windowSize || totalIncomingRight >= windowSize)
                           : ((totalIncomingLeft >= windowSize && totalIncomingRight >= windowSize) ||
                              own + batchSize > windowSize);
    if (close) {
        joiner->clearStates();
        totalIncomingLeft = totalIncomingRight = 0;
        sink.addToExecutedWindows(1);
        closed = true;
    }
    const uint64_t other = left ? totalIncomingRight : totalIncomingLeft;
    const uint64_t matches = joiner->execute(tupleBuffer, left, left ? totalIncomingLeft : totalIncomingRight, sink);
    (left ? totalIncomingLeft : totalIncomingRight) += batchSize;
    sink.addToCounter(matches, closed);
    sink.addToPossibleMatches(batchSize * other);
}

}  // namespace Eager
}  // namespace Windowing
