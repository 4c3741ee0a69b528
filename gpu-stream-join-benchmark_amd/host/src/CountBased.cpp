// Lazy and eager count-based windowing.
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <thread>

#include "sjb/utils/Pinned.hpp"
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
    nSlots = this->ringBufferSize / batchSize;
    const uint64_t bytes = nSlots * batchSize * sizeof(Tuple);
    leftRingBuffer = static_cast<Tuple *>(sjb_host::pinnedAlloc(bytes, &leftPinned));
    try {
        rightRingBuffer = static_cast<Tuple *>(sjb_host::pinnedAlloc(bytes, &rightPinned));
    } catch (...) {
        sjb_host::pinnedFree(leftRingBuffer, leftPinned);
        throw;
    }
    for (auto &sd : sides) sd.released.assign(nSlots, 0);
    for (auto &j : js) joiners.push_back(j);
}

CountBased::~CountBased() {
    sjb_host::pinnedFree(leftRingBuffer, leftPinned);
    sjb_host::pinnedFree(rightRingBuffer, rightPinned);
}

void CountBased::onIncoming(Tuple *tupleBuffer, bool left, uint64_t i) {
    Side &sd = sides[left ? 0 : 1];
    const uint64_t slot = i % nSlots, gen = i / nSlots;
    {
        // wait until the slot's previous batch (index i - nSlots) is joined
        std::unique_lock<std::mutex> g(mtx);
        for (uint64_t trial = 0; sd.released[slot] != gen; trial++) {
            if (trial > ringBufferMaxTries) throw std::runtime_error("Ring buffer is full");
            // a system_clock deadline (pthread_cond_timedwait), not wait_for
            // (pthread_cond_clockwait, which ThreadSanitizer builds of this
            // toolchain do not intercept)
            slotCv.wait_until(g, std::chrono::system_clock::now() + std::chrono::microseconds(1));
        }
    }
    std::memcpy((left ? leftRingBuffer : rightRingBuffer) + slot * batchSize, tupleBuffer,
                batchSize * sizeof(Tuple));
    {
        std::lock_guard<std::mutex> g(mtx);
        sd.ready[i / (windowSize / batchSize)]++;
    }
    execute();
}

void CountBased::onIncomingLeft(Tuple *tupleBuffer) {
    uint64_t i;
    {
        std::lock_guard<std::mutex> g(mtx);
        i = sides[0].arrivals++;
    }
    onIncoming(tupleBuffer, true, i);
}

void CountBased::onIncomingRight(Tuple *tupleBuffer) {
    uint64_t i;
    {
        std::lock_guard<std::mutex> g(mtx);
        i = sides[1].arrivals++;
    }
    onIncoming(tupleBuffer, false, i);
}

void CountBased::onIncomingLeft(Tuple *tupleBuffer, uint64_t batchIndex) { onIncoming(tupleBuffer, true, batchIndex); }
void CountBased::onIncomingRight(Tuple *tupleBuffer, uint64_t batchIndex) {
    onIncoming(tupleBuffer, false, batchIndex);
}

// CountBased.cuh:103-166, in window order
void CountBased::execute() {
    const uint64_t k = windowSize / batchSize;
    for (;;) {
        uint64_t w;
        {
            std::lock_guard<std::mutex> g(mtx);
            auto l = sides[0].ready.find(nextWindow), r = sides[1].ready.find(nextWindow);
            if (l == sides[0].ready.end() || r == sides[1].ready.end() || l->second < k || r->second < k) return;
            sides[0].ready.erase(l);
            sides[1].ready.erase(r);
            w = nextWindow++;
        }
        std::vector<uint64_t> li(k), ri(k);
        for (uint64_t j = 0; j < k; j++) li[j] = ri[j] = ((w * k + j) % nSlots) * batchSize;
        std::shared_ptr<Executor::BaseLazyExecutor> joiner;
        {
            std::unique_lock<std::mutex> g(joinerMtx);
            joinerCv.wait(g, [&] { return !joiners.empty(); });
            joiner = joiners.front();
            joiners.pop_front();
        }
        ExecutionStatistic es;
        try {
            es = joiner->execute(leftRingBuffer, rightRingBuffer, li, ri, sink);
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
            for (uint64_t j = 0; j < k; j++) {
                const uint64_t slot = (w * k + j) % nSlots;
                sides[0].released[slot]++;
                sides[1].released[slot]++;
            }
        }
        slotCv.notify_all();
        {
            std::lock_guard<std::mutex> g(joinerMtx);
            joiners.push_back(joiner);
        }
        joinerCv.notify_one();
    }
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

void CountBased::inTurn(Tuple *tupleBuffer, bool left, uint64_t turn) {
    {
        std::unique_lock<std::mutex> g(turnMtx);
        turnCv.wait(g, [&] { return nextTurn == turn || turnFailed; });
        if (turnFailed) throw std::runtime_error("an earlier batch of the stream failed");
    }
    try {
        onIncoming(tupleBuffer, left);
    } catch (...) {
        // the batches behind this one would wait for a turn that never
        // comes: release them all (each throws)
        {
            std::lock_guard<std::mutex> g(turnMtx);
            turnFailed = true;
        }
        turnCv.notify_all();
        throw;
    }
    {
        std::lock_guard<std::mutex> g(turnMtx);
        nextTurn++;
    }
    turnCv.notify_all();
}

}  // namespace Eager
}  // namespace Windowing
