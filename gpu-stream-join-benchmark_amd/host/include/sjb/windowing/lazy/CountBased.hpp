// Lazy tumbling count window (include/sjb/windowing/lazy/CountBased.cuh:19-196):
// incoming batches are copied into ring-buffer slots; once both sides hold
// the windowSize/batchSize batches of the next window, that window's slots
// are joined by a pooled executor, then released.  Thread-safe (sources may
// push from several threads); a batch that waits longer than
// ringBufferMaxTries for its slot throws instead of exit(1)
// (CountBased.cuh:58-61).
//
// Windows are canonical: batch i of a side lives in slot i mod nSlots and
// window w is batches [w k, (w+1) k) of each side (k = windowSize /
// batchSize), joined in window order.  With the indexed onIncoming* calls
// i is the batch's position in its stream, so window w is always stream
// tuples [w W, (w+1) W) whatever the interleaving of concurrent source
// threads; the index-less calls number batches in arrival order, as the
// reference does (CountBased.cuh:112-125).  A batch takes its slot only
// after the batch nSlots before it has been joined, and that batch belongs
// to an earlier window, so windows always complete in order.
#pragma once
#include <condition_variable>
#include <deque>
#include <map>
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
    void onIncomingLeft(Tuple *tupleBuffer, uint64_t batchIndex) override;
    void onIncomingRight(Tuple *tupleBuffer, uint64_t batchIndex) override;
    void onEndOfStream() override {}
    // Joins every window whose batches have all arrived (in window order).
    void execute();
    ~CountBased() override;
    CountBased(const CountBased &) = delete;
    CountBased &operator=(const CountBased &) = delete;

    // ring buffers of nSlots * batchSize tuples per side, page-locked when a
    // device is present (include/sjb/windowing/lazy/CountBased.cuh:35-36,
    // :178-179), so the GPU executors copy windows from them asynchronously
    Tuple *leftRingBuffer{};
    Tuple *rightRingBuffer{};
    bool leftPinned = false, rightPinned = false;

private:
    struct Side {
        std::vector<uint64_t> released;      // per slot: times it was freed (= generation it serves)
        std::map<uint64_t, uint64_t> ready;  // window -> batches of it written
        uint64_t arrivals = 0;               // next index of an index-less batch
    };
    void onIncoming(Tuple *tupleBuffer, bool left, uint64_t batchIndex);

    std::mutex mtx;
    std::condition_variable slotCv;
    Side sides[2];
    uint64_t nSlots = 0;
    uint64_t nextWindow = 0;  // next window to join
    std::mutex joinerMtx;
    std::condition_variable joinerCv;
    std::deque<std::shared_ptr<Executor::BaseLazyExecutor>> joiners;
    uint64_t ringBufferSize, ringBufferMaxTries, windowSize, batchSize;
};
}  // namespace Lazy
}  // namespace Windowing
