// Eager count window (include/sjb/windowing/eager/CountBased.cuh:36-76):
// every batch is joined on arrival against the opposite side's window state.
//
// Window-close rule.  The reference clears the state when EITHER side
// already holds >= windowSize tuples at the start of a batch
// (CountBased.cuh:39-45), which with the L,R,L,R source order shifts every
// right window by one batch.  CANONICAL (default) clears when BOTH sides are
// full, so window w is L[wW, (w+1)W) x R[wW, (w+1)W) like the lazy windows;
// REFERENCE keeps the reference rule.  Under CANONICAL a side that is full
// while the other is not also closes the window (its state cannot grow).
#pragma once
#include <atomic>
#include <memory>
#include <mutex>
#include <string>

#include "sjb/windowing/BaseWindowing.hpp"
#include "sjb/windowing/eager/executor/BaseEagerExecutor.hpp"

namespace Windowing {
namespace Eager {
enum CLOSE_RULE : uint8_t { CANONICAL = 0, REFERENCE = 1 };

class CountBased : public BaseWindowing {
public:
    CountBased(Sink &sink, uint64_t windowSize, uint64_t batchSize,
               std::shared_ptr<Executor::BaseEagerExecutor> joiner, std::string algorithmName,
               CLOSE_RULE rule = CANONICAL);
    void onIncomingLeft(Tuple *tupleBuffer) override { onIncoming(tupleBuffer, true); }
    void onIncomingRight(Tuple *tupleBuffer) override { onIncoming(tupleBuffer, false); }
    void onEndOfStream() override { joiner->onEndOfStream(); }

private:
    void onIncoming(Tuple *tupleBuffer, bool left);
    std::mutex mtx;  // one batch at a time: the executor state has a single writer
    uint64_t totalIncomingLeft = 0, totalIncomingRight = 0;
    uint64_t windowSize, batchSize;
    std::shared_ptr<Executor::BaseEagerExecutor> joiner;
    CLOSE_RULE rule;
};
}  // namespace Eager
}  // namespace Windowing
