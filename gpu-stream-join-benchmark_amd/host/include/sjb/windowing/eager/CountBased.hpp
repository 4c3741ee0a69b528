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
//
// Batch order.  The index-less calls join batches in arrival order (as the
// reference does).  The indexed calls join batch i of the left stream as
// the (2i)-th batch and batch i of the right stream as the (2i+1)-th,
// whatever order concurrent source threads deliver them in: a batch waits
// until every batch before it in that order has been joined.  One source
// should not mix the two kinds of call.
#pragma once
#include <atomic>
#include <condition_variable>
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
    void onIncomingLeft(Tuple *tupleBuffer, uint64_t batchIndex) override {
        inTurn(tupleBuffer, true, 2 * batchIndex);
    }
    void onIncomingRight(Tuple *tupleBuffer, uint64_t batchIndex) override {
        inTurn(tupleBuffer, false, 2 * batchIndex + 1);
    }
    void onEndOfStream() override { joiner->onEndOfStream(); }

private:
    void onIncoming(Tuple *tupleBuffer, bool left);
    void inTurn(Tuple *tupleBuffer, bool left, uint64_t turn);
    std::mutex turnMtx;
    std::condition_variable turnCv;
    uint64_t nextTurn = 0;  // position (in L0,R0,L1,R1,... order) of the next batch to join
    bool turnFailed = false;  // a batch threw: the ones waiting behind it give up
    std::mutex mtx;  // one batch at a time: the executor state has a single writer
    uint64_t totalIncomingLeft = 0, totalIncomingRight = 0;
    uint64_t windowSize, batchSize;
    std::shared_ptr<Executor::BaseEagerExecutor> joiner;
    CLOSE_RULE rule;
};
}  // namespace Eager
}  // namespace Windowing
