// Eager (per-batch) operator interface: the signature of
// include/sjb/windowing/eager/executor/BaseEagerExecutor.hpp:16-20.
#pragma once
#include <cstdint>

#include "sjb/sink/Sink.h"
#include "sjb/utils/Tuple.hpp"

namespace Windowing {
namespace Eager {
namespace Executor {
class BaseEagerExecutor {
public:
    virtual ~BaseEagerExecutor() = default;
    virtual uint64_t execute(Tuple *tupleBuffer, bool isLeftSide, uint64_t windowTupleCount, Sink &sink) = 0;
    virtual void clearStates() = 0;
    virtual void onEndOfStream() = 0;

private:
    bool endOfStream = false;
};
}  // namespace Executor
}  // namespace Eager
}  // namespace Windowing
