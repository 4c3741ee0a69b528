// Source (include/sjb/source/Source.hpp, src/sjb/source/Source.cpp:12-34):
// pushes batch i of the left stream then batch i of the right stream to the
// windowing, for i over count / bufferSize batches.  sourceThreads == 1 keeps
// the strict L0,R0,L1,R1,... order; more threads pull batch indexes from a
// shared counter like the reference's parallel_for.  Every batch is pushed
// with its index (BaseWindowing::onIncomingLeft(tuples, i)), so the
// windowing places it by stream position and windows are the same for any
// thread count.
#pragma once
#include <cstdint>
#include <memory>

#include "sjb/utils/Tuple.hpp"
#include "sjb/windowing/BaseWindowing.hpp"

class Source {
public:
    enum SIDE { LEFT = 0, RIGHT = 1 };
    Source(uint64_t count, uint64_t bufferSize);
    void start(Tuple *leftTuples, Tuple *rightTuples, uint64_t sourceThreads,
               const std::shared_ptr<Windowing::BaseWindowing> &windowing);

private:
    uint64_t count, bufferSize;
};
