// Sink — result channel shared by the executors.  Same public surface as
// the reference's Sink (include/sjb/sink/Sink.h:23-126): match counter,
// possible matches, per-window progress, phase-time totals and the public
// members sinkBuffer (a ring of sinkBufferSize result tuples), counterGPU
// (one counter that the Atomic executors add to; callers reset it with
// counterGPU[0] = 0, as src/experiment/Runner.cpp:44 does).  Both are
// page-locked host memory when a device is present (else ordinary memory,
// which the C ABI stages device output into).
// addToStatistics (a no-op in the reference, src/sjb/sink/Sink.cpp:65-67)
// accumulates the per-window checksums that ExecutionStatistic carries.
#pragma once
#include <atomic>
#include <chrono>
#include <cstdint>
#include <mutex>
#include <vector>

#include "sjb/utils/ExecutionStatistic.hpp"
#include "sjb/utils/ResultTuple.hpp"

struct Progress {
    uint64_t currentMatches;
    bool isWindowClosed;
    uint64_t timestamp;
};

class Sink {
public:
    explicit Sink(uint64_t sinkBufferSize = 1ull << 20);
    virtual ~Sink();
    Sink(const Sink &) = delete;
    Sink &operator=(const Sink &) = delete;

    void addToCounter(uint64_t currentMatches, bool isWindowClosed);
    void addToPossibleMatches(uint64_t count);
    uint64_t getMatchCount(bool useGPUCounter = false);
    const std::atomic<uint64_t> &getPossibleMatches() const;
    void addToBuildTime(uint64_t time);
    void addToProbeTime(uint64_t time);
    void addToSortTime(uint64_t time);
    void addToMergeTime(uint64_t time);
    void addToExecutedWindows(uint64_t n);
    // adds the statistic's checksums to the run's totals (getChecksum1/2)
    void addToStatistics(ExecutionStatistic statistic);
    uint64_t getSinkBufferSize() const;
    uint64_t getAverageSortTime();
    uint64_t getAverageMergeTime();
    uint64_t getAverageBuildTime();
    uint64_t getAverageProbeTime();
    std::vector<Progress> getProgress();
    void markProcessingStart();
    // store one result tuple (bitmap decode path of the reference)
    void incrementCounterAndStore(uint32_t lkey, uint32_t rkey, uint32_t lVal, uint32_t rVal,
                                  uint64_t lTs, uint64_t rTs);
    // count `tupleCount` results and copy the first min(tupleCount,
    // sinkBufferSize) of them to the start of sinkBuffer
    void incrementCounterAndStore(ResultTuple *resultTuple, uint64_t tupleCount);
    const std::atomic<uint64_t> &getTotalExecutedWindows() const;
    // run totals of the checksums handed in through addToStatistics
    uint64_t getChecksum1() const { return checksum1.load(); }
    uint64_t getChecksum2() const { return checksum2.load(); }

    // Sink ring buffer of results (include/sjb/sink/Sink.h:92-95)
    ResultTuple *sinkBuffer = nullptr;
    uint64_t sinkBufferSize;
    unsigned long long *counterGPU = nullptr;

private:
    std::atomic<uint64_t> counter{0};
    std::atomic<uint64_t> possibleMatches{0};
    std::atomic<uint64_t> totalBuildTime{0}, totalProbeTime{0}, totalSortTime{0}, totalMergeTime{0};
    std::atomic<uint64_t> totalExecutedWindows{0};
    std::atomic<uint64_t> checksum1{0}, checksum2{0};
    bool bufferPinned = false, counterPinned = false;
    std::chrono::time_point<std::chrono::high_resolution_clock> processingStartTime;
    std::mutex progressMutex;
    std::vector<Progress> progress;
};
