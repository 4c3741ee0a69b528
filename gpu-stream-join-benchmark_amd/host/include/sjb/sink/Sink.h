// Sink — result channel shared by the executors.  Same public surface as
// the reference's Sink (include/sjb/sink/Sink.h:23-126): match counter,
// possible matches, per-window progress, phase-time totals, a result-tuple
// ring.  The ring lives in host memory (the C ABI stages device output);
// `counterGPU` is an ordinary counter that the Atomic executors add to.
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
    ~Sink();

    void addToCounter(uint64_t currentMatches, bool isWindowClosed);
    void addToPossibleMatches(uint64_t count);
    uint64_t getMatchCount(bool useGPUCounter = false);
    const std::atomic<uint64_t> &getPossibleMatches() const;
    void addToBuildTime(uint64_t time);
    void addToProbeTime(uint64_t time);
    void addToSortTime(uint64_t time);
    void addToMergeTime(uint64_t time);
    void addToExecutedWindows(uint64_t n);
    void addToChecksums(uint64_t c1, uint64_t c2);
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
    // account `tupleCount` results already written into sinkBuffer
    void incrementCounterAndStore(ResultTuple *resultTuple, uint64_t tupleCount);
    const std::atomic<uint64_t> &getTotalExecutedWindows() const;
    uint64_t getChecksum1() const { return checksum1.load(); }
    uint64_t getChecksum2() const { return checksum2.load(); }

    std::vector<ResultTuple> sinkBuffer;
    uint64_t sinkBufferSize;
    std::atomic<uint64_t> counterGPU{0};

private:
    std::atomic<uint64_t> counter{0};
    std::atomic<uint64_t> possibleMatches{0};
    std::atomic<uint64_t> totalBuildTime{0}, totalProbeTime{0}, totalSortTime{0}, totalMergeTime{0};
    std::atomic<uint64_t> totalExecutedWindows{0};
    std::atomic<uint64_t> checksum1{0}, checksum2{0};
    std::chrono::time_point<std::chrono::high_resolution_clock> processingStartTime;
    std::mutex progressMutex;
    std::vector<Progress> progress;
};
