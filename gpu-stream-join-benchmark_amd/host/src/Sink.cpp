// Sink (src/sjb/sink/Sink.cpp:6-125).
#include "sjb/sink/Sink.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <new>

#include "sjb/utils/Pinned.hpp"

// sinkBuffer and counterGPU are page-locked when a device is present
// (Sink.cpp:8-9 allocates both with cudaMallocHost), so device kernels
// write pair rows straight into the ring.
Sink::Sink(uint64_t sinkBufferSize) : sinkBufferSize(sinkBufferSize) {
    sinkBuffer = static_cast<ResultTuple *>(
        sjb_host::pinnedAlloc((sinkBufferSize ? sinkBufferSize : 1) * sizeof(ResultTuple), &bufferPinned));
    try {
        counterGPU = static_cast<unsigned long long *>(
            sjb_host::pinnedAlloc(sizeof(unsigned long long), &counterPinned));
    } catch (...) {
        sjb_host::pinnedFree(sinkBuffer, bufferPinned);
        throw;
    }
}

Sink::~Sink() {
    sjb_host::pinnedFree(sinkBuffer, bufferPinned);
    sjb_host::pinnedFree(counterGPU, counterPinned);
}

void Sink::addToCounter(uint64_t currentMatches, bool isWindowClosed) {
    const uint64_t now = counter.fetch_add(currentMatches) + currentMatches;
    const auto t = std::chrono::high_resolution_clock::now();
    const auto us = std::chrono::duration_cast<std::chrono::microseconds>(t - processingStartTime).count();
    std::lock_guard<std::mutex> g(progressMutex);
    progress.push_back(Progress{now, isWindowClosed, uint64_t(us)});
}

uint64_t Sink::getMatchCount(bool useGPUCounter) {
    return useGPUCounter ? __atomic_load_n(counterGPU, __ATOMIC_RELAXED) : counter.load();
}
const std::atomic<uint64_t> &Sink::getPossibleMatches() const { return possibleMatches; }
void Sink::addToPossibleMatches(uint64_t c) { possibleMatches.fetch_add(c); }
void Sink::addToBuildTime(uint64_t t) { totalBuildTime.fetch_add(t); }
void Sink::addToProbeTime(uint64_t t) { totalProbeTime.fetch_add(t); }
void Sink::addToSortTime(uint64_t t) { totalSortTime.fetch_add(t); }
void Sink::addToMergeTime(uint64_t t) { totalMergeTime.fetch_add(t); }
void Sink::addToExecutedWindows(uint64_t n) { totalExecutedWindows.fetch_add(n); }
void Sink::addToStatistics(ExecutionStatistic statistic) {
    checksum1.fetch_add(statistic.checksum1);
    checksum2.fetch_add(statistic.checksum2Valid ? statistic.checksum2 : 0);
}
uint64_t Sink::getSinkBufferSize() const { return sinkBufferSize; }
const std::atomic<uint64_t> &Sink::getTotalExecutedWindows() const { return totalExecutedWindows; }

// The reference divides by the executed-window count even when it is 0
// (src/sjb/sink/Sink.cpp:69-83); these return 0 then.
static uint64_t avg(const std::atomic<uint64_t> &tot, const std::atomic<uint64_t> &n) {
    const uint64_t d = n.load();
    return d ? tot.load() / d : 0;
}
uint64_t Sink::getAverageSortTime() { return avg(totalSortTime, totalExecutedWindows); }
uint64_t Sink::getAverageMergeTime() { return avg(totalMergeTime, totalExecutedWindows); }
uint64_t Sink::getAverageBuildTime() { return avg(totalBuildTime, totalExecutedWindows); }
uint64_t Sink::getAverageProbeTime() { return avg(totalProbeTime, totalExecutedWindows); }

std::vector<Progress> Sink::getProgress() {
    std::lock_guard<std::mutex> g(progressMutex);
    return progress;
}

void Sink::markProcessingStart() { processingStartTime = std::chrono::high_resolution_clock::now(); }

void Sink::incrementCounterAndStore(uint32_t lkey, uint32_t rkey, uint32_t lVal, uint32_t rVal, uint64_t lTs,
                                    uint64_t rTs) {
    const uint64_t pos = counter.fetch_add(1);
    if (sinkBufferSize == 0) return;
    ResultTuple &r = sinkBuffer[pos % sinkBufferSize];
    r.lkey = lkey;
    r.rkey = rkey;
    r.lVal = lVal;
    r.rVal = rVal;
    r.lTs = lTs;
    r.rTs = rTs;
}

void Sink::incrementCounterAndStore(ResultTuple *resultTuple, uint64_t tupleCount) {
    counter.fetch_add(tupleCount);
    const uint64_t n = std::min(tupleCount, sinkBufferSize);
    if (n && resultTuple != sinkBuffer) std::memcpy(sinkBuffer, resultTuple, n * sizeof(ResultTuple));
}
