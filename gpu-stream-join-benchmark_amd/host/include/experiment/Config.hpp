// Experiment configuration structs (include/experiment/Config.hpp:6-65).
#pragma once
#include <cstdint>

#include "sjb/utils/DataGenerator.hpp"

struct CPUHJConfig {
    uint64_t nProheThreads = 1;
};
struct CPUSMJConfig {
    uint64_t nSortThreads = 1;
};
struct EagerJoinConfig {
    bool usePK = false;  // persistent-kernel eager executor
};
struct DataConfig {
    uint64_t dataSize;
    float zipfTheta;
    uint8_t tsMode;
    uint64_t batchSize;
    uint64_t numDistinctKey;
    uint64_t timestampUpperBound = 1000;
};
struct QueryConfig {
    uint64_t windowSize;
};
struct AlgorithmConfig {
    uint8_t underlyingAlgorithm;
    uint8_t device;
    uint8_t progressiveness;
    uint8_t resultWritingMethod = 0;
    CPUHJConfig cpuhjConfig;
    CPUSMJConfig cpusmjConfig;
    EagerJoinConfig eagerJoinConfig;
};
struct EngineConfig {
    uint64_t ringBufferSize;
    uint64_t ringBufferMaxTries;
    uint64_t joinerThreads;
    uint64_t nSourceThreads;
    uint64_t maxTupleInWindow;
};
struct ExperimentConfig {
    DataConfig dataConfig;
    QueryConfig queryConfig;
    AlgorithmConfig algorithmConfig;
    EngineConfig engineConfig;
    uint64_t precision = 4096;
    uint64_t repeat = 1;
    bool measurePower = false;
    bool extendedCsv = false;      // append count/checksum/throughput columns
    bool searchRate = true;        // false: run dataSize once (no rate search)
    // bound on the tuples per source stream (both streams are held in host
    // memory for the run): the rate search stops there and reports a capped
    // result; a run above it is refused.  2^30 tuples (2 x 16 GiB) is the
    // largest size this driver has run on the MI355X host; a search step at
    // 2^31 ended the process there, for a reason not found.
    uint64_t maxDataSize = uint64_t(1) << 30;
};
