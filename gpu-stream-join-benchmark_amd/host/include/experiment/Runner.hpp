// Runner (include/experiment/Runner.hpp:9-20, src/experiment/Runner.cpp):
// one run = generate both streams, build Source/Sink/Windowing, time
// Source::start; the Throughput experiment searches the largest dataSize
// whose median run time stays below 1000 ms (doubling then bisection,
// Runner.cpp:138-181) and appends one CSV row per configuration.
#pragma once
#include <cstdint>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "experiment/Config.hpp"

class Runner {
public:
    struct RunResult {
        std::string row;
        uint64_t msElapsed = 0;
        uint64_t matches = 0;
        uint64_t checksum1 = 0;
        double selectivity = 0;
    };
    RunResult runExperimentWithConfig(ExperimentConfig config, bool trackProgress = false);
    void runExperimentsAndWriteResultToFile(std::vector<ExperimentConfig> configs, std::string outputFileName);
    void runProgressivenessExperimentAndWriteResultToFile(std::vector<ExperimentConfig> configs,
                                                          std::string outputFileName);
    static double computeMedian(std::vector<uint64_t> v);
    static std::string csvHeader(bool extended);
};
