// Runner (src/experiment/Runner.cpp:14-233).
#include "experiment/Runner.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "sjb/sink/Sink.h"
#include "sjb/source/Source.hpp"
#include "sjb/utils/Pinned.hpp"
#include "sjb/utils/PowerMeasurementAgent.hpp"
#include "sjb/windowing/WindowingFactory.hpp"

std::string Runner::csvHeader(bool extended) {
    std::string h =
        "algorithm_name,write_result,ingestion_rate,batch_size,zipf_theta,distinct_key,window_size,"
        "probe_threads,joiner_threads,selectivity,time";
    if (extended) h += ",match_count,checksum1,input_tuples_per_s";
    return h + "\n";
}

// Runner.cpp:14-104
Runner::RunResult Runner::runExperimentWithConfig(ExperimentConfig config, bool trackProgress) {
    const DataConfig &d = config.dataConfig;
    if (config.maxDataSize && d.dataSize > config.maxDataSize)
        throw std::invalid_argument("data size " + std::to_string(d.dataSize) + " tuples per stream exceeds maxDataSize " +
                                    std::to_string(config.maxDataSize));
    // page-locked source streams (Runner.cpp:18-33), so GPU executors copy
    // batches and windows from them asynchronously
    sjb_host::PinnedArray<Tuple> left(d.dataSize), right(d.dataSize);
    std::thread gl([&] {
        DataGenerator::generate(left.data(), d.dataSize, d.zipfTheta, d.tsMode, d.numDistinctKey, 123,
                                d.timestampUpperBound);
    });
    DataGenerator::generate(right.data(), d.dataSize, d.zipfTheta, d.tsMode, d.numDistinctKey, 456,
                            d.timestampUpperBound);
    gl.join();

    Source source(d.dataSize, d.batchSize);
    Sink sink;
    sink.counterGPU[0] = 0;  // Runner.cpp:44
    auto windowing = Windowing::WindowingFactory::create(config.algorithmConfig, config.dataConfig,
                                                         config.queryConfig, config.engineConfig, sink);
    PowerMeasurementAgent pma;  // Runner.cpp:36-62: polls while the source runs
    if (config.measurePower) pma.start();
    const auto t0 = std::chrono::high_resolution_clock::now();
    sink.markProcessingStart();
    source.start(left.data(), right.data(), config.engineConfig.nSourceThreads, windowing);
    const auto t1 = std::chrono::high_resolution_clock::now();
    if (config.measurePower) pma.end();

    RunResult r;
    r.msElapsed = uint64_t(std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count());
    const double secs = std::chrono::duration<double>(t1 - t0).count();
    r.matches = sink.getMatchCount(config.algorithmConfig.resultWritingMethod == Windowing::Atomic);
    r.checksum1 = sink.getChecksum1();
    const uint64_t possible = sink.getPossibleMatches().load();
    r.selectivity = possible ? 100.0 * double(r.matches) / double(possible) : 0.0;
    std::ostringstream out;
    if (trackProgress) {
        for (const auto &p : sink.getProgress())
            out << windowing->getAlgorithmName() << "," << d.batchSize << "," << r.selectivity << ","
                << p.isWindowClosed << "," << p.currentMatches << "," << p.timestamp << "\n";
    } else {
        out << windowing->getAlgorithmName() << "," << unsigned(config.algorithmConfig.resultWritingMethod) << ","
            << d.dataSize << "," << d.batchSize << "," << d.zipfTheta << "," << d.numDistinctKey << ","
            << config.queryConfig.windowSize << "," << config.algorithmConfig.cpuhjConfig.nProheThreads << ","
            << config.engineConfig.joinerThreads << "," << r.selectivity << "," << r.msElapsed;
        if (config.extendedCsv)
            out << "," << r.matches << "," << r.checksum1 << "," << (secs > 0 ? 2.0 * double(d.dataSize) / secs : 0.0);
        out << "\n";
    }
    r.row = out.str();
    fprintf(stderr, "[runner] %s matches=%llu selectivity=%.6f%% elapsed=%llu ms\n",
            windowing->getAlgorithmName().c_str(), (unsigned long long)r.matches, r.selectivity,
            (unsigned long long)r.msElapsed);
    return r;
}

double Runner::computeMedian(std::vector<uint64_t> v) {
    if (v.empty()) return 0;
    std::sort(v.begin(), v.end());
    if (v.size() % 2 == 0) return (double(v[v.size() / 2 - 1]) + double(v[v.size() / 2])) / 2.0;
    return double(v[v.size() / 2]);
}

// Rate search of Runner.cpp:106-191: the largest dataSize whose median
// run time is below 1000 ms (doubling until unsustained, then bisection
// until the step is within `precision` or the median is 1000 +- 50 ms).
void Runner::runExperimentsAndWriteResultToFile(std::vector<ExperimentConfig> configs, std::string outputFileName) {
    std::ofstream f(outputFileName, std::ios_base::app);
    if (f.tellp() == 0) f << csvHeader(!configs.empty() && configs[0].extendedCsv);
    for (size_t i = 0; i < configs.size(); i++) {
        fprintf(stderr, "[runner] configuration %zu of %zu\n", i + 1, configs.size());
        ExperimentConfig cur = configs[i];
        RunResult result;
        if (!cur.searchRate) {
            std::vector<uint64_t> ms;
            for (uint64_t rep = 0; rep < std::max<uint64_t>(cur.repeat, 1); rep++) {
                result = runExperimentWithConfig(cur);
                ms.push_back(result.msElapsed);
            }
            f << result.row;
            continue;
        }
        uint64_t delta = UINT64_MAX;
        uint64_t latestUnsustained = 0;
        uint64_t currentRate = cur.dataConfig.dataSize, latestSustained = currentRate, nextRate = currentRate;
        while (delta > cur.precision) {
            std::vector<uint64_t> ms;
            for (uint64_t rep = 0; rep < std::max<uint64_t>(cur.repeat, 1); rep++) {
                result = runExperimentWithConfig(cur);
                ms.push_back(result.msElapsed);
            }
            const double median = computeMedian(ms);
            if (median < 1000) {
                latestSustained = currentRate;
                nextRate = latestUnsustained == 0 ? nextRate * 2 : (latestUnsustained + currentRate) / 2;
                if (cur.maxDataSize && nextRate > cur.maxDataSize) {
                    if (currentRate >= cur.maxDataSize) {
                        fprintf(stderr, "[runner] sustained at maxDataSize (%llu tuples per stream): search capped\n",
                                (unsigned long long)cur.maxDataSize);
                        break;
                    }
                    nextRate = cur.maxDataSize;
                }
            } else {
                latestUnsustained = currentRate;
                nextRate = (latestSustained + currentRate) / 2;
            }
            delta = currentRate > nextRate ? currentRate - nextRate : nextRate - currentRate;
            currentRate = nextRate;
            cur.dataConfig.dataSize = currentRate;
            if (median > 950 && median < 1050) break;
        }
        f << result.row;
    }
}

void Runner::runProgressivenessExperimentAndWriteResultToFile(std::vector<ExperimentConfig> configs,
                                                              std::string outputFileName) {
    for (size_t i = 0; i < configs.size(); i++) {
        RunResult r = runExperimentWithConfig(configs[i], true);
        std::ofstream f(outputFileName, std::ios_base::app);
        if (f.tellp() == 0) f << "algorithm_name,batch_size,selectivity,isWindowClosed,match_count,timestamp\n";
        f << r.row;
    }
}
