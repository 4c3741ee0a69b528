// runExperiment <yaml> <out.csv> (executable/runExperiment.cpp:14-198):
// reads the experiment YAML (same keys and values), expands the cartesian
// product of the lists, applies the INTRA/INTER/HYBRID thread mapping and
// runs the Throughput or cumulativePercentage experiment.
//
// Optional keys beyond the reference's: dataSizeTuples (initial data size
// per stream), precisionTuples, sourceThreads, extendedCsv ("Yes" appends
// match_count, checksum1, input_tuples_per_s), rateSearch ("No" runs the
// initial data size once), usePK ("Yes": persistent-kernel eager executor),
// maxDataSizeTuples (bound on the tuples per stream, see ExperimentConfig).
#include <algorithm>
#include <cstdio>
#include <exception>
#include <iostream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "experiment/Config.hpp"
#include "experiment/Runner.hpp"
#include "sjb/windowing/BaseWindowing.hpp"
#include "yaml_lite.hpp"

static constexpr uint64_t KiB = 1024 / sizeof(Tuple);
static constexpr uint64_t MiB = KiB * 1024;

static const std::vector<std::string> &need(const yaml_lite::Doc &d, const std::string &k) {
    auto it = d.find(k);
    if (it == d.end()) throw std::invalid_argument("missing key: " + k);
    return it->second;
}

static std::string opt(const yaml_lite::Doc &d, const std::string &k, const std::string &def) {
    auto it = d.find(k);
    return it == d.end() || it->second.empty() ? def : it->second[0];
}

static bool yesno(const std::string &v, const std::string &what) {
    if (v == "Yes") return true;
    if (v == "No") return false;
    throw std::invalid_argument("Unknown parameter for " + what + ": " + v);
}

int main(int argc, char **argv) {
    if (argc < 3) {
        std::cout << "Please run using ./runExperiment <path_to_yaml_config> <outputFileName>." << std::endl;
        return 1;
    }
    try {
        // defaults (runExperiment.cpp:23-38)
        ExperimentConfig base;
        base.dataConfig = {2560 * MiB, 0.0f, DataGenerator::STEADY, 128 * KiB, 1000, 1000};
        base.algorithmConfig = {Windowing::HJ, Windowing::CPU, Windowing::EAGER, Windowing::NoOutput};
        base.engineConfig = {1024 * MiB, 128, 1, 32, 32 * MiB};
        base.queryConfig = {1 * MiB};
        base.precision = 32 * KiB;

        const yaml_lite::Doc y = yaml_lite::load_file(argv[1]);
        std::vector<Windowing::PARALLELIZATION_STRATEGY> parStrats;
        for (const auto &s : need(y, "parStrats")) {
            if (s == "INTRA") parStrats.push_back(Windowing::INTRA);
            else if (s == "INTER") parStrats.push_back(Windowing::INTER);
            else if (s == "HYBRID") parStrats.push_back(Windowing::HYBRID);
            else throw std::invalid_argument("Unknown parallelization strategy: " + s);
        }
        std::vector<Windowing::ALGORITHM_TYPE> algs;
        for (const auto &s : need(y, "algTypes")) {
            if (s == "NLJ") algs.push_back(Windowing::NLJ);
            else if (s == "HJ") algs.push_back(Windowing::HJ);
            else if (s == "SMJ") algs.push_back(Windowing::SMJ);
            else throw std::invalid_argument("Unknown algorithm type: " + s);
        }
        std::vector<Windowing::DEVICE_TYPE> devs;
        for (const auto &s : need(y, "deviceTypes")) {
            if (s == "CPU") devs.push_back(Windowing::CPU);
            else if (s == "GPU") devs.push_back(Windowing::GPU);
            else throw std::invalid_argument("Unknown device type: " + s);
        }
        std::vector<Windowing::PROGRESSIVENESS_TYPE> progs;
        for (const auto &s : need(y, "progressivenessModes")) {
            if (s == "LAZY") progs.push_back(Windowing::LAZY);
            else if (s == "EAGER") progs.push_back(Windowing::EAGER);
            else throw std::invalid_argument("Unknown progressiveness mode: " + s);
        }
        std::vector<Windowing::RESULT_WRITING_METHOD> methods;
        for (const auto &s : need(y, "resultMaterializationMethod")) {
            if (s == "NoOutput") methods.push_back(Windowing::NoOutput);
            else if (s == "CountKernel") methods.push_back(Windowing::CountKernel);
            else if (s == "EstimatedSelectivity") methods.push_back(Windowing::EstimatedSelectivity);
            else if (s == "Bitmap") methods.push_back(Windowing::Bitmap);
            else if (s == "Atomic") methods.push_back(Windowing::Atomic);
            else throw std::invalid_argument("Unknown result materialization method: " + s);
        }
        std::vector<uint64_t> keys, windows, batches;
        std::vector<float> thetas;
        for (const auto &s : need(y, "distinctKeys")) keys.push_back(std::stoull(s));
        for (const auto &s : need(y, "zipfThetas")) thetas.push_back(std::stof(s));
        for (const auto &s : need(y, "windowSizesMB")) windows.push_back(std::stoull(s) * MiB);
        for (const auto &s : need(y, "batchSizeKB")) batches.push_back(std::stoull(s) * KiB);
        base.measurePower = yesno(opt(y, "measurePowerUsage", "No"), "measure power");
        base.repeat = std::stoull(opt(y, "repeat", "1"));
        base.extendedCsv = yesno(opt(y, "extendedCsv", "No"), "extendedCsv");
        base.searchRate = yesno(opt(y, "rateSearch", "Yes"), "rateSearch");
        base.algorithmConfig.eagerJoinConfig.usePK = yesno(opt(y, "usePK", "No"), "usePK");
        if (y.count("dataSizeTuples")) base.dataConfig.dataSize = std::stoull(opt(y, "dataSizeTuples", "0"));
        if (y.count("precisionTuples")) base.precision = std::stoull(opt(y, "precisionTuples", "0"));
        if (y.count("maxDataSizeTuples")) base.maxDataSize = std::stoull(opt(y, "maxDataSizeTuples", "0"));
        if (y.count("sourceThreads")) base.engineConfig.nSourceThreads = std::stoull(opt(y, "sourceThreads", "1"));
        const uint64_t hw = std::max(1u, std::thread::hardware_concurrency());

        std::vector<ExperimentConfig> configs;
        for (auto alg : algs)
            for (auto dev : devs)
                for (auto prog : progs)
                    for (auto D : keys)
                        for (auto m : methods)
                            for (auto W : windows)
                                for (auto B : batches)
                                    for (auto th : thetas)
                                        for (auto ps : parStrats) {
                                            ExperimentConfig c = base;
                                            c.dataConfig.batchSize = B;
                                            c.dataConfig.zipfTheta = th;
                                            c.algorithmConfig.progressiveness = prog;
                                            c.algorithmConfig.device = dev;
                                            c.algorithmConfig.underlyingAlgorithm = alg;
                                            c.algorithmConfig.resultWritingMethod = m;
                                            c.dataConfig.numDistinctKey = D;
                                            c.queryConfig.windowSize = W;
                                            c.engineConfig.maxTupleInWindow = W;
                                            // runExperiment.cpp:162-175
                                            if (ps == Windowing::INTRA) {
                                                c.engineConfig.joinerThreads = 1;
                                                c.algorithmConfig.cpuhjConfig.nProheThreads = hw;
                                                c.algorithmConfig.cpusmjConfig.nSortThreads = hw;
                                            } else if (ps == Windowing::INTER) {
                                                c.engineConfig.joinerThreads = hw;
                                                c.algorithmConfig.cpuhjConfig.nProheThreads = 1;
                                                c.algorithmConfig.cpusmjConfig.nSortThreads = 1;
                                            } else {
                                                c.engineConfig.joinerThreads = 4;
                                                c.algorithmConfig.cpuhjConfig.nProheThreads = std::max<uint64_t>(1, hw / 4);
                                                c.algorithmConfig.cpusmjConfig.nSortThreads = std::max<uint64_t>(1, hw / 4);
                                            }
                                            // GPU joiners each own a device context and
                                            // stream, so windows joined by different
                                            // joiners overlap (one window's host-to-device
                                            // copy with another's join); past 4 contexts
                                            // they only add device memory (the copy
                                            // engines and the device are shared)
                                            if (dev == Windowing::GPU)
                                                c.engineConfig.joinerThreads =
                                                    std::min<uint64_t>(c.engineConfig.joinerThreads, 4);
                                            configs.push_back(c);
                                        }
        Runner runner;
        const std::string metric = opt(y, "measuredMetric", "Throughput");
        if (metric == "Throughput")
            runner.runExperimentsAndWriteResultToFile(configs, argv[2]);
        else if (metric == "cumulativePercentage")
            runner.runProgressivenessExperimentAndWriteResultToFile(configs, argv[2]);
        else
            throw std::invalid_argument("Unknown measured metrics: " + metric);
    } catch (const std::exception &e) {
        std::cerr << "runExperiment: " << e.what() << std::endl;
        return 2;
    }
    return 0;
}
