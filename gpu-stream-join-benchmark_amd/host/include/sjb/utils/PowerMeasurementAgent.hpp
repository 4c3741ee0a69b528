// PowerMeasurementAgent — GPU power polling during a run, as the reference's
// agent does with NVML (src/sjb/utils/PowerMeasurementAgent.cpp:50-86):
// start() spawns a thread that appends "<ms since epoch>,<watts>" lines to
// power_data.txt (in the working directory) until end().  Readings come from
// the AMD SMI library (libamd_smi.so, loaded at run time): the socket power
// of the first GPU, in watts.  When the library or a GPU is missing, start()
// reports why and returns false, and nothing is written.
#pragma once
#include <atomic>
#include <cstdint>
#include <string>
#include <thread>

class PowerMeasurementAgent {
public:
    explicit PowerMeasurementAgent(std::string outputFilename = "power_data.txt",
                                   uint32_t intervalMicros = 1000);
    ~PowerMeasurementAgent();
    bool start();
    void end();
    uint64_t samples() const { return nSamples.load(); }
    const std::string &error() const { return err; }

private:
    void poll();
    std::string outputFilename;
    uint32_t intervalMicros;
    std::atomic<bool> running{false};
    std::atomic<uint64_t> nSamples{0};
    std::thread thread;
    std::string err;
};
