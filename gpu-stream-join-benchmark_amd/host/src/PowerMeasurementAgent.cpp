// PowerMeasurementAgent over the AMD SMI C library (replaces the NVML
// polling of src/sjb/utils/PowerMeasurementAgent.cpp:50-136).  The library
// is opened with dlopen so the host layer has no link-time dependency on it.
#include "sjb/utils/PowerMeasurementAgent.hpp"

#include <dlfcn.h>

#include <chrono>
#include <cstdio>
#include <vector>

#include <amd_smi/amdsmi.h>

namespace {

struct SmiApi {
    void *so = nullptr;
    decltype(&amdsmi_init) init = nullptr;
    decltype(&amdsmi_shut_down) shut_down = nullptr;
    decltype(&amdsmi_get_socket_handles) sockets = nullptr;
    decltype(&amdsmi_get_processor_handles) processors = nullptr;
    decltype(&amdsmi_get_power_info) power = nullptr;
    bool load(std::string &err) {
        for (const char *name : {"libamd_smi.so", "/opt/rocm/lib/libamd_smi.so"}) {
            so = dlopen(name, RTLD_NOW | RTLD_LOCAL);
            if (so) break;
        }
        if (!so) {
            err = "libamd_smi.so not found";
            return false;
        }
        init = reinterpret_cast<decltype(init)>(dlsym(so, "amdsmi_init"));
        shut_down = reinterpret_cast<decltype(shut_down)>(dlsym(so, "amdsmi_shut_down"));
        sockets = reinterpret_cast<decltype(sockets)>(dlsym(so, "amdsmi_get_socket_handles"));
        processors = reinterpret_cast<decltype(processors)>(dlsym(so, "amdsmi_get_processor_handles"));
        power = reinterpret_cast<decltype(power)>(dlsym(so, "amdsmi_get_power_info"));
        if (!init || !shut_down || !sockets || !processors || !power) {
            err = "libamd_smi.so lacks a needed symbol";
            return false;
        }
        return true;
    }
};

// watts of one reading: current socket power (MI300 and later), else the
// average, else the generic field; < 0 when none is reported
double watts(const amdsmi_power_info_t &p) {
    if (p.current_socket_power != UINT32_MAX && p.current_socket_power != 0) return p.current_socket_power;
    if (p.average_socket_power != UINT32_MAX && p.average_socket_power != 0) return p.average_socket_power;
    if (p.socket_power != UINT64_MAX && p.socket_power != UINT32_MAX) return double(p.socket_power);
    return -1.0;
}

}  // namespace

PowerMeasurementAgent::PowerMeasurementAgent(std::string outputFilename, uint32_t intervalMicros)
    : outputFilename(std::move(outputFilename)), intervalMicros(intervalMicros) {}

PowerMeasurementAgent::~PowerMeasurementAgent() { end(); }

bool PowerMeasurementAgent::start() {
    if (running.load()) return true;
    err.clear();
    // probe once on this thread so failures are reported to the caller
    SmiApi api;
    if (!api.load(err)) {
        fprintf(stderr, "[power] %s: no power data\n", err.c_str());
        return false;
    }
    bool ok = false;
    if (api.init(AMDSMI_INIT_AMD_GPUS) == AMDSMI_STATUS_SUCCESS) {
        uint32_t ns = 0;
        if (api.sockets(&ns, nullptr) == AMDSMI_STATUS_SUCCESS && ns > 0) ok = true;
        api.shut_down();
    }
    dlclose(api.so);
    if (!ok) {
        err = "AMD SMI found no GPU";
        fprintf(stderr, "[power] %s: no power data\n", err.c_str());
        return false;
    }
    running.store(true);
    thread = std::thread([this] { poll(); });
    return true;
}

void PowerMeasurementAgent::end() {
    if (!running.exchange(false)) return;
    if (thread.joinable()) thread.join();
}

void PowerMeasurementAgent::poll() {
    SmiApi api;
    if (!api.load(err) || api.init(AMDSMI_INIT_AMD_GPUS) != AMDSMI_STATUS_SUCCESS) {
        running.store(false);
        return;
    }
    amdsmi_processor_handle gpu = nullptr;
    uint32_t ns = 0;
    if (api.sockets(&ns, nullptr) == AMDSMI_STATUS_SUCCESS && ns > 0) {
        std::vector<amdsmi_socket_handle> socks(ns);
        if (api.sockets(&ns, socks.data()) == AMDSMI_STATUS_SUCCESS) {
            for (uint32_t s = 0; s < ns && !gpu; s++) {
                uint32_t np = 0;
                if (api.processors(socks[s], &np, nullptr) != AMDSMI_STATUS_SUCCESS || np == 0) continue;
                std::vector<amdsmi_processor_handle> procs(np);
                if (api.processors(socks[s], &np, procs.data()) == AMDSMI_STATUS_SUCCESS && np > 0) gpu = procs[0];
            }
        }
    }
    FILE *fp = gpu ? fopen(outputFilename.c_str(), "a+") : nullptr;
    while (fp && running.load()) {
        amdsmi_power_info_t info{};
        if (api.power(gpu, &info) == AMDSMI_STATUS_SUCCESS) {
            const double w = watts(info);
            if (w >= 0) {
                const auto now = std::chrono::system_clock::now().time_since_epoch();
                fprintf(fp, "%lld,%.3lf\n",
                        (long long)std::chrono::duration_cast<std::chrono::milliseconds>(now).count(), w);
                nSamples.fetch_add(1);
            }
        }
        std::this_thread::sleep_for(std::chrono::microseconds(intervalMicros));
    }
    if (fp) fclose(fp);
    api.shut_down();
    dlclose(api.so);
}
