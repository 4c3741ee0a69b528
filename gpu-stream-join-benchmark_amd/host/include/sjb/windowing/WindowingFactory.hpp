// WindowingFactory (include/sjb/windowing/WindowingFactory.hpp,
// src/sjb/windowing/WindowingFactory.cpp:52-298): maps (progressiveness,
// device, algorithm, result method) to a windowing with its executors and
// the reference's CSV algorithm name (e.g. "lazy_gpu_hj_nm").  Unsupported
// combinations throw std::runtime_error("Not Implemented") as in the
// reference.
#pragma once
#include <memory>

#include "experiment/Config.hpp"
#include "sjb/windowing/BaseWindowing.hpp"

namespace Windowing {
class WindowingFactory {
public:
    static std::shared_ptr<BaseWindowing> create(AlgorithmConfig algorithmConfig, DataConfig dataConfig,
                                                 QueryConfig queryConfig, EngineConfig engineConfig,
                                                 Sink &sink);
    static double getEstimatedSelectivity(uint64_t numDistinctKey) { return 1.0 / double(numDistinctKey); }
};
}  // namespace Windowing
