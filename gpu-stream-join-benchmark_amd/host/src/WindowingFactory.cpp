// WindowingFactory (src/sjb/windowing/WindowingFactory.cpp:52-298): same
// combinations, executor classes and CSV names.  Differences: the eager GPU
// NLJ NoOutput branch is reachable (the reference's missing `else` makes it
// throw, WindowingFactory.cpp:275-284), and eager GPU HJ with usePK selects
// the persistent-kernel executor (the reference comments it out and throws,
// :245-251).
#include "sjb/windowing/WindowingFactory.hpp"

#include <stdexcept>

#include "sjb/windowing/eager/CountBased.hpp"
#include "sjb/windowing/eager/executor/Executors.hpp"
#include "sjb/windowing/lazy/CountBased.hpp"
#include "sjb/windowing/lazy/executor/Executors.hpp"

namespace Windowing {

std::shared_ptr<BaseWindowing> WindowingFactory::create(AlgorithmConfig a, DataConfig d, QueryConfig q,
                                                       EngineConfig e, Sink &sink) {
    using namespace Lazy::Executor;
    std::string name = "NOTSET";
    if (a.progressiveness == LAZY) {
        std::vector<std::shared_ptr<BaseLazyExecutor>> joiners;
        const uint64_t nj = e.joinerThreads ? e.joinerThreads : 1;
        const double es = getEstimatedSelectivity(d.numDistinctKey);
        if (a.device == CPU) {
            for (uint64_t i = 0; i < nj; i++) {
                if (a.underlyingAlgorithm == NLJ) {
                    name = "lazy_cpu_nlj_nm";
                    joiners.push_back(std::make_shared<CPU_NLJ>(d.batchSize));
                } else if (a.underlyingAlgorithm == HJ) {
                    name = "lazy_cpu_hj_nm";
                    joiners.push_back(std::make_shared<CPU_HJ>(d.numDistinctKey, d.batchSize,
                                                               a.cpuhjConfig.nProheThreads, e.maxTupleInWindow));
                } else if (a.underlyingAlgorithm == SMJ) {
                    name = "lazy_cpu_smj";
                    joiners.push_back(std::make_shared<CPU_SMJ>(d.numDistinctKey, d.batchSize,
                                                                a.cpusmjConfig.nSortThreads));
                } else {
                    throw std::runtime_error("Not Implemented");
                }
            }
        } else if (a.device == GPU) {
            for (uint64_t i = 0; i < nj; i++) {
                const uint8_t m = a.resultWritingMethod;
                if (a.underlyingAlgorithm == NLJ) {
                    switch (m) {
                        case NoOutput:
                            name = "lazy_gpu_nlj_nm";
                            joiners.push_back(std::make_shared<GPU_NLJ_NoOutput>(d.batchSize, e.maxTupleInWindow));
                            break;
                        case CountKernel:
                            name = "lazy_gpu_nlj_ck";
                            joiners.push_back(std::make_shared<GPU_NLJ_CountKernel>(d.batchSize, e.maxTupleInWindow));
                            break;
                        case EstimatedSelectivity:
                            name = "lazy_gpu_nlj_es";
                            joiners.push_back(std::make_shared<GPU_NLJ_EstimatedSelectivity>(
                                d.batchSize, e.maxTupleInWindow, es));
                            break;
                        case Bitmap:
                            name = "lazy_gpu_nlj_bm";
                            joiners.push_back(std::make_shared<GPU_NLJ_Bitmap>(d.batchSize, e.maxTupleInWindow));
                            break;
                        default:
                            throw std::runtime_error("Not implemented");
                    }
                } else if (a.underlyingAlgorithm == HJ) {
                    if (m == NoOutput) {
                        name = "lazy_gpu_hj_nm";
                        joiners.push_back(std::make_shared<GPU_HJ_NoOutput>(d.numDistinctKey, d.batchSize, e.maxTupleInWindow));
                    } else if (m == CountKernel) {
                        name = "lazy_gpu_hj_ck";
                        joiners.push_back(std::make_shared<GPU_HJ_CountKernel>(d.numDistinctKey, d.batchSize, e.maxTupleInWindow));
                    } else if (m == EstimatedSelectivity) {
                        name = "lazy_gpu_hj_es";
                        joiners.push_back(std::make_shared<GPU_HJ_EstimatedSelectivity>(es, d.numDistinctKey, d.batchSize,
                                                                                        e.maxTupleInWindow));
                    } else if (m == Atomic) {
                        name = "lazy_gpu_hj_at";
                        joiners.push_back(std::make_shared<GPU_HJ_Atomic>(d.numDistinctKey, d.batchSize, e.maxTupleInWindow));
                    } else {
                        throw std::runtime_error("Not Implemented");
                    }
                } else if (a.underlyingAlgorithm == SMJ) {
                    if (m == NoOutput) {
                        name = "lazy_gpu_smj_nm";
                        joiners.push_back(std::make_shared<GPU_SMJ_NoOutput>(d.numDistinctKey, d.batchSize, e.maxTupleInWindow));
                    } else if (m == CountKernel) {
                        name = "lazy_gpu_smj_ck";
                        joiners.push_back(std::make_shared<GPU_SMJ_CountKernel>(d.numDistinctKey, d.batchSize, e.maxTupleInWindow));
                    } else if (m == EstimatedSelectivity) {
                        name = "lazy_gpu_smj_es";
                        joiners.push_back(std::make_shared<GPU_SMJ_EstimatedSelectivity>(d.numDistinctKey, d.batchSize,
                                                                                         e.maxTupleInWindow, es));
                    } else if (m == Atomic) {
                        name = "lazy_gpu_smj_at";
                        joiners.push_back(std::make_shared<GPU_SMJ_Atomic>(d.numDistinctKey, d.batchSize, e.maxTupleInWindow));
                    } else {
                        throw std::runtime_error("Not Implemented");
                    }
                } else {
                    throw std::runtime_error("Not Implemented");
                }
            }
        } else {
            throw std::runtime_error("Not Implemented");
        }
        return std::make_shared<Lazy::CountBased>(sink, e.ringBufferSize, e.ringBufferMaxTries, q.windowSize,
                                                  d.batchSize, joiners, name);
    }
    // EAGER
    using namespace Eager::Executor;
    std::shared_ptr<BaseEagerExecutor> joiner;
    if (a.device == CPU) {
        if (a.underlyingAlgorithm != HJ) throw std::runtime_error("Not Implemented");
        name = "eager_cpu_hj_nm";
        joiner = std::make_shared<::Windowing::Eager::Executor::CPU_HJ>(e.maxTupleInWindow, d.numDistinctKey, d.batchSize);
    } else if (a.device == GPU) {
        const uint8_t m = a.resultWritingMethod;
        if (a.underlyingAlgorithm == HJ) {
            if (a.eagerJoinConfig.usePK) {
                if (m != NoOutput) throw std::runtime_error("Not Implemented");
                name = "eager_gpu_hj_nm_pk";
                joiner = std::make_shared<GPU_HJ_PK>(d.numDistinctKey, q.windowSize, d.batchSize);
            } else if (m == NoOutput) {
                name = "eager_gpu_hj_nm";
                joiner = std::make_shared<::Windowing::Eager::Executor::GPU_HJ>(d.numDistinctKey, q.windowSize, d.batchSize);
            } else if (m == Atomic) {
                name = "eager_gpu_hj_at";
                joiner = std::make_shared<::Windowing::Eager::Executor::GPU_HJ_Atomic>(d.numDistinctKey, q.windowSize, d.batchSize);
            } else {
                throw std::runtime_error("Not Implemented");
            }
        } else if (a.underlyingAlgorithm == NLJ) {
            if (a.eagerJoinConfig.usePK) throw std::runtime_error("Not Implemented");
            if (m == NoOutput) {
                name = "eager_gpu_nlj_nm";
                joiner = std::make_shared<::Windowing::Eager::Executor::GPU_NLJ>(d.batchSize, q.windowSize);
            } else if (m == Bitmap) {
                name = "eager_gpu_nlj_bm";
                joiner = std::make_shared<GPU_NLJ_BM>(d.batchSize, q.windowSize);
            } else {
                throw std::runtime_error("Not Implemented");
            }
        } else {
            throw std::runtime_error("Not Implemented");
        }
    } else {
        throw std::runtime_error("Not Implemented");
    }
    return std::make_shared<Eager::CountBased>(sink, q.windowSize, d.batchSize, joiner, name);
}

}  // namespace Windowing
