// ExecutionStatistic — per-window statistics, the fields of
// include/sjb/utils/ExecutionStatistic.hpp:6-20 plus the checksums.
#pragma once
#include <cstdint>

class ExecutionStatistic {
public:
    ExecutionStatistic() = default;
    uint64_t resultCount{};
    uint64_t buildTime{};
    uint64_t probeTime{};
    uint64_t sortTime{};
    uint64_t mergeTime{};
    // additions: order-independent checksums over (r.key, r.val, s.val)
    uint64_t checksum1{};
    uint64_t checksum2{};
    bool checksum2Valid{false};
};
