// Synthetic stream generator with the reference's signature
// (include/sjb/utils/DataGenerator.hpp:9-16).  Keys come from the C ABI's
// deterministic generator (uniform, or Zipf(theta) over [0, distinctKey));
// val is an independent uniform u32 (the reference copies the key into
// val, src/sjb/utils/DataGenerator.cpp:66-67); ts follows the STEADY rule
// ts = floor(i / floor(count / ts_upperbound)) (DataGenerator.cpp:33-47) or a
// BURST profile whose density per millisecond is sin(t mod 360 deg) + 1.
#pragma once
#include <cstdint>

#include "sjb/utils/Tuple.hpp"

class DataGenerator {
public:
    enum TS_MODE : uint8_t { STEADY = 0, BURST = 1 };
    static void generate(Tuple *tuples, uint64_t count, float zipf_theta, uint8_t tsMode,
                         uint64_t distinctKey, uint32_t seed, uint64_t ts_upperbound);
};
