// DataGenerator (src/sjb/utils/DataGenerator.cpp:14-74) over the C ABI
// generator; see DataGenerator.hpp for the differences.
#include "sjb/utils/DataGenerator.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "sjb/utils/Error.hpp"

void DataGenerator::generate(Tuple *tuples, uint64_t count, float zipf_theta, uint8_t tsMode,
                             uint64_t distinctKey, uint32_t seed, uint64_t ts_upperbound) {
    if (count == 0) return;
    sjbCheck(sjb_generate(reinterpret_cast<sjb_tuple_t *>(tuples), count,
                          zipf_theta > 0.0f ? SJB_DIST_ZIPF : SJB_DIST_UNIFORM, distinctKey ? distinctKey : 1,
                          double(zipf_theta), seed, 0),
             "sjb_generate");
    if (ts_upperbound == 0) ts_upperbound = 1;
    if (tsMode == STEADY) {
        const double perMs = std::max(1.0, std::floor(double(count) / double(ts_upperbound)));
        for (uint64_t i = 0; i < count; i++) tuples[i].ts = uint64_t(std::floor(double(i) / perMs));
    } else if (tsMode == BURST) {
        // density over milliseconds t in [0, U): sin(t mod 360 degrees) + 1;
        // timestamps are the quantiles (i + 0.5) / count of that density
        std::vector<double> cdf(ts_upperbound + 1, 0.0);
        for (uint64_t t = 0; t < ts_upperbound; t++)
            cdf[t + 1] = cdf[t] + std::sin(double(t % 360) * M_PI / 180.0) + 1.0;
        uint64_t t = 0;
        for (uint64_t i = 0; i < count; i++) {
            const double u = (double(i) + 0.5) / double(count) * cdf[ts_upperbound];
            while (t + 1 < ts_upperbound && cdf[t + 1] < u) t++;
            tuples[i].ts = t;
        }
    } else {
        throw std::invalid_argument("Unknown TS_MODE");
    }
}
