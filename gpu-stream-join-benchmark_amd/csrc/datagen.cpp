// datagen.cpp — deterministic synthetic streams (host).
//
// The reference generates keys on the GPU from curand with a Zipf kernel
// (src/sjb/utils/RandomGenerator.cu:12-115) and sets val = key
// (src/sjb/utils/DataGenerator.cpp:66-67); curand's stream is a closed
// library and cannot be reproduced, so this generator is the build's own:
//   * counter-based SplitMix64 (Steele, Lea, Flood 2014): the i-th tuple's
//     draws are mix64 of (seed, i, draw) — any thread split gives the same
//     bytes;
//   * uniform keys in [0, key_range) by multiply-shift;
//   * Zipf(theta) ranks in [1, key_range] by rejection-inversion (Hoermann &
//     Derflinger, "Rejection-inversion to generate variates from monotone
//     discrete distributions", 1996), valid for any theta > 0 including
//     theta >= 1, then scrambled by a fixed bijection of [0, key_range) so
//     hot keys are not the smallest ones;
//   * val: an independent uniform u32 (so checksums see val != key);
//   * ts = ts_base + index (monotone).
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "sjb.h"

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

inline uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// draw d of tuple i of stream `seed`
inline uint64_t draw(uint64_t seed, uint64_t i, uint64_t d) {
    return mix(mix(seed * 0xD1B54A32D192ED03ULL + d * 0xA0761D6478BD642FULL) + (i + 1) * kGolden);
}

inline double u01(uint64_t x) { return double(x >> 11) * (1.0 / 9007199254740992.0); }

// Bijection of [0, range): multiply/xorshift rounds on the enclosing
// power-of-two domain, cycle-walking until the value falls inside range.
struct Scramble {
    uint64_t range, mask;
    int bits;
    explicit Scramble(uint64_t r) : range(r) {
        bits = 1;
        while ((1ULL << bits) < r) bits++;
        mask = bits >= 64 ? ~0ULL : (1ULL << bits) - 1;
    }
    uint64_t once(uint64_t x) const {
        x = (x * 0x9E3779B97F4A7C15ULL) & mask;
        x ^= x >> ((bits + 1) / 2);
        x = (x * 0xC2B2AE3D27D4EB4FULL) & mask;
        x ^= x >> ((bits + 2) / 3);
        return x;
    }
    uint64_t operator()(uint64_t x) const {
        do {
            x = once(x);
        } while (x >= range);
        return x;
    }
};

// Rejection-inversion Zipf sampler over {1..n}, P(k) ~ k^-s.
struct Zipf {
    double s, n, hx1, hn, sval;
    static double helper1(double x) { return fabs(x) > 1e-8 ? log1p(x) / x : 1 - x * (0.5 - x * (1.0 / 3 - 0.25 * x)); }
    static double helper2(double x) { return fabs(x) > 1e-8 ? expm1(x) / x : 1 + x * 0.5 * (1 + x * (1.0 / 3) * (1 + 0.25 * x)); }
    double h(double x) const { return exp(-s * log(x)); }
    double H(double x) const {
        const double lx = log(x);
        return helper2((1 - s) * lx) * lx;
    }
    double Hinv(double x) const {
        double t = x * (1 - s);
        if (t < -1) t = -1;
        return exp(helper1(t) * x);
    }
    Zipf(double s_, uint64_t n_) : s(s_), n(double(n_)) {
        hx1 = H(1.5) - 1.0;
        hn = H(n + 0.5);
        sval = 2 - Hinv(H(2.5) - h(2));
    }
    template <class U>
    uint64_t sample(U &&uniform) const {
        for (;;) {
            const double u = hn + uniform() * (hx1 - hn);
            const double x = Hinv(u);
            double k = floor(x + 0.5);
            if (k < 1) k = 1;
            if (k > n) k = n;
            if (k - x <= sval || u >= H(k + 0.5) - h(k)) return uint64_t(k);
        }
    }
};

}  // namespace

extern "C" int sjb_generate(sjb_tuple_t *out, uint64_t n, int dist, uint64_t key_range, double theta,
                            uint64_t seed, uint64_t ts_base) {
    if ((n && !out) || key_range == 0 || key_range > (1ULL << 32)) return SJB_E_INVALID;
    if (dist != SJB_DIST_UNIFORM && dist != SJB_DIST_ZIPF) return SJB_E_INVALID;
    if (dist == SJB_DIST_ZIPF && !(theta >= 0.0)) return SJB_E_INVALID;
    const bool zipf = dist == SJB_DIST_ZIPF && theta > 0.0;
    const Zipf z(zipf ? theta : 1.0, key_range);
    const Scramble sc(key_range);
    auto work = [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            uint64_t key;
            if (zipf) {
                uint64_t d = 2;
                const uint64_t rank = z.sample([&] { return u01(draw(seed, i, d++)); });
                key = sc(rank - 1);
            } else {
                key = ((draw(seed, i, 0) >> 32) * key_range) >> 32;
            }
            out[i].key = uint32_t(key);
            out[i].val = uint32_t(draw(seed, i, 1));
            out[i].ts = ts_base + i;
        }
    };
    unsigned hw = std::thread::hardware_concurrency();
    unsigned T = std::max(1u, std::min(hw ? hw : 1u, 16u));
    if (n < (1u << 16)) T = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++) th.emplace_back(work, n * t / T, n * (t + 1) / T);
    for (auto &x : th) x.join();
    return SJB_OK;
}

extern "C" int sjb_generate_mix(sjb_tuple_t *out, uint64_t n, uint64_t window, double zipf_frac,
                                double theta, uint64_t seed, uint64_t ts_base) {
    if ((n && !out) || !(zipf_frac >= 0.0 && zipf_frac <= 1.0) || !(theta > 0.0)) return SJB_E_INVALID;
    constexpr int kComp = 8;
    const double two32 = 4294967296.0;
    double cum[kComp], mean[kComp], sigma[kComp];
    double wsum = 0;
    for (int j = 0; j < kComp; j++) wsum += 1 + (5 * j) % 4;
    double acc = 0;
    const double drift = 0.01 * two32 * double(window / 64);
    for (int j = 0; j < kComp; j++) {
        acc += (1 + (5 * j) % 4) / wsum;
        cum[j] = acc;
        mean[j] = fmod((0.06 + 0.12 * j) * two32 + drift, two32);
        sigma[j] = (0.004 + 0.002 * (j % 4)) * two32;
    }
    const Zipf z(theta, 1ULL << 32);
    const Scramble sc(1ULL << 32);
    auto work = [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            uint64_t key;
            if (u01(draw(seed, i, 2)) < zipf_frac) {
                uint64_t d = 8;
                key = sc(z.sample([&] { return u01(draw(seed, i, d++)); }) - 1);
            } else {
                const double c = u01(draw(seed, i, 3));
                int j = 0;
                while (j < kComp - 1 && c >= cum[j]) j++;
                // Box-Muller from two further draws
                const double u1 = (double(draw(seed, i, 4) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
                const double u2 = u01(draw(seed, i, 5));
                const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
                double v = fmod(floor(mean[j] + sigma[j] * g + 0.5), two32);
                if (v < 0) v += two32;
                key = uint64_t(v) & 0xFFFFFFFFULL;
            }
            out[i].key = uint32_t(key);
            out[i].val = uint32_t(draw(seed, i, 1));
            out[i].ts = ts_base + i;
        }
    };
    unsigned hw = std::thread::hardware_concurrency();
    unsigned T = std::max(1u, std::min(hw ? hw : 1u, 16u));
    if (n < (1u << 16)) T = 1;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < T; t++) th.emplace_back(work, n * t / T, n * (t + 1) / T);
    for (auto &x : th) x.join();
    return SJB_OK;
}
