// rmi.hip — the learned partitioner: a two-stage RMI over the key CDF,
// fitted on the GPU from a sample of earlier windows.
//
// Fit (sjb_rmi_fit):
//   1. k_rmi_sample + k_sort_keys: 16384 strided keys of the training
//      tuples, bitonic-sorted in LDS by one workgroup.
//   2. host: min/max key (normalisation) and the heavy keys — the most
//      frequent sample keys (at most 32, each at least heavy_min times);
//      grouped joins aggregate them apart, so they are removed from the
//      sample before the CDF is fitted.
//   3. k_make_xt: x = (key - kmin) * scale, label t = (rank + 0.5) / n.
//   4. root: linear least squares (k_linear_fit) or a 1-32-32-1 ReLU MLP
//      trained by full-batch Adam with this file's forward/backward kernel
//      (k_mlp_grad) and update kernel (k_adam); the trained MLP is then
//      sampled at kRootTable+1 points (k_mlp_table, made monotone).
//   5. leaves: per-leaf least squares of t on x over the samples the root
//      routes to each of the 1024 leaves (k_leaf_stats, k_leaf_fit).
//   6. host check against the plain radix path: the learned partition of
//      the sample must not be more imbalanced than MSD radix on the key
//      bits (x 1.25); otherwise the model is marked unusable and joins with
//      SJB_F_LEARNED fall back to the radix / hash path.
//
// Any partition function of the key alone keeps the join exact (equal keys
// always meet), so the model only changes balance, never results.
#include <math.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "sjb_device.hpp"
#include "sjb_internal.hpp"

namespace sjb {

constexpr int kFitSample = 16384;   // keys sampled per fit
constexpr int kFitMax = 65536;      // samples of sjb_mlp_grad
constexpr int kGradTile = 128;      // samples per k_mlp_grad workgroup
// parameter offsets of the MLP (W2 row-major: W2[i][j], i = output unit)
constexpr int oW1 = 0, oB1 = 32, oW2 = 64, oB2 = 64 + 1024, oW3 = oB2 + 32, oB3 = oW3 + 32;
static_assert(oB3 + 1 == kMlpParams, "MLP parameter layout");

__global__ void k_rmi_sample(const sjb_tuple_t *__restrict__ T, uint64_t n, uint32_t ns,
                             uint32_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ns) out[i] = T[(uint64_t(i) * n) / ns].key;
}

// Sorts kFitSample keys in LDS (one workgroup, bitonic).
__global__ __launch_bounds__(1024) void k_sort_keys(uint32_t *__restrict__ keys) {
    __shared__ uint32_t s[kFitSample];
    for (int i = threadIdx.x; i < kFitSample; i += 1024) s[i] = keys[i];
    __syncthreads();
    for (int k = 2; k <= kFitSample; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < kFitSample; i += 1024) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const uint32_t a = s[i], b = s[ixj];
                    if ((a > b) == ((i & k) == 0)) {
                        s[i] = b;
                        s[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < kFitSample; i += 1024) keys[i] = s[i];
}

__global__ void k_make_xt(const uint32_t *__restrict__ keys, uint32_t n, uint32_t kmin, float scale,
                          float *__restrict__ x, float *__restrict__ t) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        x[i] = float(keys[i] - kmin) * scale;
        t[i] = (float(i) + 0.5f) / float(n);
    }
}

// ---------------------------------------------------------------------------
// MLP y = w3 . relu(W2 relu(W1 x + b1) + b2) + b3, loss = mean (y - t)^2
// ---------------------------------------------------------------------------
// Forward and backward of one tile of kGradTile samples per workgroup:
// threads [0, kGradTile) run one sample each and store h1, h2 and the
// back-propagated dz2 = dy * w3 * [h2 > 0], dz1 = (W2^T dz2) * [h1 > 0]
// (h > 0 <=> z > 0 for h = relu(z)); then every thread reduces a few of
// the 1153 parameter gradients over the tile and adds them to grad[] (a
// sum over tiles, so grad is the full-batch gradient).  loss_acc gains
// sum (y - t)^2.
__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ P, const float *__restrict__ X,
                                                  const float *__restrict__ T, uint32_t n,
                                                  float *__restrict__ grad, double *__restrict__ loss_acc) {
    __shared__ float sP[kMlpParams];
    __shared__ float sH1[kGradTile][kMlpH + 1], sH2[kGradTile][kMlpH + 1];
    __shared__ float sDZ1[kGradTile][kMlpH + 1], sDZ2[kGradTile][kMlpH + 1];
    __shared__ float sX[kGradTile], sDY[kGradTile];
    __shared__ double sL[4];
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) sP[i] = P[i];
    __syncthreads();
    const int s = threadIdx.x;
    double le = 0.0;
    if (s < kGradTile) {
        const uint32_t g = blockIdx.x * kGradTile + s;
        const bool valid = g < n;
        const float x = valid ? X[g] : 0.f, t = valid ? T[g] : 0.f;
        float h1[kMlpH], h2[kMlpH];
#pragma unroll
        for (int j = 0; j < kMlpH; j++) h1[j] = fmaxf(fmaf(sP[oW1 + j], x, sP[oB1 + j]), 0.f);
        float y = sP[oB3];
#pragma unroll
        for (int i = 0; i < kMlpH; i++) {
            float z = sP[oB2 + i];
#pragma unroll
            for (int j = 0; j < kMlpH; j++) z = fmaf(sP[oW2 + i * kMlpH + j], h1[j], z);
            h2[i] = fmaxf(z, 0.f);
            y = fmaf(sP[oW3 + i], h2[i], y);
        }
        const float e = y - t;
        const float dy = valid ? 2.f * e / float(n) : 0.f;
        le = valid ? double(e) * double(e) : 0.0;
        float dz1[kMlpH];
#pragma unroll
        for (int j = 0; j < kMlpH; j++) dz1[j] = 0.f;
#pragma unroll
        for (int i = 0; i < kMlpH; i++) {
            const float dz2 = h2[i] > 0.f ? dy * sP[oW3 + i] : 0.f;
            sDZ2[s][i] = dz2;
            sH2[s][i] = h2[i];
#pragma unroll
            for (int j = 0; j < kMlpH; j++) dz1[j] = fmaf(dz2, sP[oW2 + i * kMlpH + j], dz1[j]);
        }
#pragma unroll
        for (int j = 0; j < kMlpH; j++) {
            sH1[s][j] = h1[j];
            sDZ1[s][j] = h1[j] > 0.f ? dz1[j] : 0.f;
        }
        sX[s] = x;
        sDY[s] = dy;
    }
    // loss: wave sums, then one atomic per workgroup
    for (int off = 32; off > 0; off >>= 1) le += __shfl_xor(le, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) sL[threadIdx.x / kWave] = le;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss_acc, sL[0] + sL[1] + sL[2] + sL[3]);
    for (int e = threadIdx.x; e < kMlpParams; e += blockDim.x) {
        float acc = 0.f;
        if (e < oB1) {
            for (int q = 0; q < kGradTile; q++) acc = fmaf(sDZ1[q][e - oW1], sX[q], acc);
        } else if (e < oW2) {
            for (int q = 0; q < kGradTile; q++) acc += sDZ1[q][e - oB1];
        } else if (e < oB2) {
            const int i = (e - oW2) / kMlpH, j = (e - oW2) % kMlpH;
            for (int q = 0; q < kGradTile; q++) acc = fmaf(sDZ2[q][i], sH1[q][j], acc);
        } else if (e < oW3) {
            for (int q = 0; q < kGradTile; q++) acc += sDZ2[q][e - oB2];
        } else if (e < oB3) {
            for (int q = 0; q < kGradTile; q++) acc = fmaf(sDY[q], sH2[q][e - oW3], acc);
        } else {
            for (int q = 0; q < kGradTile; q++) acc += sDY[q];
        }
        atomicAdd(&grad[e], acc);
    }
}

// One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on every parameter;
// clears the gradient, records mean loss of the step and clears its sum.
__global__ __launch_bounds__(1024) void k_adam(float *__restrict__ P, float *__restrict__ G,
                                               float *__restrict__ M, float *__restrict__ V, float lr,
                                               float c1, float c2, double *__restrict__ loss_acc,
                                               float *__restrict__ loss_out, uint32_t n) {
    for (int e = threadIdx.x; e < kMlpParams; e += blockDim.x) {
        const float g = G[e];
        const float m = 0.9f * M[e] + 0.1f * g;
        const float v = 0.999f * V[e] + 0.001f * g * g;
        M[e] = m;
        V[e] = v;
        P[e] -= lr * (m / c1) / (sqrtf(v / c2) + 1e-8f);
        G[e] = 0.f;
    }
    if (threadIdx.x == 0) {
        *loss_out = float(*loss_acc / double(n));
        *loss_acc = 0.0;
    }
}

__device__ float mlp_eval(const float *__restrict__ P, float x) {
    float h1[kMlpH];
#pragma unroll
    for (int j = 0; j < kMlpH; j++) h1[j] = fmaxf(fmaf(P[oW1 + j], x, P[oB1 + j]), 0.f);
    float y = P[oB3];
    for (int i = 0; i < kMlpH; i++) {
        float z = P[oB2 + i];
#pragma unroll
        for (int j = 0; j < kMlpH; j++) z = fmaf(P[oW2 + i * kMlpH + j], h1[j], z);
        y = fmaf(P[oW3 + i], fmaxf(z, 0.f), y);
    }
    return y;
}

// Samples the trained MLP at x = i / kRootTable into the model's root table,
// clamped to [0, 1) and made non-decreasing (a running maximum).
__global__ __launch_bounds__(1024) void k_mlp_table(const float *__restrict__ P, RmiParams *__restrict__ m) {
    __shared__ float s[kRootTable + 1];
    for (int i = threadIdx.x; i <= kRootTable; i += blockDim.x)
        s[i] = fminf(fmaxf(mlp_eval(P, float(i) / float(kRootTable)), 0.f), 0.99999994f);
    __syncthreads();
    if (threadIdx.x == 0) {
        float run = 0.f;
        for (int i = 0; i <= kRootTable; i++) {
            run = fmaxf(run, s[i]);
            m->table[i] = run;
        }
    }
}

// Linear root by least squares of t on x (double sums in one workgroup).
__global__ __launch_bounds__(1024) void k_linear_fit(const float *__restrict__ X, const float *__restrict__ T,
                                                     uint32_t n, RmiParams *__restrict__ m) {
    __shared__ double s[4][1024 / kWave];
    double a[4] = {0, 0, 0, 0};
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const double x = X[i], t = T[i];
        a[0] += x;
        a[1] += t;
        a[2] += x * x;
        a[3] += x * t;
    }
    for (int k = 0; k < 4; k++) {
        for (int off = 32; off > 0; off >>= 1) a[k] += __shfl_xor(a[k], off, kWave);
        if ((threadIdx.x & (kWave - 1)) == 0) s[k][threadIdx.x / kWave] = a[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[4] = {0, 0, 0, 0};
        for (int k = 0; k < 4; k++)
            for (int w = 0; w < 1024 / kWave; w++) r[k] += s[k][w];
        const double mx = r[0] / n, mt = r[1] / n;
        const double var = r[2] / n - mx * mx, cov = r[3] / n - mx * mt;
        const double slope = var > 0 ? cov / var : 0.0;
        m->root_a = float(slope);
        m->root_b = float(mt - slope * mx);
    }
}

__device__ __forceinline__ int root_leaf(const RmiParams *__restrict__ m, float x) {
    float r;
    if (m->root_kind == ROOT_LINEAR) {
        r = fmaf(m->root_a, x, m->root_b);
    } else {
        const float u = fminf(fmaxf(x, 0.f), 1.f) * float(kRootTable);
        const int i = min(int(u), kRootTable - 1);
        r = fmaf(m->table[i + 1] - m->table[i], u - float(i), m->table[i]);
    }
    return int(fminf(fmaxf(r * float(kRmiLeaves), 0.f), float(kRmiLeaves - 1)));
}

// Per-leaf sums (n, x, t, xx, xt) of the samples the root routes there.
__global__ void k_leaf_stats(const float *__restrict__ X, const float *__restrict__ T, uint32_t n,
                             const RmiParams *__restrict__ m, double *__restrict__ st) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = X[i], t = T[i];
    double *q = st + 5 * root_leaf(m, X[i]);
    atomicAdd(q, 1.0);
    atomicAdd(q + 1, x);
    atomicAdd(q + 2, t);
    atomicAdd(q + 3, x * x);
    atomicAdd(q + 4, x * t);
}

__global__ void k_leaf_fit(const double *__restrict__ st, RmiParams *__restrict__ m) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= kRmiLeaves) return;
    const double *q = st + 5 * l;
    const double c = q[0];
    float a = 0.f, b = (float(l) + 0.5f) / float(kRmiLeaves);
    if (c >= 1.0) {
        const double mx = q[1] / c, mt = q[2] / c;
        const double var = q[3] / c - mx * mx, cov = q[4] / c - mx * mt;
        const double slope = (c >= 2.0 && var > 1e-18) ? cov / var : 0.0;
        a = float(slope);
        b = float(mt - slope * mx);
    }
    m->leaf_a[l] = a;
    m->leaf_b[l] = b;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// Each buffer is allocated on its own first use: the key sample alone is
// also needed without a fit (sampled heavy keys, drift judgement).
static int fit_buffers(sjb_ctx *c, uint32_t steps) {
    int rc;
    if ((!c->d_fit_keys && (rc = dev_alloc(c, (void **)&c->d_fit_keys, kFitSample * sizeof(uint32_t)))) ||
        (!c->d_fit_x && (rc = dev_alloc(c, (void **)&c->d_fit_x, kFitMax * sizeof(float)))) ||
        (!c->d_fit_t && (rc = dev_alloc(c, (void **)&c->d_fit_t, kFitMax * sizeof(float)))) ||
        (!c->d_mlp && (rc = dev_alloc(c, (void **)&c->d_mlp, 4 * kMlpParams * sizeof(float)))) ||
        (!c->d_fit_acc && (rc = dev_alloc(c, (void **)&c->d_fit_acc, (1 + 5 * kRmiLeaves) * sizeof(double)))))
        return rc;
    if (c->loss_hist_cap < steps + 1) {
        void *p = c->d_loss_hist;
        rc = dev_regrow(c, &p, c->d_loss_hist ? c->loss_hist_cap * sizeof(float) : 0, (steps + 1) * sizeof(float));
        c->d_loss_hist = static_cast<float *>(p);
        c->loss_hist_cap = c->d_loss_hist ? steps + 1 : 0;
        if (rc) return rc;
    }
    return SJB_OK;
}

// SplitMix64 stream for the MLP's initial weights.
struct InitRng {
    uint64_t s;
    double u01() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return (double(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    }
    double normal() { return sqrt(-2.0 * log(u01())) * cos(6.283185307179586 * u01()); }
};

// He-style initialisation for a 1-D input: first-layer kinks spread
// uniformly over x in [0, 1).
static void mlp_init(std::vector<float> &P, uint32_t seed) {
    InitRng r{0x5EEDULL * (seed + 1)};
    P.assign(kMlpParams, 0.f);
    for (int j = 0; j < kMlpH; j++) {
        const double w = r.normal() * sqrt(2.0);
        P[oW1 + j] = float(w);
        P[oB1 + j] = float(-w * r.u01());
    }
    for (int i = 0; i < kMlpH * kMlpH; i++) P[oW2 + i] = float(r.normal() * sqrt(2.0 / kMlpH));
    for (int i = 0; i < kMlpH; i++) {
        P[oB2 + i] = 0.01f;
        P[oW3 + i] = float(r.normal() * sqrt(1.0 / kMlpH));
    }
    P[oB3] = 0.f;
}

// Host restatement of rmi_partition (partition.hip) for the fit check.
static uint32_t host_partition(const RmiParams &m, uint32_t k, uint32_t F) {
    const float x = k < m.kmin ? 0.f : float(k - m.kmin) * m.scale;
    float r;
    if (m.root_kind == ROOT_LINEAR) {
        r = fmaf(m.root_a, x, m.root_b);
    } else {
        const float u = fminf(fmaxf(x, 0.f), 1.f) * float(kRootTable);
        const int i = std::min(int(u), kRootTable - 1);
        r = fmaf(m.table[i + 1] - m.table[i], u - float(i), m.table[i]);
    }
    r = fminf(fmaxf(r * float(kRmiLeaves), 0.f), float(kRmiLeaves - 1));
    float cdf = fmaf(m.leaf_a[int(r)], x, m.leaf_b[int(r)]);
    cdf = fminf(fmaxf(cdf * float(F), 0.f), float(F - 1));
    return uint32_t(cdf);
}

static double imbalance(const std::vector<uint32_t> &hist) {
    uint64_t mx = 0, tot = 0;
    for (uint32_t h : hist) {
        mx = std::max<uint64_t>(mx, h);
        tot += h;
    }
    return tot ? double(mx) * double(hist.size()) / double(tot) : 1.0;
}

// Heavy-key lookup table of m.heavy[0, n_heavy) (see RmiParams).
static void heavy_table_build(RmiParams &m) {
    for (int s = 0; s < 64; s++) {
        m.hslot_key[s] = 0;
        m.hslot_idx[s] = -1;
    }
    for (int w = 0; w < 128; w++) m.hfilter[w] = 0;
    for (uint32_t i = 0; i < m.n_heavy && i < uint32_t(kMaxHeavy); i++) {
        const uint32_t k = m.heavy[i];
        const uint32_t x = k * kHashMul;
        uint32_t s = x >> 26;
        while (m.hslot_idx[s] >= 0 && m.hslot_key[s] != k) s = (s + 1) & 63;
        m.hslot_key[s] = k;
        m.hslot_idx[s] = int32_t(i);
        m.hfilter[x >> 25] |= 1u << ((x >> 20) & 31);
    }
}


// Sample keys that the model partitions (heavy keys left out when peeled).
static std::vector<uint32_t> model_keys(const RmiParams &m, const std::vector<uint32_t> &keys, bool peel) {
    std::vector<uint32_t> rest;
    rest.reserve(keys.size());
    for (uint32_t k : keys) {
        bool heavy = false;
        for (uint32_t h = 0; peel && h < m.n_heavy; h++) heavy |= m.heavy[h] == k;
        if (!heavy) rest.push_back(k);
    }
    return rest;
}

// Partition imbalance (max / mean over Fv parts) of `rest` under model m
// and under MSD radix on the keys' range.
static void sample_imbalance(const RmiParams &m, const std::vector<uint32_t> &rest, double *learned,
                             double *radix) {
    const uint32_t Fv = 256;
    uint32_t kmin = rest[0], kmax = rest[0];
    for (uint32_t k : rest) {
        kmin = std::min(kmin, k);
        kmax = std::max(kmax, k);
    }
    const uint32_t span = kmax - kmin;
    const int bits = span ? 32 - __builtin_clz(span) : 0;
    const int shift = bits > 8 ? bits - 8 : 0;
    std::vector<uint32_t> hl(Fv, 0), hr(Fv, 0);
    for (uint32_t k : rest) {
        hl[host_partition(m, k, Fv)]++;
        hr[std::min<uint32_t>((k - kmin) >> shift, Fv - 1)]++;
    }
    *learned = imbalance(hl);
    *radix = imbalance(hr);
}

// Evenly spaced sample of up to kFitSample keys of T[0, n), copied to the
// host (uses the fit's key buffer; needs a fitted model).
int rmi_sample_keys(sjb_ctx *c, const sjb_tuple_t *T, uint64_t n, std::vector<uint32_t> &keys) {
    if (!c->d_fit_keys || n == 0) return SJB_E_INVALID;
    const uint32_t ns = uint32_t(std::min<uint64_t>(n, kFitSample));
    k_rmi_sample<<<(ns + 255) / 256, 256, 0, c->stream>>>(T, n, ns, c->d_fit_keys);
    SJB_HIP_TRY(hipGetLastError());
    keys.resize(ns);
    SJB_HIP_TRY(hipMemcpyAsync(keys.data(), c->d_fit_keys, size_t(ns) * 4, hipMemcpyDeviceToHost, c->stream));
    SJB_HIP_TRY(hipStreamSynchronize(c->stream));
    return SJB_OK;
}

// Heavy keys of a batch without a model (the radix baseline with peeling):
// the keys occurring at least kPeelMin times in an evenly spaced sample of
// up to kFitSample keys of T[0, n), the kMaxHeavy most frequent first, with
// their lookup table; no CDF (out->root_kind etc. unused).  Returns the
// number of heavy keys in *n_heavy.
constexpr uint32_t kPeelMin = 3;
int sampled_heavy(sjb_ctx *c, const sjb_tuple_t *T, uint64_t n, RmiParams *out, uint32_t *n_heavy) {
    std::memset(out, 0, sizeof(*out));
    *n_heavy = 0;
    if (n == 0) return SJB_OK;
    if (!c->d_fit_keys) {
        int rc = dev_alloc(c, (void **)&c->d_fit_keys, kFitSample * sizeof(uint32_t));
        if (rc) return rc;
    }
    std::vector<uint32_t> keys;
    int rc = rmi_sample_keys(c, T, n, keys);
    if (rc) return rc;
    std::sort(keys.begin(), keys.end());
    std::vector<std::pair<uint32_t, uint32_t>> runs;  // (count, key)
    for (size_t i = 0; i < keys.size();) {
        size_t j = i;
        while (j < keys.size() && keys[j] == keys[i]) j++;
        if (j - i >= kPeelMin) runs.emplace_back(uint32_t(j - i), keys[i]);
        i = j;
    }
    std::stable_sort(runs.begin(), runs.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
    if (runs.size() > size_t(kMaxHeavy)) runs.resize(kMaxHeavy);
    out->n_heavy = uint32_t(runs.size());
    for (uint32_t h = 0; h < out->n_heavy; h++) out->heavy[h] = runs[h].second;
    heavy_table_build(*out);
    *n_heavy = out->n_heavy;
    return SJB_OK;
}

// Refit of model m to the sorted keys `rest`, without training: the root
// becomes their empirical CDF over x (fraction of keys below x = i /
// kRootTable, interpolated linearly), and leaf l the chord of that root
// between the x where it reaches l / kRmiLeaves and the x where it reaches
// (l + 1) / kRmiLeaves, so the two stages follow the keys' quantiles to
// within 1 / kRmiLeaves.  Heavy keys and their tables are left as they are.
static void refit_to(RmiParams &m, const std::vector<uint32_t> &rest) {
    const size_t ns = rest.size();
    m.kmin = rest.front();
    m.scale = float(1.0 / (double(rest.back() - rest.front()) + 1.0));
    m.root_kind = ROOT_TABLE;
    std::vector<float> xs(ns);
    for (size_t j = 0; j < ns; j++) xs[j] = float(rest[j] - m.kmin) * m.scale;
    for (int i = 0; i <= kRootTable; i++) {
        const float xi = float(i) / float(kRootTable);
        const size_t below = size_t(std::lower_bound(xs.begin(), xs.end(), xi) - xs.begin());
        m.table[i] = float(double(below) / double(ns));
    }
    // x at which the (non-decreasing, piecewise linear) root reaches v;
    // called with non-decreasing v, so the segment search resumes
    int i = 0;
    auto root_inv = [&](double v) {
        if (v <= double(m.table[0])) return 0.0;
        while (i < kRootTable - 1 && double(m.table[i + 1]) < v) i++;
        const double lo = m.table[i], hi = m.table[i + 1];
        const double f = hi > lo ? std::min(1.0, std::max(0.0, (v - lo) / (hi - lo))) : 0.0;
        return (double(i) + f) / double(kRootTable);
    };
    for (int l = 0; l < kRmiLeaves; l++) {
        const double v0 = double(l) / kRmiLeaves, v1 = double(l + 1) / kRmiLeaves;
        const double x0 = root_inv(v0), x1 = root_inv(v1);
        if (x1 - x0 > 1e-12) {
            const double a = (v1 - v0) / (x1 - x0);
            m.leaf_a[l] = float(a);
            m.leaf_b[l] = float(v0 - a * x0);
        } else {
            m.leaf_a[l] = 0.f;
            m.leaf_b[l] = float((l + 0.5) / kRmiLeaves);
        }
    }
}

// Drift judgement of the learned CDF on the keys a batch will partition (a
// key sample; heavy keys left out when they are peeled).  The sample's odd
// half is scored (partition imbalance, max / mean over 256 parts) under the
// model, under MSD radix on its key range and, when refits are allowed,
// under a refit of the model to the even half.  Returns
//   kDriftRefit  the model is more than 1.5x as imbalanced as the refit,
//                and the refit within 1.25x of radix: *out = the model
//                refitted to the whole sample;
//   kDriftRanges the model (no refit allowed or the refit is no good) is
//                more than 1.25x as imbalanced as radix;
//   kDriftNone   otherwise (the model is used as it is).
int rmi_judge(const sjb_ctx *c, const std::vector<uint32_t> &keys, bool peel, bool allow_refit,
              RmiParams *out) {
    if (c->rmi_host.size() != sizeof(RmiParams) || keys.size() < 512) return kDriftNone;
    const RmiParams &m = *reinterpret_cast<const RmiParams *>(c->rmi_host.data());
    std::vector<uint32_t> rest = model_keys(m, keys, peel);
    if (rest.size() < 512) return kDriftNone;
    std::vector<uint32_t> even, odd;
    even.reserve(rest.size() / 2 + 1);
    odd.reserve(rest.size() / 2);
    for (size_t j = 0; j < rest.size(); j++) (j & 1 ? odd : even).push_back(rest[j]);
    double il, ir;
    sample_imbalance(m, odd, &il, &ir);
    if (allow_refit) {
        std::memcpy(out, &m, sizeof(RmiParams));
        std::sort(even.begin(), even.end());
        refit_to(*out, even);
        double ifit, unused;
        sample_imbalance(*out, odd, &ifit, &unused);
        if (il > 1.5 * ifit && ifit <= 1.25 * ir) {
            std::sort(rest.begin(), rest.end());
            refit_to(*out, rest);
            return kDriftRefit;
        }
    }
    return il > 1.25 * ir ? kDriftRanges : kDriftNone;
}
}  // namespace sjb

using namespace sjb;

extern "C" {

int sjb_rmi_fit(sjb_ctx_t *c, const sjb_tuple_t *d_train, uint64_t n_train, const sjb_rmi_config_t *cfg,
                sjb_rmi_report_t *rep) {
    if (!c || !cfg || !rep || n_train == 0 || !device_accessible(d_train)) return SJB_E_INVALID;
    if (cfg->root != SJB_RMI_LINEAR && cfg->root != SJB_RMI_MLP) return SJB_E_INVALID;
    if (cfg->root == SJB_RMI_MLP && (cfg->steps == 0 || cfg->steps > 100000 || !(cfg->lr > 0.f)))
        return SJB_E_INVALID;
    SJB_HIP_TRY(hipSetDevice(c->device));
    std::memset(rep, 0, sizeof(*rep));
    const uint32_t steps = cfg->root == SJB_RMI_MLP ? cfg->steps : 0;
    int rc = fit_buffers(c, steps);
    if (rc) return rc;
    c->rmi_valid = false;
    c->rmi_nheavy = 0;
    hipStream_t s = c->stream;
    SJB_HIP_TRY(hipEventRecord(c->ev0, s));

    // 1. sample + sort
    k_rmi_sample<<<kFitSample / 256, 256, 0, s>>>(d_train, n_train, kFitSample, c->d_fit_keys);
    k_sort_keys<<<1, 1024, 0, s>>>(c->d_fit_keys);
    SJB_HIP_TRY(hipGetLastError());
    std::vector<uint32_t> keys(kFitSample);
    SJB_HIP_TRY(hipMemcpyAsync(keys.data(), c->d_fit_keys, kFitSample * 4, hipMemcpyDeviceToHost, s));
    SJB_HIP_TRY(hipStreamSynchronize(s));

    // 2. heavy keys: the most frequent sample keys (runs of the sorted sample)
    std::vector<std::pair<uint32_t, uint32_t>> runs;  // (count, key)
    for (uint32_t i = 0; i < uint32_t(kFitSample);) {
        uint32_t j = i;
        while (j < uint32_t(kFitSample) && keys[j] == keys[i]) j++;
        if (cfg->heavy_min && j - i >= cfg->heavy_min) runs.emplace_back(j - i, keys[i]);
        i = j;
    }
    std::stable_sort(runs.begin(), runs.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
    if (runs.size() > size_t(kMaxHeavy)) runs.resize(kMaxHeavy);
    RmiParams hm;
    std::memset(&hm, 0, sizeof(hm));
    hm.n_heavy = uint32_t(runs.size());
    for (uint32_t h = 0; h < hm.n_heavy; h++) hm.heavy[h] = runs[h].second;
    heavy_table_build(hm);
    std::vector<uint32_t> rest;
    rest.reserve(kFitSample);
    for (uint32_t k : keys) {
        bool heavy = false;
        for (uint32_t h = 0; h < hm.n_heavy; h++) heavy |= hm.heavy[h] == k;
        if (!heavy) rest.push_back(k);
    }
    if (rest.size() < 64) return SJB_E_INVALID;  // nothing left to model
    const uint32_t ns = uint32_t(rest.size());
    hm.kmin = rest.front();
    hm.scale = float(1.0 / (double(rest.back() - rest.front()) + 1.0));
    hm.root_kind = cfg->root == SJB_RMI_MLP ? ROOT_TABLE : ROOT_LINEAR;
    // header (kmin, scale, root_kind, n_heavy) and heavy keys
    SJB_HIP_TRY(hipMemcpyAsync(c->d_rmi, &hm, 16, hipMemcpyHostToDevice, s));
    SJB_HIP_TRY(hipMemcpyAsync(c->d_rmi->heavy, hm.heavy, sizeof(hm.heavy), hipMemcpyHostToDevice, s));
    SJB_HIP_TRY(hipMemcpyAsync(c->d_rmi->hslot_key, hm.hslot_key,
                               sizeof(hm.hslot_key) + sizeof(hm.hslot_idx) + sizeof(hm.hfilter),
                               hipMemcpyHostToDevice, s));
    SJB_HIP_TRY(hipMemcpyAsync(c->d_fit_keys, rest.data(), size_t(ns) * 4, hipMemcpyHostToDevice, s));
    k_make_xt<<<(ns + 255) / 256, 256, 0, s>>>(c->d_fit_keys, ns, hm.kmin, hm.scale, c->d_fit_x, c->d_fit_t);
    SJB_HIP_TRY(hipGetLastError());

    // 4. root
    float *P = c->d_mlp, *G = P + kMlpParams, *M = G + kMlpParams, *V = M + kMlpParams;
    std::vector<float> init;  // outlives the asynchronous upload (synchronised below)
    if (cfg->root == SJB_RMI_LINEAR) {
        k_linear_fit<<<1, 1024, 0, s>>>(c->d_fit_x, c->d_fit_t, ns, c->d_rmi);
    } else {
        mlp_init(init, cfg->seed);
        SJB_HIP_TRY(hipMemcpyAsync(P, init.data(), kMlpParams * 4, hipMemcpyHostToDevice, s));
        if ((rc = launch_fill_u32(s, reinterpret_cast<uint32_t *>(G), 0u, 3 * kMlpParams))) return rc;
        if ((rc = launch_fill_u64(s, reinterpret_cast<uint64_t *>(c->d_fit_acc), 0, 1))) return rc;
        const uint32_t blocks = (ns + kGradTile - 1) / kGradTile;
        double b1 = 1.0, b2 = 1.0;
        for (uint32_t st = 1; st <= steps; st++) {
            b1 *= 0.9;
            b2 *= 0.999;
            // cosine decay to 10% of lr
            const double lr = cfg->lr * (0.1 + 0.45 * (1.0 + cos(3.141592653589793 * st / steps)));
            k_mlp_grad<<<blocks, 256, 0, s>>>(P, c->d_fit_x, c->d_fit_t, ns, G, c->d_fit_acc);
            k_adam<<<1, 1024, 0, s>>>(P, G, M, V, float(lr), float(1.0 - b1), float(1.0 - b2), c->d_fit_acc,
                                      c->d_loss_hist + (st - 1), ns);
        }
        // loss after the last update
        k_mlp_grad<<<blocks, 256, 0, s>>>(P, c->d_fit_x, c->d_fit_t, ns, G, c->d_fit_acc);
        k_adam<<<1, 1024, 0, s>>>(P, G, M, V, 0.f, 1.f, 1.f, c->d_fit_acc, c->d_loss_hist + steps, ns);
        k_mlp_table<<<1, 1024, 0, s>>>(P, c->d_rmi);
    }
    SJB_HIP_TRY(hipGetLastError());

    // 5. leaves
    double *st = c->d_fit_acc + 1;
    if ((rc = launch_fill_u64(s, reinterpret_cast<uint64_t *>(st), 0, 5 * kRmiLeaves))) return rc;
    k_leaf_stats<<<(ns + 255) / 256, 256, 0, s>>>(c->d_fit_x, c->d_fit_t, ns, c->d_rmi, st);
    k_leaf_fit<<<kRmiLeaves / 256, 256, 0, s>>>(st, c->d_rmi);
    SJB_HIP_TRY(hipGetLastError());
    SJB_HIP_TRY(hipEventRecord(c->ev1, s));

    // 6. check against radix on the sample
    RmiParams m;
    SJB_HIP_TRY(hipMemcpyAsync(&m, c->d_rmi, sizeof(m), hipMemcpyDeviceToHost, s));
    if (steps) {
        std::vector<float> lh(steps + 1);
        SJB_HIP_TRY(hipMemcpyAsync(lh.data(), c->d_loss_hist, (steps + 1) * 4, hipMemcpyDeviceToHost, s));
        SJB_HIP_TRY(hipStreamSynchronize(s));
        rep->loss_init = lh[0];
        rep->loss_final = lh[steps];
    }
    SJB_HIP_TRY(hipStreamSynchronize(s));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    const uint32_t Fv = std::max<uint32_t>(16, std::min<uint32_t>(1024, ns / 64));
    std::vector<uint32_t> hl(Fv, 0), hr(Fv, 0);
    uint32_t used = 0;
    {
        std::vector<uint8_t> leaf_used(kRmiLeaves, 0);
        const uint32_t span = rest.back() - rest.front();
        const int bits = span ? 32 - __builtin_clz(span) : 0;
        int lf = 0;
        while ((1u << (lf + 1)) <= Fv) lf++;
        const int shift = bits > lf ? bits - lf : 0;
        for (uint32_t k : rest) {
            hl[host_partition(m, k, Fv)]++;
            hr[std::min<uint32_t>((k - rest.front()) >> shift, (1u << lf) - 1)]++;
        }
        hr.resize(1u << lf);
        for (uint32_t i = 0; i < kRmiLeaves; i++) leaf_used[i] = m.leaf_a[i] != 0.f;
        for (uint8_t u : leaf_used) used += u;
    }
    rep->imbalance_learned = imbalance(hl);
    rep->imbalance_radix = imbalance(hr);
    rep->fit_ms = ms;
    rep->n_heavy = m.n_heavy;
    rep->leaves_used = used;
    rep->n_sample = ns;
    rep->fallback = rep->imbalance_learned > 1.25 * rep->imbalance_radix ? 1u : 0u;
    c->rmi_valid = rep->fallback == 0;
    c->rmi_nheavy = c->rmi_valid ? m.n_heavy : 0;
    c->rmi_host.resize(sizeof(RmiParams));
    std::memcpy(c->rmi_host.data(), &m, sizeof(RmiParams));
    return SJB_OK;
}


int sjb_rmi_clear(sjb_ctx_t *c) {
    if (!c) return SJB_E_INVALID;
    c->rmi_valid = false;
    c->rmi_nheavy = 0;
    return SJB_OK;
}

int sjb_mlp_grad(sjb_ctx_t *c, const float *params, const float *x, const float *t, uint32_t n,
                 float *grad, double *loss) {
    if (!c || !params || !x || !t || !grad || n == 0 || n > uint32_t(kFitMax)) return SJB_E_INVALID;
    SJB_HIP_TRY(hipSetDevice(c->device));
    int rc = fit_buffers(c, 1);
    if (rc) return rc;
    hipStream_t s = c->stream;
    float *P = c->d_mlp, *G = P + kMlpParams;
    SJB_HIP_TRY(hipMemcpyAsync(P, params, kMlpParams * 4, hipMemcpyHostToDevice, s));
    SJB_HIP_TRY(hipMemcpyAsync(c->d_fit_x, x, size_t(n) * 4, hipMemcpyHostToDevice, s));
    SJB_HIP_TRY(hipMemcpyAsync(c->d_fit_t, t, size_t(n) * 4, hipMemcpyHostToDevice, s));
    if ((rc = launch_fill_u32(s, reinterpret_cast<uint32_t *>(G), 0u, kMlpParams))) return rc;
    if ((rc = launch_fill_u64(s, reinterpret_cast<uint64_t *>(c->d_fit_acc), 0, 1))) return rc;
    k_mlp_grad<<<(n + kGradTile - 1) / kGradTile, 256, 0, s>>>(P, c->d_fit_x, c->d_fit_t, n, G, c->d_fit_acc);
    SJB_HIP_TRY(hipGetLastError());
    double acc = 0;
    SJB_HIP_TRY(hipMemcpyAsync(grad, G, kMlpParams * 4, hipMemcpyDeviceToHost, s));
    SJB_HIP_TRY(hipMemcpyAsync(&acc, c->d_fit_acc, 8, hipMemcpyDeviceToHost, s));
    SJB_HIP_TRY(hipStreamSynchronize(s));
    if (loss) *loss = acc / double(n);
    return SJB_OK;
}

}  // extern "C"
