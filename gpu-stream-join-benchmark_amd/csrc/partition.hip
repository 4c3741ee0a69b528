// partition.hip — bounded-fanout partitioning of both window sides.
//
// Replaces the reference's D-bucket counting sort (computeHistogram +
// host prefix sum + buildHashTable, src/sjb/windowing/lazy/executor/GPU/
// GPU_HJ.cu:9-39 and GPU_HJ_NoOutput.cu:53-76; smjHistogramKernel +
// thrust::exclusive_scan + smjSortKernel, GPU_SMJ.cu:9-32) with a
// fanout F = 2^logF that is independent of the key range:
//   1. k_hist    one workgroup per 16384-tuple chunk builds an LDS
//                histogram over F partitions and stores it partition-major
//                (hist[p * nchunks + chunk]),
//   2. k_scan_*  exclusive scan of that matrix on the device (no host round
//                trip): entry (p, chunk) becomes the chunk's first slot in
//                partition p,
//   3. k_scatter re-reads the chunk and writes 8-byte records (key,
//                payload) to their slots; ranks inside a chunk come from
//                LDS atomics.
// Both sides go through one launch of each kernel: blocks [0, ncL) take
// left chunks and [ncL, ncL+ncR) right chunks, and the scan runs over the
// concatenated matrix, so left records land in d_part[0, nL) and right
// records in d_part[nL, nL+nR).
#include <algorithm>

#include "sjb_device.hpp"
#include "sjb_internal.hpp"

namespace sjb {

constexpr int kItems = kChunk / kPartThreads;  // 32 tuples per thread
constexpr int kScanThreads = 512;
constexpr int kScanItems = 8;
constexpr int kScanSeg = kScanThreads * kScanItems;  // 4096 counters per block

// Partition function of one key.  HASH: top logF bits of the Fibonacci hash
// (bijective, so equal keys and only equal keys collide before truncation).
// RANGE: (key - kmin) >> shift, i.e. MSD radix on the key digits above the
// window's minimum.  LEARNED: floor(F * cdf(key)) from the RMI.
template <int KIND>
struct PartFn {
    uint32_t logF, fmask, kmin, shift;
    const RmiParams *rmi;
    __device__ __forceinline__ void init(int logF_, const PartParams *pp, const RmiParams *rmi_) {
        logF = logF_;
        fmask = (1u << logF_) - 1u;
        rmi = rmi_;
        if (KIND == PART_RANGE) {
            kmin = pp->kmin;
            shift = pp->shift;
        } else {
            kmin = 0;
            shift = 32 - logF_;
        }
    }
    __device__ __forceinline__ uint32_t operator()(uint32_t k) const {
        if (KIND == PART_HASH) return key_hash(k) >> shift;
        if (KIND == PART_RANGE) {
            uint32_t p = (k - kmin) >> shift;
            return p > fmask ? fmask : p;
        }
        // LEARNED
        float x = float(k);
        float l = fmaf(rmi->root_a, x, rmi->root_b);
        int leaf = int(l);
        leaf = leaf < 0 ? 0 : (leaf >= kRmiLeaves ? kRmiLeaves - 1 : leaf);
        float cdf = fmaf(rmi->leaf_a[leaf], x, rmi->leaf_b[leaf]);
        int p = int(cdf * float(fmask + 1));
        p = p < 0 ? 0 : (p > int(fmask) ? int(fmask) : p);
        return uint32_t(p);
    }
};

// ---------------------------------------------------------------------------
// min / max key over both sides (range partitioning parameters)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kPartThreads) void k_minmax_partial(
    const sjb_tuple_t *__restrict__ L, uint64_t nL, const sjb_tuple_t *__restrict__ R,
    uint64_t nR, uint32_t *__restrict__ partial) {
    __shared__ uint32_t smin[kPartThreads / kWave], smax[kPartThreads / kWave];
    uint32_t mn = 0xFFFFFFFFu, mx = 0;
    const uint64_t n = nL + nR;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x) {
        uint32_t k = i < nL ? L[i].key : R[i - nL].key;
        mn = min(mn, k);
        mx = max(mx, k);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mn = min(mn, (uint32_t)__shfl_xor((int)mn, off, kWave));
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, kWave));
    }
    const int wid = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        smin[wid] = mn;
        smax[wid] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kPartThreads / kWave; w++) {
            mn = min(mn, smin[w]);
            mx = max(mx, smax[w]);
        }
        mn = min(mn, smin[0]);
        mx = max(mx, smax[0]);
        partial[2 * blockIdx.x] = mn;
        partial[2 * blockIdx.x + 1] = mx;
    }
}

__global__ void k_minmax_final(const uint32_t *__restrict__ partial, int nblocks, int logF,
                               PartParams *__restrict__ pp) {
    if (threadIdx.x != 0) return;
    uint32_t mn = 0xFFFFFFFFu, mx = 0;
    for (int b = 0; b < nblocks; b++) {
        mn = min(mn, partial[2 * b]);
        mx = max(mx, partial[2 * b + 1]);
    }
    if (mn > mx) mn = mx = 0;
    uint32_t span = mx - mn;
    int bits = span ? 32 - __clz(span) : 0;
    pp->kmin = mn;
    pp->shift = bits > logF ? uint32_t(bits - logF) : 0u;
}

int launch_minmax(sjb_ctx *c, const sjb_tuple_t *L, uint64_t nL, const sjb_tuple_t *R,
                  uint64_t nR, int logF) {
    uint64_t n = nL + nR;
    int blocks = int(std::min<uint64_t>((n + kChunk - 1) / kChunk, 1024));
    if (blocks < 1) blocks = 1;
    // The scan-block buffer doubles as scratch for the partial min/max.
    k_minmax_partial<<<blocks, kPartThreads, 0, c->stream>>>(L, nL, R, nR, c->d_scan_blocks);
    SJB_HIP_TRY(hipGetLastError());
    k_minmax_final<<<1, 64, 0, c->stream>>>(c->d_scan_blocks, blocks, logF, c->d_pparams);
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

// ---------------------------------------------------------------------------
// histogram
// ---------------------------------------------------------------------------
template <int KIND>
__global__ __launch_bounds__(kPartThreads) void k_hist(
    const sjb_tuple_t *__restrict__ L, uint64_t nL, int ncL, const sjb_tuple_t *__restrict__ R,
    uint64_t nR, int ncR, int logF, const PartParams *__restrict__ pp,
    const RmiParams *__restrict__ rmi, uint32_t *__restrict__ hist) {
    extern __shared__ uint32_t lds_hist[];
    const int F = 1 << logF;
    PartFn<KIND> part;
    part.init(logF, pp, rmi);
    const bool left = int(blockIdx.x) < ncL;
    const int chunk = left ? blockIdx.x : blockIdx.x - ncL;
    const sjb_tuple_t *in = left ? L : R;
    const uint64_t n = left ? nL : nR;
    const int nc = left ? ncL : ncR;
    uint32_t *out = left ? hist : hist + uint64_t(F) * ncL;

    for (int p = threadIdx.x; p < F; p += kPartThreads) lds_hist[p] = 0;
    __syncthreads();
    const uint64_t base = uint64_t(chunk) * kChunk;
#pragma unroll 8
    for (int j = 0; j < kItems; j++) {
        uint64_t i = base + uint64_t(j) * kPartThreads + threadIdx.x;
        if (i < n) atomicAdd(&lds_hist[part(in[i].key)], 1u);
    }
    __syncthreads();
    for (int p = threadIdx.x; p < F; p += kPartThreads) out[uint64_t(p) * nc + chunk] = lds_hist[p];
}

// ---------------------------------------------------------------------------
// exclusive scan of a u32 array (reduce -> scan of block sums -> apply)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t *__restrict__ a,
                                                              uint64_t n,
                                                              uint32_t *__restrict__ block_sums) {
    __shared__ uint32_t sw[kScanThreads / kWave];
    const uint64_t base = uint64_t(blockIdx.x) * kScanSeg;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        uint64_t i = base + uint64_t(j) * kScanThreads + threadIdx.x;
        if (i < n) s += a[i];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) sw[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kScanThreads / kWave; w++) t += sw[w];
        block_sums[blockIdx.x] = t;
    }
}

// Single workgroup: exclusive scan of nb block sums in place.
__global__ __launch_bounds__(1024) void k_scan_top(uint32_t *__restrict__ bs, int nb) {
    __shared__ uint32_t wsum[1024 / kWave];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        int i = base + threadIdx.x;
        uint32_t v = i < nb ? bs[i] : 0;
        uint32_t tot;
        uint32_t ex = wave_excl_scan(v, &tot);
        if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x / kWave] = tot;
        __syncthreads();
        uint32_t wpre = 0, all = 0;
        for (int w = 0; w < 1024 / kWave; w++) {
            if (w < int(threadIdx.x / kWave)) wpre += wsum[w];
            all += wsum[w];
        }
        if (i < nb) bs[i] = carry + wpre + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry += all;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(uint32_t *__restrict__ a, uint64_t n,
                                                             const uint32_t *__restrict__ block_sums) {
    __shared__ uint32_t wsum[kScanThreads / kWave];
    const uint64_t base = uint64_t(blockIdx.x) * kScanSeg;
    // Each thread owns kScanItems consecutive counters.
    uint32_t v[kScanItems];
    uint32_t local = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        uint64_t i = base + uint64_t(threadIdx.x) * kScanItems + j;
        v[j] = i < n ? a[i] : 0;
        local += v[j];
    }
    uint32_t tot;
    uint32_t ex = wave_excl_scan(local, &tot);
    if ((threadIdx.x & (kWave - 1)) == 0) wsum[threadIdx.x / kWave] = tot;
    __syncthreads();
    uint32_t wpre = 0;
    for (int w = 0; w < int(threadIdx.x / kWave); w++) wpre += wsum[w];
    uint32_t run = block_sums[blockIdx.x] + wpre + ex;
#pragma unroll
    for (int j = 0; j < kScanItems; j++) {
        uint64_t i = base + uint64_t(threadIdx.x) * kScanItems + j;
        if (i < n) a[i] = run;
        run += v[j];
    }
}

static int scan_inplace(sjb_ctx *c, uint32_t *a, uint64_t n) {
    int nb = int((n + kScanSeg - 1) / kScanSeg);
    k_scan_reduce<<<nb, kScanThreads, 0, c->stream>>>(a, n, c->d_scan_blocks);
    SJB_HIP_TRY(hipGetLastError());
    k_scan_top<<<1, 1024, 0, c->stream>>>(c->d_scan_blocks, nb);
    SJB_HIP_TRY(hipGetLastError());
    k_scan_apply<<<nb, kScanThreads, 0, c->stream>>>(a, n, c->d_scan_blocks);
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

// ---------------------------------------------------------------------------
// scatter
// ---------------------------------------------------------------------------
template <int KIND, int PAY>
__global__ __launch_bounds__(kPartThreads) void k_scatter(
    const sjb_tuple_t *__restrict__ L, uint64_t nL, int ncL, const sjb_tuple_t *__restrict__ R,
    uint64_t nR, int ncR, int logF, const PartParams *__restrict__ pp,
    const RmiParams *__restrict__ rmi, const uint32_t *__restrict__ hist, Rec *__restrict__ part_out) {
    extern __shared__ uint32_t lds_cur[];
    const int F = 1 << logF;
    PartFn<KIND> part;
    part.init(logF, pp, rmi);
    const bool left = int(blockIdx.x) < ncL;
    const int chunk = left ? blockIdx.x : blockIdx.x - ncL;
    const sjb_tuple_t *in = left ? L : R;
    const uint64_t n = left ? nL : nR;
    const int nc = left ? ncL : ncR;
    const uint32_t *h = left ? hist : hist + uint64_t(F) * ncL;

    for (int p = threadIdx.x; p < F; p += kPartThreads) lds_cur[p] = h[uint64_t(p) * nc + chunk];
    __syncthreads();
    const uint64_t base = uint64_t(chunk) * kChunk;
#pragma unroll 8
    for (int j = 0; j < kItems; j++) {
        uint64_t i = base + uint64_t(j) * kPartThreads + threadIdx.x;
        if (i < n) {
            const uint2 kv = *reinterpret_cast<const uint2 *>(&in[i]);
            uint32_t pos = atomicAdd(&lds_cur[part(kv.x)], 1u);
            part_out[pos] = make_uint2(kv.x, PAY == PAY_VAL ? kv.y : uint32_t(i));
        }
    }
}

int choose_logF(int alg, uint64_t n_build, int requested) {
    if (requested > 0) return std::min(std::max(requested, 1), kMaxLogF);
    // Aim for half-full LDS tiles: expected partition = cap / 2.
    const uint64_t target = alg == SJB_ALG_SMJ ? kSmjCap / 2 : kHjCap / 2;
    int logF = 1;
    while (logF < kMaxLogF && (n_build >> logF) > target) logF++;
    return logF;
}

template <int KIND>
static int partition_kind(sjb_ctx *c, const WindowPlan &p, const sjb_tuple_t *L,
                          const sjb_tuple_t *R) {
    const int F = 1 << p.logF;
    const size_t lds = size_t(F) * sizeof(uint32_t);
    const int nblk = p.ncL + p.ncR;
    k_hist<KIND><<<nblk, kPartThreads, lds, c->stream>>>(L, p.nL, p.ncL, R, p.nR, p.ncR, p.logF,
                                                         c->d_pparams, c->d_rmi, c->d_hist);
    SJB_HIP_TRY(hipGetLastError());
    int rc = scan_inplace(c, c->d_hist, uint64_t(F) * nblk);
    if (rc) return rc;
    if (p.payload == PAY_VAL)
        k_scatter<KIND, PAY_VAL><<<nblk, kPartThreads, lds, c->stream>>>(
            L, p.nL, p.ncL, R, p.nR, p.ncR, p.logF, c->d_pparams, c->d_rmi, c->d_hist, c->d_part);
    else
        k_scatter<KIND, PAY_IDX><<<nblk, kPartThreads, lds, c->stream>>>(
            L, p.nL, p.ncL, R, p.nR, p.ncR, p.logF, c->d_pparams, c->d_rmi, c->d_hist, c->d_part);
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

int launch_partition(sjb_ctx *c, const WindowPlan &p, const sjb_tuple_t *L, const sjb_tuple_t *R) {
    if (uint64_t(1u << p.logF) * uint64_t(p.ncL + p.ncR) > c->hist_cap) return SJB_E_CAPACITY;
    switch (p.kind) {
        case PART_HASH: return partition_kind<PART_HASH>(c, p, L, R);
        case PART_RANGE: return partition_kind<PART_RANGE>(c, p, L, R);
        case PART_LEARNED: return partition_kind<PART_LEARNED>(c, p, L, R);
    }
    return SJB_E_INVALID;
}

// ---------------------------------------------------------------------------
// fills (used instead of hipMemsetAsync so a captured graph never holds a
// memset node)
// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u64(uint64_t *p, uint64_t v, uint64_t n) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

int launch_fill_u32(hipStream_t s, uint32_t *p, uint32_t v, uint64_t n) {
    if (n == 0) return SJB_OK;
    int blocks = int(std::min<uint64_t>((n + 255) / 256, 4096));
    k_fill_u32<<<blocks, 256, 0, s>>>(p, v, n);
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}
int launch_fill_u64(hipStream_t s, uint64_t *p, uint64_t v, uint64_t n) {
    if (n == 0) return SJB_OK;
    int blocks = int(std::min<uint64_t>((n + 255) / 256, 4096));
    k_fill_u64<<<blocks, 256, 0, s>>>(p, v, n);
    SJB_HIP_TRY(hipGetLastError());
    return SJB_OK;
}

}  // namespace sjb
