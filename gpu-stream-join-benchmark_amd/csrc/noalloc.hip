// noalloc.hip — sjb_lazy_join_noalloc: one HJ / SMJ window (count and
// checksums) that allocates no memory.
//
// It runs the same partition, join and reduction kernels as
// sjb_lazy_join_async, but their scratch memory (partition records, chunk
// offsets, per-partition partials, the key sample's range bounds) is carved out
// of one file-scope device array, `scratch`, instead of a context's
// hipMalloc'd buffers.  Each GPU holds its own copy of the array for the
// life of the process; its device address is looked up once per device with
// hipGetSymbolAddress and cached.  The kernels clear what must start at zero
// themselves (no memset nodes), so the call can be captured in a HIP graph.
// Calls that overlap on one GPU share the array and race.
#include <string.h>

#include <mutex>

#include "sjb_device.hpp"
#include "sjb_internal.hpp"

namespace sjb {

namespace {

constexpr uint64_t kNaMax = SJB_NOALLOC_MAX_TUPLES;  // tuples per side
constexpr uint64_t kNaChunks = (kNaMax + kChunk - 1) / kChunk;
constexpr uint64_t kNaOffsWords = ((uint64_t(1) << kMaxLogF) + 1) * 2 * kNaChunks;
constexpr uint64_t kNaPartials = uint64_t(1) << kMaxLogF;  // one partial (3 u64) per join workgroup
constexpr uint64_t kNaMinmaxWords = 4096;

constexpr uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }
constexpr uint64_t kOffPart = 0;
constexpr uint64_t kOffOffs = align256(kOffPart + 2 * kNaMax * sizeof(Rec));
constexpr uint64_t kOffPartials = align256(kOffOffs + kNaOffsWords * sizeof(uint32_t));
constexpr uint64_t kOffMinmax = align256(kOffPartials + 3 * kNaPartials * sizeof(uint64_t));
constexpr uint64_t kNaBytes = align256(kOffMinmax + kNaMinmaxWords * sizeof(uint32_t));

}  // namespace

// The scratch array.  In host code this name is a stand-in for the symbol,
// not a device address: only the address from hipGetSymbolAddress is used.
static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {

constexpr int kNaMaxDevices = 64;
struct NaDevice {
    unsigned char *ptr = nullptr;  // device address of `scratch` on this device
    int n_cu = 0;
    bool ready = false;
};
NaDevice na_dev[kNaMaxDevices];
std::mutex na_mtx;

// Once per device (that device current): the array's address, the CU
// count, the gfx950 check and the partition kernels' LDS limit.
int na_setup(int dev, NaDevice **out) {
    std::lock_guard<std::mutex> g(na_mtx);
    NaDevice &d = na_dev[dev];
    if (!d.ready) {
        hipDeviceProp_t prop;
        SJB_HIP_TRY(hipGetDeviceProperties(&prop, dev));
        if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return SJB_E_NODEVICE;
        void *p = nullptr;
        SJB_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return SJB_E_HIP;
        d.ptr = static_cast<unsigned char *>(p);
        d.n_cu = prop.multiProcessorCount;
        partition_setup(nullptr);
        d.ready = true;
    }
    *out = &d;
    return SJB_OK;
}

}  // namespace

}  // namespace sjb

using namespace sjb;

extern "C" int sjb_lazy_join_noalloc(int algorithm, uint32_t flags, const sjb_tuple_t *d_left,
                                     uint64_t n_left, const sjb_tuple_t *d_right, uint64_t n_right,
                                     sjb_window_result_t *d_result, void *stream) {
    if (algorithm != SJB_ALG_HJ && algorithm != SJB_ALG_SMJ) return SJB_E_UNSUPPORTED;
    if (flags & ~uint32_t(SJB_F_C2 | SJB_F_GROUPED)) return SJB_E_UNSUPPORTED;
    if (!d_result || (n_left && !d_left) || (n_right && !d_right)) return SJB_E_INVALID;
    if (n_left > kNaMax || n_right > kNaMax) return SJB_E_CAPACITY;
    int dev = 0, ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return SJB_E_NODEVICE;
    }
    SJB_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaMaxDevices) return SJB_E_UNSUPPORTED;
    NaDevice *d = nullptr;
    int rc = na_setup(dev, &d);
    if (rc) return rc;
    // A view of a context whose buffers are slices of this device's array:
    // only the fields the HJ / SMJ count-and-checksum window path reads are
    // set (no pair output, no learned model, no overlap, no timing).  It
    // owns nothing and is never passed to sjb_close.
    sjb_ctx c;
    c.device = dev;
    c.stream = static_cast<hipStream_t>(stream);
    c.n_cu = d->n_cu;
    c.cap = kNaMax;
    c.offs_cap = kNaOffsWords;
    c.small_chunks = false;  // the offsets slice holds kChunk chunks only
    c.d_part = reinterpret_cast<Rec *>(d->ptr + kOffPart);
    c.d_offs = reinterpret_cast<uint32_t *>(d->ptr + kOffOffs);
    c.d_partials = reinterpret_cast<uint64_t *>(d->ptr + kOffPartials);
    c.d_scratch = reinterpret_cast<uint32_t *>(d->ptr + kOffMinmax);
    int logF = 0;
    return window_enqueue(&c, algorithm, flags, d_left, n_left, d_right, n_right, 0, nullptr, 0, 0, d_result,
                          &logF, nullptr);
}
