// batch.hip — many uneven windows per call (count/checksums).
//
// The reference joins one window per executor call
// (include/sjb/windowing/lazy/CountBased.cuh:135-138).  sjb_lazy_join_batch
// takes n_windows windows at once; every window is enqueued back to back on
// the ctx stream with no host sync, results land in a device array that is
// copied back once.
#include <algorithm>

#include "sjb_device.hpp"
#include "sjb_internal.hpp"

namespace sjb {

int batch_join(sjb_ctx *c, int alg, uint32_t flags, const sjb_tuple_t *L, const uint64_t *loff,
               const sjb_tuple_t *R, const uint64_t *roff, uint32_t n_windows,
               sjb_window_result_t *results) {
    if (n_windows == 0) return SJB_OK;
    for (uint32_t w = 0; w < n_windows; w++) {
        if (loff[w + 1] < loff[w] || roff[w + 1] < roff[w]) return SJB_E_INVALID;
        if (loff[w + 1] - loff[w] > c->cap || roff[w + 1] - roff[w] > c->cap) return SJB_E_CAPACITY;
    }
    SJB_HIP_TRY(hipSetDevice(c->device));
    if ((loff[n_windows] && !device_accessible(L)) || (roff[n_windows] && !device_accessible(R)))
        return SJB_E_INVALID;
    if (c->batch_res_cap < n_windows) {
        void *p = nullptr;
        int rc = dev_alloc(c, &p, size_t(n_windows) * sizeof(sjb_window_result_t));
        if (rc) return rc;
        c->d_batch_res = static_cast<sjb_window_result_t *>(p);
        c->batch_res_cap = n_windows;
    }
    for (uint32_t w = 0; w < n_windows; w++) {
        int logF;
        int rc = window_enqueue(c, alg, flags, L + loff[w], loff[w + 1] - loff[w], R + roff[w],
                                roff[w + 1] - roff[w], 0, nullptr, 0, 0, c->d_batch_res + w, &logF,
                                false);
        if (rc) return rc;
    }
    SJB_HIP_TRY(hipMemcpyAsync(results, c->d_batch_res, size_t(n_windows) * sizeof(sjb_window_result_t),
                               hipMemcpyDeviceToHost, c->stream));
    SJB_HIP_TRY(hipStreamSynchronize(c->stream));
    return SJB_OK;
}

}  // namespace sjb
