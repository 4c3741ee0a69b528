// Host buffers the device can copy from asynchronously: page-locked memory
// from sjb_host_alloc (the reference allocates its ring buffers and sink with
// cudaMallocHost, include/sjb/windowing/lazy/CountBased.cuh:35-36,
// src/sjb/sink/Sink.cpp:8-9).  Without a device (CPU-only runs) the memory
// comes from calloc instead; either way it is zeroed.
#pragma once
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>

#include "sjb.h"

namespace sjb_host {

inline void *pinnedAlloc(uint64_t bytes, bool *pinned) {
    void *p = nullptr;
    if (sjb_host_alloc(bytes, &p) == SJB_OK && p) {
        std::memset(p, 0, bytes);
        *pinned = true;
        return p;
    }
    *pinned = false;
    p = std::calloc(bytes ? bytes : 1, 1);
    if (!p) throw std::bad_alloc();
    return p;
}

inline void pinnedFree(void *p, bool pinned) {
    if (!p) return;
    if (pinned)
        sjb_host_free(p);
    else
        std::free(p);
}

// n zeroed elements of T in page-locked memory when a device is present
template <class T>
class PinnedArray {
public:
    explicit PinnedArray(uint64_t n) : n_(n) { p_ = static_cast<T *>(pinnedAlloc(n * sizeof(T), &pinned_)); }
    ~PinnedArray() { pinnedFree(p_, pinned_); }
    PinnedArray(const PinnedArray &) = delete;
    PinnedArray &operator=(const PinnedArray &) = delete;
    T *data() { return p_; }
    const T *data() const { return p_; }
    uint64_t size() const { return n_; }
    bool pinned() const { return pinned_; }

private:
    T *p_ = nullptr;
    uint64_t n_;
    bool pinned_ = false;
};

}  // namespace sjb_host
