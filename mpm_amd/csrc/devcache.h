// devcache.h -- per-device cache of released device blocks.
//
// hipFree waits until every stream of the device has drained.  A server
// that frees a finished scene (or the staging buffers of a new one) while
// batches run therefore stalls behind all queued batch steps, and the GPU
// then idles while the host builds the next scene.  Solvers and batches
// take their device memory from this cache instead: a released block is
// kept for the next request of its size class, and blocks go back to the
// driver only when a device's cached bytes exceed the limit, on
// mpm_release_cache(), or to retry an allocation that failed.
//
// A block is released only after the streams that used it are done (the
// callers synchronise the owning stream first), so a reused block needs no
// event.  All functions are thread-safe.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <unordered_map>
#include <utility>
#include <vector>

namespace mpm {
namespace devcache {

constexpr int kMaxDevices = 64;

struct State {
    std::mutex mu;
    std::unordered_map<void*, std::pair<int, size_t>> live;  // block -> (device, class bytes)
    std::multimap<size_t, void*> idle[kMaxDevices];          // class bytes -> released block
    size_t cached[kMaxDevices] = {};
    size_t live_bytes[kMaxDevices] = {};
    size_t limit = size_t(16) << 30;  // cached bytes kept per device
    uint64_t hits = 0, misses = 0;
};
inline State& state() {
    static State* s = new State;  // never destroyed: releases may come during process exit
    return *s;
}

// Size classes: 4 KiB minimum, then 8 classes per power of two (a block is
// at most 12.5% larger than its request).
inline size_t class_bytes(size_t n) {
    if (n <= 4096) return 4096;
    const int k = 63 - __builtin_clzll((unsigned long long)(n - 1));  // 2^k < n <= 2^(k+1)
    const size_t step = (size_t)1 << (k - 2);                          // 2^(k+1) / 8
    return (n + step - 1) / step * step;
}

// hipFree of blocks owned by device dev (made current for the call).
inline void free_on(int dev, const std::vector<void*>& blocks) {
    if (blocks.empty()) return;
    int cur = -1;
    (void)hipGetDevice(&cur);
    if (cur != dev) (void)hipSetDevice(dev);
    for (void* p : blocks) (void)hipFree(p);
    if (cur != dev && cur >= 0) (void)hipSetDevice(cur);
}

// Take idle blocks of device dev (largest first) until at most keep bytes
// stay cached; the caller frees them outside the lock.
inline std::vector<void*> take_excess(State& st, int dev, size_t keep) {
    std::vector<void*> out;
    auto& m = st.idle[dev];
    while (st.cached[dev] > keep && !m.empty()) {
        auto it = std::prev(m.end());
        st.cached[dev] -= it->first;
        out.push_back(it->second);
        m.erase(it);
    }
    return out;
}

// A block of at least n bytes on the current device; *got = its class bytes.
inline hipError_t alloc(void** p, size_t n, size_t* got) {
    *p = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    const size_t cls = class_bytes(n);
    State& st = state();
    {
        std::lock_guard<std::mutex> g(st.mu);
        auto& m = st.idle[dev];
        auto it = m.lower_bound(cls);
        if (it != m.end() && it->first <= cls + cls / 4) {  // a cached block at most 25% larger
            *p = it->second;
            *got = it->first;
            st.cached[dev] -= it->first;
            st.live_bytes[dev] += it->first;
            st.live[*p] = {dev, it->first};
            m.erase(it);
            ++st.hits;
            return hipSuccess;
        }
    }
    e = hipMalloc(p, cls);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        std::vector<void*> blocks;
        {
            std::lock_guard<std::mutex> g(st.mu);
            blocks = take_excess(st, dev, 0);
        }
        free_on(dev, blocks);
        e = hipMalloc(p, cls);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    std::lock_guard<std::mutex> g(st.mu);
    st.live[*p] = {dev, cls};
    st.live_bytes[dev] += cls;
    ++st.misses;
    *got = cls;
    return hipSuccess;
}

// Return a block to the cache (every stream that used it must be done).
// Returns its class bytes (0 for a pointer the cache did not hand out,
// which is freed directly).
inline size_t release(void* p) {
    if (!p) return 0;
    State& st = state();
    std::vector<void*> excess;
    int dev = 0;
    size_t cls = 0;
    {
        std::lock_guard<std::mutex> g(st.mu);
        auto it = st.live.find(p);
        if (it != st.live.end()) {
            dev = it->second.first;
            cls = it->second.second;
            st.live.erase(it);
            st.live_bytes[dev] -= cls;
            st.idle[dev].emplace(cls, p);
            st.cached[dev] += cls;
            if (st.cached[dev] > st.limit) excess = take_excess(st, dev, st.limit);
        }
    }
    if (cls == 0) {
        (void)hipFree(p);
        return 0;
    }
    free_on(dev, excess);
    return cls;
}

// Hand every cached block of every device back to the driver.
inline void release_all() {
    State& st = state();
    for (int d = 0; d < kMaxDevices; ++d) {
        std::vector<void*> blocks;
        {
            std::lock_guard<std::mutex> g(st.mu);
            blocks = take_excess(st, d, 0);
        }
        free_on(d, blocks);
    }
}

inline void set_limit(size_t bytes) {
    State& st = state();
    std::vector<std::vector<void*>> excess(kMaxDevices);
    {
        std::lock_guard<std::mutex> g(st.mu);
        st.limit = bytes;
        for (int d = 0; d < kMaxDevices; ++d) excess[d] = take_excess(st, d, bytes);
    }
    for (int d = 0; d < kMaxDevices; ++d) free_on(d, excess[d]);
}

}  // namespace devcache
}  // namespace mpm
