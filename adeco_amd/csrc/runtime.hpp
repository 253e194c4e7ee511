// Host-runtime internals shared by the translation units of libadeco_amd.so.
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstddef>
#include <mutex>
#include <string>

#include "../../include/adeco.h"

namespace adeco {

extern thread_local std::string g_last_error;
extern thread_local double g_last_kernel_ms;

adeco_status fail(adeco_status st, const std::string& msg);
adeco_status hip_fail(hipError_t e, const char* what);
adeco_status check_device(int dev);
bool valid_bbox(const float* b);

// A grow-only device allocation reused across calls.
struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    adeco_status reserve(size_t nbytes);
    void free_all();
    template <class T> T* as() const { return static_cast<T*>(ptr); }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf();
};

// Timings of the calling thread's last deposit: whole deposit (events on the launch stream),
// the dominant kernel alone, (tet, candidate) pairs and exact recomputes.
struct LastTimings {
    double total_ms = 0, main_ms = 0;
    double candidates = 0, fallbacks = 0, partials = 0;
};
extern thread_local LastTimings g_timings;

struct DeviceState {
    int id = -1;
    bool ready = false;
    std::mutex mutex;  // serialises calls that share this device's buffers and stream
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evk0 = nullptr, evk1 = nullptr;
    size_t total_mem = 0;
    DevBuf pos, grid, stats, scratch;
    // the batched deposit of a tet soup (boundary calls in the batched form, adeco_deposit_soup):
    // the soup, tet ids, grid descriptor, candidate space, tile map and counters, reused by every
    // call on this device (calls hold the mutex)
    DevBuf soup, soup_ids, soup_desc, soup_work, soup_tiles, soup_ctr;
    adeco_status ensure(int dev);
};

// A named roctx range over a host scope (rocprofv3 --marker-trace shows it); next() ends it and
// starts the following phase.  It measures host time, so it covers a phase's GPU work where the
// phase ends with a synchronisation.
struct TraceRange {
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    void next(const char* name) {
        roctxRangePop();
        roctxRangePushA(name);
    }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

DeviceState& device_state(int dev);
size_t memory_budget(const DeviceState& ds);
adeco_status timed_begin(DeviceState& ds);
adeco_status timed_end(DeviceState& ds);

}  // namespace adeco
