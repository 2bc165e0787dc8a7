// Per-kernel timers: HIP events recorded on the launching stream around the
// hot kernels when timing is switched on (the role of the reference's
// TIME_RECORDER, SPONGE/utils/control/time_recorder.hpp:10-42, without its
// device synchronisation on every start / stop).
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "timing.h"

namespace {

struct Timer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    hipEvent_t open = nullptr;
    double total_ms = 0.0;
    long long count = 0;
};

std::mutex g_mu;
std::map<std::string, Timer> g_timers;
std::vector<hipEvent_t> g_pool;

hipEvent_t take_event() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int drain(Timer& t) {
    for (auto& pr : t.pending) {
        hipError_t e = hipEventSynchronize(pr.second);
        if (e != hipSuccess) return (int)e;
        float ms = 0.f;
        e = hipEventElapsedTime(&ms, pr.first, pr.second);
        if (e != hipSuccess) return (int)e;
        t.total_ms += ms;
        t.count += 1;
        g_pool.push_back(pr.first);
        g_pool.push_back(pr.second);
    }
    t.pending.clear();
    return 0;
}

}  // namespace

namespace sa {

bool g_timing = false;

void timer_begin(const char* name, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_mu);
    Timer& t = g_timers[name];
    hipEvent_t e = take_event();
    if (!e) return;
    if (hipEventRecord(e, s) != hipSuccess) {
        g_pool.push_back(e);
        return;
    }
    t.open = e;
}

void timer_end(const char* name, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_mu);
    Timer& t = g_timers[name];
    if (!t.open) return;
    hipEvent_t e = take_event();
    if (!e) return;
    if (hipEventRecord(e, s) != hipSuccess) {
        g_pool.push_back(e);
        return;
    }
    t.pending.emplace_back(t.open, e);
    t.open = nullptr;
    // bound the number of outstanding events
    if (t.pending.size() > 4096) drain(t);
}

}  // namespace sa

extern "C" int sponge_amd_set_kernel_timing(int enable) {
    sa::g_timing = enable != 0;
    return 0;
}

extern "C" int sponge_amd_kernel_timing(const char* name, double* total_ms, long long* launches) {
    if (!name) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_timers.find(name);
    if (it == g_timers.end()) {
        if (total_ms) *total_ms = 0.0;
        if (launches) *launches = 0;
        return 0;
    }
    int rc = drain(it->second);
    if (rc) return rc;
    if (total_ms) *total_ms = it->second.total_ms;
    if (launches) *launches = it->second.count;
    return 0;
}

extern "C" int sponge_amd_reset_kernel_timing(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& kv : g_timers) {
        int rc = drain(kv.second);
        if (rc) return rc;
        kv.second.total_ms = 0.0;
        kv.second.count = 0;
    }
    return 0;
}
