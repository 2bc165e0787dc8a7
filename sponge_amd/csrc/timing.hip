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
std::vector<std::string> g_filter;  // names timed; empty = all

// the segmented capture open on this thread, if any
thread_local sa::SegmentedCapture* t_capture = nullptr;
thread_local hipStream_t t_capture_stream = nullptr;

bool selected(const char* name) {
    if (g_filter.empty()) return true;
    for (const auto& f : g_filter)
        if (f == name) return true;
    return false;
}


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

// A timer boundary inside the segmented capture: close the current segment
// and open the next; the mark goes between them.
void capture_cut(const char* name, bool begin) {
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(t_capture_stream, &g) != hipSuccess) return;
    t_capture->items.push_back(sa::CaptureItem{g, nullptr, std::string(), false});
    t_capture->items.push_back(sa::CaptureItem{nullptr, nullptr, std::string(name), begin});
    (void)hipStreamBeginCapture(t_capture_stream, hipStreamCaptureModeThreadLocal);
}

void record_begin(const char* name, hipStream_t s) {
    hipEvent_t e = take_event();
    if (!e) return;
    if (hipEventRecord(e, s) != hipSuccess) {
        g_pool.push_back(e);
        return;
    }
    g_timers[name].open = e;
}

void record_end(const char* name, hipStream_t s) {
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

}  // namespace

namespace sa {

bool g_timing = false;

void timer_begin(const char* name, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!selected(name)) return;
    if (t_capture && s == t_capture_stream) {
        capture_cut(name, true);
        return;
    }
    record_begin(name, s);
}

void timer_end(const char* name, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!selected(name)) return;
    if (t_capture && s == t_capture_stream) {
        capture_cut(name, false);
        return;
    }
    record_end(name, s);
}

int capture_begin(SegmentedCapture& c, hipStream_t s) {
    capture_destroy(c);
    c.timing = g_timing;
    SA_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    t_capture = &c;
    t_capture_stream = s;
    return 0;
}

int capture_end(SegmentedCapture& c, hipStream_t s) {
    t_capture = nullptr;
    t_capture_stream = nullptr;
    hipGraph_t g = nullptr;
    SA_CHECK(hipStreamEndCapture(s, &g));
    c.items.push_back(CaptureItem{g, nullptr, std::string(), false});
    // drop empty segments, instantiate the others
    std::vector<CaptureItem> kept;
    for (auto& it : c.items) {
        if (!it.graph) {
            kept.push_back(it);
            continue;
        }
        size_t n = 0;
        SA_CHECK(hipGraphGetNodes(it.graph, nullptr, &n));
        if (n == 0) {
            (void)hipGraphDestroy(it.graph);
            continue;
        }
        SA_CHECK(hipGraphInstantiate(&it.exec, it.graph, nullptr, nullptr, 0));
        kept.push_back(it);
    }
    c.items.swap(kept);
    return 0;
}

int capture_launch(const SegmentedCapture& c, hipStream_t s) {
    for (const auto& it : c.items) {
        if (it.exec) {
            SA_CHECK(hipGraphLaunch(it.exec, s));
        } else {
            std::lock_guard<std::mutex> lk(g_mu);
            if (it.begin) record_begin(it.timer.c_str(), s);
            else record_end(it.timer.c_str(), s);
        }
    }
    return 0;
}

void capture_destroy(SegmentedCapture& c) {
    for (auto& it : c.items) {
        if (it.exec) (void)hipGraphExecDestroy(it.exec);
        if (it.graph) (void)hipGraphDestroy(it.graph);
    }
    c.items.clear();
}


}  // namespace sa

extern "C" int sponge_amd_set_kernel_timing(int enable) {
    sa::g_timing = enable != 0;
    return 0;
}

extern "C" int sponge_amd_set_kernel_timing_filter(const char* names) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_filter.clear();
    if (!names) return 0;
    std::string cur;
    for (const char* p = names;; ++p) {
        if (*p == ',' || *p == 0) {
            if (!cur.empty()) g_filter.push_back(cur);
            cur.clear();
            if (*p == 0) break;
        } else {
            cur.push_back(*p);
        }
    }
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
