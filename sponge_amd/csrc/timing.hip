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

struct Captured {
    std::string name;
    hipEvent_t begin, end;
};
std::vector<Captured> g_captured;  // timers recorded into a capturing stream
std::map<std::string, hipEvent_t> g_capture_open;

bool selected(const char* name) {
    if (g_filter.empty()) return true;
    for (const auto& f : g_filter)
        if (f == name) return true;
    return false;
}

bool capturing(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) return false;
    return cs == hipStreamCaptureStatusActive;
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

}  // namespace

namespace sa {

bool g_timing = false;

// inside a capture the record must be an external event-record node (a plain
// record there only marks a dependency point of the capture)
hipError_t record(hipEvent_t e, hipStream_t s, bool cap) {
    return cap ? hipEventRecordWithFlags(e, s, hipEventRecordExternal) : hipEventRecord(e, s);
}

void timer_begin(const char* name, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!selected(name)) return;
    const bool cap = capturing(s);
    hipEvent_t e = take_event();
    if (!e) return;
    if (record(e, s, cap) != hipSuccess) {
        g_pool.push_back(e);
        return;
    }
    if (cap) g_capture_open[name] = e;
    else g_timers[name].open = e;
}

void timer_end(const char* name, hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!selected(name)) return;
    const bool cap = capturing(s);
    hipEvent_t b = nullptr;
    if (cap) {
        auto it = g_capture_open.find(name);
        if (it == g_capture_open.end()) return;
        b = it->second;
        g_capture_open.erase(it);
    } else {
        b = g_timers[name].open;
        if (!b) return;
    }
    hipEvent_t e = take_event();
    if (!e) return;
    if (record(e, s, cap) != hipSuccess) {
        g_pool.push_back(e);
        return;
    }
    if (cap) {
        g_captured.push_back({name, b, e});
        return;
    }
    Timer& t = g_timers[name];
    t.pending.emplace_back(b, e);
    t.open = nullptr;
    // bound the number of outstanding events
    if (t.pending.size() > 4096) drain(t);
}

int graph_timers_collect(hipGraph_t g, GraphTimers& out) {
    std::lock_guard<std::mutex> lk(g_mu);
    out.nodes.clear();
    out.timing = g_timing;
    std::vector<Captured> cap;
    cap.swap(g_captured);
    g_capture_open.clear();
    if (cap.empty()) return 0;
    size_t n = 0;
    hipError_t e = hipGraphGetNodes(g, nullptr, &n);
    if (e != hipSuccess) return (int)e;
    std::vector<hipGraphNode_t> nodes(n);
    e = hipGraphGetNodes(g, nodes.data(), &n);
    if (e != hipSuccess) return (int)e;
    for (const auto& c : cap) {
        GraphTimer gt;
        gt.name = c.name;
        for (auto nd : nodes) {
            hipGraphNodeType ty;
            if (hipGraphNodeGetType(nd, &ty) != hipSuccess || ty != hipGraphNodeTypeEventRecord) continue;
            hipEvent_t ev = nullptr;
            if (hipGraphEventRecordNodeGetEvent(nd, &ev) != hipSuccess) continue;
            if (ev == c.begin) gt.begin = nd;
            if (ev == c.end) gt.end = nd;
        }
        // the capture events stay referenced by the graph: not returned to the pool
        if (gt.begin && gt.end) out.nodes.push_back(gt);
    }
    return 0;
}

int graph_timers_arm(hipGraphExec_t exec, const GraphTimers& t) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (const auto& gt : t.nodes) {
        hipEvent_t b = take_event(), e = take_event();
        if (!b || !e) return (int)hipErrorOutOfMemory;
        hipError_t he = hipGraphExecEventRecordNodeSetEvent(exec, gt.begin, b);
        if (he == hipSuccess) he = hipGraphExecEventRecordNodeSetEvent(exec, gt.end, e);
        if (he != hipSuccess) {
            g_pool.push_back(b);
            g_pool.push_back(e);
            return (int)he;
        }
        Timer& tm = g_timers[gt.name];
        tm.pending.emplace_back(b, e);
        if (tm.pending.size() > 4096) {
            int rc = drain(tm);
            if (rc) return rc;
        }
    }
    return 0;
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
