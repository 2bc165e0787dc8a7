// Optional per-kernel event timers (see timing.hip).
#pragma once

#include "common.h"

#include <string>
#include <vector>

namespace sa {
extern bool g_timing;
void timer_begin(const char* name, hipStream_t s);
void timer_end(const char* name, hipStream_t s);

// Timers inside HIP graphs: the events a timer records while its stream is
// capturing become event-record nodes of the graph.  The graph's owner takes
// the (name, begin event, end event) triples after capture, finds the nodes,
// and before every launch points them at fresh events, so each replay of the
// graph is timed like an eager launch.
struct GraphTimer {
    std::string name;
    hipGraphNode_t begin = nullptr, end = nullptr;
};
struct GraphTimers {
    std::vector<GraphTimer> nodes;
    bool timing = false;  // whether timers were on when the graph was captured
};
// after hipStreamEndCapture: find the event-record nodes of the timers
// recorded during the capture; clears the capture record
int graph_timers_collect(hipGraph_t g, GraphTimers& out);
// before hipGraphLaunch: fresh events into the nodes, registered as pending
int graph_timers_arm(hipGraphExec_t exec, const GraphTimers& t);

// Records events around a launch when timing is on.
struct ScopedTimer {
    const char* name;
    hipStream_t s;
    ScopedTimer(const char* n, hipStream_t st) : name(n), s(st) {
        if (g_timing) timer_begin(name, s);
    }
    ~ScopedTimer() {
        if (g_timing) timer_end(name, s);
    }
};
}  // namespace sa
