// Optional per-kernel event timers (see timing.hip).
#pragma once

#include "common.h"

#include <string>
#include <vector>

namespace sa {
extern bool g_timing;
void timer_begin(const char* name, hipStream_t s);
void timer_end(const char* name, hipStream_t s);

// Timers inside a captured sequence.  While a SegmentedCapture is open on a
// stream, each timer boundary ends the capture there and starts a new one, so
// the captured work becomes a list of graphs with the timer marks between
// them.  Replaying the list launches the graphs in order on the stream and
// records ordinary timer events at the marks: every replay is timed like an
// eager launch, and with timers off the list is a single graph.
struct CaptureItem {
    hipGraph_t graph = nullptr;        // a segment, or
    hipGraphExec_t exec = nullptr;
    std::string timer;                 // a timer mark
    bool begin = false;
};
struct SegmentedCapture {
    std::vector<CaptureItem> items;
    bool timing = false;               // timers were on during the capture
};
int capture_begin(SegmentedCapture& c, hipStream_t s);
int capture_end(SegmentedCapture& c, hipStream_t s);  // also instantiates
int capture_launch(const SegmentedCapture& c, hipStream_t s);
void capture_destroy(SegmentedCapture& c);

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
