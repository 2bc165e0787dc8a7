// Optional per-kernel event timers (see timing.hip).
#pragma once

#include "common.h"

namespace sa {
extern bool g_timing;
void timer_begin(const char* name, hipStream_t s);
void timer_end(const char* name, hipStream_t s);

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
