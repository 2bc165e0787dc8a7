// Batched driver: many independent systems (each a sponge_md with its own
// neighbor list, PME grid and tuned FFT tiling) advanced together with ONE
// dispatch per kernel of the MD step.
//
// Every kernel of the force-only step is a device body behind SA_LAUNCH
// (batch_launch.h).  The driver records one step of each system -- the same
// host code path that launches it alone, with a LaunchRecorder installed --
// and checks that every system produced the same sequence of kernels.  For
// each position of that sequence it keeps, in one device buffer, the prefix
// of the systems' workgroup counts and each system's packed arguments; one
// launch of k_batched then covers every system (workgroup b runs the body for
// the system whose range holds b).  k steps of these launches are captured
// into one HIP graph and replayed.  No counterpart in the reference (one
// process runs one system there).
#include <algorithm>
#include <array>
#include <chrono>
#include <map>
#include <tuple>
#include <vector>

#include "internal.h"
#include "batch_launch.h"
#include "timing.h"

using namespace sa;

thread_local sa::LaunchRecorder* sa::g_recorder = nullptr;

namespace {

// one kernel of the recorded step, over every system
struct BatchPos {
    const void* fn = nullptr;
    int lane = 0;       // 1: on the side stream
    bool join = false;  // after the side stream's launches so far
    dim3 block;
    size_t shmem = 0;      // the largest of the systems'
    int total_blocks = 0;  // sum over the systems
    int stride = 0;        // bytes per system in the argument array
    size_t start_off = 0;  // device buffer offsets
    size_t args_off = 0;
};

}  // namespace

struct sponge_batch {
    std::vector<sponge_md*> md;
    hipStream_t stream = nullptr;  // the batched launches
    hipStream_t side = nullptr;    // their reciprocal-space chain
    hipEvent_t fork = nullptr, join = nullptr;
    std::vector<hipEvent_t> done;  // per system: its earlier work
    hipEvent_t after = nullptr;    // end of a run
    std::vector<BatchPos> pos;
    char* d_buf = nullptr;         // starts and argument tuples of every position
    bool recorded = false;
    std::vector<int> generation;   // each system's list generation at the recording
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = 0;
};

namespace {

void batch_free_graph(sponge_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_steps = 0;
}

void batch_free_record(sponge_batch* b) {
    batch_free_graph(b);
    if (b->d_buf) (void)hipFree(b->d_buf);
    b->d_buf = nullptr;
    b->pos.clear();
    b->recorded = false;
}

void batch_free(sponge_batch* b) {
    batch_free_record(b);
    for (auto e : b->done)
        if (e) (void)hipEventDestroy(e);
    b->done.clear();
    if (b->after) (void)hipEventDestroy(b->after);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->join) (void)hipEventDestroy(b->join);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->side) (void)hipStreamDestroy(b->side);
    b->after = b->fork = b->join = nullptr;
    b->stream = b->side = nullptr;
}

// One step of system m recorded instead of launched.  The system's stream is
// captured meanwhile: a kernel or copy of the step that does not go through
// SA_LAUNCH would land in that capture, and is reported.
int record_system(sponge_md* m, LaunchRecorder& rec) {
    hipStream_t s = md_stream(m);
    SA_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    g_recorder = &rec;
    const int rc = md_record_steps(m, 1);
    g_recorder = nullptr;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    size_t stray = 0;
    if (g) {
        (void)hipGraphGetNodes(g, nullptr, &stray);
        (void)hipGraphDestroy(g);
    }
    (void)hipGetLastError();
    if (rc) return rc;
    if (e != hipSuccess) return (int)e;
    return stray ? SPONGE_AMD_ERROR_NOT_IMPLEMENTED : 0;
}

int batch_record(sponge_batch* b) {
    batch_free_record(b);
    const int n = (int)b->md.size();
    // the tile spread's workgroup size must be one for every system
    int waves = 1 << 30;
    for (sponge_md* m : b->md) {
        sponge_nl* nl = nullptr;
        sponge_lj* lj = nullptr;
        sponge_pme* pme = nullptr;
        const int rc = sponge_md_handles(m, &nl, &lj, &pme);
        if (rc) return rc;
        waves = std::min(waves, pme_spread_waves(pme));
    }
    std::vector<LaunchRecorder> rec(n);
    for (int i = 0; i < n; ++i) {
        sponge_md* m = b->md[i];
        sponge_nl* nl = nullptr;
        sponge_lj* lj = nullptr;
        sponge_pme* pme = nullptr;
        int rc = sponge_md_handles(m, &nl, &lj, &pme);
        if (!rc) pme_set_spread_waves(pme, waves);
        if (!rc) rc = record_system(m, rec[i]);
        if (rc) return rc;
    }
    // every system must have recorded the same kernels with the same blocks
    const size_t np = rec[0].launches.size();
    for (int i = 1; i < n; ++i) {
        if (rec[i].launches.size() != np) return SPONGE_AMD_ERROR_NOT_IMPLEMENTED;
        for (size_t p = 0; p < np; ++p) {
            const RecordedLaunch& a = rec[0].launches[p];
            const RecordedLaunch& c = rec[i].launches[p];
            if (a.fn != c.fn || a.block.x != c.block.x || a.block.y != c.block.y ||
                a.block.z != c.block.z || a.args.size() != c.args.size() ||
                a.lane != c.lane || a.join != c.join)
                return SPONGE_AMD_ERROR_NOT_IMPLEMENTED;
        }
    }
    // one device buffer: per position the block prefix [n+1] and the tuples
    std::vector<char> host;
    auto align16 = [](size_t x) { return (x + 15) / 16 * 16; };
    b->pos.resize(np);
    for (size_t p = 0; p < np; ++p) {
        BatchPos& bp = b->pos[p];
        const RecordedLaunch& r0 = rec[0].launches[p];
        bp.fn = r0.fn;
        bp.lane = r0.lane;
        bp.join = r0.join;
        bp.block = r0.block;
        bp.stride = (int)align16(r0.args.size());
        std::vector<int> start(n + 1, 0);
        for (int i = 0; i < n; ++i) {
            const RecordedLaunch& r = rec[i].launches[p];
            start[i + 1] = start[i] + r.blocks;
            bp.shmem = std::max(bp.shmem, r.shmem);
        }
        bp.total_blocks = start[n];
        bp.start_off = align16(host.size());
        host.resize(bp.start_off + sizeof(int) * (n + 1));
        std::memcpy(host.data() + bp.start_off, start.data(), sizeof(int) * (n + 1));
        bp.args_off = align16(host.size());
        host.resize(bp.args_off + (size_t)bp.stride * n, 0);
        for (int i = 0; i < n; ++i) {
            const RecordedLaunch& r = rec[i].launches[p];
            std::memcpy(host.data() + bp.args_off + (size_t)bp.stride * i, r.args.data(),
                        r.args.size());
        }
    }
    if (host.empty()) return SPONGE_AMD_ERROR_NOT_IMPLEMENTED;
    if (hipMalloc(&b->d_buf, host.size()) != hipSuccess) {
        b->d_buf = nullptr;
        return SPONGE_AMD_ERROR_MALLOC_FAILED;
    }
    SA_CHECK(hipMemcpyAsync(b->d_buf, host.data(), host.size(), hipMemcpyHostToDevice, b->stream));
    SA_CHECK(hipStreamSynchronize(b->stream));
    b->generation.clear();
    for (sponge_md* m : b->md) b->generation.push_back(md_generation(m));
    b->recorded = true;
    return 0;
}

// k batched steps on b->stream (eager, or into the capture in progress).  The
// launches recorded on the side lane (the reciprocal chain that reads only the
// coordinates) run on b->side, forked at the start of each step and joined
// before the first launch that reads their results, as the single-system
// step forks them.
int batch_issue(sponge_batch* b, int k) {
    const int n = (int)b->md.size();
    bool any_side = false;
    for (const BatchPos& bp : b->pos) any_side |= bp.lane == 1;
    for (int step = 0; step < k; ++step) {
        bool open = false;  // side launches of this step not yet joined
        if (any_side) {
            SA_CHECK(hipEventRecord(b->fork, b->stream));
            SA_CHECK(hipStreamWaitEvent(b->side, b->fork, 0));
        }
        for (const BatchPos& bp : b->pos) {
            if (bp.join && open) {
                SA_CHECK(hipEventRecord(b->join, b->side));
                SA_CHECK(hipStreamWaitEvent(b->stream, b->join, 0));
                open = false;
            }
            const int* start = reinterpret_cast<const int*>(b->d_buf + bp.start_off);
            const char* args = b->d_buf + bp.args_off;
            int nn = n, stride = bp.stride;
            void* kargs[] = {(void*)&start, (void*)&nn, (void*)&args, (void*)&stride};
            SA_CHECK(hipLaunchKernel(bp.fn, dim3(bp.total_blocks), bp.block, kargs, bp.shmem,
                                     bp.lane == 1 ? b->side : b->stream));
            open |= bp.lane == 1;
        }
        if (open) {  // no recorded consumer: joined at the step's end
            SA_CHECK(hipEventRecord(b->join, b->side));
            SA_CHECK(hipStreamWaitEvent(b->stream, b->join, 0));
        }
    }
    return 0;
}

int batch_capture(sponge_batch* b, int k) {
    batch_free_graph(b);
    SA_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = batch_issue(b, k);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return rc ? rc : (int)e;
    }
    b->graph = g;
    SA_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_steps = k;
    return 0;
}

}  // namespace

extern "C" {

int sponge_batch_initial(sponge_batch** out, int n_systems, sponge_md* const* systems) {
    if (!out || n_systems <= 0 || !systems) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    *out = nullptr;
    for (int i = 0; i < n_systems; ++i)
        if (!systems[i]) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    sponge_batch* b = new sponge_batch();
    b->md.assign(systems, systems + n_systems);
    b->done.assign(n_systems, nullptr);
    bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&b->after, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&b->join, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; i < n_systems && ok; ++i)
        ok = hipEventCreateWithFlags(&b->done[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        batch_free(b);
        delete b;
        return SPONGE_AMD_ERROR_MALLOC_FAILED;
    }
    *out = b;
    return 0;
}

int sponge_batch_autotune(sponge_batch* b, int reps, int* tiles) {
    if (!b || reps < 1) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    // one tuning per distinct grid shape
    std::map<std::tuple<int, int, int>, std::vector<int>> done;
    for (size_t i = 0; i < b->md.size(); ++i) {
        sponge_md* m = b->md[i];
        sponge_nl* nl = nullptr;
        sponge_lj* lj = nullptr;
        sponge_pme* pme = nullptr;
        int rc = sponge_md_handles(m, &nl, &lj, &pme);
        if (rc) return rc;
        const auto key = std::make_tuple(pme->fx, pme->fy, pme->fz);
        auto it = done.find(key);
        int t[3];
        if (it == done.end()) {
            float* crd = nullptr;
            float* charge = nullptr;
            rc = md_state_arrays(m, &crd, &charge);
            if (!rc) rc = sponge_pme_autotune(pme, crd, charge, reps, t);
            if (rc) return rc;
            done[key] = {t[0], t[1], t[2]};
        } else {
            for (int a = 0; a < 3; ++a) t[a] = it->second[a];
            rc = sponge_pme_set_fft_tiling(pme, t);
            if (rc) return rc;
        }
    }
    batch_free_record(b);  // the tilings are launch arguments: record again
    // Refinement for the batched launches: a single system's tuning favours
    // short lines per workgroup (a small grid alone has few lines to spread
    // over the GPU); in one launch over every system longer ones may win.
    // Per axis, the per-shape choices are timed against one common value of
    // {2, 4, 8, 16} on batched steps (steps_per_launch 10: one warm launch,
    // then `reps` timed ones), and the fastest is kept.
    const size_t n = b->md.size();
    std::vector<std::array<int, 3>> cur(n), best(n);
    for (size_t i = 0; i < n; ++i) {
        sponge_nl* nl = nullptr;
        sponge_lj* lj = nullptr;
        sponge_pme* pme = nullptr;
        int rc = sponge_md_handles(b->md[i], &nl, &lj, &pme);
        if (rc) return rc;
        const auto& v = done[std::make_tuple(pme->fx, pme->fy, pme->fz)];
        cur[i] = {v[0], v[1], v[2]};
    }
    auto apply = [&](const std::vector<std::array<int, 3>>& t) -> int {
        for (size_t i = 0; i < n; ++i) {
            sponge_nl* nl = nullptr;
            sponge_lj* lj = nullptr;
            sponge_pme* pme = nullptr;
            int rc = sponge_md_handles(b->md[i], &nl, &lj, &pme);
            if (!rc) rc = sponge_pme_set_fft_tiling(pme, t[i].data());
            if (rc) return rc;
        }
        batch_free_record(b);
        return 0;
    };
    auto time_steps = [&](double* ms) -> int {
        constexpr int kSpl = 10;
        int rc = sponge_batch_run(b, kSpl, kSpl);
        if (!rc) rc = sponge_batch_synchronize(b);
        if (rc) return rc;
        const auto t0 = std::chrono::steady_clock::now();
        rc = sponge_batch_run(b, reps * kSpl, kSpl);
        if (!rc) rc = sponge_batch_synchronize(b);
        *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return rc;
    };
    for (int a = 0; a < 3; ++a) {
        double best_ms = 0;
        int rc = apply(cur);
        if (!rc) rc = time_steps(&best_ms);
        if (rc) return rc;
        best = cur;
        for (int v : {2, 4, 8, 16}) {
            std::vector<std::array<int, 3>> cand = cur;
            for (size_t i = 0; i < n; ++i) cand[i][a] = v;
            double ms = 0;
            rc = apply(cand);
            if (!rc) rc = time_steps(&ms);
            if (rc) return rc;
            if (ms < best_ms) {
                best_ms = ms;
                best = cand;
            }
        }
        cur = best;
    }
    int rc = apply(cur);
    if (rc) return rc;
    if (tiles)
        for (size_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) tiles[3 * i + a] = cur[i][a];
    return 0;
}

int sponge_batch_run(sponge_batch* b, int steps, int steps_per_launch) {
    if (!b || steps < 0 || steps_per_launch < 1) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (g_timing) return SPONGE_AMD_ERROR_CONFLICTING_COMMAND;  // timers cannot join this capture
    int done_steps = 0;
    // the first step of a system builds its list; other schedules (an impulse
    // interval, a fixed refresh interval, the surrogate) run system by system
    bool all = true;
    for (sponge_md* m : b->md) all = all && md_graphable(m);
    if (!all) {
        for (sponge_md* m : b->md) {
            const int rc = sponge_md_run(m, 1, 0, 0);
            if (rc) return rc;
        }
        done_steps = 1;
        for (sponge_md* m : b->md)
            if (!md_graphable(m)) {
                for (sponge_md* q : b->md) {
                    const int rc = sponge_md_run(q, steps - done_steps, 0, 0);
                    if (rc) return rc;
                }
                return 0;
            }
    }
    while (done_steps < steps) {
        if (!b->recorded) {
            // the systems' earlier work completes before the recording reads
            // their state
            for (sponge_md* m : b->md) SA_CHECK(hipStreamSynchronize(md_stream(m)));
            const int rc = batch_record(b);
            if (rc) return rc;
        }
        const int left = steps - done_steps;
        const int k = std::min(left, steps_per_launch);
        const int launches = left / steps_per_launch;
        // the batched work follows each system's earlier work on its stream
        for (size_t i = 0; i < b->md.size(); ++i) {
            SA_CHECK(hipEventRecord(b->done[i], md_stream(b->md[i])));
            SA_CHECK(hipStreamWaitEvent(b->stream, b->done[i], 0));
        }
        int advanced = 0;
        if (launches > 0) {
            if (b->graph_steps != steps_per_launch) {
                const int rc = batch_capture(b, steps_per_launch);
                if (rc) return rc;
            }
            // one graph launch per steps_per_launch steps; overflow checks (a
            // host synchronisation) at the launch boundaries
            SA_CHECK(hipGraphLaunch(b->exec, b->stream));
            advanced = steps_per_launch;
        } else {
            const int rc = batch_issue(b, k);
            if (rc) return rc;
            advanced = k;
        }
        SA_CHECK(hipEventRecord(b->after, b->stream));
        for (sponge_md* m : b->md) {
            SA_CHECK(hipStreamWaitEvent(md_stream(m), b->after, 0));
            const long long first = md_step_count(m);
            md_advance(m, advanced);
            // each system's neighbor-list overflow check over these steps
            const int rc = md_overflow_check(m, first, md_step_count(m) - 1);
            if (rc) return rc;
        }
        done_steps += advanced;
        for (size_t i = 0; i < b->md.size(); ++i)
            if (md_generation(b->md[i]) != b->generation[i]) {
                batch_free_record(b);  // a re-initialised list: record again
                break;
            }
    }
    return 0;
}

int sponge_batch_synchronize(sponge_batch* b) {
    if (!b) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    SA_CHECK(hipStreamSynchronize(b->stream));
    for (sponge_md* m : b->md) SA_CHECK(hipStreamSynchronize(md_stream(m)));
    return 0;
}

int sponge_batch_clear(sponge_batch* b) {
    if (!b) return 0;
    (void)sponge_batch_synchronize(b);
    batch_free(b);
    delete b;
    return 0;
}

}  // extern "C"
