// Batched driver: many independent systems (each a sponge_md with its own
// stream, neighbor list, PME grid and tuned FFT tiling) stepped together by
// one HIP graph.  The capture forks every system's stream from one origin
// stream, records k MD steps of each system on its own stream and joins them
// back, so one graph launch advances every system by k steps and the
// systems' kernels run concurrently as independent branches of the graph.
// No counterpart in the reference (one process runs one system there).
#include <map>
#include <tuple>
#include <vector>

#include "internal.h"
#include "timing.h"

using namespace sa;

struct sponge_batch {
    std::vector<sponge_md*> md;
    hipStream_t stream = nullptr;  // origin of the capture and of the launches
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> done;  // per system: its work before a launch
    hipEvent_t after = nullptr;    // end of a launch
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = 0;
    std::vector<int> generation;  // each system's list generation at the capture
};

namespace {

void batch_free_graph(sponge_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_steps = 0;
}

void batch_free(sponge_batch* b) {
    batch_free_graph(b);
    for (auto e : b->done)
        if (e) (void)hipEventDestroy(e);
    b->done.clear();
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->after) (void)hipEventDestroy(b->after);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    b->fork = b->after = nullptr;
    b->stream = nullptr;
}

int capture(sponge_batch* b, int k) {
    batch_free_graph(b);
    SA_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = (int)hipEventRecord(b->fork, b->stream);
    for (size_t i = 0; i < b->md.size() && !rc; ++i) {
        hipStream_t s = md_stream(b->md[i]);
        rc = (int)hipStreamWaitEvent(s, b->fork, 0);
        if (!rc) rc = md_record_steps(b->md[i], k);
        if (!rc) rc = (int)hipEventRecord(b->done[i], s);
        if (!rc) rc = (int)hipStreamWaitEvent(b->stream, b->done[i], 0);
    }
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
    b->generation.clear();
    for (sponge_md* m : b->md) b->generation.push_back(md_generation(m));
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
              hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&b->after, hipEventDisableTiming) == hipSuccess;
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
        if (tiles)
            for (int a = 0; a < 3; ++a) tiles[3 * i + a] = t[a];
    }
    batch_free_graph(b);
    return 0;
}

int sponge_batch_run(sponge_batch* b, int steps, int steps_per_launch) {
    if (!b || steps < 0 || steps_per_launch < 1) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (g_timing) return SPONGE_AMD_ERROR_CONFLICTING_COMMAND;  // timers cannot cut this capture
    for (sponge_md* m : b->md)
        if (!md_graphable(m)) {
            // first steps build the lists; other schedules run system by system
            for (sponge_md* q : b->md) {
                const int rc = sponge_md_run(q, steps, 0, 0);
                if (rc) return rc;
            }
            return 0;
        }
    const int launches = steps / steps_per_launch, rest = steps - launches * steps_per_launch;
    if (launches > 0) {
        if (b->graph_steps != steps_per_launch) {
            const int rc = capture(b, steps_per_launch);
            if (rc) return rc;
        }
        // the launches follow each system's earlier work on its own stream
        for (size_t i = 0; i < b->md.size(); ++i) {
            SA_CHECK(hipEventRecord(b->done[i], md_stream(b->md[i])));
            SA_CHECK(hipStreamWaitEvent(b->stream, b->done[i], 0));
        }
        for (int l = 0; l < launches; ++l) SA_CHECK(hipGraphLaunch(b->exec, b->stream));
        SA_CHECK(hipEventRecord(b->after, b->stream));
        for (sponge_md* m : b->md) {
            SA_CHECK(hipStreamWaitEvent(md_stream(m), b->after, 0));
            const long long first = md_step_count(m);
            md_advance(m, launches * steps_per_launch);
            // each system's neighbor-list overflow check over the replayed steps
            const int rc = md_overflow_check(m, first, md_step_count(m) - 1);
            if (rc) return rc;
        }
        for (size_t i = 0; i < b->md.size(); ++i)
            if (md_generation(b->md[i]) != b->generation[i]) {
                batch_free_graph(b);  // a re-initialised list: capture again
                break;
            }
    }
    for (sponge_md* m : b->md)
        if (rest) {
            const int rc = sponge_md_run(m, rest, 0, 0);
            if (rc) return rc;
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
