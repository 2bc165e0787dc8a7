// One dispatch per kernel over many systems.
//
// Every kernel of the MD step is a device function `body(int block, args...)`
// behind a named __global__ wrapper that passes blockIdx.x.  Its launch site
// goes through SA_LAUNCH, which either launches the wrapper (the single-system
// path) or, while a LaunchRecorder is installed on the calling thread, records
// the launch instead: the batched kernel of the body, the system's block count,
// block size, dynamic LDS and its arguments packed into a DevTuple of the
// body's parameter types.  The batched driver records one step per system
// (same code path for each), then issues, for each position of the step, ONE
// launch of k_batched: workgroup b finds its system from the prefix of the
// systems' block counts and calls the same body with that system's argument
// tuple and its local block index.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>
#include <type_traits>
#include <utility>
#include <vector>

namespace sa {

// Tuple of kernel arguments with one layout on the host and the device.
template <typename... T>
struct DevTuple {};
template <typename H, typename... T>
struct DevTuple<H, T...> {
    H head;
    DevTuple<T...> tail;
    __host__ __device__ DevTuple() {}
    template <typename A, typename... B>
    __host__ DevTuple(A&& a, B&&... b) : head(static_cast<H>(a)), tail(std::forward<B>(b)...) {}
};

template <int I>
struct TupleGet {
    template <typename H, typename... T>
    __device__ __forceinline__ static const auto& get(const DevTuple<H, T...>& t) {
        return TupleGet<I - 1>::get(t.tail);
    }
};
template <>
struct TupleGet<0> {
    template <typename H, typename... T>
    __device__ __forceinline__ static const H& get(const DevTuple<H, T...>& t) {
        return t.head;
    }
};

// A body's parameters are stored by value in its tuple; a body that takes a
// structure by const reference reads it in place (from the kernel arguments,
// or from the batched launch's argument blob), never through a private copy.
template <typename F>
struct BodyTraits;
template <typename... P>
struct BodyTraits<void (*)(int, P...)> {
    using Tuple = DevTuple<std::decay_t<P>...>;
    static constexpr int arity = sizeof...(P);
};

template <auto F, typename Tup, size_t... I>
__device__ __forceinline__ void call_body(int block, const Tup& t, std::index_sequence<I...>) {
    F(block, TupleGet<I>::get(t)...);
}

// system of workgroup b and its block index within the system: the last
// system whose first block is <= b (bisection over the prefix)
__device__ __forceinline__ int batch_locate(const int* __restrict__ start, int n, int b,
                                            int& local) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    local = b - start[lo];
    return lo;
}

// Wavefronts per SIMD that a body's batched kernel is compiled for: 8 caps
// it at 64 VGPRs (a body whose speed depends on full occupancy specialises
// this); 1 leaves the register budget to the compiler.
template <auto F>
struct BodyWaves {
    static constexpr int value = 1;
};

template <int LB, auto F>
__global__ void __launch_bounds__(LB) __attribute__((amdgpu_waves_per_eu(BodyWaves<F>::value)))
k_batched(const int* __restrict__ start, int n, const char* __restrict__ blobs, int stride) {
    using Tr = BodyTraits<decltype(F)>;
    int local = 0;
    const int sys = batch_locate(start, n, (int)blockIdx.x, local);
    const auto& t = *reinterpret_cast<const typename Tr::Tuple*>(blobs + (size_t)sys * stride);
    call_body<F>(local, t, std::make_index_sequence<Tr::arity>{});
}

struct RecordedLaunch {
    const void* fn = nullptr;  // k_batched<LB, body> (host stub)
    int blocks = 0;            // this system's workgroups
    dim3 block;
    size_t shmem = 0;
    std::vector<char> args;    // the body's DevTuple
    int lane = 0;              // 0: the step's stream; 1: its side stream
    bool join = false;         // waits for the side stream's launches so far
};

// lane / join_next: the recording code marks the launches that may run on a
// side stream beside the rest of the step (lane 1), and the first launch that
// needs their results (join_next, consumed by the next recorded launch)
struct LaunchRecorder {
    std::vector<RecordedLaunch> launches;
    int lane = 0;
    bool join_next = false;
};

// the recorder of the calling thread (null: launches run)
extern thread_local LaunchRecorder* g_recorder;

template <int LB, auto F, typename... A>
void record_launch(dim3 grid, dim3 block, size_t shmem, A&&... a) {
    using Tr = BodyTraits<decltype(F)>;
    const typename Tr::Tuple t(std::forward<A>(a)...);
    RecordedLaunch r;
    r.fn = reinterpret_cast<const void*>(&k_batched<LB, F>);
    r.blocks = (int)(grid.x * grid.y * grid.z);
    r.block = block;
    r.shmem = shmem;
    r.args.resize(sizeof(t));
    std::memcpy(r.args.data(), &t, sizeof(t));
    r.lane = g_recorder->lane;
    r.join = g_recorder->join_next;
    g_recorder->join_next = false;
    g_recorder->launches.push_back(std::move(r));
}

}  // namespace sa

// Launch KERNEL (the named wrapper of BODY, both parenthesised when they hold
// template arguments), or record it for the batched driver.
#define SA_LAUNCH(LB, BODY, KERNEL, GRID, BLOCK, SH, STREAM, ...)                    \
    do {                                                                            \
        if (::sa::g_recorder)                                                       \
            ::sa::record_launch<LB, BODY>(dim3(GRID), dim3(BLOCK), (SH), __VA_ARGS__); \
        else                                                                        \
            hipLaunchKernelGGL(KERNEL, dim3(GRID), dim3(BLOCK), (SH), (STREAM), __VA_ARGS__); \
    } while (0)
