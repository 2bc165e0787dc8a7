// Shared device / host helpers of the gfx950 hot path.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>

#include "../../include/sponge_amd.h"

namespace sa {

constexpr int kWave = 64;  // CDNA wavefront

// LTMatrix3 {a11, a21, a22, a31, a32, a33} (SPONGE/utils/vector.hpp:245-252)
struct LT6 {
    float a11, a21, a22, a31, a32, a33;
};

inline LT6 lt_from(const float* p) { return {p[0], p[1], p[2], p[3], p[4], p[5]}; }

inline bool lt_is_orthorhombic(const LT6& c) {
    return c.a21 == 0.0f && c.a31 == 0.0f && c.a32 == 0.0f;
}

#define SA_CHECK(expr)                                          \
    do {                                                        \
        hipError_t sa_err_ = (expr);                            \
        if (sa_err_ != hipSuccess) return (int)sa_err_;         \
    } while (0)

#define SA_LAUNCH_CHECK() SA_CHECK(hipGetLastError())

template <typename T>
inline int dmalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return SPONGE_AMD_ERROR_MALLOC_FAILED;
    }
    return 0;
}

template <typename T>
inline void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

inline int grid_for(long long n, int block) { return (int)((n + block - 1) / block); }

// Minimum-image displacement a - b in the exact operation order of the
// reference's Get_Displacement<Periodic> (SPONGE/utils/boundary.h:22-31):
// dr - floorf(dr * rcell + 0.5f) * cell, with VECTOR * LTMatrix3 =
// {x a11 + y a21 + z a31, y a22 + z a32, z a33}.  FMA contraction is off so
// the neighbor-list criterion is reproduced bit for bit.
template <bool ORTHO>
__device__ __forceinline__ float3 disp_pbc(float3 a, float3 b, const LT6& c, const LT6& r) {
#pragma clang fp contract(off)
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    float3 o;
    if (ORTHO) {
        // the zero off-diagonal products of the general form add exact zeros
        float sx = floorf(dx * r.a11 + 0.5f);
        float sy = floorf(dy * r.a22 + 0.5f);
        float sz = floorf(dz * r.a33 + 0.5f);
        o.x = dx - sx * c.a11;
        o.y = dy - sy * c.a22;
        o.z = dz - sz * c.a33;
    } else {
        float sx = floorf(dx * r.a11 + dy * r.a21 + dz * r.a31 + 0.5f);
        float sy = floorf(dy * r.a22 + dz * r.a32 + 0.5f);
        float sz = floorf(dz * r.a33 + 0.5f);
        o.x = dx - (sx * c.a11 + sy * c.a21 + sz * c.a31);
        o.y = dy - (sy * c.a22 + sz * c.a32);
        o.z = dz - sz * c.a33;
    }
    return o;
}

// The periodic image of x nearest the point C (orthorhombic box).
__device__ __forceinline__ float3 near_image(float3 x, float3 C, const LT6& c, const LT6& r) {
    return make_float3(x.x - c.a11 * rintf((x.x - C.x) * r.a11),
                       x.y - c.a22 * rintf((x.y - C.y) * r.a22),
                       x.z - c.a33 * rintf((x.z - C.z) * r.a33));
}

// r^2 as the reference's VECTOR dot product ((x*x + y*y) + z*z), no FMA.
__device__ __forceinline__ float norm2_exact(float3 d) {
#pragma clang fp contract(off)
    return d.x * d.x + d.y * d.y + d.z * d.z;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// number of set bits of mask in lanes below this one
__device__ __forceinline__ int lanes_below(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, kWave));
    return v;
}

// Scaled complementary error function erfcx(x) = exp(x^2) erfc(x) on
// 0 <= x <= kErfcxMax, as t P(t) with t = 1 / (1 + 0.4 x) and P of degree 7,
// coefficients fitted by iteratively reweighted least squares to minimise the
// largest relative error (2.9e-8 for the fit in FP64, under 4e-7 evaluated in
// FP32; generator and pins: tests/golden/make_erfcx_fit.py,
// tests/test_erfcx_fit.py).  The Ewald direct term then needs one
// exponential, exp(-x^2), for both erfc(x) and its derivative.  The rcp is the
// hardware reciprocal (1 ulp).
constexpr float kErfcxMax = 4.5f;
constexpr float kErfcxT = 0.4f;  // t = 1 / (1 + kErfcxT x)
// erfcx from t itself (callers that form t = 1 / (1 + kErfcxT x) their own way)
__device__ __forceinline__ float erfcx_poly_t(float t) {
    float p = 7.614902268e-02f;
    p = fmaf(p, t, -4.040216792e-01f);
    p = fmaf(p, t, 7.048484256e-01f);
    p = fmaf(p, t, -3.981623732e-01f);
    p = fmaf(p, t, 4.256499008e-01f);
    p = fmaf(p, t, 1.327632050e-01f);
    p = fmaf(p, t, 2.379624228e-01f);
    p = fmaf(p, t, 2.248110469e-01f);
    return p * t;
}
__device__ __forceinline__ float erfcx_poly(float x) {
    return erfcx_poly_t(__builtin_amdgcn_rcpf(1.0f + kErfcxT * x));
}

// Lower-triangle packed pair index (Get_LJ_Type, Lennard_Jones_force.h:17-28):
// max(a,b) * (max(a,b)+1) / 2 + min(a,b).
__host__ __device__ __forceinline__ int lj_pair_index(int a, int b) {
    int hi = a > b ? a : b, lo = a > b ? b : a;
    return (hi * (hi + 1) >> 1) + lo;
}

// Fill kernel used instead of hipMemsetAsync so every step can be captured
// in a HIP graph.
template <typename T>
__global__ void k_fill(T* p, T v, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <typename T>
inline int fill_async(T* p, T v, long long n, hipStream_t s) {
    if (n <= 0) return 0;
    int blocks = grid_for(n, 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill<T>, dim3(blocks), dim3(256), 0, s, p, v, n);
    SA_LAUNCH_CHECK();
    return 0;
}

// Smallest float t with sqrtf(t) >= rc: for r2 >= 0, (sqrtf(r2) < rc) <=>
// (r2 < t).  sqrtf is correctly rounded and monotone, so this reproduces the
// reference's strict "dr_abs < cutoff" test on r2 without a square root.
inline float sqrt_threshold(float rc) {
    float t = rc * rc;
    // walk to the boundary; a few ulps at most
    for (int it = 0; it < 64 && std::sqrt(t) >= rc; ++it) t = std::nextafter(t, 0.0f);
    for (int it = 0; it < 64 && std::sqrt(t) < rc; ++it) t = std::nextafter(t, 1e30f);
    return t;
}

}  // namespace sa
