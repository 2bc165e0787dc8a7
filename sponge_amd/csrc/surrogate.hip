// Learned reciprocal-force surrogate: a per-atom MLP over local charge-
// environment features that stands in for the PME reciprocal force on the
// steps between full PME evaluations.  It has no counterpart in the
// reference; it sits beside PM_force (SPONGE/PM_force/PM_force.cpp), whose
// PME_Reciprocal_Force_With_Energy_And_Virial (:1070-1169) stays the
// fallback and the reference force every sampled step is checked against.
//
// Features of atom i (partners j on the neighbor list with r < Rf; d = r_j -
// r_i, u = r / Rf, fc = (1 - u^2)^2, w_j = q_j fc, x = 2u - 1, T_k Chebyshev):
//   R_k = sum w_j T_k(x), k < 32;  P_a = sum w_j T_a(x) d / r, a < 5;
//   S = sum w_j d d^T / r^2, Q = S - R_0 / 3 I;
//   angular = [P_a . P_b, a <= b] (15) + |Q|_F^2 (1).
// Basis vectors: U_0 = sum q_j fc phi'(r) d / r over the list partners,
// U_1 = sum q_j phi'(r) d / r over the excluded partners, U_2 = P_0, with
// phi(r) = erf(beta r) / r.  Predicted reciprocal force of atom i:
//   F_i = q_i (c_0 U_0 + c_1 U_1 + c_2 U_2),  c = MLP(normalised features),
// MLP 48 -> 64 -> 64 -> 3 with SiLU.  The coefficients (1, 1, 0) are the
// near-field reciprocal force itself (the physics baseline the network
// starts from and is compared with).
//
// Layout in HBM: features [N][48] FP32 and basis [N][3] float4 in the neighbor
// list's sorted order; model parameters FP32 (master) with BF16 copies in the
// MFMA operand order; training frames (features, basis, target force + charge)
// kept resident, [F][N] each.  The MLP runs on v_mfma_f32_16x16x16_bf16: a
// wave owns 16 atoms and computes every layer transposed (outputs x atoms),
// so that one layer's accumulator fragment is the next layer's B operand
// without a shuffle.
#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "internal.h"
#include "timing.h"

using namespace sa;

namespace {

constexpr int kRadial = 32, kVec = 5, kAngular = 16, kFeat = kRadial + kAngular;
constexpr int kHidden = 64, kOut = 3;
constexpr int oW1 = 0, oB1 = oW1 + kHidden * kFeat, oW2 = oB1 + kHidden;
constexpr int oB2 = oW2 + kHidden * kHidden, oW3 = oB2 + kHidden, oB3 = oW3 + kOut * kHidden;
constexpr int kParams = oB3 + kOut;
constexpr int kModel = kParams + 2 * kFeat;  // + feature mean and inverse std
constexpr int kFeatStage = 3584;
constexpr float kTwoOverSqrtPiF = 1.1283791670955126f;

}  // namespace

struct sponge_rsur {
    int N = 0;
    float rf = 0, beta = 0;
    sponge_rsur_options opt{};
    hipStream_t stream = nullptr;
    int stage = 0;
    const int* sidx = nullptr;     // sorted -> local order of the last feature call's list
    const float4* pos_last = nullptr;  // sorted positions + charges the last features used
    float4* d_pos = nullptr;       // [N] sorted x, y, z, charge
    float* d_feat = nullptr;       // [N][48] sorted order
    float4* d_basis = nullptr;     // [N][3] sorted order
    float4* d_uex = nullptr;       // [N] exclusion field, local order
    float* d_params = nullptr;     // [kParams] FP32 master copy
    float* d_norm = nullptr;       // [96] feature mean, inverse std
    unsigned short* d_w1 = nullptr;   // [64][48] bf16
    unsigned short* d_w2 = nullptr;   // [64][64]
    unsigned short* d_w2t = nullptr;  // [64][64] transposed
    unsigned short* d_w3 = nullptr;   // [16][64], rows >= 3 zero
    unsigned short* d_w3t = nullptr;  // [64][16], columns >= 3 zero
    float* d_grad = nullptr;       // [kParams]
    float* d_m = nullptr;          // Adam moments
    float* d_v = nullptr;
    double* d_acc = nullptr;       // [0] loss, [1..3] sums, [8..104) feature stats, [104..110) evaluation
    long long adam_t = 0;
    int n_frames = 0;
    float* d_ffeat = nullptr;      // [F][N][48]
    float4* d_fbasis = nullptr;    // [F][N][3]
    float4* d_ftarget = nullptr;   // [F][N] target force, w = charge
};

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {
    // round to nearest even (finite inputs)
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 v;
    v[0] = (short)f2bf(a);
    v[1] = (short)f2bf(b);
    v[2] = (short)f2bf(c);
    v[3] = (short)f2bf(d);
    return v;
}

__device__ __forceinline__ bf16x4 load4(const unsigned short* p) {
    const uint2 u = *reinterpret_cast<const uint2*>(p);
    bf16x4 v;
    v[0] = (short)(u.x & 0xffffu);
    v[1] = (short)(u.x >> 16);
    v[2] = (short)(u.y & 0xffffu);
    v[3] = (short)(u.y >> 16);
    return v;
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// Wavefront sum with DPP row operations (VALU only, no LDS traffic): pairs,
// quads, half-rows and rows by quad_perm / mirrors, then row_bcast:15 and
// row_bcast:31 carry the row sums up to lane 63, read back as a uniform value.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_read(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, true));
}

__device__ __forceinline__ float wave_total(float v) {
    v += dpp_read<0xb1, 0xf>(v);   // quad_perm [1, 0, 3, 2]
    v += dpp_read<0x4e, 0xf>(v);   // quad_perm [2, 3, 0, 1]
    v += dpp_read<0x141, 0xf>(v);  // row_half_mirror
    v += dpp_read<0x140, 0xf>(v);  // row_mirror
    v += dpp_read<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_read<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// d/dr [erf(beta r) / r] / r, times the caller's weights
__device__ __forceinline__ float dphi_over_r(float r, float rinv, float beta) {
    const float x = beta * r;
    return (kTwoOverSqrtPiF * x * __expf(-x * x) - erff(x)) * rinv * rinv * rinv;
}

__global__ void k_rs_pack(int N, const int* __restrict__ sidx, const float* __restrict__ crd,
                          const float* __restrict__ charge, float4* __restrict__ pos) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const int i = sidx[p];
    pos[p] = make_float4(crd[3 * i], crd[3 * i + 1], crd[3 * i + 2], charge[i]);
}

// U_1: the reciprocal pair field of the excluded partners (both directions
// of the reference's half exclusion list)
template <bool ORTHO>
__global__ void k_rs_exfield(int N, const float* __restrict__ crd, const float* __restrict__ charge,
                             const int* __restrict__ ex_start, const int* __restrict__ ex_num,
                             const int* __restrict__ ex_list, LT6 c, LT6 r, float beta,
                             float4* __restrict__ uex) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int n = ex_num[i];
    if (n <= 0) return;
    const float3 pi = make_float3(crd[3 * i], crd[3 * i + 1], crd[3 * i + 2]);
    const float qi = charge[i];
    const int* lst = ex_list + ex_start[i];
    float fx = 0.f, fy = 0.f, fz = 0.f;
    for (int e = 0; e < n; ++e) {
        const int j = lst[e];
        const float3 pj = make_float3(crd[3 * j], crd[3 * j + 1], crd[3 * j + 2]);
        const float3 d = disp_pbc<ORTHO>(pj, pi, c, r);
        const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
        const float rinv = 1.0f / sqrtf(r2);
        const float g = dphi_over_r(r2 * rinv, rinv, beta);
        const float qj = charge[j];
        fx += qj * g * d.x;
        fy += qj * g * d.y;
        fz += qj * g * d.z;
        unsafeAtomicAdd(&uex[j].x, -qi * g * d.x);
        unsafeAtomicAdd(&uex[j].y, -qi * g * d.y);
        unsafeAtomicAdd(&uex[j].z, -qi * g * d.z);
    }
    unsafeAtomicAdd(&uex[i].x, fx);
    unsafeAtomicAdd(&uex[i].y, fy);
    unsafeAtomicAdd(&uex[i].z, fz);
}

// Features and basis vectors, one workgroup per neighbor-list cell: the
// cell's candidate stream (positions and charges) is staged in LDS once; one
// wavefront per row walks the row's 16-bit stream positions, one partner per
// lane and pass, and the per-lane sums are reduced across the wavefront.
template <bool ORTHO, bool STAGED>
__device__ __forceinline__ void rs_row(int p, int lane, int n, const unsigned short* __restrict__ row,
                                       const float4* src, int nr, const int* r_first,
                                       const int* r_off, float4 pi4, const LT6& c, const LT6& r,
                                       float rf2, float inv_rf, float beta,
                                       float* __restrict__ feat, float4* __restrict__ basis,
                                       const float4* __restrict__ uex_i) {
    const float3 pi = make_float3(pi4.x, pi4.y, pi4.z);
    float R[kRadial];
#pragma unroll
    for (int k = 0; k < kRadial; ++k) R[k] = 0.f;
    float P[kVec][3];
#pragma unroll
    for (int a = 0; a < kVec; ++a) P[a][0] = P[a][1] = P[a][2] = 0.f;
    float S0 = 0.f, S1 = 0.f, S2 = 0.f, S3 = 0.f, S4 = 0.f, S5 = 0.f;
    float U0 = 0.f, U1 = 0.f, U2 = 0.f;
    // the next pass's entry is loaded before the current one is used, so
    // that the list's memory latency overlaps the arithmetic
    int cur = n > 0 ? (int)row[lane < n ? lane : n - 1] : 0;
    for (int e = lane; e - lane < n; e += kWave) {
        const bool has = e < n;
        const int en = e + kWave;
        int nxt = (int)row[en < n ? en : n - 1];
        const int t = cur;
        const int q = STAGED ? t : stream_to_sorted(t, nr, r_first, r_off);
        const float4 pj = src[q];
        const float3 d = disp_pbc<ORTHO>(make_float3(pj.x, pj.y, pj.z), pi, c, r);
        const float r2 = norm2_exact(d);
        const bool in = has && r2 < rf2;
        const float rinv = in ? __builtin_amdgcn_rsqf(r2) : 0.f;
        const float rr = r2 * rinv;
        const float u = rr * inv_rf;
        const float fb = fmaf(-u, u, 1.0f);
        const float fc = fb * fb;
        const float w = in ? pj.w * fc : 0.f;
        const float x = 2.0f * u - 1.0f;
        // Chebyshev recurrence T_{k+1} = 2x T_k - T_{k-1}
        float tm = 1.0f, tk = x;
        R[0] += w;
        R[1] = fmaf(w, x, R[1]);
        float wr = w * rinv;
        P[0][0] = fmaf(wr, d.x, P[0][0]);
        P[0][1] = fmaf(wr, d.y, P[0][1]);
        P[0][2] = fmaf(wr, d.z, P[0][2]);
        float a = wr * x;
        P[1][0] = fmaf(a, d.x, P[1][0]);
        P[1][1] = fmaf(a, d.y, P[1][1]);
        P[1][2] = fmaf(a, d.z, P[1][2]);
#pragma unroll
        for (int k = 2; k < kRadial; ++k) {
            const float tn = fmaf(2.0f * x, tk, -tm);
            tm = tk;
            tk = tn;
            R[k] = fmaf(w, tn, R[k]);
            if (k < kVec) {
                const float b = wr * tn;
                P[k][0] = fmaf(b, d.x, P[k][0]);
                P[k][1] = fmaf(b, d.y, P[k][1]);
                P[k][2] = fmaf(b, d.z, P[k][2]);
            }
        }
        const float s = wr * rinv;
        S0 = fmaf(s * d.x, d.x, S0);
        S1 = fmaf(s * d.x, d.y, S1);
        S2 = fmaf(s * d.y, d.y, S2);
        S3 = fmaf(s * d.x, d.z, S3);
        S4 = fmaf(s * d.y, d.z, S4);
        S5 = fmaf(s * d.z, d.z, S5);
        const float g = in ? pj.w * fc * dphi_over_r(rr, rinv, beta) : 0.f;
        U0 = fmaf(g, d.x, U0);
        U1 = fmaf(g, d.y, U1);
        U2 = fmaf(g, d.z, U2);
        asm volatile("" : "+v"(nxt));
        cur = nxt;
    }
#pragma unroll
    for (int k = 0; k < kRadial; ++k) R[k] = wave_total(R[k]);
#pragma unroll
    for (int a = 0; a < kVec; ++a) {
        P[a][0] = wave_total(P[a][0]);
        P[a][1] = wave_total(P[a][1]);
        P[a][2] = wave_total(P[a][2]);
    }
    S0 = wave_total(S0);
    S1 = wave_total(S1);
    S2 = wave_total(S2);
    S3 = wave_total(S3);
    S4 = wave_total(S4);
    S5 = wave_total(S5);
    U0 = wave_total(U0);
    U1 = wave_total(U1);
    U2 = wave_total(U2);
    if (lane == 0) {
        float* f = feat + (size_t)kFeat * p;
#pragma unroll
        for (int k = 0; k < kRadial; ++k) f[k] = R[k];
        int o = kRadial;
#pragma unroll
        for (int a = 0; a < kVec; ++a)
#pragma unroll
            for (int b = a; b < kVec; ++b)
                f[o++] = P[a][0] * P[b][0] + P[a][1] * P[b][1] + P[a][2] * P[b][2];
        const float tr = R[0] * (1.0f / 3.0f);
        const float q0 = S0 - tr, q2 = S2 - tr, q5 = S5 - tr;
        f[o] = q0 * q0 + q2 * q2 + q5 * q5 + 2.0f * (S1 * S1 + S3 * S3 + S4 * S4);
        float4* bs = basis + (size_t)3 * p;
        bs[0] = make_float4(U0, U1, U2, 0.f);
        const float4 ue = *uex_i;
        bs[1] = make_float4(ue.x, ue.y, ue.z, 0.f);
        bs[2] = make_float4(P[0][0], P[0][1], P[0][2], 0.f);
    }
}

template <bool ORTHO>
__global__ void __launch_bounds__(512) k_rs_features(
    const int* __restrict__ start, const int* __restrict__ nranges, int max_ranges,
    const int* __restrict__ g_rfirst, const int* __restrict__ g_roff,
    const int* __restrict__ list_count, const unsigned short* __restrict__ list, int cap,
    const float4* __restrict__ pos, const int* __restrict__ sidx, LT6 c, LT6 r, float rf2,
    float inv_rf, float beta, int stage, const float4* __restrict__ uex, float* __restrict__ feat,
    float4* __restrict__ basis) {
    extern __shared__ float4 cpos[];
    __shared__ int r_first[kMaxRanges];
    __shared__ int r_off[kMaxRanges + 1];
    const int cell = blockIdx.x;
    const int a0 = start[cell], a1 = start[cell + 1];
    if (a0 == a1) return;
    const int nr = nranges[cell];
    for (int k = threadIdx.x; k < nr; k += blockDim.x) r_first[k] = g_rfirst[(size_t)cell * max_ranges + k];
    for (int k = threadIdx.x; k <= nr; k += blockDim.x) r_off[k] = g_roff[(size_t)cell * (max_ranges + 1) + k];
    __syncthreads();
    const int total = r_off[nr];
    const bool staged = total <= stage;
    if (staged) stage_stream(nr, r_first, r_off, total, [&](int t, int q) { cpos[t] = pos[q]; });
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave, nwave = blockDim.x / kWave;
    for (int p = a0 + wave; p < a1; p += nwave) {
        const int n = list_count[p];
        const unsigned short* row = list + (size_t)p * cap;
        const float4* ue = uex + sidx[p];
        if (staged)
            rs_row<ORTHO, true>(p, lane, n, row, cpos, nr, r_first, r_off, pos[p], c, r, rf2, inv_rf,
                                beta, feat, basis, ue);
        else
            rs_row<ORTHO, false>(p, lane, n, row, pos, nr, r_first, r_off, pos[p], c, r, rf2, inv_rf,
                                 beta, feat, basis, ue);
    }
}

// ----------------------------------------------------------------------------
// MLP on MFMA.  Lane l of a wave: atom column l & 15 of the wave's 16-atom
// tile, k-group kq = l >> 4.  Operand fragments of v_mfma_f32_16x16x16_bf16:
// A (M x K): row l & 15, k = 4 kq + j;  B (K x N): column l & 15, k = 4 kq + j;
// C (M x N): column l & 15, row 4 kq + j  (j = 0..3, the vector element).
// ----------------------------------------------------------------------------

struct Fwd {
    f32x4 z1[4], z2[4], c;
    bf16x4 x[3], h1[4], h2[4];
};

// Forward pass of one tile (rows p0 .. p0+15 of feat; rows >= N read as zero).
__device__ __forceinline__ void mlp_forward(int p0, int N, int lane, const float* __restrict__ feat,
                                            const float* __restrict__ norm,
                                            const unsigned short* __restrict__ w1,
                                            const unsigned short* __restrict__ w2,
                                            const unsigned short* __restrict__ w3,
                                            const float* __restrict__ params, Fwd& f) {
    const int col = lane & 15, kq = lane >> 4;
    const int a = p0 + col;
    const bool ok = a < N;
#pragma unroll
    for (int kc = 0; kc < 3; ++kc) {
        const int k = kc * 16 + kq * 4;
        float4 v = ok ? *reinterpret_cast<const float4*>(feat + (size_t)kFeat * a + k)
                      : make_float4(0.f, 0.f, 0.f, 0.f);
        if (!ok) {
            // padding rows: normalised zeros
            v = make_float4(norm[k], norm[k + 1], norm[k + 2], norm[k + 3]);
        }
        f.x[kc] = pack4((v.x - norm[k]) * norm[kFeat + k], (v.y - norm[k + 1]) * norm[kFeat + k + 1],
                        (v.z - norm[k + 2]) * norm[kFeat + k + 2],
                        (v.w - norm[k + 3]) * norm[kFeat + k + 3]);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int o = t * 16 + kq * 4;
        f32x4 acc = {params[oB1 + o], params[oB1 + o + 1], params[oB1 + o + 2], params[oB1 + o + 3]};
#pragma unroll
        for (int kc = 0; kc < 3; ++kc)
            acc = mfma16(load4(w1 + (size_t)(t * 16 + col) * kFeat + kc * 16 + kq * 4), f.x[kc], acc);
        f.z1[t] = acc;
        f.h1[t] = pack4(silu(acc[0]), silu(acc[1]), silu(acc[2]), silu(acc[3]));
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int o = t * 16 + kq * 4;
        f32x4 acc = {params[oB2 + o], params[oB2 + o + 1], params[oB2 + o + 2], params[oB2 + o + 3]};
#pragma unroll
        for (int kc = 0; kc < 4; ++kc)
            acc = mfma16(load4(w2 + (size_t)(t * 16 + col) * kHidden + kc * 16 + kq * 4), f.h1[kc], acc);
        f.z2[t] = acc;
        f.h2[t] = pack4(silu(acc[0]), silu(acc[1]), silu(acc[2]), silu(acc[3]));
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    if (kq == 0) {
        acc[0] = params[oB3];
        acc[1] = params[oB3 + 1];
        acc[2] = params[oB3 + 2];
    }
#pragma unroll
    for (int kc = 0; kc < 4; ++kc)
        acc = mfma16(load4(w3 + (size_t)col * kHidden + kc * 16 + kq * 4), f.h2[kc], acc);
    f.c = acc;  // lanes kq == 0: c_0, c_1, c_2 of atom p0 + col
}

// Inference: frc[local] += q (c . U), over tiles of 16 sorted atoms.
__global__ void __launch_bounds__(256) k_rs_predict(
    int N, int ntiles, const float* __restrict__ feat, const float4* __restrict__ basis,
    const float4* __restrict__ pos, const int* __restrict__ sidx, const float* __restrict__ norm,
    const unsigned short* __restrict__ w1, const unsigned short* __restrict__ w2,
    const unsigned short* __restrict__ w3, const float* __restrict__ params, float* __restrict__ frc,
    int assign, const float* __restrict__ anchor) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    for (int tile = wave; tile < ntiles; tile += nw) {
        Fwd f;
        mlp_forward(tile * 16, N, lane, feat, norm, w1, w2, w3, params, f);
        const int a = tile * 16 + (lane & 15);
        if ((lane >> 4) == 0 && a < N) {
            const float q = pos[a].w;
            const float4 u0 = basis[3 * (size_t)a], u1 = basis[3 * (size_t)a + 1],
                         u2 = basis[3 * (size_t)a + 2];
            float fx = q * (f.c[0] * u0.x + f.c[1] * u1.x + f.c[2] * u2.x);
            float fy = q * (f.c[0] * u0.y + f.c[1] * u1.y + f.c[2] * u2.y);
            float fz = q * (f.c[0] * u0.z + f.c[1] * u1.z + f.c[2] * u2.z);
            const size_t i = (size_t)sidx[a];
            if (anchor) {
                fx += anchor[3 * i];
                fy += anchor[3 * i + 1];
                fz += anchor[3 * i + 2];
            }
            float* o = frc + 3 * i;
            if (assign) {
                o[0] = fx;
                o[1] = fy;
                o[2] = fz;
            } else {
                o[0] += fx;
                o[1] += fy;
                o[2] += fz;
            }
        }
    }
}

// Training pass over one frame: forward, loss, backward, and the weight and
// bias gradients accumulated over the wave's tiles, then added to grad.
// One wavefront per workgroup (the LDS transposes are wave-private).
__device__ __forceinline__ void to_lds(float (*T)[17], const f32x4& v, int lane) {
    const int col = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) T[kq * 4 + j][col] = v[j];
}

// A fragment (row l & 15, k = 4 kq + j) of the transpose of a C-layout tile
// written by to_lds: element (row = l & 15, k) = T[row][k].
__device__ __forceinline__ bf16x4 from_lds_t(float (*T)[17], int lane) {
    const int col = lane & 15, kq = lane >> 4;
    return pack4(T[col][kq * 4], T[col][kq * 4 + 1], T[col][kq * 4 + 2], T[col][kq * 4 + 3]);
}

__global__ void __launch_bounds__(64) k_rs_train(
    int N, int ntiles, const float* __restrict__ feat, const float4* __restrict__ basis,
    const float4* __restrict__ target, const float* __restrict__ norm,
    const unsigned short* __restrict__ w1, const unsigned short* __restrict__ w2,
    const unsigned short* __restrict__ w2t, const unsigned short* __restrict__ w3,
    const unsigned short* __restrict__ w3t, const float* __restrict__ params,
    float* __restrict__ grad, double* __restrict__ loss) {
    __shared__ float T[16][17];
    const int lane = threadIdx.x;
    const int col = lane & 15, kq = lane >> 4;
    const float scale = 2.0f / (3.0f * (float)N);
    f32x4 gW1[4][3], gW2[4][4], gW3[4];
    float gb1[4][4], gb2[4][4], gb3[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            gb1[t][j] = gb2[t][j] = 0.f;
            gW3[t][j] = 0.f;
#pragma unroll
            for (int kc = 0; kc < 3; ++kc) gW1[t][kc][j] = 0.f;
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) gW2[t][kc][j] = 0.f;
        }
        gb3[t] = 0.f;
    }
    float lsum = 0.f;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int p0 = tile * 16;
        Fwd f;
        mlp_forward(p0, N, lane, feat, norm, w1, w2, w3, params, f);
        // loss and dL/dc (lanes kq == 0 own the atom's c)
        const int a = p0 + col;
        float dc0 = 0.f, dc1 = 0.f, dc2 = 0.f;
        if (kq == 0 && a < N) {
            const float4 u0 = basis[3 * (size_t)a], u1 = basis[3 * (size_t)a + 1],
                         u2 = basis[3 * (size_t)a + 2];
            const float4 tg = target[a];
            const float q = tg.w;
            const float rx = q * (f.c[0] * u0.x + f.c[1] * u1.x + f.c[2] * u2.x) - tg.x;
            const float ry = q * (f.c[0] * u0.y + f.c[1] * u1.y + f.c[2] * u2.y) - tg.y;
            const float rz = q * (f.c[0] * u0.z + f.c[1] * u1.z + f.c[2] * u2.z) - tg.z;
            lsum += rx * rx + ry * ry + rz * rz;
            dc0 = scale * q * (rx * u0.x + ry * u0.y + rz * u0.z);
            dc1 = scale * q * (rx * u1.x + ry * u1.y + rz * u1.z);
            dc2 = scale * q * (rx * u2.x + ry * u2.y + rz * u2.z);
        }
        gb3[0] += dc0;
        gb3[1] += dc1;
        gb3[2] += dc2;
        // dC as a B operand (k = output 4 kq + j, column = atom): lanes kq == 0
        const bf16x4 dcb = pack4(dc0, dc1, dc2, 0.f);
        // dW3 += dC . H2^T: A = dC (output x atom), B = H2^T (atom x hidden)
        f32x4 dcC = {dc0, dc1, dc2, 0.f};  // C layout: row = output 4 kq + j, column = atom
        __syncthreads();
        to_lds(T, dcC, lane);
        __syncthreads();
        const bf16x4 dcA = from_lds_t(T, lane);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 h = {0.f, 0.f, 0.f, 0.f};
            // H2 tile t in C layout (row = hidden, column = atom) from its bf16 fragment
            h[0] = __uint_as_float(((unsigned)(unsigned short)f.h2[t][0]) << 16);
            h[1] = __uint_as_float(((unsigned)(unsigned short)f.h2[t][1]) << 16);
            h[2] = __uint_as_float(((unsigned)(unsigned short)f.h2[t][2]) << 16);
            h[3] = __uint_as_float(((unsigned)(unsigned short)f.h2[t][3]) << 16);
            __syncthreads();
            to_lds(T, h, lane);
            __syncthreads();
            gW3[t] = mfma16(dcA, from_lds_t(T, lane), gW3[t]);
        }
        // dH2 = W3^T dC, dZ2 = dH2 * silu'(Z2)
        f32x4 dz2[4];
        bf16x4 dz2b[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma16(load4(w3t + (size_t)(t * 16 + col) * 16 + kq * 4), dcb, acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[j] *= dsilu(f.z2[t][j]);
                gb2[t][j] += acc[j];
            }
            dz2[t] = acc;
            dz2b[t] = pack4(acc[0], acc[1], acc[2], acc[3]);
        }
        // dW2 += dZ2 . H1^T
#pragma unroll
        for (int to = 0; to < 4; ++to) {
            __syncthreads();
            to_lds(T, dz2[to], lane);
            __syncthreads();
            const bf16x4 A = from_lds_t(T, lane);
#pragma unroll
            for (int ti = 0; ti < 4; ++ti) {
                f32x4 h;
                h[0] = __uint_as_float(((unsigned)(unsigned short)f.h1[ti][0]) << 16);
                h[1] = __uint_as_float(((unsigned)(unsigned short)f.h1[ti][1]) << 16);
                h[2] = __uint_as_float(((unsigned)(unsigned short)f.h1[ti][2]) << 16);
                h[3] = __uint_as_float(((unsigned)(unsigned short)f.h1[ti][3]) << 16);
                __syncthreads();
                to_lds(T, h, lane);
                __syncthreads();
                gW2[to][ti] = mfma16(A, from_lds_t(T, lane), gW2[to][ti]);
            }
        }
        // dH1 = W2^T dZ2, dZ1 = dH1 * silu'(Z1); dW1 += dZ1 . X
        bf16x4 xb[3];
#pragma unroll
        for (int kc = 0; kc < 3; ++kc) {
            // B = X (atom x feature): column = feature kc*16 + col, k = atom 4 kq + j
            float v[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int aa = p0 + kq * 4 + j, k = kc * 16 + col;
                v[j] = aa < N ? (feat[(size_t)kFeat * aa + k] - norm[k]) * norm[kFeat + k] : 0.f;
            }
            xb[kc] = pack4(v[0], v[1], v[2], v[3]);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kc = 0; kc < 4; ++kc)
                acc = mfma16(load4(w2t + (size_t)(t * 16 + col) * kHidden + kc * 16 + kq * 4), dz2b[kc], acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[j] *= dsilu(f.z1[t][j]);
                gb1[t][j] += acc[j];
            }
            __syncthreads();
            to_lds(T, acc, lane);
            __syncthreads();
            const bf16x4 A = from_lds_t(T, lane);
#pragma unroll
            for (int kc = 0; kc < 3; ++kc) gW1[t][kc] = mfma16(A, xb[kc], gW1[t][kc]);
        }
    }
    // flush: C-layout element j of lane l is (row 4 kq + j, column col) of its tile
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int o = t * 16 + kq * 4 + j;
#pragma unroll
            for (int kc = 0; kc < 3; ++kc)
                unsafeAtomicAdd(&grad[oW1 + o * kFeat + kc * 16 + col], gW1[t][kc][j]);
#pragma unroll
            for (int kc = 0; kc < 4; ++kc)
                unsafeAtomicAdd(&grad[oW2 + o * kHidden + kc * 16 + col], gW2[t][kc][j]);
            const int o3 = kq * 4 + j;
            if (o3 < kOut) unsafeAtomicAdd(&grad[oW3 + o3 * kHidden + t * 16 + col], gW3[t][j]);
        }
    // biases: sum the 16 atom columns of each row
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float s1 = gb1[t][j], s2 = gb2[t][j];
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                s1 += __shfl_xor(s1, off, 64);
                s2 += __shfl_xor(s2, off, 64);
            }
            if (col == 0) {
                unsafeAtomicAdd(&grad[oB1 + t * 16 + kq * 4 + j], s1);
                unsafeAtomicAdd(&grad[oB2 + t * 16 + kq * 4 + j], s2);
            }
        }
#pragma unroll
    for (int m = 0; m < kOut; ++m) {
        float s = gb3[m];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) unsafeAtomicAdd(&grad[oB3 + m], s);
    }
    const float ls = wave_sum(lsum);
    if (lane == 0) atomicAdd(loss, (double)ls / (3.0 * (double)N));
}

// BF16 copies of the weights in the kernels' operand order
__device__ __forceinline__ void cast_param(int i, float w, unsigned short* w1, unsigned short* w2,
                                           unsigned short* w2t, unsigned short* w3,
                                           unsigned short* w3t) {
    const unsigned short b = f2bf(w);
    if (i < oB1) {
        w1[i] = b;
    } else if (i >= oW2 && i < oB2) {
        const int k = i - oW2, o = k / kHidden, in = k % kHidden;
        w2[k] = b;
        w2t[in * kHidden + o] = b;
    } else if (i >= oW3 && i < oB3) {
        const int k = i - oW3, o = k / kHidden, in = k % kHidden;
        w3[o * kHidden + in] = b;
        w3t[in * 16 + o] = b;
    }
}

__global__ void k_rs_cast(const float* __restrict__ params, unsigned short* w1, unsigned short* w2,
                          unsigned short* w2t, unsigned short* w3, unsigned short* w3t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kParams) cast_param(i, params[i], w1, w2, w2t, w3, w3t);
}

__global__ void k_rs_adam(float* __restrict__ params, float* __restrict__ grad, float* __restrict__ m,
                          float* __restrict__ v, float lr, float c1, float c2, unsigned short* w1,
                          unsigned short* w2, unsigned short* w2t, unsigned short* w3,
                          unsigned short* w3t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    const float g = grad[i];
    grad[i] = 0.f;
    const float mi = 0.9f * m[i] + 0.1f * g;
    const float vi = 0.999f * v[i] + 0.001f * g * g;
    m[i] = mi;
    v[i] = vi;
    const float w = params[i] - lr * (mi * c1) / (sqrtf(vi * c2) + 1e-8f);
    params[i] = w;
    cast_param(i, w, w1, w2, w2t, w3, w3t);
}

// feature sums over stored frames (FP64): acc[0..47] sum, acc[48..95] sum of squares
__global__ void __launch_bounds__(256) k_rs_stats(long long rows, const float* __restrict__ feat,
                                                  double* __restrict__ acc) {
    __shared__ double s1[256], s2[256];
    const int k = threadIdx.x % kFeat;
    const int lanes = (256 / kFeat) * kFeat;  // 240 threads, 5 rows per pass
    double a = 0, b = 0;
    if (threadIdx.x < lanes) {
        for (long long rr = (long long)blockIdx.x * 5 + threadIdx.x / kFeat; rr < rows;
             rr += (long long)gridDim.x * 5) {
            const double x = feat[rr * kFeat + k];
            a += x;
            b += x * x;
        }
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    if (threadIdx.x < kFeat) {
        for (int t = threadIdx.x + kFeat; t < lanes; t += kFeat) {
            a += s1[t];
            b += s2[t];
        }
        atomicAdd(&acc[threadIdx.x], a);
        atomicAdd(&acc[kFeat + threadIdx.x], b);
    }
}

__global__ void k_rs_norm(long long rows, const double* __restrict__ acc, float* __restrict__ norm) {
    const int k = threadIdx.x;
    if (k >= kFeat) return;
    const double mu = acc[k] / (double)rows;
    const double var = std::max(acc[kFeat + k] / (double)rows - mu * mu, 0.0);
    const double sd = std::sqrt(var);
    norm[k] = (float)mu;
    norm[kFeat + k] = sd > 1e-12 * (1.0 + std::fabs(mu)) ? (float)(1.0 / sd) : 1.0f;
}

// FP64 sums of a force array: [0..2] components (for the mean)
__global__ void __launch_bounds__(256) k_rs_sum3(int N, const float* __restrict__ f,
                                                 double* __restrict__ out) {
    __shared__ double s[3][256];
    double a = 0, b = 0, c = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        a += f[3 * i];
        b += f[3 * i + 1];
        c += f[3 * i + 2];
    }
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    s[2][threadIdx.x] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int m = 0; m < 3; ++m) s[m][threadIdx.x] += s[m][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int m = 0; m < 3; ++m) atomicAdd(&out[m], s[m][0]);
}

// store the current features / basis and the target (rec - mean) with the
// charge, in atom order (frames of different list builds line up)
__global__ void k_rs_store(int N, const int* __restrict__ sidx, const float* __restrict__ feat,
                           const float4* __restrict__ basis, const float4* __restrict__ pos,
                           const float* __restrict__ rec, const double* __restrict__ sums,
                           float* __restrict__ ffeat, float4* __restrict__ fbasis,
                           float4* __restrict__ ftarget) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const int i = sidx[p];
    for (int k = 0; k < kFeat; ++k) ffeat[(size_t)kFeat * i + k] = feat[(size_t)kFeat * p + k];
    for (int m = 0; m < 3; ++m) fbasis[3 * (size_t)i + m] = basis[3 * (size_t)p + m];
    const float mx = (float)(sums[0] / N), my = (float)(sums[1] / N), mz = (float)(sums[2] / N);
    ftarget[i] = make_float4(rec[3 * i] - mx, rec[3 * i + 1] - my, rec[3 * i + 2] - mz, pos[p].w);
}

// model prediction q (c . U) of one atom of a stored frame (lanes kq == 0)
__device__ __forceinline__ float3 frame_pred(const Fwd& f, const float4* basis, const float4* target,
                                             int a) {
    const float4 u0 = basis[3 * (size_t)a], u1 = basis[3 * (size_t)a + 1], u2 = basis[3 * (size_t)a + 2];
    const float q = target[a].w;
    return make_float3(q * (f.c[0] * u0.x + f.c[1] * u1.x + f.c[2] * u2.x),
                       q * (f.c[0] * u0.y + f.c[1] * u1.y + f.c[2] * u2.y),
                       q * (f.c[0] * u0.z + f.c[1] * u1.z + f.c[2] * u2.z));
}

__device__ __forceinline__ float3 frame_base(const float4* basis, const float4* target, int a) {
    const float4 u0 = basis[3 * (size_t)a], u1 = basis[3 * (size_t)a + 1];
    const float q = target[a].w;
    return make_float3(q * (u0.x + u1.x), q * (u0.y + u1.y), q * (u0.z + u1.z));
}

__device__ __forceinline__ float sq3(float3 a, float4 t) {
    const float x = a.x - t.x, y = a.y - t.y, z = a.z - t.z;
    return x * x + y * y + z * z;
}

// evaluation sums over a stored frame f (and, with lag > 0, the frame
// f - lag it is anchored on): [0] |model - T|^2, [1] |T|^2, [2] |baseline -
// T|^2 (c = 1, 1, 0); anchored on the earlier frame's exact target:
// [3] |model + (T' - model') - T|^2, [4] the same for the baseline,
// [5] |T' - T|^2 (the stale force)
__global__ void __launch_bounds__(256) k_rs_eval(
    int N, int ntiles, const float* __restrict__ feat, const float4* __restrict__ basis,
    const float4* __restrict__ target, const float* __restrict__ feat0,
    const float4* __restrict__ basis0, const float4* __restrict__ target0,
    const float* __restrict__ norm, const unsigned short* __restrict__ w1,
    const unsigned short* __restrict__ w2, const unsigned short* __restrict__ w3,
    const float* __restrict__ params, double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int nw = (gridDim.x * blockDim.x) >> 6;
    float e[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int tile = wave; tile < ntiles; tile += nw) {
        Fwd f;
        mlp_forward(tile * 16, N, lane, feat, norm, w1, w2, w3, params, f);
        Fwd f0;
        if (feat0) mlp_forward(tile * 16, N, lane, feat0, norm, w1, w2, w3, params, f0);
        const int a = tile * 16 + (lane & 15);
        if ((lane >> 4) == 0 && a < N) {
            const float4 tg = target[a];
            const float3 pm = frame_pred(f, basis, target, a), pb = frame_base(basis, target, a);
            e[0] += sq3(pm, tg);
            e[1] += tg.x * tg.x + tg.y * tg.y + tg.z * tg.z;
            e[2] += sq3(pb, tg);
            if (feat0) {
                const float4 t0 = target0[a];
                const float3 pm0 = frame_pred(f0, basis0, target0, a);
                const float3 pb0 = frame_base(basis0, target0, a);
                e[3] += sq3(make_float3(pm.x + t0.x - pm0.x, pm.y + t0.y - pm0.y, pm.z + t0.z - pm0.z), tg);
                e[4] += sq3(make_float3(pb.x + t0.x - pb0.x, pb.y + t0.y - pb0.y, pb.z + t0.z - pb0.z), tg);
                e[5] += sq3(make_float3(t0.x, t0.y, t0.z), tg);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const float v = wave_sum(e[k]);
        if (lane == 0) atomicAdd(&out[k], (double)v);
    }
}

__global__ void k_rs_permute_out(int N, const int* __restrict__ sidx, const float* __restrict__ feat,
                                 const float4* __restrict__ basis, float* __restrict__ feat_out,
                                 float* __restrict__ basis_out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const int i = sidx[p];
    if (feat_out)
        for (int k = 0; k < kFeat; ++k) feat_out[(size_t)kFeat * i + k] = feat[(size_t)kFeat * p + k];
    if (basis_out)
        for (int m = 0; m < 3; ++m) {
            const float4 b = basis[3 * (size_t)p + m];
            basis_out[9 * (size_t)i + 3 * m] = b.x;
            basis_out[9 * (size_t)i + 3 * m + 1] = b.y;
            basis_out[9 * (size_t)i + 3 * m + 2] = b.z;
        }
}

int tiles_of(int N) { return (N + 15) / 16; }

int cast_weights(sponge_rsur* h) {
    hipLaunchKernelGGL(k_rs_cast, dim3(grid_for(kParams, 256)), dim3(256), 0, h->stream, h->d_params,
                       h->d_w1, h->d_w2, h->d_w2t, h->d_w3, h->d_w3t);
    SA_LAUNCH_CHECK();
    return 0;
}

void rsur_free(sponge_rsur* h) {
    dfree(h->d_pos);
    dfree(h->d_feat);
    dfree(h->d_basis);
    dfree(h->d_uex);
    dfree(h->d_params);
    dfree(h->d_norm);
    dfree(h->d_w1);
    dfree(h->d_w2);
    dfree(h->d_w2t);
    dfree(h->d_w3);
    dfree(h->d_w3t);
    dfree(h->d_grad);
    dfree(h->d_m);
    dfree(h->d_v);
    dfree(h->d_acc);
    dfree(h->d_ffeat);
    dfree(h->d_fbasis);
    dfree(h->d_ftarget);
}

// initial parameters: normal weights scaled by 1/sqrt(fan-in), zero biases,
// output layer zero with biases (1, 1, 0): the physics baseline
std::vector<float> initial_params(unsigned seed) {
    std::vector<float> p(kParams, 0.f);
    std::mt19937 gen(seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    for (int i = 0; i < kHidden * kFeat; ++i) p[oW1 + i] = nd(gen) / std::sqrt((float)kFeat);
    for (int i = 0; i < kHidden * kHidden; ++i) p[oW2 + i] = nd(gen) / std::sqrt((float)kHidden);
    p[oB3] = 1.f;
    p[oB3 + 1] = 1.f;
    return p;
}

}  // namespace

namespace sa {

// features of the current coordinates over the list (sorted positions given)
int rsur_features_sorted(sponge_rsur* h, sponge_nl* nl, const float4* pos, const float* crd,
                         const float* charge, const float* cell, const float* rcell,
                         const int* ex_start, const int* ex_list, const int* ex_num) {
    hipStream_t s = h->stream;
    const LT6 c = lt_from(cell), r = lt_from(rcell);
    const bool ortho = lt_is_orthorhombic(c);
    int rc = fill_async(reinterpret_cast<float*>(h->d_uex), 0.f, 4LL * h->N, s);
    if (rc) return rc;
    if (ex_start && ex_num && ex_list) {
        if (ortho)
            hipLaunchKernelGGL(k_rs_exfield<true>, dim3(grid_for(h->N, 256)), dim3(256), 0, s, h->N,
                               crd, charge, ex_start, ex_num, ex_list, c, r, h->beta, h->d_uex);
        else
            hipLaunchKernelGGL(k_rs_exfield<false>, dim3(grid_for(h->N, 256)), dim3(256), 0, s, h->N,
                               crd, charge, ex_start, ex_num, ex_list, c, r, h->beta, h->d_uex);
        SA_LAUNCH_CHECK();
    }
    const size_t sh = sizeof(float4) * h->stage;
    const float rf2 = h->rf * h->rf;
    timer_begin("rsur_features", s);
#define SA_RS_FEAT(O)                                                                               \
    hipLaunchKernelGGL(k_rs_features<O>, dim3(nl->G), dim3(512), sh, s, nl->d_cell_start,            \
                       nl->d_nranges, nl->max_ranges, nl->d_rfirst, nl->d_roff, nl->d_list_count,    \
                       nl->d_list, nl->cap, pos, nl->d_sidx, c, r, rf2, 1.0f / h->rf, h->beta,      \
                       h->stage, h->d_uex, h->d_feat, h->d_basis)
    if (ortho) SA_RS_FEAT(true);
    else SA_RS_FEAT(false);
#undef SA_RS_FEAT
    SA_LAUNCH_CHECK();
    timer_end("rsur_features", s);
    h->sidx = nl->d_sidx;
    h->pos_last = pos;
    return 0;
}

int rsur_predict_sorted(sponge_rsur* h, float* frc, int assign, const float* anchor) {
    if (!h->sidx) return SPONGE_AMD_ERROR_MISSING_COMMAND;
    const float4* pos = h->pos_last;
    const int nt = tiles_of(h->N);
    const int blocks = std::min(grid_for(nt, 4), 2048);
    timer_begin("rsur_mlp", h->stream);
    hipLaunchKernelGGL(k_rs_predict, dim3(blocks), dim3(256), 0, h->stream, h->N, nt, h->d_feat,
                       h->d_basis, pos, h->sidx, h->d_norm, h->d_w1, h->d_w2, h->d_w3, h->d_params,
                       frc, assign, anchor);
    SA_LAUNCH_CHECK();
    timer_end("rsur_mlp", h->stream);
    return 0;
}

void rsur_set_stream(sponge_rsur* h, hipStream_t s) { h->stream = s; }
int rsur_atoms(const sponge_rsur* h) { return h->N; }
float rsur_cutoff(const sponge_rsur* h) { return h->rf; }

}  // namespace sa

extern "C" {

void sponge_rsur_default_options(sponge_rsur_options* opt) {
    if (!opt) return;
    opt->feature_cutoff = 10.0f;
    opt->max_frames = 0;
    opt->seed = 1;
}

int sponge_rsur_model_size(void) { return kModel; }

int sponge_rsur_initial(sponge_rsur** out, int atom_numbers, float beta,
                        const sponge_rsur_options* opt, void* stream) {
    if (!out || atom_numbers <= 0 || beta <= 0.f) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    *out = nullptr;
    sponge_rsur* h = new sponge_rsur();
    h->N = atom_numbers;
    h->beta = beta;
    if (opt) h->opt = *opt;
    else sponge_rsur_default_options(&h->opt);
    if (!(h->opt.feature_cutoff > 0.f) || h->opt.max_frames < 0) {
        delete h;
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    }
    h->rf = h->opt.feature_cutoff;
    h->stream = (hipStream_t)stream;
    h->stage = stage_entries(sizeof(float4), kFeatStage);
    const size_t N = (size_t)atom_numbers, F = (size_t)h->opt.max_frames;
    int e = 0;
    e |= dmalloc(&h->d_pos, N);
    e |= dmalloc(&h->d_feat, N * kFeat);
    e |= dmalloc(&h->d_basis, N * 3);
    e |= dmalloc(&h->d_uex, N);
    e |= dmalloc(&h->d_params, kParams);
    e |= dmalloc(&h->d_norm, 2 * kFeat);
    e |= dmalloc(&h->d_w1, kHidden * kFeat);
    e |= dmalloc(&h->d_w2, kHidden * kHidden);
    e |= dmalloc(&h->d_w2t, kHidden * kHidden);
    e |= dmalloc(&h->d_w3, 16 * kHidden);
    e |= dmalloc(&h->d_w3t, kHidden * 16);
    e |= dmalloc(&h->d_grad, kParams);
    e |= dmalloc(&h->d_m, kParams);
    e |= dmalloc(&h->d_v, kParams);
    e |= dmalloc(&h->d_acc, 16 + 2 * kFeat);
    if (F) {
        e |= dmalloc(&h->d_ffeat, F * N * kFeat);
        e |= dmalloc(&h->d_fbasis, F * N * 3);
        e |= dmalloc(&h->d_ftarget, F * N);
    }
    if (e) {
        rsur_free(h);
        delete h;
        return SPONGE_AMD_ERROR_MALLOC_FAILED;
    }
    hipStream_t s = h->stream;
    std::vector<float> p = initial_params((unsigned)h->opt.seed);
    std::vector<float> norm(2 * kFeat, 0.f);
    for (int k = 0; k < kFeat; ++k) norm[kFeat + k] = 1.f;
    int rc = 0;
    if (hipMemcpyAsync(h->d_params, p.data(), sizeof(float) * kParams, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipMemcpyAsync(h->d_norm, norm.data(), sizeof(float) * 2 * kFeat, hipMemcpyHostToDevice, s) !=
            hipSuccess)
        rc = SPONGE_AMD_ERROR_MALLOC_FAILED;
    // the bf16 copies carry zeros where no parameter maps (padding rows)
    rc |= fill_async(h->d_w3, (unsigned short)0, 16 * kHidden, s);
    rc |= fill_async(h->d_w3t, (unsigned short)0, kHidden * 16, s);
    rc |= fill_async(h->d_grad, 0.f, kParams, s);
    rc |= fill_async(h->d_m, 0.f, kParams, s);
    rc |= fill_async(h->d_v, 0.f, kParams, s);
    rc |= fill_async(h->d_uex, make_float4(0.f, 0.f, 0.f, 0.f), (long long)N, s);
    if (!rc) rc = cast_weights(h);
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = SPONGE_AMD_ERROR_MALLOC_FAILED;
    if (rc) {
        rsur_free(h);
        delete h;
        return rc;
    }
    *out = h;
    return 0;
}

int sponge_rsur_features(sponge_rsur* h, sponge_nl* nl, const float* crd, const float* cell,
                         const float* rcell, const float* charge, const int* excluded_list_start,
                         const int* excluded_list, const int* excluded_numbers) {
    if (!h || !nl || !crd || !cell || !rcell || !charge) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (nl->N != h->N || !nl->built) return SPONGE_AMD_ERROR_MISSING_COMMAND;
    // every partner within the feature cutoff is on the list between rebuilds
    if (h->rf > nl->cutoff) return SPONGE_AMD_ERROR_CONFLICTING_COMMAND;
    hipLaunchKernelGGL(k_rs_pack, dim3(grid_for(h->N, 256)), dim3(256), 0, h->stream, h->N,
                       nl->d_sidx, crd, charge, h->d_pos);
    SA_LAUNCH_CHECK();
    return rsur_features_sorted(h, nl, h->d_pos, crd, charge, cell, rcell, excluded_list_start,
                                excluded_list, excluded_numbers);
}

int sponge_rsur_export_features(sponge_rsur* h, float* feat, float* basis) {
    if (!h) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (!h->sidx) return SPONGE_AMD_ERROR_MISSING_COMMAND;
    hipLaunchKernelGGL(k_rs_permute_out, dim3(grid_for(h->N, 256)), dim3(256), 0, h->stream, h->N,
                       h->sidx, h->d_feat, h->d_basis, feat, basis);
    SA_LAUNCH_CHECK();
    return 0;
}

int sponge_rsur_predict(sponge_rsur* h, float* frc) {
    if (!h || !frc) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    return rsur_predict_sorted(h, frc, 0, nullptr);
}

int sponge_rsur_get_model(sponge_rsur* h, float* model) {
    if (!h || !model) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    SA_CHECK(hipMemcpyAsync(model, h->d_params, sizeof(float) * kParams, hipMemcpyDeviceToHost,
                            h->stream));
    SA_CHECK(hipMemcpyAsync(model + kParams, h->d_norm, sizeof(float) * 2 * kFeat,
                            hipMemcpyDeviceToHost, h->stream));
    SA_CHECK(hipStreamSynchronize(h->stream));
    return 0;
}

int sponge_rsur_set_model(sponge_rsur* h, const float* model) {
    if (!h || !model) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    SA_CHECK(hipMemcpyAsync(h->d_params, model, sizeof(float) * kParams, hipMemcpyHostToDevice,
                            h->stream));
    SA_CHECK(hipMemcpyAsync(h->d_norm, model + kParams, sizeof(float) * 2 * kFeat,
                            hipMemcpyHostToDevice, h->stream));
    int rc = cast_weights(h);
    if (rc) return rc;
    SA_CHECK(hipStreamSynchronize(h->stream));
    return 0;
}

int sponge_rsur_add_frame(sponge_rsur* h, const float* rec_frc) {
    if (!h || !rec_frc) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (!h->sidx) return SPONGE_AMD_ERROR_MISSING_COMMAND;
    if (h->n_frames >= h->opt.max_frames) return SPONGE_AMD_ERROR_OVERFLOW;
    hipStream_t s = h->stream;
    int rc = fill_async(h->d_acc + 1, 0.0, 3, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_rs_sum3, dim3(std::min(grid_for(h->N, 256), 256)), dim3(256), 0, s, h->N,
                       rec_frc, h->d_acc + 1);
    SA_LAUNCH_CHECK();
    const size_t f = (size_t)h->n_frames, N = (size_t)h->N;
    hipLaunchKernelGGL(k_rs_store, dim3(grid_for(h->N, 256)), dim3(256), 0, s, h->N, h->sidx,
                       h->d_feat, h->d_basis, h->pos_last, rec_frc, h->d_acc + 1,
                       h->d_ffeat + f * N * kFeat, h->d_fbasis + f * N * 3, h->d_ftarget + f * N);
    SA_LAUNCH_CHECK();
    h->n_frames++;
    return 0;
}

int sponge_rsur_frame_count(sponge_rsur* h, int* n) {
    if (!h || !n) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    *n = h->n_frames;
    return 0;
}

int sponge_rsur_fit_normalization(sponge_rsur* h, int n_frames) {
    if (!h || n_frames <= 0 || n_frames > h->n_frames) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    hipStream_t s = h->stream;
    double* acc = h->d_acc + 8;
    int rc = fill_async(acc, 0.0, 2 * kFeat, s);
    if (rc) return rc;
    const long long rows = (long long)n_frames * h->N;
    hipLaunchKernelGGL(k_rs_stats, dim3(1024), dim3(256), 0, s, rows, h->d_ffeat, acc);
    SA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rs_norm, dim3(1), dim3(64), 0, s, rows, acc, h->d_norm);
    SA_LAUNCH_CHECK();
    return 0;
}

int sponge_rsur_gradient(sponge_rsur* h, int frame, float* grad, double* loss) {
    if (!h || frame < 0 || frame >= h->n_frames) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    hipStream_t s = h->stream;
    int rc = fill_async(h->d_grad, 0.f, kParams, s);
    rc |= fill_async(h->d_acc, 0.0, 1, s);
    if (rc) return rc;
    const size_t f = (size_t)frame, N = (size_t)h->N;
    const int nt = tiles_of(h->N);
    hipLaunchKernelGGL(k_rs_train, dim3(std::min(nt, 1024)), dim3(64), 0, s, h->N, nt,
                       h->d_ffeat + f * N * kFeat, h->d_fbasis + f * N * 3, h->d_ftarget + f * N,
                       h->d_norm, h->d_w1, h->d_w2, h->d_w2t, h->d_w3, h->d_w3t, h->d_params,
                       h->d_grad, h->d_acc);
    SA_LAUNCH_CHECK();
    if (grad)
        SA_CHECK(hipMemcpyAsync(grad, h->d_grad, sizeof(float) * kParams, hipMemcpyDeviceToHost, s));
    if (loss) SA_CHECK(hipMemcpyAsync(loss, h->d_acc, sizeof(double), hipMemcpyDeviceToHost, s));
    SA_CHECK(hipStreamSynchronize(s));
    return fill_async(h->d_grad, 0.f, kParams, s);
}

int sponge_rsur_train(sponge_rsur* h, int n_train_frames, int epochs, float lr, double* epoch_loss) {
    if (!h || n_train_frames <= 0 || n_train_frames > h->n_frames || epochs < 0 || !(lr > 0.f))
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    hipStream_t s = h->stream;
    const size_t N = (size_t)h->N;
    const int nt = tiles_of(h->N);
    std::mt19937 gen((unsigned)h->opt.seed + 17u);
    std::vector<int> order(n_train_frames);
    for (int i = 0; i < n_train_frames; ++i) order[i] = i;
    int rc = fill_async(h->d_grad, 0.f, kParams, s);
    if (rc) return rc;
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(order.begin(), order.end(), gen);
        rc = fill_async(h->d_acc, 0.0, 1, s);
        if (rc) return rc;
        for (int k = 0; k < n_train_frames; ++k) {
            const size_t f = (size_t)order[k];
            hipLaunchKernelGGL(k_rs_train, dim3(std::min(nt, 1024)), dim3(64), 0, s, h->N, nt,
                               h->d_ffeat + f * N * kFeat, h->d_fbasis + f * N * 3,
                               h->d_ftarget + f * N, h->d_norm, h->d_w1, h->d_w2, h->d_w2t, h->d_w3,
                               h->d_w3t, h->d_params, h->d_grad, h->d_acc);
            SA_LAUNCH_CHECK();
            h->adam_t++;
            const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)h->adam_t)));
            const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)h->adam_t)));
            hipLaunchKernelGGL(k_rs_adam, dim3(grid_for(kParams, 256)), dim3(256), 0, s, h->d_params,
                               h->d_grad, h->d_m, h->d_v, lr, c1, c2, h->d_w1, h->d_w2, h->d_w2t,
                               h->d_w3, h->d_w3t);
            SA_LAUNCH_CHECK();
        }
        if (epoch_loss) {
            double l = 0;
            SA_CHECK(hipMemcpyAsync(&l, h->d_acc, sizeof(double), hipMemcpyDeviceToHost, s));
            SA_CHECK(hipStreamSynchronize(s));
            epoch_loss[ep] = l / n_train_frames;
        }
    }
    SA_CHECK(hipStreamSynchronize(s));
    return 0;
}

int sponge_rsur_evaluate(sponge_rsur* h, int frame_begin, int frame_end, int lag, double* out) {
    if (!h || !out || frame_begin < 0 || frame_end > h->n_frames || frame_begin >= frame_end ||
        lag < 0 || frame_begin - lag < 0)
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    hipStream_t s = h->stream;
    double* acc = h->d_acc + 8 + 2 * kFeat;
    int rc = fill_async(acc, 0.0, 6, s);
    if (rc) return rc;
    const size_t N = (size_t)h->N;
    const int nt = tiles_of(h->N);
    for (int f = frame_begin; f < frame_end; ++f) {
        const size_t f0 = (size_t)(f - lag);
        hipLaunchKernelGGL(k_rs_eval, dim3(std::min(grid_for(nt, 4), 2048)), dim3(256), 0, s, h->N, nt,
                           h->d_ffeat + f * N * kFeat, h->d_fbasis + f * N * 3, h->d_ftarget + f * N,
                           lag ? h->d_ffeat + f0 * N * kFeat : nullptr, h->d_fbasis + f0 * N * 3,
                           h->d_ftarget + f0 * N, h->d_norm, h->d_w1, h->d_w2, h->d_w3, h->d_params,
                           acc);
        SA_LAUNCH_CHECK();
    }
    double a[6];
    SA_CHECK(hipMemcpyAsync(a, acc, sizeof(a), hipMemcpyDeviceToHost, s));
    SA_CHECK(hipStreamSynchronize(s));
    const double ref = a[1] > 0 ? a[1] : 1.0;
    out[0] = std::sqrt(a[0] / ref);
    out[1] = std::sqrt(a[2] / ref);
    out[2] = lag ? std::sqrt(a[3] / ref) : 0.0;
    out[3] = lag ? std::sqrt(a[4] / ref) : 0.0;
    out[4] = lag ? std::sqrt(a[5] / ref) : 0.0;
    return 0;
}

int sponge_rsur_clear(sponge_rsur* h) {
    if (!h) return 0;
    (void)hipStreamSynchronize(h->stream);
    rsur_free(h);
    delete h;
    return 0;
}

}  // extern "C"
