// LJ soft core with PME direct Coulomb for alchemical lambda windows: the
// reference's LJ_SOFT_CORE module (SPONGE/Lennard_Jones_force/
// LJ_soft_core.{h:9-214, cpp:195-440, 463-700, 757-946}).  State A (types and
// tables AA / AB) is weighted 1 - lambda, state B (BA / BB) lambda; a pair
// whose repulsion vanishes in exactly one state uses the soft-core distances
// r_A = (r^6 + alpha lambda^p sigma_A)^(1/6), r_B = (r^6 + alpha (1 -
// lambda)^p sigma_B)^(1/6), sigma = A / (2B) (sigma^6 of the options when A
// or B is zero, at least sigma_min^6).
//
// Same structure as the plain pair kernel: one workgroup (16 wavefronts) per
// neighbor-list cell stages the cell's candidate stream (positions and
// charges, A / B types, charge_BA, global ids) and both square coefficient
// tables in LDS; one wavefront per row over the full 16-bit list.  Forces
// need no atomics (rows hold both directions); energies, virial and
// dU/dlambda are counted in the row of the lower global id.  Force-only
// steps take a lighter form (positions and types staged, two partners per
// lane and pass, the plain kernel's arithmetic for the pairs that are not
// soft-core): 0.88 -> 0.57 ms per step of the 16-window batch.
#include <algorithm>
#include <vector>

#include "internal.h"
#include "batch_launch.h"
#include "pair_term.h"
#include "timing.h"

using namespace sa;

struct sponge_ljsc {
    int N = 0, TA = 0, TB = 0;
    float cutoff = 0, rc2_thr = 0;
    float lambda = 0, alpha = 0.5f, p = 1.0f, sigma6 = 729.0f, sigma6_min = 0.0f;
    float long_range_factor = 0, long_range_factor_TI = 0;
    hipStream_t stream = nullptr;
    int stage = 0;
    int stage_fast = 0;            // entries staged by the force-only kernel
    float2* d_ABa = nullptr;       // [TA][TA] (A, B) of state A
    float2* d_ABb = nullptr;       // [TB][TB] of state B
    int* d_typeA = nullptr;        // [N] local order
    int* d_typeB = nullptr;
    float4* d_pos = nullptr;       // [N] sorted x, y, z, charge
    int* d_tp = nullptr;           // [N] sorted typeA | typeB << 16
    float* d_qba = nullptr;        // [N] sorted charge_BA
    int* d_gid = nullptr;          // [N] sorted global id
    float4* d_facc = nullptr;      // [N] per row
    float2* d_eacc = nullptr;      // [N] (LJ, Coulomb) per row
    float* d_vacc = nullptr;       // [N][6]
    float* d_LJ_energy_atom = nullptr;  // [N] local order
    double* d_du = nullptr;        // [0] LJ, [1] direct
    double* d_sum = nullptr;
};

namespace {

constexpr int kScStage = 3584;
constexpr float kTwoOverSqrtPiF = 1.1283791670955126f;

__device__ __forceinline__ void ljsc_pack_body(int sa_block, int N, const int* __restrict__ sidx,
                            const float4* __restrict__ spos,
                            const float* __restrict__ crd, const float* __restrict__ charge,
                            const float* __restrict__ qba, const int* __restrict__ tA,
                            const int* __restrict__ tB, float4* __restrict__ pos, int* __restrict__ tp,
                            float* __restrict__ pqba, int* __restrict__ gid) {
    const int p = sa_block * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const int i = sidx[p];
    pos[p] = make_float4(crd[3 * i], crd[3 * i + 1], crd[3 * i + 2], charge[i]);
    tp[p] = tA[i] | (tB[i] << 16);
    pqba[p] = qba ? qba[i] : 0.f;
    gid[p] = __float_as_int(spos[p].w);
}

__global__ void k_ljsc_pack(int N, const int* __restrict__ sidx, const float4* __restrict__ spos,
                            const float* __restrict__ crd, const float* __restrict__ charge,
                            const float* __restrict__ qba, const int* __restrict__ tA,
                            const int* __restrict__ tB, float4* __restrict__ pos, int* __restrict__ tp,
                            float* __restrict__ pqba, int* __restrict__ gid) {
    ljsc_pack_body(blockIdx.x, N, sidx, spos, crd, charge, qba, tA, tB, pos, tp, pqba, gid);
}

struct ScParams {
    float lambda, alpha, p, sigma6, sigma6_min, beta;
};

// Pair functions of LJ_soft_core.h: forces are F / r (the vector force is
// F d), energies the pair energies.
__device__ __forceinline__ float lj_f(float r, float A, float B) {
    const float ri2 = 1.0f / (r * r), r6 = ri2 * ri2 * ri2;
    return (B - A * r6) * r6 * ri2;
}
__device__ __forceinline__ float lj_e(float r, float A, float B) {
    const float ri2 = 1.0f / (r * r), r6 = ri2 * ri2 * ri2;
    return (0.083333333f * A * r6 - 0.166666667f * B) * r6;
}
__device__ __forceinline__ float cf_f(float qq, float r, float beta) {
    const float br = beta * r;
    return qq / (r * r * r) * (br * kTwoOverSqrtPiF * expf(-br * br) + erfcf(br));
}
__device__ __forceinline__ float cf_e(float qq, float r, float beta) { return qq * erfcf(beta * r) / r; }
__device__ __forceinline__ float sc_sigma(float A, float B, const ScParams& s) {
    return (A == 0.f || B == 0.f) ? s.sigma6 : fmaxf(0.5f * A / B, s.sigma6_min);
}
__device__ __forceinline__ float sc_dist(float sigma, float r2, float lam, const ScParams& s) {
    return powf(r2 * r2 * r2 + s.alpha * powf(lam, s.p) * sigma, 0.16666667f);
}
__device__ __forceinline__ float sc_lj_f(float r2, float rs, float A, float B) {
    const float rsi6 = 1.0f / (rs * rs * rs * rs * rs * rs);
    return r2 * r2 * (B - A * rsi6) * rsi6 * rsi6;
}
__device__ __forceinline__ float sc_cf_f(float qq, float r2, float rs, float beta) {
    const float b = rs * beta;
    const float rsi6 = 1.0f / (rs * rs * rs * rs * rs * rs);
    return qq * r2 * r2 * (expf(-b * b) * kTwoOverSqrtPiF * beta + erfcf(b) / rs) * rsi6;
}
__device__ __forceinline__ float sc_du(float F, float sigma, float rs, float lam, const ScParams& s) {
    return 0.16666667f * s.p * s.alpha * (1.0f - lam) * powf(lam + 1.17549435e-38f, s.p - 1.0f) * F *
           sigma / (rs * rs * rs * rs);
}

struct ScTerm {
    float f, elj, ecf, dlj, dcf;
};

template <bool E, bool DU>
__device__ __forceinline__ ScTerm sc_pair(float r2, float qi, float qj, float qbai, float qbaj,
                                          float2 a, float2 b, const ScParams& s) {
    ScTerm t{0.f, 0.f, 0.f, 0.f, 0.f};
    const float r = sqrtf(r2);
    const float lam = s.lambda, lam_ = 1.0f - s.lambda;
    const float qq = qi * qj;
    const float Aa = a.x, Ab = a.y, Ba = b.x, Bb = b.y;
    if (Ba * Aa != 0.f || Ba + Aa == 0.f) {
        t.f = lam_ * lj_f(r, Aa, Ab) + lam * lj_f(r, Ba, Bb) - cf_f(qq, r, s.beta);
        if (E) {
            t.ecf = cf_e(qq, r, s.beta);
            t.elj = lam_ * lj_e(r, Aa, Ab) + lam * lj_e(r, Ba, Bb);
        }
        if (DU) {
            t.dlj = lj_e(r, Ba, Bb) - lj_e(r, Aa, Ab);
            t.dcf = erfcf(s.beta * r) * (qbaj * qi + qj * qbai) / r;
        }
    } else {
        const float sA = sc_sigma(Aa, Ab, s), sB = sc_sigma(Ba, Bb, s);
        const float rA = sc_dist(sA, r2, lam, s), rB = sc_dist(sB, r2, lam_, s);
        t.f = lam_ * sc_lj_f(r2, rA, Aa, Ab) + lam * sc_lj_f(r2, rB, Ba, Bb) -
              (lam_ * sc_cf_f(qq, r2, rA, s.beta) + lam * sc_cf_f(qq, r2, rB, s.beta));
        if (E) {
            t.ecf = lam_ * cf_e(qq, rA, s.beta) + lam * cf_e(qq, rB, s.beta);
            t.elj = lam_ * lj_e(rA, Aa, Ab) + lam * lj_e(rB, Ba, Bb);
        }
        if (DU) {
            t.dlj = lj_e(rB, Ba, Bb) - lj_e(rA, Aa, Ab) + sc_du(lj_f(rA, Aa, Ab), sA, rA, lam, s) -
                    sc_du(lj_f(rB, Ba, Bb), sB, rB, lam_, s);
            const float qx = qbaj * qi + qj * qbai;
            t.dcf = cf_e(qq, rB, s.beta) - cf_e(qq, rA, s.beta) +
                    sc_du(cf_f(qq, rB, s.beta), sB, rB, lam_, s) -
                    sc_du(cf_f(qq, rA, s.beta), sA, rA, lam, s) +
                    lam * erfcf(s.beta * rB) * qx / rB + lam_ * erfcf(s.beta * rA) * qx / rA;
        }
    }
    return t;
}

// Cell geometry for the force-only kernel's cell-centre images: reach =
// cutoff + the largest distance a row atom may have moved since the list was
// built (< 0: no bound known, periodic displacements kept).
struct ShiftGeom {
    float gl;
    int Nx, Ny, Nz;
    float reach;
};

// One pair of the force-only soft-core rows at cell-centre images (plain
// displacement): the plain kernel's FAST arithmetic with the state-weighted
// coefficients (1 - lambda) A_A + lambda A_B for pairs whose LJ is present in
// both states or in neither, the reference's soft-core pair function for the
// others (a branch taken only by wavefronts that hold one).  has = false: an
// absent entry of a partial pass (r^2 = the cutoff threshold, out of range).
__device__ __forceinline__ void sc_pair_fast(const float4& pj, int tpj, const float3& pi, float qi,
                                             const float2* rowA, const float2* rowB, float lam,
                                             float lam_, float rc2_thr, const ScParams& sp,
                                             bool has, float& fx, float& fy, float& fz) {
    const float3 d = make_float3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
    float r2 = fmaf(d.z, d.z, fmaf(d.y, d.y, d.x * d.x));
    r2 = has ? r2 : rc2_thr;
    const bool in = r2 < rc2_thr;
    const float2 a = rowA[tpj & 0xffff], b = rowB[tpj >> 16];
    const bool plain = a.x * b.x != 0.f || a.x + b.x == 0.f;
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float rabs = r2 * rinv;
    const float r2inv = rinv * rinv;
    const float pw6 = r2inv * r2inv * r2inv;
    const float pw8 = pw6 * r2inv;
    const float pw3 = rinv * r2inv;
    const float cA = fmaf(lam, b.x, lam_ * a.x), cB = fmaf(lam, b.y, lam_ * a.y);
    float erfc_br, bracket;
    coul_fast(r2, rabs, sp.beta, erfc_br, bracket);
    float f = (cB - cA * pw6) * pw8 - qi * pj.w * pw3 * bracket;
    const bool soft = in && !plain;
    if (__ballot(soft)) {
        // (charge_BA enters only dU/dlambda)
        if (soft) f = sc_pair<false, false>(r2, qi, pj.w, 0.f, 0.f, a, b, sp).f;
    }
    f = in ? f : 0.f;
    fx = fmaf(f, d.x, fx);
    fy = fmaf(f, d.y, fy);
    fz = fmaf(f, d.z, fz);
}

// Force-only row of a staged cell at cell-centre images: two entries per
// lane per pass from one 32-bit word (prefetched a pass ahead), the passes in
// which all 128 entries exist without presence tests, the partial last pass
// masked.
__device__ __forceinline__ void sc_row_fast(int p, int lane, int n, const unsigned short* __restrict__ row,
                                            const float4* cpos, const int* ctp,
                                            const float4* __restrict__ pos, const float2* rowA,
                                            const float2* rowB, const LT6& c, const LT6& r,
                                            float rc2_thr, const ScParams& sp,
                                            float4* __restrict__ facc, float3 C) {
    const float lam = sp.lambda, lam_ = 1.0f - sp.lambda;
    const float4 pi4 = pos[p];
    const float3 pi = near_image(make_float3(pi4.x, pi4.y, pi4.z), C, c, r);
    float fx = 0.f, fy = 0.f, fz = 0.f;
    const int npairs = (n + 1) >> 1;
    const int full = n >> 7;
    const unsigned* row2 = reinterpret_cast<const unsigned*>(row);
    const unsigned last = npairs > 0 ? npairs - 1 : 0;
    unsigned cur = row2[(unsigned)lane < last ? (unsigned)lane : last];
    int w = lane;
    for (int k = 0; k < full; ++k, w += kWave) {
        const unsigned wn = (unsigned)(w + kWave);
        unsigned nxt = row2[wn < last ? wn : last];
        const int ta = (int)(cur & 0xffffu), tb = (int)(cur >> 16);
        sc_pair_fast(cpos[ta], ctp[ta], pi, pi4.w, rowA, rowB, lam, lam_, rc2_thr, sp, true, fx,
                     fy, fz);
        sc_pair_fast(cpos[tb], ctp[tb], pi, pi4.w, rowA, rowB, lam, lam_, rc2_thr, sp, true, fx,
                     fy, fz);
        // the loaded word is opaque from here: the compiler cannot fold the
        // prefetch back into a load at the top of the next pass
        asm volatile("" : "+v"(nxt));
        cur = nxt;
    }
    if (full * 2 * kWave < n) {
        // cur is the row's word w, or its last word for lanes past the end
        const bool has_a = w < npairs;
        const bool has_b = 2 * w + 1 < n;
        const int ta = (int)(cur & 0xffffu);
        const int tb = has_b ? (int)(cur >> 16) : ta;
        sc_pair_fast(cpos[ta], ctp[ta], pi, pi4.w, rowA, rowB, lam, lam_, rc2_thr, sp, has_a, fx,
                     fy, fz);
        sc_pair_fast(cpos[tb], ctp[tb], pi, pi4.w, rowA, rowB, lam, lam_, rc2_thr, sp, has_b, fx,
                     fy, fz);
    }
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
    if (lane == 0) facc[p] = make_float4(fx, fy, fz, 0.f);
}

template <bool ORTHO, bool F, bool E, bool V, bool DU, bool FAST = false>
__device__ __forceinline__ void ljsc_cells_body(int sa_block,
    const int* __restrict__ start, const int* __restrict__ nranges, int max_ranges,
    const int* __restrict__ g_rfirst, const int* __restrict__ g_roff,
    const int* __restrict__ list_count, const unsigned short* __restrict__ list, int cap,
    const float4* __restrict__ pos, const int* __restrict__ ptp, const float* __restrict__ pqba,
    const int* __restrict__ pgid, const float2* __restrict__ ABa, int TA,
    const float2* __restrict__ ABb, int TB, LT6 c, LT6 r, float rc2_thr, ScParams sp, int stage,
    float4* __restrict__ facc, float2* __restrict__ eacc, float* __restrict__ vacc,
    double* __restrict__ du, ShiftGeom sg) {
    extern __shared__ float4 cpos[];
    // force-only steps stage positions and types only (20 B per entry, so
    // two workgroups share a CU)
    int* ctp = reinterpret_cast<int*>(cpos + stage);
    float* cqba = reinterpret_cast<float*>(ctp + stage);
    int* cgid = reinterpret_cast<int*>(cqba + stage);
    float2* ta = FAST ? reinterpret_cast<float2*>(ctp + stage) : reinterpret_cast<float2*>(cgid + stage);
    float2* tb = ta + TA * TA;
    __shared__ int r_first[kMaxRanges];
    __shared__ int r_off[kMaxRanges + 1];
    const int cell = sa_block;
    const int a0 = start[cell], a1 = start[cell + 1];
    if (a0 == a1) return;
    const int nr = nranges[cell];
    for (int k = threadIdx.x; k < nr; k += blockDim.x) r_first[k] = g_rfirst[(size_t)cell * max_ranges + k];
    for (int k = threadIdx.x; k <= nr; k += blockDim.x) r_off[k] = g_roff[(size_t)cell * (max_ranges + 1) + k];
    for (int k = threadIdx.x; k < TA * TA; k += blockDim.x) ta[k] = ABa[k];
    for (int k = threadIdx.x; k < TB * TB; k += blockDim.x) tb[k] = ABb[k];
    __syncthreads();
    const int total = r_off[nr];
    const bool staged = total <= stage;
    // force-only steps in an orthorhombic box wider than twice (half the cell
    // + reach) on every axis: partners staged at their images nearest the
    // cell's centre C (the plain pair kernel's construction)
    float3 C = make_float3(0.f, 0.f, 0.f);
    bool shift = false;
    if (FAST && ORTHO && sg.reach >= 0.f) {
        const int ix = cell / (sg.Ny * sg.Nz), iy = (cell / sg.Nz) % sg.Ny, iz = cell % sg.Nz;
        const float gl = sg.gl;
        const float lx = ix * gl, hx = ix == sg.Nx - 1 ? c.a11 : (ix + 1) * gl;
        const float ly = iy * gl, hy = iy == sg.Ny - 1 ? c.a22 : (iy + 1) * gl;
        const float lz = iz * gl, hz = iz == sg.Nz - 1 ? c.a33 : (iz + 1) * gl;
        C = make_float3(0.5f * (lx + hx), 0.5f * (ly + hy), 0.5f * (lz + hz));
        shift = staged && 0.5f * c.a11 > sg.reach + 0.5f * (hx - lx) + 0.01f &&
                0.5f * c.a22 > sg.reach + 0.5f * (hy - ly) + 0.01f &&
                0.5f * c.a33 > sg.reach + 0.5f * (hz - lz) + 0.01f;
    }
    if (staged)
        stage_stream(nr, r_first, r_off, total, [&](int t, int q) {
            float4 v = pos[q];
            if (shift) {
                const float3 im = near_image(make_float3(v.x, v.y, v.z), C, c, r);
                v = make_float4(im.x, im.y, im.z, v.w);
            }
            cpos[t] = v;
            ctp[t] = ptp[q];
            if (!FAST) {
                cqba[t] = pqba[q];
                cgid[t] = pgid[q];
            }
        });
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave, nwave = blockDim.x / kWave;
    if (FAST) {
        // force-only steps: two entries per lane per pass, branch-free; the
        // pairs whose LJ is present in both states or in neither (all but
        // the perturbed atoms' pairs) take the plain kernel's arithmetic with
        // the state-weighted coefficients (1 - lambda) A_A + lambda A_B, the
        // soft-core pairs the reference's pair function in a branch taken
        // only by the wavefronts that hold one
        const float lam = sp.lambda, lam_ = 1.0f - sp.lambda;
        if (shift) {
            for (int p = a0 + wave; p < a1; p += nwave) {
                const int tpi = ptp[p];
                sc_row_fast(p, lane, list_count[p], list + (size_t)p * cap, cpos, ctp, pos,
                            ta + (tpi & 0xffff) * TA, tb + (tpi >> 16) * TB, c, r, rc2_thr, sp,
                            facc, C);
            }
            return;
        }
        for (int p = a0 + wave; p < a1; p += nwave) {
            const float4 pi4 = pos[p];
            const float3 pi = make_float3(pi4.x, pi4.y, pi4.z);
            const int tpi = ptp[p];
            const float2* rowA = ta + (tpi & 0xffff) * TA;
            const float2* rowB = tb + (tpi >> 16) * TB;
            const int n = list_count[p];
            const unsigned short* row = list + (size_t)p * cap;
            float fx = 0.f, fy = 0.f, fz = 0.f;
            for (int base = 0; base < n; base += 2 * kWave) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int e = base + h * kWave + lane;
                    const bool has = e < n;
                    const int t = row[has ? e : 0];
                    const int q = staged ? t : stream_to_sorted(t, nr, r_first, r_off);
                    const float4 pj = staged ? cpos[q] : pos[q];
                    const int tpj = staged ? ctp[q] : ptp[q];
                    const float3 d = disp_pbc<ORTHO>(make_float3(pj.x, pj.y, pj.z), pi, c, r);
                    const float r2 = norm2_exact(d);
                    const bool in = has && r2 < rc2_thr;
                    const float2 a = rowA[tpj & 0xffff], b = rowB[tpj >> 16];
                    const bool plain = a.x * b.x != 0.f || a.x + b.x == 0.f;
                    // plain arithmetic (values of an out-of-range or soft-core
                    // pair are discarded by the selects)
                    const float rinv = __builtin_amdgcn_rsqf(r2);
                    const float rabs = r2 * rinv;
                    const float r2inv = rinv * rinv;
                    const float pw6 = r2inv * r2inv * r2inv;
                    const float pw8 = pw6 * r2inv;
                    const float pw3 = rinv * r2inv;
                    const float cA = fmaf(lam, b.x, lam_ * a.x), cB = fmaf(lam, b.y, lam_ * a.y);
                    const float br = sp.beta * rabs;
                    const float ex = __expf(-br * br);
                    const float bracket = ex * fmaf(br, kTwoOverSqrtPiF, erfcx_poly(br));
                    float f = (cB - cA * pw6) * pw8 - pi4.w * pj.w * pw3 * bracket;
                    const bool soft = in && !plain;
                    if (__ballot(soft)) {
                        // (charge_BA enters only dU/dlambda)
                        if (soft) f = sc_pair<false, false>(r2, pi4.w, pj.w, 0.f, 0.f, a, b, sp).f;
                    }
                    f = in ? f : 0.f;
                    fx = fmaf(f, d.x, fx);
                    fy = fmaf(f, d.y, fy);
                    fz = fmaf(f, d.z, fz);
                }
            }
            fx = wave_sum(fx);
            fy = wave_sum(fy);
            fz = wave_sum(fz);
            if (lane == 0) facc[p] = make_float4(fx, fy, fz, 0.f);
        }
        return;
    }
    for (int p = a0 + wave; p < a1; p += nwave) {
        const float4 pi4 = pos[p];
        const float3 pi = make_float3(pi4.x, pi4.y, pi4.z);
        const int tpi = ptp[p], gi = pgid[p];
        const float qbai = pqba[p];
        const float2* rowA = ta + (tpi & 0xffff) * TA;
        const float2* rowB = tb + (tpi >> 16) * TB;
        const int n = list_count[p];
        const unsigned short* row = list + (size_t)p * cap;
        float fx = 0.f, fy = 0.f, fz = 0.f, elj = 0.f, ecf = 0.f, dlj = 0.f, dcf = 0.f;
        float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f, v4 = 0.f, v5 = 0.f;
        for (int e = lane; e < n; e += kWave) {
            const int t = row[e];
            const int q = staged ? t : stream_to_sorted(t, nr, r_first, r_off);
            const float4 pj = staged ? cpos[q] : pos[q];
            const int tpj = staged ? ctp[q] : ptp[q];
            const float3 d = disp_pbc<ORTHO>(make_float3(pj.x, pj.y, pj.z), pi, c, r);
            const float r2 = norm2_exact(d);
            if (!(r2 < rc2_thr)) continue;
            const float qbaj = staged ? cqba[q] : pqba[q];
            const int gj = staged ? cgid[q] : pgid[q];
            const ScTerm tm = sc_pair<E, DU>(r2, pi4.w, pj.w, qbai, qbaj, rowA[tpj & 0xffff],
                                             rowB[tpj >> 16], sp);
            if (F) {
                fx += tm.f * d.x;
                fy += tm.f * d.y;
                fz += tm.f * d.z;
            }
            if (gj > gi) {
                // the reference's half-list row: energies, virial, dU/dlambda
                elj += tm.elj;
                ecf += tm.ecf;
                dlj += tm.dlj;
                dcf += tm.dcf;
                if (V) {
                    const float lx = tm.f * d.x, ly = tm.f * d.y, lz = tm.f * d.z;
                    v0 -= lx * d.x;
                    v1 -= lx * d.y + ly * d.x;
                    v2 -= ly * d.y;
                    v3 -= lx * d.z + lz * d.x;
                    v4 -= ly * d.z + lz * d.y;
                    v5 -= lz * d.z;
                }
            }
        }
        if (F) {
            fx = wave_sum(fx);
            fy = wave_sum(fy);
            fz = wave_sum(fz);
        }
        double de = 0, dc = 0, ddl = 0, ddc = 0;
        if (E) {
            de = wave_sum((double)elj);
            dc = wave_sum((double)ecf);
        }
        if (DU) {
            ddl = wave_sum((double)dlj);
            ddc = wave_sum((double)dcf);
        }
        double w0 = 0, w1 = 0, w2 = 0, w3 = 0, w4 = 0, w5 = 0;
        if (V) {
            w0 = wave_sum((double)v0);
            w1 = wave_sum((double)v1);
            w2 = wave_sum((double)v2);
            w3 = wave_sum((double)v3);
            w4 = wave_sum((double)v4);
            w5 = wave_sum((double)v5);
        }
        if (lane == 0) {
            if (F) facc[p] = make_float4(fx, fy, fz, 0.f);
            if (E) eacc[p] = make_float2((float)de, (float)dc);
            if (V) {
                float* v = vacc + (size_t)6 * p;
                v[0] = (float)w0; v[1] = (float)w1; v[2] = (float)w2;
                v[3] = (float)w3; v[4] = (float)w4; v[5] = (float)w5;
            }
            if (DU) {
                atomicAdd(&du[0], ddl);
                atomicAdd(&du[1], ddc);
            }
        }
    }
}

}  // namespace

// The force-only kernels run at 8 wavefronts per SIMD (at most 64 VGPRs), as
// the plain pair kernel does.
template <>
struct sa::BodyWaves<&ljsc_cells_body<true, true, false, false, false, true>> {
    static constexpr int value = 8;
};
template <>
struct sa::BodyWaves<&ljsc_cells_body<false, true, false, false, false, true>> {
    static constexpr int value = 8;
};

namespace {

template <bool ORTHO, bool F, bool E, bool V, bool DU, bool FAST = false>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(FAST ? 8 : 1)))
k_ljsc_cells(
    const int* __restrict__ start, const int* __restrict__ nranges, int max_ranges,
    const int* __restrict__ g_rfirst, const int* __restrict__ g_roff,
    const int* __restrict__ list_count, const unsigned short* __restrict__ list, int cap,
    const float4* __restrict__ pos, const int* __restrict__ ptp, const float* __restrict__ pqba,
    const int* __restrict__ pgid, const float2* __restrict__ ABa, int TA,
    const float2* __restrict__ ABb, int TB, LT6 c, LT6 r, float rc2_thr, ScParams sp, int stage,
    float4* __restrict__ facc, float2* __restrict__ eacc, float* __restrict__ vacc,
    double* __restrict__ du, ShiftGeom sg) {
    ljsc_cells_body<ORTHO, F, E, V, DU, FAST>(blockIdx.x, start, nranges, max_ranges, g_rfirst, g_roff,
                                         list_count, list, cap, pos, ptp, pqba, pgid, ABa, TA, ABb,
                                         TB, c, r, rc2_thr, sp, stage, facc, eacc, vacc, du, sg);
}

// Scatter of the rows: frc +=; atom_energy += LJ + Coulomb; the module's LJ
// energy and the direct-Coulomb energy overwritten (the reference clears
// them first, LJ_soft_core.cpp:776-785); atom_virial +=.
// assign: frc = instead of += (the step driver's first force term);
// reset_flag (nullable) is cleared by the first thread (the step driver's
// list-rebuild flag, once the rows of this step have been read).
template <bool E, bool V>
__device__ __forceinline__ void ljsc_finalize_body(int sa_block, int N, const int* __restrict__ sidx,
                                const float4* __restrict__ facc,
                                const float2* __restrict__ eacc, const float* __restrict__ vacc,
                                float* __restrict__ frc, float* __restrict__ atom_energy,
                                float* __restrict__ lj_energy, float* __restrict__ direct_energy,
                                float* __restrict__ atom_virial, int assign,
                                int* __restrict__ reset_flag) {
    const int p = sa_block * blockDim.x + threadIdx.x;
    if (p == 0 && reset_flag) *reset_flag = 0;
    if (p >= N) return;
    const int i = sidx[p];
    const float4 f = facc[p];
    if (assign) {
        frc[3 * i] = f.x;
        frc[3 * i + 1] = f.y;
        frc[3 * i + 2] = f.z;
    } else {
        frc[3 * i] += f.x;
        frc[3 * i + 1] += f.y;
        frc[3 * i + 2] += f.z;
    }
    if (E) {
        const float2 e = eacc[p];
        if (atom_energy) atom_energy[i] += e.x + e.y;
        lj_energy[i] = e.x;
        if (direct_energy) direct_energy[i] = e.y;
    }
    if (V && atom_virial)
        for (int k = 0; k < 6; ++k) atom_virial[6 * i + k] += vacc[6 * (size_t)p + k];
}

template <bool E, bool V>
__global__ void k_ljsc_finalize(int N, const int* __restrict__ sidx, const float4* __restrict__ facc,
                                const float2* __restrict__ eacc, const float* __restrict__ vacc,
                                float* __restrict__ frc, float* __restrict__ atom_energy,
                                float* __restrict__ lj_energy, float* __restrict__ direct_energy,
                                float* __restrict__ atom_virial, int assign,
                                int* __restrict__ reset_flag) {
    ljsc_finalize_body<E, V>(blockIdx.x, N, sidx, facc, eacc, vacc, frc, atom_energy, lj_energy,
                             direct_energy, atom_virial, assign, reset_flag);
}

__global__ void k_ljsc_add_scalar(float* p, float v) { p[0] += v; }

__global__ void k_ljsc_add_diag(float* p, float v) {
    p[0] += v;
    p[2] += v;
    p[5] += v;
}

__global__ void k_ljsc_sum(int N, const float* __restrict__ e, double* out) {
    double s = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) s += e[i];
    s = wave_sum(s);
    if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(out, s);
}

std::vector<float2> square_table(int T, const float* A, const float* B) {
    std::vector<float2> t((size_t)T * T);
    for (int a = 0; a < T; ++a)
        for (int b = 0; b < T; ++b) {
            const int k = lj_pair_index(a, b);
            t[(size_t)a * T + b] = make_float2(A[k], B[k]);
        }
    return t;
}

void ljsc_free(sponge_ljsc* h) {
    dfree(h->d_ABa);
    dfree(h->d_ABb);
    dfree(h->d_typeA);
    dfree(h->d_typeB);
    dfree(h->d_pos);
    dfree(h->d_tp);
    dfree(h->d_qba);
    dfree(h->d_gid);
    dfree(h->d_facc);
    dfree(h->d_eacc);
    dfree(h->d_vacc);
    dfree(h->d_LJ_energy_atom);
    dfree(h->d_du);
    dfree(h->d_sum);
}

template <bool ORTHO>
int launch(sponge_ljsc* h, sponge_nl* nl, LT6 c, LT6 r, float beta, bool F, bool E, bool V,
           bool DU, float reach = -1.f) {
    const ScParams sp{h->lambda, h->alpha, h->p, h->sigma6, h->sigma6_min, beta};
    const ShiftGeom sg{nl->grid_length, nl->Nx, nl->Ny, nl->Nz, reach};
    const bool fast = F && !E && !V && !DU && beta * h->cutoff <= kErfcxMax;
    const int stage = fast ? h->stage_fast : h->stage;
    const size_t sh = (size_t)stage * (sizeof(float4) + (fast ? 1 : 3) * sizeof(int)) +
                      sizeof(float2) * ((size_t)h->TA * h->TA + (size_t)h->TB * h->TB);
#define SA_SC(f, e, v, d, fa)                                                                      \
    SA_LAUNCH(1024, (&ljsc_cells_body<ORTHO, f, e, v, d, fa>), (k_ljsc_cells<ORTHO, f, e, v, d, fa>),  \
              dim3(nl->G), dim3(1024), sh, h->stream,                                              \
                       nl->d_cell_start, nl->d_nranges, nl->max_ranges, nl->d_rfirst, nl->d_roff,   \
                       nl->d_list_count, nl->d_list, nl->cap, h->d_pos, h->d_tp, h->d_qba,          \
                       h->d_gid, h->d_ABa, h->TA, h->d_ABb, h->TB, c, r, h->rc2_thr, sp, stage, \
                       h->d_facc, h->d_eacc, h->d_vacc, h->d_du, sg)
    if (DU) SA_SC(false, false, false, true, false);
    else if (E && V) SA_SC(true, true, true, false, false);
    else if (E) SA_SC(true, true, false, false, false);
    else if (V) SA_SC(true, false, true, false, false);
    else if (fast) SA_SC(true, false, false, false, true);
    else SA_SC(true, false, false, false, false);
#undef SA_SC
    (void)F;
    SA_LAUNCH_CHECK();
    return 0;
}

int prepare(sponge_ljsc* h, sponge_nl* nl, const float* crd, const float* charge,
            const float* charge_BA) {
    if (!nl->built || nl->N != h->N) return SPONGE_AMD_ERROR_MISSING_COMMAND;
    SA_LAUNCH(256, (&ljsc_pack_body), (k_ljsc_pack), dim3(grid_for(h->N, 256)), dim3(256), 0,
              h->stream, h->N, nl->d_sidx, nl->d_spos, crd, charge, charge_BA, h->d_typeA, h->d_typeB,
                       h->d_pos, h->d_tp, h->d_qba, h->d_gid);
    SA_LAUNCH_CHECK();
    return 0;
}

}  // namespace

extern "C" {

void sponge_ljsc_default_options(sponge_ljsc_options* o) {
    if (!o) return;
    o->lambda = 0.0f;
    o->alpha = 0.5f;
    o->p = 1.0f;
    o->sigma = 3.0f;
    o->sigma_min = 0.0f;
}

int sponge_ljsc_initial(sponge_ljsc** out, int atom_numbers, int type_numbers_A,
                        int type_numbers_B, const float* LJ_AA, const float* LJ_AB,
                        const float* LJ_BA, const float* LJ_BB, const int* atom_type_A,
                        const int* atom_type_B, float cutoff, const sponge_ljsc_options* opt,
                        void* stream) {
    if (!out || atom_numbers <= 0 || type_numbers_A <= 0 || type_numbers_B <= 0 || !LJ_AA ||
        !LJ_AB || !LJ_BA || !LJ_BB || !atom_type_A || !atom_type_B || !(cutoff > 0.f))
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    *out = nullptr;
    if (type_numbers_A > 64 || type_numbers_B > 64) return SPONGE_AMD_ERROR_NOT_IMPLEMENTED;
    sponge_ljsc_options o;
    if (opt) o = *opt;
    else sponge_ljsc_default_options(&o);
    if (!(o.lambda >= 0.f && o.lambda <= 1.f) || !(o.p > 0.f)) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    const int N = atom_numbers, TA = type_numbers_A, TB = type_numbers_B;
    for (int i = 0; i < N; ++i)
        if (atom_type_A[i] < 0 || atom_type_A[i] >= TA || atom_type_B[i] < 0 || atom_type_B[i] >= TB)
            return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    sponge_ljsc* h = new sponge_ljsc();
    h->N = N;
    h->TA = TA;
    h->TB = TB;
    h->cutoff = cutoff;
    h->rc2_thr = sqrt_threshold(cutoff);
    h->lambda = o.lambda;
    h->alpha = o.alpha;
    h->p = o.p;
    h->sigma6 = powf(o.sigma, 6);
    h->sigma6_min = powf(o.sigma_min, 6);
    h->stream = (hipStream_t)stream;
    const size_t tables = sizeof(float2) * ((size_t)TA * TA + (size_t)TB * TB);
    h->stage = stage_entries(sizeof(float4) + 3 * sizeof(int), kScStage, tables);
    {
        // the force-only kernel: two workgroups per CU (160 KiB of LDS)
        const long fit = ((long)80 * 1024 - 4096 - (long)tables) / (long)(sizeof(float4) + sizeof(int));
        h->stage_fast = (int)(std::max(64L, std::min((long)kScStage, fit)) & ~63L);
    }
    // long-range factors (LJ_soft_core.cpp:83-195, 640-668): sums of the
    // lambda-mixed and B - A C6 coefficients over all ordered pairs, by type
    // histograms, in double, scaled in float
    std::vector<long long> nA(TA, 0), nB(TB, 0);
    for (int i = 0; i < N; ++i) {
        nA[atom_type_A[i]]++;
        nB[atom_type_B[i]]++;
    }
    double sA = 0, sB = 0;
    for (int a = 0; a < TA; ++a)
        for (int b = 0; b < TA; ++b) sA += (double)nA[a] * nA[b] * LJ_AB[lj_pair_index(a, b)];
    for (int a = 0; a < TB; ++a)
        for (int b = 0; b < TB; ++b) sB += (double)nB[a] * nB[b] * LJ_BB[lj_pair_index(a, b)];
    const float lam_ = 1.0f - o.lambda;
    const float scale = -2.0f / 3.0f * 3.1415926535897932f / cutoff / cutoff / cutoff / 6.0f;
    h->long_range_factor = (float)((double)lam_ * sA + (double)o.lambda * sB) * scale;
    h->long_range_factor_TI = (float)(sB - sA) * scale;
    int e = 0;
    e |= dmalloc(&h->d_ABa, (size_t)TA * TA);
    e |= dmalloc(&h->d_ABb, (size_t)TB * TB);
    e |= dmalloc(&h->d_typeA, N);
    e |= dmalloc(&h->d_typeB, N);
    e |= dmalloc(&h->d_pos, N);
    e |= dmalloc(&h->d_tp, N);
    e |= dmalloc(&h->d_qba, N);
    e |= dmalloc(&h->d_gid, N);
    e |= dmalloc(&h->d_facc, N);
    e |= dmalloc(&h->d_eacc, N);
    e |= dmalloc(&h->d_vacc, (size_t)6 * N);
    e |= dmalloc(&h->d_LJ_energy_atom, N);
    e |= dmalloc(&h->d_du, 2);
    e |= dmalloc(&h->d_sum, 1);
    if (e) {
        ljsc_free(h);
        delete h;
        return SPONGE_AMD_ERROR_MALLOC_FAILED;
    }
    const std::vector<float2> ta = square_table(TA, LJ_AA, LJ_AB), tb = square_table(TB, LJ_BA, LJ_BB);
    hipStream_t s = h->stream;
    hipError_t he = hipMemcpyAsync(h->d_ABa, ta.data(), sizeof(float2) * ta.size(), hipMemcpyHostToDevice, s);
    if (he == hipSuccess)
        he = hipMemcpyAsync(h->d_ABb, tb.data(), sizeof(float2) * tb.size(), hipMemcpyHostToDevice, s);
    if (he == hipSuccess)
        he = hipMemcpyAsync(h->d_typeA, atom_type_A, sizeof(int) * N, hipMemcpyHostToDevice, s);
    if (he == hipSuccess)
        he = hipMemcpyAsync(h->d_typeB, atom_type_B, sizeof(int) * N, hipMemcpyHostToDevice, s);
    int rc = he == hipSuccess ? fill_async(h->d_LJ_energy_atom, 0.f, N, s) : (int)he;
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = SPONGE_AMD_ERROR_MALLOC_FAILED;
    if (rc) {
        ljsc_free(h);
        delete h;
        return rc;
    }
    *out = h;
    return 0;
}

int sponge_ljsc_long_range_factors(sponge_ljsc* h, float* factor, float* factor_TI) {
    if (!h) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (factor) *factor = h->long_range_factor;
    if (factor_TI) *factor_TI = h->long_range_factor_TI;
    return 0;
}

int sponge_ljsc_direct_force(sponge_ljsc* h, sponge_nl* nl, const float* crd, const float* charge,
                             float* frc, const float* cell, const float* rcell, float pme_beta,
                             int need_atom_energy, float* atom_energy, int need_virial,
                             float* atom_virial, float* atom_direct_pme_energy) {
    if (!h || !nl || !crd || !charge || !frc || !cell || !rcell)
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    int rc = prepare(h, nl, crd, charge, nullptr);
    if (rc) return rc;
    const LT6 c = lt_from(cell), r = lt_from(rcell);
    const bool E = need_atom_energy != 0, V = need_virial != 0;
    timer_begin("ljsc_pairs", h->stream);
    rc = lt_is_orthorhombic(c) ? launch<true>(h, nl, c, r, pme_beta, true, E, V, false)
                               : launch<false>(h, nl, c, r, pme_beta, true, E, V, false);
    timer_end("ljsc_pairs", h->stream);
    if (rc) return rc;
    const dim3 g(grid_for(h->N, 256)), b(256);
#define SA_FIN(e, v)                                                                              \
    SA_LAUNCH(256, (&ljsc_finalize_body<e, v>), (k_ljsc_finalize<e, v>), g, b, 0, h->stream, h->N, \
              nl->d_sidx, h->d_facc, h->d_eacc, h->d_vacc, frc, atom_energy, h->d_LJ_energy_atom, \
              atom_direct_pme_energy, atom_virial, 0, (int*)nullptr)
    if (E && V) SA_FIN(true, true);
    else if (E) SA_FIN(true, false);
    else if (V) SA_FIN(false, true);
    else SA_FIN(false, false);
#undef SA_FIN
    SA_LAUNCH_CHECK();
    return 0;
}

int sponge_ljsc_du_dlambda(sponge_ljsc* h, sponge_nl* nl, const float* crd, const float* charge,
                           const float* charge_BA, const float* cell, const float* rcell,
                           float pme_beta, float* du_lj, float* du_direct) {
    if (!h || !nl || !crd || !charge || !cell || !rcell) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    int rc = prepare(h, nl, crd, charge, charge_BA);
    if (!rc) rc = fill_async(h->d_du, 0.0, 2, h->stream);
    if (rc) return rc;
    const LT6 c = lt_from(cell), r = lt_from(rcell);
    rc = lt_is_orthorhombic(c) ? launch<true>(h, nl, c, r, pme_beta, false, false, false, true)
                               : launch<false>(h, nl, c, r, pme_beta, false, false, false, true);
    if (rc) return rc;
    double d[2];
    SA_CHECK(hipMemcpyAsync(d, h->d_du, sizeof(d), hipMemcpyDeviceToHost, h->stream));
    SA_CHECK(hipStreamSynchronize(h->stream));
    // Get_Partial_H_Partial_Lambda_With_Columb_Direct (LJ_soft_core.cpp:828-880)
    // adds the long-range TI term to the LJ part
    const float volume = c.a11 * c.a22 * c.a33;
    if (du_lj) *du_lj = (float)d[0] + h->long_range_factor_TI / volume;
    if (du_direct) *du_direct = (float)d[1];
    return 0;
}

// LJ_SOFT_CORE::Long_Range_Correction (LJ_soft_core.cpp:928-946)
int sponge_ljsc_long_range_correction(sponge_ljsc* h, int need_pressure, float* d_virial,
                                      int need_potential, float* d_potential, float volume) {
    if (!h || !(volume > 0.f)) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (need_pressure) {
        if (!d_virial) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
        hipLaunchKernelGGL(k_ljsc_add_diag, dim3(1), dim3(1), 0, h->stream, d_virial,
                           2 * h->long_range_factor / volume);
        SA_LAUNCH_CHECK();
    }
    if (need_potential) {
        if (!d_potential) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
        hipLaunchKernelGGL(k_ljsc_add_scalar, dim3(1), dim3(1), 0, h->stream, d_potential,
                           h->long_range_factor / volume);
        SA_LAUNCH_CHECK();
    }
    return 0;
}

int sponge_ljsc_step_print(sponge_ljsc* h, float* LJ_soft_short) {
    if (!h) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    int rc = fill_async(h->d_sum, 0.0, 1, h->stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ljsc_sum, dim3(64), dim3(256), 0, h->stream, h->N, h->d_LJ_energy_atom,
                       h->d_sum);
    SA_LAUNCH_CHECK();
    double sum = 0;
    SA_CHECK(hipMemcpyAsync(&sum, h->d_sum, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    SA_CHECK(hipStreamSynchronize(h->stream));
    if (LJ_soft_short) *LJ_soft_short = (float)sum;
    return 0;
}

int sponge_ljsc_clear(sponge_ljsc* h) {
    if (!h) return 0;
    (void)hipStreamSynchronize(h->stream);
    ljsc_free(h);
    delete h;
    return 0;
}

}  // extern "C"

// The step driver's form of the direct-force call: frc assigned (the step's
// first force term), the list's rebuild flag cleared once the rows were read,
// per-atom energies only on energy steps.
int sa::ljsc_direct_force_md(sponge_ljsc* h, sponge_nl* nl, const float* crd, const float* charge,
                             float* frc, const float* cell, const float* rcell, float pme_beta,
                             int need_atom_energy, float* atom_energy,
                             float* atom_direct_pme_energy, int* reset_flag) {
    int rc = prepare(h, nl, crd, charge, nullptr);
    if (rc) return rc;
    const LT6 c = lt_from(cell), r = lt_from(rcell);
    const bool E = need_atom_energy != 0;
    // the automatic refresh rebuilds once any atom moved more than
    // skin * skin_permit: the force-only kernel may use cell-centre images
    const float reach = nl->opt.refresh_interval <= 0 ? h->cutoff + nl->skin * nl->opt.skin_permit
                                                      : -1.f;
    {
        ScopedTimer timer("ljsc_pairs", h->stream);
        rc = lt_is_orthorhombic(c)
                 ? launch<true>(h, nl, c, r, pme_beta, true, E, false, false, reach)
                 : launch<false>(h, nl, c, r, pme_beta, true, E, false, false, reach);
    }
    if (rc) return rc;
    const dim3 g(grid_for(h->N, 256)), b(256);
    if (E)
        SA_LAUNCH(256, (&ljsc_finalize_body<true, false>), (k_ljsc_finalize<true, false>), g, b, 0,
                  h->stream, h->N, nl->d_sidx, h->d_facc, h->d_eacc, h->d_vacc, frc, atom_energy,
                  h->d_LJ_energy_atom, atom_direct_pme_energy, (float*)nullptr, 1, reset_flag);
    else
        SA_LAUNCH(256, (&ljsc_finalize_body<false, false>), (k_ljsc_finalize<false, false>), g, b,
                  0, h->stream, h->N, nl->d_sidx, h->d_facc, h->d_eacc, h->d_vacc, frc,
                  atom_energy, h->d_LJ_energy_atom, atom_direct_pme_energy, (float*)nullptr, 1,
                  reset_flag);
    SA_LAUNCH_CHECK();
    return 0;
}

void sa::ljsc_set_stream(sponge_ljsc* h, hipStream_t s) { h->stream = s; }
int sa::ljsc_atoms(const sponge_ljsc* h) { return h->N; }
float sa::ljsc_long_range_factor(const sponge_ljsc* h) { return h->long_range_factor; }
