// Pair arithmetic shared by the LJ + PME direct pair kernels (the plain
// module and the solvent_LJ specialisation): the reference's pair functions
// (SPONGE/Lennard_Jones_force/Lennard_Jones_force.h:45-75) evaluated
// branch-free on one pair, and the partner sources of a cell's candidate
// stream (LDS stage or sorted HBM arrays).
#pragma once

#include "internal.h"

namespace sa {

constexpr float kTwoOverSqrtPi = 1.1283791670218446f;  // Lennard_Jones_force.h:8

// Arithmetic per in-cutoff pair on force-only steps (FAST, when beta *
// cutoff <= kErfcxMax): 1/r from the hardware rsqrt and erfc(beta r) as
// exp(-(beta r)^2) erfcx(beta r) with the polynomial erfcx, so one
// exponential serves erfc and the force bracket.  Steps that need energies or
// the virial use the reference's sqrtf, division and erfcf.
constexpr int kPairStage = 3584;
constexpr int kMaxLdsTypes = 64;  // the square (A, B) table lives in LDS: up to 64 LJ types

// Partner data of a row entry: from the block's LDS stage (every position of
// the stream staged) or, for the rare stream longer than the stage, from the
// sorted arrays in HBM.  The two sources are separate code paths chosen once
// per block, so that LDS reads stay ds_read and HBM reads global loads.
struct PairSrc {
    const float4* pos;
    const int* typ;
    const int* gid;
};

template <bool STAGED>
__device__ __forceinline__ int src_index(int t, int nr, const int* r_first, const int* r_off) {
    if (STAGED) return t;
    return stream_to_sorted(t, nr, r_first, r_off);
}

// r^-3, r^-6 and r^-8 each rounded once to FP32 from the FP32 distance, as
// the reference's powf (Lennard_Jones_force.h:45-75): the products are taken
// in FP64.  Repeated FP32 squaring of 1/r instead biases r^-6 by a few ulp,
// which the LJ energy (a difference of the r^-12 and r^-6 sums) turns into a
// ~2e-6 relative error of the total.
__device__ __forceinline__ void ref_powers(float rabs, float& pw3, float& pw6, float& pw8) {
    const double ri = 1.0 / (double)rabs;
    const double ri2 = ri * ri;
    const double ri6 = ri2 * ri2 * ri2;
    pw3 = (float)(ri2 * ri);
    pw6 = (float)ri6;
    pw8 = (float)(ri6 * ri2);
}

// erfc(beta r) and the force bracket erfc(beta r) + 2 beta r / sqrt(pi)
// exp(-beta^2 r^2) from r^2 and r (FAST arithmetic): erfc as exp(-x^2)
// erfcx(x), with beta folded into the constants (hoisted out of a row's
// loop), so a pair costs one exp2 of r^2 times a constant and one rcp.
__device__ __forceinline__ void coul_fast(float r2, float rabs, float beta, float& erfc_br,
                                          float& bracket) {
    constexpr float kLog2e = 1.4426950408889634f;
    const float t = __builtin_amdgcn_rcpf(fmaf(rabs, kErfcxT * beta, 1.0f));
    const float ex = __builtin_amdgcn_exp2f(r2 * (-kLog2e * beta * beta));
    const float ecx = erfcx_poly_t(t);
    erfc_br = ex * ecx;
    bracket = ex * fmaf(rabs, kTwoOverSqrtPi * beta, ecx);
}

// Force (and energy / virial terms) of one pair, branch-free: pairs at or
// beyond the cutoff contribute exact zeros (selects, so the non-finite
// values of a far or padding pair never reach the sums).
// LJ = false: the row atom's type has no LJ coefficient with any type (a
// water hydrogen), so the LJ term of every pair of the row is exactly zero
// and is not evaluated (ab unused).
template <bool NEED_E, bool NEED_V, bool FAST, bool LJ = true>
struct PairTerm {
    float f, lx, ly, lz, elj, ec;  // f: the force factor (force = f d)
    bool in;
    __device__ __forceinline__ PairTerm(const float3& d, float r2, float qi, float qj, float2 ab,
                                        float rc2_thr, float beta) {
        in = r2 < rc2_thr;
        // FAST: hardware rsqrt (1 ulp); otherwise the reference's sqrtf and
        // division, which energies and virials need: the LJ energy is a
        // difference of the r^-12 and r^-6 terms and amplifies a biased 1/r
        float rinv, rabs, pw3, pw6, pw8;
        if (FAST) {
            rinv = __builtin_amdgcn_rsqf(r2);
            rabs = r2 * rinv;
            const float r2inv = rinv * rinv;
            pw6 = LJ ? r2inv * r2inv * r2inv : 0.f;
            pw8 = LJ ? pw6 * r2inv : 0.f;
            pw3 = rinv * r2inv;
        } else {
            rabs = sqrtf(r2);
            rinv = 1.0f / rabs;
            ref_powers(rabs, pw3, pw6, pw8);
        }
        const float qq = qi * qj;
        float erfc_br, bracket;
        if (FAST) {
            coul_fast(r2, rabs, beta, erfc_br, bracket);
        } else {
            const float br = beta * rabs;
            erfc_br = erfcf(br);
            bracket = fmaf(br * kTwoOverSqrtPi, expf(-br * br), erfc_br);
        }
        const float flj = LJ ? (ab.y - ab.x * pw6) * pw8 : 0.f;
        const float fc = qq * pw3 * bracket;
        f = in ? flj - fc : 0.f;
        lx = f * d.x;
        ly = f * d.y;
        lz = f * d.z;
        elj = 0.f;
        ec = 0.f;
        if (NEED_E) {
            elj = in && LJ ? (0.083333333f * ab.x * pw6 - 0.166666667f * ab.y) * pw6 : 0.f;
            ec = in ? qq * erfc_br / rabs : 0.f;
        }
    }
};

}  // namespace sa
