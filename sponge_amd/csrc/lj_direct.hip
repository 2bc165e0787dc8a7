// LJ + PME direct-space pair forces for gfx950.
//
// Reference: SPONGE/Lennard_Jones_force/Lennard_Jones_force.{h,cpp}.
// Per in-cutoff pair (strict dr_abs < cutoff, Lennard_Jones_force.cpp:74):
//   F_LJ/r   = (B - A r^-6) r^-8                      (Get_LJ_Force, .h:52-56)
//   F_C/r    = qq r^-3 (2/sqrt(pi) br e^{-b^2r^2} + erfc(br))   (.h:67-74)
//   f        = F_LJ/r - F_C/r;  F_i += f dr, F_j -= f dr with dr = r_j - r_i
//   E_LJ     = (A/12 r^-6 - B/6) r^-6;  E_C = qq erfc(br)/r     (.h:45-50, 62-66)
//   virial_i -= Get_Virial_From_Force_Dis(f dr, dr)
// Energies and virial of a pair go to its row atom i (the lower global id),
// exactly as in the reference's half-list kernel.
//
// Layout: rows are in the neighbor list's cell-sorted index space and hold
// both directions of every pair, each entry a 16-bit position in the row
// cell's candidate stream (the atoms of its stencil cells, in sorted order).
// Positions and charges are gathered once per step into a float4 array in
// sorted space.  One workgroup (16 wavefronts, so that two workgroups fill a
// CU's LDS and give 8 wavefronts per SIMD to hide the list and LDS
// latencies) per cell stages that stream in LDS; one
// wavefront owns one row and reduces its force by shuffles into a plain store
// (no atomics); a final pass scatters the rows into the caller's AoS arrays.
// The kernel over a caller-provided reference-format half list (k_lj_pairs)
// keeps the reference's column updates, as hardware float atomics.
#include <mutex>
#include <vector>

#include "internal.h"
#include "pair_term.h"
#include "batch_launch.h"
#include "timing.h"

using namespace sa;

namespace {

template <bool ORTHO, bool NEED_E, bool NEED_V>
__global__ void __launch_bounds__(256) k_lj_pairs(
    int nrows, const int* __restrict__ list_count, const int* __restrict__ list, int max_nb,
    const float4* __restrict__ pos, const int* __restrict__ ptype,
    const float2* __restrict__ AB, LT6 c, LT6 r, float rc2_thr, float beta,
    float4* __restrict__ facc, float2* __restrict__ eacc, float* __restrict__ vacc) {
    const int lane = threadIdx.x & (kWave - 1);
    const int p = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (p >= nrows) return;
    const float4 pi4 = pos[p];
    const float3 pi = make_float3(pi4.x, pi4.y, pi4.z);
    const int ti = ptype[p];
    const int n = list_count[p];
    const int* row = list + (size_t)p * max_nb;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    double elj = 0, ec = 0;  // energy and virial terms summed in FP64 per lane
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    for (int k = lane; k < n; k += kWave) {
        const int q = row[k];
        const float4 pj4 = pos[q];
        const float3 d = disp_pbc<ORTHO>(make_float3(pj4.x, pj4.y, pj4.z), pi, c, r);
        const float r2 = norm2_exact(d);
        if (r2 < rc2_thr) {
            const float2 ab = AB[lj_pair_index(ti, ptype[q])];
            const float rabs = sqrtf(r2);
            float pw3, pw6, pw8;
            ref_powers(rabs, pw3, pw6, pw8);
            const float qq = pi4.w * pj4.w;
            const float br = beta * rabs;
            const float erfc_br = erfcf(br);
            const float flj = (ab.y - ab.x * pw6) * pw8;
            const float fc = qq * pw3 * (br * kTwoOverSqrtPi * expf(-br * br) + erfc_br);
            const float f = flj - fc;
            const float lx = f * d.x, ly = f * d.y, lz = f * d.z;
            fx += lx;
            fy += ly;
            fz += lz;
            unsafeAtomicAdd(&facc[q].x, -lx);
            unsafeAtomicAdd(&facc[q].y, -ly);
            unsafeAtomicAdd(&facc[q].z, -lz);
            if (NEED_V) {
                v0 -= lx * d.x;
                v1 -= lx * d.y + ly * d.x;
                v2 -= ly * d.y;
                v3 -= lx * d.z + lz * d.x;
                v4 -= ly * d.z + lz * d.y;
                v5 -= lz * d.z;
            }
            if (NEED_E) {
                elj += (0.083333333f * ab.x * pw6 - 0.166666667f * ab.y) * pw6;
                ec += qq * erfc_br / rabs;
            }
        }
    }
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
    // energies and virial: a few terms per lane in FP32, the 64 lane
    // partials summed in FP64
    double delj = 0, dec = 0, dv0 = 0, dv1 = 0, dv2 = 0, dv3 = 0, dv4 = 0, dv5 = 0;
    if (NEED_E) {
        delj = wave_sum((double)elj);
        dec = wave_sum((double)ec);
    }
    if (NEED_V) {
        dv0 = wave_sum((double)v0);
        dv1 = wave_sum((double)v1);
        dv2 = wave_sum((double)v2);
        dv3 = wave_sum((double)v3);
        dv4 = wave_sum((double)v4);
        dv5 = wave_sum((double)v5);
    }
    if (lane == 0) {
        unsafeAtomicAdd(&facc[p].x, fx);
        unsafeAtomicAdd(&facc[p].y, fy);
        unsafeAtomicAdd(&facc[p].z, fz);
        if (NEED_E) eacc[p] = make_float2((float)delj, (float)dec);
        if (NEED_V) {
            float* v = vacc + (size_t)6 * p;
            v[0] = (float)dv0; v[1] = (float)dv1; v[2] = (float)dv2;
            v[3] = (float)dv3; v[4] = (float)dv4; v[5] = (float)dv5;
        }
    }
}

// Cell-block rows: one workgroup per neighbor-list cell stages the cell's
// candidate stream -- current positions with charges, LJ types and global ids
// of every atom of its stencil -- into LDS once, with the square (A, B) table
// of the type pairs; each wavefront then owns one row (an atom of the cell)
// and reads its partners from LDS through the 16-bit stream positions of the
// list.  Forces need no atomics (rows hold both directions); energies and
// virials are counted in the row of the lower global id only, the
// reference's half-list row.
//
// One row (wavefront-wide): two adjacent partner entries per lane and pass
// (one 32-bit load: rows start 4-byte aligned since the row capacity is
// even), with the next pass's pair loaded before the current one is used so
// that the list's memory latency overlaps the arithmetic.
// SHIFT: the staged partner positions are the images nearest the cell's
// centre C (see k_lj_cells), so the displacement is a plain difference with
// the row atom's own image nearest C.
// SOLV: the solvent_LJ split of the reference (Lennard_Jones_force.cpp:55-58
// with solvent_numbers > 0): atoms from solvent_start on are the solvent
// tail, whose pairs with each other the solvent module evaluates; a solvent
// row here keeps only its partners below solvent_start.
// LJ = false: the row atom's type has no LJ with any type (the pair
// arithmetic is Coulomb only and the partners' types are not read).
template <bool ORTHO, bool NEED_E, bool NEED_V, bool FAST, bool STAGED, bool SHIFT = false,
          bool SOLV = false, bool LJ = true>
__device__ __forceinline__ void lj_row(int p, int lane, int n, const unsigned short* __restrict__ row,
                                       PairSrc src, int nr, const int* r_first, const int* r_off,
                                       const float4* __restrict__ pos, const float2* trow, int gi,
                                       const LT6& c, const LT6& r, float rc2_thr, float beta,
                                       float4* __restrict__ facc, float2* __restrict__ eacc,
                                       float* __restrict__ vacc, float3 C = float3{0.f, 0.f, 0.f},
                                       int solvent_start = 0) {
    const bool solvent_row = SOLV && gi >= solvent_start;
    const float4 pi4 = pos[p];
    const float3 pi = SHIFT ? near_image(make_float3(pi4.x, pi4.y, pi4.z), C, c, r)
                            : make_float3(pi4.x, pi4.y, pi4.z);
    float fx = 0.f, fy = 0.f, fz = 0.f;
    double elj = 0, ec = 0;  // energy and virial terms summed in FP64 per lane
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    const unsigned* row2 = reinterpret_cast<const unsigned*>(row);
    const int npairs = (n + 1) >> 1;  // 32-bit words holding the row
    const int last = npairs > 0 ? npairs - 1 : 0;
    unsigned cur = row2[lane < last ? lane : last];
    for (int w = lane; w - lane < npairs; w += kWave) {
        const int wn = w + kWave;
        unsigned nxt = row2[wn < last ? wn : last];
        const bool has_a = w < npairs;
        const bool has_b = 2 * w + 1 < n;
        const int ta = (int)(cur & 0xffffu);
        const int tb = has_b ? (int)(cur >> 16) : ta;
        const int qa = src_index<STAGED>(ta, nr, r_first, r_off);
        const int qb = src_index<STAGED>(tb, nr, r_first, r_off);
        const float4 pa = src.pos[qa];
        const float4 pb = src.pos[qb];
        const float2 aba = LJ ? trow[src.typ[qa]] : make_float2(0.f, 0.f);
        const float2 abb = LJ ? trow[src.typ[qb]] : make_float2(0.f, 0.f);
        float3 da, db;
        float r2a, r2b;
        if (SHIFT) {
            da = make_float3(pa.x - pi.x, pa.y - pi.y, pa.z - pi.z);
            db = make_float3(pb.x - pi.x, pb.y - pi.y, pb.z - pi.z);
            r2a = fmaf(da.z, da.z, fmaf(da.y, da.y, da.x * da.x));
            r2b = fmaf(db.z, db.z, fmaf(db.y, db.y, db.x * db.x));
        } else {
            da = disp_pbc<ORTHO>(make_float3(pa.x, pa.y, pa.z), pi, c, r);
            db = disp_pbc<ORTHO>(make_float3(pb.x, pb.y, pb.z), pi, c, r);
            r2a = norm2_exact(da);
            r2b = norm2_exact(db);
        }
        // an absent entry is given r2 = the cutoff threshold: out of range
        r2a = has_a ? r2a : rc2_thr;
        r2b = has_b ? r2b : rc2_thr;
        if (SOLV && solvent_row) {
            r2a = src.gid[qa] < solvent_start ? r2a : rc2_thr;
            r2b = src.gid[qb] < solvent_start ? r2b : rc2_thr;
        }
        const PairTerm<NEED_E, NEED_V, FAST, LJ> A(da, r2a, pi4.w, pa.w, aba, rc2_thr, beta);
        const PairTerm<NEED_E, NEED_V, FAST, LJ> B(db, r2b, pi4.w, pb.w, abb, rc2_thr, beta);
        fx += A.lx + B.lx;
        fy += A.ly + B.ly;
        fz += A.lz + B.lz;
        if (NEED_E || NEED_V) {
            // energies and virial in the row of the lower global id
            const bool ca = A.in && src.gid[qa] > gi;
            const bool cb = B.in && src.gid[qb] > gi;
            if (NEED_V) {
                const float ax = ca ? A.lx : 0.f, ay = ca ? A.ly : 0.f, az = ca ? A.lz : 0.f;
                const float bx = cb ? B.lx : 0.f, by = cb ? B.ly : 0.f, bz = cb ? B.lz : 0.f;
                v0 -= ax * da.x + bx * db.x;
                v1 -= (ax * da.y + ay * da.x) + (bx * db.y + by * db.x);
                v2 -= ay * da.y + by * db.y;
                v3 -= (ax * da.z + az * da.x) + (bx * db.z + bz * db.x);
                v4 -= (ay * da.z + az * da.y) + (by * db.z + bz * db.y);
                v5 -= az * da.z + bz * db.z;
            }
            if (NEED_E) {
                elj += (ca ? A.elj : 0.f) + (cb ? B.elj : 0.f);
                ec += (ca ? A.ec : 0.f) + (cb ? B.ec : 0.f);
            }
        }
        // the loaded word is opaque from here: the compiler cannot fold the
        // prefetch back into a load at the top of the next pass
        asm volatile("" : "+v"(nxt));
        cur = nxt;
    }
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
    double delj = 0, dec = 0, dv0 = 0, dv1 = 0, dv2 = 0, dv3 = 0, dv4 = 0, dv5 = 0;
    if (NEED_E) {
        delj = wave_sum((double)elj);
        dec = wave_sum((double)ec);
    }
    if (NEED_V) {
        dv0 = wave_sum((double)v0);
        dv1 = wave_sum((double)v1);
        dv2 = wave_sum((double)v2);
        dv3 = wave_sum((double)v3);
        dv4 = wave_sum((double)v4);
        dv5 = wave_sum((double)v5);
    }
    if (lane == 0) {
        facc[p] = make_float4(fx, fy, fz, 0.f);
        if (NEED_E) eacc[p] = make_float2((float)delj, (float)dec);
        if (NEED_V) {
            float* v = vacc + (size_t)6 * p;
            v[0] = (float)dv0; v[1] = (float)dv1; v[2] = (float)dv2;
            v[3] = (float)dv3; v[4] = (float)dv4; v[5] = (float)dv5;
        }
    }
}

// One pair of lj_row_fast: the force factor of PairTerm's FAST arithmetic
// summed into the row force by fused multiply-adds.  has = false gives the
// absent entry of a partial pass r^2 = the cutoff threshold (out of range).
template <bool LJ>
__device__ __forceinline__ void fast_pair(const float4& pj, float2 ab, const float3& pi, float qi,
                                          float rc2_thr, float beta, bool has, float& fx,
                                          float& fy, float& fz) {
    const float3 d = make_float3(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z);
    float r2 = fmaf(d.z, d.z, fmaf(d.y, d.y, d.x * d.x));
    r2 = has ? r2 : rc2_thr;
    const PairTerm<false, false, true, LJ> t(d, r2, qi, pj.w, ab, rc2_thr, beta);
    fx = fmaf(t.f, d.x, fx);
    fy = fmaf(t.f, d.y, fy);
    fz = fmaf(t.f, d.z, fz);
}

// Force-only row of a staged cell whose partners sit at their images nearest
// the cell's centre (lj_row with FAST, STAGED and SHIFT): the same pairs and
// pair arithmetic, with the passes in which all 128 entries exist (the row's
// first n / 128, the same count in every lane) run without presence tests;
// only the partial last pass masks its absent entries.
template <bool LJ>
__device__ __forceinline__ void lj_row_fast(int p, int lane, int n, const unsigned short* __restrict__ row,
                                            const float4* cpos, const int* ctyp,
                                            const float4* __restrict__ pos, const float2* trow,
                                            const LT6& c, const LT6& r, float rc2_thr, float beta,
                                            float4* __restrict__ facc, float3 C) {
    const float4 pi4 = pos[p];
    const float3 pi = near_image(make_float3(pi4.x, pi4.y, pi4.z), C, c, r);
    float fx = 0.f, fy = 0.f, fz = 0.f;
    const int npairs = (n + 1) >> 1;  // 32-bit words holding the row
    const int full = n >> 7;          // passes whose 2 x 64 entries all exist
    const unsigned* row2 = reinterpret_cast<const unsigned*>(row);
    // unsigned word indices: a 32-bit offset from the row's base
    const unsigned last = npairs > 0 ? npairs - 1 : 0;
    unsigned cur = row2[(unsigned)lane < last ? (unsigned)lane : last];
    int w = lane;
    auto full_pass = [&](unsigned word) {
        const int ta = (int)(word & 0xffffu), tb = (int)(word >> 16);
        const float4 pa = cpos[ta];
        const float4 pb = cpos[tb];
        const float2 aba = LJ ? trow[ctyp[ta]] : make_float2(0.f, 0.f);
        const float2 abb = LJ ? trow[ctyp[tb]] : make_float2(0.f, 0.f);
        fast_pair<LJ>(pa, aba, pi, pi4.w, rc2_thr, beta, true, fx, fy, fz);
        fast_pair<LJ>(pb, abb, pi, pi4.w, rc2_thr, beta, true, fx, fy, fz);
    };
    // two full passes per iteration (their independent arithmetic
    // interleaves; 62 VGPRs keep 8 wavefronts per SIMD, which the kernel needs:
    // separate force sums per pass cost 66 VGPRs, 7 wavefronts and 23% more
    // time), the words of both loaded ahead
    int k = 0;
    for (; k + 2 <= full; k += 2, w += 2 * kWave) {
        const unsigned wn1 = (unsigned)(w + kWave), wn2 = (unsigned)(w + 2 * kWave);
        const unsigned nxt1 = row2[wn1 < last ? wn1 : last];
        unsigned nxt2 = row2[wn2 < last ? wn2 : last];
        full_pass(cur);
        full_pass(nxt1);
        // the loaded word is opaque from here: the compiler cannot fold the
        // prefetch back into a load at the top of the next iteration
        asm volatile("" : "+v"(nxt2));
        cur = nxt2;
    }
    if (k < full) {
        const unsigned wn = (unsigned)(w + kWave);
        unsigned nxt = row2[wn < last ? wn : last];
        full_pass(cur);
        asm volatile("" : "+v"(nxt));
        cur = nxt;
        w += kWave;
    }
    if (full * 2 * kWave < n) {
        // cur is the row's word w, or its last word for lanes past the end
        const bool has_a = w < npairs;
        const bool has_b = 2 * w + 1 < n;
        const int ta = (int)(cur & 0xffffu);
        const int tb = has_b ? (int)(cur >> 16) : ta;
        const float4 pa = cpos[ta];
        const float4 pb = cpos[tb];
        const float2 aba = LJ ? trow[ctyp[ta]] : make_float2(0.f, 0.f);
        const float2 abb = LJ ? trow[ctyp[tb]] : make_float2(0.f, 0.f);
        fast_pair<LJ>(pa, aba, pi, pi4.w, rc2_thr, beta, has_a, fx, fy, fz);
        fast_pair<LJ>(pb, abb, pi, pi4.w, rc2_thr, beta, has_b, fx, fy, fz);
    }
    fx = wave_sum(fx);
    fy = wave_sum(fy);
    fz = wave_sum(fz);
    if (lane == 0) facc[p] = make_float4(fx, fy, fz, 0.f);
}

template <bool ORTHO, bool NEED_E, bool NEED_V, bool FAST, bool SOLV = false>
__device__ __forceinline__ void lj_cells_body(int sa_block, 
    const int* __restrict__ start, const int* __restrict__ nranges, int max_ranges,
    const int* __restrict__ g_rfirst, const int* __restrict__ g_roff,
    const int* __restrict__ list_count, const unsigned short* __restrict__ list, int cap,
    const float4* __restrict__ pos, const int* __restrict__ ptype, const int* __restrict__ pgid,
    const float2* __restrict__ AB_sq, int T, LT6 c, LT6 r, float rc2_thr, float beta, int stage,
    float gl, int Nx, int Ny, int Nz, float reach,
    float4* __restrict__ facc, float2* __restrict__ eacc, float* __restrict__ vacc,
    int solvent_start) {
    constexpr bool kGid = NEED_E || NEED_V || SOLV;  // global ids staged
    extern __shared__ float4 cpos[];
    int* ctyp = reinterpret_cast<int*>(cpos + stage);
    int* cgid = ctyp + stage;
    float2* cab = reinterpret_cast<float2*>(ctyp + (kGid ? 2 : 1) * stage);
    __shared__ int r_first[kMaxRanges];
    __shared__ int r_off[kMaxRanges + 1];
    const int cell = sa_block;
    const int a0 = start[cell], a1 = start[cell + 1];
    if (a0 == a1) return;
    // Force-only steps in an orthorhombic box wider than twice (half the
    // cell + reach) on every axis, reach = cutoff + the largest distance a row
    // atom may have moved since the list was built (skin * skin_permit under
    // the automatic refresh; reach < 0 when the list gives no such bound):
    // the partners are staged as their images nearest the cell's centre C.
    // A row atom then lies within h/2 + drift of C and a partner within the
    // cutoff of it within h/2 + reach of C on each axis, less than L/2, so
    // that partner's nearest image to C is its minimum image from the row
    // atom and the displacement is a plain difference.  It differs from the
    // reference arithmetic only by the rounding of the shifted coordinates
    // (a few ulp of a coordinate), and pairs beyond the cutoff stay beyond it
    // (no image is nearer than the minimum image).
    float3 C = make_float3(0.f, 0.f, 0.f);
    bool shift = false;
    if (FAST && ORTHO && reach >= 0.f) {
        const int ix = cell / (Ny * Nz), iy = (cell / Nz) % Ny, iz = cell % Nz;
        const float lx = ix * gl, hx = ix == Nx - 1 ? c.a11 : (ix + 1) * gl;
        const float ly = iy * gl, hy = iy == Ny - 1 ? c.a22 : (iy + 1) * gl;
        const float lz = iz * gl, hz = iz == Nz - 1 ? c.a33 : (iz + 1) * gl;
        C = make_float3(0.5f * (lx + hx), 0.5f * (ly + hy), 0.5f * (lz + hz));
        // 0.01 A of margin against the rounding of C and of the coordinates
        shift = 0.5f * c.a11 > reach + 0.5f * (hx - lx) + 0.01f &&
                0.5f * c.a22 > reach + 0.5f * (hy - ly) + 0.01f &&
                0.5f * c.a33 > reach + 0.5f * (hz - lz) + 0.01f;
    }
    const int nr = nranges[cell];
    for (int k = threadIdx.x; k < nr; k += blockDim.x) r_first[k] = g_rfirst[(size_t)cell * max_ranges + k];
    for (int k = threadIdx.x; k <= nr; k += blockDim.x) r_off[k] = g_roff[(size_t)cell * (max_ranges + 1) + k];
    for (int k = threadIdx.x; k < T * T; k += blockDim.x) cab[k] = AB_sq[k];
    __syncthreads();
    // types with an LJ coefficient (force-only rows of the others skip the LJ
    // arithmetic)
    __shared__ unsigned char tlj[kMaxLdsTypes];
    if (FAST)
        for (int a = threadIdx.x; a < T; a += blockDim.x) {
            bool any = false;
            for (int b = 0; b < T; ++b) {
                const float2 v = cab[a * T + b];
                any |= v.x != 0.f || v.y != 0.f;
            }
            tlj[a] = any;
        }
    const int total = r_off[nr];
    const bool staged = total <= stage;
    if (staged)
        stage_stream(nr, r_first, r_off, total, [&](int t, int q) {
            float4 v = pos[q];
            if (shift) {
                const float3 w = near_image(make_float3(v.x, v.y, v.z), C, c, r);
                v = make_float4(w.x, w.y, w.z, v.w);
            }
            cpos[t] = v;
            ctyp[t] = ptype[q];
            if (kGid) cgid[t] = pgid[q];
        });
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave, nwave = blockDim.x / kWave;
    for (int p = a0 + wave; p < a1; p += nwave) {
        const int ti = ptype[p];
        const float2* trow = cab + ti * T;
        const int gi = kGid ? pgid[p] : 0;
        const int n = list_count[p];
        const unsigned short* row = list + (size_t)p * cap;
        if (FAST && !SOLV && staged && shift) {
            if (tlj[ti])
                lj_row_fast<true>(p, lane, n, row, cpos, ctyp, pos, trow, c, r, rc2_thr, beta,
                                  facc, C);
            else
                lj_row_fast<false>(p, lane, n, row, cpos, ctyp, pos, trow, c, r, rc2_thr, beta,
                                   facc, C);
            continue;
        }
        if (FAST && !SOLV && !tlj[ti]) {
            if (staged)
                lj_row<ORTHO, false, false, true, true, false, false, false>(
                    p, lane, n, row, PairSrc{cpos, ctyp, cgid}, nr, r_first, r_off, pos, trow, gi,
                    c, r, rc2_thr, beta, facc, eacc, vacc, C, solvent_start);
            else
                lj_row<ORTHO, false, false, true, false, false, false, false>(
                    p, lane, n, row, PairSrc{pos, ptype, pgid}, nr, r_first, r_off, pos, trow, gi,
                    c, r, rc2_thr, beta, facc, eacc, vacc, C, solvent_start);
            continue;
        }
        if (staged && shift)
            lj_row<ORTHO, NEED_E, NEED_V, FAST, true, true, SOLV>(
                p, lane, n, row, PairSrc{cpos, ctyp, cgid}, nr, r_first, r_off, pos, trow, gi, c,
                r, rc2_thr, beta, facc, eacc, vacc, C, solvent_start);
        else if (staged)
            lj_row<ORTHO, NEED_E, NEED_V, FAST, true, false, SOLV>(
                p, lane, n, row, PairSrc{cpos, ctyp, cgid}, nr, r_first, r_off, pos, trow, gi, c,
                r, rc2_thr, beta, facc, eacc, vacc, C, solvent_start);
        else
            lj_row<ORTHO, NEED_E, NEED_V, FAST, false, false, SOLV>(
                p, lane, n, row, PairSrc{pos, ptype, pgid}, nr, r_first, r_off, pos, trow, gi, c,
                r, rc2_thr, beta, facc, eacc, vacc, C, solvent_start);
    }
}

}  // namespace

// The force-only kernels run at 8 wavefronts per SIMD (at most 64 VGPRs):
// at 7 the same code takes ~20% longer.
template <>
struct sa::BodyWaves<&lj_cells_body<true, false, false, true, false>> {
    static constexpr int value = 8;
};
template <>
struct sa::BodyWaves<&lj_cells_body<false, false, false, true, false>> {
    static constexpr int value = 8;
};

namespace {

template <bool ORTHO, bool NEED_E, bool NEED_V, bool FAST, bool SOLV = false>
__global__ void __launch_bounds__(1024)
__attribute__((amdgpu_waves_per_eu(FAST && !NEED_E && !NEED_V && !SOLV ? 8 : 1))) k_lj_cells(
    const int* __restrict__ start, const int* __restrict__ nranges, int max_ranges,
    const int* __restrict__ g_rfirst, const int* __restrict__ g_roff,
    const int* __restrict__ list_count, const unsigned short* __restrict__ list, int cap,
    const float4* __restrict__ pos, const int* __restrict__ ptype, const int* __restrict__ pgid,
    const float2* __restrict__ AB_sq, int T, LT6 c, LT6 r, float rc2_thr, float beta, int stage,
    float gl, int Nx, int Ny, int Nz, float reach,
    float4* __restrict__ facc, float2* __restrict__ eacc, float* __restrict__ vacc,
    int solvent_start) {
    lj_cells_body<ORTHO, NEED_E, NEED_V, FAST, SOLV>(blockIdx.x, start, nranges, max_ranges, g_rfirst, g_roff, list_count, list, cap, pos, ptype, pgid, AB_sq, T, c, r, rc2_thr, beta, stage, gl, Nx, Ny, Nz, reach, facc, eacc, vacc, solvent_start);
}

// Prune of the full rows into the inner rows that the MD driver's pair
// kernel reads: one workgroup per cell stages the positions of the cell's
// candidate stream (this step's coordinates, packed by k_nl_pack); one
// wavefront per row keeps, in row order, the entries within cutoff + delta
// (keep2 = (cutoff + delta)^2) and records the row atom's position of this
// prune.  It runs only on steps whose device flags ask for it: the list was
// rebuilt, or an atom moved more than delta / 2 since the last prune.
template <bool ORTHO>
__device__ __forceinline__ void lj_prune_cells_body(
    int sa_block, const int* __restrict__ need, const int* __restrict__ need_prune,
    const int* __restrict__ start, const int* __restrict__ nranges, int max_ranges,
    const int* __restrict__ g_rfirst, const int* __restrict__ g_roff,
    const int* __restrict__ list_count, const unsigned short* __restrict__ list, int cap,
    const float4* __restrict__ pos, const int* __restrict__ pgid, LT6 c, LT6 r, float keep2,
    int stage, int* __restrict__ ilist_count, unsigned short* __restrict__ ilist,
    float* __restrict__ prune_crd) {
    if (*need == 0 && *need_prune == 0) return;
    extern __shared__ float4 cpos[];
    __shared__ int r_first[kMaxRanges];
    __shared__ int r_off[kMaxRanges + 1];
    const int cell = sa_block;
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
        const float4 pi4 = pos[p];
        const float3 pi = make_float3(pi4.x, pi4.y, pi4.z);
        const int n = list_count[p];
        const unsigned short* row = list + (size_t)p * cap;
        unsigned short* out = ilist + (size_t)p * cap;
        int kept = 0;
        for (int k0 = 0; k0 < n; k0 += kWave) {
            const int k = k0 + lane;
            bool keep = false;
            unsigned short t = 0;
            if (k < n) {
                t = row[k];
                const float4 pj = staged ? cpos[t] : pos[stream_to_sorted(t, nr, r_first, r_off)];
                const float3 d = disp_pbc<ORTHO>(make_float3(pj.x, pj.y, pj.z), pi, c, r);
                keep = norm2_exact(d) < keep2;
            }
            const unsigned long long m = __ballot(keep);
            if (keep) out[kept + lanes_below(m)] = t;
            kept += __popcll(m);
        }
        if (lane == 0) {
            ilist_count[p] = kept;
            const int g = pgid[p];
            prune_crd[3 * g] = pi4.x;
            prune_crd[3 * g + 1] = pi4.y;
            prune_crd[3 * g + 2] = pi4.z;
        }
    }
}

template <bool ORTHO>
__global__ void __launch_bounds__(1024) k_lj_prune_cells(
    const int* __restrict__ need, const int* __restrict__ need_prune,
    const int* __restrict__ start, const int* __restrict__ nranges, int max_ranges,
    const int* __restrict__ g_rfirst, const int* __restrict__ g_roff,
    const int* __restrict__ list_count, const unsigned short* __restrict__ list, int cap,
    const float4* __restrict__ pos, const int* __restrict__ pgid, LT6 c, LT6 r, float keep2,
    int stage, int* __restrict__ ilist_count, unsigned short* __restrict__ ilist,
    float* __restrict__ prune_crd) {
    lj_prune_cells_body<ORTHO>(blockIdx.x, need, need_prune, start, nranges, max_ranges, g_rfirst,
                               g_roff, list_count, list, cap, pos, pgid, c, r, keep2, stage,
                               ilist_count, ilist, prune_crd);
}

// Scatter of the per-row accumulators to the caller's arrays: frc +=,
// atom_energy += E_LJ + E_C, LJ energy and direct-PME energy overwritten
// (the reference clears them before its kernel, Lennard_Jones_force.cpp:
// 343-349), atom_virial +=.
template <bool NEED_E, bool NEED_V>
__device__ __forceinline__ void lj_finalize_body(int sa_block, int N, const int* __restrict__ sidx, const float4* __restrict__ facc,
                              const float2* __restrict__ eacc, const float* __restrict__ vacc,
                              float* __restrict__ frc, float* __restrict__ atom_energy,
                              float* __restrict__ lj_energy, float* __restrict__ direct_energy,
                              float* __restrict__ atom_virial, int assign,
                              int* __restrict__ reset_flag, int* __restrict__ reset_flag2) {
    int p = sa_block * blockDim.x + threadIdx.x;
    if (p == 0 && reset_flag) *reset_flag = 0;
    if (p == 0 && reset_flag2) *reset_flag2 = 0;
    if (p >= N) return;
    const int i = sidx ? sidx[p] : p;
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
    if (NEED_E) {
        const float2 e = eacc[p];
        atom_energy[i] += e.x + e.y;
        if (lj_energy) lj_energy[i] = e.x;
        if (direct_energy) direct_energy[i] = e.y;
    }
    if (NEED_V) {
        const float* v = vacc + (size_t)6 * p;
        float* o = atom_virial + (size_t)6 * i;
        o[0] += v[0]; o[1] += v[1]; o[2] += v[2]; o[3] += v[3]; o[4] += v[4]; o[5] += v[5];
    }
}

template <bool NEED_E, bool NEED_V>
__global__ void k_lj_finalize(int N, const int* __restrict__ sidx, const float4* __restrict__ facc,
                              const float2* __restrict__ eacc, const float* __restrict__ vacc,
                              float* __restrict__ frc, float* __restrict__ atom_energy,
                              float* __restrict__ lj_energy, float* __restrict__ direct_energy,
                              float* __restrict__ atom_virial, int assign,
                              int* __restrict__ reset_flag, int* __restrict__ reset_flag2) {
    lj_finalize_body<NEED_E, NEED_V>(blockIdx.x, N, sidx, facc, eacc, vacc, frc, atom_energy,
                                     lj_energy, direct_energy, atom_virial, assign, reset_flag,
                                     reset_flag2);
}

// Identity-order pack for the reference-format list entry point.
__global__ void k_lj_pack_local(int N, const float* __restrict__ crd,
                                const float* __restrict__ charge, const int* __restrict__ type,
                                float4* __restrict__ pos, int* __restrict__ ptype,
                                float4* __restrict__ facc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    pos[i] = make_float4(crd[3 * i], crd[3 * i + 1], crd[3 * i + 2], charge[i]);
    ptype[i] = type[i];
    facc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_add_scalar(float* p, float v) { p[0] += v; }

__global__ void k_add_diag(float* p, float v) {
    p[0] += v;
    p[2] += v;
    p[5] += v;
}

// Sum of a float list with double accumulation (Sum_Of_List's double
// partials, SPONGE/common.cpp:329-392).
__global__ void k_sum_f32(int n, const float* __restrict__ a, double* out) {
    double s = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += a[i];
    s = wave_sum(s);
    if ((threadIdx.x & (kWave - 1)) == 0) unsafeAtomicAdd(out, s);
}

// allow_shift: the caller updated the list (CONDITIONAL or FORCED) on these
// coordinates, so every atom is within the refresh bound of its build cell
// (the step driver); the public entry point makes no such assumption.
template <bool ORTHO>
int launch_rows(sponge_lj* lj, sponge_nl* nl, LT6 c, LT6 r, float beta, int need_e, int need_v,
                hipStream_t s, bool allow_shift = false, int solvent_start = -1,
                bool inner = false) {
    const int* rows_count = inner ? nl->d_ilist_count : nl->d_list_count;
    const unsigned short* rows = inner ? nl->d_ilist : nl->d_list;
    ScopedTimer timer("lj_pairs", s);
    const bool solv = solvent_start >= 0 && solvent_start < lj->N;
    const size_t sh =
        (sizeof(float4) + ((need_e || need_v || solv) ? 2 : 1) * sizeof(int)) * lj->stage +
        sizeof(float2) * lj->T * lj->T;
    const bool fast = beta * lj->cutoff <= kErfcxMax && !need_e && !need_v;
    // how far a row atom can be from its build-time cell: the automatic
    // refresh rebuilds once any atom moved more than skin * skin_permit
    const float reach = allow_shift && nl->opt.refresh_interval <= 0
                            ? lj->cutoff + nl->skin * nl->opt.skin_permit
                            : -1.f;
#define SA_LJ_ROWS(E, V, F, SV)                                                                 \
    SA_LAUNCH(1024, (&lj_cells_body<ORTHO, E, V, F, SV>), (k_lj_cells<ORTHO, E, V, F, SV>),      \
              dim3(nl->G), dim3(1024), sh, s,                                                   \
                       nl->d_cell_start, nl->d_nranges, nl->max_ranges, nl->d_rfirst, nl->d_roff, \
                       rows_count, rows, nl->cap, lj->d_pos, lj->d_ptype, lj->d_pgid,             \
                       lj->d_AB_sq, lj->T, c, r, lj->rc2_thr, beta, lj->stage, nl->grid_length,   \
                       nl->Nx, nl->Ny, nl->Nz, reach, lj->d_facc, lj->d_eacc, lj->d_vacc,         \
                       solvent_start)
#define SA_LJ_ROWS_EVF(SV)                                      \
    if (fast) SA_LJ_ROWS(false, false, true, SV);               \
    else if (need_e && need_v) SA_LJ_ROWS(true, true, false, SV); \
    else if (need_e) SA_LJ_ROWS(true, false, false, SV);        \
    else if (need_v) SA_LJ_ROWS(false, true, false, SV);        \
    else SA_LJ_ROWS(false, false, false, SV)
    if (solv) {
        SA_LJ_ROWS_EVF(true);
    } else {
        SA_LJ_ROWS_EVF(false);
    }
#undef SA_LJ_ROWS_EVF
#undef SA_LJ_ROWS
    SA_LAUNCH_CHECK();
    return 0;
}

template <bool ORTHO>
int launch_pairs(sponge_lj* lj, int nrows, const int* cnt, const int* list, int max_nb,
                 LT6 c, LT6 r, float beta, int need_e, int need_v, hipStream_t s) {
    dim3 grid(grid_for(nrows, 4)), block(256);
    ScopedTimer timer("lj_pairs", s);
#define SA_LJ_LAUNCH(E, V)                                                                     \
    hipLaunchKernelGGL((k_lj_pairs<ORTHO, E, V>), grid, block, 0, s, nrows, cnt, list, max_nb, \
                       lj->d_pos, lj->d_ptype, lj->d_AB, c, r, lj->rc2_thr, beta, lj->d_facc,  \
                       lj->d_eacc, lj->d_vacc)
    if (need_e && need_v) SA_LJ_LAUNCH(true, true);
    else if (need_e) SA_LJ_LAUNCH(true, false);
    else if (need_v) SA_LJ_LAUNCH(false, true);
    else SA_LJ_LAUNCH(false, false);
#undef SA_LJ_LAUNCH
    SA_LAUNCH_CHECK();
    return 0;
}

int finalize(sponge_lj* lj, int N, const int* sidx, int need_e, int need_v, float* frc,
             float* atom_energy, float* atom_virial, float* direct, hipStream_t s, int assign = 0,
             int* reset_flag = nullptr, int* reset_flag2 = nullptr) {
    dim3 grid(grid_for(N, 256)), block(256);
#define SA_FIN(E, V)                                                                            \
    SA_LAUNCH(256, (&lj_finalize_body<E, V>), (k_lj_finalize<E, V>), grid, block, 0, s, N, sidx, \
              lj->d_facc,                                                                       \
                       lj->d_eacc, lj->d_vacc, frc, atom_energy, lj->d_LJ_energy_atom, direct, \
                       atom_virial, assign, reset_flag, reset_flag2)
    if (need_e && need_v) SA_FIN(true, true);
    else if (need_e) SA_FIN(true, false);
    else if (need_v) SA_FIN(false, true);
    else SA_FIN(false, false);
#undef SA_FIN
    SA_LAUNCH_CHECK();
    return 0;
}

void lj_free(sponge_lj* lj) {
    dfree(lj->d_AB);
    dfree(lj->d_AB_sq);
    dfree(lj->d_type);
    dfree(lj->d_pos);
    dfree(lj->d_ptype);
    dfree(lj->d_pgid);
    dfree(lj->d_facc);
    dfree(lj->d_eacc);
    dfree(lj->d_vacc);
    dfree(lj->d_LJ_energy_atom);
    dfree(lj->d_sum);
}

}  // namespace

extern "C" int sponge_lj_initial(sponge_lj** out, int atom_numbers, int atom_type_numbers,
                                 const float* LJ_A, const float* LJ_B, const int* atom_LJ_type,
                                 float cutoff, void* stream) {
    if (!out || atom_numbers <= 0 || atom_type_numbers <= 0 || !LJ_A || !LJ_B || !atom_LJ_type ||
        !(cutoff > 0))
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    *out = nullptr;
    // the pair kernel keeps the square (A, B) table of the types in LDS
    if (atom_type_numbers > kMaxLdsTypes) return SPONGE_AMD_ERROR_NOT_IMPLEMENTED;
    const int N = atom_numbers, T = atom_type_numbers, P = T * (T + 1) / 2;
    for (int i = 0; i < N; ++i)
        if (atom_LJ_type[i] < 0 || atom_LJ_type[i] >= T) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    sponge_lj* lj = new sponge_lj();
    lj->N = N;
    lj->T = T;
    lj->cutoff = cutoff;
    lj->rc2_thr = sqrt_threshold(cutoff);
    lj->stream = (hipStream_t)stream;
    lj->h_A.assign(LJ_A, LJ_A + P);
    lj->h_B.assign(LJ_B, LJ_B + P);
    lj->h_type.assign(atom_LJ_type, atom_LJ_type + N);
    // long-range factor (Lennard_Jones_force.cpp:208-235): C6 sum over the
    // type histogram in double, then the float scale
    std::vector<long long> count(T, 0);
    for (int i = 0; i < N; ++i) count[atom_LJ_type[i]]++;
    double c6 = 0.0;
    for (int a = 0; a < T; ++a) {
        if (!count[a]) continue;
        double inner = 0.0;
        for (int b = 0; b < T; ++b) {
            if (!count[b]) continue;
            inner += (double)count[b] * (double)LJ_B[lj_pair_index(a, b)];
        }
        c6 += (double)count[a] * inner;
    }
    float f = (float)c6;
    f *= -2.0f / 3.0f * 3.1415926535897932f / cutoff / cutoff / cutoff / 6.0f;
    lj->long_range_factor = f;
    std::vector<float2> ab(P);
    for (int k = 0; k < P; ++k) ab[k] = make_float2(LJ_A[k], LJ_B[k]);
    std::vector<float2> ab_sq((size_t)T * T);
    for (int a = 0; a < T; ++a)
        for (int b = 0; b < T; ++b) ab_sq[(size_t)a * T + b] = ab[lj_pair_index(a, b)];
    lj->T = T;
    int e = 0;
    e |= dmalloc(&lj->d_AB, P);
    e |= dmalloc(&lj->d_AB_sq, (size_t)T * T);
    e |= dmalloc(&lj->d_type, N);
    lj->stage = stage_entries(sizeof(float4) + 2 * sizeof(int), kPairStage, sizeof(float2) * T * T);
    e |= dmalloc(&lj->d_pos, N);
    e |= dmalloc(&lj->d_ptype, N);
    e |= dmalloc(&lj->d_pgid, N);
    e |= dmalloc(&lj->d_facc, N);
    e |= dmalloc(&lj->d_eacc, N);
    e |= dmalloc(&lj->d_vacc, (size_t)6 * N);
    e |= dmalloc(&lj->d_LJ_energy_atom, N);
    e |= dmalloc(&lj->d_sum, 1);
    if (e) {
        lj_free(lj);
        delete lj;
        return SPONGE_AMD_ERROR_MALLOC_FAILED;
    }
    hipError_t he = hipMemcpyAsync(lj->d_AB, ab.data(), P * sizeof(float2), hipMemcpyHostToDevice,
                                   lj->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(lj->d_AB_sq, ab_sq.data(), ab_sq.size() * sizeof(float2),
                            hipMemcpyHostToDevice, lj->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(lj->d_type, atom_LJ_type, N * sizeof(int), hipMemcpyHostToDevice,
                            lj->stream);
    int fe = 0;
    if (he == hipSuccess) {
        fe |= fill_async(lj->d_LJ_energy_atom, 0.f, N, lj->stream);
        fe |= fill_async(lj->d_eacc, make_float2(0.f, 0.f), N, lj->stream);
        fe |= fill_async(lj->d_vacc, 0.f, (long long)6 * N, lj->stream);
        he = hipStreamSynchronize(lj->stream);
    }
    if (he != hipSuccess || fe) {
        lj_free(lj);
        delete lj;
        return he != hipSuccess ? (int)he : fe;
    }
    *out = lj;
    return 0;
}

extern "C" int sponge_lj_long_range_factor(sponge_lj* lj, float* factor) {
    if (!lj || !factor) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    *factor = lj->long_range_factor;
    return 0;
}

// The MD driver's form: forces assigned (the driver's first force term of a
// step, so its force array needs no clearing) and the neighbor list's rebuild
// flag cleared once the pair kernel has consumed the list.
int sa::lj_direct_force_md(sponge_lj* lj, sponge_nl* nl, const float* crd, const float* charge,
                           float* frc, const float* cell, const float* rcell, float pme_beta,
                           int need_atom_energy, float* atom_energy, float* atom_direct_pme_energy,
                           int* reset_flag, bool* deferred) {
    if (!nl->built) return SPONGE_AMD_ERROR_MISSING_COMMAND;
    if (deferred) *deferred = false;
    hipStream_t s = lj->stream;
    const LT6 c = lt_from(cell), r = lt_from(rcell);
    int rc = nl_pack_for_pairs(nl, crd, charge, lj->d_type, lj->d_pos, lj->d_ptype, lj->d_pgid, s);
    if (rc) return rc;
    // the inner rows (the driver's automatic refresh runs the displacement
    // test that keeps them valid)
    const bool inner = nl->prune_delta > 0.f && nl->opt.refresh_interval <= 0;
    const bool ortho = lt_is_orthorhombic(c);
    if (inner) {
        ScopedTimer timer("lj_prune", s);
        const float keep = lj->cutoff + nl->prune_delta;
        const size_t sh = sizeof(float4) * lj->stage;
        if (ortho)
            SA_LAUNCH(1024, (&lj_prune_cells_body<true>), (k_lj_prune_cells<true>), dim3(nl->G),
                      dim3(1024), sh, s, nl->d_need, nl->d_need_prune, nl->d_cell_start,
                      nl->d_nranges, nl->max_ranges, nl->d_rfirst, nl->d_roff, nl->d_list_count,
                      nl->d_list, nl->cap, lj->d_pos, lj->d_pgid, c, r, keep * keep, lj->stage,
                      nl->d_ilist_count, nl->d_ilist, nl->d_prune_crd);
        else
            SA_LAUNCH(1024, (&lj_prune_cells_body<false>), (k_lj_prune_cells<false>), dim3(nl->G),
                      dim3(1024), sh, s, nl->d_need, nl->d_need_prune, nl->d_cell_start,
                      nl->d_nranges, nl->max_ranges, nl->d_rfirst, nl->d_roff, nl->d_list_count,
                      nl->d_list, nl->cap, lj->d_pos, lj->d_pgid, c, r, keep * keep, lj->stage,
                      nl->d_ilist_count, nl->d_ilist, nl->d_prune_crd);
        SA_LAUNCH_CHECK();
    }
    if (ortho)
        rc = launch_rows<true>(lj, nl, c, r, pme_beta, need_atom_energy, 0, s, true, -1, inner);
    else
        rc = launch_rows<false>(lj, nl, c, r, pme_beta, need_atom_energy, 0, s, true, -1, inner);
    if (rc) return rc;
    // a force-only step may leave the per-row forces for the caller's next
    // pass (the PME gather assigns them with the reciprocal ones)
    if (deferred && !need_atom_energy && !inner) {
        *deferred = true;
        return 0;
    }
    return finalize(lj, lj->N, nl->d_sidx, need_atom_energy, 0, frc, atom_energy, nullptr,
                    atom_direct_pme_energy, s, 1, reset_flag, inner ? nl->d_need_prune : nullptr);
}

extern "C" int sponge_lj_pme_direct_force(sponge_lj* lj, sponge_nl* nl, int atom_numbers,
                                          int local_atom_numbers, int solvent_numbers,
                                          int ghost_numbers, const float* crd,
                                          const float* charge, float* frc, const float* cell,
                                          const float* rcell, float pme_beta,
                                          int need_atom_energy, float* atom_energy,
                                          int need_virial, float* atom_virial,
                                          float* atom_direct_pme_energy) {
    if (!lj || !nl || !crd || !charge || !frc || !cell || !rcell)
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (ghost_numbers != 0) return SPONGE_AMD_ERROR_NOT_IMPLEMENTED;
    if (solvent_numbers < 0 || solvent_numbers > local_atom_numbers)
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (atom_numbers == 0 || local_atom_numbers == 0) return 0;  // :351
    if (local_atom_numbers != lj->N || nl->N != lj->N) return SPONGE_AMD_ERROR_CONFLICTING_COMMAND;
    if (!nl->built) return SPONGE_AMD_ERROR_MISSING_COMMAND;
    if ((need_atom_energy && !atom_energy) || (need_virial && !atom_virial))
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    hipStream_t s = lj->stream;
    const LT6 c = lt_from(cell), r = lt_from(rcell);
    int rc = nl_pack_for_pairs(nl, crd, charge, lj->d_type, lj->d_pos, lj->d_ptype, lj->d_pgid, s);
    if (rc) return rc;
    // rows from local_atom_numbers - solvent_numbers on are the solvent tail:
    // only their pairs with the atoms below it are computed here
    const int solvent_start = solvent_numbers > 0 ? lj->N - solvent_numbers : -1;
    if (lt_is_orthorhombic(c))
        rc = launch_rows<true>(lj, nl, c, r, pme_beta, need_atom_energy, need_virial, s, false,
                               solvent_start);
    else
        rc = launch_rows<false>(lj, nl, c, r, pme_beta, need_atom_energy, need_virial, s, false,
                                solvent_start);
    if (rc) return rc;
    return finalize(lj, lj->N, nl->d_sidx, need_atom_energy, need_virial, frc, atom_energy,
                    atom_virial, atom_direct_pme_energy, s);
}

extern "C" int sponge_lj_pme_direct_force_list(sponge_lj* lj, const int* counts,
                                               const int* serial, int max_numbers,
                                               int atom_numbers, const float* crd,
                                               const float* charge, float* frc,
                                               const float* cell, const float* rcell,
                                               float pme_beta, int need_atom_energy,
                                               float* atom_energy, int need_virial,
                                               float* atom_virial,
                                               float* atom_direct_pme_energy) {
    if (!lj || !counts || !serial || !crd || !charge || !frc || !cell || !rcell || max_numbers <= 0)
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (atom_numbers != lj->N) return SPONGE_AMD_ERROR_CONFLICTING_COMMAND;
    if ((need_atom_energy && !atom_energy) || (need_virial && !atom_virial))
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    hipStream_t s = lj->stream;
    const LT6 c = lt_from(cell), r = lt_from(rcell);
    hipLaunchKernelGGL(k_lj_pack_local, dim3(grid_for(lj->N, 256)), dim3(256), 0, s, lj->N, crd,
                       charge, lj->d_type, lj->d_pos, lj->d_ptype, lj->d_facc);
    SA_LAUNCH_CHECK();
    int rc;
    if (lt_is_orthorhombic(c))
        rc = launch_pairs<true>(lj, lj->N, counts, serial, max_numbers, c, r, pme_beta,
                                need_atom_energy, need_virial, s);
    else
        rc = launch_pairs<false>(lj, lj->N, counts, serial, max_numbers, c, r, pme_beta,
                                 need_atom_energy, need_virial, s);
    if (rc) return rc;
    return finalize(lj, lj->N, nullptr, need_atom_energy, need_virial, frc, atom_energy,
                    atom_virial, atom_direct_pme_energy, s);
}

extern "C" int sponge_lj_long_range_correction(sponge_lj* lj, int need_pressure, float* d_virial,
                                               int need_potential, float* d_potential,
                                               float volume) {
    if (!lj) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (need_pressure) {
        if (!d_virial) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
        hipLaunchKernelGGL(k_add_diag, dim3(1), dim3(1), 0, lj->stream, d_virial,
                           2 * lj->long_range_factor / volume);
        SA_LAUNCH_CHECK();
    }
    if (need_potential) {
        if (!d_potential) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
        hipLaunchKernelGGL(k_add_scalar, dim3(1), dim3(1), 0, lj->stream, d_potential,
                           lj->long_range_factor / volume);
        SA_LAUNCH_CHECK();
        lj->h_LJ_long_energy = lj->long_range_factor / volume;
    }
    return 0;
}

extern "C" int sponge_lj_step_print(sponge_lj* lj, float* LJ_short, float* LJ_long) {
    if (!lj) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    int rc = fill_async(lj->d_sum, 0.0, 1, lj->stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sum_f32, dim3(64), dim3(256), 0, lj->stream, lj->N,
                       lj->d_LJ_energy_atom, lj->d_sum);
    SA_LAUNCH_CHECK();
    double h = 0;
    SA_CHECK(hipMemcpyAsync(&h, lj->d_sum, sizeof(double), hipMemcpyDeviceToHost, lj->stream));
    SA_CHECK(hipStreamSynchronize(lj->stream));
    if (LJ_short) *LJ_short = (float)h;
    if (LJ_long) *LJ_long = lj->h_LJ_long_energy;
    return 0;
}

extern "C" float* sponge_lj_atom_energy(sponge_lj* lj) { return lj ? lj->d_LJ_energy_atom : nullptr; }

extern "C" int sponge_lj_clear(sponge_lj* lj) {
    if (!lj) return 0;
    (void)hipStreamSynchronize(lj->stream);
    lj_free(lj);
    delete lj;
    return 0;
}

// ---------------------------------------------------------------------------
// Allocation-free form of sponge_lj_pme_direct_force_list for callers that
// cannot allocate device memory (code under HIP graph capture, serving loops):
// no module handle, the LJ tables come from the caller's device arrays, and
// the pair kernel's working arrays live in one file-scope device array,
// `scratch`, of a fixed maximum size.  Each GPU holds its own copy for the
// life of the process; its device address is taken once per device with
// hipGetSymbolAddress (the array's name is a host-side stand-in, never a
// device address).  Every call on a GPU shares that GPU's array: calls that
// overlap on one GPU (other streams, other host threads) race.
namespace {

constexpr int kNoAllocMaxAtoms = SPONGE_AMD_NOALLOC_MAX_ATOMS;
constexpr int kNoAllocMaxTypes = SPONGE_AMD_NOALLOC_MAX_TYPES;
constexpr int kNoAllocMaxPairs = kNoAllocMaxTypes * (kNoAllocMaxTypes + 1) / 2;
constexpr int kMaxDevices = 64;

// byte offsets of the working arrays (each 256-byte aligned)
constexpr size_t kNaAlign = 256;
constexpr size_t na_round(size_t b) { return (b + kNaAlign - 1) / kNaAlign * kNaAlign; }
constexpr size_t kNaPos = 0;                                                    // float4[N]
constexpr size_t kNaFacc = kNaPos + na_round(sizeof(float4) * kNoAllocMaxAtoms);  // float4[N]
constexpr size_t kNaEacc = kNaFacc + na_round(sizeof(float4) * kNoAllocMaxAtoms); // float2[N]
constexpr size_t kNaVacc = kNaEacc + na_round(sizeof(float2) * kNoAllocMaxAtoms); // float[6N]
constexpr size_t kNaType = kNaVacc + na_round(sizeof(float) * 6 * kNoAllocMaxAtoms); // int[N]
constexpr size_t kNaAB = kNaType + na_round(sizeof(int) * kNoAllocMaxAtoms);      // float2[P]
constexpr size_t kNaBytes = kNaAB + na_round(sizeof(float2) * kNoAllocMaxPairs);

__device__ __align__(256) unsigned char scratch[kNaBytes];

std::mutex g_na_mutex;
void* g_na_ptr[kMaxDevices] = {};

// The device address of this device's scratch, looked up on the first call
// on the device.
int noalloc_scratch(unsigned char** out) {
    int dev = 0;
    SA_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return SPONGE_AMD_ERROR_NOT_IMPLEMENTED;
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (!p) return SPONGE_AMD_ERROR_MALLOC_FAILED;
        g_na_ptr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_na_ptr[dev]);
    return 0;
}

// Interleaved (A, B) table of the caller's two device arrays.
__global__ void k_lj_interleave_ab(int P, const float* __restrict__ A, const float* __restrict__ B,
                                   float2* __restrict__ AB) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < P) AB[k] = make_float2(A[k], B[k]);
}

}  // namespace

extern "C" int sponge_lj_pme_direct_force_list_noalloc(
    int atom_numbers, int atom_type_numbers, const float* d_LJ_A, const float* d_LJ_B,
    const int* d_atom_LJ_type, float cutoff, const int* counts, const int* serial,
    int max_numbers, const float* crd, const float* charge, float* frc, const float* cell,
    const float* rcell, float pme_beta, int need_atom_energy, float* atom_energy, int need_virial,
    float* atom_virial, float* atom_LJ_energy, float* atom_direct_pme_energy, void* stream) {
    if (!d_LJ_A || !d_LJ_B || !d_atom_LJ_type || !counts || !serial || !crd || !charge || !frc ||
        !cell || !rcell || max_numbers <= 0 || atom_numbers < 0 || atom_type_numbers <= 0 ||
        !(cutoff > 0))
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if ((need_atom_energy && !atom_energy) || (need_virial && !atom_virial))
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    // sizes beyond the fixed scratch
    if (atom_numbers > kNoAllocMaxAtoms || atom_type_numbers > kNoAllocMaxTypes)
        return SPONGE_AMD_ERROR_OVERFLOW;
    if (atom_numbers == 0) return 0;
    unsigned char* base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc) return rc;
    const int N = atom_numbers;
    const int P = atom_type_numbers * (atom_type_numbers + 1) / 2;
    float4* pos = reinterpret_cast<float4*>(base + kNaPos);
    float4* facc = reinterpret_cast<float4*>(base + kNaFacc);
    float2* eacc = reinterpret_cast<float2*>(base + kNaEacc);
    float* vacc = reinterpret_cast<float*>(base + kNaVacc);
    int* ptype = reinterpret_cast<int*>(base + kNaType);
    float2* AB = reinterpret_cast<float2*>(base + kNaAB);
    hipStream_t s = (hipStream_t)stream;
    const LT6 c = lt_from(cell), r = lt_from(rcell);
    const float rc2_thr = sqrt_threshold(cutoff);
    hipLaunchKernelGGL(k_lj_interleave_ab, dim3(grid_for(P, 256)), dim3(256), 0, s, P, d_LJ_A,
                       d_LJ_B, AB);
    SA_LAUNCH_CHECK();
    // positions, types and cleared force accumulators (every row written)
    hipLaunchKernelGGL(k_lj_pack_local, dim3(grid_for(N, 256)), dim3(256), 0, s, N, crd, charge,
                       d_atom_LJ_type, pos, ptype, facc);
    SA_LAUNCH_CHECK();
    const bool ortho = lt_is_orthorhombic(c);
    dim3 grid(grid_for(N, 4)), block(256);
#define SA_NA_PAIRS(O, E, V)                                                                  \
    hipLaunchKernelGGL((k_lj_pairs<O, E, V>), grid, block, 0, s, N, counts, serial, max_numbers, \
                       pos, ptype, AB, c, r, rc2_thr, pme_beta, facc, eacc, vacc)
#define SA_NA_PAIRS_EV(O)                                       \
    if (need_atom_energy && need_virial) SA_NA_PAIRS(O, true, true); \
    else if (need_atom_energy) SA_NA_PAIRS(O, true, false);     \
    else if (need_virial) SA_NA_PAIRS(O, false, true);          \
    else SA_NA_PAIRS(O, false, false)
    if (ortho) {
        SA_NA_PAIRS_EV(true);
    } else {
        SA_NA_PAIRS_EV(false);
    }
#undef SA_NA_PAIRS_EV
#undef SA_NA_PAIRS
    SA_LAUNCH_CHECK();
    dim3 fgrid(grid_for(N, 256));
#define SA_NA_FIN(E, V)                                                                          \
    hipLaunchKernelGGL((k_lj_finalize<E, V>), fgrid, block, 0, s, N, nullptr, facc, eacc, vacc, \
                       frc, atom_energy, atom_LJ_energy, atom_direct_pme_energy, atom_virial, 0,  \
                       nullptr, nullptr)
    if (need_atom_energy && need_virial) SA_NA_FIN(true, true);
    else if (need_atom_energy) SA_NA_FIN(true, false);
    else if (need_virial) SA_NA_FIN(false, true);
    else SA_NA_FIN(false, false);
#undef SA_NA_FIN
    SA_LAUNCH_CHECK();
    return 0;
}
