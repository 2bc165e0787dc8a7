// solvent_LJ: the water-tail specialisation of the LJ + PME direct pair
// forces (SPONGE/Lennard_Jones_force/solvent_LJ.cpp:24-147, 149-248,
// 256-432).
//
// The reference evaluates, for every water residue w of the solvent tail
// (atoms from solvent_start on, water_points atoms per residue, oxygen
// first), the pairs (a, j) of every atom a of w with every j of the OXYGEN's
// half list nl[O_w] (j > O_w): the hydrogens' own lists are not used.  Pair
// energies and virials of a water go to its oxygen's entries; the LJ module,
// called with solvent_numbers, evaluates the rows below the tail (and with
// them every pair that has a non-solvent member).
//
// The same pair set on the full-row list of this library (rows hold both
// directions, in the list's cell-sorted order): a solvent atom b of water w
// takes
//   forward : in O_w's row, every partner j of a water v > w -- the pairs
//             (a, j) for all a of w, forces on w's atoms (kept in a per-water
//             accumulator written by O_w's wavefront), energies and virial on
//             O_w;
//   reverse : in b's own row, every partner that is the oxygen O_v of a water
//             v < w (b is in nl[O_v]) -- the pairs (a, b) for all a of v,
//             forces on b.
// Pairs with a non-solvent member are the LJ module's.  No atomics: every
// pair is evaluated from both of its sides.  On a list fresh from a rebuild
// every pair within the cutoff of two waters is in the set (an atom within
// the cutoff of a hydrogen is within cutoff + |OH| < cutoff + skin of its
// oxygen); between rebuilds the set is the reference's, not the plain one.
#include <vector>

#include "internal.h"
#include "pair_term.h"
#include "timing.h"

using namespace sa;

struct sponge_solvent_lj {
    sponge_lj* lj = nullptr;
    int N = 0;
    int enabled = 0;
    int water_points = 0;
    int solvent_numbers = 0;   // atoms in the solvent tail
    int solvent_start = 0;     // first residue of the tail
    int solvent_atom = 0;      // first atom of the tail
    hipStream_t stream = nullptr;
    float4* d_facc = nullptr;  // [N] reverse forces per sorted row
    float4* d_wacc = nullptr;  // [solvent_numbers] forward forces per solvent atom
    float2* d_eacc = nullptr;  // [N] (E_LJ, E_coulomb) per sorted row
    float* d_vacc = nullptr;   // [N][6] virial per sorted row
};

namespace {

// Partner position and type of a water atom given by its global id (another
// atom of the residue of a staged oxygen): from the sorted arrays in HBM.
__device__ __forceinline__ void water_atom(int g, const int* __restrict__ rank,
                                           const float4* __restrict__ pos,
                                           const int* __restrict__ ptype, float4& x, int& t) {
    const int q = rank[g];
    x = pos[q];
    t = ptype[q];
}

// One solvent row (wavefront-wide), one partner entry per lane and pass.
template <bool ORTHO, bool NEED_E, bool NEED_V, bool FAST, bool STAGED, int WP>
__device__ __forceinline__ void solvent_row(
    int p, int lane, int n, const unsigned short* __restrict__ row, PairSrc src, int nr,
    const int* r_first, const int* r_off, const float4* __restrict__ pos,
    const int* __restrict__ ptype, const int* __restrict__ rank, const float2* cab, int T, int gi,
    int S, const LT6& c, const LT6& r, float rc2_thr, float beta, float4* __restrict__ facc,
    float4* __restrict__ wacc, float2* __restrict__ eacc, float* __restrict__ vacc) {
    const int wb = (gi - S) / WP;
    const bool is_o = gi - S == wb * WP;
    const float4 pb4 = pos[p];
    const float3 pb = make_float3(pb4.x, pb4.y, pb4.z);
    const int tb = ptype[p];
    // the row water's own atoms (forward pass of an oxygen row)
    float4 wpos[WP];
    int wtyp[WP];
    if (is_o) {
#pragma unroll
        for (int m = 0; m < WP; ++m) water_atom(gi + m, rank, pos, ptype, wpos[m], wtyp[m]);
    }
    float3 fb = make_float3(0.f, 0.f, 0.f);
    float3 fw[WP];
#pragma unroll
    for (int m = 0; m < WP; ++m) fw[m] = make_float3(0.f, 0.f, 0.f);
    double elj = 0, ec = 0;  // energy and virial terms summed in FP64 per lane
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    for (int k = lane; k < n; k += kWave) {
        const int q = src_index<STAGED>((int)row[k], nr, r_first, r_off);
        const int gj = src.gid[q];
        if (gj < S) continue;  // a pair with a non-solvent member
        const int wj = (gj - S) / WP;
        const bool oj = gj - S == wj * WP;
        if (oj && wj < wb) {
            // reverse: b is in nl[O_v]; the pairs (a, b) of every atom a of v
#pragma unroll
            for (int m = 0; m < WP; ++m) {
                float4 xa;
                int ta;
                if (m == 0) {
                    xa = src.pos[q];
                    ta = src.typ[q];
                } else {
                    water_atom(gj + m, rank, pos, ptype, xa, ta);
                }
                const float3 d = disp_pbc<ORTHO>(make_float3(xa.x, xa.y, xa.z), pb, c, r);
                const PairTerm<false, false, FAST> t(d, norm2_exact(d), pb4.w, xa.w,
                                                     cab[tb * T + ta], rc2_thr, beta);
                fb.x += t.lx;
                fb.y += t.ly;
                fb.z += t.lz;
            }
        } else if (is_o && wj > wb) {
            // forward: j is in nl[O_w]; the pairs (a, j) of every atom a of w
            const float4 xj = src.pos[q];
            const int tj = src.typ[q];
            const float3 pj = make_float3(xj.x, xj.y, xj.z);
#pragma unroll
            for (int m = 0; m < WP; ++m) {
                const float3 d = disp_pbc<ORTHO>(pj, make_float3(wpos[m].x, wpos[m].y, wpos[m].z),
                                                 c, r);
                const PairTerm<NEED_E, NEED_V, FAST> t(d, norm2_exact(d), wpos[m].w, xj.w,
                                                       cab[wtyp[m] * T + tj], rc2_thr, beta);
                fw[m].x += t.lx;
                fw[m].y += t.ly;
                fw[m].z += t.lz;
                if (NEED_E) {
                    elj += t.elj;
                    ec += t.ec;
                }
                if (NEED_V) {
                    v0 -= t.lx * d.x;
                    v1 -= t.lx * d.y + t.ly * d.x;
                    v2 -= t.ly * d.y;
                    v3 -= t.lx * d.z + t.lz * d.x;
                    v4 -= t.ly * d.z + t.lz * d.y;
                    v5 -= t.lz * d.z;
                }
            }
        }
    }
    fb.x = wave_sum(fb.x);
    fb.y = wave_sum(fb.y);
    fb.z = wave_sum(fb.z);
    if (is_o) {
#pragma unroll
        for (int m = 0; m < WP; ++m) {
            fw[m].x = wave_sum(fw[m].x);
            fw[m].y = wave_sum(fw[m].y);
            fw[m].z = wave_sum(fw[m].z);
        }
    }
    double delj = 0, dec = 0, dv0 = 0, dv1 = 0, dv2 = 0, dv3 = 0, dv4 = 0, dv5 = 0;
    if (NEED_E && is_o) {
        delj = wave_sum((double)elj);
        dec = wave_sum((double)ec);
    }
    if (NEED_V && is_o) {
        dv0 = wave_sum((double)v0);
        dv1 = wave_sum((double)v1);
        dv2 = wave_sum((double)v2);
        dv3 = wave_sum((double)v3);
        dv4 = wave_sum((double)v4);
        dv5 = wave_sum((double)v5);
    }
    if (lane == 0) {
        facc[p] = make_float4(fb.x, fb.y, fb.z, 0.f);
        if (is_o) {
#pragma unroll
            for (int m = 0; m < WP; ++m)
                wacc[gi - S + m] = make_float4(fw[m].x, fw[m].y, fw[m].z, 0.f);
        }
        if (NEED_E) eacc[p] = make_float2((float)delj, (float)dec);
        if (NEED_V) {
            float* v = vacc + (size_t)6 * p;
            v[0] = (float)dv0; v[1] = (float)dv1; v[2] = (float)dv2;
            v[3] = (float)dv3; v[4] = (float)dv4; v[5] = (float)dv5;
        }
    }
}

// One workgroup (16 wavefronts) per neighbor-list cell: the cell's candidate
// stream (positions with charges, LJ types, global ids) and the square (A, B)
// table staged in LDS as in the plain pair kernel; one wavefront per solvent
// row of the cell.
template <bool ORTHO, bool NEED_E, bool NEED_V, bool FAST, int WP>
__global__ void __launch_bounds__(1024) k_solvent_cells(
    const int* __restrict__ start, const int* __restrict__ nranges, int max_ranges,
    const int* __restrict__ g_rfirst, const int* __restrict__ g_roff,
    const int* __restrict__ list_count, const unsigned short* __restrict__ list, int cap,
    const float4* __restrict__ pos, const int* __restrict__ ptype, const int* __restrict__ pgid,
    const int* __restrict__ rank, const float2* __restrict__ AB_sq, int T, LT6 c, LT6 r,
    float rc2_thr, float beta, int stage, int S, float4* __restrict__ facc,
    float4* __restrict__ wacc, float2* __restrict__ eacc, float* __restrict__ vacc) {
    extern __shared__ float4 cpos[];
    int* ctyp = reinterpret_cast<int*>(cpos + stage);
    int* cgid = ctyp + stage;
    float2* cab = reinterpret_cast<float2*>(cgid + stage);
    __shared__ int r_first[kMaxRanges];
    __shared__ int r_off[kMaxRanges + 1];
    const int cell = blockIdx.x;
    const int a0 = start[cell], a1 = start[cell + 1];
    if (a0 == a1) return;
    const int nr = nranges[cell];
    for (int k = threadIdx.x; k < nr; k += blockDim.x) r_first[k] = g_rfirst[(size_t)cell * max_ranges + k];
    for (int k = threadIdx.x; k <= nr; k += blockDim.x) r_off[k] = g_roff[(size_t)cell * (max_ranges + 1) + k];
    for (int k = threadIdx.x; k < T * T; k += blockDim.x) cab[k] = AB_sq[k];
    __syncthreads();
    const int total = r_off[nr];
    const bool staged = total <= stage;
    if (staged)
        stage_stream(nr, r_first, r_off, total, [&](int t, int q) {
            cpos[t] = pos[q];
            ctyp[t] = ptype[q];
            cgid[t] = pgid[q];
        });
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave, nwave = blockDim.x / kWave;
    for (int p = a0 + wave; p < a1; p += nwave) {
        const int gi = pgid[p];
        if (gi < S) continue;  // a row of the LJ module
        const int n = list_count[p];
        const unsigned short* row = list + (size_t)p * cap;
        if (staged)
            solvent_row<ORTHO, NEED_E, NEED_V, FAST, true, WP>(
                p, lane, n, row, PairSrc{cpos, ctyp, cgid}, nr, r_first, r_off, pos, ptype, rank,
                cab, T, gi, S, c, r, rc2_thr, beta, facc, wacc, eacc, vacc);
        else
            solvent_row<ORTHO, NEED_E, NEED_V, FAST, false, WP>(
                p, lane, n, row, PairSrc{pos, ptype, pgid}, nr, r_first, r_off, pos, ptype, rank,
                cab, T, gi, S, c, r, rc2_thr, beta, facc, wacc, eacc, vacc);
    }
}

// Adds the solvent rows to the caller's arrays (the reference's Warp_Sum_To
// and atomicAdd updates, solvent_LJ.cpp:124-145): frc += reverse + forward
// forces; atom_energy, the LJ energy array and the direct-PME energy array +=
// (oxygens carry their water's pair energies); atom_virial +=.
template <bool NEED_E, bool NEED_V>
__global__ void k_solvent_finalize(int N, int S, const int* __restrict__ rank,
                                   const float4* __restrict__ facc, const float4* __restrict__ wacc,
                                   const float2* __restrict__ eacc, const float* __restrict__ vacc,
                                   float* __restrict__ frc, float* __restrict__ atom_energy,
                                   float* __restrict__ lj_energy, float* __restrict__ direct_energy,
                                   float* __restrict__ atom_virial) {
    const int i = S + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int p = rank[i];
    const float4 a = facc[p], b = wacc[i - S];
    frc[3 * i] += a.x + b.x;
    frc[3 * i + 1] += a.y + b.y;
    frc[3 * i + 2] += a.z + b.z;
    if (NEED_E) {
        const float2 e = eacc[p];
        atom_energy[i] += e.x + e.y;
        if (lj_energy) lj_energy[i] += e.x;
        if (direct_energy) direct_energy[i] += e.y;
    }
    if (NEED_V) {
        const float* v = vacc + (size_t)6 * p;
        float* o = atom_virial + (size_t)6 * i;
        o[0] += v[0]; o[1] += v[1]; o[2] += v[2]; o[3] += v[3]; o[4] += v[4]; o[5] += v[5];
    }
}

void solvent_free(sponge_solvent_lj* h) {
    dfree(h->d_facc);
    dfree(h->d_wacc);
    dfree(h->d_eacc);
    dfree(h->d_vacc);
}

template <bool ORTHO, int WP>
int launch_solvent(sponge_solvent_lj* h, sponge_nl* nl, LT6 c, LT6 r, float beta, int need_e,
                   int need_v, hipStream_t s) {
    sponge_lj* lj = h->lj;
    ScopedTimer timer("solvent_lj_pairs", s);
    const size_t sh = (sizeof(float4) + 2 * sizeof(int)) * lj->stage + sizeof(float2) * lj->T * lj->T;
    const bool fast = beta * lj->cutoff <= kErfcxMax && !need_e && !need_v;
#define SA_SOLV(E, V, F)                                                                          \
    hipLaunchKernelGGL((k_solvent_cells<ORTHO, E, V, F, WP>), dim3(nl->G), dim3(1024), sh, s,     \
                       nl->d_cell_start, nl->d_nranges, nl->max_ranges, nl->d_rfirst, nl->d_roff, \
                       nl->d_list_count, nl->d_list, nl->cap, lj->d_pos, lj->d_ptype, lj->d_pgid, \
                       nl->d_rank, lj->d_AB_sq, lj->T, c, r, lj->rc2_thr, beta, lj->stage,        \
                       h->solvent_atom, h->d_facc, h->d_wacc, h->d_eacc, h->d_vacc)
    if (fast) SA_SOLV(false, false, true);
    else if (need_e && need_v) SA_SOLV(true, true, false);
    else if (need_e) SA_SOLV(true, false, false);
    else if (need_v) SA_SOLV(false, true, false);
    else SA_SOLV(false, false, false);
#undef SA_SOLV
    SA_LAUNCH_CHECK();
    return 0;
}

}  // namespace

extern "C" int sponge_solvent_lj_initial(sponge_solvent_lj** out, sponge_lj* lj, int atom_numbers,
                                         int residue_numbers, const int* res_start,
                                         const float* mass, int solvent_LJ, int default_enable,
                                         void* stream) {
    if (!out || !lj || atom_numbers <= 0 || residue_numbers <= 0 || !res_start || !mass)
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (atom_numbers != lj->N) return SPONGE_AMD_ERROR_CONFLICTING_COMMAND;
    *out = nullptr;
    const int N = atom_numbers, R = residue_numbers;
    for (int k = 0; k < R; ++k) {
        const int end = k + 1 < R ? res_start[k + 1] : N;
        if (res_start[k] < 0 || end <= res_start[k] || end > N)
            return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    }
    sponge_solvent_lj* h = new sponge_solvent_lj();
    h->lj = lj;
    h->N = N;
    h->stream = (hipStream_t)stream;
    // Initial's gating (solvent_LJ.cpp:165-186): without the mdin key, on
    // by default (NVT and NPT) when there are at least 10 groups; the water
    // size is that of the last group and must be 3 or 4
    bool enable = solvent_LJ < 0 ? (default_enable && R >= 10) : solvent_LJ != 0;
    const int last_len = N - res_start[R - 1];
    if (enable) h->water_points = last_len;
    if (h->water_points != 3 && h->water_points != 4) enable = false;
    if (enable) {
        // the trailing residues of water_points atoms with water masses
        // (solvent_LJ.cpp:195-217)
        int count = 0, first = R;
        for (int k = R - 1; k >= 0; --k) {
            const int end = k + 1 < R ? res_start[k + 1] : N;
            const int a = res_start[k];
            const bool water = end - a == h->water_points && mass[a] > 15.9f && mass[a] < 16.1f &&
                               mass[a + 1] > 1.007f && mass[a + 1] < 1.009f &&
                               mass[a + 2] > 1.007f && mass[a + 2] < 1.009f;
            if (!water) break;
            count += end - a;
            first = k;
        }
        h->solvent_numbers = count;
        h->solvent_start = first;
        h->solvent_atom = N - count;
        h->enabled = count > 0;
    }
    if (h->enabled) {
        int e = 0;
        e |= dmalloc(&h->d_facc, N);
        e |= dmalloc(&h->d_wacc, h->solvent_numbers);
        e |= dmalloc(&h->d_eacc, N);
        e |= dmalloc(&h->d_vacc, (size_t)6 * N);
        if (e) {
            solvent_free(h);
            delete h;
            return SPONGE_AMD_ERROR_MALLOC_FAILED;
        }
    }
    *out = h;
    return 0;
}

extern "C" int sponge_solvent_lj_info(sponge_solvent_lj* h, int* enabled, int* solvent_numbers,
                                      int* solvent_start_residue, int* water_points) {
    if (!h) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (enabled) *enabled = h->enabled;
    if (solvent_numbers) *solvent_numbers = h->solvent_numbers;
    if (solvent_start_residue) *solvent_start_residue = h->solvent_start;
    if (water_points) *water_points = h->water_points;
    return 0;
}

extern "C" int sponge_solvent_lj_direct_force(sponge_solvent_lj* h, sponge_nl* nl,
                                              int atom_numbers, int residue_numbers,
                                              const float* crd, const float* charge, float* frc,
                                              const float* cell, const float* rcell,
                                              float pme_beta, int need_atom_energy,
                                              float* atom_energy, int need_virial,
                                              float* atom_virial, float* atom_direct_pme_energy) {
    if (!h || !nl || !crd || !charge || !frc || !cell || !rcell)
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    (void)residue_numbers;
    if (!h->enabled) return 0;  // is_initialized == 0: nothing to do (solvent_LJ.cpp:264)
    if (atom_numbers != h->N || nl->N != h->N) return SPONGE_AMD_ERROR_CONFLICTING_COMMAND;
    if (!nl->built) return SPONGE_AMD_ERROR_MISSING_COMMAND;
    if ((need_atom_energy && !atom_energy) || (need_virial && !atom_virial))
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    sponge_lj* lj = h->lj;
    hipStream_t s = h->stream;
    const LT6 c = lt_from(cell), r = lt_from(rcell);
    // the LJ module's packed sorted positions (the reference reads the LJ
    // module's crd_with_LJ_parameters_local), refreshed for these coordinates
    int rc = nl_pack_for_pairs(nl, crd, charge, lj->d_type, lj->d_pos, lj->d_ptype, lj->d_pgid, s);
    if (rc) return rc;
    const bool ortho = lt_is_orthorhombic(c);
    if (h->water_points == 3)
        rc = ortho ? launch_solvent<true, 3>(h, nl, c, r, pme_beta, need_atom_energy, need_virial, s)
                   : launch_solvent<false, 3>(h, nl, c, r, pme_beta, need_atom_energy, need_virial, s);
    else
        rc = ortho ? launch_solvent<true, 4>(h, nl, c, r, pme_beta, need_atom_energy, need_virial, s)
                   : launch_solvent<false, 4>(h, nl, c, r, pme_beta, need_atom_energy, need_virial, s);
    if (rc) return rc;
    const int n = h->solvent_numbers;
    dim3 grid(grid_for(n, 256)), block(256);
    float* lje = need_atom_energy ? lj->d_LJ_energy_atom : nullptr;
#define SA_SFIN(E, V)                                                                            \
    hipLaunchKernelGGL((k_solvent_finalize<E, V>), grid, block, 0, s, h->N, h->solvent_atom,    \
                       nl->d_rank, h->d_facc, h->d_wacc, h->d_eacc, h->d_vacc, frc, atom_energy, \
                       lje, atom_direct_pme_energy, atom_virial)
    if (need_atom_energy && need_virial) SA_SFIN(true, true);
    else if (need_atom_energy) SA_SFIN(true, false);
    else if (need_virial) SA_SFIN(false, true);
    else SA_SFIN(false, false);
#undef SA_SFIN
    SA_LAUNCH_CHECK();
    return 0;
}

extern "C" int sponge_solvent_lj_clear(sponge_solvent_lj* h) {
    if (!h) return 0;
    solvent_free(h);
    delete h;
    return 0;
}
