// Module state behind the opaque handles of include/sponge_amd.h.
#pragma once

#include <vector>

#include "common.h"

// ATOM_GROUP of the reference (SPONGE/common.h:177-182).
struct sa_atom_group {
    int atom_numbers;
    int ghost_numbers;
    int* atom_serial;
};

// Neighbor list.  Atoms are counting-sorted by cell (x-major, z fastest) on
// every rebuild; the list holds, per sorted atom p, the sorted indices q of
// all its neighbours (both directions) as 16-bit positions in the candidate
// stream of p's cell (the concatenated runs of its stencil cells); the
// reference's half list is the subset with global id of q > global id of p.
struct sponge_nl {
    int N = 0;
    float cutoff = 0, skin = 0, grid_length = 0, thr = 0;  // thr = gl*gl*4 (cutoff_skin_square)
    int Nx = 0, Ny = 0, Nz = 0, G = 0;
    int ortho = 1;
    sa::LT6 cell0{}, rcell0{};  // box the cell stencil was built for
    sponge_nl_options opt{};
    hipStream_t stream = nullptr;

    int* d_need = nullptr;        // rebuild flag of this step
    int* d_flags = nullptr;       // sticky overflow flags
    unsigned long long* d_rebuilds = nullptr;  // number of list builds
    int* d_cell_count = nullptr;  // [G]
    int* d_cell_start = nullptr;  // [G+1]
    int* d_atom_cell = nullptr;   // [N]
    int* d_atom_slot = nullptr;   // [N]
    int* d_sidx = nullptr;        // sorted -> local
    int* d_rank = nullptr;        // local -> sorted
    float4* d_spos = nullptr;     // sorted x, y, z at build time, w = global id bits
    float* d_old_crd = nullptr;   // VECTOR[N] at the last build (auto refresh)
    int2* d_ranges = nullptr;     // [G][max_ranges] runs of neighbour cells [begin, end)
    int* d_nranges = nullptr;     // [G]
    int max_ranges = 0;
    int cap = 0;                  // row capacity (both directions)
    int* d_list_count = nullptr;  // [N] per sorted atom
    unsigned short* d_list = nullptr;  // [N][cap] positions in the row cell's candidate stream
    int2* d_sexw = nullptr;       // [N] combined exclusion window of each sorted atom
    int2* d_gexw = nullptr;       // [N] combined exclusion window per global id
    const int* exw_src[3] = {nullptr, nullptr, nullptr};  // lists the windows were built from
    int* d_rfirst = nullptr;      // [G][max_ranges] first sorted index of each stream run
    int* d_roff = nullptr;        // [G][max_ranges+1] stream offset of each run
    int* d_exp_serial = nullptr;  // reference-format export storage (lazy)
    int* d_exp_count = nullptr;
    sa_atom_group* d_atom_group = nullptr;
    int built = 0;
    // inner (pruned) rows used by the MD driver's pair kernel: the entries of
    // the rows within cutoff + prune_delta at the last prune, pruned again once
    // an atom moved more than prune_delta / 2 since then or the list was rebuilt
    float prune_delta = 0.f;          // 0: the pair kernel reads the full rows
    int* d_need_prune = nullptr;      // prune flag of this step
    float* d_prune_crd = nullptr;     // VECTOR[N] at the last prune
    int* d_ilist_count = nullptr;     // [N] per sorted atom
    unsigned short* d_ilist = nullptr;  // [N][cap]
};

struct sponge_lj {
    int N = 0, T = 0;
    float cutoff = 10.0f;
    float rc2_thr = 0;            // r2 < rc2_thr  <=>  sqrtf(r2) < cutoff
    float long_range_factor = 0;
    float h_LJ_long_energy = 0;
    hipStream_t stream = nullptr;
    std::vector<float> h_A, h_B;
    std::vector<int> h_type;
    float2* d_AB = nullptr;       // interleaved (A, B) per pair type
    float2* d_AB_sq = nullptr;    // [T][T] square copy of the (A, B) table
    int* d_type = nullptr;        // [N] local order
    int stage = 0;                // candidate-stream entries staged in LDS by the pair kernel
    float4* d_pos = nullptr;      // [N] packed x, y, z, charge (list index space)
    int* d_ptype = nullptr;       // [N] LJ type (list index space)
    int* d_pgid = nullptr;        // [N] global id (list index space)
    float4* d_facc = nullptr;     // [N] force accumulator (list index space)
    float2* d_eacc = nullptr;     // [N] (E_LJ, E_coulomb) per row
    float* d_vacc = nullptr;      // [N][6] virial per row
    float* d_LJ_energy_atom = nullptr;  // [N] local order
    double* d_sum = nullptr;      // scratch for Step_Print reductions
};

struct sa_radix_plan {
    int n;
    int nst;
    int r[16];
};

struct sponge_pme {
    int N = 0;
    int fx = 0, fy = 0, fz = 0, fzc = 0;
    long long nall = 0, nfft = 0;
    float beta = 0, cutoff = 0, volume = 0, neutralizing_factor = 0;
    sa::LT6 cell{}, rcell{};
    sponge_pme_options opt{};
    hipStream_t stream = nullptr;
    sa_radix_plan plan[3]{};
    float2* d_tw[3] = {nullptr, nullptr, nullptr};  // exp(-2 pi i t / n), t < n
    float* d_Q = nullptr;         // [fx][fy][fz]
    float2* d_FQ = nullptr;       // [fzc][fx][fy]
    float* d_FBCFQ = nullptr;     // [fx][fy][fz]
    float* d_BCt = nullptr;       // [fzc][fy][fx]
    float* d_vBCt = nullptr;      // 6 planes of [fzc][fy][fx]
    double* d_acc = nullptr;      // [0] sum w BC |F|^2, [1..6] virial, [7] sum q^2, [8] sum q
    float* d_ene = nullptr;       // [0] reciprocal, [1] self (incl. neutralizing)
    float* d_direct_atom_energy = nullptr;
    float* d_correction_atom_energy = nullptr;
    double* d_sum = nullptr;
    int tl_z = 2, tl_y = 2, tl_x = 2;  // lines per block of the FFT passes
    // spread bricks: the grid cut into Bx x By x Bz bricks (the last brick of an
    // axis takes the remainder); atoms binned by the brick of their base
    // point, cap per brick, the rest in an overflow list
    int bx = 0, by = 0, bz = 0, nbx = 0, nby = 0, nbz = 0, nb = 0, bcap = 0;
    int* d_bcount = nullptr;      // [nb + 1]: atoms per brick, [nb] overflow count (zero between evaluations)
    float4* d_bw = nullptr;       // [nb][bcap] fractions x, y, z and charge
    int* d_bu = nullptr;          // [nb][bcap] packed base point ux | uy << 10 | uz << 20
    float4* d_ow = nullptr;       // [N] overflow entries
    int* d_ou = nullptr;
    int* d_bi = nullptr;          // [nb][bcap] atom index of each binned atom
    int* d_oi = nullptr;          // [N] atom index of each overflow entry
    int* d_gcount = nullptr;      // [nb + 1] the counts as the gather reads them
    float* d_tiles = nullptr;     // [nb][tile points] halo tiles of the spread
    int spread_waves = 1;         // wavefronts (tile copies) per spread workgroup
};

namespace sa {
// runs per cell stream (the stencil of a cell has at most 192 cells)
constexpr int kMaxRanges = 256;

// sorted index of position t of a cell's candidate stream, from the cell's
// run tables (first sorted index and stream offset of each run)
// (the last run k with r_off[k] <= t, by bisection: that run is non-empty)
__device__ __forceinline__ int stream_to_sorted(int t, int nr, const int* r_first, const int* r_off) {
    int lo = 0, hi = nr - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (r_off[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return r_first[lo] + (t - r_off[lo]);
}

// Stage the first `staged` positions of a cell's candidate stream into LDS:
// one wavefront per run (a run is contiguous in sorted order, so the loads
// coalesce); copy(t, q) moves sorted entry q to stream position t.
template <typename Copy>
__device__ __forceinline__ void stage_stream(int nr, const int* r_first, const int* r_off, int staged,
                                             Copy copy) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
    for (int k = wave; k < nr; k += nwave) {
        const int off = r_off[k];
        if (off >= staged) continue;
        const int len = min(r_off[k + 1], staged) - off, f = r_first[k];
        for (int u = lane; u < len; u += 64) copy(off + u, f + u);
    }
}

// Candidate-stream entries a kernel stages in LDS, given the bytes per entry:
// as many as the device's per-workgroup LDS holds beside 4 KiB of tables and
// `reserve` further bytes, at most `want`, a multiple of 64.
int stage_entries(size_t bytes_per_entry, int want, size_t reserve = 0);

// MD-driver forms of two module calls (see their definitions)
int lj_direct_force_md(sponge_lj* lj, sponge_nl* nl, const float* crd, const float* charge,
                       float* frc, const float* cell, const float* rcell, float pme_beta,
                       int need_atom_energy, float* atom_energy, float* atom_direct_pme_energy,
                       int* reset_flag, bool* deferred = nullptr);
int nl_update_impl(sponge_nl* nl, const int* atom_local, int local_atom_numbers, int ghost_numbers,
                   const float* crd, const float* cell, const float* rcell, int step, int update,
                   const int* excluded_list_start, const int* excluded_list,
                   const int* excluded_numbers, bool need_cleared);

// surrogate forms driven by the MD step (sorted positions + charges given)
int rsur_features_sorted(sponge_rsur* h, sponge_nl* nl, const float4* pos, const float* crd,
                         const float* charge, const float* cell, const float* rcell,
                         const int* ex_start, const int* ex_list, const int* ex_num);
int rsur_predict_sorted(sponge_rsur* h, float* frc, int assign, const float* anchor);
void rsur_set_stream(sponge_rsur* h, hipStream_t s);
// the reciprocal force in two halves (spread + FFT convolution; gather +
// energy) on the module's stream, and that stream
int pme_recip_begin(sponge_pme* p, const float* crd, const float* cell, const float* rcell,
                    const float* charge, int need_virial, int need_energy, int step);
// extra (nullable, [N][3]): forces the gather adds with the reciprocal ones.
// base (nullable): the gather assigns frc[i] = base[rank[i]].xyz + reciprocal
// + extra instead of adding to frc (base: the pair kernel's per-row forces
// in sorted order, rank: the list's atom -> sorted position map), and clears
// *reset (nullable) as the pair path's finalize pass would
struct GatherBase {
    const float4* base = nullptr;
    const int* rank = nullptr;
    int* reset = nullptr;
};
int pme_recip_end(sponge_pme* p, const float* rcell, const float* charge, float* frc,
                  int need_virial, int need_energy, float* d_virial, float* d_potential, int step,
                  const float* extra = nullptr, GatherBase gb = GatherBase{});
hipStream_t pme_stream(const sponge_pme* p);
bool pme_calculates_reciprocal(const sponge_pme* p);
bool pme_updates_on(const sponge_pme* p, int step);  // the reciprocal part is evaluated this step
void pme_set_stream(sponge_pme* p, hipStream_t s);
// wavefronts (tile copies) per workgroup of the tile spread, and a smaller
// count for launches that must share one workgroup size (the batched driver)
int pme_spread_waves(const sponge_pme* p);
void pme_set_spread_waves(sponge_pme* p, int waves);
int rsur_atoms(const sponge_rsur* h);
// soft-core module driven by the MD step (step-driver form of the direct
// force: frc assigned, the list's rebuild flag cleared)
int ljsc_direct_force_md(sponge_ljsc* h, sponge_nl* nl, const float* crd, const float* charge,
                         float* frc, const float* cell, const float* rcell, float pme_beta,
                         int need_atom_energy, float* atom_energy, float* atom_direct_pme_energy,
                         int* reset_flag);
void ljsc_set_stream(sponge_ljsc* h, hipStream_t s);
int ljsc_atoms(const sponge_ljsc* h);
float ljsc_long_range_factor(const sponge_ljsc* h);
float rsur_cutoff(const sponge_rsur* h);

// MD-driver hooks of the batched driver
hipStream_t md_stream(const sponge_md* m);
bool md_graphable(const sponge_md* m);
int md_record_steps(sponge_md* m, int k);
void md_advance(sponge_md* m, int k);
// the driver's overflow check over steps first..last (see sponge_md_run)
int md_overflow_check(sponge_md* m, long long first, long long last);
long long md_step_count(const sponge_md* m);
int md_generation(const sponge_md* m);  // changes when the list is re-initialised
int md_state_arrays(sponge_md* m, float** crd, float** charge);

// inner-row pruning of the list (delta > 0 on, 0 off): allocates the inner
// rows on first use and marks them for a prune on the next step
int nl_set_prune(sponge_nl* nl, float delta);

// the list's sticky overflow flags, read (synchronising) and cleared
int nl_take_flags(sponge_nl* nl, int* flags);
// Check_Overflow's recovery: *nl replaced by a fresh list with 1.1x the
// limits named by flags (the caller rebuilds anything that held the old one)
int nl_reinitialize(sponge_nl** nl, int flags);

int nl_pack_for_pairs(sponge_nl* nl, const float* crd, const float* charge, const int* type,
                      float4* pos, int* ptype, int* pgid, hipStream_t s);
}
