// Cell-grid half neighbor list for gfx950.
//
// Reference: SPONGE/neighbor_list/neighbor_list.cpp.  The pair set is the
// reference's: pair (i, j) is listed iff global(j) > global(i), the
// minimum-image r^2 (Get_Displacement<Periodic>, evaluated without FMA
// contraction) is < cutoff_skin_square = gl*gl*4 with gl = 0.5f*(cutoff+skin)
// (neighbor_list.cpp:859, 998), and j is not excluded from i
// (Delete_Excluded_Atoms_Serial_In_Neighbor_List, :736-807).
//
// The MI355X design differs from the reference's bucket-per-cell layout:
//   1. atoms are counting-sorted by cell (atomic histogram, one-block scan,
//      scatter) into a dense SoA float4 array, so every stencil column is a
//      contiguous, coalesced range of the sorted array;
//   2. the stencil of each cell (the reference's ascending +-2 neighbour cells,
//      Find_Neighor_Grids_Device :76-191) is compressed on the host into runs
//      of consecutive cell ids;
//   3. one wavefront owns one i atom and sweeps the runs 64 candidates at a
//      time; acceptance is compacted with a 64-bit ballot and mbcnt, so list
//      slots need no atomics and the exclusions are filtered inline instead
//      of by a separate swap-with-last pass.
#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "internal.h"
#include "batch_launch.h"
#include "timing.h"

using namespace sa;

int sa::stage_entries(size_t bytes_per_entry, int want, size_t reserve) {
    int dev = 0, lds = 65536;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        lds = 65536;
    const long fit = ((long)lds - 4096 - (long)reserve) / (long)bytes_per_entry;
    return (int)(std::max(64L, std::min((long)want, fit)) & ~63L);
}

namespace {

constexpr int kMaxGridNeighbors = 192;  // MAX_GRID_NEIGHBORS, neighbor_list.cpp:5

__global__ void k_nl_set(int* need, int v) { *need = v; }

// Check_Refresh (neighbor_list.cpp:307-321)
template <bool ORTHO>
__device__ __forceinline__ void nl_check_body(int sa_block, int N, const float* __restrict__ crd,
                                              const float* __restrict__ old, LT6 c, LT6 r,
                                              float permit2, int* need,
                                              const float* __restrict__ pruned, float pmove2,
                                              int* need_prune) {
    int i = sa_block * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float3 a = make_float3(crd[3 * i], crd[3 * i + 1], crd[3 * i + 2]);
    float3 b = make_float3(old[3 * i], old[3 * i + 1], old[3 * i + 2]);
    float3 d = disp_pbc<ORTHO>(a, b, c, r);
    if (norm2_exact(d) > permit2) *need = 1;
    if (pruned) {
        // the inner rows hold every pair within cutoff + delta at the prune:
        // valid while no atom moved more than delta / 2 since then
        const float3 e = disp_pbc<ORTHO>(a, make_float3(pruned[3 * i], pruned[3 * i + 1],
                                                        pruned[3 * i + 2]), c, r);
        if (norm2_exact(e) > pmove2) *need_prune = 1;
    }
}

template <bool ORTHO>
__global__ void k_nl_check(int N, const float* __restrict__ crd, const float* __restrict__ old,
                           LT6 c, LT6 r, float permit2, int* need, const float* __restrict__ pruned,
                           float pmove2, int* need_prune) {
    nl_check_body<ORTHO>(blockIdx.x, N, crd, old, c, r, permit2, need, pruned, pmove2, need_prune);
}

// Cell of an atom as Put_Atom_In_Grids wraps it (neighbor_list.cpp:357-377).
__device__ __forceinline__ int cell_of(float x, float y, float z, const LT6& c, float gl, int Nx,
                                       int Ny, int Nz) {
#pragma clang fp contract(off)
    float k3 = floorf(z / c.a33);
    z -= k3 * c.a33;
    y -= k3 * c.a32;
    float k2 = floorf(y / c.a22);
    y -= k2 * c.a22;
    x -= k3 * c.a31 + k2 * c.a21;
    x -= floorf(x / c.a11) * c.a11;
    int nx = (int)(x / gl), ny = (int)(y / gl), nz = (int)(z / gl);
    nx = nx < 0 ? 0 : (nx < Nx ? nx : Nx - 1);
    ny = ny < 0 ? 0 : (ny < Ny ? ny : Ny - 1);
    nz = nz < 0 ? 0 : (nz < Nz ? nz : Nz - 1);
    return nx * Ny * Nz + ny * Nz + nz;
}

__device__ __forceinline__ void nl_bin_body(int sa_block, const int* __restrict__ need, int N, const float* __restrict__ crd,
                         float* __restrict__ old, int need_copy, LT6 c, float gl, int Nx, int Ny,
                         int Nz, int* __restrict__ count, int* __restrict__ atom_cell,
                         int* __restrict__ atom_slot) {
    if (*need == 0) return;
    int i = sa_block * blockDim.x + threadIdx.x;
    if (i >= N) return;
    float x = crd[3 * i], y = crd[3 * i + 1], z = crd[3 * i + 2];
    if (need_copy) {
        old[3 * i] = x;
        old[3 * i + 1] = y;
        old[3 * i + 2] = z;
    }
    int cell = cell_of(x, y, z, c, gl, Nx, Ny, Nz);
    atom_cell[i] = cell;
    atom_slot[i] = atomicAdd(&count[cell], 1);
}

__global__ void k_nl_bin(const int* __restrict__ need, int N, const float* __restrict__ crd,
                         float* __restrict__ old, int need_copy, LT6 c, float gl, int Nx, int Ny,
                         int Nz, int* __restrict__ count, int* __restrict__ atom_cell,
                         int* __restrict__ atom_slot) {
    nl_bin_body(blockIdx.x, need, N, crd, old, need_copy, c, gl, Nx, Ny, Nz, count, atom_cell, atom_slot);
}

// One-block exclusive scan of the cell counts -> cell_start[G+1]; flags cells
// above max_in_grid (the reference's grid overflow, :382-386) without
// dropping atoms.
__device__ __forceinline__ void nl_scan_body(int sa_block, const int* __restrict__ need, int G,
                                                  int* __restrict__ count, int* __restrict__ start,
                                                  int max_in_grid, int* __restrict__ flags,
                                                  unsigned long long* __restrict__ rebuilds) {
    if (*need == 0) return;
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int chunk = (G + 1023) / 1024;
    const int b = t * chunk, e = min(G, b + chunk);
    int s = 0, over = 0;
    for (int c = b; c < e; ++c) {
        int v = count[c];
        s += v;
        over |= v > max_in_grid;
    }
    if (over) atomicOr(flags, SPONGE_AMD_NL_GRID_OVERFLOW);
    part[t] = s;
    __syncthreads();
    // Hillis-Steele inclusive scan over 1024 partial sums
    for (int off = 1; off < 1024; off <<= 1) {
        int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;  // exclusive prefix of this chunk
    for (int c = b; c < e; ++c) {
        start[c] = run;
        run += count[c];
        count[c] = 0;  // left clear for the next rebuild
    }
    if (t == 1023) {
        start[G] = part[1023];
        rebuilds[0] += 1;
    }
}

__global__ void __launch_bounds__(1024) k_nl_scan(const int* __restrict__ need, int G,
                                                  int* __restrict__ count, int* __restrict__ start,
                                                  int max_in_grid, int* __restrict__ flags,
                                                  unsigned long long* __restrict__ rebuilds) {
    nl_scan_body(blockIdx.x, need, G, count, start, max_in_grid, flags, rebuilds);
}

__device__ __forceinline__ void nl_scatter_body(int sa_block, const int* __restrict__ need, int N, const float* __restrict__ crd,
                             const int* __restrict__ atom_local,
                             const int* __restrict__ atom_cell, const int* __restrict__ atom_slot,
                             const int* __restrict__ start, const int2* __restrict__ gexw,
                             int* __restrict__ sidx, int* __restrict__ rank,
                             float4* __restrict__ spos, int2* __restrict__ sexw) {
    if (*need == 0) return;
    int i = sa_block * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int p = start[atom_cell[i]] + atom_slot[i];
    int gi = atom_local ? atom_local[i] : i;
    sidx[p] = i;
    rank[i] = p;
    spos[p] = make_float4(crd[3 * i], crd[3 * i + 1], crd[3 * i + 2], __int_as_float(gi));
    sexw[p] = gexw ? gexw[gi] : make_int2(1, 0);
}

__global__ void k_nl_scatter(const int* __restrict__ need, int N, const float* __restrict__ crd,
                             const int* __restrict__ atom_local,
                             const int* __restrict__ atom_cell, const int* __restrict__ atom_slot,
                             const int* __restrict__ start, const int2* __restrict__ gexw,
                             int* __restrict__ sidx, int* __restrict__ rank,
                             float4* __restrict__ spos, int2* __restrict__ sexw) {
    nl_scatter_body(blockIdx.x, need, N, crd, atom_local, atom_cell, atom_slot, start, gexw, sidx, rank, spos, sexw);
}

// Combined exclusion window of every atom: the smallest and largest id among
// the atoms of its own exclusion list and the atoms whose lists hold it.  A
// pair (i, j) with j outside i's combined window cannot be excluded, so the
// search tests the reference's exclusion rule only inside it.
__global__ void k_nl_exwin_init(int N, int2* __restrict__ gexw) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) gexw[i] = make_int2(0x7fffffff, -1);
}

__global__ void k_nl_exwin(int N, const int* __restrict__ ex_start, const int* __restrict__ ex_list,
                           const int* __restrict__ ex_num, int2* __restrict__ gexw) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= N) return;
    const int n = ex_num[a], s0 = ex_start[a];
    for (int e = s0; e < s0 + n; ++e) {
        const int b = ex_list[e];
        if (b < 0 || b >= N) continue;
        atomicMin(&gexw[a].x, b);
        atomicMax(&gexw[a].y, b);
        atomicMin(&gexw[b].x, a);
        atomicMax(&gexw[b].y, a);
    }
}

// Candidate stream of a cell: its stencil runs of consecutive cells, i.e.
// contiguous ranges [first_k, first_k + len_k) of the sorted array,
// concatenated.  Built in LDS by the block (loads in parallel, scan by one
// thread over at most kMaxRanges runs).

__device__ __forceinline__ int cell_stream(int cell, const int* __restrict__ start,
                                           const int2* __restrict__ ranges,
                                           const int* __restrict__ nranges, int max_ranges,
                                           int* r_first, int* r_off) {
    const int nr = nranges[cell];
    for (int k = threadIdx.x; k < nr; k += blockDim.x) {
        const int2 cr = ranges[(size_t)cell * max_ranges + k];
        const int f = start[cr.x];
        r_first[k] = f;
        r_off[k + 1] = start[cr.y] - f;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        r_off[0] = 0;
        for (int k = 0; k < nr; ++k) r_off[k + 1] += r_off[k];
    }
    __syncthreads();
    return nr;
}

// One workgroup (8 wavefronts) per cell.  The cell's candidate stream (sorted
// positions with the global id in w) is staged into LDS in chunks of
// kSearchChunk positions; each wavefront takes up to ROWS i atoms of the cell
// (kSearchRows = 4; 2 in the batched driver's launch) and sweeps every chunk,
// two candidates per lane per pass, testing each against all of its rows (one
// LDS read feeds up to ROWS independent tests).  Small chunks keep the workgroup's LDS small, so that
// several workgroups share a CU: the search is bound by latency (dependent
// compares, ballots and stores), which only more resident wavefronts hide.
// The pair test is the reference's Find_Neighbors_Gridly criterion
// (Get_Displacement, r^2 < cutoff_skin_square) and, inside the row's
// combined exclusion window, its exclusion rule.  Rows hold both directions
// as 16-bit stream positions, in ascending order.
//
// In an orthorhombic box wider than twice (list radius + half a cell) on
// every axis, the chunk holds each candidate's image nearest the cell's
// centre C, and the rows' images nearest C are taken too: a candidate within
// the list radius of a row atom then has that image at its minimum-image
// displacement, which is a plain difference (3 subtractions and 3 FMAs
// instead of the reference's 18 operations).  That r^2 differs from the
// reference's by at most `margin` (rounding of the shifted coordinates and of
// the two formulas; bound below); a candidate decided by it is decided the
// same way by the reference, and the few within the margin of the threshold
// are tested again with the reference arithmetic on the raw coordinates, so
// the list is the reference's set bit for bit.
// stream positions staged at a time: 32 KiB of LDS, three workgroups per CU
// (measured on the 95,000-atom water step: chunks of 512 / 1024 / 2048 and the
// whole stream, 3,584 entries: 557 / 519 / 490 / 552 us per rebuild)
constexpr int kSearchChunk = 2048;
constexpr int kSearchRows = 4;
constexpr int kSearchThreads = 512;  // workgroup size of the search

// Bound on |r2_ref - r2_shift| for a pair within the list radius: per axis
// the reference displacement fl(fl(a-b) - fl(s L)) and the shifted one
// fl(a' - b'), a' = fl(a - fl(k L)), both differ from the exact minimum-image
// displacement D by at most u (6 M + 7 L + 2 |D|) together (u = 2^-24,
// M = the largest coordinate magnitude, |D| <= 13 near the threshold), and
// |d|^2 vs |d'|^2 adds delta (2 sqrt(3 thr) + 3 delta) plus 6 u thr of the
// two r^2 sums; taken x4.
__host__ __device__ inline float search_margin(float M, float L, float thr) {
    const float u = 5.9604645e-08f;
    const float rt = sqrtf(thr);
    const float delta = u * (6.f * M + 7.f * L + 2.f * rt + 2.f);
    return 4.f * (delta * (2.f * sqrtf(3.f) * rt + 3.f * delta) + 6.f * u * thr);
}

// The reference's exclusion rule for the unordered pair {a, b}, applied in the
// half-list row of the lower id lo (Delete_Excluded_Atoms_Serial_In_Neighbor_
// List, neighbor_list.cpp:736-807): the higher id must lie inside the window
// [list[first], list[last]] of lo's exclusion list and be one of its entries.

__device__ __forceinline__ bool excluded_pair_inl(int a, int b, const int* __restrict__ ex_start,
                                                  const int* __restrict__ ex_list,
                                                  const int* __restrict__ ex_num) {
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    const int n = ex_num[lo];
    if (n <= 0) return false;
    const int s0 = ex_start[lo];
    if (hi < ex_list[s0] || hi > ex_list[s0 + n - 1]) return false;
    bool hit = false;
    for (int e = s0; e < s0 + n; ++e) hit |= ex_list[e] == hi;
    return hit;
}

template <bool ORTHO, int ROWS = kSearchRows>
__device__ __forceinline__ void nl_search_cells_body(int sa_block, 
    const int* __restrict__ need, const float4* __restrict__ spos, const int2* __restrict__ sexw,
    const int* __restrict__ start, const int2* __restrict__ ranges,
    const int* __restrict__ nranges, int max_ranges, LT6 c, LT6 r, float thr,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list,
    const int* __restrict__ ex_num, int cap, int* __restrict__ list_count,
    unsigned short* __restrict__ list, int* __restrict__ g_rfirst, int* __restrict__ g_roff,
    int* __restrict__ flags, float gl, int Nx, int Ny, int Nz) {
    if (*need == 0) return;
    __shared__ float4 cand[kSearchChunk];
    __shared__ int r_first[kMaxRanges];
    __shared__ int r_off[kMaxRanges + 1];
    __shared__ unsigned s_maxbits;
    const int cell = sa_block;
    const int nr = cell_stream(cell, start, ranges, nranges, max_ranges, r_first, r_off);
    const int total = r_off[nr];
    // the stream tables are kept for the pair and export kernels
    for (int k = threadIdx.x; k < nr; k += blockDim.x) g_rfirst[(size_t)cell * max_ranges + k] = r_first[k];
    for (int k = threadIdx.x; k <= nr; k += blockDim.x) g_roff[(size_t)cell * (max_ranges + 1) + k] = r_off[k];
    const int a0 = start[cell], a1 = start[cell + 1];
    if (total > 65535) {
        // positions no longer fit 16 bits: empty rows and a raised flag, which
        // the step driver turns into an error at its next overflow check
        for (int p = a0 + (int)threadIdx.x; p < a1; p += blockDim.x) list_count[p] = 0;
        if (threadIdx.x == 0) atomicOr(flags, SPONGE_AMD_NL_STREAM_OVERFLOW | SPONGE_AMD_NL_LIST_OVERFLOW);
        return;
    }
    // cell-centre images: box wide enough on every axis (uniform per block)
    float3 C = make_float3(0.f, 0.f, 0.f);
    bool shift = false;
    if (ORTHO) {
        const int ix = cell / (Ny * Nz), iy = (cell / Nz) % Ny, iz = cell % Nz;
        const float lx = ix * gl, hx = ix == Nx - 1 ? c.a11 : (ix + 1) * gl;
        const float ly = iy * gl, hy = iy == Ny - 1 ? c.a22 : (iy + 1) * gl;
        const float lz = iz * gl, hz = iz == Nz - 1 ? c.a33 : (iz + 1) * gl;
        C = make_float3(0.5f * (lx + hx), 0.5f * (ly + hy), 0.5f * (lz + hz));
        const float reach = sqrtf(thr) + 0.01f;
        shift = 0.5f * c.a11 > reach + 0.5f * (hx - lx) && 0.5f * c.a22 > reach + 0.5f * (hy - ly) &&
                0.5f * c.a33 > reach + 0.5f * (hz - lz);
    }
    const float Lmax = fmaxf(c.a11, fmaxf(c.a22, c.a33));
    // the largest coordinate magnitude of the cell's atoms (part of the
    // margin of every chunk)
    float mrow = 0.f;
    if (shift) {
        for (int p = a0 + (int)threadIdx.x; p < a1; p += blockDim.x) {
            const float4 v = spos[p];
            mrow = fmaxf(mrow, fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z))));
        }
    }
    // running maximum of the coordinate magnitudes staged so far (only
    // grows, so each chunk's margin covers the chunk)
    if (threadIdx.x == 0) s_maxbits = 0u;
    const int lane = threadIdx.x & (kWave - 1);
    // the wavefront index as a scalar: row bases and counters stay in SGPRs
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int nwave = blockDim.x / kWave;
    const int rows = a1 - a0;
    // every wavefront runs the same row-block and chunk loops (the barriers
    // of the staging); a wavefront without rows in a block only stages
    for (int k0 = 0; k0 * nwave < rows; k0 += ROWS) {
        float3 pi[ROWS], pc[ROWS];
        int gi[ROWS], cnt[ROWS];
        int2 wi[ROWS];
        bool act[ROWS];
#pragma unroll
        for (int u = 0; u < ROWS; ++u) {
            const int row = wave + (k0 + u) * nwave;
            act[u] = row < rows;
            const int p = a0 + (act[u] ? row : 0);
            const float4 v = spos[p];
            pi[u] = make_float3(v.x, v.y, v.z);
            pc[u] = shift ? near_image(pi[u], C, c, r) : pi[u];
            gi[u] = __float_as_int(v.w);
            // the exclusion window widened to hold the row atom itself: the
            // row's own stream entry takes the window's rare path, which
            // drops it by id (the sweep tests no ids otherwise)
            const int2 w = sexw[p];
            wi[u] = w.x > w.y ? make_int2(gi[u], gi[u])
                              : make_int2(min(w.x, gi[u]), max(w.y, gi[u]));
            cnt[u] = 0;
        }
        int na = 0;
#pragma unroll
        for (int u = 0; u < ROWS; ++u) na += act[u] ? 1 : 0;
        for (int c0 = 0; c0 < total; c0 += kSearchChunk) {
            const int clen = min(kSearchChunk, total - c0);
            __syncthreads();  // the previous chunk is no longer read
            // stage stream positions [c0, c0 + clen): one wavefront per run
            float mx = mrow;
            for (int k = wave; k < nr; k += nwave) {
                const int lo = max(r_off[k], c0), hi = min(r_off[k + 1], c0 + clen);
                for (int t = lo + lane; t < hi; t += kWave) {
                    const float4 v = spos[r_first[k] + (t - r_off[k])];
                    if (shift) {
                        mx = fmaxf(mx, fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z))));
                        const float3 w = near_image(make_float3(v.x, v.y, v.z), C, c, r);
                        cand[t - c0] = make_float4(w.x, w.y, w.z, v.w);
                    } else {
                        cand[t - c0] = v;
                    }
                }
            }
            if (shift) {
                mx = wave_max(mx);
                if (lane == 0) atomicMax(&s_maxbits, __float_as_uint(mx));
            }
            __syncthreads();
            float thr_lo = thr, thr_hi = thr;
            if (shift) {
                const float m = search_margin(__uint_as_float(s_maxbits), Lmax, thr);
                thr_lo = thr - m;
                thr_hi = thr + m;
            }
            // the reference test on raw coordinates (stream position t)
            auto exact = [&](int u, int t, float3 pj) {
                if (shift) {
                    const float4 raw = spos[stream_to_sorted(t, nr, r_first, r_off)];
                    pj = make_float3(raw.x, raw.y, raw.z);
                }
                return norm2_exact(disp_pbc<ORTHO>(pi[u], pj, c, r)) < thr;
            };
            // Two candidates per lane per pass (stream positions t and
            // t + 64): the two halves' tests are independent.  The sweep is
            // compiled per count of active rows (uniform per wavefront), so
            // empty row slots cost nothing.
            auto sweep = [&](auto na_tag, auto shift_tag) {
                constexpr int NA = decltype(na_tag)::value;
                constexpr bool SH = decltype(shift_tag)::value;
                constexpr int H = 2;
                unsigned short* rowp[NA];
#pragma unroll
                for (int u = 0; u < NA; ++u) rowp[u] = list + (size_t)(a0 + wave + (k0 + u) * nwave) * cap;
                // a row's hits of one half-pass appended in stream order (mask
                // m of the lanes holding ok): one uniform capacity test unless
                // the row is about to overflow
                auto append = [&](int u, int h, unsigned long long m, bool okv, int t) {
                    const int after = cnt[u] + (int)__popcll(m);
                    if (okv) {
                        const unsigned pos = __builtin_amdgcn_mbcnt_hi(
                            (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, (unsigned)cnt[u]));
                        if (after <= cap || pos < (unsigned)cap) rowp[u][pos] = (unsigned short)t;
                    }
                    cnt[u] = after;
                };
                for (int base = 0; base < clen; base += H * kWave) {
                    float3 pj[H];
                    int gj[H], tt[H];
                    bool in[H];
#pragma unroll
                    for (int h = 0; h < H; ++h) {
                        const int loc = base + h * kWave + lane;
                        in[h] = loc < clen;
                        tt[h] = c0 + loc;
                        const float4 v = cand[in[h] ? loc : 0];
                        // SH: a lane past the chunk tests a far point (r2
                        // ~1e36, beyond every threshold; no test of in)
                        pj[h] = SH && !in[h] ? make_float3(1e18f, 1e18f, 1e18f)
                                             : make_float3(v.x, v.y, v.z);
                        gj[h] = __float_as_int(v.w);
                    }
                    bool ok[H][NA], amb[H][NA], chk[H][NA];
                    bool rare = false;
#pragma unroll
                    for (int h = 0; h < H; ++h) {
#pragma unroll
                        for (int u = 0; u < NA; ++u) {
                            if (SH) {
                                const float dx = pj[h].x - pc[u].x, dy = pj[h].y - pc[u].y,
                                            dz = pj[h].z - pc[u].z;
                                const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                                ok[h][u] = r2 < thr_lo;
                                amb[h][u] = (r2 < thr_hi) & !ok[h][u];
                            } else {
                                // Get_Displacement(crd_i, crd_j) as in Find_Neighbors_Gridly
                                ok[h][u] = in[h] & (norm2_exact(disp_pbc<ORTHO>(pi[u], pj[h], c, r)) < thr);
                                amb[h][u] = false;
                            }
                            // inside the row's window (one unsigned compare;
                            // the rare path applies the exclusion rule only
                            // to the candidates found within the list radius)
                            chk[h][u] = (unsigned)(gj[h] - wi[u].x) <= (unsigned)(wi[u].y - wi[u].x);
                            rare |= amb[h][u] | chk[h][u];
                        }
                    }
                    if (__ballot(rare)) {
#pragma unroll
                        for (int h = 0; h < H; ++h) {
#pragma unroll
                            for (int u = 0; u < NA; ++u) {
                                if (SH && amb[h][u]) ok[h][u] = exact(u, tt[h], pj[h]);
                                if (chk[h][u] && ok[h][u])
                                    ok[h][u] = gj[h] != gi[u] &&
                                               !excluded_pair_inl(gi[u], gj[h], ex_start, ex_list, ex_num);
                            }
                        }
#pragma unroll
                        for (int u = 0; u < NA; ++u)
#pragma unroll
                            for (int h = 0; h < H; ++h) append(u, h, __ballot(ok[h][u]), ok[h][u], tt[h]);
                    } else {
                        // no exclusion or rounding case in this pass: the
                        // compares' masks are final
#pragma unroll
                        for (int u = 0; u < NA; ++u)
#pragma unroll
                            for (int h = 0; h < H; ++h) append(u, h, __ballot(ok[h][u]), ok[h][u], tt[h]);
                    }
                }
            };
            static_assert(ROWS == 4 || ROWS == 2, "the sweep dispatch covers 2 or 4 rows");
            using I1 = std::integral_constant<int, 1>;
            using I2 = std::integral_constant<int, 2>;
            using I3 = std::integral_constant<int, 3>;
            using I4 = std::integral_constant<int, 4>;
            auto dispatch = [&](auto shift_tag) {
                if constexpr (ROWS == 4) {
                    if (na == 4) sweep(I4{}, shift_tag);
                    else if (na == 3) sweep(I3{}, shift_tag);
                    else if (na == 2) sweep(I2{}, shift_tag);
                    else if (na == 1) sweep(I1{}, shift_tag);
                } else {
                    if (na == 2) sweep(I2{}, shift_tag);
                    else if (na == 1) sweep(I1{}, shift_tag);
                }
            };
            if (shift) dispatch(std::true_type{});
            else dispatch(std::false_type{});
        }
        if (lane == 0) {
#pragma unroll
            for (int u = 0; u < ROWS; ++u) {
                if (!act[u]) continue;
                const int p = a0 + wave + (k0 + u) * nwave;
                list_count[p] = cnt[u] < cap ? cnt[u] : cap;
                if (cnt[u] > cap) atomicOr(flags, SPONGE_AMD_NL_LIST_OVERFLOW);
            }
        }
    }
}

template <bool ORTHO>
__global__ void __launch_bounds__(kSearchThreads) k_nl_search_cells(
    const int* __restrict__ need, const float4* __restrict__ spos, const int2* __restrict__ sexw,
    const int* __restrict__ start, const int2* __restrict__ ranges,
    const int* __restrict__ nranges, int max_ranges, LT6 c, LT6 r, float thr,
    const int* __restrict__ ex_start, const int* __restrict__ ex_list,
    const int* __restrict__ ex_num, int cap, int* __restrict__ list_count,
    unsigned short* __restrict__ list, int* __restrict__ g_rfirst, int* __restrict__ g_roff,
    int* __restrict__ flags, float gl, int Nx, int Ny, int Nz) {
    nl_search_cells_body<ORTHO>(blockIdx.x, need, spos, sexw, start, ranges, nranges, max_ranges, c, r, thr, ex_start, ex_list, ex_num, cap, list_count, list, g_rfirst, g_roff, flags, gl, Nx, Ny, Nz);
}


// Reference-format export: the half rows (global(j) > global(i)) in local
// ids, rows by local atom; one workgroup per cell, one wavefront per row,
// ballot compaction.
__global__ void __launch_bounds__(256) k_nl_export(
    const int* __restrict__ sidx, const float4* __restrict__ spos, const int* __restrict__ start,
    const int* __restrict__ nranges, int max_ranges, const int* __restrict__ g_rfirst,
    const int* __restrict__ g_roff, const int* __restrict__ list_count,
    const unsigned short* __restrict__ list, int cap, int* __restrict__ counts,
    int* __restrict__ serial, int max_out, sa_atom_group* group, int* __restrict__ flags) {
    __shared__ int r_first[kMaxRanges];
    __shared__ int r_off[kMaxRanges + 1];
    const int cell = blockIdx.x;
    const int nr = nranges[cell];
    for (int k = threadIdx.x; k < nr; k += blockDim.x) r_first[k] = g_rfirst[(size_t)cell * max_ranges + k];
    for (int k = threadIdx.x; k <= nr; k += blockDim.x) r_off[k] = g_roff[(size_t)cell * (max_ranges + 1) + k];
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave, nwave = blockDim.x / kWave;
    for (int p = start[cell] + wave; p < start[cell + 1]; p += nwave) {
        const int i = sidx[p];
        const int gi = __float_as_int(spos[p].w);
        const int n = list_count[p];
        const unsigned short* in = list + (size_t)p * cap;
        int* o = serial + (size_t)i * max_out;
        int count = 0;
        for (int base = 0; base < n; base += kWave) {
            const int k = base + lane;
            int q = 0;
            bool keep = false;
            if (k < n) {
                q = stream_to_sorted(in[k], nr, r_first, r_off);
                keep = __float_as_int(spos[q].w) > gi;
            }
            const unsigned long long m = __ballot(keep);
            if (keep) {
                const int pos = count + lanes_below(m);
                if (pos < max_out) o[pos] = sidx[q];
            }
            count += __popcll(m);
        }
        if (lane == 0) {
            const int nout = count < max_out ? count : max_out;
            counts[i] = nout;
            if (group) {
                group[i].atom_numbers = nout;
                group[i].ghost_numbers = 0;
                group[i].atom_serial = o;
            }
            if (count > max_out) atomicOr(flags, SPONGE_AMD_NL_LIST_OVERFLOW);
        }
    }
}

__global__ void k_nl_sum_counts(int N, const int* __restrict__ cnt, unsigned long long* out) {
    unsigned long long s = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) s += cnt[i];
    s = wave_sum(s);
    if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(out, s);
}

// Neighbour cells of every cell, as the enumeration branch of
// Find_Neighor_Grids_Device (neighbor_list.cpp:96-181): for each of the 27
// periodic images the cells with offset in [-2, 2] (the low bound widened to
// -3 along x / y for sheared images), sorted ascending and de-duplicated;
// then compressed into runs of consecutive ids.
int build_stencil(sponge_nl* nl, std::vector<int2>& ranges, std::vector<int>& nranges,
                  int& max_ranges) {
    const int Nx = nl->Nx, Ny = nl->Ny, Nz = nl->Nz;
    const LT6 cell = nl->cell0;
    const float gl = nl->grid_length;
    const int64_t dxy = (int64_t)(cell.a21 / gl), dxz = (int64_t)(cell.a31 / gl),
                  dyz = (int64_t)(cell.a32 / gl);
    std::vector<std::vector<int2>> all(nl->G);
    max_ranges = 0;
    std::vector<int> cand;
    for (int g = 0; g < nl->G; ++g) {
        const int64_t gx = g / ((int64_t)Ny * Nz), gy = (g % ((int64_t)Ny * Nz)) / Nz, gz = g % Nz;
        cand.clear();
        for (int k = 0; k < 27; ++k) {
            const int dx = k / 9 - 1, dy = (k % 9) / 3 - 1, dz = k % 3 - 1;
            const int cx = (double)dy * cell.a21 + (double)dz * cell.a31 == 0.0 ? -2 : -3;
            const int cy = (double)dz * cell.a32 == 0.0 ? -2 : -3;
            const int64_t sx = (int64_t)dx * Nx + dy * dxy + dz * dxz;
            const int64_t sy = (int64_t)dy * Ny + dz * dyz;
            const int64_t sz = (int64_t)dz * Nz;
            const int64_t lox = std::max<int64_t>(gx + sx - 2, 0), hix = std::min<int64_t>(gx + sx - cx, Nx - 1);
            const int64_t loy = std::max<int64_t>(gy + sy - 2, 0), hiy = std::min<int64_t>(gy + sy - cy, Ny - 1);
            const int64_t loz = std::max<int64_t>(gz + sz - 2, 0), hiz = std::min<int64_t>(gz + sz + 2, Nz - 1);
            for (int64_t jx = lox; jx <= hix; ++jx)
                for (int64_t jy = loy; jy <= hiy; ++jy)
                    for (int64_t jz = loz; jz <= hiz; ++jz) cand.push_back((int)((jx * Ny + jy) * Nz + jz));
        }
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
        if ((int)cand.size() > kMaxGridNeighbors) return SPONGE_AMD_ERROR_OVERFLOW;
        auto& rr = all[g];
        for (size_t a = 0; a < cand.size();) {
            size_t b = a + 1;
            while (b < cand.size() && cand[b] == cand[b - 1] + 1) ++b;
            rr.push_back(make_int2(cand[a], cand[b - 1] + 1));
            a = b;
        }
        max_ranges = std::max(max_ranges, (int)rr.size());
    }
    ranges.assign((size_t)nl->G * max_ranges, make_int2(0, 0));
    nranges.assign(nl->G, 0);
    for (int g = 0; g < nl->G; ++g) {
        nranges[g] = (int)all[g].size();
        for (size_t k = 0; k < all[g].size(); ++k) ranges[(size_t)g * max_ranges + k] = all[g][k];
    }
    return 0;
}

void nl_free(sponge_nl* nl) {
    dfree(nl->d_need);
    dfree(nl->d_flags);
    dfree(nl->d_rebuilds);
    dfree(nl->d_cell_count);
    dfree(nl->d_cell_start);
    dfree(nl->d_atom_cell);
    dfree(nl->d_atom_slot);
    dfree(nl->d_sidx);
    dfree(nl->d_rank);
    dfree(nl->d_spos);
    dfree(nl->d_old_crd);
    dfree(nl->d_ranges);
    dfree(nl->d_nranges);
    dfree(nl->d_list_count);
    dfree(nl->d_list);
    dfree(nl->d_sexw);
    dfree(nl->d_gexw);
    dfree(nl->d_rfirst);
    dfree(nl->d_roff);
    dfree(nl->d_exp_serial);
    dfree(nl->d_exp_count);
    dfree(nl->d_atom_group);
    dfree(nl->d_need_prune);
    dfree(nl->d_prune_crd);
    dfree(nl->d_ilist_count);
    dfree(nl->d_ilist);
}

}  // namespace

// the batched search's register budget: its LDS allows 6 wavefronts per SIMD
template <>
struct sa::BodyWaves<&nl_search_cells_body<true, 2>> {
    static constexpr int value = 6;
};

extern "C" void sponge_nl_default_options(sponge_nl_options* o) {
    o->refresh_interval = 0;
    o->skin_permit = 0.5f;
    o->max_neighbor_numbers = 1200;
    o->max_atom_in_grid_numbers = 150;
    o->throw_error_when_overflow = 0;
    o->check_overflow_interval = 150;
}

extern "C" int sponge_nl_initial(sponge_nl** out, int atom_numbers, float cutoff, float skin,
                                 const float* cell, const float* rcell,
                                 const sponge_nl_options* opt, void* stream) {
    if (!out || !cell || !rcell || atom_numbers <= 0 || !(cutoff > 0) || !(skin >= 0))
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    *out = nullptr;
    sponge_nl* nl = new sponge_nl();
    if (opt) nl->opt = *opt;
    else sponge_nl_default_options(&nl->opt);
    if (nl->opt.max_neighbor_numbers <= 0 || nl->opt.max_atom_in_grid_numbers <= 0) {
        delete nl;
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    }
    nl->N = atom_numbers;
    nl->cutoff = cutoff;
    nl->skin = skin;
    nl->stream = (hipStream_t)stream;
    nl->cell0 = lt_from(cell);
    nl->rcell0 = lt_from(rcell);
    nl->ortho = lt_is_orthorhombic(nl->cell0) ? 1 : 0;
    nl->grid_length = (cutoff + skin) * 0.5f;                       // :948
    nl->thr = nl->grid_length * nl->grid_length * 4.0f;             // :859
    nl->Nx = (int)floorf(nl->cell0.a11 / nl->grid_length);          // :198-200
    nl->Ny = (int)floorf(nl->cell0.a22 / nl->grid_length);
    nl->Nz = (int)floorf(nl->cell0.a33 / nl->grid_length);
    if (nl->Nx <= 2 || nl->Ny <= 2 || nl->Nz <= 2) {  // :949-954
        delete nl;
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    }
    nl->G = nl->Nx * nl->Ny * nl->Nz;
    std::vector<int2> ranges;
    std::vector<int> nranges;
    int rc = build_stencil(nl, ranges, nranges, nl->max_ranges);
    if (rc) {
        delete nl;
        return rc;
    }
    const int N = nl->N, G = nl->G;
    // rows hold both directions: twice the reference's per-atom half capacity
    nl->cap = 2 * nl->opt.max_neighbor_numbers;
    if (nl->max_ranges > kMaxRanges) {
        delete nl;
        return SPONGE_AMD_ERROR_OVERFLOW;
    }
    int e = 0;
    e |= dmalloc(&nl->d_need, 1);
    e |= dmalloc(&nl->d_flags, 1);
    e |= dmalloc(&nl->d_rebuilds, 1);
    e |= dmalloc(&nl->d_cell_count, G);
    e |= dmalloc(&nl->d_cell_start, G + 1);
    e |= dmalloc(&nl->d_atom_cell, N);
    e |= dmalloc(&nl->d_atom_slot, N);
    e |= dmalloc(&nl->d_sidx, N);
    e |= dmalloc(&nl->d_rank, N);
    e |= dmalloc(&nl->d_spos, N);
    e |= dmalloc(&nl->d_old_crd, (size_t)3 * N);
    e |= dmalloc(&nl->d_ranges, ranges.size());
    e |= dmalloc(&nl->d_nranges, G);
    e |= dmalloc(&nl->d_list_count, N);
    e |= dmalloc(&nl->d_list, (size_t)N * nl->cap);
    e |= dmalloc(&nl->d_sexw, N);
    e |= dmalloc(&nl->d_gexw, N);
    e |= dmalloc(&nl->d_rfirst, (size_t)G * nl->max_ranges);
    e |= dmalloc(&nl->d_roff, (size_t)G * (nl->max_ranges + 1));
    if (e) {
        nl_free(nl);
        delete nl;
        return SPONGE_AMD_ERROR_MALLOC_FAILED;
    }
    hipStream_t s = nl->stream;
    hipError_t he = hipMemcpyAsync(nl->d_ranges, ranges.data(), ranges.size() * sizeof(int2),
                                   hipMemcpyHostToDevice, s);
    if (he == hipSuccess)
        he = hipMemcpyAsync(nl->d_nranges, nranges.data(), nranges.size() * sizeof(int),
                            hipMemcpyHostToDevice, s);
    int fe = 0;
    if (he == hipSuccess) {
        fe |= fill_async(nl->d_flags, 0, 1, s);
        fe |= fill_async(nl->d_rebuilds, 0ull, 1, s);
        fe |= fill_async(nl->d_cell_count, 0, G, s);
        fe |= fill_async(nl->d_list_count, 0, N, s);
        he = hipStreamSynchronize(s);
    }
    if (he != hipSuccess || fe) {
        nl_free(nl);
        delete nl;
        return he != hipSuccess ? (int)he : fe;
    }
    *out = nl;
    return 0;
}

extern "C" int sponge_nl_update(sponge_nl* nl, const int* atom_local, int local_atom_numbers,
                                int ghost_numbers, const float* crd, const float* cell,
                                const float* rcell, int step, int update,
                                const int* excluded_list_start, const int* excluded_list,
                                const int* excluded_numbers) {
    return nl_update_impl(nl, atom_local, local_atom_numbers, ghost_numbers, crd, cell, rcell, step,
                          update, excluded_list_start, excluded_list, excluded_numbers, false);
}

// need_cleared: the caller guarantees the rebuild flag is already 0 (the MD
// driver clears it after the pair kernel), so the automatic-refresh check
// only has to raise it.
int sa::nl_update_impl(sponge_nl* nl, const int* atom_local, int local_atom_numbers,
                       int ghost_numbers, const float* crd, const float* cell, const float* rcell,
                       int step, int update, const int* excluded_list_start,
                       const int* excluded_list, const int* excluded_numbers, bool need_cleared) {
    if (!nl || !crd || !cell || !rcell) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (ghost_numbers != 0) return SPONGE_AMD_ERROR_NOT_IMPLEMENTED;
    if (local_atom_numbers <= 0) return 0;  // :980
    if (local_atom_numbers != nl->N) return SPONGE_AMD_ERROR_CONFLICTING_COMMAND;
    if ((excluded_numbers != nullptr) != (excluded_list_start != nullptr))
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    const LT6 c = lt_from(cell), r = lt_from(rcell);
    // the cell stencil belongs to the box of Initial (the reference rebuilds
    // the module when the box changes, SPONGE/main.cpp:2104-2110)
    if (memcmp(&c, &nl->cell0, sizeof(LT6)) != 0) return SPONGE_AMD_ERROR_CONFLICTING_COMMAND;
    const int N = nl->N;
    hipStream_t s = nl->stream;
    const bool auto_refresh = nl->opt.refresh_interval <= 0;
    // need flag (neighbor_list.cpp:983-998)
    if (update == SPONGE_AMD_NL_FORCED || !nl->built) {
        hipLaunchKernelGGL(k_nl_set, dim3(1), dim3(1), 0, s, nl->d_need, 1);
    } else if (auto_refresh) {
        if (!need_cleared) hipLaunchKernelGGL(k_nl_set, dim3(1), dim3(1), 0, s, nl->d_need, 0);
        const float permit2 = nl->skin * nl->skin * nl->opt.skin_permit * nl->opt.skin_permit;
        const bool prune = nl->prune_delta > 0.f;
        const float half = 0.5f * nl->prune_delta - 1e-3f;  // margin for the rounding of both tests
        const float pmove2 = half * half;
        const float* pruned = prune ? nl->d_prune_crd : nullptr;
        int* need_prune = prune ? nl->d_need_prune : nullptr;
        if (nl->ortho)
            SA_LAUNCH(256, (&nl_check_body<true>), (k_nl_check<true>), dim3(grid_for(N, 256)),
                      dim3(256), 0, s, N, crd, nl->d_old_crd, c, r, permit2, nl->d_need, pruned,
                      pmove2, need_prune);
        else
            SA_LAUNCH(256, (&nl_check_body<false>), (k_nl_check<false>), dim3(grid_for(N, 256)),
                      dim3(256), 0, s, N, crd, nl->d_old_crd, c, r, permit2, nl->d_need, pruned,
                      pmove2, need_prune);
    } else {
        hipLaunchKernelGGL(k_nl_set, dim3(1), dim3(1), 0, s, nl->d_need,
                           ((step + 1) % nl->opt.refresh_interval == 0) ? 1 : 0);
    }
    SA_LAUNCH_CHECK();
    SA_LAUNCH(256, (&nl_bin_body), (k_nl_bin), dim3(grid_for(N, 256)), dim3(256), 0, s, nl->d_need, N, crd,
                       nl->d_old_crd, auto_refresh ? 1 : 0, c, nl->grid_length, nl->Nx, nl->Ny,
                       nl->Nz, nl->d_cell_count, nl->d_atom_cell, nl->d_atom_slot);
    SA_LAUNCH_CHECK();
    SA_LAUNCH(1024, (&nl_scan_body), (k_nl_scan), dim3(1), dim3(1024), 0, s, nl->d_need, nl->G, nl->d_cell_count,
                       nl->d_cell_start, nl->opt.max_atom_in_grid_numbers, nl->d_flags, nl->d_rebuilds);
    SA_LAUNCH_CHECK();
    // combined exclusion windows: the exclusion lists are topology, so the
    // windows are derived once per set of list pointers (and again on a
    // forced update)
    const bool have_ex = excluded_numbers != nullptr;
    if (have_ex && (update == SPONGE_AMD_NL_FORCED || nl->exw_src[0] != excluded_list_start ||
                    nl->exw_src[1] != excluded_list || nl->exw_src[2] != excluded_numbers)) {
        hipLaunchKernelGGL(k_nl_exwin_init, dim3(grid_for(N, 256)), dim3(256), 0, s, N, nl->d_gexw);
        hipLaunchKernelGGL(k_nl_exwin, dim3(grid_for(N, 256)), dim3(256), 0, s, N,
                           excluded_list_start, excluded_list, excluded_numbers, nl->d_gexw);
        SA_LAUNCH_CHECK();
        nl->exw_src[0] = excluded_list_start;
        nl->exw_src[1] = excluded_list;
        nl->exw_src[2] = excluded_numbers;
    }
    SA_LAUNCH(256, (&nl_scatter_body), (k_nl_scatter), dim3(grid_for(N, 256)), dim3(256), 0, s, nl->d_need, N, crd,
                       atom_local, nl->d_atom_cell, nl->d_atom_slot, nl->d_cell_start,
                       have_ex ? nl->d_gexw : nullptr, nl->d_sidx, nl->d_rank, nl->d_spos,
                       nl->d_sexw);
    SA_LAUNCH_CHECK();
    {
        ScopedTimer timer("nl_search", s);
        if (nl->ortho) {
            // the batched search at two rows per wavefront: fewer scalar
            // registers, so its batched kernel runs at 6 wavefronts per SIMD
            // (at 4 rows it needed 108 VGPRs and ran at 4)
            if (g_recorder) {
                SA_LAUNCH(kSearchThreads, (&nl_search_cells_body<true, 2>), (k_nl_search_cells<true>), dim3(nl->G), dim3(kSearchThreads), 0, s, nl->d_need,
                                   nl->d_spos, nl->d_sexw, nl->d_cell_start, nl->d_ranges,
                                   nl->d_nranges, nl->max_ranges, c, r, nl->thr, excluded_list_start,
                                   excluded_list, excluded_numbers, nl->cap, nl->d_list_count,
                                   nl->d_list, nl->d_rfirst, nl->d_roff, nl->d_flags,
                                   nl->grid_length, nl->Nx, nl->Ny, nl->Nz);
            } else {
                SA_LAUNCH(kSearchThreads, (&nl_search_cells_body<true>), (k_nl_search_cells<true>), dim3(nl->G), dim3(kSearchThreads), 0, s, nl->d_need,
                                   nl->d_spos, nl->d_sexw, nl->d_cell_start, nl->d_ranges,
                                   nl->d_nranges, nl->max_ranges, c, r, nl->thr, excluded_list_start,
                                   excluded_list, excluded_numbers, nl->cap, nl->d_list_count,
                                   nl->d_list, nl->d_rfirst, nl->d_roff, nl->d_flags,
                                   nl->grid_length, nl->Nx, nl->Ny, nl->Nz);
            }
        } else {
            SA_LAUNCH(kSearchThreads, (&nl_search_cells_body<false>), (k_nl_search_cells<false>), dim3(nl->G), dim3(kSearchThreads), 0, s, nl->d_need,
                               nl->d_spos, nl->d_sexw, nl->d_cell_start, nl->d_ranges,
                               nl->d_nranges, nl->max_ranges, c, r, nl->thr, excluded_list_start,
                               excluded_list, excluded_numbers, nl->cap, nl->d_list_count,
                               nl->d_list, nl->d_rfirst, nl->d_roff, nl->d_flags,
                               nl->grid_length, nl->Nx, nl->Ny, nl->Nz);
        }
        SA_LAUNCH_CHECK();
    }
    nl->built = 1;
    return 0;
}

int sa::nl_set_prune(sponge_nl* nl, float delta) {
    if (!(delta >= 0.f)) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (delta > 0.f && !nl->d_ilist) {
        int e = 0;
        e |= dmalloc(&nl->d_need_prune, 1);
        e |= dmalloc(&nl->d_prune_crd, (size_t)3 * nl->N);
        e |= dmalloc(&nl->d_ilist_count, nl->N);
        e |= dmalloc(&nl->d_ilist, (size_t)nl->N * nl->cap);
        if (e) return SPONGE_AMD_ERROR_MALLOC_FAILED;
        const int rc = fill_async(nl->d_ilist_count, 0, nl->N, nl->stream);
        if (rc) return rc;
    }
    nl->prune_delta = delta;
    if (delta > 0.f) return fill_async(nl->d_need_prune, 1, 1, nl->stream);
    return 0;
}

int sa::nl_take_flags(sponge_nl* nl, int* flags) {
    int h = 0;
    SA_CHECK(hipMemcpyAsync(&h, nl->d_flags, sizeof(int), hipMemcpyDeviceToHost, nl->stream));
    SA_CHECK(hipStreamSynchronize(nl->stream));
    if (h) {
        int rc = fill_async(nl->d_flags, 0, 1, nl->stream);
        if (rc) return rc;
    }
    *flags = h;
    return 0;
}

extern "C" int sponge_nl_check_overflow(sponge_nl* nl, int* flags) {
    if (!nl) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    int h = 0;
    const int rc = nl_take_flags(nl, &h);
    if (rc) return rc;
    if (flags) *flags = h;
    if ((h & SPONGE_AMD_NL_STREAM_OVERFLOW) || (h && nl->opt.throw_error_when_overflow))
        return SPONGE_AMD_ERROR_OVERFLOW;
    return 0;
}

extern "C" int sponge_nl_get_options(sponge_nl* nl, sponge_nl_options* opt) {
    if (!nl || !opt) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    *opt = nl->opt;
    return 0;
}

// Check_Overflow's recovery (neighbor_list.cpp:1027-1103): a new list of the
// same atoms, cutoff, skin and box with 1.1x the limits whose flags were
// raised; the old one is released.  The new list is built on its first update.
int sa::nl_reinitialize(sponge_nl** nl, int flags) {
    sponge_nl* old = *nl;
    sponge_nl_options o = old->opt;
    if (flags & SPONGE_AMD_NL_LIST_OVERFLOW)
        o.max_neighbor_numbers = (int)(o.max_neighbor_numbers * 1.1);
    if (flags & SPONGE_AMD_NL_GRID_OVERFLOW)
        o.max_atom_in_grid_numbers = (int)(o.max_atom_in_grid_numbers * 1.1);
    sponge_nl* fresh = nullptr;
    SA_CHECK(hipStreamSynchronize(old->stream));
    const int rc = sponge_nl_initial(&fresh, old->N, old->cutoff, old->skin, &old->cell0.a11,
                                     &old->rcell0.a11, &o, old->stream);
    if (rc) return rc;
    sponge_nl_clear(old);
    *nl = fresh;
    return 0;
}

extern "C" int sponge_nl_export_half_list(sponge_nl* nl, int* counts, int* serial,
                                          int max_numbers) {
    if (!nl || !counts || !serial || max_numbers <= 0) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (!nl->built) return SPONGE_AMD_ERROR_MISSING_COMMAND;
    hipLaunchKernelGGL(k_nl_export, dim3(nl->G), dim3(256), 0, nl->stream, nl->d_sidx, nl->d_spos,
                       nl->d_cell_start, nl->d_nranges, nl->max_ranges, nl->d_rfirst, nl->d_roff,
                       nl->d_list_count, nl->d_list, nl->cap, counts, serial, max_numbers,
                       (sa_atom_group*)nullptr, nl->d_flags);
    SA_LAUNCH_CHECK();
    return 0;
}

extern "C" int sponge_nl_atom_group(sponge_nl* nl, void** d_nl) {
    if (!nl || !d_nl) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (!nl->built) return SPONGE_AMD_ERROR_MISSING_COMMAND;
    const int M = nl->opt.max_neighbor_numbers;
    if (!nl->d_atom_group && dmalloc(&nl->d_atom_group, nl->N)) return SPONGE_AMD_ERROR_MALLOC_FAILED;
    if (!nl->d_exp_serial) {
        int e = dmalloc(&nl->d_exp_serial, (size_t)nl->N * M);
        e |= dmalloc(&nl->d_exp_count, nl->N);
        if (e) return SPONGE_AMD_ERROR_MALLOC_FAILED;
    }
    hipLaunchKernelGGL(k_nl_export, dim3(nl->G), dim3(256), 0, nl->stream, nl->d_sidx, nl->d_spos,
                       nl->d_cell_start, nl->d_nranges, nl->max_ranges, nl->d_rfirst, nl->d_roff,
                       nl->d_list_count, nl->d_list, nl->cap, nl->d_exp_count, nl->d_exp_serial, M,
                       nl->d_atom_group, nl->d_flags);
    SA_LAUNCH_CHECK();
    *d_nl = nl->d_atom_group;
    return 0;
}

extern "C" int sponge_nl_atom_group_address(sponge_nl* nl, void** d_nl) {
    if (!nl || !d_nl) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (!nl->d_atom_group && dmalloc(&nl->d_atom_group, nl->N)) return SPONGE_AMD_ERROR_MALLOC_FAILED;
    *d_nl = nl->d_atom_group;
    return 0;
}

extern "C" int sponge_nl_info(sponge_nl* nl, long long* list_entries, int* grid_dims) {
    if (!nl) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (grid_dims) {
        grid_dims[0] = nl->Nx;
        grid_dims[1] = nl->Ny;
        grid_dims[2] = nl->Nz;
    }
    if (list_entries) {
        unsigned long long* d = nullptr;
        if (dmalloc(&d, 1)) return SPONGE_AMD_ERROR_MALLOC_FAILED;
        unsigned long long h = 0;
        int rc = fill_async(d, 0ull, 1, nl->stream);
        if (!rc) {
            hipLaunchKernelGGL(k_nl_sum_counts, dim3(256), dim3(256), 0, nl->stream, nl->N,
                               nl->d_list_count, d);
            hipError_t he = hipGetLastError();
            if (he == hipSuccess) he = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, nl->stream);
            if (he == hipSuccess) he = hipStreamSynchronize(nl->stream);
            rc = (int)he;
        }
        dfree(d);
        if (rc) return rc;
        *list_entries = (long long)(h / 2);  // each unordered pair is stored twice
    }
    return 0;
}

extern "C" int sponge_nl_rebuild_count(sponge_nl* nl, long long* count) {
    if (!nl || !count) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    unsigned long long h = 0;
    SA_CHECK(hipMemcpyAsync(&h, nl->d_rebuilds, sizeof(h), hipMemcpyDeviceToHost, nl->stream));
    SA_CHECK(hipStreamSynchronize(nl->stream));
    *count = (long long)h;
    return 0;
}

extern "C" int sponge_nl_clear(sponge_nl* nl) {
    if (!nl) return 0;
    (void)hipStreamSynchronize(nl->stream);
    nl_free(nl);
    delete nl;
    return 0;
}

// Gather of the current positions, charges and LJ types into the sorted
// order of the last build (the role of Copy_Crd_And_Charge_To_New_Crd,
// Lennard_Jones_force.cpp:17-27), clearing the sorted force accumulator.
__device__ __forceinline__ void nl_pack_body(int sa_block, int N, const int* __restrict__ sidx, const float4* __restrict__ spos,
                          const float* __restrict__ crd, const float* __restrict__ charge,
                          const int* __restrict__ type, float4* __restrict__ pos,
                          int* __restrict__ ptype, int* __restrict__ pgid) {
    int p = sa_block * blockDim.x + threadIdx.x;
    if (p >= N) return;
    int i = sidx[p];
    pos[p] = make_float4(crd[3 * i], crd[3 * i + 1], crd[3 * i + 2], charge[i]);
    ptype[p] = type[i];
    pgid[p] = __float_as_int(spos[p].w);
}

__global__ void k_nl_pack(int N, const int* __restrict__ sidx, const float4* __restrict__ spos,
                          const float* __restrict__ crd, const float* __restrict__ charge,
                          const int* __restrict__ type, float4* __restrict__ pos,
                          int* __restrict__ ptype, int* __restrict__ pgid) {
    nl_pack_body(blockIdx.x, N, sidx, spos, crd, charge, type, pos, ptype, pgid);
}

int sa::nl_pack_for_pairs(sponge_nl* nl, const float* crd, const float* charge, const int* type,
                          float4* pos, int* ptype, int* pgid, hipStream_t s) {
    SA_LAUNCH(256, (&nl_pack_body), (k_nl_pack), dim3(grid_for(nl->N, 256)), dim3(256), 0, s, nl->N, nl->d_sidx,
                       nl->d_spos, crd, charge, type, pos, ptype, pgid);
    SA_LAUNCH_CHECK();
    return 0;
}
