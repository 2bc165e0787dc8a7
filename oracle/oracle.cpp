// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of SPONGE's nonbonded hot path (neighbor list, LJ + PME
// direct pair sum with exclusions, PME reciprocal) used as the checker for the
// HIP implementation in sponge_amd/.  Only tests/, the smoke check and the
// CPU-baseline leg of bench.py may load it; the product never links it.
//
// Two layers live here:
//   * orc_*   : an FP32 restatement of the reference CPU path -- same formulas,
//               same per-pair operation order, same constants.  Each function
//               cites the reference file:line it follows.
//   * orc64_* : an independent FP64 oracle -- brute-force minimum-image pair
//               sets, direct sums in double, the exact Ewald k-space sum and a
//               smooth PME in double -- used to pin the restatement.
// The FFT is our own mixed-radix transform in double precision (the reference
// delegates it to FFTW / hipFFT, which are absent here).
//
// Build: oracle/Makefile (IEEE build with -ffp-contract=off, the checker) and
// a second build with the reference's release flags (-O3 -ffast-math
// -march=native, cmake/utils/common.cmake:5-7) used only as the timed CPU
// baseline.
//
// Pinning status is recorded next to each function ("pinned by: ...").

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API extern "C" __attribute__((visibility("default")))

// The checker build is serial and deterministic.  The baseline build
// (-DORC_PARALLEL) restores the reference's OpenMP structure: parallel loops
// over atoms or cells and an atomic per scattered update
// (Lennard_Jones_force.cpp:55-58, common.cpp:26-39, PM_force.cpp:765-814).
#ifdef ORC_PARALLEL
#define ORC_PRAGMA(x) _Pragma(#x)
#define ORC_ATOMIC _Pragma("omp atomic")
#else
#define ORC_PRAGMA(x)
#define ORC_ATOMIC
#endif

namespace {

// ---------------------------------------------------------------------------
// Small vector types in the reference layout (SPONGE/utils/vector.hpp:30-368).
// LTMatrix3 = {a11, a21, a22, a31, a32, a33}.
struct V3 {
    float x, y, z;
};
struct LT {
    float a11, a21, a22, a31, a32, a33;
};

inline V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 vadd(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 vscale(float s, V3 v) { return {s * v.x, s * v.y, s * v.z}; }
inline float vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// VECTOR * LTMatrix3 (vector.hpp:317-322)
inline V3 vmulm(V3 v, LT m) {
    return {v.x * m.a11 + v.y * m.a21 + v.z * m.a31, v.y * m.a22 + v.z * m.a32,
            v.z * m.a33};
}
// MultiplyTranspose (vector.hpp:311-316)
inline V3 mult_t(V3 v, LT m) {
    return {v.x * m.a11, v.x * m.a21 + v.y * m.a22,
            v.x * m.a31 + v.y * m.a32 + v.z * m.a33};
}
inline V3 vfloor(V3 v) { return {floorf(v.x), floorf(v.y), floorf(v.z)}; }
// Get_Displacement<Periodic> (SPONGE/utils/boundary.h:22-31)
inline V3 disp_pbc(V3 a, V3 b, LT cell, LT rcell) {
    V3 dr = vsub(a, b);
    V3 s = vmulm(dr, rcell);
    s = {s.x + 0.5f, s.y + 0.5f, s.z + 0.5f};
    return vsub(dr, vmulm(vfloor(s), cell));
}
// Get_Virial_From_Force_Dis (vector.hpp:323-334)
inline LT virial_fd(V3 a, V3 b) {
    LT m;
    m.a11 = a.x * b.x;
    m.a21 = a.x * b.y + a.y * b.x;
    m.a22 = a.y * b.y;
    m.a31 = a.x * b.z + a.z * b.x;
    m.a32 = a.y * b.z + a.z * b.y;
    m.a33 = a.z * b.z;
    return m;
}
inline LT lt_sub(LT a, LT b) {
    return {a.a11 - b.a11, a.a21 - b.a21, a.a22 - b.a22,
            a.a31 - b.a31, a.a32 - b.a32, a.a33 - b.a33};
}
inline LT lt_add(LT a, LT b) {
    return {a.a11 + b.a11, a.a21 + b.a21, a.a22 + b.a22,
            a.a31 + b.a31, a.a32 + b.a32, a.a33 + b.a33};
}
inline LT lt_scale(float s, LT a) {
    return {s * a.a11, s * a.a21, s * a.a22, s * a.a31, s * a.a32, s * a.a33};
}

inline V3 ld3(const float* p, int i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
inline LT ld_lt(const float* p) { return {p[0], p[1], p[2], p[3], p[4], p[5]}; }

// Get_LJ_Type (SPONGE/Lennard_Jones_force/Lennard_Jones_force.h:17-28)
inline int lj_type_pair(int a, int b) {
    int y = (b - a);
    int x = y >> 31;
    y = (y ^ x) - x;
    x = b + a;
    int z = (x + y) >> 1;
    x = (x - y) >> 1;
    return (z * (z + 1) >> 1) + x;
}

const float TWO_DIVIDED_BY_SQRT_PI = 1.1283791670218446f;  // Lennard_Jones_force.h:8
const float PME_PI = 3.1415926f;                             // PM_force.cpp:17

// ---------------------------------------------------------------------------
// FFT: own mixed-radix complex DFT in double precision (recursive
// decimation in time over the factors 2, 3, 4, 5, 7 and a direct DFT for any
// other prime).  Pinned by: numpy.fft golden vectors (tests/golden/fft_*.npz).
using cd = std::complex<double>;

void dft_rec(const cd* in, cd* out, int n, int stride, int sign, std::vector<cd>& tmp) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    int r = 0;
    for (int f : {4, 2, 3, 5, 7}) {
        if (n % f == 0) { r = f; break; }
    }
    if (r == 0) {  // prime > 7: direct DFT
        for (int k = 0; k < n; ++k) {
            cd s = 0;
            for (int j = 0; j < n; ++j) {
                double ang = sign * 2.0 * M_PI * (double)((long long)j * k % n) / n;
                s += in[(size_t)j * stride] * cd(cos(ang), sin(ang));
            }
            out[k] = s;
        }
        return;
    }
    int m = n / r;
    // sub-transforms of the r decimated sequences into out[q*m .. q*m+m)
    for (int q = 0; q < r; ++q) dft_rec(in + (size_t)q * stride, out + (size_t)q * m, m, stride * r, sign, tmp);
    std::vector<cd> buf(out, out + n);
    for (int k = 0; k < m; ++k) {
        for (int p = 0; p < r; ++p) {
            int kk = k + p * m;
            cd s = 0;
            for (int q = 0; q < r; ++q) {
                double ang = sign * 2.0 * M_PI * (double)(((long long)q * kk) % n) / n;
                s += buf[(size_t)q * m + k] * cd(cos(ang), sin(ang));
            }
            out[kk] = s;
        }
    }
}

void fft1d(cd* data, int n, int stride, int sign) {
    std::vector<cd> in(n), out(n), tmp;
    for (int i = 0; i < n; ++i) in[i] = data[(size_t)i * stride];
    dft_rec(in.data(), out.data(), n, 1, sign, tmp);
    for (int i = 0; i < n; ++i) data[(size_t)i * stride] = out[i];
}

// full complex 3D transform in place, layout [nx][ny][nz], z fastest
// (lines are independent, so the parallel loops do not change the result)
void fft3d_c(cd* a, int nx, int ny, int nz, int sign) {
#pragma omp parallel for collapse(2)
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y) fft1d(a + ((size_t)x * ny + y) * nz, nz, 1, sign);
#pragma omp parallel for collapse(2)
    for (int x = 0; x < nx; ++x)
        for (int z = 0; z < nz; ++z) fft1d(a + (size_t)x * ny * nz + z, ny, nz, sign);
#pragma omp parallel for collapse(2)
    for (int y = 0; y < ny; ++y)
        for (int z = 0; z < nz; ++z) fft1d(a + (size_t)y * nz + z, nx, ny * nz, sign);
}

// ---------------------------------------------------------------------------
// B-spline pieces of the reciprocal path (PM_force.cpp:20-64)
const float PME_Ma[4] = {(float)(1.0 / 6.0), -0.5f, 0.5f, (float)(-1.0 / 6.0)};
const float PME_Mb[4] = {0.f, 0.5f, -1.f, 0.5f};
const float PME_Mc[4] = {0.f, 0.5f, 0.f, -0.5f};
const float PME_Md[4] = {0.f, (float)(1.0 / 6.0), (float)(4.0 / 6.0), (float)(1.0 / 6.0)};
const float PME_dMa[4] = {0.5f, -1.5f, 1.5f, -0.5f};
const float PME_dMb[4] = {0.f, 1.f, -2.f, 1.f};
const float PME_dMc[4] = {0.f, 0.5f, 0.f, -0.5f};

// M_ (PM_force.cpp:29-39): cardinal B-spline of order n by recursion, float
float bspline_M(float u, int n) {
    if (n == 2) {
        if (u > 2 || u < 0) return 0;
        return 1 - fabsf(u - 1);
    }
    return u / (n - 1) * bspline_M(u, n - 1) + (n - u) / (n - 1) * bspline_M(u - 1, n - 1);
}

// getb (PM_force.cpp:42-64) with expc / divc of SPONGE/utils/fft.hpp:26-51
float getb(int k, int NFFT, int B_order) {
    float t2r = 0, t2i = 0;
    float ang0 = 2 * (B_order - 1) * PME_PI * k / NFFT;
    float rr = cosf(ang0), ri = sinf(ang0);
    for (int kk = 0; kk < (B_order - 1); kk++) {
        float ang = 2.0f * PME_PI * k / NFFT * kk;
        float cr = cosf(ang), ci = sinf(ang);
        float tf = bspline_M((float)(kk + 1), B_order);
        t2r += tf * cr;
        t2i += tf * ci;
    }
    float denom = t2r * t2r + t2i * t2i;
    float resr = (rr * t2r + ri * t2i) / denom;
    float resi = (ri * t2r - rr * t2i) / denom;
    return resr * resr + resi * resi;
}

}  // namespace

// ===========================================================================
// Scalar procedures
// ===========================================================================

// Get_Beta (PM_force.cpp:246-274): bisection on erfc(beta rc)/rc = tol with
// float variables and the double erfc.  Pinned by: known answers
// beta(8)=0.348639, beta(10)=0.275106, beta(12)=0.226635 at tol 1e-5.
ORC_API float orc_get_beta(float cutoff, float tolerance) {
    float beta, low, high, tempf;
    int ilow, ihigh;
    high = 1.0;
    ihigh = 1;
    while (1) {
        tempf = erfc(high * cutoff) / cutoff;
        if (tempf <= tolerance) break;
        high *= 2;
        ihigh++;
    }
    ihigh += 50;
    low = 0.0;
    beta = 0.0f;
    for (ilow = 1; ilow < ihigh; ilow++) {
        beta = (low + high) / 2;
        tempf = erfc(beta * cutoff) / cutoff;
        if (tempf >= tolerance)
            low = beta;
        else
            high = beta;
    }
    return beta;
}

// Check_2357_Factor (common.cpp:521-557).  Pinned by: brute-force factoring.
ORC_API int orc_check_2357(int number) {
    if (number <= 0) return 0;
    for (int f : {2, 3, 5, 7})
        while (number % f == 0) number /= f;
    return number == 1;
}

// Get_Fft_Patameter (common.cpp:639-659).  Pinned by: the survey's known
// answers 28 / 64 / 100 for L = 26.9 / 62.1 / 99 A at 1 A spacing.
ORC_API int orc_get_fft_parameter(float length) {
    int tempi = (int)ceil(length + 3) >> 2 << 2;
    if (tempi >= 60 && tempi <= 68) tempi = 64;
    else if (tempi >= 120 && tempi <= 136) tempi = 128;
    else if (tempi >= 240 && tempi <= 272) tempi = 256;
    else if (tempi >= 480 && tempi <= 544) tempi = 512;
    else if (tempi >= 960 && tempi <= 1088) tempi = 1024;
    while (1) {
        if (orc_check_2357(tempi)) return tempi;
        tempi += 4;
    }
}

// Get_Displacement<Periodic> exposed for the boundary known-answer tests
// (tests/boundary/test_boundary_policy.cpp:89-170 of the reference).
ORC_API void orc_displacement(const float* a3, const float* b3, const float* cell6,
                              const float* rcell6, float* out3) {
    V3 d = disp_pbc(ld3(a3, 0), ld3(b3, 0), ld_lt(cell6), ld_lt(rcell6));
    out3[0] = d.x;
    out3[1] = d.y;
    out3[2] = d.z;
}

// ===========================================================================
// Neighbor list (SPONGE/neighbor_list/neighbor_list.cpp)
// ===========================================================================

// Cell grid dimensions (neighbor_list.cpp:193-204): Nx = floorf(a11 / gl)
// with gl = (cutoff + skin) * 0.5f.
ORC_API void orc_nl_grid_dims(const float* cell6, float cutoff, float skin, int* dims3) {
    LT cell = ld_lt(cell6);
    float gl = (cutoff + skin) * 0.5f;
    dims3[0] = (int)floorf(cell.a11 / gl);
    dims3[1] = (int)floorf(cell.a22 / gl);
    dims3[2] = (int)floorf(cell.a33 / gl);
}

// Neighbor cells of one cell (Find_Neighbor_Grids_Scan, neighbor_list.cpp:8-69):
// ascending unique list of cells within +-2 over the 27 periodic images.
// Returns the count or -1 when more than max_out.
ORC_API int orc_nl_cell_neighbors(int grid_i, int Nx, int Ny, int Nz, const float* cell6,
                                  float grid_length, int max_out, int* out) {
    LT cell = ld_lt(cell6);
    const int64_t dxy = (int64_t)(cell.a21 / grid_length);
    const int64_t dxz = (int64_t)(cell.a31 / grid_length);
    const int64_t dyz = (int64_t)(cell.a32 / grid_length);
    const int64_t NyNz = (int64_t)Ny * Nz;
    const int64_t gx = grid_i / NyNz, gy = (grid_i % NyNz) / Nz, gz = grid_i % Nz;
    int local = 0;
    int grid_j = 0;
    for (int64_t jx = 0; jx < Nx; ++jx)
        for (int64_t jy = 0; jy < Ny; ++jy)
            for (int64_t jz = 0; jz < Nz; ++jz, ++grid_j) {
                for (int k = 0; k < 27; ++k) {
                    int dx = k / 9 - 1, dy = (k % 9) / 3 - 1, dz = k % 3 - 1;
                    int64_t tx = gx - jx + (int64_t)dx * Nx + dy * dxy + dz * dxz;
                    int64_t ty = gy - jy + (int64_t)dy * Ny + dz * dyz;
                    int64_t tz = gz - jz + (int64_t)dz * Nz;
                    int cx = (double)dy * cell.a21 + (double)dz * cell.a31 == 0.0 ? -2 : -3;
                    int cy = (double)dz * cell.a32 == 0.0 ? -2 : -3;
                    if (tx <= 2 && tx >= cx && ty <= 2 && ty >= cy && tz <= 2 && tz >= -2) {
                        if (local >= max_out) return -1;
                        out[local++] = grid_j;
                        break;
                    }
                }
            }
    return local;
}

// Cell of one atom (Put_Atom_In_Grids, neighbor_list.cpp:347-377)
static int cell_of(V3 c, LT cell, float gl, int Nx, int Ny, int Nz) {
    float k3 = floorf(c.z / cell.a33);
    c.z -= k3 * cell.a33;
    c.y -= k3 * cell.a32;
    float k2 = floorf(c.y / cell.a22);
    c.y -= k2 * cell.a22;
    c.x -= k3 * cell.a31 + k2 * cell.a21;
    c.x -= floorf(c.x / cell.a11) * cell.a11;
    int nx = (int)(c.x / gl), ny = (int)(c.y / gl), nz = (int)(c.z / gl);
    nx = nx < 0 ? 0 : (nx < Nx ? nx : Nx - 1);
    ny = ny < 0 ? 0 : (ny < Ny ? ny : Ny - 1);
    nz = nz < 0 ? 0 : (nz < Nz ? nz : Nz - 1);
    return nx * Ny * Nz + ny * Nz + nz;
}

ORC_API int orc_nl_cell_of(const float* crd3, const float* cell6, float cutoff, float skin,
                           const int* dims3) {
    float gl = (cutoff + skin) * 0.5f;
    return cell_of(ld3(crd3, 0), ld_lt(cell6), gl, dims3[0], dims3[1], dims3[2]);
}

// Exclusion test of Delete_Excluded_Atoms_Serial_In_Neighbor_List
// (neighbor_list.cpp:736-807) for a local j: j is dropped iff it lies in
// [list[first], list[last]] of i's list and equals one of its entries.
static inline bool excluded(int gi, int gj, const int* excl_start, const int* excl_list,
                            const int* excl_num) {
    if (!excl_num) return false;
    int n = excl_num[gi];
    if (n <= 0) return false;
    int s = excl_start[gi];
    if (gj < excl_list[s] || gj > excl_list[s + n - 1]) return false;
    for (int k = s; k < s + n; ++k)
        if (excl_list[k] == gj) return true;
    return false;
}

// The reference CPU build of the half list: cell binning (Put_Atom_In_Grids,
// :347-417), the CPU Find_Neighbors_Gridly (:637-733) and the exclusion
// removal (:736-807).  Output: counts[N], serial[N * max_nb] (local indices;
// order as produced by the cell walk).  Return value: bit 0 = a cell bucket
// overflowed, bit 1 = a neighbor list overflowed (entries beyond max_nb are
// dropped, as the reference does), -1 = box too small (Nx, Ny or Nz <= 2).
// Pinned by: orc_nl_bruteforce (same FP32 criterion, no cell grid) and
// hand-made edge cases in tests/test_oracle_nl.py.
ORC_API int orc_nl_build(int N, const float* crd, const float* cell6, const float* rcell6,
                         float cutoff, float skin, int max_nb, int max_in_grid,
                         const int* atom_local, const int* excl_start, const int* excl_list,
                         const int* excl_num, int* counts, int* serial) {
    LT cell = ld_lt(cell6), rcell = ld_lt(rcell6);
    const float gl = 0.5f * (cutoff + skin);
    int dims[3];
    orc_nl_grid_dims(cell6, cutoff, skin, dims);
    const int Nx = dims[0], Ny = dims[1], Nz = dims[2];
    if (Nx <= 2 || Ny <= 2 || Nz <= 2) return -1;
    const int G = Nx * Ny * Nz;
    const int MAXG = 192;  // MAX_GRID_NEIGHBORS, neighbor_list.cpp:5
    std::vector<int> ngrid(G), grids((size_t)G * MAXG);
    for (int g = 0; g < G; ++g) ngrid[g] = orc_nl_cell_neighbors(g, Nx, Ny, Nz, cell6, gl, MAXG, &grids[(size_t)g * MAXG]);
    std::vector<int> bucket_n(G, 0), bucket((size_t)G * max_in_grid);
    std::vector<V3> bucket_crd((size_t)G * max_in_grid);
    int flags = 0;
    for (int i = 0; i < N; ++i) {
        V3 c = ld3(crd, i);
        int g = cell_of(c, cell, gl, Nx, Ny, Nz);
        int k1 = bucket_n[g]++;
        if (k1 >= max_in_grid) {
            flags |= 1;
            bucket_n[g]--;
        } else {
            bucket[(size_t)g * max_in_grid + k1] = i;
            bucket_crd[(size_t)g * max_in_grid + k1] = c;
        }
        counts[i] = 0;
    }
    const float thr = gl * gl * 4.0f;
    int flag2 = 0;
    ORC_PRAGMA(omp parallel for schedule(dynamic) reduction(| : flag2))
    for (int gi = 0; gi < G; ++gi) {
        for (int jj = 0; jj < ngrid[gi]; ++jj) {
            int gj = grids[(size_t)gi * MAXG + jj];
            for (int a = 0; a < bucket_n[gi]; ++a) {
                int ai = bucket[(size_t)gi * max_in_grid + a];
                int gli = atom_local ? atom_local[ai] : ai;
                V3 ci = bucket_crd[(size_t)gi * max_in_grid + a];
                for (int b = 0; b < bucket_n[gj]; ++b) {
                    int aj = bucket[(size_t)gj * max_in_grid + b];
                    int glj = atom_local ? atom_local[aj] : aj;
                    if (glj <= gli) continue;
                    V3 dr = disp_pbc(ci, bucket_crd[(size_t)gj * max_in_grid + b], cell, rcell);
                    float dr2 = vdot(dr, dr);
                    if (dr2 < thr) {
                        if (counts[ai] < max_nb) serial[(size_t)ai * max_nb + counts[ai]++] = aj;
                        else flag2 |= 2;
                    }
                }
            }
        }
    }
    flags |= flag2;
    // exclusion removal with swap-with-last compaction
    ORC_PRAGMA(omp parallel for)
    for (int ai = 0; ai < N; ++ai) {
        int gli = atom_local ? atom_local[ai] : ai;
        int n = counts[ai];
        int* s = serial + (size_t)ai * max_nb;
        for (int k = 0; k < n; ++k) {
            int glj = atom_local ? atom_local[s[k]] : s[k];
            if (excluded(gli, glj, excl_start, excl_list, excl_num)) {
                n--;
                s[k] = s[n];
                k--;
            }
        }
        counts[ai] = n;
    }
    return flags;
}

// Independent O(N^2) list with the same FP32 criterion (j > i by global id,
// minimum-image r^2 < ((rc+skin)*0.5f)^2 * 4.0f, exclusions removed).  Output
// is sorted ascending per atom.  Returns the number of pairs or -1 when an
// atom exceeds max_nb.
ORC_API long long orc_nl_bruteforce(int N, const float* crd, const float* cell6,
                                    const float* rcell6, float cutoff, float skin, int max_nb,
                                    const int* atom_local, const int* excl_start,
                                    const int* excl_list, const int* excl_num, int* counts,
                                    int* serial) {
    LT cell = ld_lt(cell6), rcell = ld_lt(rcell6);
    const float gl = 0.5f * (cutoff + skin);
    const float thr = gl * gl * 4.0f;
    long long total = 0;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total) reduction(| : bad)
    for (int i = 0; i < N; ++i) {
        int gi = atom_local ? atom_local[i] : i;
        V3 ci = ld3(crd, i);
        int c = 0;
        for (int j = 0; j < N; ++j) {
            int gj = atom_local ? atom_local[j] : j;
            if (gj <= gi) continue;
            V3 dr = disp_pbc(ci, ld3(crd, j), cell, rcell);
            float dr2 = vdot(dr, dr);
            if (dr2 < thr && !excluded(gi, gj, excl_start, excl_list, excl_num)) {
                if (c < max_nb) serial[(size_t)i * max_nb + c] = j;
                else bad = 1;
                c++;
            }
        }
        counts[i] = c < max_nb ? c : max_nb;
        total += c;
    }
    return bad ? -1 : total;
}

// ===========================================================================
// LJ + PME direct pair sum (Lennard_Jones_force.cpp:43-137, CPU branch)
// ===========================================================================
//
// Accumulates (+=) into frc[N*3], atom_energy[N], atom_virial[N*6]; writes
// (+=) the per-atom LJ and direct-Coulomb energies.  The list is the half
// list of orc_nl_build: counts[N], serial[N*max_nb] of local indices.
// Per-pair operation order follows the reference (powf, erfcf, expf,
// norm3df = sqrtf(x*x+y*y+z*z) from common.cpp:14-22).
// Pinned by: orc64_direct (FP64 closed forms over brute-force pair sets).
ORC_API void orc_lj_direct(int local_n, int solvent_n, const int* counts, const int* serial,
                           int max_nb, const float* crd, const float* charge,
                           const int* lj_type, const float* LJ_A, const float* LJ_B,
                           const float* cell6, const float* rcell6, float cutoff, float pme_beta,
                           int need_energy, int need_virial, float* frc, float* atom_energy,
                           float* atom_virial, float* atom_direct_cf_energy,
                           float* atom_LJ_ene) {
    LT cell = ld_lt(cell6), rcell = ld_lt(rcell6);
    ORC_PRAGMA(omp parallel for schedule(dynamic))
    for (int atom_i = 0; atom_i < local_n - solvent_n; ++atom_i) {
        V3 frc_record = {0, 0, 0};
        LT virial = {0, 0, 0, 0, 0, 0};
        float energy_lj = 0, energy_coulomb = 0;
        V3 r1 = ld3(crd, atom_i);
        float q1 = charge[atom_i];
        int t1 = lj_type[atom_i];
        const int* s = serial + (size_t)atom_i * max_nb;
        for (int j = 0; j < counts[atom_i]; ++j) {
            int atom_j = s[j];
            float ij_factor = atom_j < local_n ? 1.0f : 0.5f;
            V3 r2 = ld3(crd, atom_j);
            V3 dr = disp_pbc(r2, r1, cell, rcell);
            float dr_abs = sqrtf(dr.x * dr.x + dr.y * dr.y + dr.z * dr.z);
            if (dr_abs < cutoff) {
                int tp = lj_type_pair(t1, lj_type[atom_j]);
                float A = LJ_A[tp], B = LJ_B[tp];
                float q2 = charge[atom_j];
                // Get_LJ_Force (Lennard_Jones_force.h:52-56)
                float frc_abs = (B - A * powf(dr_abs, -6.0f)) * powf(dr_abs, -8.0f);
                // Get_Direct_Coulomb_Force (Lennard_Jones_force.h:67-74)
                float beta_dr = pme_beta * dr_abs;
                float frc_cf_abs = q1 * q2 * powf(dr_abs, -3.0f) *
                                   (beta_dr * TWO_DIVIDED_BY_SQRT_PI * expf(-beta_dr * beta_dr) +
                                    erfcf(beta_dr));
                frc_abs = frc_abs - frc_cf_abs;
                V3 frc_lin = vscale(frc_abs, dr);
                frc_record = vadd(frc_record, frc_lin);
                if (atom_j < local_n) {
                    ORC_ATOMIC frc[3 * atom_j] -= frc_lin.x;
                    ORC_ATOMIC frc[3 * atom_j + 1] -= frc_lin.y;
                    ORC_ATOMIC frc[3 * atom_j + 2] -= frc_lin.z;
                }
                if (need_virial) virial = lt_sub(virial, lt_scale(ij_factor, virial_fd(frc_lin, dr)));
                if (need_energy) {
                    float dr_6 = powf(dr_abs, -6.0f);
                    energy_lj += ij_factor * ((0.083333333f * A * dr_6 - 0.166666667f * B) * dr_6);
                    energy_coulomb += ij_factor * (q1 * q2 * erfcf(pme_beta * dr_abs) / dr_abs);
                }
            }
        }
        ORC_ATOMIC frc[3 * atom_i] += frc_record.x;
        ORC_ATOMIC frc[3 * atom_i + 1] += frc_record.y;
        ORC_ATOMIC frc[3 * atom_i + 2] += frc_record.z;
        if (need_energy) {
            atom_energy[atom_i] += energy_lj + energy_coulomb;
            if (atom_LJ_ene) atom_LJ_ene[atom_i] += energy_lj;
            if (atom_direct_cf_energy) atom_direct_cf_energy[atom_i] += energy_coulomb;
        }
        if (need_virial) {
            float* v = atom_virial + 6 * atom_i;
            v[0] += virial.a11; v[1] += virial.a21; v[2] += virial.a22;
            v[3] += virial.a31; v[4] += virial.a32; v[5] += virial.a33;
        }
    }
}

// LJ soft core with direct Coulomb, restating
// Lennard_Jones_And_Direct_Coulomb_Soft_Core_CUDA (SPONGE/Lennard_Jones_force/
// LJ_soft_core.cpp:195-440) and its pair functions (LJ_soft_core.h:33-214):
// state A (types typeA, tables AA/AB) weighted 1 - lambda, state B (typeB,
// BA/BB) weighted lambda; a pair whose repulsion vanishes in exactly one
// state takes soft-core distances r_A = (r^6 + alpha lambda^p sigma_A)^(1/6),
// r_B = (r^6 + alpha (1 - lambda)^p sigma_B)^(1/6) with sigma = A / (2B)
// (sigma6 when A or B is zero, at least sigma6_min).  Outputs are
// accumulated: forces, per-atom energy (LJ + Coulomb), LJ and direct-Coulomb
// energies, virial, and the dU/dlambda totals du[0] (LJ), du[1] (direct).
// Pinned by: the plain kernel (identical states), the end points lambda = 0
// / 1 with a dummy state in closed form, and dU/dlambda against central
// differences of the FP64 energy (tests/test_oracle_softcore.py).
ORC_API void orc_ljsc_direct(int local_n, const int* counts, const int* serial, int max_nb,
                             const float* crd, const float* charge, const float* charge_BA,
                             const int* typeA, const int* typeB, const float* AA, const float* AB,
                             const float* BA, const float* BB, const float* cell6,
                             const float* rcell6, float cutoff, float pme_beta, float lambda,
                             float alpha, float p, float sigma6, float sigma6_min, int need_energy,
                             int need_virial, int need_du, float* frc, float* atom_energy,
                             float* atom_virial, float* atom_direct_cf_energy, float* atom_LJ_ene,
                             double* du) {
    LT cell = ld_lt(cell6), rcell = ld_lt(rcell6);
    const float lambda_ = 1.0f - lambda;
    auto lj_force = [](float r, float A, float B) { return (B - A * powf(r, -6.0f)) * powf(r, -8.0f); };
    auto lj_energy = [](float r, float A, float B) {
        float r6 = powf(r, -6.0f);
        return (0.083333333f * A * r6 - 0.166666667f * B) * r6;
    };
    auto cf_force = [&](float qq, float r) {
        float br = pme_beta * r;
        return qq * powf(r, -3.0f) * (br * TWO_DIVIDED_BY_SQRT_PI * expf(-br * br) + erfcf(br));
    };
    auto cf_energy = [&](float qq, float r) { return qq * erfcf(pme_beta * r) / r; };
    auto sc_sigma = [&](float A, float B) {
        return (A == 0 || B == 0) ? sigma6 : fmaxf(0.5f * A / B, sigma6_min);
    };
    auto sc_dist = [&](float sigma, float r, float lam) {
        float alp = alpha * powf(lam, p);
        return powf(powf(r, 6.0f) + alp * sigma, 0.16666667f);
    };
    auto sc_lj_force = [](float r, float rs, float A, float B) {
        return powf(r, 4.0f) * (B - A * powf(rs, -6.0f)) * powf(rs, -12.0f);
    };
    auto sc_cf_force = [&](float qq, float r, float rs) {
        float b = rs * pme_beta;
        return qq * powf(r, 4.0f) *
               (expf(-b * b) * TWO_DIVIDED_BY_SQRT_PI * pme_beta + erfcf(b) / rs) * powf(rs, -6.0f);
    };
    auto sc_du = [&](float F, float sigma, float rs, float lam) {
        return 0.16666667f * p * alpha * (1 - lam) * powf(lam + FLT_MIN, p - 1.0f) * F *
               powf(rs, -4.0f) * sigma;
    };
    double du_lj = 0, du_cf = 0;
    ORC_PRAGMA(omp parallel for schedule(dynamic) reduction(+ : du_lj, du_cf))
    for (int atom_i = 0; atom_i < local_n; ++atom_i) {
        V3 frc_record = {0, 0, 0};
        LT virial = {0, 0, 0, 0, 0, 0};
        float energy_lj = 0, energy_coulomb = 0, d_lj = 0, d_cf = 0;
        V3 r1 = ld3(crd, atom_i);
        const float q1 = charge[atom_i], qba1 = charge_BA ? charge_BA[atom_i] : 0.f;
        const int* s = serial + (size_t)atom_i * max_nb;
        for (int j = 0; j < counts[atom_i]; ++j) {
            int atom_j = s[j];
            V3 dr = disp_pbc(ld3(crd, atom_j), r1, cell, rcell);
            float dr_abs = sqrtf(dr.x * dr.x + dr.y * dr.y + dr.z * dr.z);
            if (!(dr_abs < cutoff)) continue;
            const float q2 = charge[atom_j], qba2 = charge_BA ? charge_BA[atom_j] : 0.f;
            const float qq = q1 * q2;
            const int ta = lj_type_pair(typeA[atom_i], typeA[atom_j]);
            const int tb = lj_type_pair(typeB[atom_i], typeB[atom_j]);
            const float Aa = AA[ta], Ab = AB[ta], Ba = BA[tb], Bb = BB[tb];
            float frc_abs, e_lj = 0, e_cf = 0;
            if (Ba * Aa != 0 || Ba + Aa == 0) {
                frc_abs = lambda_ * lj_force(dr_abs, Aa, Ab) + lambda * lj_force(dr_abs, Ba, Bb);
                frc_abs = frc_abs - cf_force(qq, dr_abs);
                if (need_energy) {
                    e_cf = cf_energy(qq, dr_abs);
                    e_lj = lambda_ * lj_energy(dr_abs, Aa, Ab) + lambda * lj_energy(dr_abs, Ba, Bb);
                }
                if (need_du) {
                    d_lj += lj_energy(dr_abs, Ba, Bb) - lj_energy(dr_abs, Aa, Ab);
                    d_cf += erfcf(pme_beta * dr_abs) * (qba2 * q1 + q2 * qba1) / dr_abs;
                }
            } else {
                const float sA = sc_sigma(Aa, Ab), sB = sc_sigma(Ba, Bb);
                const float rA = sc_dist(sA, dr_abs, lambda), rB = sc_dist(sB, dr_abs, lambda_);
                frc_abs = lambda_ * sc_lj_force(dr_abs, rA, Aa, Ab) + lambda * sc_lj_force(dr_abs, rB, Ba, Bb);
                frc_abs = frc_abs - (lambda_ * sc_cf_force(qq, dr_abs, rA) + lambda * sc_cf_force(qq, dr_abs, rB));
                if (need_energy) {
                    e_cf = lambda_ * cf_energy(qq, rA) + lambda * cf_energy(qq, rB);
                    e_lj = lambda_ * lj_energy(rA, Aa, Ab) + lambda * lj_energy(rB, Ba, Bb);
                }
                if (need_du) {
                    d_lj += lj_energy(rB, Ba, Bb) - lj_energy(rA, Aa, Ab);
                    d_lj += sc_du(lj_force(rA, Aa, Ab), sA, rA, lambda) -
                            sc_du(lj_force(rB, Ba, Bb), sB, rB, lambda_);
                    d_cf += cf_energy(qq, rB) - cf_energy(qq, rA);
                    d_cf += sc_du(cf_force(qq, rB), sB, rB, lambda_) - sc_du(cf_force(qq, rA), sA, rA, lambda);
                    const float eB = erfcf(pme_beta * rB) * (qba2 * q1 + q2 * qba1) / rB;
                    const float eA = erfcf(pme_beta * rA) * (qba2 * q1 + q2 * qba1) / rA;
                    d_cf += lambda * eB + lambda_ * eA;
                }
            }
            V3 frc_lin = vscale(frc_abs, dr);
            frc_record = vadd(frc_record, frc_lin);
            ORC_ATOMIC frc[3 * atom_j] -= frc_lin.x;
            ORC_ATOMIC frc[3 * atom_j + 1] -= frc_lin.y;
            ORC_ATOMIC frc[3 * atom_j + 2] -= frc_lin.z;
            if (need_virial) virial = lt_sub(virial, virial_fd(frc_lin, dr));
            energy_lj += e_lj;
            energy_coulomb += e_cf;
        }
        ORC_ATOMIC frc[3 * atom_i] += frc_record.x;
        ORC_ATOMIC frc[3 * atom_i + 1] += frc_record.y;
        ORC_ATOMIC frc[3 * atom_i + 2] += frc_record.z;
        if (need_energy) {
            atom_energy[atom_i] += energy_lj + energy_coulomb;
            if (atom_LJ_ene) atom_LJ_ene[atom_i] += energy_lj;
            if (atom_direct_cf_energy) atom_direct_cf_energy[atom_i] += energy_coulomb;
        }
        if (need_virial) {
            float* v = atom_virial + 6 * atom_i;
            v[0] += virial.a11; v[1] += virial.a21; v[2] += virial.a22;
            v[3] += virial.a31; v[4] += virial.a32; v[5] += virial.a33;
        }
        du_lj += d_lj;
        du_cf += d_cf;
    }
    if (need_du) {
        du[0] += du_lj;
        du[1] += du_cf;
    }
}

// Pair statistics for the roofline: in-cutoff pairs and list entries.
ORC_API void orc_pair_counts(int N, const int* counts, const int* serial, int max_nb,
                             const float* crd, const float* cell6, const float* rcell6,
                             float cutoff, long long* n_in, long long* n_list) {
    LT cell = ld_lt(cell6), rcell = ld_lt(rcell6);
    long long a = 0, b = 0;
#pragma omp parallel for reduction(+ : a, b)
    for (int i = 0; i < N; ++i) {
        V3 r1 = ld3(crd, i);
        for (int j = 0; j < counts[i]; ++j) {
            V3 dr = disp_pbc(ld3(crd, serial[(size_t)i * max_nb + j]), r1, cell, rcell);
            float d = sqrtf(dr.x * dr.x + dr.y * dr.y + dr.z * dr.z);
            b++;
            if (d < cutoff) a++;
        }
    }
    *n_in = a;
    *n_list = b;
}

// LJ long-range correction factor (Lennard_Jones_force.cpp:208-235): C6 sum
// over the type histogram in double, then the float scale.  Pinned by: the
// closed form -(2/3) pi / rc^3 / 6 * sum_ab n_a n_b B_ab in tests.
ORC_API float orc_lj_long_range_factor(int N, int n_types, const int* lj_type,
                                       const float* LJ_B, float cutoff) {
    std::vector<long long> count(n_types, 0);
    for (int i = 0; i < N; ++i) count[lj_type[i]]++;
    double c6 = 0;
    for (int a = 0; a < n_types; ++a) {
        if (!count[a]) continue;
        double inner = 0;
        for (int b = 0; b < n_types; ++b) {
            if (!count[b]) continue;
            inner += (double)count[b] * (double)LJ_B[lj_type_pair(a, b)];
        }
        c6 += (double)count[a] * inner;
    }
    float f = (float)c6;
    f *= -2.0f / 3.0f * 3.1415926535897932f / cutoff / cutoff / cutoff / 6.0f;
    return f;
}

// ===========================================================================
// PME excluded-pair correction (PM_force.cpp:1745-1817, factor 1 on one rank)
// Pinned by: orc64_excluded (closed form -qq erf(beta r)/r in double).
// ===========================================================================
ORC_API void orc_pme_excluded(int N, const float* crd, const float* charge, const float* cell6,
                              const float* rcell6, float beta, const int* excl_start,
                              const int* excl_list, const int* excl_num, int need_energy,
                              int need_virial, float* frc, float* atom_ene, float* this_ene,
                              float* atom_virial) {
    LT cell = ld_lt(cell6), rcell = ld_lt(rcell6);
    const float factor = 1.0f;
    for (int i = 0; i < N; ++i) {
        int n = excl_num[i];
        if (n <= 0) continue;
        int s = excl_start[i];
        float qi = charge[i];
        V3 r1 = ld3(crd, i);
        V3 frc_record = {0, 0, 0};
        LT vir = {0, 0, 0, 0, 0, 0};
        float ene_lin = 0.;
        for (int k = s; k < s + n; ++k) {
            int j = excl_list[k];
            V3 r2 = ld3(crd, j);
            float qj = charge[j];
            V3 dr = disp_pbc(r2, r1, cell, rcell);
            float dr2 = dr.x * dr.x + dr.y * dr.y + dr.z * dr.z;
            float dr_abs = sqrtf(dr2);
            float beta_dr = beta * dr_abs;
            float frc_abs = beta_dr * TWO_DIVIDED_BY_SQRT_PI * expf(-beta_dr * beta_dr) + erfcf(beta_dr);
            frc_abs = (frc_abs - 1.f) / dr2 / dr_abs;
            frc_abs = -qi * qj * frc_abs;
            V3 frc_lin = vscale(frc_abs, dr);
            frc[3 * j] -= frc_lin.x;
            frc[3 * j + 1] -= frc_lin.y;
            frc[3 * j + 2] -= frc_lin.z;
            frc_record = vadd(frc_record, frc_lin);
            if (need_energy) ene_lin -= factor * qi * qj * erff(beta_dr) / dr_abs;
            if (need_virial) vir = lt_sub(vir, lt_scale(factor, virial_fd(frc_lin, dr)));
        }
        frc[3 * i] += frc_record.x;
        frc[3 * i + 1] += frc_record.y;
        frc[3 * i + 2] += frc_record.z;
        if (need_energy) {
            atom_ene[i] += ene_lin;
            if (this_ene) this_ene[i] = ene_lin;
        }
        if (need_virial) {
            float* v = atom_virial + 6 * i;
            v[0] += vir.a11; v[1] += vir.a21; v[2] += vir.a22;
            v[3] += vir.a31; v[4] += vir.a32; v[5] += vir.a33;
        }
    }
}

// ===========================================================================
// PME reciprocal (PM_force.cpp:295-654 set-up, 699-1168 per step)
// ===========================================================================

struct orc_pme {
    int fx, fy, fz, nfft, nall;
    float beta, volume, neutralizing_factor;
    LT cell, rcell;
    std::vector<float> BC;      // [fx][fy][fz/2+1]
    std::vector<float> vBC;     // 6 per entry
    std::vector<float> Q, FBCFQ;
    std::vector<unsigned> uxyz; // 3 per atom, 1<<30 = not yet set
    std::vector<float> frxyz;
    std::vector<int> near;      // 64 per atom
    std::vector<float> force_backup;
    std::vector<cd> FQ;         // half spectrum [fx][fy][fz/2+1]
    int n_atoms;
};

// Initial's BC / virial_BC tables (PM_force.cpp:545-611).
ORC_API orc_pme* orc_pme_create(int N, int fx, int fy, int fz, float beta, const float* cell6,
                                const float* rcell6) {
    orc_pme* p = new orc_pme();
    p->n_atoms = N;
    p->fx = fx; p->fy = fy; p->fz = fz;
    p->nall = fx * fy * fz;
    p->nfft = fx * fy * (fz / 2 + 1);
    p->beta = beta;
    p->cell = ld_lt(cell6);
    p->rcell = ld_lt(rcell6);
    p->volume = p->cell.a11 * p->cell.a22 * p->cell.a33;
    p->neutralizing_factor = -0.5 * 3.1415926535897932f / (beta * beta * p->volume);
    p->BC.assign(p->nfft, 0.f);
    p->vBC.assign((size_t)p->nfft * 6, 0.f);
    std::vector<float> B1(fx), B2(fy), B3(fz);
    for (int k = 0; k < fx; ++k) B1[k] = getb(k, fx, 4);
    for (int k = 0; k < fy; ++k) B2[k] = getb(k, fy, 4);
    for (int k = 0; k < fz; ++k) B3[k] = getb(k, fz, 4);
    float mprefactor = PME_PI * PME_PI / beta / beta;
    for (int kx = 0; kx < fx; ++kx) {
        float kxrp = kx > fx / 2 ? kx - fx : kx;
        for (int ky = 0; ky < fy; ++ky) {
            float kyrp = ky > fy / 2 ? ky - fy : ky;
            for (int kz = 0; kz <= fz / 2; ++kz) {
                float kzrp = kz;
                V3 m = mult_t({kxrp, kyrp, kzrp}, p->rcell);
                float msq = vdot(m, m);
                int index = kx * fy * (fz / 2 + 1) + ky * (fz / 2 + 1) + kz;
                float bc;
                LT v;
                if (kx + ky + kz == 0) {
                    bc = 0;
                    v = {0, 0, 0, 0, 0, 0};
                } else {
                    bc = (float)1.0 / PME_PI / msq * expf(-mprefactor * msq) / p->volume;
                    float t = 2 / msq * (1 + mprefactor * msq);
                    v.a11 = 1 - t * m.x * m.x;
                    v.a21 = 0 - t * m.y * m.x;
                    v.a22 = 1 - t * m.y * m.y;
                    v.a31 = 0 - t * m.z * m.x;
                    v.a32 = 0 - t * m.z * m.y;
                    v.a33 = 1 - t * m.z * m.z;
                }
                float bc0 = B1[kx] * B2[ky] * B3[kz];
                bc *= bc0;
                v = lt_scale(0.5f * bc, v);
                p->BC[index] = bc;
                float* o = &p->vBC[(size_t)index * 6];
                o[0] = v.a11; o[1] = v.a21; o[2] = v.a22; o[3] = v.a31; o[4] = v.a32; o[5] = v.a33;
            }
        }
    }
    p->Q.assign(p->nall, 0.f);
    p->FBCFQ.assign(p->nall, 0.f);
    p->FQ.assign(p->nfft, 0);
    p->uxyz.assign((size_t)3 * N, 1u << 30);
    p->frxyz.assign((size_t)3 * N, 0.f);
    p->near.assign((size_t)64 * N, 0);
    p->force_backup.assign((size_t)3 * N, 0.f);
    return p;
}

ORC_API void orc_pme_destroy(orc_pme* p) { delete p; }

ORC_API void orc_pme_tables(orc_pme* p, float* BC, float* vBC) {
    if (BC) memcpy(BC, p->BC.data(), sizeof(float) * p->nfft);
    if (vBC) memcpy(vBC, p->vBC.data(), sizeof(float) * 6 * p->nfft);
}

static void sum_to_float_via_double(const float* a, int n, float* out) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += a[i];
    *out = (float)s;
}

// PME_Reciprocal_Force_With_Energy_And_Virial (PM_force.cpp:1070-1168).
// frc (+=), virial6 (+= at index 0 semantics: a single LTMatrix3), potential
// (+= self + reciprocal).  Also reports the pieces in out_energies[3] =
// {reciprocal, self (incl. neutralizing), unused}.
// Pinned by: orc64_ewald_recip (exact Ewald sum, FP64) as the grid refines,
// numpy.fft for the transform, and partition-of-unity checks of the weights.
namespace {
// wall time of the pieces of the last orc_pme_reciprocal evaluation:
// [0] base points + spread, [1] R2C + BC + C2R, [2] gather (the CPU
// baseline's per-module split)
double g_recip_seconds[3] = {0, 0, 0};
inline double wall_now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
}  // namespace

ORC_API void orc_recip_seconds(double* out3) {
    for (int k = 0; k < 3; ++k) out3[k] = g_recip_seconds[k];
}

ORC_API void orc_pme_reciprocal(orc_pme* p, const float* crd, const float* charge, float* frc,
                                int need_virial, int need_energy, float* virial6,
                                float* potential, int step, int update_interval,
                                float* out_energies) {
    const int N = p->n_atoms, fx = p->fx, fy = p->fy, fz = p->fz;
    const int PME_Nin = fy * fz;
    float reciprocal_ene = 0, self_ene = 0;
    if (step % update_interval == 0) {
        std::fill(p->Q.begin(), p->Q.end(), 0.f);
        const double t_spread0 = wall_now();
        // PME_Atom_Near (PM_force.cpp:699-762)
        ORC_PRAGMA(omp parallel for)
        for (int atom = 0; atom < N; ++atom) {
            p->force_backup[3 * atom] = p->force_backup[3 * atom + 1] = p->force_backup[3 * atom + 2] = 0.f;
            V3 frac = vmulm(ld3(crd, atom), p->rcell);
            frac = vsub(frac, vfloor(frac));
            if (!std::isfinite(frac.x) || !std::isfinite(frac.y) || !std::isfinite(frac.z)) frac = {0, 0, 0};
            int u[3];
            float fr[3];
            float fc[3] = {frac.x, frac.y, frac.z};
            int dims[3] = {fx, fy, fz};
            for (int d = 0; d < 3; ++d) {
                fc[d] *= dims[d];
                int t = (int)fc[d];
                t = t < 0 ? 0 : (t < dims[d] ? t : dims[d] - 1);
                u[d] = t;
                fr[d] = fc[d] - t;
                fr[d] = fr[d] - floorf(fr[d]);
            }
            p->frxyz[3 * atom] = fr[0];
            p->frxyz[3 * atom + 1] = fr[1];
            p->frxyz[3 * atom + 2] = fr[2];
            unsigned* uu = &p->uxyz[3 * atom];
            if ((unsigned)u[0] != uu[0] || (unsigned)u[1] != uu[1] || (unsigned)u[2] != uu[2]) {
                uu[0] = u[0]; uu[1] = u[1]; uu[2] = u[2];
                int* tn = &p->near[(size_t)64 * atom];
                for (int k = 0; k < 64; ++k) {
                    int kx = k / 16, ky = (k - 16 * kx) / 4, kz = k % 4;
                    kx = u[0] - kx; if (kx < 0) kx += fx; if (kx >= fx) kx -= fx;
                    ky = u[1] - ky; if (ky < 0) ky += fy; if (ky >= fy) ky -= fy;
                    kz = u[2] - kz; if (kz < 0) kz += fz; if (kz >= fz) kz -= fz;
                    tn[k] = kx * PME_Nin + ky * fz + kz;
                }
            }
        }
        // PME_Q_Spread (PM_force.cpp:765-814)
        ORC_PRAGMA(omp parallel for)
        for (int atom = 0; atom < N; ++atom) {
            const int* tn = &p->near[(size_t)64 * atom];
            float fxr = p->frxyz[3 * atom], fyr = p->frxyz[3 * atom + 1], fzr = p->frxyz[3 * atom + 2];
            float q = charge[atom];
            for (int k = 0; k < 64; ++k) {
                int kx = k / 16;
                float t = fxr, t2 = t * t;
                float w = PME_Ma[kx] * t * t2 + PME_Mb[kx] * t2 + PME_Mc[kx] * t + PME_Md[kx];
                float tq = q * w;
                kx = (k - kx * 16) / 4;
                t = fyr; t2 = t * t;
                w = PME_Ma[kx] * t * t2 + PME_Mb[kx] * t2 + PME_Mc[kx] * t + PME_Md[kx];
                tq = tq * w;
                kx = k % 4;
                t = fzr; t2 = t * t;
                w = PME_Ma[kx] * t * t2 + PME_Mb[kx] * t2 + PME_Mc[kx] * t + PME_Md[kx];
                tq = tq * w;
                unsigned idx = (unsigned)tn[k];
                if (idx < (unsigned)p->nall) {
                    ORC_ATOMIC p->Q[idx] += tq;
                }
            }
        }
        const double t_fft0 = wall_now();
        // R2C (unnormalised, half spectrum along z)
        std::vector<cd> full((size_t)p->nall);
        for (int i = 0; i < p->nall; ++i) full[i] = p->Q[i];
        fft3d_c(full.data(), fx, fy, fz, -1);
        const int fzc = fz / 2 + 1;
        for (int x = 0; x < fx; ++x)
            for (int y = 0; y < fy; ++y)
                for (int z = 0; z < fzc; ++z)
                    p->FQ[((size_t)x * fy + y) * fzc + z] = cd((float)full[((size_t)x * fy + y) * fz + z].real(),
                                                               (float)full[((size_t)x * fy + y) * fz + z].imag());
        // PME_Sum_Virial (PM_force.cpp:1024-1068), reference weights
        if (need_virial) {
            double v[6] = {0, 0, 0, 0, 0, 0};
            for (int index = 0; index < p->nfft; ++index) {
                int nz = index % fzc;
                float factor = (nz == 0 || nz == fzc - 1) ? 0.5f : 1.0f;
                float re = (float)p->FQ[index].real(), im = (float)p->FQ[index].imag();
                float s = factor * (re * re + im * im);
                for (int c = 0; c < 6; ++c) v[c] -= (double)(s * p->vBC[(size_t)index * 6 + c]);
            }
            for (int c = 0; c < 6; ++c) virial6[c] += (float)v[c];
        }
        // PME_BCFQ (PM_force.cpp:817-826)
        for (int index = 0; index < p->nfft; ++index) {
            float re = (float)p->FQ[index].real(), im = (float)p->FQ[index].imag();
            p->FQ[index] = cd(re * p->BC[index], im * p->BC[index]);
        }
        // C2R: Hermitian completion then inverse complex transform
        for (int x = 0; x < fx; ++x)
            for (int y = 0; y < fy; ++y)
                for (int z = 0; z < fz; ++z) {
                    cd val;
                    if (z < fzc) val = p->FQ[((size_t)x * fy + y) * fzc + z];
                    else {
                        int xm = (fx - x) % fx, ym = (fy - y) % fy, zm = fz - z;
                        val = std::conj(p->FQ[((size_t)xm * fy + ym) * fzc + zm]);
                    }
                    full[((size_t)x * fy + y) * fz + z] = val;
                }
        fft3d_c(full.data(), fx, fy, fz, +1);
        for (int i = 0; i < p->nall; ++i) p->FBCFQ[i] = (float)full[i].real();
        const double t_gather0 = wall_now();
        // PME_Final (PM_force.cpp:829-893)
        ORC_PRAGMA(omp parallel for)
        for (int atom = 0; atom < N; ++atom) {
            const int* tn = &p->near[(size_t)64 * atom];
            float q = charge[atom];
            float frx = p->frxyz[3 * atom], fry = p->frxyz[3 * atom + 1], frz = p->frxyz[3 * atom + 2];
            V3 nv = {0, 0, 0};
            for (int k = 0; k < 64; ++k) {
                unsigned idx = (unsigned)tn[k];
                if (idx >= (unsigned)p->nall) continue;
                float dqf = -p->FBCFQ[idx] * q;
                int kx = k / 16;
                float t = frx, t2 = t * t;
                float wx = PME_Ma[kx] * t * t2 + PME_Mb[kx] * t2 + PME_Mc[kx] * t + PME_Md[kx];
                float dx = PME_dMa[kx] * t2 + PME_dMb[kx] * t + PME_dMc[kx];
                kx = (k - kx * 16) / 4;
                t = fry; t2 = t * t;
                float wy = PME_Ma[kx] * t * t2 + PME_Mb[kx] * t2 + PME_Mc[kx] * t + PME_Md[kx];
                float dy = PME_dMa[kx] * t2 + PME_dMb[kx] * t + PME_dMc[kx];
                kx = k % 4;
                t = frz; t2 = t * t;
                float wz = PME_Ma[kx] * t * t2 + PME_Mb[kx] * t2 + PME_Mc[kx] * t + PME_Md[kx];
                float dz = PME_dMa[kx] * t2 + PME_dMb[kx] * t + PME_dMc[kx];
                V3 dq = {dx * wy * wz * fx, dy * wx * wz * fy, dz * wx * wy * fz};
                dq = vscale(dqf, mult_t(dq, p->rcell));
                nv = vadd(nv, dq);
            }
            p->force_backup[3 * atom] += nv.x;
            p->force_backup[3 * atom + 1] += nv.y;
            p->force_backup[3 * atom + 2] += nv.z;
        }
        const double t_end = wall_now();
        g_recip_seconds[0] = t_fft0 - t_spread0;
        g_recip_seconds[1] = t_gather0 - t_fft0;
        g_recip_seconds[2] = t_end - t_gather0;
        // device_add_force (PM_force.cpp:1014-1022)
        for (int i = 0; i < 3 * N; ++i) frc[i] = frc[i] + (float)update_interval * p->force_backup[i];
    }
    if (need_energy) {
        // PME_Energy_Product then x0.5 (PM_force.cpp:896-919, 1142-1147);
        // the sum is taken in double here.
        double e = 0;
        for (int i = 0; i < p->nall; ++i) e += (double)p->Q[i] * (double)p->FBCFQ[i];
        reciprocal_ene = (float)e * 0.5f;
        // self and neutralizing (PM_force.cpp:1149-1163)
        std::vector<float> qsq(N);
        for (int i = 0; i < N; ++i) qsq[i] = charge[i] * charge[i];
        float sq, sumq;
        sum_to_float_via_double(qsq.data(), N, &sq);
        self_ene = sq * (-p->beta / sqrtf(PME_PI));
        sum_to_float_via_double(charge, N, &sumq);
        self_ene += p->neutralizing_factor * sumq * sumq;
        if (potential) potential[0] += self_ene + reciprocal_ene;
    }
    if (out_energies) {
        out_energies[0] = reciprocal_ene;
        out_energies[1] = self_ene;
    }
}

// Access to the intermediate grids for kernel-by-kernel parity.
ORC_API void orc_pme_grids(orc_pme* p, float* Q, float* FBCFQ, float* FQ_interleaved,
                           int* near, float* frxyz) {
    if (Q) memcpy(Q, p->Q.data(), sizeof(float) * p->nall);
    if (FBCFQ) memcpy(FBCFQ, p->FBCFQ.data(), sizeof(float) * p->nall);
    if (FQ_interleaved)
        for (int i = 0; i < p->nfft; ++i) {
            FQ_interleaved[2 * i] = (float)p->FQ[i].real();
            FQ_interleaved[2 * i + 1] = (float)p->FQ[i].imag();
        }
    if (near) memcpy(near, p->near.data(), sizeof(int) * 64 * p->n_atoms);
    if (frxyz) memcpy(frxyz, p->frxyz.data(), sizeof(float) * 3 * p->n_atoms);
}

// ===========================================================================
// FFT entry points (double in, double out).  Pinned by numpy.fft fixtures.
// ===========================================================================
ORC_API void orc_fft3d_r2c(int nx, int ny, int nz, const double* in, double* out_interleaved) {
    std::vector<cd> a((size_t)nx * ny * nz);
    for (size_t i = 0; i < a.size(); ++i) a[i] = in[i];
    fft3d_c(a.data(), nx, ny, nz, -1);
    int nzc = nz / 2 + 1;
    for (int x = 0; x < nx; ++x)
        for (int y = 0; y < ny; ++y)
            for (int z = 0; z < nzc; ++z) {
                cd v = a[((size_t)x * ny + y) * nz + z];
                size_t o = ((size_t)x * ny + y) * nzc + z;
                out_interleaved[2 * o] = v.real();
                out_interleaved[2 * o + 1] = v.imag();
            }
}

ORC_API void orc_fft1d(int n, double* data_interleaved, int sign) {
    std::vector<cd> a(n);
    for (int i = 0; i < n; ++i) a[i] = cd(data_interleaved[2 * i], data_interleaved[2 * i + 1]);
    fft1d(a.data(), n, 1, sign);
    for (int i = 0; i < n; ++i) {
        data_interleaved[2 * i] = a[i].real();
        data_interleaved[2 * i + 1] = a[i].imag();
    }
}

// ===========================================================================
// FP64 oracle
// ===========================================================================

static inline void disp64(const double* a, const double* b, const double* box, double* d) {
    for (int k = 0; k < 3; ++k) {
        double x = a[k] - b[k];
        d[k] = x - floor(x / box[k] + 0.5) * box[k];
    }
}

// Direct space in double over all minimum-image pairs i<j with r < cutoff,
// minus the excluded pairs: LJ (A/12 r^-12 - B/6 r^-6) + qq erfc(beta r)/r.
// Accumulates frc (+=, double) and returns e_lj, e_coul and the virial
// (6 components, reference sign convention: -sum f_lin (x) dr with the
// symmetrised off-diagonals of Get_Virial_From_Force_Dis).
ORC_API void orc64_direct(int N, const double* crd, const double* charge, const int* lj_type,
                          const double* LJ_A, const double* LJ_B, const double* box,
                          double cutoff, double beta, const int* excl_start, const int* excl_list,
                          const int* excl_num, double* frc, double* energies2, double* virial6) {
    double elj = 0, ec = 0;
    double vir[6] = {0, 0, 0, 0, 0, 0};
    std::vector<double> f((size_t)3 * N, 0.0);
    const double tsp = 2.0 / sqrt(M_PI);
#pragma omp parallel
    {
        std::vector<double> fl((size_t)3 * N, 0.0);
        double lelj = 0, lec = 0, lv[6] = {0, 0, 0, 0, 0, 0};
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < N; ++i) {
            for (int j = i + 1; j < N; ++j) {
                bool ex = false;
                if (excl_num) {
                    for (int k = excl_start[i]; k < excl_start[i] + excl_num[i]; ++k)
                        if (excl_list[k] == j) ex = true;
                    for (int k = excl_start[j]; k < excl_start[j] + excl_num[j]; ++k)
                        if (excl_list[k] == i) ex = true;
                }
                if (ex) continue;
                double d[3];
                disp64(&crd[3 * j], &crd[3 * i], box, d);
                double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                double r = sqrt(r2);
                if (r >= cutoff) continue;
                int tp = lj_type_pair(lj_type[i], lj_type[j]);
                double A = LJ_A[tp], B = LJ_B[tp];
                double r6 = 1.0 / (r2 * r2 * r2);
                double qq = charge[i] * charge[j];
                double br = beta * r;
                double fr = (B - A * r6) * r6 / r2 - qq / (r2 * r) * (br * tsp * exp(-br * br) + erfc(br));
                lelj += (A / 12.0 * r6 - B / 6.0) * r6;
                lec += qq * erfc(br) / r;
                double fl3[3] = {fr * d[0], fr * d[1], fr * d[2]};
                for (int k = 0; k < 3; ++k) {
                    fl[3 * i + k] += fl3[k];
                    fl[3 * j + k] -= fl3[k];
                }
                lv[0] -= fl3[0] * d[0];
                lv[1] -= fl3[0] * d[1] + fl3[1] * d[0];
                lv[2] -= fl3[1] * d[1];
                lv[3] -= fl3[0] * d[2] + fl3[2] * d[0];
                lv[4] -= fl3[1] * d[2] + fl3[2] * d[1];
                lv[5] -= fl3[2] * d[2];
            }
        }
#pragma omp critical
        {
            for (size_t k = 0; k < f.size(); ++k) f[k] += fl[k];
            elj += lelj;
            ec += lec;
            for (int k = 0; k < 6; ++k) vir[k] += lv[k];
        }
    }
    for (size_t k = 0; k < f.size(); ++k) frc[k] += f[k];
    energies2[0] = elj;
    energies2[1] = ec;
    for (int k = 0; k < 6; ++k) virial6[k] = vir[k];
}

// Excluded-pair correction in double: E = -sum qq erf(beta r)/r over the
// exclusion list, forces from its gradient.
ORC_API void orc64_excluded(int N, const double* crd, const double* charge, const double* box,
                            double beta, const int* excl_start, const int* excl_list,
                            const int* excl_num, double* frc, double* energy, double* virial6) {
    double e = 0;
    double vir[6] = {0, 0, 0, 0, 0, 0};
    const double tsp = 2.0 / sqrt(M_PI);
    for (int i = 0; i < N; ++i)
        for (int k = excl_start[i]; k < excl_start[i] + excl_num[i]; ++k) {
            int j = excl_list[k];
            double d[3];
            disp64(&crd[3 * j], &crd[3 * i], box, d);
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], r = sqrt(r2);
            double br = beta * r, qq = charge[i] * charge[j];
            e -= qq * erf(br) / r;
            // dE/dr = qq [erf(br)/r^2 - tsp beta exp(-b^2r^2)/r]; F_i = -dE/dr_i
            double fr = -qq * (br * tsp * exp(-br * br) + erfc(br) - 1.0) / (r2 * r);
            double fl3[3] = {fr * d[0], fr * d[1], fr * d[2]};
            for (int c = 0; c < 3; ++c) {
                frc[3 * i + c] += fl3[c];
                frc[3 * j + c] -= fl3[c];
            }
            vir[0] -= fl3[0] * d[0];
            vir[1] -= fl3[0] * d[1] + fl3[1] * d[0];
            vir[2] -= fl3[1] * d[1];
            vir[3] -= fl3[0] * d[2] + fl3[2] * d[0];
            vir[4] -= fl3[1] * d[2] + fl3[2] * d[1];
            vir[5] -= fl3[2] * d[2];
        }
    *energy = e;
    for (int k = 0; k < 6; ++k) virial6[k] = vir[k];
}

// Exact Ewald reciprocal sum in double over integer k with |k_d| <= kmax_d:
// E = 1/(2 pi V) sum_{m != 0} exp(-pi^2 m^2/beta^2)/m^2 |S(m)|^2,
// m = k / L (orthorhombic).  Forces F_j = (2 q_j / V) sum g(m) m Im(conj(S) e^{i th_j}).
// Also returns self -beta/sqrt(pi) sum q^2 and neutralizing -pi/(2 beta^2 V)(sum q)^2.
ORC_API void orc64_ewald_recip(int N, const double* crd, const double* charge, const double* box,
                               double beta, const int* kmax3, double* frc, double* energies3) {
    const double V = box[0] * box[1] * box[2];
    double E = 0;
    std::vector<double> f((size_t)3 * N, 0.0);
    const int kx0 = kmax3[0], ky0 = kmax3[1], kz0 = kmax3[2];
#pragma omp parallel
    {
        std::vector<double> fl((size_t)3 * N, 0.0);
        std::vector<double> c(N), s(N);
        double El = 0;
#pragma omp for schedule(dynamic) collapse(2)
        for (int kx = -kx0; kx <= kx0; ++kx)
            for (int ky = -ky0; ky <= ky0; ++ky)
                for (int kz = -kz0; kz <= kz0; ++kz) {
                    if (kx == 0 && ky == 0 && kz == 0) continue;
                    double m[3] = {kx / box[0], ky / box[1], kz / box[2]};
                    double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
                    double g = exp(-M_PI * M_PI * m2 / (beta * beta)) / m2;
                    double Sr = 0, Si = 0;
                    for (int j = 0; j < N; ++j) {
                        double th = 2 * M_PI * (m[0] * crd[3 * j] + m[1] * crd[3 * j + 1] + m[2] * crd[3 * j + 2]);
                        c[j] = cos(th);
                        s[j] = sin(th);
                        Sr += charge[j] * c[j];
                        Si += charge[j] * s[j];
                    }
                    El += g * (Sr * Sr + Si * Si) / (2 * M_PI * V);
                    for (int j = 0; j < N; ++j) {
                        // Im(conj(S) e^{i th}) = Sr*s - Si*c
                        double im = Sr * s[j] - Si * c[j];
                        double a = 2 * charge[j] / V * g * im;
                        for (int d = 0; d < 3; ++d) fl[3 * j + d] += a * m[d];
                    }
                }
#pragma omp critical
        {
            E += El;
            for (size_t k = 0; k < f.size(); ++k) f[k] += fl[k];
        }
    }
    for (size_t k = 0; k < f.size(); ++k) frc[k] += f[k];
    double sq = 0, sum = 0;
    for (int j = 0; j < N; ++j) {
        sq += charge[j] * charge[j];
        sum += charge[j];
    }
    energies3[0] = E;
    energies3[1] = -beta / sqrt(M_PI) * sq;
    energies3[2] = -M_PI / (2 * beta * beta * V) * sum * sum;
}

// Smooth PME in double with the same order-4 B-splines and the exact
// B-spline moduli: E_rec and forces.  Converges to orc64_ewald_recip as the
// grid refines; the FP32 restatement must agree with it to float rounding.
ORC_API void orc64_pme(int N, const double* crd, const double* charge, const double* box,
                       double beta, int fx, int fy, int fz, double* frc, double* energy,
                       double* virial6) {
    const int dims[3] = {fx, fy, fz};
    const size_t nall = (size_t)fx * fy * fz;
    std::vector<cd> Q(nall, 0.0);
    auto weights = [](double f, double* w, double* dw) {
        // offsets 0..3 below the base point u, as PME_Ma..Md / dMa..dMc
        w[0] = f * f * f / 6;
        w[1] = (-3 * f * f * f + 3 * f * f + 3 * f + 1) / 6;
        w[2] = (3 * f * f * f - 6 * f * f + 4) / 6;
        w[3] = (1 - f) * (1 - f) * (1 - f) / 6;
        dw[0] = f * f / 2;
        dw[1] = (-3 * f * f + 2 * f + 1) / 2;
        dw[2] = (3 * f * f - 4 * f) / 2;
        dw[3] = -(1 - f) * (1 - f) / 2;
    };
    std::vector<int> base((size_t)3 * N);
    std::vector<double> W((size_t)12 * N), DW((size_t)12 * N);
    for (int a = 0; a < N; ++a)
        for (int d = 0; d < 3; ++d) {
            double fr = crd[3 * a + d] / box[d];
            fr -= floor(fr);
            double t = fr * dims[d];
            int u = (int)t;
            if (u >= dims[d]) u = dims[d] - 1;
            base[3 * a + d] = u;
            weights(t - u, &W[12 * a + 4 * d], &DW[12 * a + 4 * d]);
        }
    for (int a = 0; a < N; ++a)
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) {
                    int x = ((base[3 * a] - i) % fx + fx) % fx;
                    int y = ((base[3 * a + 1] - j) % fy + fy) % fy;
                    int z = ((base[3 * a + 2] - k) % fz + fz) % fz;
                    Q[((size_t)x * fy + y) * fz + z] += charge[a] * W[12 * a + i] * W[12 * a + 4 + j] * W[12 * a + 8 + k];
                }
    std::vector<cd> F = Q;
    fft3d_c(F.data(), fx, fy, fz, -1);
    const double V = box[0] * box[1] * box[2];
    auto bmod = [](int k, int n) {
        cd s = 0;
        for (int kk = 0; kk < 3; ++kk) {
            double Mv = kk == 1 ? 4.0 / 6.0 : 1.0 / 6.0;
            double ang = 2 * M_PI * (double)k * kk / n;
            s += Mv * cd(cos(ang), sin(ang));
        }
        return 1.0 / std::norm(s);
    };
    double E = 0;
    double vir[6] = {0, 0, 0, 0, 0, 0};
    for (int x = 0; x < fx; ++x)
        for (int y = 0; y < fy; ++y)
            for (int z = 0; z < fz; ++z) {
                size_t idx = ((size_t)x * fy + y) * fz + z;
                if (x == 0 && y == 0 && z == 0) {
                    F[idx] = 0;
                    continue;
                }
                double mx = (x > fx / 2 ? x - fx : x) / box[0];
                double my = (y > fy / 2 ? y - fy : y) / box[1];
                double mz = (z > fz / 2 ? z - fz : z) / box[2];
                double m2 = mx * mx + my * my + mz * mz;
                double bc = exp(-M_PI * M_PI * m2 / (beta * beta)) / (M_PI * m2 * V) * bmod(x, fx) * bmod(y, fy) * bmod(z, fz);
                E += 0.5 * bc * std::norm(F[idx]);
                // the reference's reciprocal virial: over the half spectrum
                // kz <= fz/2 with weight 1/2 on kz = 0 and kz = fz/2,
                // -w |F|^2 BC/2 (I - 2 (1 + pi^2 m^2/beta^2) m m^T / m^2)
                // (PM_force.cpp:586-606, 1024-1068), here in double
                if (z <= fz / 2) {
                    double w = (z == 0 || z == fz / 2) ? 0.5 : 1.0;
                    double t = 2.0 / m2 * (1.0 + M_PI * M_PI * m2 / (beta * beta));
                    double s = -w * std::norm(F[idx]) * 0.5 * bc;
                    vir[0] += s * (1 - t * mx * mx);
                    vir[1] += s * (0 - t * my * mx);
                    vir[2] += s * (1 - t * my * my);
                    vir[3] += s * (0 - t * mz * mx);
                    vir[4] += s * (0 - t * mz * my);
                    vir[5] += s * (1 - t * mz * mz);
                }
                F[idx] *= bc;
            }
    if (virial6)
        for (int k = 0; k < 6; ++k) virial6[k] = vir[k];
    fft3d_c(F.data(), fx, fy, fz, +1);
    for (int a = 0; a < N; ++a) {
        double g[3] = {0, 0, 0};
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j)
                for (int k = 0; k < 4; ++k) {
                    int x = ((base[3 * a] - i) % fx + fx) % fx;
                    int y = ((base[3 * a + 1] - j) % fy + fy) % fy;
                    int z = ((base[3 * a + 2] - k) % fz + fz) % fz;
                    double phi = F[((size_t)x * fy + y) * fz + z].real();
                    const double* w = &W[12 * a];
                    const double* dw = &DW[12 * a];
                    g[0] += phi * dw[i] * w[4 + j] * w[8 + k] * fx / box[0];
                    g[1] += phi * w[i] * dw[4 + j] * w[8 + k] * fy / box[1];
                    g[2] += phi * w[i] * w[4 + j] * dw[8 + k] * fz / box[2];
                }
        for (int d = 0; d < 3; ++d) frc[3 * a + d] += -charge[a] * g[d];
    }
    *energy = E;
}

ORC_API double orc_sum_f32(const float* a, int n) {
    double s = 0;
    for (int i = 0; i < n; ++i) s += a[i];
    return s;
}

ORC_API void orc_set_num_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

ORC_API int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// ---------------------------------------------------------------------------
// NVE leap-frog + SETTLE of rigid waters: one MD_Iteration in the reference's
// NVE order (SPONGE/main.cpp:1705-1777): remember each triangle's B - A and
// C - A, leap-frog every atom (MD_Iteration_Leap_Frog, SPONGE/MD_core/
// nve.hpp:3-16: v += dt F / m, x += dt v), then SETTLE of every triangle
// (settle_triangle + SETTLE_DO_TRIANGLE, SPONGE/constrain/settle.cpp:100-258)
// with the NVE factors v_factor = x_factor = 1.

// Canonical triangle of side lengths rAB, rAC, rBC with its centre of mass at
// the origin, following Get_Rabcde_From_SSS (settle.cpp:51-94): out =
// {ra, rb, rc, rd, re}, A = (0, ra), B = (rc, rb), C = (rd, re).
// Pinned by: the side lengths and the centre of mass of the canonical
// triangle (tests/test_oracle_settle.py).
ORC_API void orc_settle_geometry(float rAB, float rAC, float rBC, float mA, float mB, float mC,
                                 float* out) {
    const float m_sum = mA + mB + mC;
    // A at the origin, B on the -x axis, C above the x axis (law of cosines)
    const float cos_a = (rBC * rBC - rAC * rAC - rAB * rAB) / (2 * rAC * rAB);
    const float sin_a = sqrtf(1.0f - cos_a * cos_a);
    float px[3] = {0.0f, -rAB, rAC * cos_a};
    float py[3] = {0.0f, 0.0f, rAC * sin_a};
    const float ox = (px[1] * mB + px[2] * mC) / m_sum;
    const float oy = py[2] * mC / m_sum;
    for (int k = 0; k < 3; ++k) {
        px[k] -= ox;
        py[k] -= oy;
    }
    // rotate A onto the y axis
    const float c = 1.0f / sqrtf(1.0f + px[0] * px[0] / py[0] / py[0]);
    const float s = c * px[0] / py[0];
    float ra = px[0] * s + py[0] * c;
    float rc = px[1] * c - py[1] * s, rb = px[1] * s + py[1] * c;
    float rd = px[2] * c - py[2] * s, re = px[2] * s + py[2] * c;
    if (ra < 0) {  // reflect y so that A lies on +y
        ra = -ra;
        rb = -rb;
        re = -re;
    }
    out[0] = ra;
    out[1] = rb;
    out[2] = rc;
    out[3] = rd;
    out[4] = re;
}

namespace {
inline V3 vcross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V3 vunit(V3 a) { return vscale(1.0f / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z), a); }
// x ex + y ey + z ez, component by component in the reference's order
inline V3 from_frame(V3 p, V3 ex, V3 ey, V3 ez) {
    return {p.x * ex.x + p.y * ey.x + p.z * ez.x, p.x * ex.y + p.y * ey.y + p.z * ez.y,
            p.x * ex.z + p.y * ey.z + p.z * ez.z};
}
}  // namespace

// One NVE step in place: crd, vel (float[3N]) from frc (float[3N]) and the
// inverse masses; waters are (A, B, C) triples of atom ids, g the canonical
// triangle of orc_settle_geometry, mass the masses SETTLE weights with.
// Pinned by: the FP64 SHAKE solution of the same constraint equations
// (orc64_nve_shake), constraint lengths, conservation of each water's linear
// momentum by the correction (tests/test_oracle_settle.py).
ORC_API void orc_nve_settle(int N, float* crd, float* vel, const float* frc, const float* inv_mass,
                            const float* mass, int n_water, const int* water, const float* g,
                            const float* cell6, const float* rcell6, float dt) {
    const LT cell = ld_lt(cell6), rcell = ld_lt(rcell6);
    std::vector<V3> old_b(n_water), old_c(n_water);
    for (int w = 0; w < n_water; ++w) {  // remember_triangle_BA_CA (settle.cpp:5-25)
        const int a = water[3 * w], b = water[3 * w + 1], c = water[3 * w + 2];
        old_b[w] = disp_pbc(ld3(crd, b), ld3(crd, a), cell, rcell);
        old_c[w] = disp_pbc(ld3(crd, c), ld3(crd, a), cell, rcell);
    }
    for (int i = 0; i < N; ++i) {  // MD_Iteration_Leap_Frog
        const V3 acc = vscale(inv_mass[i], ld3(frc, i));
        const V3 v = vadd(ld3(vel, i), vscale(dt, acc));
        const V3 x = vadd(ld3(crd, i), vscale(dt, v));
        for (int d = 0; d < 3; ++d) {
            vel[3 * i + d] = (&v.x)[d];
            crd[3 * i + d] = (&x.x)[d];
        }
    }
    const float ra = g[0], rb = g[1], rc = g[2], rd = g[3], re = g[4];
    const float v_scale = 1.0f / dt / 1.0f;  // exp_gamma / dt / half_exp_gamma_plus_half
    for (int w = 0; w < n_water; ++w) {
        const int ia = water[3 * w], ib = water[3 * w + 1], ic = water[3 * w + 2];
        const float mA = mass[ia], mB = mass[ib], mC = mass[ic];
        // settle_triangle: positions relative to the centre of mass, formed
        // through the absolute position of A as the reference does
        V3 pa = ld3(crd, ia);
        V3 pb = disp_pbc(ld3(crd, ib), pa, cell, rcell);
        V3 pc = disp_pbc(ld3(crd, ic), pa, cell, rcell);
        const V3 com = vadd(vscale(1.0f / (mA + mB + mC), vadd(vscale(mB, pb), vscale(mC, pc))), pa);
        pa = vsub(pa, com);
        pb = vadd(pb, pa);
        pc = vadd(pc, pa);
        // SETTLE_DO_TRIANGLE: frame from the old plane and the new A
        const V3 b0 = old_b[w], c0 = old_c[w];
        V3 ez = vcross(b0, c0);
        V3 ex = vcross(pa, ez);
        V3 ey = vcross(ez, ex);
        ex = vunit(ex);
        ey = vunit(ey);
        ez = vunit(ez);
        const float b0x = vdot(ex, b0), b0y = vdot(ey, b0);
        const float c0x = vdot(ex, c0), c0y = vdot(ey, c0);
        const float a1z = vdot(ez, pa);
        const V3 b1 = {vdot(ex, pb), vdot(ey, pb), vdot(ez, pb)};
        const V3 c1 = {vdot(ex, pc), vdot(ey, pc), vdot(ez, pc)};
        const float sin_phi = a1z / ra;
        const float cos_phi = sqrtf(1.0f - sin_phi * sin_phi);
        const float sin_psi = (b1.z - c1.z - (rb - re) * sin_phi) / ((rd - rc) * cos_phi);
        const float cos_psi = sqrtf(1.0f - sin_psi * sin_psi);
        const V3 a2 = {0.0f, ra * cos_phi, a1z};
        const V3 b2 = {rc * cos_psi, rb * cos_phi + rc * sin_psi * sin_phi, b1.z};
        const V3 c2 = {rd * cos_psi, re * cos_phi + rd * sin_psi * sin_phi, c1.z};
        const float alpha = b2.x * b0x + c2.x * c0x + b2.y * b0y + c2.y * c0y;
        const float beta = -b2.x * b0y - c2.x * c0y + b2.y * b0x + c2.y * c0x;
        const float gamma = b1.y * b0x - b1.x * b0y + c1.y * c0x - c1.x * c0y;
        const float ab2 = alpha * alpha + beta * beta;
        const float sin_t = (alpha * gamma - beta * sqrtf(ab2 - gamma * gamma)) / ab2;
        const float cos_t = sqrtf(1.0f - sin_t * sin_t);
        const V3 a3 = {-a2.y * sin_t, a2.y * cos_t, a2.z};
        const V3 b3 = {b2.x * cos_t - b2.y * sin_t, b2.x * sin_t + b2.y * cos_t, b2.z};
        const V3 c3 = {c2.x * cos_t - c2.y * sin_t, c2.x * sin_t + c2.y * cos_t, c2.z};
        const V3 wa = from_frame(a3, ex, ey, ez);
        const V3 wb = from_frame(b3, ex, ey, ez);
        const V3 wc = from_frame(c3, ex, ey, ez);
        const V3 dva = vscale(v_scale, vsub(wa, pa));
        const V3 dvb = vscale(v_scale, vsub(wb, pb));
        const V3 dvc = vscale(v_scale, vsub(wc, pc));
        const V3 na = vadd(wa, com), nb = vadd(wb, com), nc = vadd(wc, com);
        const V3 va = vadd(ld3(vel, ia), dva), vb = vadd(ld3(vel, ib), dvb),
                 vc = vadd(ld3(vel, ic), dvc);
        for (int d = 0; d < 3; ++d) {
            crd[3 * ia + d] = (&na.x)[d];
            crd[3 * ib + d] = (&nb.x)[d];
            crd[3 * ic + d] = (&nc.x)[d];
            vel[3 * ia + d] = (&va.x)[d];
            vel[3 * ib + d] = (&vb.x)[d];
            vel[3 * ic + d] = (&vc.x)[d];
        }
    }
}

// FP64 oracle of the same step: leap-frog, then the constraint equations of
// each water (|A-B| = dAB, |A-C| = dAC, |B-C| = dBC with the correction of
// each atom along the old bond vectors, weighted by the inverse masses --
// the equations SETTLE solves in closed form) solved by SHAKE iteration to
// a relative tolerance tol; velocities corrected by the displacement / dt.
// Returns the largest iteration count used (max_iter when not converged).
// Pinned by: exact constraint lengths after the step and the closed form of
// a force-free rotating water (tests/test_oracle_settle.py).
ORC_API int orc64_nve_shake(int N, double* crd, double* vel, const double* frc,
                            const double* inv_mass, int n_water, const int* water, double dAB,
                            double dAC, double dBC, const double* box, double dt, int max_iter,
                            double tol) {
    auto mimg = [&](double d, int k) { return d - box[k] * std::floor(d / box[k] + 0.5); };
    std::vector<double> old(9 * (size_t)n_water);
    for (int w = 0; w < n_water; ++w) {
        const int id[3] = {water[3 * w], water[3 * w + 1], water[3 * w + 2]};
        // old bond vectors A-B, A-C, B-C
        for (int k = 0; k < 3; ++k) {
            const double ab = mimg(crd[3 * id[0] + k] - crd[3 * id[1] + k], k);
            const double ac = mimg(crd[3 * id[0] + k] - crd[3 * id[2] + k], k);
            old[9 * w + k] = ab;
            old[9 * w + 3 + k] = ac;
            old[9 * w + 6 + k] = ac - ab;  // B - C = (A - C) - (A - B)
        }
    }
    for (int i = 0; i < N; ++i)
        for (int k = 0; k < 3; ++k) {
            vel[3 * i + k] += dt * inv_mass[i] * frc[3 * i + k];
            crd[3 * i + k] += dt * vel[3 * i + k];
        }
    int worst = 0;
    for (int w = 0; w < n_water; ++w) {
        const int id[3] = {water[3 * w], water[3 * w + 1], water[3 * w + 2]};
        // unwrapped positions relative to A's
        double p[3][3], p0[3][3];
        for (int k = 0; k < 3; ++k) {
            p[0][k] = crd[3 * id[0] + k];
            p[1][k] = p[0][k] + mimg(crd[3 * id[1] + k] - p[0][k], k);
            p[2][k] = p[0][k] + mimg(crd[3 * id[2] + k] - p[0][k], k);
        }
        std::memcpy(p0, p, sizeof(p));
        const int ca[3] = {0, 0, 1}, cb[3] = {1, 2, 2};
        const double d2[3] = {dAB * dAB, dAC * dAC, dBC * dBC};
        int it = 0;
        for (; it < max_iter; ++it) {
            double worst_rel = 0;
            for (int c = 0; c < 3; ++c) {
                const int a = ca[c], b = cb[c];
                const double* o = &old[9 * w + 3 * c];
                double r[3], r2 = 0, ro = 0;
                for (int k = 0; k < 3; ++k) {
                    r[k] = p[a][k] - p[b][k];
                    r2 += r[k] * r[k];
                    ro += r[k] * o[k];
                }
                worst_rel = std::max(worst_rel, std::fabs(r2 - d2[c]) / d2[c]);
                const double ima = inv_mass[id[a]], imb = inv_mass[id[b]];
                const double gg = (d2[c] - r2) / (2.0 * (ima + imb) * ro);
                for (int k = 0; k < 3; ++k) {
                    p[a][k] += gg * ima * o[k];
                    p[b][k] -= gg * imb * o[k];
                }
            }
            if (worst_rel < tol) break;
        }
        worst = std::max(worst, it);
        for (int a = 0; a < 3; ++a)
            for (int k = 0; k < 3; ++k) {
                vel[3 * id[a] + k] += (p[a][k] - p0[a][k]) / dt;
                crd[3 * id[a] + k] = p[a][k];
            }
    }
    return worst;
}

// orc64_direct over a half list (counts[N], serial[N*max_nb]) that holds
// every non-excluded pair within the cutoff -- orc_nl_build's list, equal to
// the brute-force set (tests/test_oracle.py) -- for systems where the O(N^2)
// enumeration of orc64_direct is too slow.  Same sums in double.
// Pinned by: orc64_direct on the same system (tests/test_oracle.py).
ORC_API void orc64_direct_list(int N, const double* crd, const double* charge, const int* lj_type,
                               const double* LJ_A, const double* LJ_B, const double* box,
                               double cutoff, double beta, const int* counts, const int* serial,
                               int max_nb, double* frc, double* energies2, double* virial6) {
    double elj = 0, ec = 0;
    double vir[6] = {0, 0, 0, 0, 0, 0};
    std::vector<double> f((size_t)3 * N, 0.0);
    const double tsp = 2.0 / sqrt(M_PI);
#pragma omp parallel
    {
        std::vector<double> fl((size_t)3 * N, 0.0);
        double lelj = 0, lec = 0, lv[6] = {0, 0, 0, 0, 0, 0};
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < N; ++i) {
            const int* row = serial + (size_t)i * max_nb;
            for (int k = 0; k < counts[i]; ++k) {
                const int j = row[k];
                double d[3];
                disp64(&crd[3 * j], &crd[3 * i], box, d);
                const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                const double r = sqrt(r2);
                if (r >= cutoff) continue;
                const int tp = lj_type_pair(lj_type[i], lj_type[j]);
                const double A = LJ_A[tp], B = LJ_B[tp];
                const double r6 = 1.0 / (r2 * r2 * r2);
                const double qq = charge[i] * charge[j];
                const double br = beta * r;
                const double fr =
                    (B - A * r6) * r6 / r2 - qq / (r2 * r) * (br * tsp * exp(-br * br) + erfc(br));
                lelj += (A / 12.0 * r6 - B / 6.0) * r6;
                lec += qq * erfc(br) / r;
                const double g[3] = {fr * d[0], fr * d[1], fr * d[2]};
                for (int c = 0; c < 3; ++c) {
                    fl[3 * i + c] += g[c];
                    fl[3 * j + c] -= g[c];
                }
                lv[0] -= g[0] * d[0];
                lv[1] -= g[0] * d[1] + g[1] * d[0];
                lv[2] -= g[1] * d[1];
                lv[3] -= g[0] * d[2] + g[2] * d[0];
                lv[4] -= g[1] * d[2] + g[2] * d[1];
                lv[5] -= g[2] * d[2];
            }
        }
#pragma omp critical
        {
            for (size_t k = 0; k < f.size(); ++k) f[k] += fl[k];
            elj += lelj;
            ec += lec;
            for (int k = 0; k < 6; ++k) vir[k] += lv[k];
        }
    }
    for (size_t k = 0; k < f.size(); ++k) frc[k] += f[k];
    energies2[0] = elj;
    energies2[1] = ec;
    for (int k = 0; k < 6; ++k) virial6[k] = vir[k];
}

// solvent_LJ, restating Lennard_Jones_And_Direct_Coulomb_Device<WAT_POINTS>
// of SPONGE/Lennard_Jones_force/solvent_LJ.cpp:24-147: for every water of the
// solvent tail (atoms from solvent_atom on, water_points atoms per water,
// oxygen first) and every j of the OXYGEN's half list, the pairs (a, j) of
// each atom a of the water with the generic pair functions; forces on a and
// j; energies (LJ + Coulomb, LJ, Coulomb) and virial on the oxygen.
// Accumulates (+=) like orc_lj_direct; called after orc_lj_direct with
// solvent_n = local_n - solvent_atom.
// Pinned by: on a list fresh from a build the two together equal the plain
// orc_lj_direct (every pair within the cutoff of two waters is in the
// oxygens' lists); on a stale list, an FP64 sum over the pair set spelled out
// in numpy (tests/test_oracle_solvent.py).
ORC_API void orc_solvent_lj(int local_n, int solvent_atom, int water_points, const int* counts,
                            const int* serial, int max_nb, const float* crd, const float* charge,
                            const int* lj_type, const float* LJ_A, const float* LJ_B,
                            const float* cell6, const float* rcell6, float cutoff, float pme_beta,
                            int need_energy, int need_virial, float* frc, float* atom_energy,
                            float* atom_virial, float* atom_direct_cf_energy,
                            float* atom_LJ_ene) {
    const LT cell = ld_lt(cell6), rcell = ld_lt(rcell6);
    const int n_water = (local_n - solvent_atom) / water_points;
    ORC_PRAGMA(omp parallel for schedule(dynamic))
    for (int w = 0; w < n_water; ++w) {
        const int oxygen = solvent_atom + w * water_points;
        V3 own[4] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        LT virial = {0, 0, 0, 0, 0, 0};
        float e_lj = 0, e_c = 0;
        const int* row = serial + (size_t)oxygen * max_nb;
        for (int k = 0; k < counts[oxygen]; ++k) {
            const int atom_j = row[k];
            const float ij_factor = atom_j < local_n ? 1.0f : 0.5f;
            const V3 r2 = ld3(crd, atom_j);
            const float q2 = charge[atom_j];
            const int t2 = lj_type[atom_j];
            V3 on_j = {0, 0, 0};
            for (int a = 0; a < water_points; ++a) {
                const int atom_a = oxygen + a;
                const V3 dr = disp_pbc(r2, ld3(crd, atom_a), cell, rcell);
                const float dr_abs = sqrtf(dr.x * dr.x + dr.y * dr.y + dr.z * dr.z);
                if (!(dr_abs < cutoff)) continue;
                const int tp = lj_type_pair(lj_type[atom_a], t2);
                const float A = LJ_A[tp], B = LJ_B[tp];
                const float q1 = charge[atom_a];
                float frc_abs = (B - A * powf(dr_abs, -6.0f)) * powf(dr_abs, -8.0f);
                const float beta_dr = pme_beta * dr_abs;
                frc_abs -= q1 * q2 * powf(dr_abs, -3.0f) *
                           (beta_dr * TWO_DIVIDED_BY_SQRT_PI * expf(-beta_dr * beta_dr) +
                            erfcf(beta_dr));
                const V3 frc_lin = vscale(frc_abs, dr);
                own[a] = vadd(own[a], frc_lin);
                on_j = vsub(on_j, frc_lin);
                if (need_virial) virial = lt_sub(virial, lt_scale(ij_factor, virial_fd(frc_lin, dr)));
                if (need_energy) {
                    const float dr_6 = powf(dr_abs, -6.0f);
                    e_lj += ij_factor * ((0.083333333f * A * dr_6 - 0.166666667f * B) * dr_6);
                    e_c += ij_factor * (q1 * q2 * erfcf(pme_beta * dr_abs) / dr_abs);
                }
            }
            if (atom_j < local_n) {
                ORC_ATOMIC frc[3 * atom_j] += on_j.x;
                ORC_ATOMIC frc[3 * atom_j + 1] += on_j.y;
                ORC_ATOMIC frc[3 * atom_j + 2] += on_j.z;
            }
        }
        for (int a = 0; a < water_points; ++a) {
            ORC_ATOMIC frc[3 * (oxygen + a)] += own[a].x;
            ORC_ATOMIC frc[3 * (oxygen + a) + 1] += own[a].y;
            ORC_ATOMIC frc[3 * (oxygen + a) + 2] += own[a].z;
        }
        if (need_energy) {
            atom_energy[oxygen] += e_lj + e_c;
            if (atom_LJ_ene) atom_LJ_ene[oxygen] += e_lj;
            if (atom_direct_cf_energy) atom_direct_cf_energy[oxygen] += e_c;
        }
        if (need_virial) {
            float* v = atom_virial + 6 * oxygen;
            v[0] += virial.a11; v[1] += virial.a21; v[2] += virial.a22;
            v[3] += virial.a31; v[4] += virial.a32; v[5] += virial.a33;
        }
    }
}
