// PME reciprocal part and excluded-pair correction for gfx950.
//
// Reference: SPONGE/PM_force/PM_force.cpp.
//   Initial          :295-654   beta, grid, B-spline moduli, BC / virial_BC
//   PME_Atom_Near    :699-762   fractional coordinates, base point u, f
//   PME_Q_Spread     :765-814   order-4 cardinal B-spline charge spreading
//   R2C, PME_Sum_Virial, PME_BCFQ, C2R   :1102-1125
//   PME_Final        :829-893   force gather, device_add_force :1014-1022
//   energies         :1142-1166
//   MPI_PME_Excluded_Force_With_Atom_Energy_Correction :1745-1817
//
// The 3D FFT is our own: five LDS-staged passes of a mixed-radix (4,2,3,5,7)
// Stockham transform.  The complex half spectrum is kept as [kz][x][y] so
// every pass reads and writes whole 128-byte segments:
//   z-R2C (two real lines per complex transform) -> y-FFT ->
//   x-FFT + BC multiply + energy / virial sums + inverse x-FFT ->
//   inverse y-FFT -> z-C2R (two real lines per complex transform).
// The stencil indices of PME_Atom_Near are recomputed on the fly in spread
// and gather (they are a pure function of the coordinates), so no 64-int
// per-atom table is stored.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"
#include "batch_launch.h"
#include "timing.h"

using namespace sa;

namespace {

constexpr float kPI = 3.1415926f;                      // PM_force.cpp:17
constexpr float kTwoOverSqrtPi = 1.1283791670218446f;  // PM_force.cpp:19

__constant__ float cMa[4] = {(float)(1.0 / 6.0), -0.5f, 0.5f, (float)(-1.0 / 6.0)};
__constant__ float cMb[4] = {0.f, 0.5f, -1.f, 0.5f};
__constant__ float cMc[4] = {0.f, 0.5f, 0.f, -0.5f};
__constant__ float cMd[4] = {0.f, (float)(1.0 / 6.0), (float)(4.0 / 6.0), (float)(1.0 / 6.0)};

// ---------------------------------------------------------------------------
// complex helpers
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float s, float2 a) { return make_float2(s * a.x, s * a.y); }
// multiply by SIGN * i
template <int SIGN>
__device__ __forceinline__ float2 mul_si(float2 a) {
    return SIGN < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// Small DFTs X_p = sum_q v_q exp(SIGN 2 pi i p q / R)
template <int R, int SIGN>
struct Dft;

template <int SIGN>
struct Dft<2, SIGN> {
    __device__ __forceinline__ static void run(float2* v) {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int SIGN>
struct Dft<4, SIGN> {
    __device__ __forceinline__ static void run(float2* v) {
        float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        float2 t2 = cadd(v[1], v[3]), t3 = mul_si<SIGN>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <int SIGN>
struct Dft<3, SIGN> {
    __device__ __forceinline__ static void run(float2* v) {
        const float h = 0.86602540378443864676f;  // sqrt(3)/2
        float2 s = cadd(v[1], v[2]);
        float2 d = mul_si<SIGN>(csub(v[1], v[2]));
        float2 m = csub(v[0], cscale(0.5f, s));
        v[0] = cadd(v[0], s);
        v[1] = cadd(m, cscale(h, d));
        v[2] = csub(m, cscale(h, d));
    }
};

// cos / sin of 2 pi k / R for the R = 5, 7 butterflies
template <int R>
struct Roots;
template <>
struct Roots<5> {
    static constexpr float c[5] = {1.000000000f, 0.309016994f, -0.809016994f, -0.809016994f,
                                   0.309016994f};
    static constexpr float s[5] = {0.000000000f, 0.951056516f, 0.587785252f, -0.587785252f,
                                   -0.951056516f};
};
template <>
struct Roots<7> {
    static constexpr float c[7] = {1.000000000f,  0.623489802f,  -0.222520934f, -0.900968868f,
                                   -0.900968868f, -0.222520934f, 0.623489802f};
    static constexpr float s[7] = {0.000000000f,  0.781831482f,  0.974927912f, 0.433883739f,
                                   -0.433883739f, -0.974927912f, -0.781831482f};
};

// R = 5, 7: direct small DFT
template <int R, int SIGN>
struct Dft {
    __device__ __forceinline__ static void run(float2* v) {
        float cs[R], sn[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            cs[k] = Roots<R>::c[k];
            sn[k] = SIGN * Roots<R>::s[k];
        }
        float2 o[R];
#pragma unroll
        for (int p = 0; p < R; ++p) {
            float2 acc = v[0];
#pragma unroll
            for (int q = 1; q < R; ++q) {
                const int k = (p * q) % R;
                acc = cadd(acc, cmul(v[q], make_float2(cs[k], sn[k])));
            }
            o[p] = acc;
        }
#pragma unroll
        for (int p = 0; p < R; ++p) v[p] = o[p];
    }
};

// One radix-R Stockham stage over nl lines of length n (line stride ld) held
// in LDS: out[(j/Ns)*Ns*R + j%Ns + r*Ns] = DFT_R(in[j + r*n/R] * w^(r*(j%Ns))).
template <int R, int SIGN>
__device__ __forceinline__ void stockham_stage(const float2* a, float2* b, int nl, int n, int ld,
                                               int Ns, const float2* __restrict__ tw) {
    const int nb = n / R;
    const int total = nl * nb;
    const int step = n / (Ns * R);
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
        const int line = t / nb, j = t - line * nb;
        const float2* in = a + line * ld;
        float2* out = b + line * ld;
        float2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = in[j + r * nb];
        const int k = j % Ns;
#pragma unroll
        for (int r = 1; r < R; ++r) {
            float2 w = tw[r * k * step];
            if (SIGN > 0) w.y = -w.y;
            v[r] = cmul(v[r], w);
        }
        Dft<R, SIGN>::run(v);
        const int idx = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) out[idx + r * Ns] = v[r];
    }
}

// Transform nl lines in place (result returned; ping-pong between a and b).
// tw: the line length's twiddles, staged in LDS by the caller.
template <int SIGN>
__device__ __forceinline__ float2* fft_lines(float2* a, float2* b, int nl, int n, int ld, const sa_radix_plan& pl,
                             const float2* __restrict__ tw) {
    int Ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = pl.r[s];
        switch (R) {
            case 4: stockham_stage<4, SIGN>(a, b, nl, n, ld, Ns, tw); break;
            case 2: stockham_stage<2, SIGN>(a, b, nl, n, ld, Ns, tw); break;
            case 3: stockham_stage<3, SIGN>(a, b, nl, n, ld, Ns, tw); break;
            case 5: stockham_stage<5, SIGN>(a, b, nl, n, ld, Ns, tw); break;
            default: stockham_stage<7, SIGN>(a, b, nl, n, ld, Ns, tw); break;
        }
        __syncthreads();
        float2* t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

constexpr int kMaxFftLen = 512;  // largest grid dimension (checked at init)

// z pass, real -> half spectrum: block = (y block, x)
__device__ __forceinline__ void fft_z_r2c_body(int sa_block, const float* __restrict__ Q,
                                                   float2* __restrict__ FQ, int fx, int fy, int fz,
                                                   int fzc, int TL, const sa_radix_plan& pl,
                                                   const float2* __restrict__ tw) {
    extern __shared__ float2 sm[];
    __shared__ float2 stw[kMaxFftLen];
    const int gy = (fy + TL - 1) / TL;  // y blocks per x line (1D grid: x-major)
    const int x = sa_block / gy, y0 = (sa_block - x * gy) * TL;
    const int nl = min(TL, fy - y0), npair = (nl + 1) >> 1;
    float2* a = sm;
    float2* b = sm + npair * fz;
    const float* src = Q + ((size_t)x * fy + y0) * fz;
    for (int k = threadIdx.x; k < fz; k += blockDim.x) stw[k] = tw[k];
    for (int e = threadIdx.x; e < npair * fz; e += blockDim.x) {
        const int m = e / fz, z = e - m * fz;
        const float v0 = src[(size_t)(2 * m) * fz + z];
        const float v1 = (2 * m + 1 < nl) ? src[(size_t)(2 * m + 1) * fz + z] : 0.f;
        a[e] = make_float2(v0, v1);
    }
    __syncthreads();
    const float2* res = fft_lines<-1>(a, b, npair, fz, fz, pl, stw);
    for (int e = threadIdx.x; e < fzc * nl; e += blockDim.x) {
        const int kz = e / nl, l = e - kz * nl;
        const int m = l >> 1;
        const float2 Z = res[m * fz + kz];
        const float2 Zm = res[m * fz + (kz == 0 ? 0 : fz - kz)];
        float2 v;
        if ((l & 1) == 0) v = make_float2(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
        else v = make_float2(0.5f * (Z.y + Zm.y), -0.5f * (Z.x - Zm.x));
        FQ[((size_t)kz * fx + x) * fy + y0 + l] = v;
    }
}

__global__ void __launch_bounds__(256) k_fft_z_r2c(const float* __restrict__ Q,
                                                   float2* __restrict__ FQ, int fx, int fy, int fz,
                                                   int fzc, int TL, sa_radix_plan pl,
                                                   const float2* __restrict__ tw) {
    fft_z_r2c_body(blockIdx.x, Q, FQ, fx, fy, fz, fzc, TL, pl, tw);
}

// y pass over contiguous lines [kz][x][0..fy)
template <int SIGN>
__device__ __forceinline__ void fft_y_body(int sa_block, float2* __restrict__ FQ, int nlines, int fy, int TL,
                                               const sa_radix_plan& pl, const float2* __restrict__ tw) {
    extern __shared__ float2 sm[];
    __shared__ float2 stw[kMaxFftLen];
    const int l0 = sa_block * TL;
    const int nl = min(TL, nlines - l0);
    float2* a = sm;
    float2* b = sm + TL * fy;
    float2* base = FQ + (size_t)l0 * fy;
    for (int k = threadIdx.x; k < fy; k += blockDim.x) stw[k] = tw[k];
    for (int e = threadIdx.x; e < nl * fy; e += blockDim.x) a[e] = base[e];
    __syncthreads();
    const float2* res = fft_lines<SIGN>(a, b, nl, fy, fy, pl, stw);
    for (int e = threadIdx.x; e < nl * fy; e += blockDim.x) base[e] = res[e];
}

template <int SIGN>
__global__ void __launch_bounds__(256) k_fft_y(float2* __restrict__ FQ, int nlines, int fy, int TL,
                                               sa_radix_plan pl, const float2* __restrict__ tw) {
    fft_y_body<SIGN>(blockIdx.x, FQ, nlines, fy, TL, pl, tw);
}

// x pass with the reciprocal-space convolution: block = (y block, kz).
// acc[0] += sum w(kz) BC |F|^2 (w = 1 on kz = 0 and kz = fz/2 for even fz,
// 2 otherwise: the full-spectrum sum of the reference's real-space
// 0.5 * sum Q * FBCFQ, PM_force.cpp:896-919); acc[1..6] -= factor |F|^2
// virial_BC with factor 0.5 on kz = 0 and kz = fzc-1 (PME_Sum_Virial,
// :1024-1068, including its odd-fz weighting).
// ACC: accumulate the energy (and, with need_virial, the virial) sums; off on
// force-only evaluations.
template <bool ACC>
__device__ __forceinline__ void fft_x_conv_body(int sa_block, 
    float2* __restrict__ FQ, const float* __restrict__ BCt, const float* __restrict__ vBCt,
    long long nfft, int fx, int fy, int fz, int fzc, int TL, const sa_radix_plan& pl,
    const float2* __restrict__ tw, int need_virial, double* __restrict__ acc) {
    extern __shared__ float2 sm[];
    __shared__ double red[7][4];
    __shared__ float2 stw[kMaxFftLen];
    const int gy = (fy + TL - 1) / TL;  // y blocks per kz plane (1D grid: kz-major)
    const int kz = sa_block / gy, y0 = (sa_block - kz * gy) * TL;
    const int nl = min(TL, fy - y0);
    float2* a = sm;
    float2* b = sm + TL * fx;
    for (int k = threadIdx.x; k < fx; k += blockDim.x) stw[k] = tw[k];
    for (int e = threadIdx.x; e < fx * nl; e += blockDim.x) {
        const int xi = e / nl, l = e - xi * nl;
        a[l * fx + xi] = FQ[((size_t)kz * fx + xi) * fy + y0 + l];
    }
    __syncthreads();
    float2* res = fft_lines<-1>(a, b, nl, fx, fx, pl, stw);
    float2* other = (res == a) ? b : a;
    const double wE = (kz == 0 || 2 * kz == fz) ? 1.0 : 2.0;
    const float vfac = (kz == 0 || kz == fzc - 1) ? 0.5f : 1.0f;
    double e_acc = 0.0;
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int e = threadIdx.x; e < nl * fx; e += blockDim.x) {
        const int l = e / fx, kx = e - l * fx;
        const size_t ti = ((size_t)kz * fy + y0 + l) * fx + kx;
        const float2 F = res[e];
        const float bc = BCt[ti];
        if (ACC) {
            const float mag = F.x * F.x + F.y * F.y;
            e_acc += wE * (double)bc * (double)mag;
            if (need_virial) {
                const float s = vfac * mag;
#pragma unroll
                for (int c = 0; c < 6; ++c) v[c] -= (double)(s * vBCt[(size_t)c * nfft + ti]);
            }
        }
        res[e] = make_float2(F.x * bc, F.y * bc);
    }
    __syncthreads();
    const float2* out = fft_lines<+1>(res, other, nl, fx, fx, pl, stw);
    for (int e = threadIdx.x; e < fx * nl; e += blockDim.x) {
        const int xi = e / nl, l = e - xi * nl;
        FQ[((size_t)kz * fx + xi) * fy + y0 + l] = out[l * fx + xi];
    }
    if (!ACC) return;
    // block reduction of the energy and virial sums
    e_acc = wave_sum(e_acc);
    if (need_virial)
        for (int c = 0; c < 6; ++c) v[c] = wave_sum(v[c]);
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    if (lane == 0) {
        red[0][w] = e_acc;
        for (int c = 0; c < 6; ++c) red[1 + c][w] = v[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x / kWave;
        double s0 = 0;
        for (int k = 0; k < nw; ++k) s0 += red[0][k];
        unsafeAtomicAdd(&acc[0], s0);
        if (need_virial)
            for (int c = 0; c < 6; ++c) {
                double s = 0;
                for (int k = 0; k < nw; ++k) s += red[1 + c][k];
                unsafeAtomicAdd(&acc[1 + c], s);
            }
    }
}

template <bool ACC>
__global__ void __launch_bounds__(256) k_fft_x_conv(
    float2* __restrict__ FQ, const float* __restrict__ BCt, const float* __restrict__ vBCt,
    long long nfft, int fx, int fy, int fz, int fzc, int TL, sa_radix_plan pl,
    const float2* __restrict__ tw, int need_virial, double* __restrict__ acc) {
    fft_x_conv_body<ACC>(blockIdx.x, FQ, BCt, vBCt, nfft, fx, fy, fz, fzc, TL, pl, tw, need_virial, acc);
}

// z pass, half spectrum -> real (unnormalised C2R): block = (y block, x)
__device__ __forceinline__ void fft_z_c2r_body(int sa_block, const float2* __restrict__ FQ,
                                                   float* __restrict__ out, int fx, int fy, int fz,
                                                   int fzc, int TL, const sa_radix_plan& pl,
                                                   const float2* __restrict__ tw) {
    extern __shared__ float2 sm[];
    __shared__ float2 stw[kMaxFftLen];
    const int gy = (fy + TL - 1) / TL;  // y blocks per x line (1D grid: x-major)
    const int x = sa_block / gy, y0 = (sa_block - x * gy) * TL;
    const int nl = min(TL, fy - y0), npair = (nl + 1) >> 1;
    for (int k = threadIdx.x; k < fz; k += blockDim.x) stw[k] = tw[k];
    float2* h = sm;                       // [nl][fzc]
    float2* a = sm + TL * fzc;            // [npair][fz]
    float2* b = a + ((TL + 1) >> 1) * fz;
    for (int e = threadIdx.x; e < fzc * nl; e += blockDim.x) {
        const int kz = e / nl, l = e - kz * nl;
        float2 v = FQ[((size_t)kz * fx + x) * fy + y0 + l];
        // the imaginary parts of the DC and Nyquist terms do not enter a
        // real-to-complex inverse
        if (kz == 0 || 2 * kz == fz) v.y = 0.f;
        h[l * fzc + kz] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < npair * fz; e += blockDim.x) {
        const int m = e / fz, k = e - m * fz;
        float2 A, B = make_float2(0.f, 0.f);
        if (k < fzc) A = h[(2 * m) * fzc + k];
        else {
            A = h[(2 * m) * fzc + fz - k];
            A.y = -A.y;
        }
        if (2 * m + 1 < nl) {
            if (k < fzc) B = h[(2 * m + 1) * fzc + k];
            else {
                B = h[(2 * m + 1) * fzc + fz - k];
                B.y = -B.y;
            }
        }
        a[e] = make_float2(A.x - B.y, A.y + B.x);
    }
    __syncthreads();
    const float2* res = fft_lines<+1>(a, b, npair, fz, fz, pl, stw);
    float* dst = out + ((size_t)x * fy + y0) * fz;
    for (int e = threadIdx.x; e < nl * fz; e += blockDim.x) {
        const int l = e / fz, z = e - l * fz;
        const float2 v = res[(l >> 1) * fz + z];
        dst[e] = (l & 1) ? v.y : v.x;
    }
}

__global__ void __launch_bounds__(256) k_fft_z_c2r(const float2* __restrict__ FQ,
                                                   float* __restrict__ out, int fx, int fy, int fz,
                                                   int fzc, int TL, sa_radix_plan pl,
                                                   const float2* __restrict__ tw) {
    fft_z_c2r_body(blockIdx.x, FQ, out, fx, fy, fz, fzc, TL, pl, tw);
}

// ---------------------------------------------------------------------------
// PME_Atom_Near (PM_force.cpp:706-733): base point u and fraction f
struct Near {
    int ux, uy, uz;
    float fx, fy, fz;
};

// FMA contraction stays off here: fused, frac * n - t would subtract the
// truncated base from the unrounded product, and an atom on a grid plane
// (frac * n rounding up to an integer) would get f = 1 - eps on the base
// below instead of f = 0 on its own plane: a charge one grid step away.
__device__ __forceinline__ void near_axis(float frac, int n, int& u, float& f) {
#pragma clang fp contract(off)
    frac *= n;
    int t = (int)frac;
    t = t < 0 ? 0 : (t < n ? t : n - 1);
    u = t;
    f = frac - t;
    f = f - floorf(f);
}

__device__ __forceinline__ Near atom_near(const float* __restrict__ crd, int i, const LT6& rc,
                                          int fx, int fy, int fz) {
#pragma clang fp contract(off)
    const float x = crd[3 * i], y = crd[3 * i + 1], z = crd[3 * i + 2];
    // VECTOR * LTMatrix3 then frac - floorf(frac)
    float sx = x * rc.a11 + y * rc.a21 + z * rc.a31;
    float sy = y * rc.a22 + z * rc.a32;
    float sz = z * rc.a33;
    sx = sx - floorf(sx);
    sy = sy - floorf(sy);
    sz = sz - floorf(sz);
    if (!isfinite(sx) || !isfinite(sy) || !isfinite(sz)) sx = sy = sz = 0.f;
    Near o;
    near_axis(sx, fx, o.ux, o.fx);
    near_axis(sy, fy, o.uy, o.fy);
    near_axis(sz, fz, o.uz, o.fz);
    return o;
}

// The weight polynomials of one stencil index k, held in registers by a
// lane for its whole kernel (the lane's k never changes), so the hot loops
// do not reload the coefficient tables.
struct BspCoef {
    float a, b, c, d;     // M: ((a f + b) f + c) f + d
    __device__ __forceinline__ explicit BspCoef(int k) : a(cMa[k]), b(cMb[k]), c(cMc[k]), d(cMd[k]) {}
    __device__ __forceinline__ float w(float f) const { return fmaf(fmaf(fmaf(a, f, b), f, c), f, d); }
};


// Tile spread.  The charge grid is cut into bricks of Bx x By x Bz points (the
// last brick of an axis takes the remainder, Bx..2Bx-1 points).  An atom
// spreads onto the points u - k (k = 0..3 per axis) below its base point u;
// so every atom whose base point lies in a brick spreads inside that brick's
// halo tile, the brick extended by 3 points downwards on each axis.  One
// workgroup per brick spreads the atoms binned in it into the tile in LDS and
// stores the tile; a second pass sums, for every grid point, the entries of
// the (at most 8) tiles covering it in a fixed order.  Neither pass uses an
// atomic or needs the grid cleared, and the grid is the same on every run
// for the same binning.  Same weights and order as PME_Q_Spread
// (PM_force.cpp:787-812): q Mx(kx) My(ky) Mz(kz).

// brick of grid coordinate x on one axis (the last brick takes the remainder)
__device__ __forceinline__ int brick_of(int x, int B, int nbk) {
    const int h = x / B;
    return h < nbk ? h : nbk - 1;
}

// Bin the atoms by the brick of their base point (and clear the FFT energy /
// virial accumulators of this evaluation).  Slots: an LDS histogram of the
// block's atoms per brick, then one global atomic per distinct brick of the
// block, all issued at once (one round trip instead of a chain); for grids
// with more bricks than the LDS table holds, one atomic per distinct brick of
// each wavefront.
constexpr int kBinLds = 4096;  // bricks tracked in LDS by the bin kernel

__device__ __forceinline__ void pme_bin_body(int sa_block, int N, const float* __restrict__ crd,
                                                 const float* __restrict__ charge, LT6 rc, int fx,
                                                 int fy, int fz, int bx, int by, int bz, int nbx,
                                                 int nby, int nbz, int nb, int bcap,
                                                 int* __restrict__ bcount, float4* __restrict__ bw,
                                                 int* __restrict__ bu, int* __restrict__ bi,
                                                 float4* __restrict__ ow, int* __restrict__ ou,
                                                 int* __restrict__ oi, double* __restrict__ acc) {
    __shared__ int lcount[kBinLds];
    __shared__ int lbase[kBinLds];
    const int i = sa_block * blockDim.x + threadIdx.x;
    if (i < 7 && sa_block == 0) acc[i] = 0.0;
    const bool live = i < N;
    int h = -1, u = 0;
    float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
    if (live) {
        const Near nr = atom_near(crd, i, rc, fx, fy, fz);
        h = (brick_of(nr.ux, bx, nbx) * nby + brick_of(nr.uy, by, nby)) * nbz +
            brick_of(nr.uz, bz, nbz);
        w = make_float4(nr.fx, nr.fy, nr.fz, charge[i]);
        u = nr.ux | (nr.uy << 10) | (nr.uz << 20);
    }
    int slot = 0;
    if (nb <= kBinLds) {
        if (live) lcount[h] = 0;
        __syncthreads();
        int local = 0;
        if (live) local = atomicAdd(&lcount[h], 1);
        __syncthreads();
        if (live && local == 0) lbase[h] = atomicAdd(&bcount[h], lcount[h]);
        __syncthreads();
        if (live) slot = lbase[h] + local;
    } else {
        unsigned long long todo = __ballot(live);
        while (todo) {
            const int leader = __builtin_ctzll(todo);
            const int hb = __builtin_amdgcn_readlane(h, leader);
            const unsigned long long peers = __ballot(h == hb);
            int base = 0;
            if (__lane_id() == leader) base = atomicAdd(&bcount[hb], __popcll(peers));
            base = __builtin_amdgcn_readlane(base, leader);
            if (h == hb) slot = base + lanes_below(peers);
            todo &= ~peers;
        }
    }
    if (!live) return;
    if (slot < bcap) {
        bw[(size_t)h * bcap + slot] = w;
        bu[(size_t)h * bcap + slot] = u;
        bi[(size_t)h * bcap + slot] = i;
    } else {
        const int o = atomicAdd(&bcount[nb], 1);
        ow[o] = w;
        ou[o] = u;
        oi[o] = i;
    }
}

__global__ void __launch_bounds__(256) k_pme_bin(int N, const float* __restrict__ crd,
                                                 const float* __restrict__ charge, LT6 rc, int fx,
                                                 int fy, int fz, int bx, int by, int bz, int nbx,
                                                 int nby, int nbz, int nb, int bcap,
                                                 int* __restrict__ bcount, float4* __restrict__ bw,
                                                 int* __restrict__ bu, int* __restrict__ bi,
                                                 float4* __restrict__ ow, int* __restrict__ ou,
                                                 int* __restrict__ oi, double* __restrict__ acc) {
    pme_bin_body(blockIdx.x, N, crd, charge, rc, fx, fy, fz, bx, by, bz, nbx, nby, nbz, nb, bcap, bcount, bw, bu, bi, ow, ou, oi, acc);
}

// Brick of a workgroup of the spread, tile-reduce and gather launches (one
// workgroup per brick).  Workgroups go to the 8 XCDs in turn, so workgroup i
// takes brick (i % 8) * (nb / 8) + i / 8: each XCD gets a contiguous range of
// bricks (x-major order), whose neighbours' tiles mostly sit in the same L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int brick_for_block(int i, int nb) {
    if (nb % kXcds) return i;
    return (i % kXcds) * (nb / kXcds) + i / kXcds;
}

__device__ __forceinline__ float lane_f(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// k / d for 0 <= k < 2^16 and 1 <= d <= 64 from inv = rcp(d) (relative error
// below 2^-22): (k + 0.5) / d lies at least 0.5 / d from an integer and the
// product errs by less than 2^16 / d * 2^-21, so truncation is exact
__device__ __forceinline__ int div_small(int k, float inv) {
    return (int)(((float)k + 0.5f) * inv);
}

// Tile geometry (host and device).  A tile is sized for the largest brick,
// m = B or the remainder brick: (mx + 3) x (my + 3) x (mz + 3) points, dense
// in global memory.  In LDS the z stride is 1, the y stride mz + 3 rounded up
// to 4 and the x stride is rounded up to 16 mod 32 floats, which puts the 64
// stencil offsets kx*sx + ky*sy + kz of one atom into distinct banks per half
// wavefront for mz = 8.
struct TileGeom {
    int tx, ty, tz;  // tile extent (points)
    int sy, sx;      // LDS strides (floats)
    __host__ __device__ int lds_floats() const { return tx * sx; }
    __host__ __device__ int points() const { return tx * ty * tz; }
};

inline TileGeom tile_geom(int mx, int my, int mz) {
    TileGeom g;
    g.tx = mx + 3;
    g.ty = my + 3;
    g.tz = mz + 3;
    g.sy = (g.tz + 3) / 4 * 4;
    const int p = g.ty * g.sy;
    g.sx = p + ((16 - p % 32) + 32) % 32;
    return g;
}

// One workgroup per brick.  Each lane of a wavefront owns one stencil offset
// (kx, ky, kz); the brick's atoms are taken one at a time with their base
// point in scalar registers, so an atom costs the three weight polynomials and
// one LDS read-add-write per lane, every lane in range.  Every wavefront
// accumulates into its own copy of the tile: the LDS operations of one
// wavefront complete in order, so the plain read-add-write needs no atomic (an
// LDS float atomic costs about three cycles per lane on CDNA4); the copies are
// summed when the tile is stored.
__device__ __forceinline__ void pme_spread_tiles_body(int sa_block, 
    int bx, int by, int bz, int nbx, int nby, int nbz, int nb, int bcap, TileGeom tg,
    const int* __restrict__ bcount, const float4* __restrict__ bw, const int* __restrict__ bu,
    const float4* __restrict__ ow, const int* __restrict__ ou, float* __restrict__ tiles) {
    extern __shared__ float acc[];
    const int b = brick_for_block(sa_block, nb);
    const int cz = b % nbz, cy = (b / nbz) % nby, cx = b / (nbz * nby);
    const int X0 = cx * bx, Y0 = cy * by, Z0 = cz * bz;
    const int nwave = blockDim.x / kWave;
    const int tsize = tg.lds_floats();
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    // the first chunk of the brick's slots is loaded alongside its count
    // (slots below bcap always exist; those past the count are ignored)
    const int g0 = lane * nwave + wave;
    float4 w0 = make_float4(0.f, 0.f, 0.f, 0.f);
    int u0 = 0;
    if (g0 < bcap) {
        w0 = bw[(size_t)b * bcap + g0];
        u0 = bu[(size_t)b * bcap + g0];
    }
    const int n_own = min(bcount[b], bcap);
    const int n_over = bcount[nb];
    for (int k = threadIdx.x; k < nwave * tsize; k += blockDim.x) acc[k] = 0.f;
    __syncthreads();
    float* __restrict__ mine = acc + wave * tsize;
    const int kx = lane >> 4, ky = (lane >> 2) & 3, kz = lane & 3;
    int loff = kx * tg.sx + ky * tg.sy + kz;
    // keep the tile-relative base point scalar: without this the compiler
    // folds the lane offset back into per-lane multiplies
    asm volatile("" : "+v"(loff));
    const BspCoef Cx(kx), Cy(ky), Cz(kz);
    // tile corner: (X0 - 3, Y0 - 3, Z0 - 3)
    const int sbase = (3 - X0) * tg.sx + (3 - Y0) * tg.sy + (3 - Z0);
    const int total = n_own + n_over;
    // atoms dealt to the wavefronts in turn (a brick holds about one
    // wavefront's worth)
    for (int base = 0; base < total; base += nwave * kWave) {
        const int g = base + lane * nwave + wave;
        float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
        int u = 0;
        bool hit = false;
        if (g < n_own) {
            if (base == 0) {
                w = w0;
                u = u0;
            } else {
                w = bw[(size_t)b * bcap + g];
                u = bu[(size_t)b * bcap + g];
            }
            hit = w.w != 0.f;  // an uncharged atom adds nothing
        } else if (g < total) {
            w = ow[g - n_own];
            u = ou[g - n_own];
            hit = w.w != 0.f && brick_of(u & 1023, bx, nbx) == cx &&
                  brick_of((u >> 10) & 1023, by, nby) == cy && brick_of(u >> 20, bz, nbz) == cz;
        }
        unsigned long long m = __ballot(hit);
        while (m) {
            const int src = __builtin_ctzll(m);
            m &= m - 1;
            const int us = __builtin_amdgcn_readlane(u, src);
            const int a0 = sbase + (us & 1023) * tg.sx + ((us >> 10) & 1023) * tg.sy + (us >> 20);
            const float wfx = lane_f(w.x, src);
            const float wfy = lane_f(w.y, src);
            const float wfz = lane_f(w.z, src);
            const float q = lane_f(w.w, src);
            float tq = q * Cx.w(wfx);
            tq = tq * Cy.w(wfy);
            tq = tq * Cz.w(wfz);
            mine[a0 - loff] += tq;
        }
    }
    __syncthreads();
    float* __restrict__ out = tiles + (size_t)b * tg.points();
    const int npts = tg.points();
    const float inv_tz = __builtin_amdgcn_rcpf((float)tg.tz);
    const float inv_ty = __builtin_amdgcn_rcpf((float)tg.ty);
    for (int k = threadIdx.x; k < npts; k += blockDim.x) {
        const int r = div_small(k, inv_tz), z = k - r * tg.tz;
        const int x = div_small(r, inv_ty), y = r - x * tg.ty;
        const int a = x * tg.sx + y * tg.sy + z;
        float v = acc[a];
        for (int w = 1; w < nwave; ++w) v += acc[w * tsize + a];
        out[k] = v;
    }
}

__global__ void __launch_bounds__(256) k_pme_spread_tiles(
    int bx, int by, int bz, int nbx, int nby, int nbz, int nb, int bcap, TileGeom tg,
    const int* __restrict__ bcount, const float4* __restrict__ bw, const int* __restrict__ bu,
    const float4* __restrict__ ow, const int* __restrict__ ou, float* __restrict__ tiles) {
    pme_spread_tiles_body(blockIdx.x, bx, by, bz, nbx, nby, nbz, nb, bcap, tg, bcount, bw, bu, ow, ou, tiles);
}

// Grid point (x, y, z) of brick c at local (lx, ly, lz) sums, per axis, its
// own tile at l + 3 and, for the top 3 points of the brick, the tile of the
// brick above at l - extent + 3 (the same tile when the axis has one brick).
// It also hands the bin counts to the gather (gcount) and clears them for the
// next evaluation's binning (the spread, their other reader, has finished).
__device__ __forceinline__ void pme_tiles_reduce_body(int sa_block, int fx, int fy, int fz, int bx, int by,
                                                          int bz, int nbx, int nby, int nbz,
                                                          int nb, TileGeom tg,
                                                          const float* __restrict__ tiles,
                                                          float* __restrict__ Q,
                                                          int* __restrict__ bcount,
                                                          int* __restrict__ gcount) {
    const int b = brick_for_block(sa_block, nb);
    if (threadIdx.x == 0) {
        gcount[b] = bcount[b];
        bcount[b] = 0;
        if (b == 0) {
            gcount[nb] = bcount[nb];
            bcount[nb] = 0;
        }
    }
    const int cz = b % nbz, cy = (b / nbz) % nby, cx = b / (nbz * nby);
    const int X0 = cx * bx, Y0 = cy * by, Z0 = cz * bz;
    const int ex = cx == nbx - 1 ? fx - X0 : bx;
    const int ey = cy == nby - 1 ? fy - Y0 : by;
    const int ez = cz == nbz - 1 ? fz - Z0 : bz;
    const int ax = cx + 1 == nbx ? 0 : cx + 1;
    const int ay = cy + 1 == nby ? 0 : cy + 1;
    const int az = cz + 1 == nbz ? 0 : cz + 1;
    const size_t tn = (size_t)tg.points();
    const int npts = ex * ey * ez;
    const float inv_ez = __builtin_amdgcn_rcpf((float)ez);
    const float inv_ey = __builtin_amdgcn_rcpf((float)ey);
#pragma unroll 2
    for (int k = threadIdx.x; k < npts; k += blockDim.x) {
        const int r = div_small(k, inv_ez), lz = k - r * ez;
        const int lx = div_small(r, inv_ey), ly = r - lx * ey;
        // the up to 8 covering tiles, loaded independently, summed in a fixed order
        const bool hx = lx >= ex - 3, hy = ly >= ey - 3, hz = lz >= ez - 3;
        // (an absent tile reads the own tile's entry and adds zero, so the
        // loads stay unconditional and all in flight together)
        float t[8];
        size_t at[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int dx = c >> 2, dy = (c >> 1) & 1, dz = c & 1;
            const int bxi = dx ? ax : cx, tx = dx ? lx - ex + 3 : lx + 3;
            const int byi = dy ? ay : cy, ty = dy ? ly - ey + 3 : ly + 3;
            const int bzi = dz ? az : cz, tz = dz ? lz - ez + 3 : lz + 3;
            at[c] = (size_t)((bxi * nby + byi) * nbz + bzi) * tn + (tx * tg.ty + ty) * tg.tz + tz;
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int dx = c >> 2, dy = (c >> 1) & 1, dz = c & 1;
            const bool ok = (!dx || hx) && (!dy || hy) && (!dz || hz);
            const float v = tiles[ok ? at[c] : at[0]];
            t[c] = ok ? v : 0.f;
        }
        const float v = ((t[0] + t[1]) + (t[2] + t[3])) + ((t[4] + t[5]) + (t[6] + t[7]));
        Q[((size_t)(X0 + lx) * fy + (Y0 + ly)) * fz + (Z0 + lz)] = v;
    }
}

__global__ void __launch_bounds__(256) k_pme_tiles_reduce(int fx, int fy, int fz, int bx, int by,
                                                          int bz, int nbx, int nby, int nbz,
                                                          int nb, TileGeom tg,
                                                          const float* __restrict__ tiles,
                                                          float* __restrict__ Q,
                                                          int* __restrict__ bcount,
                                                          int* __restrict__ gcount) {
    pme_tiles_reduce_body(blockIdx.x, fx, fy, fz, bx, by, bz, nbx, nby, nbz, nb, tg, tiles, Q, bcount, gcount);
}

// B-spline weight polynomials M_k and their derivatives dM_k as compile-time
// constants (M_k the same as cMa..cMd) for the thread-per-atom gather
struct BspK {
    static constexpr float a[4] = {(float)(1.0 / 6.0), -0.5f, 0.5f, (float)(-1.0 / 6.0)};
    static constexpr float b[4] = {0.f, 0.5f, -1.f, 0.5f};
    static constexpr float c[4] = {0.f, 0.5f, 0.f, -0.5f};
    static constexpr float d[4] = {0.f, (float)(1.0 / 6.0), (float)(4.0 / 6.0), (float)(1.0 / 6.0)};
    static constexpr float da[4] = {0.5f, -1.5f, 1.5f, -0.5f};
    static constexpr float db[4] = {0.f, 1.f, -2.f, 1.f};
    static constexpr float dc[4] = {0.f, 0.5f, 0.f, -0.5f};
    __device__ static __forceinline__ float w(int k, float f) {
        return fmaf(fmaf(fmaf(a[k], f, b[k]), f, c[k]), f, d[k]);
    }
    __device__ static __forceinline__ float dw(int k, float f) {
        return fmaf(fmaf(da[k], f, db[k]), f, dc[k]);
    }
};

// PME_Final + device_add_force: frc += update_interval * F (PM_force.cpp:
// 852-892, 1014-1022).  One workgroup per brick: the brick's halo tile of the
// convolved grid (the points the stencils of its atoms reach) is read into
// LDS with coalesced loads, then each thread takes one of the brick's atoms, as
// binned by the spread: its 4 x 4 x 4 stencil is read from the tile as 16 rows
// of 4 points and contracted with the weights in registers (per row the z sums
// with M_z and dM_z, then the x / y factors).
constexpr int kGatherThreads = 256;
constexpr int kGatherLoads = 8;

__device__ __forceinline__ void pme_gather_tiles_body(int sa_block, 
    int fx, int fy, int fz, int bx, int by, int bz, int nbx, int nby, int nbz, int nb, int bcap,
    TileGeom tg, const int* __restrict__ gcount, const float4* __restrict__ bw,
    const int* __restrict__ bu, const int* __restrict__ bi, const float4* __restrict__ ow,
    const int* __restrict__ ou, const int* __restrict__ oi, const float* __restrict__ phi, LT6 rc,
    float update_interval, float* __restrict__ frc, const float* __restrict__ extra,
    const float4* __restrict__ base, const int* __restrict__ base_rank, int* __restrict__ reset) {
    extern __shared__ float tile[];
    if (reset && sa_block == 0 && threadIdx.x == 0) *reset = 0;
    const int b = brick_for_block(sa_block, nb);
    const int cz = b % nbz, cy = (b / nbz) % nby, cx = b / (nbz * nby);
    const int X0 = cx * bx, Y0 = cy * by, Z0 = cz * bz;
    const int TX = (cx == nbx - 1 ? fx - X0 : bx) + 3;
    const int TY = (cy == nby - 1 ? fy - Y0 : by) + 3;
    const int TZ = (cz == nbz - 1 ? fz - Z0 : bz) + 3;
    const int npts = TX * TY * TZ;
    const float inv_tz = __builtin_amdgcn_rcpf((float)TZ);
    const float inv_ty = __builtin_amdgcn_rcpf((float)TY);
    // batches of kGatherLoads loads per thread in flight before their LDS writes
    for (int k0 = threadIdx.x; k0 < npts; k0 += blockDim.x * kGatherLoads) {
        float v[kGatherLoads];
        int at[kGatherLoads];
#pragma unroll
        for (int j = 0; j < kGatherLoads; ++j) {
            // points past the tile repeat its last one (loaded, not stored),
            // so the loads stay unconditional and all in flight together
            const int kk = k0 + j * (int)blockDim.x;
            const int k = kk < npts ? kk : npts - 1;
            const int r = div_small(k, inv_tz), tz = k - r * TZ;
            const int tx = div_small(r, inv_ty), ty = r - tx * TY;
            int gx = X0 - 3 + tx, gy = Y0 - 3 + ty, gz = Z0 - 3 + tz;
            gx += gx < 0 ? fx : 0;
            gy += gy < 0 ? fy : 0;
            gz += gz < 0 ? fz : 0;
            v[j] = phi[((size_t)gx * fy + gy) * fz + gz];
            at[j] = kk < npts ? tx * tg.sx + ty * tg.sy + tz : -1;
        }
#pragma unroll
        for (int j = 0; j < kGatherLoads; ++j)
            if (at[j] >= 0) tile[at[j]] = v[j];
    }
    __syncthreads();
    const int n_own = min(gcount[b], bcap);
    const int total = n_own + gcount[nb];
    for (int g = threadIdx.x; g < total; g += blockDim.x) {
        float4 w;
        int u, i;
        if (g < n_own) {
            w = bw[(size_t)b * bcap + g];
            u = bu[(size_t)b * bcap + g];
            i = bi[(size_t)b * bcap + g];
        } else {
            w = ow[g - n_own];
            u = ou[g - n_own];
            i = oi[g - n_own];
            if (brick_of(u & 1023, bx, nbx) != cx || brick_of((u >> 10) & 1023, by, nby) != cy ||
                brick_of(u >> 20, bz, nbz) != cz)
                continue;
        }
        // base point in the tile: stencil point k lies at t - k
        const int rx = (u & 1023) - X0 + 3, ry = ((u >> 10) & 1023) - Y0 + 3, rz = (u >> 20) - Z0 + 3;
        float wx[4], wy[4], wz[4], dx[4], dy[4], dz[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            wx[k] = BspK::w(k, w.x);
            wy[k] = BspK::w(k, w.y);
            wz[k] = BspK::w(k, w.z);
            dx[k] = BspK::dw(k, w.x);
            dy[k] = BspK::dw(k, w.y);
            dz[k] = BspK::dw(k, w.z);
        }
        float gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const float* row = tile + (rx - a) * tg.sx + (ry - c) * tg.sy + rz;
                float sw = 0.f, sd = 0.f;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const float v = row[-k];
                    sw = fmaf(v, wz[k], sw);
                    sd = fmaf(v, dz[k], sd);
                }
                gx = fmaf(dx[a] * wy[c], sw, gx);
                gy = fmaf(wx[a] * dy[c], sw, gy);
                gz = fmaf(wx[a] * wy[c], sd, gz);
            }
        }
        gx *= (float)fx;
        gy *= (float)fy;
        gz *= (float)fz;
        const float dq = -w.w;
        // MultiplyTranspose(g, rcell) = {gx a11, gx a21 + gy a22, gx a31 + gy a32 + gz a33}
        const float ox = dq * (gx * rc.a11);
        const float oy = dq * (gx * rc.a21 + gy * rc.a22);
        const float oz = dq * (gx * rc.a31 + gy * rc.a32 + gz * rc.a33);
        float ex = 0.f, ey = 0.f, ez = 0.f;
        if (extra) {
            ex = extra[3 * i];
            ey = extra[3 * i + 1];
            ez = extra[3 * i + 2];
        }
        // every atom is in exactly one brick's list (own slots or overflow)
        float f0x, f0y, f0z;
        if (base) {
            const float4 f0 = base[base_rank[i]];
            f0x = f0.x;
            f0y = f0.y;
            f0z = f0.z;
        } else {
            f0x = frc[3 * i];
            f0y = frc[3 * i + 1];
            f0z = frc[3 * i + 2];
        }
        frc[3 * i] = f0x + update_interval * ox + ex;
        frc[3 * i + 1] = f0y + update_interval * oy + ey;
        frc[3 * i + 2] = f0z + update_interval * oz + ez;
    }
}

__global__ void __launch_bounds__(kGatherThreads) k_pme_gather_tiles(
    int fx, int fy, int fz, int bx, int by, int bz, int nbx, int nby, int nbz, int nb, int bcap,
    TileGeom tg, const int* __restrict__ gcount, const float4* __restrict__ bw,
    const int* __restrict__ bu, const int* __restrict__ bi, const float4* __restrict__ ow,
    const int* __restrict__ ou, const int* __restrict__ oi, const float* __restrict__ phi, LT6 rc,
    float update_interval, float* __restrict__ frc, const float* __restrict__ extra,
    const float4* __restrict__ base, const int* __restrict__ base_rank, int* __restrict__ reset) {
    pme_gather_tiles_body(blockIdx.x, fx, fy, fz, bx, by, bz, nbx, nby, nbz, nb, bcap, tg, gcount, bw, bu, bi, ow, ou, oi, phi, rc, update_interval, frc, extra, base, base_rank, reset);
}

__global__ void k_pme_charge_sums(int N, const float* __restrict__ q, double* __restrict__ acc) {
    double s2 = 0, s1 = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const float v = q[i];
        s2 += (double)(v * v);
        s1 += (double)v;
    }
    s2 = wave_sum(s2);
    s1 = wave_sum(s1);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        unsafeAtomicAdd(&acc[7], s2);
        unsafeAtomicAdd(&acc[8], s1);
    }
}

// Self, neutralizing and reciprocal energies added to the potential
// (PM_force.cpp:1142-1166).
__global__ void k_pme_energy_final(const double* __restrict__ acc, float* __restrict__ ene,
                                   float* __restrict__ potential, float self_factor,
                                   float neutralizing_factor) {
    const float rec = (float)acc[0] * 0.5f;
    float self = (float)acc[7];
    self = self * self_factor;
    const float sumq = (float)acc[8];
    self += neutralizing_factor * sumq * sumq;
    ene[0] = rec;
    ene[1] = self;
    if (potential) potential[0] += self + rec;
}

__global__ void k_pme_virial_add(const double* __restrict__ acc, float* __restrict__ virial) {
    const int c = threadIdx.x;
    if (c < 6) virial[c] += (float)acc[1 + c];
}

// g(x) = 2/sqrt(pi) x e^{-x^2} - erf(x)
//      = 2/sqrt(pi) sum_{n>=1} (-1)^n 2n / ((2n+1) n!) x^{2n+1},
// series below x = 0.75 (truncation < 1e-9 relative), direct form above.
__device__ __forceinline__ float excl_bracket(float x) {
    if (x < 0.75f) {
        const float x2 = x * x;
        float s = 2.62450659e-7f;         // n = 10
        s = s * x2 - 2.61069340e-6f;      // n = 9
        s = s * x2 + 2.33426704e-5f;      // n = 8
        s = s * x2 - 1.85185185e-4f;      // n = 7
        s = s * x2 + 1.28205128e-3f;      // n = 6
        s = s * x2 - 7.57575758e-3f;      // n = 5
        s = s * x2 + 3.70370370e-2f;      // n = 4
        s = s * x2 - 1.42857143e-1f;      // n = 3
        s = s * x2 + 4.00000000e-1f;      // n = 2
        s = s * x2 - 6.66666667e-1f;      // n = 1
        return kTwoOverSqrtPi * s * x2 * x;
    }
    return x * kTwoOverSqrtPi * expf(-x * x) - erff(x);
}

// MPI_PME_Excluded_Force_With_Atom_Energy_Correction with factor 1
// (PM_force.cpp:1744-1817), one thread per local atom.
template <bool ORTHO>
__device__ __forceinline__ void pme_excluded_body(int sa_block, int N, const float* __restrict__ crd,
                               const float* __restrict__ charge, LT6 c, LT6 r, float beta,
                               const int* __restrict__ ex_start, const int* __restrict__ ex_list,
                               const int* __restrict__ ex_num, float* __restrict__ frc,
                               int need_e, float* __restrict__ atom_ene,
                               float* __restrict__ corr_ene, int need_v,
                               float* __restrict__ atom_virial) {
    const int i = sa_block * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int n = ex_num[i];
    if (n <= 0) return;
    const int s = ex_start[i];
    const float qi = charge[i];
    const float3 r1 = make_float3(crd[3 * i], crd[3 * i + 1], crd[3 * i + 2]);
    float fx = 0, fy = 0, fz = 0, e = 0;
    float v0 = 0, v1 = 0, v2 = 0, v3 = 0, v4 = 0, v5 = 0;
    for (int k = s; k < s + n; ++k) {
        const int j = ex_list[k];
        const float3 r2 = make_float3(crd[3 * j], crd[3 * j + 1], crd[3 * j + 2]);
        const float qj = charge[j];
        const float3 d = disp_pbc<ORTHO>(r2, r1, c, r);
        const float dr2 = norm2_exact(d);
        const float dr_abs = sqrtf(dr2);
        const float br = beta * dr_abs;
        // (2/sqrt(pi) x e^{-x^2} + erfc(x) - 1) / r^3; the bracket cancels to
        // a few percent of its terms for bonded distances, so it is summed as
        // its odd Taylor series there
        float fa = excl_bracket(br) / dr2 / dr_abs;
        fa = -qi * qj * fa;
        const float lx = fa * d.x, ly = fa * d.y, lz = fa * d.z;
        unsafeAtomicAdd(&frc[3 * j], -lx);
        unsafeAtomicAdd(&frc[3 * j + 1], -ly);
        unsafeAtomicAdd(&frc[3 * j + 2], -lz);
        fx += lx;
        fy += ly;
        fz += lz;
        if (need_e) e -= qi * qj * erff(br) / dr_abs;
        if (need_v) {
            v0 -= lx * d.x;
            v1 -= lx * d.y + ly * d.x;
            v2 -= ly * d.y;
            v3 -= lx * d.z + lz * d.x;
            v4 -= ly * d.z + lz * d.y;
            v5 -= lz * d.z;
        }
    }
    unsafeAtomicAdd(&frc[3 * i], fx);
    unsafeAtomicAdd(&frc[3 * i + 1], fy);
    unsafeAtomicAdd(&frc[3 * i + 2], fz);
    if (need_e) {
        unsafeAtomicAdd(&atom_ene[i], e);
        corr_ene[i] = e;
    }
    if (need_v) {
        float* o = atom_virial + (size_t)6 * i;
        unsafeAtomicAdd(&o[0], v0);
        unsafeAtomicAdd(&o[1], v1);
        unsafeAtomicAdd(&o[2], v2);
        unsafeAtomicAdd(&o[3], v3);
        unsafeAtomicAdd(&o[4], v4);
        unsafeAtomicAdd(&o[5], v5);
    }
}

template <bool ORTHO>
__global__ void k_pme_excluded(int N, const float* __restrict__ crd,
                               const float* __restrict__ charge, LT6 c, LT6 r, float beta,
                               const int* __restrict__ ex_start, const int* __restrict__ ex_list,
                               const int* __restrict__ ex_num, float* __restrict__ frc,
                               int need_e, float* __restrict__ atom_ene,
                               float* __restrict__ corr_ene, int need_v,
                               float* __restrict__ atom_virial) {
    pme_excluded_body<ORTHO>(blockIdx.x, N, crd, charge, c, r, beta, ex_start, ex_list, ex_num, frc, need_e, atom_ene, corr_ene, need_v, atom_virial);
}

__global__ void k_sum2_f32(int n, const float* __restrict__ a, const float* __restrict__ b,
                           double* out) {
    double sa_ = 0, sb = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        sa_ += a[i];
        sb += b[i];
    }
    sa_ = wave_sum(sa_);
    sb = wave_sum(sb);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        unsafeAtomicAdd(&out[0], sa_);
        unsafeAtomicAdd(&out[1], sb);
    }
}

// ---------------------------------------------------------------------------
// host set-up

// getb (PM_force.cpp:42-64): |B(k)|^2 = 1 / |sum_kk M4(kk+1) e^{2 pi i k kk/N}|^2
// evaluated in float with the reference's expc / divc (SPONGE/utils/fft.hpp:26-51).
float bspline_M(float u, int n) {
    if (n == 2) {
        if (u > 2 || u < 0) return 0;
        return 1 - fabsf(u - 1);
    }
    return u / (n - 1) * bspline_M(u, n - 1) + (n - u) / (n - 1) * bspline_M(u - 1, n - 1);
}

float getb(int k, int NFFT, int order) {
    float t2r = 0, t2i = 0;
    const float ang0 = 2 * (order - 1) * kPI * k / NFFT;
    const float rr = cosf(ang0), ri = sinf(ang0);
    for (int kk = 0; kk < order - 1; ++kk) {
        const float ang = 2.0f * kPI * k / NFFT * kk;
        const float tf = bspline_M((float)(kk + 1), order);
        t2r += tf * cosf(ang);
        t2i += tf * sinf(ang);
    }
    const float den = t2r * t2r + t2i * t2i;
    const float re = (rr * t2r + ri * t2i) / den, im = (ri * t2r - rr * t2i) / den;
    return re * re + im * im;
}

bool make_plan(int n, sa_radix_plan* p) {
    p->n = n;
    p->nst = 0;
    int m = n;
    for (int f : {4, 2, 3, 5, 7})
        while (m % f == 0) {
            if (p->nst >= 16) return false;
            p->r[p->nst++] = f;
            m /= f;
        }
    return m == 1;
}

int lines_per_block(int n) {
    int t = 2048 / n;
    if (t > 16) t = 16;
    if (t < 2) t = 2;
    return t & ~1;
}

// BC and virial_BC (PM_force.cpp:545-611; recomputed for a new box as
// up_box_bc does, :1171-1238) in the [kz][ky][kx] order of the convolution
// pass, from the module's cell, rcell, volume and beta.
void bc_tables(const sponge_pme* p, std::vector<float>& bct, std::vector<float>& vbct) {
    const int fx = p->fx, fy = p->fy, fz = p->fz;
    std::vector<float> B1(fx), B2(fy), B3(fz);
    for (int k = 0; k < fx; ++k) B1[k] = getb(k, fx, 4);
    for (int k = 0; k < fy; ++k) B2[k] = getb(k, fy, 4);
    for (int k = 0; k < fz; ++k) B3[k] = getb(k, fz, 4);
    bct.assign(p->nfft, 0.f);
    vbct.assign((size_t)6 * p->nfft, 0.f);
    const float mprefactor = kPI * kPI / p->beta / p->beta;
    const LT6& rc = p->rcell;
    for (int kx = 0; kx < fx; ++kx) {
        const float kxrp = kx > fx / 2 ? kx - fx : kx;
        for (int ky = 0; ky < fy; ++ky) {
            const float kyrp = ky > fy / 2 ? ky - fy : ky;
            for (int kz = 0; kz <= fz / 2; ++kz) {
                const float kzrp = kz;
                // m = MultiplyTranspose({kx, ky, kz}, rcell)
                const float mx = kxrp * rc.a11;
                const float my = kxrp * rc.a21 + kyrp * rc.a22;
                const float mz = kxrp * rc.a31 + kyrp * rc.a32 + kzrp * rc.a33;
                const float msq = mx * mx + my * my + mz * mz;
                float bc, v[6];
                if (kx + ky + kz == 0) {
                    bc = 0;
                    for (auto& t : v) t = 0;
                } else {
                    bc = (float)1.0 / kPI / msq * expf(-mprefactor * msq) / p->volume;
                    const float t = 2 / msq * (1 + mprefactor * msq);
                    v[0] = 1 - t * mx * mx;
                    v[1] = 0 - t * my * mx;
                    v[2] = 1 - t * my * my;
                    v[3] = 0 - t * mz * mx;
                    v[4] = 0 - t * mz * my;
                    v[5] = 1 - t * mz * mz;
                }
                bc *= B1[kx] * B2[ky] * B3[kz];
                const size_t ti = ((size_t)kz * fy + ky) * fx + kx;
                bct[ti] = bc;
                for (int c = 0; c < 6; ++c) vbct[(size_t)c * p->nfft + ti] = 0.5f * bc * v[c];
            }
        }
    }
}

void pme_free(sponge_pme* p) {
    for (auto& t : p->d_tw) dfree(t);
    dfree(p->d_Q);
    dfree(p->d_FQ);
    dfree(p->d_FBCFQ);
    dfree(p->d_BCt);
    dfree(p->d_vBCt);
    dfree(p->d_acc);
    dfree(p->d_ene);
    dfree(p->d_direct_atom_energy);
    dfree(p->d_correction_atom_energy);
    dfree(p->d_sum);
    dfree(p->d_bcount);
    dfree(p->d_bw);
    dfree(p->d_bu);
    dfree(p->d_ow);
    dfree(p->d_ou);
    dfree(p->d_tiles);
    dfree(p->d_bi);
    dfree(p->d_oi);
    dfree(p->d_gcount);
}

// Brick edge on one axis: the divisor of the grid size closest to 8 among
// 3..16; for a larger size without one, 8 (the last brick then takes the
// remainder, 9..15 points).
int brick_edge(int f) {
    int best = f > 16 ? 8 : f, score = 1 << 30;
    for (int d = 3; d <= 16 && d <= f; ++d) {
        if (f % d) continue;
        const int sc = d > 8 ? d - 8 : 8 - d;
        if (sc < score) {
            score = sc;
            best = d;
        }
    }
    return best;
}

// tile geometry of the spread: sized for the largest brick of each axis
TileGeom pme_tile_geom(const sponge_pme* p) {
    return tile_geom(std::max(p->bx, p->fx - (p->nbx - 1) * p->bx),
                     std::max(p->by, p->fy - (p->nby - 1) * p->by),
                     std::max(p->bz, p->fz - (p->nbz - 1) * p->bz));
}

}  // namespace

// Register budgets of the batched FFT passes: at 6 wavefronts per SIMD they
// fit the standalone kernels' ~77 VGPRs (left free, the batched kernels took
// 100-118 and ran at 4 wavefronts).  The energy / virial convolution keeps
// the compiler's budget (it would spill).
template <>
struct sa::BodyWaves<&fft_z_r2c_body> {
    static constexpr int value = 6;
};
template <>
struct sa::BodyWaves<&fft_y_body<-1>> {
    static constexpr int value = 6;
};
template <>
struct sa::BodyWaves<&fft_y_body<1>> {
    static constexpr int value = 6;
};
template <>
struct sa::BodyWaves<&fft_x_conv_body<false>> {
    static constexpr int value = 6;
};
template <>
struct sa::BodyWaves<&fft_z_c2r_body> {
    static constexpr int value = 6;
};

extern "C" void sponge_pme_default_options(sponge_pme_options* o) {
    o->fftx = o->ffty = o->fftz = -1;
    o->grid_spacing = 1.0f;
    o->direct_tolerance = 0.00001f;
    o->update_interval = 1;
    o->calculate_reciprocal_part = 1;
    o->calculate_excluded_part = 1;
}

extern "C" float sponge_amd_get_beta(float cutoff, float tolerance) {
    // Get_Beta (PM_force.cpp:246-274).  The direct-space tail erfc(b rc) / rc
    // falls as b grows: an upper end is found by doubling from b = 1, then the
    // bracket [0, upper] is halved 50 times more than the doublings needed,
    // and the last midpoint tried is beta.  The tail is evaluated with the
    // library erfc and kept in float, as are b and the bracket, so that beta
    // comes out bit-identical to the reference's.
    auto tail = [cutoff](float b) -> float { return erfc(b * cutoff) / cutoff; };
    float upper = 1.0f;
    int doublings = 0;
    while (!(tail(upper) <= tolerance)) {
        upper *= 2.0f;
        ++doublings;
    }
    float lower = 0.0f, mid = 0.0f;
    for (int halving = 0; halving < doublings + 50; ++halving) {
        mid = (lower + upper) / 2;
        if (tail(mid) >= tolerance) lower = mid;
        else upper = mid;
    }
    return mid;
}

static int check_2357(int n) {
    if (n <= 0) return 0;
    for (int f : {2, 3, 5, 7})
        while (n % f == 0) n /= f;
    return n == 1;
}

extern "C" int sponge_amd_get_fft_parameter(float length) {
    // Get_Fft_Patameter (SPONGE/common.cpp:639-659)
    int t = (int)ceil(length + 3) >> 2 << 2;
    if (t >= 60 && t <= 68) t = 64;
    else if (t >= 120 && t <= 136) t = 128;
    else if (t >= 240 && t <= 272) t = 256;
    else if (t >= 480 && t <= 544) t = 512;
    else if (t >= 960 && t <= 1088) t = 1024;
    while (!check_2357(t)) t += 4;
    return t;
}

extern "C" int sponge_pme_initial(sponge_pme** out, int atom_numbers, const float* cell,
                                  const float* rcell, float cutoff, const sponge_pme_options* opt,
                                  void* stream) {
    if (!out || !cell || !rcell || atom_numbers <= 0 || !(cutoff > 0))
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    *out = nullptr;
    sponge_pme* p = new sponge_pme();
    if (opt) p->opt = *opt;
    else sponge_pme_default_options(&p->opt);
    p->N = atom_numbers;
    p->cutoff = cutoff;
    p->cell = lt_from(cell);
    p->rcell = lt_from(rcell);
    p->stream = (hipStream_t)stream;
    if (p->opt.update_interval <= 0 || !(p->opt.grid_spacing > 0)) {
        delete p;
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    }
    // grid (PM_force.cpp:383-394): box_length / grid_spacing when not given,
    // box_length being the lengths of the cell vectors (MD_core.cpp:257-262);
    // any lower-triangular cell: fractional coordinates, the BC tables and
    // the gathered forces go through the full rcell
    const LT6& cl = p->cell;
    const bool ortho = lt_is_orthorhombic(cl);
    const float lx = cl.a11;
    const float ly = ortho ? cl.a22 : sqrtf(cl.a21 * cl.a21 + cl.a22 * cl.a22);
    const float lz = ortho ? cl.a33 : sqrtf(cl.a31 * cl.a31 + cl.a32 * cl.a32 + cl.a33 * cl.a33);
    p->fx = p->opt.fftx > 0 ? p->opt.fftx : sponge_amd_get_fft_parameter(lx / p->opt.grid_spacing);
    p->fy = p->opt.ffty > 0 ? p->opt.ffty : sponge_amd_get_fft_parameter(ly / p->opt.grid_spacing);
    p->fz = p->opt.fftz > 0 ? p->opt.fftz : sponge_amd_get_fft_parameter(lz / p->opt.grid_spacing);
    const int dims[3] = {p->fx, p->fy, p->fz};
    for (int d = 0; d < 3; ++d) {
        if (dims[d] < 4 || dims[d] > kMaxFftLen || !make_plan(dims[d], &p->plan[d])) {
            delete p;
            return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
        }
    }
    const int fx = p->fx, fy = p->fy, fz = p->fz;
    p->fzc = fz / 2 + 1;
    p->nall = (long long)fx * fy * fz;
    p->nfft = (long long)fx * fy * p->fzc;
    p->volume = p->cell.a11 * p->cell.a22 * p->cell.a33;
    p->beta = sponge_amd_get_beta(cutoff, p->opt.direct_tolerance);
    p->neutralizing_factor = -0.5 * 3.1415926535897932f / (p->beta * p->beta * p->volume);
    p->tl_z = lines_per_block(fz);
    p->tl_y = lines_per_block(fy);
    p->tl_x = lines_per_block(fx);

    std::vector<float> bct, vbct;
    bc_tables(p, bct, vbct);
    int e = 0;
    for (int d = 0; d < 3; ++d) e |= dmalloc(&p->d_tw[d], dims[d]);
    e |= dmalloc(&p->d_Q, p->nall);
    e |= dmalloc(&p->d_FQ, p->nfft);
    e |= dmalloc(&p->d_FBCFQ, p->nall);
    e |= dmalloc(&p->d_BCt, p->nfft);
    e |= dmalloc(&p->d_vBCt, (size_t)6 * p->nfft);
    e |= dmalloc(&p->d_acc, 9);
    e |= dmalloc(&p->d_ene, 2);
    e |= dmalloc(&p->d_direct_atom_energy, p->N);
    e |= dmalloc(&p->d_correction_atom_energy, p->N);
    e |= dmalloc(&p->d_sum, 2);
    p->bx = brick_edge(p->fx);
    p->by = brick_edge(p->fy);
    p->bz = brick_edge(p->fz);
    p->nbx = p->fx / p->bx;
    p->nby = p->fy / p->by;
    p->nbz = p->fz / p->bz;
    p->nb = p->nbx * p->nby * p->nbz;
    {
        // four times the mean occupancy plus a margin; denser bins go to the
        // overflow list
        const long long mean = ((long long)p->N + p->nb - 1) / p->nb;
        p->bcap = (int)(((4 * mean + 64) + 63) / 64 * 64);
    }
    e |= dmalloc(&p->d_bcount, (size_t)p->nb + 1);
    e |= dmalloc(&p->d_bw, (size_t)p->nb * p->bcap);
    e |= dmalloc(&p->d_bu, (size_t)p->nb * p->bcap);
    e |= dmalloc(&p->d_ow, (size_t)p->N);
    e |= dmalloc(&p->d_ou, (size_t)p->N);
    e |= dmalloc(&p->d_bi, (size_t)p->nb * p->bcap);
    e |= dmalloc(&p->d_oi, (size_t)p->N);
    e |= dmalloc(&p->d_gcount, (size_t)p->nb + 1);
    {
        // tile copies per spread workgroup: as many as let every brick's
        // workgroup be resident at once (160 KB of LDS per CU, 256 CUs), 1..4
        const TileGeom tg = pme_tile_geom(p);
        const long long per_cu = (p->nb + 255) / 256;
        const long long bytes = (long long)tg.lds_floats() * (long long)sizeof(float);
        const long long fit = (160LL * 1024) / (per_cu * bytes);
        p->spread_waves = (int)std::max(1LL, std::min(4LL, fit));
        e |= dmalloc(&p->d_tiles, (size_t)p->nb * tg.points());
    }
    if (e) {
        pme_free(p);
        delete p;
        return SPONGE_AMD_ERROR_MALLOC_FAILED;
    }
    hipStream_t s = p->stream;
    hipError_t he = hipSuccess;
    for (int d = 0; d < 3 && he == hipSuccess; ++d) {
        std::vector<float2> tw(dims[d]);
        for (int t = 0; t < dims[d]; ++t) {
            const double a = -2.0 * M_PI * (double)t / dims[d];
            tw[t] = make_float2((float)cos(a), (float)sin(a));
        }
        he = hipMemcpyAsync(p->d_tw[d], tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice, s);
        if (he == hipSuccess) he = hipStreamSynchronize(s);
    }
    if (he == hipSuccess)
        he = hipMemcpyAsync(p->d_BCt, bct.data(), bct.size() * sizeof(float), hipMemcpyHostToDevice, s);
    if (he == hipSuccess)
        he = hipMemcpyAsync(p->d_vBCt, vbct.data(), vbct.size() * sizeof(float), hipMemcpyHostToDevice, s);
    int fe = 0;
    if (he == hipSuccess) {
        fe |= fill_async(p->d_acc, 0.0, 9, s);
        fe |= fill_async(p->d_bcount, 0, p->nb + 1, s);
        fe |= fill_async(p->d_ene, 0.f, 2, s);
        fe |= fill_async(p->d_Q, 0.f, p->nall, s);
        fe |= fill_async(p->d_FBCFQ, 0.f, p->nall, s);
        fe |= fill_async(p->d_direct_atom_energy, 0.f, p->N, s);
        fe |= fill_async(p->d_correction_atom_energy, 0.f, p->N, s);
        he = hipStreamSynchronize(s);
    }
    if (he != hipSuccess || fe) {
        pme_free(p);
        delete p;
        return he != hipSuccess ? (int)he : fe;
    }
    *out = p;
    return 0;
}

extern "C" int sponge_pme_info(sponge_pme* p, float* beta, int* fft) {
    if (!p) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (beta) *beta = p->beta;
    if (fft) {
        fft[0] = p->fx;
        fft[1] = p->fy;
        fft[2] = p->fz;
    }
    return 0;
}

extern "C" float* sponge_pme_direct_atom_energy(sponge_pme* p) {
    return p ? p->d_direct_atom_energy : nullptr;
}

// (the energy / virial accumulators acc[0..6] are cleared by k_pme_bin)
static int pme_fft_convolve(sponge_pme* p, int need_virial, int need_energy) {
    hipStream_t s = p->stream;
    const int fx = p->fx, fy = p->fy, fz = p->fz, fzc = p->fzc;
    {
        const int TL = p->tl_z;
        const size_t sh = (size_t)2 * ((TL + 1) / 2) * fz * sizeof(float2);
        ScopedTimer timer("pme_fft_z_r2c", s);
        SA_LAUNCH(256, (&fft_z_r2c_body), (k_fft_z_r2c), dim3(grid_for(fy, TL) * fx), dim3(256),
                  sh, s, p->d_Q, p->d_FQ, fx, fy, fz, fzc, TL, p->plan[2], p->d_tw[2]);
        SA_LAUNCH_CHECK();
    }
    {
        const int TL = p->tl_y, nlines = fzc * fx;
        const size_t sh = (size_t)2 * TL * fy * sizeof(float2);
        ScopedTimer timer("pme_fft_y", s);
        SA_LAUNCH(256, (&fft_y_body<-1>), (k_fft_y<-1>), dim3(grid_for(nlines, TL)), dim3(256), sh,
                  s, p->d_FQ, nlines, fy, TL, p->plan[1], p->d_tw[1]);
        SA_LAUNCH_CHECK();
    }
    {
        const int TL = p->tl_x;
        const size_t sh = (size_t)2 * TL * fx * sizeof(float2);
        ScopedTimer timer("pme_fft_x_conv", s);
        const dim3 grid(grid_for(fy, TL) * fzc);
        if (need_energy || need_virial)
            SA_LAUNCH(256, (&fft_x_conv_body<true>), (k_fft_x_conv<true>), grid, dim3(256), sh, s,
                      p->d_FQ, p->d_BCt, p->d_vBCt, p->nfft, fx, fy, fz, fzc, TL, p->plan[0],
                      p->d_tw[0], need_virial, p->d_acc);
        else
            SA_LAUNCH(256, (&fft_x_conv_body<false>), (k_fft_x_conv<false>), grid, dim3(256), sh, s,
                      p->d_FQ, p->d_BCt, p->d_vBCt, p->nfft, fx, fy, fz, fzc, TL, p->plan[0],
                      p->d_tw[0], need_virial, p->d_acc);
        SA_LAUNCH_CHECK();
    }
    {
        const int TL = p->tl_y, nlines = fzc * fx;
        const size_t sh = (size_t)2 * TL * fy * sizeof(float2);
        ScopedTimer timer("pme_ifft_y", s);
        SA_LAUNCH(256, (&fft_y_body<+1>), (k_fft_y<+1>), dim3(grid_for(nlines, TL)), dim3(256), sh,
                  s, p->d_FQ, nlines, fy, TL, p->plan[1], p->d_tw[1]);
        SA_LAUNCH_CHECK();
    }
    {
        const int TL = p->tl_z;
        const size_t sh = ((size_t)TL * fzc + (size_t)2 * ((TL + 1) / 2) * fz) * sizeof(float2);
        ScopedTimer timer("pme_fft_z_c2r", s);
        SA_LAUNCH(256, (&fft_z_c2r_body), (k_fft_z_c2r), dim3(grid_for(fy, TL) * fx), dim3(256),
                  sh, s, p->d_FQ, p->d_FBCFQ, fx, fy, fz, fzc, TL, p->plan[2], p->d_tw[2]);
        SA_LAUNCH_CHECK();
    }
    return 0;
}

// The reciprocal force in two halves, so that a caller can run the first (bin,
// spread, FFT convolution: it reads only the coordinates and charges) on a
// stream of its own beside other work and join before the second (gather,
// energy), which adds to the force array.  Both use the module's stream.
int sa::pme_spread_waves(const sponge_pme* p) { return p->spread_waves; }

void sa::pme_set_spread_waves(sponge_pme* p, int waves) {
    if (waves >= 1 && waves <= p->spread_waves) p->spread_waves = waves;
}

int sa::pme_recip_begin(sponge_pme* p, const float* crd, const float* cell, const float* rcell,
                        const float* charge, int need_virial, int need_energy, int step) {
    const LT6 c = lt_from(cell), r = lt_from(rcell);
    if (memcmp(&c, &p->cell, sizeof(LT6)) != 0) return SPONGE_AMD_ERROR_CONFLICTING_COMMAND;
    if (step % p->opt.update_interval != 0) return 0;
    hipStream_t s = p->stream;
    const int N = p->N;
    {
        ScopedTimer timer("pme_spread", s);
        SA_LAUNCH(256, (&pme_bin_body), (k_pme_bin), dim3(grid_for(N, 256)), dim3(256), 0, s, N, crd, charge, r, p->fx, p->fy, p->fz, p->bx, p->by, p->bz, p->nbx, p->nby, p->nbz, p->nb,
                           p->bcap, p->d_bcount, p->d_bw, p->d_bu, p->d_bi, p->d_ow, p->d_ou,
                           p->d_oi, p->d_acc);
        SA_LAUNCH_CHECK();
        const TileGeom tg = pme_tile_geom(p);
        const size_t sh = sizeof(float) * (size_t)p->spread_waves * tg.lds_floats();
        SA_LAUNCH(256, (&pme_spread_tiles_body), (k_pme_spread_tiles), dim3(p->nb), dim3(kWave * p->spread_waves), sh, s,
                           p->bx, p->by, p->bz, p->nbx, p->nby, p->nbz, p->nb, p->bcap, tg,
                           p->d_bcount, p->d_bw, p->d_bu, p->d_ow, p->d_ou, p->d_tiles);
        SA_LAUNCH_CHECK();
        SA_LAUNCH(256, (&pme_tiles_reduce_body), (k_pme_tiles_reduce), dim3(p->nb), dim3(256), 0, s, p->fx, p->fy, p->fz,
                           p->bx, p->by, p->bz, p->nbx, p->nby, p->nbz, p->nb, tg, p->d_tiles,
                           p->d_Q, p->d_bcount, p->d_gcount);
        SA_LAUNCH_CHECK();
    }
    // the energy sum is kept for the energy of later steps between
    // reciprocal updates, so it is accumulated on every update then
    return pme_fft_convolve(p, need_virial, need_energy || p->opt.update_interval > 1);
}

int sa::pme_recip_end(sponge_pme* p, const float* rcell, const float* charge, float* frc,
                      int need_virial, int need_energy, float* d_virial, float* d_potential,
                      int step, const float* extra, GatherBase gb) {
    const LT6 r = lt_from(rcell);
    hipStream_t s = p->stream;
    const int N = p->N;
    if (step % p->opt.update_interval == 0) {
        if (need_virial) {
            hipLaunchKernelGGL(k_pme_virial_add, dim3(1), dim3(64), 0, s, p->d_acc, d_virial);
            SA_LAUNCH_CHECK();
        }
        ScopedTimer timer("pme_gather", s);
        const TileGeom tg = pme_tile_geom(p);
        SA_LAUNCH(256, (&pme_gather_tiles_body), (k_pme_gather_tiles), dim3(p->nb), dim3(kGatherThreads),
                           sizeof(float) * tg.lds_floats(), s, p->fx, p->fy, p->fz, p->bx, p->by,
                           p->bz, p->nbx, p->nby, p->nbz, p->nb, p->bcap, tg, p->d_gcount, p->d_bw,
                           p->d_bu, p->d_bi, p->d_ow, p->d_ou, p->d_oi, p->d_FBCFQ, r,
                           (float)p->opt.update_interval, frc, extra, gb.base, gb.rank, gb.reset);
        SA_LAUNCH_CHECK();
    }
    if (need_energy) {
        int rc = fill_async(p->d_acc + 7, 0.0, 2, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_pme_charge_sums, dim3(64), dim3(256), 0, s, N, charge, p->d_acc);
        SA_LAUNCH_CHECK();
        const float self_factor = -p->beta / sqrtf(kPI);
        hipLaunchKernelGGL(k_pme_energy_final, dim3(1), dim3(1), 0, s, p->d_acc, p->d_ene,
                           d_potential, self_factor, p->neutralizing_factor);
        SA_LAUNCH_CHECK();
    }
    return 0;
}

hipStream_t sa::pme_stream(const sponge_pme* p) { return p->stream; }
bool sa::pme_calculates_reciprocal(const sponge_pme* p) { return p->opt.calculate_reciprocal_part != 0; }
bool sa::pme_updates_on(const sponge_pme* p, int step) { return step % p->opt.update_interval == 0; }
void sa::pme_set_stream(sponge_pme* p, hipStream_t s) { p->stream = s; }

extern "C" int sponge_pme_reciprocal_force(sponge_pme* p, const float* crd, const float* cell,
                                           const float* rcell, const float* charge, float* frc,
                                           int need_virial, int need_energy, float* d_virial,
                                           float* d_potential, int step) {
    if (!p || !crd || !cell || !rcell || !charge || !frc) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (!p->opt.calculate_reciprocal_part) return 0;
    if ((need_virial && !d_virial) || (need_energy && !d_potential))
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    int rc = sa::pme_recip_begin(p, crd, cell, rcell, charge, need_virial, need_energy, step);
    if (rc) return rc;
    return sa::pme_recip_end(p, rcell, charge, frc, need_virial, need_energy, d_virial,
                             d_potential, step);
}

extern "C" int sponge_pme_excluded_force(sponge_pme* p, int local_atom_numbers,
                                         const int* atom_local, const int* atom_local_id,
                                         const float* crd, const float* cell, const float* rcell,
                                         const float* charge, const int* excluded_list_start,
                                         const int* excluded_list, const int* excluded_numbers,
                                         float* frc, int need_energy, float* atom_energy,
                                         int need_virial, float* atom_virial) {
    (void)atom_local;
    (void)atom_local_id;
    if (!p || !crd || !cell || !rcell || !charge || !frc) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (!p->opt.calculate_excluded_part) return 0;
    if (!excluded_list_start || !excluded_list || !excluded_numbers)
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if ((need_energy && !atom_energy) || (need_virial && !atom_virial))
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (local_atom_numbers <= 0) return 0;
    if (local_atom_numbers > p->N) return SPONGE_AMD_ERROR_CONFLICTING_COMMAND;
    hipStream_t s = p->stream;
    if (need_energy) {
        int rc = fill_async(p->d_correction_atom_energy, 0.f, local_atom_numbers, s);
        if (rc) return rc;
    }
    const LT6 c = lt_from(cell), r = lt_from(rcell);
    dim3 grid(grid_for(local_atom_numbers, 256)), block(256);
    if (lt_is_orthorhombic(c))
        SA_LAUNCH(256, (&pme_excluded_body<true>), (k_pme_excluded<true>), grid, block, 0, s, local_atom_numbers, crd, charge, c, r, p->beta, excluded_list_start, excluded_list, excluded_numbers, frc,
                           need_energy, atom_energy, p->d_correction_atom_energy, need_virial,
                           atom_virial);
    else
        SA_LAUNCH(256, (&pme_excluded_body<false>), (k_pme_excluded<false>), grid, block, 0, s, local_atom_numbers, crd, charge, c, r, p->beta, excluded_list_start, excluded_list,
                           excluded_numbers, frc, need_energy, atom_energy,
                           p->d_correction_atom_energy, need_virial, atom_virial);
    SA_LAUNCH_CHECK();
    return 0;
}

extern "C" int sponge_pme_step_print(sponge_pme* p, float* PM, float* direct, float* reciprocal,
                                     float* self_energy, float* correction) {
    if (!p) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    hipStream_t s = p->stream;
    int rc = fill_async(p->d_sum, 0.0, 2, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sum2_f32, dim3(64), dim3(256), 0, s, p->N, p->d_direct_atom_energy,
                       p->d_correction_atom_energy, p->d_sum);
    SA_LAUNCH_CHECK();
    double hs[2];
    float he[2];
    SA_CHECK(hipMemcpyAsync(hs, p->d_sum, sizeof(hs), hipMemcpyDeviceToHost, s));
    SA_CHECK(hipMemcpyAsync(he, p->d_ene, sizeof(he), hipMemcpyDeviceToHost, s));
    SA_CHECK(hipStreamSynchronize(s));
    const float dir = (float)hs[0], corr = (float)hs[1];
    if (direct) *direct = dir;
    if (correction) *correction = corr;
    if (reciprocal) *reciprocal = he[0];
    if (self_energy) *self_energy = he[1];
    if (PM) *PM = dir + he[0] + he[1] + corr;
    return 0;
}

extern "C" int sponge_pme_grids(sponge_pme* p, float* Q, float* FBCFQ) {
    if (!p) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (Q) SA_CHECK(hipMemcpyAsync(Q, p->d_Q, p->nall * sizeof(float), hipMemcpyDeviceToDevice, p->stream));
    if (FBCFQ)
        SA_CHECK(hipMemcpyAsync(FBCFQ, p->d_FBCFQ, p->nall * sizeof(float), hipMemcpyDeviceToDevice,
                                p->stream));
    return 0;
}

// Particle_Mesh::Update_Box (PM_force.cpp:1240-1256): the BC and virial_BC
// tables and the neutralizing factor of a new box (same grid); synchronises.
extern "C" int sponge_pme_update_box(sponge_pme* p, const float* cell, const float* rcell) {
    if (!p || !cell || !rcell) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    const LT6 c = lt_from(cell);
    p->cell = c;
    p->rcell = lt_from(rcell);
    p->volume = p->cell.a11 * p->cell.a22 * p->cell.a33;
    p->neutralizing_factor = -0.5 * 3.1415926535897932f / (p->beta * p->beta * p->volume);
    std::vector<float> bct, vbct;
    bc_tables(p, bct, vbct);
    SA_CHECK(hipMemcpyAsync(p->d_BCt, bct.data(), bct.size() * sizeof(float), hipMemcpyHostToDevice,
                            p->stream));
    SA_CHECK(hipMemcpyAsync(p->d_vBCt, vbct.data(), vbct.size() * sizeof(float),
                            hipMemcpyHostToDevice, p->stream));
    SA_CHECK(hipStreamSynchronize(p->stream));
    return 0;
}

// Autotuner of the FFT passes' tiling: lines per workgroup of the z, y and
// x passes, each chosen in turn (the others held) as the fastest of a few
// candidates by HIP-event timing of whole reciprocal-force calls on the
// module's stream.  Every tiling computes the same per-line transforms, so
// forces do not depend on the choice.
extern "C" int sponge_pme_autotune(sponge_pme* p, const float* crd, const float* charge, int reps,
                                   int* chosen) {
    if (!p || !crd || !charge || reps < 1) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    hipStream_t s = p->stream;
    float* frc = nullptr;
    int rc = dmalloc(&frc, (size_t)3 * p->N);
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        dfree(frc);
        return SPONGE_AMD_ERROR_MALLOC_FAILED;
    }
    const float* cell = &p->cell.a11;
    const float* rcell = &p->rcell.a11;
    float best_ms = 0.f;
    auto measure = [&](float* ms) -> int {
        int r = sponge_pme_reciprocal_force(p, crd, cell, rcell, charge, frc, 0, 0, nullptr, nullptr, 0);
        if (r) return r;
        SA_CHECK(hipEventRecord(e0, s));
        for (int k = 0; k < reps && !r; ++k)
            r = sponge_pme_reciprocal_force(p, crd, cell, rcell, charge, frc, 0, 0, nullptr, nullptr, 0);
        if (r) return r;
        SA_CHECK(hipEventRecord(e1, s));
        SA_CHECK(hipEventSynchronize(e1));
        SA_CHECK(hipEventElapsedTime(ms, e0, e1));
        return 0;
    };
    int* axes[3] = {&p->tl_z, &p->tl_y, &p->tl_x};
    const int cand[4] = {2, 4, 8, 16};
    rc = measure(&best_ms);
    for (int a = 0; a < 3 && !rc; ++a) {
        const int keep = *axes[a];
        int best = keep;
        for (int c : cand) {
            if (c == keep) continue;
            *axes[a] = c;
            float ms = 0.f;
            rc = measure(&ms);
            if (rc) break;
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
        *axes[a] = best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    dfree(frc);
    if (!rc && chosen) {
        chosen[0] = p->tl_z;
        chosen[1] = p->tl_y;
        chosen[2] = p->tl_x;
    }
    return rc;
}

extern "C" int sponge_pme_set_fft_tiling(sponge_pme* p, const int* tiles) {
    if (!p || !tiles) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    // even line counts up to 16 (the candidates of the autotuner)
    for (int a = 0; a < 3; ++a)
        if (tiles[a] < 2 || tiles[a] > 16 || (tiles[a] & 1)) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    p->tl_z = tiles[0];
    p->tl_y = tiles[1];
    p->tl_x = tiles[2];
    return 0;
}

extern "C" int sponge_pme_clear(sponge_pme* p) {
    if (!p) return 0;
    (void)hipStreamSynchronize(p->stream);
    pme_free(p);
    delete p;
    return 0;
}

extern "C" const char* sponge_amd_version(void) { return "sponge_amd 0.1.0 (gfx950)"; }
