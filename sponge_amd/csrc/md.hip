// NVE step driver for rigid water (+ ions) on one GPU: the caller side of the
// hot path (Main_Calculate_Force + Main_Iteration of SPONGE/main.cpp:1355-1836
// for mode NVE with SETTLE), so that a whole MD step -- neighbor-list check /
// rebuild, excluded-pair correction, LJ + PME direct, long-range correction,
// PME reciprocal, leap-frog and SETTLE -- runs on one stream with no host
// synchronisation and can be captured into one HIP graph.
//
// Leap-frog (SPONGE/MD_core/nve.hpp:3-16): v += dt F / m; x += dt v.
// SETTLE: the analytic rigid-triangle constraint of Miyamoto & Kollman
// (J. Comput. Chem. 13, 952, 1992), written here from the paper: rotate the
// canonical (COM-centred, in-plane) triangle by phi, psi and theta so that the
// constraint displacements are consistent with the previous geometry; the
// velocity correction is (x_constrained - x_unconstrained) / dt, as the
// reference applies it (SPONGE/constrain/settle.cpp:109-214, 216-258).
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"
#include "batch_launch.h"
#include "timing.h"

using namespace sa;

// reciprocal part of a step: PME, the surrogate, or the surrogate checked against PME
enum class StepKind { Pme, Surrogate, Checked };

struct sponge_md {
    int N = 0, n_water = 0;
    float dt = 0;  // internal time units
    float cutoff = 0, skin = 0;
    LT6 cell{}, rcell{};
    float volume = 0;
    hipStream_t stream = nullptr;
    sponge_nl* nl = nullptr;
    sponge_lj* lj = nullptr;
    sponge_pme* pme = nullptr;
    sponge_ljsc* sc = nullptr;  // soft-core module of a lambda window (not owned), or none
    float* d_crd = nullptr;
    float* d_vel = nullptr;
    float* d_frc = nullptr;
    float* d_inv_mass = nullptr;
    float* d_mass = nullptr;
    float* d_charge = nullptr;
    float* d_atom_energy = nullptr;
    float* d_atom_virial = nullptr;
    float* d_potential = nullptr;  // [0] potential of energy steps
    float* d_virial = nullptr;
    int* d_excl_start = nullptr;
    int* d_excl_num = nullptr;
    int* d_excl_list = nullptr;
    int4* d_water = nullptr;       // O, H1, H2 per water
    int n_free = 0;                // atoms outside any water
    int* d_free = nullptr;
    float geom[5] = {0, 0, 0, 0, 0};  // canonical ya, xb, yb, xc, yc
    float mO = 0, mH1 = 0, mH2 = 0;
    double* d_ek = nullptr;
    long long step = 0;
    SegmentedCapture capture;      // the captured steps (one graph, or segments between timers)
    int capture_len = 0;           // steps in the captured cycle
    std::vector<StepKind> capture_kinds;  // reciprocal part of each captured step
    // the same for kGraphCycles cycles in one graph (long runs: fewer graph
    // launches, no gap between the steps of a launch)
    SegmentedCapture capture_long;
    int capture_long_len = 0;
    std::vector<StepKind> capture_long_kinds;
    bool need_cleared = false;     // the pair kernel of the last step cleared the rebuild flag
    // learned multiple-time-step schedule (surrogate between PME steps)
    sponge_rsur* rsur = nullptr;
    int sur_period = 0, sur_check = 0;
    float sur_bound = 0;
    long long sur_steps = 0;
    unsigned long long sur_fallbacks_seen = 0;  // fallbacks read back by the host
    int sur_disabled = 0;          // a check failed: plain PME until the surrogate is set again
    int collect = 0;               // store a training frame on this step
    sponge_rsur* collect_into = nullptr;
    float* d_rec = nullptr;        // [3N] PME reciprocal force of a checked / collected step
    float* d_sur = nullptr;        // [3N] surrogate force of a checked step
    float* d_anchor = nullptr;     // [3N] PME minus surrogate force of the last PME step
    bool anchor_valid = false;
    // 1: impulse schedule (the surrogate's force every step, the slow part
    // period x (F_PME - F_surrogate) as a kick on the PME steps: the
    // reference's update_interval splitting with the surrogate as the fast
    // part); 0: anchored (F_surrogate(t) + F_PME(t0) - F_surrogate(t0))
    int sur_mode = 1;
    double* d_sacc = nullptr;      // [0..2] sum of rec, [3] error^2, [4] reference^2, [5] last error
    unsigned long long* d_scount = nullptr;  // [0] checks, [1] fallbacks
    int use_graph = 0;
    int own_stream = 0;  // the driver created its stream (graph capture needs one)
    // the reciprocal half (spread + FFT convolution) runs on a stream of its
    // own beside the list and pair work of the step, joined before the gather
    hipStream_t recip_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int overlap_mode = -1;  // 1 on, 0 off, -1 from the atom count
    float* d_exfrc = nullptr;  // [N][3] excluded-pair correction computed beside the pair work
    int nl_generation = 0;  // bumped when an overflow re-initialises the list
    float prune_delta = SPONGE_AMD_MD_LIST_PRUNE_DELTA;  // inner-row buffer of the pair kernel
    int overflow_flags = 0;  // every overflow flag seen by the driver's checks
};

namespace {

// cycles of the step schedule captured into one graph for long runs
constexpr int kGraphCycles = 8;

void drop_graphs(sponge_md* m) {
    capture_destroy(m->capture);
    m->capture_len = 0;
    capture_destroy(m->capture_long);
    m->capture_long_len = 0;
}

}  // namespace

namespace {

__device__ __forceinline__ float3 ld3(const float* p, int i) {
    return make_float3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}
__device__ __forceinline__ void st3(float* p, int i, float3 v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
}
__device__ __forceinline__ float3 f3add(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 f3sub(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 f3scale(float s, float3 a) { return make_float3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float f3dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 f3cross(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float3 f3unit(float3 a) { return f3scale(rsqrtf(f3dot(a, a)), a); }

// Canonical geometry of the rigid triangle (COM at the origin, molecule in the
// xy plane): A = (0, ya), B = (xb, yb), C = (xc, yc), masses mA, mB, mC.
struct Triangle {
    float ya, xb, yb, xc, yc, mA, mB, mC;
};

// SETTLE of one water: A1 is the unconstrained position of A, rB and rC the
// unconstrained B - A and C - A (minimum image), B0, C0 the previous B - A and
// C - A.  Writes the constrained positions and adds (x_constrained - x1) / dt
// to the velocities va, vb, vc.
//
// The COM-relative coordinates are formed from the relative vectors only,
// a1 = -(mB rB + mC rC) / M, as the reference does (settle.cpp:234-240):
// taken from absolute positions (~100 A in FP32, ulp 7.6e-6 A) they carry a
// shape error that the velocity correction divides by dt every step, and a
// rigid water then heats up steadily (+12% kinetic energy in 600 steps of
// 2 fs at 99 A, force-free).
__device__ __forceinline__ void settle_one(const int4 t, float3 A1, float3 rB, float3 rC, float3 B0,
                                           float3 C0, float3 va, float3 vb, float3 vc,
                                           const Triangle& g, float inv_dt, float* __restrict__ crd,
                                           float* __restrict__ vel) {
    const float ya = g.ya, xb = g.xb, yb = g.yb, xc = g.xc, yc = g.yc;
    const float mA = g.mA, mB = g.mB, mC = g.mC;
    const float M = mA + mB + mC;
    const float3 a1 = f3scale(-1.0f / M, f3add(f3scale(mB, rB), f3scale(mC, rC)));
    const float3 b1 = f3add(rB, a1), c1 = f3add(rC, a1);
    const float3 com = f3sub(A1, a1);
    // frame: ez normal to the old plane, ex normal to ez and the new A
    const float3 ez = f3unit(f3cross(B0, C0));
    const float3 ex = f3unit(f3cross(a1, ez));
    const float3 ey = f3cross(ez, ex);
    const float xb0 = f3dot(ex, B0), yb0 = f3dot(ey, B0);
    const float xc0 = f3dot(ex, C0), yc0 = f3dot(ey, C0);
    const float za = f3dot(ez, a1);
    const float xb1 = f3dot(ex, b1), yb1 = f3dot(ey, b1), zb = f3dot(ez, b1);
    const float xc1 = f3dot(ex, c1), yc1 = f3dot(ey, c1), zc = f3dot(ez, c1);
    // tilt out of the old plane: phi about x, psi about y
    const float sphi = za / ya;
    const float cphi = sqrtf(fmaxf(0.f, 1.0f - sphi * sphi));
    const float spsi = (zb - zc - (yb - yc) * sphi) / ((xc - xb) * cphi);
    const float cpsi = sqrtf(fmaxf(0.f, 1.0f - spsi * spsi));
    const float ya2 = ya * cphi;
    const float xb2 = xb * cpsi, yb2 = yb * cphi + xb * spsi * sphi;
    const float xc2 = xc * cpsi, yc2 = yc * cphi + xc * spsi * sphi;
    // in-plane rotation theta from the angular-momentum condition
    const float alpha = mB * (xb2 * xb0 + yb2 * yb0) + mC * (xc2 * xc0 + yc2 * yc0);
    const float beta = mB * (yb2 * xb0 - xb2 * yb0) + mC * (yc2 * xc0 - xc2 * yc0);
    const float gamma = mB * (yb1 * xb0 - xb1 * yb0) + mC * (yc1 * xc0 - xc1 * yc0);
    const float ab2 = alpha * alpha + beta * beta;
    const float sth = (alpha * gamma - beta * sqrtf(fmaxf(0.f, ab2 - gamma * gamma))) / ab2;
    const float cth = sqrtf(fmaxf(0.f, 1.0f - sth * sth));
    const float xa3 = -ya2 * sth, ya3 = ya2 * cth;
    const float xb3 = xb2 * cth - yb2 * sth, yb3 = xb2 * sth + yb2 * cth;
    const float xc3 = xc2 * cth - yc2 * sth, yc3 = xc2 * sth + yc2 * cth;
    const float3 a3 = f3add(f3add(f3scale(xa3, ex), f3scale(ya3, ey)), f3scale(za, ez));
    const float3 b3 = f3add(f3add(f3scale(xb3, ex), f3scale(yb3, ey)), f3scale(zb, ez));
    const float3 c3 = f3add(f3add(f3scale(xc3, ex), f3scale(yc3, ey)), f3scale(zc, ez));
    st3(crd, t.x, f3add(com, a3));
    st3(crd, t.y, f3add(com, b3));
    st3(crd, t.z, f3add(com, c3));
    st3(vel, t.x, f3add(va, f3scale(inv_dt, f3sub(a3, a1))));
    st3(vel, t.y, f3add(vb, f3scale(inv_dt, f3sub(b3, b1))));
    st3(vel, t.z, f3add(vc, f3scale(inv_dt, f3sub(c3, c1))));
}

// One MD integration step in one launch: a thread per water remembers its
// geometry, advances its three atoms by leap-frog and applies SETTLE; a
// thread per atom outside any water advances it by leap-frog.  The same
// arithmetic as k_settle_remember + k_leap_frog + k_settle.
__device__ __forceinline__ void leap(float3& x, float3& v, float3 f, float im, float dt) {
    v = f3add(v, f3scale(dt, f3scale(im, f)));
    x = f3add(x, f3scale(dt, v));
}

__device__ __forceinline__ void integrate_body(int sa_block, int n_water, const int4* __restrict__ w,
                                                   int n_free, const int* __restrict__ free_atoms,
                                                   float* __restrict__ crd, float* __restrict__ vel,
                                                   const float* __restrict__ frc,
                                                   const float* __restrict__ inv_mass, float dt,
                                                   LT6 c, LT6 r, Triangle g, float inv_dt) {
    const int k = sa_block * blockDim.x + threadIdx.x;
    if (k < n_water) {
        const int4 t = w[k];
        float3 xa = ld3(crd, t.x), xb = ld3(crd, t.y), xc = ld3(crd, t.z);
        const float3 B0 = disp_pbc<true>(xb, xa, c, r);
        const float3 C0 = disp_pbc<true>(xc, xa, c, r);
        float3 va = ld3(vel, t.x), vb = ld3(vel, t.y), vc = ld3(vel, t.z);
        leap(xa, va, ld3(frc, t.x), inv_mass[t.x], dt);
        leap(xb, vb, ld3(frc, t.y), inv_mass[t.y], dt);
        leap(xc, vc, ld3(frc, t.z), inv_mass[t.z], dt);
        settle_one(t, xa, disp_pbc<true>(xb, xa, c, r), disp_pbc<true>(xc, xa, c, r), B0, C0, va,
                   vb, vc, g, inv_dt, crd, vel);
    } else if (k - n_water < n_free) {
        const int i = free_atoms[k - n_water];
        float3 x = ld3(crd, i), v = ld3(vel, i);
        leap(x, v, ld3(frc, i), inv_mass[i], dt);
        st3(vel, i, v);
        st3(crd, i, x);
    }
}

__global__ void __launch_bounds__(256) k_integrate(int n_water, const int4* __restrict__ w,
                                                   int n_free, const int* __restrict__ free_atoms,
                                                   float* __restrict__ crd, float* __restrict__ vel,
                                                   const float* __restrict__ frc,
                                                   const float* __restrict__ inv_mass, float dt,
                                                   LT6 c, LT6 r, Triangle g, float inv_dt) {
    integrate_body(blockIdx.x, n_water, w, n_free, free_atoms, crd, vel, frc, inv_mass, dt, c, r, g, inv_dt);
}

__global__ void k_add_vec(int n, const float* __restrict__ a, float* __restrict__ frc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) frc[i] += a[i];
}

// check of a surrogate step against PME: acc[3] += |sur - (rec - mean)|^2,
// acc[4] += |rec - mean|^2 (mean from acc[0..2], the sums of rec)
__global__ void __launch_bounds__(256) k_sur_check(int N, const float* __restrict__ sur,
                                                   const float* __restrict__ anchor,
                                                   const float* __restrict__ rec,
                                                   double* __restrict__ acc) {
    __shared__ double s[2][256];
    const double mx = acc[0] / N, my = acc[1] / N, mz = acc[2] / N;
    double e = 0, r = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const double px = rec[3 * i] - mx, py = rec[3 * i + 1] - my, pz = rec[3 * i + 2] - mz;
        double sx = sur[3 * i], sy = sur[3 * i + 1], sz = sur[3 * i + 2];
        if (anchor) {
            sx += anchor[3 * i];
            sy += anchor[3 * i + 1];
            sz += anchor[3 * i + 2];
        }
        const double dx = sx - px, dy = sy - py, dz = sz - pz;
        e += dx * dx + dy * dy + dz * dz;
        r += px * px + py * py + pz * pz;
    }
    s[0][threadIdx.x] = e;
    s[1][threadIdx.x] = r;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s[0][threadIdx.x] += s[0][threadIdx.x + w];
            s[1][threadIdx.x] += s[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&acc[3], s[0][0]);
        atomicAdd(&acc[4], s[1][0]);
    }
}

// fallback decision of a checked step: PME's force when the relative RMS
// error exceeds the bound (or is not a number), the surrogate's otherwise
__global__ void k_sur_select(int N, const float* __restrict__ sur, const float* __restrict__ rec,
                             double* __restrict__ acc, float bound, float* __restrict__ frc,
                             unsigned long long* __restrict__ count) {
    const double err = acc[4] > 0 ? sqrt(acc[3] / acc[4]) : 0.0;
    const bool use_pme = !(err <= (double)bound);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 3 * N) frc[i] += use_pme ? rec[i] : sur[i];
    if (i == 0) {
        count[0] += 1;
        count[1] += use_pme ? 1 : 0;
        acc[5] = err;
    }
}

// anchor = PME - surrogate of this step; the force takes PME's (anchored
// schedule, kick = 1) or the surrogate's plus period x the anchor (impulse
// schedule, kick = period)
__global__ void k_anchor(int n, const float* __restrict__ rec, const float* __restrict__ sur,
                         float* __restrict__ anchor, float* __restrict__ frc, float kick) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = rec[i] - sur[i];
    if (anchor) anchor[i] = a;
    frc[i] += sur[i] + kick * a;
}

__global__ void k_kinetic(int N, const float* __restrict__ vel, const float* __restrict__ mass,
                          double* out) {
    double s = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
        const float3 v = ld3(vel, i);
        s += 0.5 * (double)mass[i] * (double)(v.x * v.x + v.y * v.y + v.z * v.z);
    }
    s = wave_sum(s);
    if ((threadIdx.x & (kWave - 1)) == 0) unsafeAtomicAdd(out, s);
}

__global__ void k_sum_energy(int N, const float* __restrict__ e, double* out) {
    double s = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) s += e[i];
    s = wave_sum(s);
    if ((threadIdx.x & (kWave - 1)) == 0) unsafeAtomicAdd(out, s);
}

// canonical in-plane coordinates of a triangle from its three side lengths
// and masses: COM at the origin, A on the +y axis, and (B - A) x (C - A)
// along +z, the handedness the SETTLE frame (ez ~ B0 x C0) assumes
void triangle_geometry(double dAB, double dAC, double dBC, double mA, double mB, double mC,
                       float* g) {
    // A at the origin, B on -x, C from the law of cosines
    const double bx = -dAB, by = 0.0;
    const double cosA = (dBC * dBC - dAC * dAC - dAB * dAB) / (2 * dAC * dAB);
    const double cx = dAC * cosA, cy = dAC * std::sqrt(1.0 - cosA * cosA);
    const double M = mA + mB + mC;
    const double ox = (mB * bx + mC * cx) / M, oy = (mB * by + mC * cy) / M;
    double ax_ = -ox, ay_ = -oy, bx_ = bx - ox, by_ = by - oy, cx_ = cx - ox, cy_ = cy - oy;
    // rotate so that A lies on the +y axis
    const double ang = std::atan2(ax_, ay_);
    const double cs = std::cos(ang), sn = std::sin(ang);
    auto rot = [&](double x, double y, double& X, double& Y) {
        X = x * cs - y * sn;
        Y = x * sn + y * cs;
    };
    double X, Y;
    rot(ax_, ay_, X, Y);
    g[0] = (float)Y;
    // B sits on -x and C above the x axis before the rotation, so
    // (B - A) x (C - A) points along -z; the reflection x -> -x (which keeps A
    // on the +y axis) restores +z
    rot(bx_, by_, X, Y);
    g[1] = (float)(-X);
    g[2] = (float)Y;
    rot(cx_, cy_, X, Y);
    g[3] = (float)(-X);
    g[4] = (float)Y;
}

void md_free(sponge_md* m) {
    drop_graphs(m);
    sponge_nl_clear(m->nl);
    sponge_lj_clear(m->lj);
    sponge_pme_clear(m->pme);
    m->nl = nullptr;
    m->lj = nullptr;
    m->pme = nullptr;
    dfree(m->d_crd);
    dfree(m->d_vel);
    dfree(m->d_frc);
    dfree(m->d_inv_mass);
    dfree(m->d_mass);
    dfree(m->d_charge);
    dfree(m->d_atom_energy);
    dfree(m->d_atom_virial);
    dfree(m->d_potential);
    dfree(m->d_virial);
    dfree(m->d_excl_start);
    dfree(m->d_excl_num);
    dfree(m->d_excl_list);
    dfree(m->d_water);
    dfree(m->d_free);
    dfree(m->d_ek);
    dfree(m->d_rec);
    dfree(m->d_sur);
    dfree(m->d_anchor);
    dfree(m->d_sacc);
    dfree(m->d_scount);
    if (m->own_stream && m->stream) (void)hipStreamDestroy(m->stream);
    m->stream = nullptr;
    dfree(m->d_exfrc);
    if (m->recip_stream) (void)hipStreamDestroy(m->recip_stream);
    if (m->ev_fork) (void)hipEventDestroy(m->ev_fork);
    if (m->ev_join) (void)hipEventDestroy(m->ev_join);
    m->recip_stream = nullptr;
    m->ev_fork = m->ev_join = nullptr;
}

// FP64 sums of a VECTOR array's components into out[0..2]
__global__ void __launch_bounds__(256) k_sum_vec3(int N, const float* __restrict__ f,
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
            for (int k = 0; k < 3; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) atomicAdd(&out[k], s[k][0]);
}

// Reciprocal part of a step under the learned schedule: PME on every
// period-th step and on energy steps; the surrogate on the others, of which
// the first after the PME step of every check-th period is also evaluated
// with PME (the fallback check).  The surrogate's prediction is anchored on
// the last PME step: F(t) = F_net(t) + (F_PME(t0) - F_net(t0)), so the part
// of the reciprocal force its local features cannot see (the far field) is
// carried over from the exact evaluation and only its change is learned.
StepKind step_kind(const sponge_md* m, long long step, int need_energy) {
    if (!m->rsur || need_energy || m->sur_period < 2 || step % m->sur_period == 0 ||
        !m->anchor_valid)
        return StepKind::Pme;
    if (m->sur_check > 0 && step % m->sur_period == 1 && (step / m->sur_period) % m->sur_check == 0)
        return StepKind::Checked;
    return StepKind::Surrogate;
}

// Steps in one period of the schedule (the unit a graph captures).
int schedule_cycle(const sponge_md* m) {
    if (!m->rsur || m->sur_period < 2) return 1;
    return m->sur_check > 0 ? m->sur_period * m->sur_check : m->sur_period;
}

// One MD step (force evaluation + integration) on m->stream (with the
// reciprocal half of a plain PME step forked onto m->recip_stream).
int md_step(sponge_md* m, long long step, int need_energy, StepKind* kind_out = nullptr,
            bool allow_fork = true) {
    hipStream_t s = m->stream;
    const int N = m->N;
    int rc = 0;
    if (need_energy) {
        rc |= fill_async(m->d_atom_energy, 0.f, N, s);
        rc |= fill_async(m->d_potential, 0.f, 1, s);
    }
    if (rc) return rc;
    const float* cell = &m->cell.a11;
    const float* rcell = &m->rcell.a11;
    const StepKind kind = step_kind(m, step, need_energy);
    if (kind_out) *kind_out = kind;
    // plain PME step: the reciprocal half that reads only the coordinates is
    // forked onto its own stream at the start of the step (its kernels are
    // latency-bound and fill the GPU beside the list and pair kernels).
    // Timers cut captures on the main stream, so with timers on the step
    // stays on one stream; so does a step recorded into the batched driver's
    // capture, where m->stream is itself a stream forked from the capture's
    // origin (the fork is made only from an origin or an eager stream).
    const bool wanted = m->overlap_mode > 0 ||
                        (m->overlap_mode < 0 && m->N >= SPONGE_AMD_MD_OVERLAP_MIN_ATOMS);
    const bool overlap = wanted && allow_fork && kind == StepKind::Pme && !m->rsur &&
                         !m->collect && !g_timing && pme_calculates_reciprocal(m->pme);
    // on a force-only step the excluded-pair correction (which needs only the
    // coordinates and the exclusion lists) joins the side stream too, into its
    // own array that the gather adds; energy steps keep it on the MD stream,
    // where the per-atom energies are summed
    const bool ex_side = overlap && !need_energy && m->d_exfrc && pme_updates_on(m->pme, (int)step);
    if (overlap) {
        SA_CHECK(hipEventRecord(m->ev_fork, s));
        SA_CHECK(hipStreamWaitEvent(m->recip_stream, m->ev_fork, 0));
        pme_set_stream(m->pme, m->recip_stream);
        rc = pme_recip_begin(m->pme, m->d_crd, cell, rcell, m->d_charge, 0, need_energy, (int)step);
        if (!rc && ex_side) rc = fill_async(m->d_exfrc, 0.f, 3 * N, m->recip_stream);
        if (!rc && ex_side)
            rc = sponge_pme_excluded_force(m->pme, N, nullptr, nullptr, m->d_crd, cell, rcell,
                                           m->d_charge, m->d_excl_start, m->d_excl_list,
                                           m->d_excl_num, m->d_exfrc, 0, nullptr, 0, nullptr);
        pme_set_stream(m->pme, s);
        if (rc) return rc;
        SA_CHECK(hipEventRecord(m->ev_join, m->recip_stream));
    }
    // a step recorded for the batched driver marks the same reciprocal chain
    // for the side stream of the batched step (the driver forks it at the
    // step's start and joins it before the gather)
    const bool rec_side = g_recorder && !allow_fork && kind == StepKind::Pme && !m->rsur &&
                          !m->collect && pme_calculates_reciprocal(m->pme) &&
                          pme_updates_on(m->pme, (int)step);
    if (rec_side) {
        g_recorder->lane = 1;
        rc = pme_recip_begin(m->pme, m->d_crd, cell, rcell, m->d_charge, 0, need_energy, (int)step);
        g_recorder->lane = 0;
        if (rc) return rc;
    }
    rc = nl_update_impl(m->nl, nullptr, N, 0, m->d_crd, cell, rcell, (int)step,
                        SPONGE_AMD_NL_CONDITIONAL, m->d_excl_start, m->d_excl_list, m->d_excl_num,
                        m->need_cleared);
    if (rc) return rc;
    // LJ + PME direct first: it assigns the force array (no clearing pass) and
    // clears the list's rebuild flag once the list has been used this step.
    // The terms are summed in another order than Main_Calculate_Force's, which
    // changes only the rounding of the total force.
    bool deferred = false;
    if (m->sc) {
        ljsc_set_stream(m->sc, s);
        rc = ljsc_direct_force_md(m->sc, m->nl, m->d_crd, m->d_charge, m->d_frc, cell, rcell,
                                  m->pme->beta, need_energy, m->d_atom_energy,
                                  m->pme->d_direct_atom_energy, m->nl->d_need);
    } else {
        // with the excluded-pair correction on the side stream the gather is
        // the first pass to write the force array: it assigns the pair
        // kernel's per-row forces with the reciprocal ones (no finalize pass)
        rc = lj_direct_force_md(m->lj, m->nl, m->d_crd, m->d_charge, m->d_frc, cell, rcell,
                                m->pme->beta, need_energy, m->d_atom_energy,
                                m->pme->d_direct_atom_energy, m->nl->d_need,
                                ex_side ? &deferred : nullptr);
    }
    if (rc) return rc;
    m->need_cleared = true;
    if (!ex_side) {
        rc = sponge_pme_excluded_force(m->pme, N, nullptr, nullptr, m->d_crd, cell, rcell,
                                       m->d_charge, m->d_excl_start, m->d_excl_list,
                                       m->d_excl_num, m->d_frc, need_energy, m->d_atom_energy, 0,
                                       nullptr);
        if (rc) return rc;
    }
    if (need_energy) {
        rc = m->sc ? sponge_ljsc_long_range_correction(m->sc, 0, nullptr, 1, m->d_potential,
                                                       m->volume)
                   : sponge_lj_long_range_correction(m->lj, 0, nullptr, 1, m->d_potential,
                                                     m->volume);
        if (rc) return rc;
    }
    const int N3 = 3 * N;
    if (overlap) {
        SA_CHECK(hipStreamWaitEvent(s, m->ev_join, 0));
        GatherBase gb;
        if (deferred) {
            gb.base = m->lj->d_facc;
            gb.rank = m->nl->d_rank;
            gb.reset = m->nl->d_need;
        }
        rc = pme_recip_end(m->pme, rcell, m->d_charge, m->d_frc, 0, need_energy, nullptr,
                           need_energy ? m->d_potential : nullptr, (int)step,
                           ex_side ? m->d_exfrc : nullptr, gb);
        if (rc) return rc;
    } else if (rec_side) {
        g_recorder->join_next = true;
        rc = pme_recip_end(m->pme, rcell, m->d_charge, m->d_frc, 0, need_energy, nullptr,
                           need_energy ? m->d_potential : nullptr, (int)step);
        if (rc) return rc;
    } else if (kind == StepKind::Pme && m->rsur && !m->collect) {
        // exact step of the learned schedule: the anchor of the next steps
        rc = fill_async(m->d_rec, 0.f, N3, s);
        if (!rc)
            rc = sponge_pme_reciprocal_force(m->pme, m->d_crd, cell, rcell, m->d_charge, m->d_rec, 0,
                                             need_energy, nullptr,
                                             need_energy ? m->d_potential : nullptr, (int)step);
        if (!rc)
            rc = rsur_features_sorted(m->rsur, m->nl, m->lj->d_pos, m->d_crd, m->d_charge, cell,
                                      rcell, m->d_excl_start, m->d_excl_list, m->d_excl_num);
        if (!rc) rc = rsur_predict_sorted(m->rsur, m->d_sur, 1, nullptr);
        if (rc) return rc;
        // the impulse schedule kicks only on the period's PME step: an energy
        // step between two of them takes the surrogate's force
        const bool on_period = step % m->sur_period == 0;
        const float kick = m->sur_mode == 0 ? 1.0f : (on_period ? (float)m->sur_period : 0.0f);
        hipLaunchKernelGGL(k_anchor, dim3(grid_for(N3, 256)), dim3(256), 0, s, N3, m->d_rec, m->d_sur,
                           on_period || m->sur_mode == 0 ? m->d_anchor : nullptr, m->d_frc, kick);
        SA_LAUNCH_CHECK();
        m->anchor_valid = true;
    } else if (kind == StepKind::Pme && !m->collect) {
        rc = sponge_pme_reciprocal_force(m->pme, m->d_crd, cell, rcell, m->d_charge, m->d_frc, 0,
                                         need_energy, nullptr, need_energy ? m->d_potential : nullptr,
                                         (int)step);
        if (rc) return rc;
    } else if (kind == StepKind::Pme) {
        // a training frame: the reciprocal force alone, then the features of
        // the same coordinates
        rc = fill_async(m->d_rec, 0.f, N3, s);
        if (!rc)
            rc = sponge_pme_reciprocal_force(m->pme, m->d_crd, cell, rcell, m->d_charge, m->d_rec, 0,
                                             need_energy, nullptr,
                                             need_energy ? m->d_potential : nullptr, (int)step);
        if (!rc)
            rc = rsur_features_sorted(m->collect_into, m->nl, m->lj->d_pos, m->d_crd, m->d_charge,
                                      cell, rcell, m->d_excl_start, m->d_excl_list, m->d_excl_num);
        if (!rc) rc = sponge_rsur_add_frame(m->collect_into, m->d_rec);
        if (rc) return rc;
        hipLaunchKernelGGL(k_add_vec, dim3(grid_for(N3, 256)), dim3(256), 0, s, N3, m->d_rec, m->d_frc);
        SA_LAUNCH_CHECK();
    } else {
        // the surrogate stands in for PME; the list's sorted positions and
        // charges were packed by the pair kernel of this step
        rc = rsur_features_sorted(m->rsur, m->nl, m->lj->d_pos, m->d_crd, m->d_charge, cell, rcell,
                                  m->d_excl_start, m->d_excl_list, m->d_excl_num);
        if (rc) return rc;
        // the anchored schedule adds F_PME(t0) - F_surrogate(t0); the impulse
        // schedule applies the surrogate alone (the slow part was the kick)
        const float* anchor = m->sur_mode == 0 ? m->d_anchor : nullptr;
        if (kind == StepKind::Surrogate) {
            rc = rsur_predict_sorted(m->rsur, m->d_frc, 0, anchor);
            if (rc) return rc;
        } else {
            rc = rsur_predict_sorted(m->rsur, m->d_sur, 1, anchor);
            if (!rc) rc = fill_async(m->d_rec, 0.f, N3, s);
            if (!rc) rc = fill_async(m->d_sacc, 0.0, 5, s);
            if (!rc)
                rc = sponge_pme_reciprocal_force(m->pme, m->d_crd, cell, rcell, m->d_charge, m->d_rec,
                                                 0, 0, nullptr, nullptr, (int)step);
            if (rc) return rc;
            const int blocks = std::min(grid_for(N, 256), 256);
            hipLaunchKernelGGL(k_sum_vec3, dim3(blocks), dim3(256), 0, s, N, m->d_rec, m->d_sacc);
            // the error is that of the anchored force in both schedules
            hipLaunchKernelGGL(k_sur_check, dim3(blocks), dim3(256), 0, s, N, m->d_sur,
                               anchor ? nullptr : m->d_anchor, m->d_rec,
                               m->d_sacc);
            hipLaunchKernelGGL(k_sur_select, dim3(grid_for(N3, 256)), dim3(256), 0, s, N, m->d_sur,
                               m->d_rec, m->d_sacc, m->sur_bound, m->d_frc, m->d_scount);
            SA_LAUNCH_CHECK();
        }
    }
    // integration (Main_Iteration, SPONGE/main.cpp:1702-1774): leap-frog and
    // SETTLE in one launch
    const Triangle g{m->geom[0], m->geom[1], m->geom[2], m->geom[3], m->geom[4], m->mO, m->mH1, m->mH2};
    const int threads = m->n_water + m->n_free;
    if (threads > 0) {
        SA_LAUNCH(256, (&integrate_body), (k_integrate), dim3(grid_for(threads, 256)), dim3(256), 0, s, m->n_water,
                           m->d_water, m->n_free, m->d_free, m->d_crd, m->d_vel, m->d_frc,
                           m->d_inv_mass, m->dt, m->cell, m->rcell, g, 1.0f / m->dt);
        SA_LAUNCH_CHECK();
    }
    return 0;
}

// After steps first..last of the learned schedule: when one of them was a
// checked step, the fallback counter is read (a host synchronisation); a new
// fallback switches the surrogate off (plain PME from the next step on) until
// sponge_md_set_surrogate installs a model again.
int md_check_surrogate(sponge_md* m, long long first, long long last) {
    if (!m->rsur || m->sur_check <= 0 || m->sur_period < 2) return 0;
    bool checked = false;
    for (long long t = first; t <= last && !checked; ++t)
        checked = t % m->sur_period == 1 && (t / m->sur_period) % m->sur_check == 0;
    if (!checked) return 0;
    unsigned long long cnt[2] = {0, 0};
    SA_CHECK(hipMemcpyAsync(cnt, m->d_scount, sizeof(cnt), hipMemcpyDeviceToHost, m->stream));
    SA_CHECK(hipStreamSynchronize(m->stream));
    if (cnt[1] > m->sur_fallbacks_seen) {
        m->sur_fallbacks_seen = cnt[1];
        m->rsur = nullptr;
        m->sur_disabled = 1;
        m->anchor_valid = false;
        drop_graphs(m);
    }
    return 0;
}

// NEIGHBOR_LIST::Check_Overflow at the reference's cadence (neighbor_list.cpp:
// 1027-1103), once steps first..last have been issued: when one of them has
// (step + 1) % check_overflow_interval == 0 the list's flags are read (a host
// synchronisation).  An overflow re-initialises the list with 1.1x the limit
// that overflowed (it is rebuilt by the next step's update, and any captured
// step that held the old list is dropped), or is an error with
// throw_error_when_overflow or when a cell's stream outgrew the 16-bit rows.
int md_check_overflow(sponge_md* m, long long first, long long last) {
    const int iv = m->nl->opt.check_overflow_interval;
    if (iv <= 0 || (last + 1) / iv * iv < first + 1) return 0;
    int flags = 0;
    int rc = nl_take_flags(m->nl, &flags);
    if (rc || !flags) return rc;
    m->overflow_flags |= flags;
    if ((flags & SPONGE_AMD_NL_STREAM_OVERFLOW) || m->nl->opt.throw_error_when_overflow)
        return SPONGE_AMD_ERROR_OVERFLOW;
    rc = nl_reinitialize(&m->nl, flags);
    if (!rc) rc = nl_set_prune(m->nl, m->prune_delta);
    if (rc) return rc;
    drop_graphs(m);
    m->need_cleared = false;
    m->nl_generation++;
    return 0;
}

}  // namespace

extern "C" int sponge_md_initial(sponge_md** out, int atom_numbers, const float* crd,
                                 const float* vel, const float* mass, const float* charge,
                                 int atom_type_numbers, const float* LJ_A, const float* LJ_B,
                                 const int* atom_LJ_type, const int* excluded_list_start,
                                 const int* excluded_numbers, const int* excluded_list,
                                 int n_water, const int* water_atoms, const float* box_length,
                                 float cutoff, float skin, float dt_ps,
                                 const sponge_nl_options* nl_opt,
                                 const sponge_pme_options* pme_opt, void* stream) {
    if (!out || atom_numbers <= 0 || !crd || !vel || !mass || !charge || !box_length ||
        !excluded_list_start || !excluded_numbers || (n_water > 0 && !water_atoms) || !(dt_ps > 0))
        return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    *out = nullptr;
    sponge_md* m = new sponge_md();
    const int N = atom_numbers;
    m->N = N;
    m->n_water = n_water;
    m->dt = dt_ps * 20.455f;  // CONSTANT_TIME_CONVERTION, SPONGE/MD_core/MD_core.cpp:132
    m->cutoff = cutoff;
    m->skin = skin;
    m->stream = (hipStream_t)stream;
    if (!m->stream) {
        // a stream of its own, at the default priority: the NULL stream
        // cannot be captured into a graph
        if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
            delete m;
            return SPONGE_AMD_ERROR_MALLOC_FAILED;
        }
        m->own_stream = 1;
    }
    stream = m->stream;
    // default priority, like every stream of this library
    if (hipStreamCreateWithFlags(&m->recip_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&m->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&m->ev_join, hipEventDisableTiming) != hipSuccess ||
        dmalloc(&m->d_exfrc, (size_t)3 * N)) {
        md_free(m);
        delete m;
        return SPONGE_AMD_ERROR_MALLOC_FAILED;
    }
    const float one = 1.0f;
    m->cell = {box_length[0], 0.f, box_length[1], 0.f, 0.f, box_length[2]};
    m->rcell = {one / box_length[0], 0.f, one / box_length[1], 0.f, 0.f, one / box_length[2]};
    m->volume = box_length[0] * box_length[1] * box_length[2];
    long long E = 0;
    for (int i = 0; i < N; ++i) E += excluded_numbers[i];
    std::vector<float> inv_mass(N);
    for (int i = 0; i < N; ++i) inv_mass[i] = mass[i] > 0 ? 1.0f / mass[i] : 0.0f;
    std::vector<int4> water(n_water);
    if (n_water > 0) {
        for (int k = 0; k < n_water; ++k) {
            water[k] = make_int4(water_atoms[3 * k], water_atoms[3 * k + 1], water_atoms[3 * k + 2], 0);
            if (water[k].x < 0 || water[k].x >= N || water[k].y < 0 || water[k].y >= N ||
                water[k].z < 0 || water[k].z >= N) {
                delete m;
                return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
            }
        }
        // constraint lengths from the first water as given (rigid model)
        const int a = water[0].x, b = water[0].y, c = water[0].z;
        auto dist = [&](int i, int j) {
            double d2 = 0;
            for (int q = 0; q < 3; ++q) {
                double d = (double)crd[3 * i + q] - crd[3 * j + q];
                d2 += d * d;
            }
            return std::sqrt(d2);
        };
        m->mO = mass[a];
        m->mH1 = mass[b];
        m->mH2 = mass[c];
        triangle_geometry(dist(a, b), dist(a, c), dist(b, c), mass[a], mass[b], mass[c], m->geom);
    }
    // atoms outside any water are advanced by leap-frog alone; a water atom
    // listed twice would be advanced twice
    std::vector<int> in_water(N, 0);
    for (int k = 0; k < n_water; ++k) {
        in_water[water[k].x]++;
        in_water[water[k].y]++;
        in_water[water[k].z]++;
    }
    std::vector<int> free_atoms;
    for (int i = 0; i < N; ++i) {
        if (in_water[i] > 1) {
            delete m;
            return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
        }
        if (!in_water[i]) free_atoms.push_back(i);
    }
    m->n_free = (int)free_atoms.size();
    int e = 0;
    e |= dmalloc(&m->d_crd, (size_t)3 * N);
    e |= dmalloc(&m->d_vel, (size_t)3 * N);
    e |= dmalloc(&m->d_frc, (size_t)3 * N);
    e |= dmalloc(&m->d_inv_mass, N);
    e |= dmalloc(&m->d_mass, N);
    e |= dmalloc(&m->d_charge, N);
    e |= dmalloc(&m->d_atom_energy, N);
    e |= dmalloc(&m->d_atom_virial, (size_t)6 * N);
    e |= dmalloc(&m->d_potential, 1);
    e |= dmalloc(&m->d_virial, 6);
    e |= dmalloc(&m->d_excl_start, N);
    e |= dmalloc(&m->d_excl_num, N);
    e |= dmalloc(&m->d_excl_list, E > 0 ? E : 1);
    e |= dmalloc(&m->d_water, n_water > 0 ? n_water : 1);
    e |= dmalloc(&m->d_free, free_atoms.size());
    e |= dmalloc(&m->d_ek, 2);
    if (e) {
        md_free(m);
        delete m;
        return SPONGE_AMD_ERROR_MALLOC_FAILED;
    }
    hipStream_t s = m->stream;
    hipError_t he = hipSuccess;
    auto cp = [&](void* d, const void* h, size_t bytes) {
        if (he == hipSuccess && bytes) he = hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s);
    };
    cp(m->d_crd, crd, sizeof(float) * 3 * N);
    cp(m->d_vel, vel, sizeof(float) * 3 * N);
    cp(m->d_inv_mass, inv_mass.data(), sizeof(float) * N);
    cp(m->d_mass, mass, sizeof(float) * N);
    cp(m->d_charge, charge, sizeof(float) * N);
    cp(m->d_excl_start, excluded_list_start, sizeof(int) * N);
    cp(m->d_excl_num, excluded_numbers, sizeof(int) * N);
    if (E > 0) cp(m->d_excl_list, excluded_list, sizeof(int) * E);
    if (n_water > 0) cp(m->d_water, water.data(), sizeof(int4) * n_water);
    if (m->n_free > 0) cp(m->d_free, free_atoms.data(), sizeof(int) * m->n_free);
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    int rc = he != hipSuccess ? (int)he : 0;
    if (!rc) rc = sponge_nl_initial(&m->nl, N, cutoff, skin, &m->cell.a11, &m->rcell.a11, nl_opt, stream);
    if (!rc) rc = nl_set_prune(m->nl, m->prune_delta);
    if (!rc) rc = sponge_lj_initial(&m->lj, N, atom_type_numbers, LJ_A, LJ_B, atom_LJ_type, cutoff, stream);
    if (!rc) rc = sponge_pme_initial(&m->pme, N, &m->cell.a11, &m->rcell.a11, cutoff, pme_opt, stream);
    if (rc) {
        md_free(m);
        delete m;
        return rc;
    }
    *out = m;
    return 0;
}

extern "C" int sponge_md_run(sponge_md* m, int steps, int energy_interval, int use_graph) {
    if (!m || steps < 0) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    const int cyc = schedule_cycle(m);
    int k = 0;
    while (k < steps) {
        const long long st = m->step;
        // graph replay needs the fixed default schedule (automatic list
        // refresh, no impulse interval on the reciprocal part) and replays
        // whole cycles of the learned schedule that hold no energy step
        bool graph = use_graph && m->nl->built && m->nl->opt.refresh_interval <= 0 &&
                     m->pme->opt.update_interval == 1 && st % cyc == 0 && steps - k >= cyc;
        for (long long t = st; graph && energy_interval > 0 && t < st + cyc; ++t)
            if (t % energy_interval == 0) graph = false;
        if (graph) {
            // kGraphCycles cycles per launch while the run has that many left
            // (and no energy step among them), else one cycle per launch; the
            // learned schedule keeps one cycle per launch, so that a failed
            // check switches the surrogate off before the next cycle
            bool long_run = !m->rsur && steps - k >= kGraphCycles * cyc;
            for (long long t = st; long_run && energy_interval > 0 && t < st + kGraphCycles * cyc; ++t)
                if (t % energy_interval == 0) long_run = false;
            SegmentedCapture& cap = long_run ? m->capture_long : m->capture;
            int& cap_len = long_run ? m->capture_long_len : m->capture_len;
            std::vector<StepKind>& kinds = long_run ? m->capture_long_kinds : m->capture_kinds;
            const int len = long_run ? kGraphCycles * cyc : cyc;
            // timers switched on or off, or another cycle, since the capture
            if (!cap.items.empty() && (cap.timing != g_timing || cap_len != len))
                capture_destroy(cap);
            if (cap.items.empty()) {
                const int rc0 = capture_begin(cap, m->stream);
                if (rc0) return rc0;
                int rc = 0;
                kinds.clear();
                for (int c = 0; c < len && !rc; ++c) {
                    StepKind kd = StepKind::Pme;
                    rc = md_step(m, st + c, 0, &kd);
                    kinds.push_back(kd);
                }
                const int rc2 = capture_end(cap, m->stream);
                if (rc || rc2) {
                    capture_destroy(cap);
                    (void)hipGetLastError();
                    return rc ? rc : rc2;
                }
                cap_len = len;
            }
            int rc = capture_launch(cap, m->stream);
            if (rc) return rc;
            for (StepKind kd : kinds)
                if (kd != StepKind::Pme) m->sur_steps++;
            m->step += len;
            k += len;
            rc = md_check_overflow(m, st, st + len - 1);
            if (!rc) rc = md_check_surrogate(m, st, st + len - 1);
            if (rc) return rc;
        } else {
            const int need_e = energy_interval > 0 && st % energy_interval == 0;
            StepKind kd = StepKind::Pme;
            int rc = md_step(m, st, need_e, &kd);
            if (rc) return rc;
            if (kd != StepKind::Pme) m->sur_steps++;
            m->step++;
            k++;
            rc = md_check_overflow(m, st, st);
            if (!rc && kd == StepKind::Checked) rc = md_check_surrogate(m, st, st);
            if (rc) return rc;
        }
    }
    return 0;
}

namespace {

int ensure_surrogate_buffers(sponge_md* m) {
    int e = 0;
    if (!m->d_rec) e |= dmalloc(&m->d_rec, (size_t)3 * m->N);
    if (!m->d_sur) e |= dmalloc(&m->d_sur, (size_t)3 * m->N);
    if (!m->d_anchor) e |= dmalloc(&m->d_anchor, (size_t)3 * m->N);
    if (!m->d_sacc) e |= dmalloc(&m->d_sacc, 8);
    if (!m->d_scount) {
        e |= dmalloc(&m->d_scount, 2);
        if (!e) e = fill_async(m->d_scount, 0ull, 2, m->stream);
        if (!e) e = fill_async(m->d_sacc, 0.0, 8, m->stream);
    }
    return e;
}

int check_surrogate(const sponge_md* m, const sponge_rsur* rsur) {
    if (rsur_atoms(rsur) != m->N) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    // every partner within the feature cutoff must be on the list
    if (rsur_cutoff(rsur) > m->cutoff) return SPONGE_AMD_ERROR_CONFLICTING_COMMAND;
    return 0;
}

}  // namespace

namespace sa {

// hooks of the batched driver (batch.hip): a system's stream, whether its
// schedule can be captured, the capture of k steps on its stream, and the
// step counter advanced by replays
hipStream_t md_stream(const sponge_md* m) { return m->stream; }

bool md_graphable(const sponge_md* m) {
    return m->nl->built && m->nl->opt.refresh_interval <= 0 && m->pme->opt.update_interval == 1 &&
           !m->rsur;
}

int md_record_steps(sponge_md* m, int k) {
    for (int c = 0; c < k; ++c) {
        const int rc = md_step(m, m->step + c, 0, nullptr, false);
        if (rc) return rc;
    }
    return 0;
}

void md_advance(sponge_md* m, int k) { m->step += k; }

int md_overflow_check(sponge_md* m, long long first, long long last) {
    return md_check_overflow(m, first, last);
}

long long md_step_count(const sponge_md* m) { return m->step; }
int md_generation(const sponge_md* m) { return m->nl_generation; }

int md_state_arrays(sponge_md* m, float** crd, float** charge) {
    *crd = m->d_crd;
    *charge = m->d_charge;
    return 0;
}

}  // namespace sa

extern "C" int sponge_md_set_list_prune(sponge_md* m, float delta) {
    if (!m || !(delta >= 0.f) || delta > m->skin) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    const int rc = nl_set_prune(m->nl, delta);
    if (rc) return rc;
    m->prune_delta = delta;
    drop_graphs(m);  // the captured step reads the other rows
    return 0;
}

extern "C" int sponge_md_request_list_update(sponge_md* m) {
    if (!m) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (!m->nl->built) return 0;  // the first step builds it anyway
    return fill_async(m->nl->d_need, 1, 1, m->stream);
}

extern "C" int sponge_md_set_reciprocal_overlap(sponge_md* m, int mode) {
    if (!m || mode < -1 || mode > 1) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    m->overlap_mode = mode;
    drop_graphs(m);
    return 0;
}

// FFT tiling of the driver's PME module tuned on its current coordinates
// (sponge_pme_autotune); the captured step is rebuilt with the new tiling.
extern "C" int sponge_md_autotune(sponge_md* m, int reps, int* tiles) {
    if (!m || reps < 1) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    drop_graphs(m);
    return sponge_pme_autotune(m->pme, m->d_crd, m->d_charge, reps, tiles);
}

extern "C" int sponge_md_set_soft_core(sponge_md* m, sponge_ljsc* sc) {
    if (!m) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (sc && (ljsc_atoms(sc) != m->N || m->rsur)) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    m->sc = sc;
    drop_graphs(m);  // the captured step calls the other module
    return 0;
}

extern "C" int sponge_md_set_surrogate(sponge_md* m, sponge_rsur* rsur, int period,
                                       int check_interval, float bound) {
    if (!m) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (rsur && m->sc) return SPONGE_AMD_ERROR_NOT_IMPLEMENTED;
    drop_graphs(m);
    m->anchor_valid = false;
    if (!rsur) {
        m->rsur = nullptr;
        return 0;
    }
    if (period < 1 || check_interval < 0 || !(bound > 0.f)) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    int rc = check_surrogate(m, rsur);
    if (!rc) rc = ensure_surrogate_buffers(m);
    if (rc) return rc;
    rsur_set_stream(rsur, m->stream);
    // fallbacks so far, so that only a new one switches the surrogate off
    unsigned long long cnt[2] = {0, 0};
    SA_CHECK(hipMemcpyAsync(cnt, m->d_scount, sizeof(cnt), hipMemcpyDeviceToHost, m->stream));
    SA_CHECK(hipStreamSynchronize(m->stream));
    m->sur_fallbacks_seen = cnt[1];
    m->sur_disabled = 0;
    m->rsur = rsur;
    m->anchor_valid = false;
    m->sur_period = period;
    m->sur_check = check_interval;
    m->sur_bound = bound;
    return 0;
}

extern "C" int sponge_md_set_surrogate_mode(sponge_md* m, int mode) {
    if (!m || (mode != 0 && mode != 1)) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    m->sur_mode = mode;
    drop_graphs(m);
    m->anchor_valid = false;  // the next step is a PME step of the new schedule
    return 0;
}

extern "C" int sponge_md_surrogate_stats(sponge_md* m, long long* out, double* last_error) {
    if (!m) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    unsigned long long cnt[2] = {0, 0};
    double err = 0;
    if (m->d_scount) {
        SA_CHECK(hipMemcpyAsync(cnt, m->d_scount, sizeof(cnt), hipMemcpyDeviceToHost, m->stream));
        SA_CHECK(hipMemcpyAsync(&err, m->d_sacc + 5, sizeof(double), hipMemcpyDeviceToHost, m->stream));
        SA_CHECK(hipStreamSynchronize(m->stream));
    }
    if (out) {
        out[0] = m->sur_steps;
        out[1] = (long long)cnt[0];
        out[2] = (long long)cnt[1];
        out[3] = m->sur_disabled;
    }
    if (last_error) *last_error = err;
    return 0;
}

extern "C" int sponge_md_collect_frames(sponge_md* m, sponge_rsur* rsur, int n_frames, int stride) {
    if (!m || !rsur || n_frames < 0 || stride < 1) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    int rc = check_surrogate(m, rsur);
    if (!rc) rc = ensure_surrogate_buffers(m);
    if (rc) return rc;
    rsur_set_stream(rsur, m->stream);
    // plain PME steps while collecting
    sponge_rsur* saved = m->rsur;
    m->rsur = nullptr;
    m->collect_into = rsur;
    for (int f = 0; f < n_frames && !rc; ++f) {
        for (int k = 0; k < stride && !rc; ++k) {
            m->collect = k == stride - 1;
            rc = md_step(m, m->step, 0);
            if (!rc) rc = md_check_overflow(m, m->step, m->step);
            if (!rc) m->step++;
        }
    }
    m->collect = 0;
    m->collect_into = nullptr;
    m->rsur = saved;
    if (!rc) SA_CHECK(hipStreamSynchronize(m->stream));
    return rc;
}

extern "C" int sponge_md_energies(sponge_md* m, double* out) {
    // out: [0] potential (last energy step), [1] kinetic, [2] LJ_short,
    // [3] LJ_long, [4] PM, [5] PM_direct, [6] PM_reciprocal, [7] PM_self,
    // [8] PM_correction, [9] sum of per-atom energies of the last energy step
    if (!m || !out) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    hipStream_t s = m->stream;
    int rc = fill_async(m->d_ek, 0.0, 2, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_kinetic, dim3(64), dim3(256), 0, s, m->N, m->d_vel, m->d_mass, m->d_ek);
    hipLaunchKernelGGL(k_sum_energy, dim3(64), dim3(256), 0, s, m->N, m->d_atom_energy, m->d_ek + 1);
    SA_LAUNCH_CHECK();
    double ek[2];
    float pot = 0;
    SA_CHECK(hipMemcpyAsync(ek, m->d_ek, sizeof(ek), hipMemcpyDeviceToHost, s));
    SA_CHECK(hipMemcpyAsync(&pot, m->d_potential, sizeof(float), hipMemcpyDeviceToHost, s));
    SA_CHECK(hipStreamSynchronize(s));
    float lj_short = 0, lj_long = 0, pm[5] = {0, 0, 0, 0, 0};
    if (m->sc) {
        ljsc_set_stream(m->sc, s);
        rc = sponge_ljsc_step_print(m->sc, &lj_short);
        lj_long = ljsc_long_range_factor(m->sc) / m->volume;
    } else {
        rc = sponge_lj_step_print(m->lj, &lj_short, &lj_long);
    }
    if (!rc) rc = sponge_pme_step_print(m->pme, &pm[0], &pm[1], &pm[2], &pm[3], &pm[4]);
    if (rc) return rc;
    out[0] = (double)pot + ek[1];
    out[1] = ek[0];
    out[2] = lj_short;
    out[3] = lj_long;
    for (int k = 0; k < 5; ++k) out[4 + k] = pm[k];
    out[9] = ek[1];
    return 0;
}

extern "C" int sponge_md_get_state(sponge_md* m, float* crd, float* vel, float* frc) {
    if (!m) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    hipStream_t s = m->stream;
    const size_t b = sizeof(float) * 3 * m->N;
    if (crd) SA_CHECK(hipMemcpyAsync(crd, m->d_crd, b, hipMemcpyDeviceToHost, s));
    if (vel) SA_CHECK(hipMemcpyAsync(vel, m->d_vel, b, hipMemcpyDeviceToHost, s));
    if (frc) SA_CHECK(hipMemcpyAsync(frc, m->d_frc, b, hipMemcpyDeviceToHost, s));
    SA_CHECK(hipStreamSynchronize(s));
    return 0;
}

extern "C" int sponge_md_set_state(sponge_md* m, const float* crd, const float* vel) {
    if (!m) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    hipStream_t s = m->stream;
    const size_t b = sizeof(float) * 3 * m->N;
    if (crd) SA_CHECK(hipMemcpyAsync(m->d_crd, crd, b, hipMemcpyHostToDevice, s));
    if (vel) SA_CHECK(hipMemcpyAsync(m->d_vel, vel, b, hipMemcpyHostToDevice, s));
    SA_CHECK(hipStreamSynchronize(s));
    return 0;
}

extern "C" int sponge_md_handles(sponge_md* m, sponge_nl** nl, sponge_lj** lj, sponge_pme** pme) {
    if (!m) return SPONGE_AMD_ERROR_VALUE_ERROR_COMMAND;
    if (nl) *nl = m->nl;
    if (lj) *lj = m->lj;
    if (pme) *pme = m->pme;
    return 0;
}

extern "C" int sponge_md_clear(sponge_md* m) {
    if (!m) return 0;
    (void)hipStreamSynchronize(m->stream);
    md_free(m);
    delete m;
    return 0;
}
