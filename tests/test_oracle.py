"""Pins the oracle (CPU) against known answers, closed forms, brute force and
library routines, so that a plausible mistake in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails here."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from sponge_amd import system as S

HERE = os.path.dirname(os.path.abspath(__file__))


# --------------------------------------------------------------------- boundary
def test_periodic_displacement_known_answers():
    # tests/boundary/test_boundary_policy.cpp:91-104 of the reference
    cell = np.array([10, 0, 20, 0, 0, 30], np.float32)
    rcell = np.array([0.1, 0, 0.05, 0, 0, 1.0 / 30.0], np.float32)
    out = np.zeros(3, np.float32)
    O.lib().orc_displacement(np.array([9, 0, 0], np.float32), np.array([1, 0, 0], np.float32),
                             cell, rcell, out)
    assert out.tolist() == [-2.0, 0.0, 0.0]
    # the updated box of :115-119
    cell[0], rcell[0] = 20.0, 0.05
    O.lib().orc_displacement(np.array([9, 0, 0], np.float32), np.array([1, 0, 0], np.float32),
                             cell, rcell, out)
    assert out.tolist() == [8.0, 0.0, 0.0]


# ------------------------------------------------------------ scalar procedures
@pytest.mark.parametrize("rc,beta", [(8.0, 0.348639), (10.0, 0.275106), (12.0, 0.226635)])
def test_beta_known_answers(rc, beta):
    b = O.get_beta(rc, 1e-5)
    assert abs(b - beta) < 2e-6
    # the defining equation erfc(beta rc)/rc = tol holds at float resolution
    assert abs(math.erfc(b * rc) / rc - 1e-5) < 1e-5 * 1e-4


def test_fft_sizes():
    assert [O.get_fft_parameter(L) for L in (26.9, 62.1, 99.0)] == [28, 64, 100]
    for n in range(1, 600):
        m = n
        for f in (2, 3, 5, 7):
            while m % f == 0:
                m //= f
        assert O.lib().orc_check_2357(n) == (m == 1)


# ----------------------------------------------------------------------- FFT
def test_fft_against_numpy_golden():
    g = np.load(os.path.join(HERE, "golden", "fft_golden.npz"))
    k = 0
    while f"x{k}" in g:
        X = O.fft3d_r2c(g[f"x{k}"])
        assert np.abs(X - g[f"X{k}"]).max() < 1e-9 * np.abs(g[f"X{k}"]).max()
        k += 1
    assert k >= 5


@pytest.mark.parametrize("n", [2, 3, 5, 7, 12, 27, 49, 64, 96, 100, 11, 13])
def test_fft1d_against_numpy(n):
    x = np.random.default_rng(n).normal(size=n) + 1j * np.random.default_rng(n + 1).normal(size=n)
    assert np.abs(O.fft1d(x, -1) - np.fft.fft(x)).max() < 1e-10 * n
    assert np.abs(O.fft1d(x, +1) - np.fft.ifft(x) * n).max() < 1e-10 * n


# ------------------------------------------------------------- neighbor list
def _numpy_half_list(s, crd=None):
    """Independent float32 restatement of the list criterion in numpy
    (IEEE float32 elementwise, no FMA): j > i, min-image r^2 < gl*gl*4."""
    crd = s.crd if crd is None else crd
    c = crd.astype(np.float32)
    L = s.box.astype(np.float32)
    r = np.float32(1.0) / L
    gl = np.float32(0.5) * (np.float32(s.cutoff) + np.float32(s.skin))
    thr = gl * gl * np.float32(4.0)
    excl = set()
    for i in range(s.n_atoms):
        for k in range(s.excl_start[i], s.excl_start[i] + s.excl_num[i]):
            excl.add((i, int(s.excl_list[k])))
    keys = []
    N = len(c)
    for i in range(N):
        d = c[i][None, :] - c[i + 1:]
        d = d - np.floor(d * r + np.float32(0.5)).astype(np.float32) * L
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        js = np.nonzero(r2 < thr)[0] + i + 1
        keys.extend(i * N + int(j) for j in js if (i, int(j)) not in excl)
    return np.array(sorted(keys), np.int64)


def test_bruteforce_list_matches_numpy_restatement():
    s = S.small_water_box(n_side=5, spacing=3.1, cutoff=5.0)
    c, ser, total = O.nl_bruteforce(s)
    want = _numpy_half_list(s)
    assert total == len(want)
    assert np.array_equal(O.pair_array(c, ser), want)


@pytest.mark.parametrize("n_side,spacing,cutoff", [(5, 3.1, 5.0), (6, 3.1, 6.5), (4, 3.2, 3.0)])
def test_cell_list_equals_bruteforce_small_grids(n_side, spacing, cutoff):
    """Cell grids of 3, 4, 5 cells per axis (stencil wraps onto itself)."""
    s = S.small_water_box(n_side=n_side, spacing=spacing, cutoff=cutoff)
    dims = O.nl_grid_dims(s.cell(), s.cutoff, s.skin)
    assert dims.min() >= 3 and dims.max() <= 5
    c, ser, flags = O.nl_build(s)
    cb, sb, total = O.nl_bruteforce(s)
    assert flags == 0
    assert np.array_equal(O.pair_array(c, ser), O.pair_array(cb, sb))


def test_cell_list_equals_bruteforce_cfg1_perturbed():
    s = S.config(1)
    rng = np.random.default_rng(5)
    crd = (s.crd + rng.normal(scale=0.4, size=s.crd.shape)).astype(np.float32)
    crd[::7] += s.box  # unwrapped images
    c, ser, flags = O.nl_build(s, crd=crd)
    cb, sb, total = O.nl_bruteforce(s, crd=crd)
    assert np.array_equal(O.pair_array(c, ser), O.pair_array(cb, sb))
    n_in, n_list = O.pair_counts(s, c, ser, crd=crd)
    assert n_list == total and 0 < n_in < n_list


def test_box_too_small_and_overflow_flags():
    s = S.small_water_box(n_side=3, spacing=3.0, cutoff=4.0)  # 9 A box, gl 3 A -> 3 cells
    c, ser, flags = O.nl_build(s)
    assert flags == 0
    s2 = S.small_water_box(n_side=3, spacing=3.0, cutoff=6.0)  # gl 4 A -> 2 cells
    _, _, flags = O.nl_build(s2)
    assert flags == -1
    # tiny capacities raise the reference's overflow flags
    _, _, flags = O.nl_build(s, max_nb=3)
    assert flags & 2
    _, _, flags = O.nl_build(s, max_in_grid=2)
    assert flags & 1


def test_exclusion_prefilter_uses_first_and_last_entries():
    """The reference drops j only inside [list[first], list[last]]
    (neighbor_list.cpp:751-768): an unsorted exclusion list that is not
    ascending opens an empty window, and the restatement reproduces that."""
    s = S.small_water_box(n_side=3, spacing=3.0, cutoff=4.0)
    base = O.nl_bruteforce(s)[2]
    # atom 0 (O of water 0) lists (2, 1): the window [2, 1] is empty, so
    # neither hydrogen is excluded from its list any more
    s.excl_list[0], s.excl_list[1] = 2, 1
    cb, sb, total = O.nl_bruteforce(s)
    pairs = O.pair_sets(cb, sb)
    assert (0, 1) in pairs and (0, 2) in pairs
    assert total == base + 2
    c, ser, _ = O.nl_build(s)
    assert np.array_equal(O.pair_array(c, ser), O.pair_array(cb, sb))


# ---------------------------------------------------------------- direct space
def test_direct64_two_atom_closed_form():
    """FP64 oracle against the closed forms for one pair."""
    L = 30.0
    r = 3.4
    crd = np.array([[1.0, 2.0, 3.0], [1.0 + r, 2.0, 3.0]])
    s = S.System(name="pair", box=np.array([L] * 3, np.float32), crd=crd.astype(np.float32),
                 vel=np.zeros((2, 3), np.float32), charge=np.array([1.3, -0.7], np.float32),
                 mass=np.ones(2, np.float32), lj_type=np.array([0, 1], np.int32),
                 lj_A=np.array([1.0e5, 2.0e5, 3.0e5], np.float32),
                 lj_B=np.array([100.0, 200.0, 300.0], np.float32),
                 excl_start=np.zeros(2, np.int32), excl_num=np.zeros(2, np.int32),
                 excl_list=np.zeros(0, np.int32), n_waters=0, cutoff=10.0)
    beta = 0.3
    out = O.direct64(s, beta)
    r = float(s.crd[1, 0]) - float(s.crd[0, 0])  # the float32 coordinates are exact inputs
    A, B = 2.0e5, 200.0  # pair index of (0, 1) is 1*2/2 + 0 = 1
    q = float(np.float32(1.3)) * float(np.float32(-0.7))
    e_lj = A / 12 * r ** -12 - B / 6 * r ** -6
    e_c = q * math.erfc(beta * r) / r
    assert abs(out["lj"] - e_lj) < 1e-12 * abs(e_lj)
    assert abs(out["coulomb"] - e_c) < 1e-12 * abs(e_c)
    # F_1 = -dE/dx_1 with E(r): along +x for atom 1
    dEdr = -A * r ** -13 + B * r ** -7 + q * (-2 / math.sqrt(math.pi) * beta * math.exp(-(beta * r) ** 2) / r
                                              - math.erfc(beta * r) / r ** 2)
    assert abs(out["frc"][1, 0] - (-dEdr)) < 1e-9 * abs(dEdr)
    assert abs(out["frc"][0, 0] + out["frc"][1, 0]) < 1e-12
    # virial_xx = r_x F_x of the pair (= -dE/d eps)
    assert abs(out["virial"][0] - r * (-dEdr)) < 1e-9 * abs(r * dEdr)


def test_direct_restatement_vs_fp64_cfg1():
    s = S.config(1)
    beta = O.get_beta(s.cutoff, s.tolerance)
    c, ser, _ = O.nl_build(s)
    r32 = O.lj_direct(s, c, ser, beta)
    r64 = O.direct64(s, beta)
    f32, f64 = r32["frc"].astype(np.float64), r64["frc"]
    assert np.sqrt(((f32 - f64) ** 2).sum() / (f64 ** 2).sum()) < 1e-6
    assert abs(r32["lj_energy"].sum(dtype=np.float64) - r64["lj"]) < 2e-6 * abs(r64["lj"])
    assert abs(r32["direct_energy"].sum(dtype=np.float64) - r64["coulomb"]) < 1e-5 * abs(r64["coulomb"])


def test_direct64_list_equals_brute_force():
    # the list form sums the same pairs as the O(N^2) enumeration
    s = S.config(1)
    beta = O.get_beta(s.cutoff, s.tolerance)
    c, ser, _ = O.nl_build(s)
    a = O.direct64(s, beta)
    b = O.direct64_list(s, beta, c, ser)
    assert np.abs(a["frc"] - b["frc"]).max() < 1e-9 * np.abs(a["frc"]).max()
    assert abs(a["lj"] - b["lj"]) < 1e-10 * abs(a["lj"])
    assert abs(a["coulomb"] - b["coulomb"]) < 1e-10 * abs(a["coulomb"])
    assert np.abs(a["virial"] - b["virial"]).max() < 1e-10 * np.abs(a["virial"]).max()


def test_direct_virial_is_minus_strain_derivative():
    """virial_aa = -dE/d(eps_aa) under an affine stretch of box and atoms."""
    s = S.small_water_box(n_side=4, spacing=3.1, cutoff=5.5)
    beta = 0.4
    base = O.direct64(s, beta)
    eps = 1e-7
    for a, comp in ((0, 0), (1, 2), (2, 5)):
        es = []
        for sgn in (+1, -1):
            t = S.small_water_box(n_side=4, spacing=3.1, cutoff=5.5)
            scale = np.ones(3)
            scale[a] += sgn * eps
            t.box = (s.box.astype(np.float64) * scale).astype(np.float64)
            r = O.direct64(t, beta, crd=s.crd.astype(np.float64) * scale)
            es.append(r["lj"] + r["coulomb"])
        fd = -(es[0] - es[1]) / (2 * eps)
        assert abs(base["virial"][comp] - fd) < 1e-4 * abs(fd) + 1e-6


def test_long_range_factor_closed_form():
    s = S.config(3)
    counts = np.bincount(s.lj_type, minlength=s.n_types).astype(np.float64)
    tot = 0.0
    for a in range(s.n_types):
        for b in range(s.n_types):
            tot += counts[a] * counts[b] * float(s.lj_B[S.lj_pair_index(a, b)])
    want = np.float32(tot) * np.float32(-2.0 / 3.0 * math.pi / s.cutoff ** 3 / 6.0)
    assert abs(O.lj_long_range_factor(s) - want) <= 1e-6 * abs(want)


# ------------------------------------------------------------------ exclusions
def test_excluded64_closed_form_and_restatement():
    s = S.config(1)
    beta = O.get_beta(s.cutoff, s.tolerance)
    r64 = O.excluded64(s, beta)
    # closed form: -sum qq erf(beta r)/r over exclusions
    e = 0.0
    for i in range(s.n_atoms):
        for k in range(s.excl_start[i], s.excl_start[i] + s.excl_num[i]):
            j = s.excl_list[k]
            d = s.crd[j].astype(np.float64) - s.crd[i]
            d -= np.round(d / s.box) * s.box
            r = math.sqrt((d * d).sum())
            e -= float(s.charge[i]) * float(s.charge[j]) * math.erf(beta * r) / r
    assert abs(r64["energy"] - e) < 1e-10 * abs(e)
    r32 = O.pme_excluded(s, beta)
    f32, f64 = r32["frc"].astype(np.float64), r64["frc"]
    assert np.sqrt(((f32 - f64) ** 2).sum() / (f64 ** 2).sum()) < 1e-5
    assert abs(r32["atom_energy"].sum(dtype=np.float64) - e) < 1e-6 * abs(e)


# ------------------------------------------------------------------ reciprocal
def _nacl(n=2, a=5.64):
    pts, q = [], []
    for i in range(2 * n):
        for j in range(2 * n):
            for k in range(2 * n):
                pts.append([i * a / 2, j * a / 2, k * a / 2])
                q.append(1.0 if (i + j + k) % 2 == 0 else -1.0)
    return np.array(pts) + 0.1, np.array(q), np.array([n * a] * 3)


def test_ewald_madelung_constant():
    """Direct + reciprocal + self of an NaCl crystal = -M q^2/r0 per ion pair,
    M = 1.747565 (the Madelung constant of rock salt)."""
    pts, q, box = _nacl()
    beta = 0.9
    s = S.System(name="nacl", box=box.astype(np.float32), crd=pts.astype(np.float32),
                 vel=np.zeros_like(pts, np.float32), charge=q.astype(np.float32),
                 mass=np.ones(len(q), np.float32), lj_type=np.zeros(len(q), np.int32),
                 lj_A=np.zeros(1, np.float32), lj_B=np.zeros(1, np.float32),
                 excl_start=np.zeros(len(q), np.int32), excl_num=np.zeros(len(q), np.int32),
                 excl_list=np.zeros(0, np.int32), n_waters=0, cutoff=float(box[0] / 2 * 0.99))
    d = O.direct64(s, beta, crd=pts)
    r = O.ewald_recip64(pts, q, box, beta, 12)
    E = d["coulomb"] + r["reciprocal"] + r["self_energy"] + r["neutralizing"]
    M = -E / (len(q) / 2) * (5.64 / 2)
    assert abs(M - 1.747565) < 2e-6
    assert np.abs(d["frc"] + r["frc"]).max() < 1e-9  # symmetric sites feel no force


def test_pme64_converges_to_ewald():
    pts, q, box = _nacl()
    rng = np.random.default_rng(1)
    pts = pts + rng.normal(scale=0.3, size=pts.shape)
    beta = 0.5
    ref = O.ewald_recip64(pts, q, box, beta, 10)
    errs = []
    for n in (16, 32, 64):
        p = O.pme64(pts, q, box, beta, (n, n, n))
        errs.append(np.sqrt(((p["frc"] - ref["frc"]) ** 2).sum() / (ref["frc"] ** 2).sum()))
    # order-4 interpolation: about an order of magnitude per halving of h
    assert errs[0] > 5 * errs[1] > 25 * errs[2] and errs[2] < 1e-4


def test_pme_restatement_vs_fp64_and_charge_conservation():
    s = S.config(1)
    beta = O.get_beta(s.cutoff, s.tolerance)
    p = O.PME(s, s.fft, beta)
    r = p.reciprocal()
    g = p.grids()
    # order-4 B-spline weights form a partition of unity: spreading conserves charge
    assert abs(g["Q"].sum(dtype=np.float64) - s.charge.sum(dtype=np.float64)) < \
        1e-6 * np.abs(g["Q"]).sum(dtype=np.float64)
    r64 = O.pme64(s.crd, s.charge, s.box, beta, s.fft)
    f32, f64 = r["frc"].astype(np.float64), r64["frc"]
    assert np.sqrt(((f32 - f64) ** 2).sum() / (f64 ** 2).sum()) < 2e-5
    assert abs(r["reciprocal"] - r64["reciprocal"]) < 1e-5 * abs(r64["reciprocal"])
    # self energy: -beta/sqrt(pi) sum q^2 plus the neutralizing term
    q = s.charge.astype(np.float64)
    V = float(np.prod(s.box.astype(np.float64)))
    self_e = -beta / math.sqrt(math.pi) * (q * q).sum() - math.pi / (2 * beta ** 2 * V) * q.sum() ** 2
    assert abs(r["self_energy"] - self_e) < 1e-6 * abs(self_e)


def _scaled_pme_energy(s, beta, fft, a, eps):
    scale = np.ones(3)
    scale[a] += eps
    return O.pme64(s.crd.astype(np.float64) * scale, s.charge, s.box.astype(np.float64) * scale,
                   beta, fft)["reciprocal"]


def _minus_strain_derivative(s, beta, a, eps=1e-6):
    return -(_scaled_pme_energy(s, beta, s.fft, a, eps) -
             _scaled_pme_energy(s, beta, s.fft, a, -eps)) / (2 * eps)


def test_pme_virial_relation_to_strain_derivative_even_grid():
    """The reference's PME_Sum_Virial (PM_force.cpp:1024-1068) accumulates
    -sum_half w |FQ|^2 virial_BC with virial_BC = BC/2 (I - 2(1 + pi^2 m^2 /
    beta^2) m m^T / m^2) (:586-606).  Since E_rec = 1/2 sum_full BC |FQ|^2,
    that is -1/2 of the strain derivative -dE_rec/d eps, while the direct-space
    virial equals -dE/d eps (test_direct_virial_is_minus_strain_derivative).
    The restatement reproduces the reference's value; the factor -1/2 is a
    property of the reference and only enters its pressure."""
    s = S.small_water_box(n_side=4, spacing=3.1, cutoff=5.5, fft=(16, 16, 16))
    beta = O.get_beta(s.cutoff, s.tolerance)
    r = O.PME(s, s.fft, beta).reciprocal()
    for a, comp in ((0, 0), (1, 2), (2, 5)):
        fd = _minus_strain_derivative(s, beta, a)
        assert abs(r["virial"][comp] - (-0.5 * fd)) < 2e-4 * abs(fd), (comp, r["virial"][comp], fd)


def test_pme_virial_odd_grid_quirk():
    """On an odd fz the reference also halves plane nz = fz/2 (PM_force.cpp:
    1036-1038), which is not self-conjugate there: the relation above then no
    longer holds.  The restatement keeps the quirk."""
    # a coarse grid, where the highest kz plane carries weight
    for fft, odd in (((10, 10, 9), True), ((10, 10, 10), False)):
        s = S.small_water_box(n_side=4, spacing=3.1, cutoff=5.5, fft=fft)
        beta = O.get_beta(s.cutoff, s.tolerance)
        r = O.PME(s, s.fft, beta).reciprocal()
        for a, comp in ((0, 0), (2, 5)):
            fd = _minus_strain_derivative(s, beta, a)
            dev = abs(r["virial"][comp] - (-0.5 * fd)) / abs(fd)
            assert (dev > 3e-3) if odd else (dev < 2e-4), (fft, comp, dev)
