"""Synthetic periodic water / ion systems in SPONGE's internal units.

Units follow the reference (``SPONGE/common.h:119-138``): lengths in Angstrom,
energies in kcal/mol, masses in amu, charges pre-multiplied by 18.2223 so that
``q_i q_j / r`` is kcal/mol, and time in 1/20.455 ps.  LJ pair tables are kept
in the in-memory convention of the reference loader
(``SPONGE/xponge/load/native/lj.hpp:47,57``): the file stores ``4 eps sigma^12``
and ``4 eps sigma^6`` and the loader multiplies them by 12 and 6, so that
``E_LJ = A/12 r^-12 - B/6 r^-6``.

Every system is built from a fixed seed, so the same inputs can be regenerated
anywhere (this module has no dependency on the GPU library).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

CHARGE_SCALE = 18.2223          # SPONGE/common.h:134
TIME_CONVERSION = 20.455        # SPONGE/common.h:138
KB = 0.00198716                 # SPONGE/common.h:124, kcal/mol/K
KJ_PER_KCAL = 4.184

OH_BOND = 0.9572                # Angstrom
HOH_ANGLE = 104.52              # degrees
Q_O, Q_H = -0.834, 0.417
M_O, M_H = 16.00, 1.008
M_NA, M_CL = 22.99, 35.45

# (sigma Angstrom, epsilon kcal/mol) per LJ type; H carries no LJ.
LJ_O = (3.1506, 0.6364 / KJ_PER_KCAL)
LJ_H = (0.0, 0.0)
LJ_NA = (3.33, 0.0116 / KJ_PER_KCAL)
LJ_CL = (4.40, 0.4186 / KJ_PER_KCAL)


def lj_pair_index(a: int, b: int) -> int:
    """Lower-triangle packed pair index, as ``Get_LJ_Type``
    (``SPONGE/Lennard_Jones_force/Lennard_Jones_force.h:17-28``)."""
    hi, lo = max(a, b), min(a, b)
    return hi * (hi + 1) // 2 + lo


@dataclasses.dataclass
class System:
    """One periodic system, arrays in the reference's host layout."""
    name: str
    box: np.ndarray               # (3,) float32 orthorhombic box lengths
    crd: np.ndarray               # (N,3) float32
    vel: np.ndarray               # (N,3) float32, Angstrom per internal time unit
    charge: np.ndarray            # (N,) float32, scaled by 18.2223
    mass: np.ndarray              # (N,) float32
    lj_type: np.ndarray           # (N,) int32
    lj_A: np.ndarray              # (T(T+1)/2,) float32, 12 * 4 eps sigma^12
    lj_B: np.ndarray              # (T(T+1)/2,) float32, 6 * 4 eps sigma^6
    excl_start: np.ndarray        # (N,) int32
    excl_num: np.ndarray          # (N,) int32
    excl_list: np.ndarray         # (E,) int32, global ids j > i, ascending per atom
    n_waters: int
    cutoff: float
    skin: float = 2.0
    fft: Optional[tuple] = None   # explicit PME grid, None -> auto sizing
    tolerance: float = 1e-5
    dt_ps: float = 0.002
    seed: int = 0
    water_start: int = 0          # first atom of the waters (3 atoms each, O first)

    @property
    def n_atoms(self) -> int:
        return int(self.crd.shape[0])

    @property
    def n_types(self) -> int:
        t = int(round((math.sqrt(8 * len(self.lj_A) + 1) - 1) / 2))
        return t

    def residue_starts(self) -> np.ndarray:
        """First atom of every residue: each water (3 atoms) and each ion
        (1 atom) is one residue, in atom order."""
        N, w0, nw = self.n_atoms, self.water_start, self.n_waters
        starts = [i for i in range(w0)]
        starts += list(range(w0, w0 + 3 * nw, 3))
        starts += list(range(w0 + 3 * nw, N))
        return np.asarray(starts, np.int32)

    def cell(self) -> np.ndarray:
        """LTMatrix3 {a11,a21,a22,a31,a32,a33} of the box (float32)."""
        return np.array([self.box[0], 0, self.box[1], 0, 0, self.box[2]], np.float32)

    def rcell(self) -> np.ndarray:
        """Inverse cell computed in float32 as the reference's
        ``Invert_Lower_Triangular_Cell`` (``SPONGE/utils/boundary.h:59-70``)."""
        one = np.float32(1.0)
        return np.array([one / self.box[0], 0, one / self.box[1], 0, 0, one / self.box[2]],
                        np.float32)


def _lj_tables(params: List[tuple]):
    t = len(params)
    npair = t * (t + 1) // 2
    A = np.zeros(npair, np.float64)
    B = np.zeros(npair, np.float64)
    for a in range(t):
        for b in range(a + 1):
            sa, ea = params[a]
            sb, eb = params[b]
            eps = math.sqrt(ea * eb)
            sig = 0.5 * (sa + sb)
            if eps == 0.0:
                continue
            # file values (4 eps sigma^n) are rounded to float32 first, as the
            # loader reads them with fscanf("%f"), then scaled in float32
            a_file = np.float32(4.0 * eps * sig ** 12)
            b_file = np.float32(4.0 * eps * sig ** 6)
            A[lj_pair_index(a, b)] = a_file * np.float32(12.0)
            B[lj_pair_index(a, b)] = b_file * np.float32(6.0)
    return A.astype(np.float32), B.astype(np.float32)


def _random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotation matrices from unit quaternions, (n,3,3)."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _water_body() -> np.ndarray:
    """O, H1, H2 positions in the molecule frame (O at the origin)."""
    half = math.radians(HOH_ANGLE) / 2
    return np.array([[0.0, 0.0, 0.0],
                     [OH_BOND * math.sin(half), 0.0, OH_BOND * math.cos(half)],
                     [-OH_BOND * math.sin(half), 0.0, OH_BOND * math.cos(half)]])


def _min_image(d: np.ndarray, box: np.ndarray) -> np.ndarray:
    return d - np.floor(d / box + 0.5) * box


def _place_waters(rng, sites, box, min_dist=1.6, max_tries=200):
    """Seeded random orientations, redrawn for a molecule while any of its atoms
    sits closer than ``min_dist`` to an atom of an already placed molecule."""
    body = _water_body()
    n = len(sites)
    pos = np.zeros((n, 3, 3))
    rot = _random_rotations(rng, n)
    pos[:] = sites[:, None, :] + np.einsum("nij,kj->nki", rot, body)
    # neighbours on the lattice are the only ones that can clash; check with a
    # coarse cell grid over oxygen sites
    cell_len = 4.0
    ncell = np.maximum((box // cell_len).astype(int), 1)
    cidx = (np.floor((sites % box) / box * ncell).astype(int)) % ncell
    flat = (cidx[:, 0] * ncell[1] + cidx[:, 1]) * ncell[2] + cidx[:, 2]
    buckets = {}
    for m, c in enumerate(flat):
        buckets.setdefault(int(c), []).append(m)
    placed = np.zeros(n, bool)
    for m in range(n):
        cx, cy, cz = cidx[m]
        neigh = []
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    c = (((cx + dx) % ncell[0]) * ncell[1] + (cy + dy) % ncell[1]) * ncell[2] \
                        + (cz + dz) % ncell[2]
                    neigh.extend(buckets.get(int(c), []))
        neigh = [k for k in set(neigh) if placed[k] and k != m]
        for _ in range(max_tries):
            if not neigh:
                break
            other = pos[neigh].reshape(-1, 3)
            d = _min_image(pos[m][:, None, :] - other[None, :, :], box)
            if np.sqrt((d * d).sum(-1)).min() >= min_dist:
                break
            r = _random_rotations(rng, 1)[0]
            pos[m] = sites[m] + body @ r.T
        placed[m] = True
    return pos


def _rigid_velocities(rng, pos, masses, temperature):
    """Maxwell-Boltzmann velocities projected onto rigid-body motion of each
    molecule (COM translation + rotation), so they satisfy the constraints."""
    n = pos.shape[0]
    m = masses[None, :, None]
    sigma = np.sqrt(KB * temperature / masses)
    v = rng.normal(size=pos.shape) * sigma[None, :, None]
    M = masses.sum()
    vcom = (v * m).sum(1) / M
    com = (pos * m).sum(1) / M
    r = pos - com[:, None, :]
    L = (np.cross(r, v - vcom[:, None, :]) * m).sum(1)
    inertia = np.zeros((n, 3, 3))
    for k in range(pos.shape[1]):
        rk = r[:, k, :]
        rr = (rk * rk).sum(-1)
        inertia += masses[k] * (rr[:, None, None] * np.eye(3)[None] - rk[:, :, None] * rk[:, None, :])
    omega = np.linalg.solve(inertia, L[..., None])[..., 0]
    return vcom[:, None, :] + np.cross(omega[:, None, :], r)


def build_water_ion_box(name, n_sites_grid, box_len, n_sites=None, n_na=0, n_cl=0,
                        cutoff=10.0, skin=2.0, fft=None, seed=20260217, temperature=300.0):
    """Water (+ion) box: lattice sites in a cubic box, seeded random water
    orientations, ions on seeded random sites, Maxwell-Boltzmann velocities."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = n_sites_grid
    box = np.array([box_len] * 3, np.float64)
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    g = g.reshape(-1, 3).astype(np.float64)
    sites = (g + 0.5) * (box / np.array([nx, ny, nz]))
    if n_sites is not None and n_sites < len(sites):
        keep = np.sort(rng.permutation(len(sites))[:n_sites])
        sites = sites[keep]
    n_ion = n_na + n_cl
    ion_idx = np.sort(rng.permutation(len(sites))[:n_ion]) if n_ion else np.zeros(0, int)
    is_ion = np.zeros(len(sites), bool)
    is_ion[ion_idx] = True
    ion_sites = sites[is_ion]
    rng.shuffle(ion_sites)
    water_sites = sites[~is_ion]
    nw = len(water_sites)

    wpos = _place_waters(rng, water_sites, box)
    wmass = np.array([M_O, M_H, M_H])
    wvel = _rigid_velocities(rng, wpos, wmass, temperature)

    na_pos, cl_pos = ion_sites[:n_na], ion_sites[n_na:]
    N = 3 * nw + n_ion
    crd = np.concatenate([wpos.reshape(-1, 3), na_pos, cl_pos])
    vel = np.concatenate([
        wvel.reshape(-1, 3),
        rng.normal(size=(n_na, 3)) * math.sqrt(KB * temperature / M_NA),
        rng.normal(size=(n_cl, 3)) * math.sqrt(KB * temperature / M_CL)])
    mass = np.concatenate([np.tile(wmass, nw), np.full(n_na, M_NA), np.full(n_cl, M_CL)])
    # remove the total COM momentum
    vel -= (vel * mass[:, None]).sum(0) / mass.sum()

    qO = np.float32(Q_O * CHARGE_SCALE)
    qH = np.float32(Q_H * CHARGE_SCALE)
    charge = np.concatenate([np.tile(np.array([qO, qH, qH], np.float32), nw),
                             np.full(n_na, np.float32(CHARGE_SCALE), np.float32),
                             np.full(n_cl, np.float32(-CHARGE_SCALE), np.float32)])
    if n_ion:
        params = [LJ_O, LJ_H, LJ_NA, LJ_CL]
    else:
        params = [LJ_O, LJ_H]
    A, B = _lj_tables(params)
    lj_type = np.concatenate([np.tile(np.array([0, 1, 1], np.int32), nw),
                              np.full(n_na, 2, np.int32), np.full(n_cl, 3, np.int32)])
    # exclusions: O -> {H1, H2}, H1 -> {H2}, H2 -> {}
    excl_num = np.concatenate([np.tile(np.array([2, 1, 0], np.int32), nw),
                               np.zeros(n_ion, np.int32)])
    excl_start = np.concatenate([[0], np.cumsum(excl_num)[:-1]]).astype(np.int32)
    base = 3 * np.arange(nw, dtype=np.int32)
    excl_list = np.stack([base + 1, base + 2, base + 2], 1).reshape(-1).astype(np.int32)
    # positions stay unwrapped as generated (inside the box)
    return System(name=name, box=box.astype(np.float32), crd=crd.astype(np.float32),
                  vel=vel.astype(np.float32), charge=charge, mass=mass.astype(np.float32),
                  lj_type=lj_type, lj_A=A, lj_B=B, excl_start=excl_start,
                  excl_num=excl_num, excl_list=excl_list, n_waters=nw,
                  cutoff=float(cutoff), skin=float(skin), fft=fft, seed=seed)


def solvent_tail(s: System) -> System:
    """The same system with the ions first and the waters after them: the
    layout the reference's solvent_LJ expects (solute, then the water tail,
    SPONGE/Lennard_Jones_force/solvent_LJ.cpp:195-217)."""
    nw, N = s.n_waters, s.n_atoms
    order = np.concatenate([np.arange(3 * nw, N), np.arange(3 * nw)])  # new -> old
    new_of = np.empty(N, np.int64)
    new_of[order] = np.arange(N)
    excl_num = s.excl_num[order].astype(np.int32)
    lists = []
    for old in order:
        a, n = s.excl_start[old], s.excl_num[old]
        lists.append(np.sort(new_of[s.excl_list[a:a + n]]).astype(np.int32))
    excl_list = np.concatenate(lists) if lists else np.zeros(0, np.int32)
    excl_start = np.concatenate([[0], np.cumsum(excl_num)[:-1]]).astype(np.int32)
    return dataclasses.replace(s, name=s.name + "_tail", crd=s.crd[order].copy(),
                               vel=s.vel[order].copy(), charge=s.charge[order].copy(),
                               mass=s.mass[order].copy(), lj_type=s.lj_type[order].copy(),
                               excl_start=excl_start, excl_num=excl_num,
                               excl_list=excl_list.astype(np.int32),
                               water_start=N - 3 * nw)


def config(idx: int, seed: int = None) -> System:
    """The benchmark configurations (box, cutoff and grid as specified):

    1: 648 waters, 9x9x8 lattice, L=26.9 A, rc 8 A, PME 27^3
    2: 8000 waters, 20^3 lattice, L=62.1 A, rc 10 A, PME 64^3
    3: 31000 waters + 1000 Na+ + 1000 Cl- (95,000 atoms), L=99 A, rc 10 A, PME 96^3
    """
    s = 20260217 + idx if seed is None else seed
    if idx == 1:
        return build_water_ion_box("water648", (9, 9, 8), 26.9, cutoff=8.0, fft=(27, 27, 27), seed=s)
    if idx == 2:
        return build_water_ion_box("water8000", (20, 20, 20), 62.1, cutoff=10.0, fft=(64, 64, 64),
                                   seed=s)
    if idx == 3:
        return build_water_ion_box("water_ions_95k", (33, 33, 33), 99.0, n_sites=33000,
                                   n_na=1000, n_cl=1000, cutoff=10.0, fft=(96, 96, 96), seed=s)
    raise ValueError(f"unknown config {idx}")


def small_water_box(n_side=4, spacing=3.1, cutoff=None, seed=7, n_na=0, n_cl=0, fft=None):
    """Tiny test systems (n_side^3 lattice sites)."""
    L = n_side * spacing
    rc = cutoff if cutoff is not None else min(8.0, 0.5 * L / 1.5 * 2 - 2.0)
    return build_water_ion_box(f"water{n_side}", (n_side,) * 3, L, n_na=n_na, n_cl=n_cl,
                               cutoff=rc, fft=fft, seed=seed)


WATER_NUMBER_DENSITY = 1.0 / 18.015 * 0.602214076  # molecules per A^3 at 1.0 g/cm^3


def water_box(n_waters, cutoff=9.0, skin=2.0, seed=7, name=None):
    """Cubic box of n_waters rigid waters at 1.0 g/cm^3 (lattice sites,
    seeded orientations, 300 K velocities); PME grid from the automatic
    ~1 A sizing."""
    L = (n_waters / WATER_NUMBER_DENSITY) ** (1.0 / 3.0)
    m = int(math.ceil(n_waters ** (1.0 / 3.0)))
    return build_water_ion_box(name or f"water{n_waters}", (m, m, m), L, n_sites=n_waters,
                               cutoff=cutoff, skin=skin, fft=None, seed=seed)


def batched_boxes(n_boxes=48, lo=500, hi=6000, seed=20260304, cutoff=9.0):
    """The batched workload: n_boxes water boxes whose sizes are drawn
    log-uniformly (seeded) between lo and hi waters, each at 1.0 g/cm^3 with
    its own PME grid at ~1 A and the given cutoff."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(lo), math.log(hi), n_boxes)).astype(int)
    return [water_box(int(n), cutoff=cutoff, seed=seed + 1 + k) for k, n in enumerate(sizes)]


M_SOLUTE = 12.011  # amu per solute atom (free LJ / charge sites)


@dataclasses.dataclass
class SoftCoreState:
    """The A / B states of a lambda-window set (LJ_SOFT_CORE's inputs,
    SPONGE/Lennard_Jones_force/LJ_soft_core.cpp:463-560): per-atom types of
    each state, both tables in the in-memory convention, charge_B - charge_A,
    and the soft-core options."""
    typeA: np.ndarray
    typeB: np.ndarray
    AA: np.ndarray
    AB: np.ndarray
    BA: np.ndarray
    BB: np.ndarray
    charge_A: np.ndarray
    charge_BA: np.ndarray
    alpha: float = 0.5
    p: float = 1.0
    sigma: float = 3.0

    def charges(self, lam: float) -> np.ndarray:
        """Charges at charge coupling lam: q_A + lam (q_B - q_A)."""
        return (self.charge_A.astype(np.float64) + lam * self.charge_BA).astype(np.float32)


def lambda_windows(n_windows=16, n_waters=2000, n_solute=24, cutoff=9.0, skin=2.0,
                   seed=20260305, temperature=300.0):
    """The lambda-window workload: one n_solute-atom LJ / charge solute
    (seeded: |q| <= 0.5 e with zero net charge, sigma 2.5-4.0 A, epsilon
    0.05-0.30 kcal/mol, Lorentz-Berthelot with the water oxygen) at the centre
    of a box of n_waters rigid waters at 1.0 g/cm^3, decoupled over n_windows
    windows at evenly spaced lambda in [0, 1]: the B state turns every solute
    atom into a dummy (no LJ, no charge); soft core alpha 0.5, p 1, sigma 3 A.
    The decoupling follows the usual two-stage path: over lambda in [0, 1/2]
    the charges are switched off (charges q_A (1 - 2 lambda)) with the LJ
    coupled (lambda_lj = 0), over [1/2, 1] the LJ (lambda_lj = 2 lambda - 1)
    with the charges off.  Switching both at once lets a water hydrogen (no
    LJ) reach a charged solute atom whose repulsion is already softened,
    where the reference's pair term for pairs without LJ in either state,
    plain Coulomb (LJ_soft_core.cpp:252), diverges.  Returns (systems,
    lambda_lj, state): systems[k] carries the charges of window k and
    lambda_lj[k] is its soft-core module's lambda_lj."""
    rng = np.random.default_rng(seed)
    n_sites = n_waters + n_solute
    L = (n_sites / WATER_NUMBER_DENSITY) ** (1.0 / 3.0)
    m = int(math.ceil(n_sites ** (1.0 / 3.0)))
    box = np.array([L] * 3, np.float64)
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1)
    sites = (g.reshape(-1, 3).astype(np.float64) + 0.5) * (L / m)
    sites = sites[np.sort(rng.permutation(len(sites))[:n_sites])]
    # the solute: the sites nearest the box centre
    d = np.linalg.norm(sites - 0.5 * box, axis=1)
    order = np.argsort(d, kind="stable")
    solute_sites = sites[np.sort(order[:n_solute])]
    water_sites = sites[np.sort(order[n_solute:])]
    wpos = _place_waters(rng, water_sites, box)
    wmass = np.array([M_O, M_H, M_H])
    wvel = _rigid_velocities(rng, wpos, wmass, temperature)
    nw = len(water_sites)
    N = 3 * nw + n_solute
    crd = np.concatenate([wpos.reshape(-1, 3), solute_sites])
    vel = np.concatenate([wvel.reshape(-1, 3),
                          rng.normal(size=(n_solute, 3)) * math.sqrt(KB * temperature / M_SOLUTE)])
    mass = np.concatenate([np.tile(wmass, nw), np.full(n_solute, M_SOLUTE)])
    vel -= (vel * mass[:, None]).sum(0) / mass.sum()
    q = rng.uniform(-0.5, 0.5, n_solute)
    q -= q.mean()
    q *= min(1.0, 0.5 / np.abs(q).max())
    qO = np.float32(Q_O * CHARGE_SCALE)
    qH = np.float32(Q_H * CHARGE_SCALE)
    charge_A = np.concatenate([np.tile(np.array([qO, qH, qH], np.float32), nw),
                               (q * CHARGE_SCALE).astype(np.float32)])
    solute_lj = [(float(rng.uniform(2.5, 4.0)), float(rng.uniform(0.05, 0.30)))
                 for _ in range(n_solute)]
    params = [LJ_O, LJ_H] + solute_lj
    AA, AB = _lj_tables(params)
    BA, BB = _lj_tables(params + [(0.0, 0.0)])  # + the dummy type
    typeA = np.concatenate([np.tile(np.array([0, 1, 1], np.int32), nw),
                            np.arange(2, 2 + n_solute, dtype=np.int32)])
    typeB = typeA.copy()
    typeB[3 * nw:] = 2 + n_solute
    charge_BA = np.zeros(N, np.float32)
    charge_BA[3 * nw:] = -charge_A[3 * nw:]
    excl_num = np.concatenate([np.tile(np.array([2, 1, 0], np.int32), nw),
                               np.zeros(n_solute, np.int32)])
    excl_start = np.concatenate([[0], np.cumsum(excl_num)[:-1]]).astype(np.int32)
    base = 3 * np.arange(nw, dtype=np.int32)
    excl_list = np.stack([base + 1, base + 2, base + 2], 1).reshape(-1).astype(np.int32)
    state = SoftCoreState(typeA=typeA, typeB=typeB, AA=AA, AB=AB, BA=BA, BB=BB,
                          charge_A=charge_A, charge_BA=charge_BA)
    lambdas = np.linspace(0.0, 1.0, n_windows)
    lam_q = np.minimum(1.0, 2.0 * lambdas)
    lam_lj = np.maximum(0.0, 2.0 * lambdas - 1.0)
    systems = [System(name=f"lambda{k}", box=box.astype(np.float32), crd=crd.astype(np.float32),
                      vel=vel.astype(np.float32), charge=state.charges(float(lam_q[k])),
                      mass=mass.astype(np.float32), lj_type=typeA, lj_A=AA, lj_B=AB,
                      excl_start=excl_start, excl_num=excl_num, excl_list=excl_list,
                      n_waters=nw, cutoff=float(cutoff), skin=float(skin), fft=None, seed=seed)
               for k in range(n_windows)]
    return systems, lam_lj, state
