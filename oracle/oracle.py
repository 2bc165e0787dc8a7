"""ORACLE -- TEST INFRASTRUCTURE ONLY.

ctypes front-end of oracle/_build/liboracle.so (the checker) and
liboracle_fast.so (the timed CPU baseline).  Only tests/, the smoke check and
the CPU-baseline leg of bench.py use it; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_BUILD = os.path.join(HERE, "_build")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)


def _load(name):
    path = os.path.join(_BUILD, name)
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    lib.orc_get_beta.restype = C.c_float
    lib.orc_get_beta.argtypes = [C.c_float, C.c_float]
    lib.orc_get_fft_parameter.restype = C.c_int
    lib.orc_get_fft_parameter.argtypes = [C.c_float]
    lib.orc_check_2357.argtypes = [C.c_int]
    lib.orc_displacement.argtypes = [_f32p] * 5
    lib.orc_nl_grid_dims.argtypes = [_f32p, C.c_float, C.c_float, _i32p]
    lib.orc_nl_cell_neighbors.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _f32p, C.c_float,
                                          C.c_int, _i32p]
    lib.orc_nl_cell_of.argtypes = [_f32p, _f32p, C.c_float, C.c_float, _i32p]
    lib.orc_nl_build.argtypes = [C.c_int, _f32p, _f32p, _f32p, C.c_float, C.c_float, C.c_int,
                                 C.c_int, C.c_void_p, _i32p, _i32p, _i32p, _i32p, _i32p]
    lib.orc_nl_bruteforce.restype = C.c_longlong
    lib.orc_nl_bruteforce.argtypes = [C.c_int, _f32p, _f32p, _f32p, C.c_float, C.c_float,
                                      C.c_int, C.c_void_p, _i32p, _i32p, _i32p, _i32p, _i32p]
    lib.orc_lj_direct.argtypes = [C.c_int, C.c_int, _i32p, _i32p, C.c_int, _f32p, _f32p, _i32p,
                                  _f32p, _f32p, _f32p, _f32p, C.c_float, C.c_float, C.c_int,
                                  C.c_int, _f32p, _f32p, _f32p, _f32p, _f32p]
    lib.orc_ljsc_direct.argtypes = [C.c_int, _i32p, _i32p, C.c_int, _f32p, _f32p, _f32p, _i32p,
                                    _i32p, _f32p, _f32p, _f32p, _f32p, _f32p, _f32p, C.c_float,
                                    C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                    C.c_float, C.c_int, C.c_int, C.c_int, _f32p, _f32p, _f32p,
                                    _f32p, _f32p, _f64p]
    lib.orc_pair_counts.argtypes = [C.c_int, _i32p, _i32p, C.c_int, _f32p, _f32p, _f32p, C.c_float,
                                    C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    lib.orc_lj_long_range_factor.restype = C.c_float
    lib.orc_lj_long_range_factor.argtypes = [C.c_int, C.c_int, _i32p, _f32p, C.c_float]
    lib.orc_pme_excluded.argtypes = [C.c_int, _f32p, _f32p, _f32p, _f32p, C.c_float, _i32p, _i32p,
                                     _i32p, C.c_int, C.c_int, _f32p, _f32p, _f32p, _f32p]
    lib.orc_pme_create.restype = C.c_void_p
    lib.orc_pme_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, _f32p, _f32p]
    lib.orc_pme_destroy.argtypes = [C.c_void_p]
    lib.orc_pme_tables.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.orc_pme_reciprocal.argtypes = [C.c_void_p, _f32p, _f32p, _f32p, C.c_int, C.c_int, _f32p,
                                       _f32p, C.c_int, C.c_int, _f32p]
    lib.orc_pme_grids.argtypes = [C.c_void_p] + [C.c_void_p] * 5
    lib.orc_fft3d_r2c.argtypes = [C.c_int, C.c_int, C.c_int, _f64p, _f64p]
    lib.orc_fft1d.argtypes = [C.c_int, _f64p, C.c_int]
    lib.orc64_direct.argtypes = [C.c_int, _f64p, _f64p, _i32p, _f64p, _f64p, _f64p, C.c_double,
                                 C.c_double, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p]
    lib.orc64_excluded.argtypes = [C.c_int, _f64p, _f64p, _f64p, C.c_double, _i32p, _i32p, _i32p,
                                   _f64p, _f64p, _f64p]
    lib.orc64_ewald_recip.argtypes = [C.c_int, _f64p, _f64p, _f64p, C.c_double, _i32p, _f64p,
                                      _f64p]
    lib.orc64_pme.argtypes = [C.c_int, _f64p, _f64p, _f64p, C.c_double, C.c_int, C.c_int, C.c_int,
                              _f64p, C.POINTER(C.c_double), _f64p]
    lib.orc_sum_f32.restype = C.c_double
    lib.orc_sum_f32.argtypes = [_f32p, C.c_int]
    lib.orc_num_threads.restype = C.c_int
    lib.orc_set_num_threads.argtypes = [C.c_int]
    lib.orc_recip_seconds.argtypes = [_f64p]
    lib.orc64_direct_list.argtypes = [C.c_int, _f64p, _f64p, _i32p, _f64p, _f64p, _f64p,
                                      C.c_double, C.c_double, _i32p, _i32p, C.c_int, _f64p,
                                      _f64p, _f64p]
    lib.orc_solvent_lj.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, C.c_int, _f32p, _f32p,
                                   _i32p, _f32p, _f32p, _f32p, _f32p, C.c_float, C.c_float,
                                   C.c_int, C.c_int, _f32p, _f32p, _f32p, _f32p, _f32p]
    lib.orc_settle_geometry.argtypes = [C.c_float] * 6 + [_f32p]
    lib.orc_nve_settle.argtypes = [C.c_int, _f32p, _f32p, _f32p, _f32p, _f32p, C.c_int, _i32p,
                                   _f32p, _f32p, _f32p, C.c_float]
    lib.orc64_nve_shake.restype = C.c_int
    lib.orc64_nve_shake.argtypes = [C.c_int, _f64p, _f64p, _f64p, _f64p, C.c_int, _i32p,
                                    C.c_double, C.c_double, C.c_double, _f64p, C.c_double,
                                    C.c_int, C.c_double]
    return lib


_libs = {}


def lib(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _libs:
        _libs[name] = _load(name)
    return _libs[name]


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------------------
def get_beta(cutoff, tol=1e-5):
    return lib().orc_get_beta(cutoff, tol)


def get_fft_parameter(length):
    return lib().orc_get_fft_parameter(length)


def nl_grid_dims(cell, cutoff, skin):
    d = np.zeros(3, np.int32)
    lib().orc_nl_grid_dims(f32(cell), cutoff, skin, d)
    return d


def nl_build(sys_, crd=None, max_nb=1200, max_in_grid=150, fast=False):
    """Reference-algorithm half list: (counts[N], serial[N, max_nb], flags)."""
    crd = sys_.crd if crd is None else crd
    N = len(crd)
    counts = np.zeros(N, np.int32)
    serial = np.zeros(N * max_nb, np.int32)
    flags = lib(fast).orc_nl_build(N, f32(crd), sys_.cell(), sys_.rcell(), sys_.cutoff,
                                   sys_.skin, max_nb, max_in_grid, None, i32(sys_.excl_start),
                                   i32(sys_.excl_list) if len(sys_.excl_list) else np.zeros(1, np.int32),
                                   i32(sys_.excl_num), counts, serial)
    return counts, serial.reshape(N, max_nb), flags


def nl_bruteforce(sys_, crd=None, max_nb=1200):
    crd = sys_.crd if crd is None else crd
    N = len(crd)
    counts = np.zeros(N, np.int32)
    serial = np.zeros(N * max_nb, np.int32)
    total = lib().orc_nl_bruteforce(N, f32(crd), sys_.cell(), sys_.rcell(), sys_.cutoff, sys_.skin,
                                    max_nb, None, i32(sys_.excl_start),
                                    i32(sys_.excl_list) if len(sys_.excl_list) else np.zeros(1, np.int32),
                                    i32(sys_.excl_num), counts, serial)
    return counts, serial.reshape(N, max_nb), total


def pair_sets(counts, serial):
    """Set of (i, j) pairs of a half list."""
    out = set()
    for i in range(len(counts)):
        for j in serial[i, :counts[i]]:
            out.add((i, int(j)))
    return out


def pair_array(counts, serial):
    """Sorted (P,2) int64 array of the half list's pairs (vectorised)."""
    N, M = serial.shape
    mask = np.arange(M)[None, :] < counts[:, None]
    ii = np.broadcast_to(np.arange(N)[:, None], (N, M))[mask]
    jj = serial[mask]
    key = ii.astype(np.int64) * N + jj.astype(np.int64)
    key.sort()
    return key


def lj_direct(sys_, counts, serial, beta, crd=None, need_energy=True, need_virial=True,
              fast=False, solvent_n=0):
    crd = sys_.crd if crd is None else crd
    N = len(crd)
    max_nb = serial.shape[1]
    frc = np.zeros(3 * N, np.float32)
    e = np.zeros(N, np.float32)
    v = np.zeros(6 * N, np.float32)
    edir = np.zeros(N, np.float32)
    elj = np.zeros(N, np.float32)
    lib(fast).orc_lj_direct(N, int(solvent_n), i32(counts), i32(serial.reshape(-1)), max_nb, f32(crd),
                            sys_.charge, i32(sys_.lj_type), sys_.lj_A, sys_.lj_B, sys_.cell(),
                            sys_.rcell(), sys_.cutoff, beta, int(need_energy), int(need_virial),
                            frc, e, v, edir, elj)
    return dict(frc=frc.reshape(N, 3), atom_energy=e, atom_virial=v.reshape(N, 6),
                direct_energy=edir, lj_energy=elj)


def solvent_lj(sys_, counts, serial, beta, solvent_atom, water_points=3, crd=None,
               need_energy=True, need_virial=True):
    """The solvent_LJ pairs of the water tail (from atom solvent_atom on)
    over the reference-format half list; same output layout as lj_direct."""
    crd = sys_.crd if crd is None else crd
    N = len(crd)
    max_nb = serial.shape[1]
    frc = np.zeros(3 * N, np.float32)
    e = np.zeros(N, np.float32)
    v = np.zeros(6 * N, np.float32)
    edir = np.zeros(N, np.float32)
    elj = np.zeros(N, np.float32)
    lib().orc_solvent_lj(N, int(solvent_atom), int(water_points), i32(counts),
                         i32(serial.reshape(-1)), max_nb, f32(crd), sys_.charge,
                         i32(sys_.lj_type), sys_.lj_A, sys_.lj_B, sys_.cell(), sys_.rcell(),
                         sys_.cutoff, beta, int(need_energy), int(need_virial), frc, e, v, edir,
                         elj)
    return dict(frc=frc.reshape(N, 3), atom_energy=e, atom_virial=v.reshape(N, 6),
                direct_energy=edir, lj_energy=elj)


def ljsc_direct(sys_, counts, serial, beta, sc, crd=None, need_energy=True, need_virial=True,
                need_du=True):
    """Soft-core LJ + direct Coulomb on a half list.  `sc` holds typeA, typeB,
    AA, AB, BA, BB, charge_BA, lambda, alpha, p, sigma6, sigma6_min."""
    crd = sys_.crd if crd is None else crd
    N = len(crd)
    frc = np.zeros(3 * N, np.float32)
    e = np.zeros(N, np.float32)
    v = np.zeros(6 * N, np.float32)
    edir = np.zeros(N, np.float32)
    elj = np.zeros(N, np.float32)
    du = np.zeros(2, np.float64)
    qba = sc.get("charge_BA")
    lib().orc_ljsc_direct(N, i32(counts), i32(serial.reshape(-1)), serial.shape[1], f32(crd),
                          f32(sys_.charge), f32(qba if qba is not None else np.zeros(N)),
                          i32(sc["typeA"]), i32(sc["typeB"]), f32(sc["AA"]), f32(sc["AB"]),
                          f32(sc["BA"]), f32(sc["BB"]), sys_.cell(), sys_.rcell(), sys_.cutoff,
                          beta, sc["lambda"], sc["alpha"], sc["p"], sc["sigma6"],
                          sc["sigma6_min"], int(need_energy), int(need_virial), int(need_du), frc,
                          e, v, edir, elj, du)
    return dict(frc=frc.reshape(N, 3), atom_energy=e, atom_virial=v.reshape(N, 6),
                direct_energy=edir, lj_energy=elj, du_lj=du[0], du_direct=du[1])


def pair_counts(sys_, counts, serial, crd=None):
    crd = sys_.crd if crd is None else crd
    a = C.c_longlong()
    b = C.c_longlong()
    lib().orc_pair_counts(len(crd), i32(counts), i32(serial.reshape(-1)), serial.shape[1],
                          f32(crd), sys_.cell(), sys_.rcell(), sys_.cutoff, C.byref(a), C.byref(b))
    return a.value, b.value


def lj_long_range_factor(sys_):
    return lib().orc_lj_long_range_factor(sys_.n_atoms, sys_.n_types, i32(sys_.lj_type),
                                          sys_.lj_B, sys_.cutoff)


def pme_excluded(sys_, beta, crd=None, need_energy=True, need_virial=True):
    crd = sys_.crd if crd is None else crd
    N = len(crd)
    frc = np.zeros(3 * N, np.float32)
    e = np.zeros(N, np.float32)
    this = np.zeros(N, np.float32)
    v = np.zeros(6 * N, np.float32)
    lib().orc_pme_excluded(N, f32(crd), sys_.charge, sys_.cell(), sys_.rcell(), beta,
                           i32(sys_.excl_start), i32(sys_.excl_list), i32(sys_.excl_num),
                           int(need_energy), int(need_virial), frc, e, this, v)
    return dict(frc=frc.reshape(N, 3), atom_energy=e, correction_energy=this,
                atom_virial=v.reshape(N, 6))


class PME:
    """Reference-algorithm PME reciprocal part (one instance per system)."""

    def __init__(self, sys_, fft, beta, fast=False):
        self.lib = lib(fast)
        self.N = sys_.n_atoms
        self.fft = tuple(int(x) for x in fft)
        self.beta = beta
        self.h = self.lib.orc_pme_create(self.N, *self.fft, beta, sys_.cell(), sys_.rcell())
        self.sys = sys_

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_pme_destroy(self.h)
            self.h = None

    def tables(self):
        fx, fy, fz = self.fft
        nfft = fx * fy * (fz // 2 + 1)
        bc = np.zeros(nfft, np.float32)
        vbc = np.zeros(6 * nfft, np.float32)
        self.lib.orc_pme_tables(self.h, bc.ctypes.data, vbc.ctypes.data)
        return bc, vbc.reshape(nfft, 6)

    def reciprocal(self, crd=None, need_energy=True, need_virial=True, step=0, update_interval=1):
        crd = self.sys.crd if crd is None else crd
        frc = np.zeros(3 * self.N, np.float32)
        vir = np.zeros(6, np.float32)
        pot = np.zeros(1, np.float32)
        en = np.zeros(3, np.float32)
        self.lib.orc_pme_reciprocal(self.h, f32(crd), self.sys.charge, frc, int(need_virial),
                                    int(need_energy), vir, pot, step, update_interval, en)
        return dict(frc=frc.reshape(self.N, 3), virial=vir, potential=float(pot[0]),
                    reciprocal=float(en[0]), self_energy=float(en[1]))

    def last_seconds(self):
        """Wall time of the last evaluation's pieces: spread, FFT
        convolution, gather."""
        t = np.zeros(3)
        self.lib.orc_recip_seconds(t)
        return dict(spread=t[0], fft_convolution=t[1], gather=t[2])

    def grids(self):
        fx, fy, fz = self.fft
        nall = fx * fy * fz
        nfft = fx * fy * (fz // 2 + 1)
        Q = np.zeros(nall, np.float32)
        F = np.zeros(nall, np.float32)
        FQ = np.zeros(2 * nfft, np.float32)
        near = np.zeros(64 * self.N, np.int32)
        fr = np.zeros(3 * self.N, np.float32)
        self.lib.orc_pme_grids(self.h, Q.ctypes.data, F.ctypes.data, FQ.ctypes.data,
                               near.ctypes.data, fr.ctypes.data)
        return dict(Q=Q, FBCFQ=F, FQ=FQ.reshape(nfft, 2), near=near.reshape(self.N, 64),
                    frxyz=fr.reshape(self.N, 3))


def fft3d_r2c(a):
    a = f64(a)
    nx, ny, nz = a.shape
    out = np.zeros(nx * ny * (nz // 2 + 1) * 2, np.float64)
    lib().orc_fft3d_r2c(nx, ny, nz, a.reshape(-1), out)
    o = out.reshape(nx, ny, nz // 2 + 1, 2)
    return o[..., 0] + 1j * o[..., 1]


def fft1d(x, sign=-1):
    x = np.asarray(x, np.complex128)
    buf = f64(np.stack([x.real, x.imag], -1).reshape(-1))
    lib().orc_fft1d(len(x), buf, sign)
    b = buf.reshape(-1, 2)
    return b[:, 0] + 1j * b[:, 1]


# FP64 oracle --------------------------------------------------------------
def direct64(sys_, beta, crd=None, cutoff=None):
    crd = sys_.crd if crd is None else crd
    N = len(crd)
    frc = np.zeros(3 * N)
    en = np.zeros(2)
    vir = np.zeros(6)
    lib().orc64_direct(N, f64(crd).reshape(-1), f64(sys_.charge), i32(sys_.lj_type),
                       f64(sys_.lj_A), f64(sys_.lj_B), f64(sys_.box),
                       float(sys_.cutoff if cutoff is None else cutoff), float(beta),
                       i32(sys_.excl_start), i32(sys_.excl_list) if len(sys_.excl_list) else np.zeros(1, np.int32),
                       i32(sys_.excl_num), frc, en, vir)
    return dict(frc=frc.reshape(N, 3), lj=en[0], coulomb=en[1], virial=vir)


def direct64_list(sys_, beta, counts, serial, crd=None):
    """FP64 direct sums over a half list holding every pair within the
    cutoff (orc_nl_build's), for systems too large for direct64."""
    crd = sys_.crd if crd is None else crd
    N = len(crd)
    frc = np.zeros(3 * N)
    en = np.zeros(2)
    vir = np.zeros(6)
    serial = np.ascontiguousarray(serial, np.int32).reshape(N, -1)
    lib().orc64_direct_list(N, f64(crd).reshape(-1), f64(sys_.charge), i32(sys_.lj_type),
                            f64(sys_.lj_A), f64(sys_.lj_B), f64(sys_.box), float(sys_.cutoff),
                            float(beta), i32(counts), serial.reshape(-1), serial.shape[1], frc,
                            en, vir)
    return dict(frc=frc.reshape(N, 3), lj=en[0], coulomb=en[1], virial=vir)


def excluded64(sys_, beta, crd=None):
    crd = sys_.crd if crd is None else crd
    N = len(crd)
    frc = np.zeros(3 * N)
    e = np.zeros(1)
    vir = np.zeros(6)
    lib().orc64_excluded(N, f64(crd).reshape(-1), f64(sys_.charge), f64(sys_.box), float(beta),
                         i32(sys_.excl_start), i32(sys_.excl_list), i32(sys_.excl_num), frc, e, vir)
    return dict(frc=frc.reshape(N, 3), energy=e[0], virial=vir)


def ewald_recip64(crd, charge, box, beta, kmax):
    crd = f64(crd)
    N = len(crd)
    frc = np.zeros(3 * N)
    en = np.zeros(3)
    lib().orc64_ewald_recip(N, crd.reshape(-1), f64(charge), f64(box), float(beta),
                            i32(np.broadcast_to(np.asarray(kmax), (3,))), frc, en)
    return dict(frc=frc.reshape(N, 3), reciprocal=en[0], self_energy=en[1], neutralizing=en[2])


def pme64(crd, charge, box, beta, fft):
    crd = f64(crd)
    N = len(crd)
    frc = np.zeros(3 * N)
    e = C.c_double()
    vir = np.zeros(6)
    lib().orc64_pme(N, crd.reshape(-1), f64(charge), f64(box), float(beta), *[int(x) for x in fft],
                    frc, C.byref(e), vir)
    return dict(frc=frc.reshape(N, 3), reciprocal=e.value, virial=vir)


TIME_CONVERSION = 20.455  # ps -> internal time units, SPONGE/common.h:138


def water_triples(sys_):
    """(A, B, C) = (O, H1, H2) atom ids of every water."""
    w0 = getattr(sys_, "water_start", 0)
    return np.arange(w0, w0 + 3 * sys_.n_waters, dtype=np.int32).reshape(-1, 3)


def water_sides(sys_, crd=None):
    """Side lengths |OH1|, |OH2|, |H1H2| of the first water, in double (the
    step driver's rigid geometry comes from the same coordinates)."""
    crd = np.asarray(sys_.crd if crd is None else crd, np.float64).reshape(-1, 3)
    w0 = getattr(sys_, "water_start", 0)
    a, b, c = crd[w0], crd[w0 + 1], crd[w0 + 2]
    return (float(np.linalg.norm(a - b)), float(np.linalg.norm(a - c)),
            float(np.linalg.norm(b - c)))


def settle_geometry(sides, masses):
    g = np.zeros(5, np.float32)
    lib().orc_settle_geometry(*[float(x) for x in sides], *[float(m) for m in masses], g)
    return g


def nve_settle(sys_, crd, vel, frc, dt_ps=None, sides=None):
    """One NVE leap-frog + SETTLE step in the reference's FP32 arithmetic:
    returns the new (crd, vel)."""
    N = sys_.n_atoms
    dt = np.float32((sys_.dt_ps if dt_ps is None else dt_ps) * np.float32(TIME_CONVERSION))
    m = f32(sys_.mass)
    im = np.where(m > 0, np.float32(1.0) / np.where(m > 0, m, 1), 0).astype(np.float32)
    w = water_triples(sys_)
    g = settle_geometry(sides or water_sides(sys_), m[w[0]] if len(w) else (16.0, 1.008, 1.008))
    c = f32(crd).reshape(-1).copy()
    v = f32(vel).reshape(-1).copy()
    lib().orc_nve_settle(N, c, v, f32(frc).reshape(-1), im, m, len(w), i32(w).reshape(-1), g,
                         sys_.cell(), sys_.rcell(), float(dt))
    return c.reshape(N, 3), v.reshape(N, 3)


def nve_shake64(sys_, crd, vel, frc, dt_ps=None, sides=None, max_iter=500, tol=1e-13):
    """FP64 leap-frog + SHAKE-iterated constraints of the same step: returns
    (crd, vel, iterations)."""
    N = sys_.n_atoms
    dt = (sys_.dt_ps if dt_ps is None else dt_ps) * TIME_CONVERSION
    m = np.asarray(sys_.mass, np.float64)
    im = np.where(m > 0, 1.0 / np.where(m > 0, m, 1), 0.0)
    w = water_triples(sys_)
    dab, dac, dbc = sides or water_sides(sys_)
    c = f64(crd).reshape(-1).copy()
    v = f64(vel).reshape(-1).copy()
    it = lib().orc64_nve_shake(N, c, v, f64(frc).reshape(-1), im, len(w), i32(w).reshape(-1),
                               dab, dac, dbc, f64(sys_.box), dt, max_iter, tol)
    return c.reshape(N, 3), v.reshape(N, 3), it
