"""ctypes binding of include/sponge_amd.h plus thin Python mirrors of the
reference's module classes (NEIGHBOR_LIST, LENNARD_JONES_INFORMATION,
Particle_Mesh) over device tensors.

Device memory and streams come from PyTorch; every computation runs in the
HIP library sponge_amd/libsponge_amd.so.  Importing fails loudly when the
library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsponge_amd.so")

SUCCESS = 0
ERROR_NAMES = {
    1001: "NotImplemented", 1002: "BadFileFormat", 1003: "ConflictingCommand",
    1004: "MissingCommand", 1005: "TypeErrorCommand", 1006: "ValueErrorCommand",
    1007: "SimulationBreakDown", 1008: "MallocFailed", 1009: "Overflow", 1010: "OpenFileFailed",
}
NL_CONDITIONAL, NL_FORCED = 0, 1


class SpongeError(RuntimeError):
    def __init__(self, code, where):
        name = ERROR_NAMES.get(code, "hipError" if code < 1000 else "unknown")
        super().__init__(f"{where} failed with {code} ({name})")
        self.code = code


def _check(rc, where):
    if rc != 0:
        raise SpongeError(rc, where)


class NLOptions(C.Structure):
    _fields_ = [("refresh_interval", C.c_int), ("skin_permit", C.c_float),
                ("max_neighbor_numbers", C.c_int), ("max_atom_in_grid_numbers", C.c_int),
                ("throw_error_when_overflow", C.c_int), ("check_overflow_interval", C.c_int)]


class PMEOptions(C.Structure):
    _fields_ = [("fftx", C.c_int), ("ffty", C.c_int), ("fftz", C.c_int),
                ("grid_spacing", C.c_float), ("direct_tolerance", C.c_float),
                ("update_interval", C.c_int), ("calculate_reciprocal_part", C.c_int),
                ("calculate_excluded_part", C.c_int)]


_P = C.c_void_p
_F = C.c_float
_I = C.c_int
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int)

# name -> (restype, argtypes)
SIGNATURES = {
    "sponge_amd_version": (C.c_char_p, []),
    "sponge_amd_get_beta": (_F, [_F, _F]),
    "sponge_amd_get_fft_parameter": (_I, [_F]),
    "sponge_amd_set_kernel_timing": (_I, [_I]),
    "sponge_amd_set_kernel_timing_filter": (_I, [C.c_char_p]),
    "sponge_amd_kernel_timing": (_I, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
    "sponge_amd_reset_kernel_timing": (_I, []),
    "sponge_nl_default_options": (None, [C.POINTER(NLOptions)]),
    "sponge_nl_initial": (_I, [C.POINTER(_P), _I, _F, _F, _FP, _FP, C.POINTER(NLOptions), _P]),
    "sponge_nl_update": (_I, [_P, _P, _I, _I, _P, _FP, _FP, _I, _I, _P, _P, _P]),
    "sponge_nl_check_overflow": (_I, [_P, C.POINTER(_I)]),
    "sponge_nl_get_options": (_I, [_P, C.POINTER(NLOptions)]),
    "sponge_nl_export_half_list": (_I, [_P, _P, _P, _I]),
    "sponge_nl_atom_group": (_I, [_P, C.POINTER(_P)]),
    "sponge_nl_atom_group_address": (_I, [_P, C.POINTER(_P)]),
    "sponge_nl_info": (_I, [_P, C.POINTER(C.c_longlong), C.POINTER(_I)]),
    "sponge_nl_rebuild_count": (_I, [_P, C.POINTER(C.c_longlong)]),
    "sponge_nl_clear": (_I, [_P]),
    "sponge_lj_initial": (_I, [C.POINTER(_P), _I, _I, _FP, _FP, C.POINTER(_I), _F, _P]),
    "sponge_lj_long_range_factor": (_I, [_P, _FP]),
    "sponge_lj_pme_direct_force": (_I, [_P, _P, _I, _I, _I, _I, _P, _P, _P, _FP, _FP, _F, _I, _P,
                                        _I, _P, _P]),
    "sponge_lj_pme_direct_force_list": (_I, [_P, _P, _P, _I, _I, _P, _P, _P, _FP, _FP, _F, _I, _P,
                                             _I, _P, _P]),
    "sponge_lj_pme_direct_force_list_noalloc": (_I, [_I, _I, _P, _P, _P, _F, _P, _P, _I, _P, _P,
                                                     _P, _FP, _FP, _F, _I, _P, _I, _P, _P, _P,
                                                     _P]),
    "sponge_solvent_lj_initial": (_I, [C.POINTER(C.c_void_p), _P, _I, _I, _IP, _FP, _I, _I, _P]),
    "sponge_solvent_lj_info": (_I, [_P, _IP, _IP, _IP, _IP]),
    "sponge_solvent_lj_direct_force": (_I, [_P, _P, _I, _I, _P, _P, _P, _FP, _FP, _F, _I, _P, _I,
                                            _P, _P]),
    "sponge_solvent_lj_clear": (_I, [_P]),
    "sponge_lj_long_range_correction": (_I, [_P, _I, _P, _I, _P, _F]),
    "sponge_lj_step_print": (_I, [_P, _FP, _FP]),
    "sponge_lj_atom_energy": (_P, [_P]),
    "sponge_lj_clear": (_I, [_P]),
    "sponge_pme_default_options": (None, [C.POINTER(PMEOptions)]),
    "sponge_pme_initial": (_I, [C.POINTER(_P), _I, _FP, _FP, _F, C.POINTER(PMEOptions), _P]),
    "sponge_pme_info": (_I, [_P, _FP, C.POINTER(_I)]),
    "sponge_pme_direct_atom_energy": (_P, [_P]),
    "sponge_pme_reciprocal_force": (_I, [_P, _P, _FP, _FP, _P, _P, _I, _I, _P, _P, _I]),
    "sponge_pme_excluded_force": (_I, [_P, _I, _P, _P, _P, _FP, _FP, _P, _P, _P, _P, _P, _I, _P,
                                       _I, _P]),
    "sponge_pme_step_print": (_I, [_P, _FP, _FP, _FP, _FP, _FP]),
    "sponge_pme_grids": (_I, [_P, _P, _P]),
    "sponge_pme_clear": (_I, [_P]),
    "sponge_md_initial": (_I, [C.POINTER(_P), _I, _FP, _FP, _FP, _FP, _I, _FP, _FP, C.POINTER(_I),
                               C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), _I, C.POINTER(_I),
                               _FP, _F, _F, _F, C.POINTER(NLOptions), C.POINTER(PMEOptions), _P]),
    "sponge_md_run": (_I, [_P, _I, _I, _I]),
    "sponge_md_autotune": (_I, [_P, _I, C.POINTER(_I)]),
    "sponge_md_set_reciprocal_overlap": (_I, [_P, _I]),
    "sponge_md_set_soft_core": (_I, [_P, _P]),
    "sponge_md_set_surrogate": (_I, [_P, _P, _I, _I, _F]),
    "sponge_md_surrogate_stats": (_I, [_P, C.POINTER(C.c_longlong), C.POINTER(C.c_double)]),
    "sponge_md_collect_frames": (_I, [_P, _P, _I, _I]),
    "sponge_pme_update_box": (_I, [_P, _FP, _FP]),
    "sponge_pme_autotune": (_I, [_P, _P, _P, _I, C.POINTER(_I)]),
    "sponge_pme_set_fft_tiling": (_I, [_P, C.POINTER(_I)]),
    "sponge_batch_initial": (_I, [C.POINTER(_P), _I, C.POINTER(_P)]),
    "sponge_batch_autotune": (_I, [_P, _I, C.POINTER(_I)]),
    "sponge_batch_run": (_I, [_P, _I, _I]),
    "sponge_batch_synchronize": (_I, [_P]),
    "sponge_batch_clear": (_I, [_P]),
    "sponge_ljsc_default_options": (None, [C.c_void_p]),
    "sponge_ljsc_initial": (_I, [C.POINTER(_P), _I, _I, _I, _FP, _FP, _FP, _FP, C.POINTER(_I),
                                 C.POINTER(_I), _F, C.c_void_p, _P]),
    "sponge_ljsc_long_range_factors": (_I, [_P, _FP, _FP]),
    "sponge_ljsc_direct_force": (_I, [_P, _P, _P, _P, _P, _FP, _FP, _F, _I, _P, _I, _P, _P]),
    "sponge_ljsc_du_dlambda": (_I, [_P, _P, _P, _P, _P, _FP, _FP, _F, _FP, _FP]),
    "sponge_ljsc_long_range_correction": (_I, [_P, _I, _P, _I, _P, _F]),
    "sponge_ljsc_step_print": (_I, [_P, _FP]),
    "sponge_ljsc_clear": (_I, [_P]),
    "sponge_rsur_default_options": (None, [C.c_void_p]),
    "sponge_rsur_model_size": (_I, []),
    "sponge_rsur_initial": (_I, [C.POINTER(_P), _I, _F, C.c_void_p, _P]),
    "sponge_rsur_features": (_I, [_P, _P, _P, _FP, _FP, _P, _P, _P, _P]),
    "sponge_rsur_export_features": (_I, [_P, _P, _P]),
    "sponge_rsur_predict": (_I, [_P, _P]),
    "sponge_rsur_get_model": (_I, [_P, _FP]),
    "sponge_rsur_set_model": (_I, [_P, _FP]),
    "sponge_rsur_add_frame": (_I, [_P, _P]),
    "sponge_rsur_frame_count": (_I, [_P, C.POINTER(_I)]),
    "sponge_rsur_fit_normalization": (_I, [_P, _I]),
    "sponge_rsur_gradient": (_I, [_P, _I, _FP, C.POINTER(C.c_double)]),
    "sponge_rsur_train": (_I, [_P, _I, _I, _F, C.POINTER(C.c_double)]),
    "sponge_rsur_evaluate": (_I, [_P, _I, _I, _I, C.POINTER(C.c_double)]),
    "sponge_rsur_clear": (_I, [_P]),
    "sponge_md_energies": (_I, [_P, C.POINTER(C.c_double)]),
    "sponge_md_get_state": (_I, [_P, _FP, _FP, _FP]),
    "sponge_md_set_state": (_I, [_P, _FP, _FP]),
    "sponge_md_request_list_update": (_I, [_P]),
    "sponge_md_set_list_prune": (_I, [_P, _F]),
    "sponge_md_set_surrogate_mode": (_I, [_P, _I]),
    "sponge_md_handles": (_I, [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    "sponge_md_clear": (_I, [_P]),
}

_lib = None


def load():
    """Load the HIP library (raises when it is missing or lacks a symbol)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m sponge_amd.build` "
                           "(there is no CPU fallback)")
    # PyTorch-ROCm ships its own HIP runtime with the same soname; loading it
    # first makes this library bind to that one runtime instead of pulling a
    # second copy from /opt/rocm into the process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)  # AttributeError when a declared symbol is not exported
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def set_kernel_timing(enable):
    _check(load().sponge_amd_set_kernel_timing(int(enable)), "set_kernel_timing")


def set_kernel_timing_filter(names=None):
    """Time only the named kernels (an iterable of timer names; None = all)."""
    arg = None if not names else ",".join(names).encode()
    _check(load().sponge_amd_set_kernel_timing_filter(arg), "set_kernel_timing_filter")


def kernel_timing(name):
    """(total milliseconds, launches) of one kernel timer (synchronises)."""
    ms = C.c_double()
    n = C.c_longlong()
    _check(load().sponge_amd_kernel_timing(name.encode(), C.byref(ms), C.byref(n)), "kernel_timing")
    return ms.value, n.value


def reset_kernel_timing():
    _check(load().sponge_amd_reset_kernel_timing(), "reset_kernel_timing")


def _fp(a):
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(_FP)


def _ptr(t):
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def _stream_ptr(stream):
    if stream is None:
        return None
    return C.c_void_p(stream.cuda_stream)


class Box:
    """Boundary of the reference (cell and rcell LTMatrix3) for an
    orthorhombic box, rcell inverted in float32 as
    Invert_Lower_Triangular_Cell (SPONGE/utils/boundary.h:59-70)."""

    def __init__(self, lengths):
        L = np.asarray(lengths, np.float32)
        one = np.float32(1.0)
        self.cell = np.array([L[0], 0, L[1], 0, 0, L[2]], np.float32)
        self.rcell = np.array([one / L[0], 0, one / L[1], 0, 0, one / L[2]], np.float32)
        self.volume = float(np.float32(L[0] * L[1]) * L[2])

    @classmethod
    def from_cell(cls, cell, rcell):
        """A general lower-triangular cell {a11, a21, a22, a31, a32, a33} and
        its inverse, as given (the neighbor list, the pair kernels and the
        PME module take any such cell)."""
        b = cls.__new__(cls)
        b.cell = np.ascontiguousarray(cell, np.float32).copy()
        b.rcell = np.ascontiguousarray(rcell, np.float32).copy()
        b.volume = float(np.float32(b.cell[0] * b.cell[2]) * b.cell[5])
        return b

    def args(self):
        return self.cell.ctypes.data_as(_FP), self.rcell.ctypes.data_as(_FP)


class NeighborList:
    """NEIGHBOR_LIST (SPONGE/neighbor_list/neighbor_list.h:7-127)."""

    def __init__(self, atom_numbers, cutoff, skin, box: Box, stream=None, **opts):
        self.lib = load()
        o = NLOptions()
        self.lib.sponge_nl_default_options(C.byref(o))
        for k, v in opts.items():
            setattr(o, k, v)
        self.opt = o
        h = C.c_void_p()
        _check(self.lib.sponge_nl_initial(C.byref(h), int(atom_numbers), float(cutoff), float(skin),
                                          *box.args(), C.byref(o), _stream_ptr(stream)),
               "sponge_nl_initial")
        self.h = h
        self.N = int(atom_numbers)
        self.box = box

    def update(self, crd, step=0, update=NL_CONDITIONAL, excl=None, atom_local=None):
        es, el, en = excl if excl is not None else (None, None, None)
        _check(self.lib.sponge_nl_update(self.h, _ptr(atom_local), self.N, 0, _ptr(crd),
                                         *self.box.args(), int(step), int(update), _ptr(es),
                                         _ptr(el), _ptr(en)), "sponge_nl_update")

    def check_overflow(self):
        f = C.c_int()
        rc = self.lib.sponge_nl_check_overflow(self.h, C.byref(f))
        return rc, f.value

    def options(self):
        o = NLOptions()
        _check(self.lib.sponge_nl_get_options(self.h, C.byref(o)), "sponge_nl_get_options")
        return {k: getattr(o, k) for k, _ in NLOptions._fields_}

    def export_half_list(self, counts, serial, max_numbers):
        _check(self.lib.sponge_nl_export_half_list(self.h, _ptr(counts), _ptr(serial),
                                                   int(max_numbers)), "sponge_nl_export_half_list")

    def info(self):
        n = C.c_longlong()
        d = (C.c_int * 3)()
        _check(self.lib.sponge_nl_info(self.h, C.byref(n), d), "sponge_nl_info")
        return n.value, tuple(d)

    def clear(self):
        if self.h:
            self.lib.sponge_nl_clear(self.h)
            self.h = None

    __del__ = clear


class LennardJones:
    """LENNARD_JONES_INFORMATION (Lennard_Jones_force.h:87-173)."""

    def __init__(self, atom_numbers, type_numbers, A, B, types, cutoff, stream=None):
        self.lib = load()
        A, pA = _fp(A)
        B, pB = _fp(B)
        t = np.ascontiguousarray(types, np.int32)
        h = C.c_void_p()
        _check(self.lib.sponge_lj_initial(C.byref(h), int(atom_numbers), int(type_numbers), pA, pB,
                                          t.ctypes.data_as(C.POINTER(C.c_int)), float(cutoff),
                                          _stream_ptr(stream)), "sponge_lj_initial")
        self.h = h
        self.N = int(atom_numbers)

    def long_range_factor(self):
        f = C.c_float()
        _check(self.lib.sponge_lj_long_range_factor(self.h, C.byref(f)), "long_range_factor")
        return f.value

    def direct_force(self, nl, crd, charge, frc, box, beta, atom_energy=None, atom_virial=None,
                     direct_energy=None, solvent_numbers=0):
        _check(self.lib.sponge_lj_pme_direct_force(
            self.h, nl.h, self.N, self.N, int(solvent_numbers), 0, _ptr(crd), _ptr(charge),
            _ptr(frc), *box.args(),
            float(beta), int(atom_energy is not None), _ptr(atom_energy),
            int(atom_virial is not None), _ptr(atom_virial), _ptr(direct_energy)),
            "sponge_lj_pme_direct_force")

    def direct_force_list(self, counts, serial, max_numbers, crd, charge, frc, box, beta,
                          atom_energy=None, atom_virial=None, direct_energy=None):
        _check(self.lib.sponge_lj_pme_direct_force_list(
            self.h, _ptr(counts), _ptr(serial), int(max_numbers), self.N, _ptr(crd), _ptr(charge),
            _ptr(frc), *box.args(), float(beta), int(atom_energy is not None), _ptr(atom_energy),
            int(atom_virial is not None), _ptr(atom_virial), _ptr(direct_energy)),
            "sponge_lj_pme_direct_force_list")

    def long_range_correction(self, need_pressure, virial, need_potential, potential, volume):
        _check(self.lib.sponge_lj_long_range_correction(self.h, int(need_pressure), _ptr(virial),
                                                        int(need_potential), _ptr(potential),
                                                        float(volume)), "long_range_correction")

    def step_print(self):
        a, b = C.c_float(), C.c_float()
        _check(self.lib.sponge_lj_step_print(self.h, C.byref(a), C.byref(b)), "lj step_print")
        return a.value, b.value

    def atom_energy_ptr(self):
        return self.lib.sponge_lj_atom_energy(self.h)

    def clear(self):
        if self.h:
            self.lib.sponge_lj_clear(self.h)
            self.h = None

    __del__ = clear


class SolventLJ:
    """SOLVENT_LENNARD_JONES (SPONGE/Lennard_Jones_force/solvent_LJ.h:9-50)
    over a LennardJones module: the water tail's own pairs."""

    def __init__(self, lj, s, solvent_LJ=-1, default_enable=True, stream=None):
        self.lib = load()
        self.lj = lj
        rs = np.ascontiguousarray(s.residue_starts(), np.int32)
        m = np.ascontiguousarray(s.mass, np.float32)
        self.R = len(rs)
        h = C.c_void_p()
        _check(self.lib.sponge_solvent_lj_initial(
            C.byref(h), lj.h, int(s.n_atoms), self.R, rs.ctypes.data_as(_IP),
            m.ctypes.data_as(_FP), int(solvent_LJ), int(bool(default_enable)),
            _stream_ptr(stream)), "sponge_solvent_lj_initial")
        self.h = h
        self.N = int(s.n_atoms)

    def info(self):
        v = [C.c_int() for _ in range(4)]
        _check(self.lib.sponge_solvent_lj_info(self.h, *[C.byref(x) for x in v]),
               "sponge_solvent_lj_info")
        return dict(enabled=bool(v[0].value), solvent_numbers=v[1].value,
                    solvent_start=v[2].value, water_points=v[3].value)

    def direct_force(self, nl, crd, charge, frc, box, beta, atom_energy=None, atom_virial=None,
                     direct_energy=None):
        _check(self.lib.sponge_solvent_lj_direct_force(
            self.h, nl.h, self.N, self.R, _ptr(crd), _ptr(charge), _ptr(frc), *box.args(),
            float(beta), int(atom_energy is not None), _ptr(atom_energy),
            int(atom_virial is not None), _ptr(atom_virial), _ptr(direct_energy)),
            "sponge_solvent_lj_direct_force")

    def clear(self):
        if self.h:
            self.lib.sponge_solvent_lj_clear(self.h)
            self.h = None

    __del__ = clear


class ParticleMesh:
    """Particle_Mesh (SPONGE/PM_force/PM_force.h:152-209)."""

    def __init__(self, atom_numbers, box: Box, cutoff, stream=None, **opts):
        self.lib = load()
        o = PMEOptions()
        self.lib.sponge_pme_default_options(C.byref(o))
        for k, v in opts.items():
            setattr(o, k, v)
        h = C.c_void_p()
        _check(self.lib.sponge_pme_initial(C.byref(h), int(atom_numbers), *box.args(),
                                           float(cutoff), C.byref(o), _stream_ptr(stream)),
               "sponge_pme_initial")
        self.h = h
        self.N = int(atom_numbers)
        self.box = box
        b = C.c_float()
        f = (C.c_int * 3)()
        _check(self.lib.sponge_pme_info(self.h, C.byref(b), f), "sponge_pme_info")
        self.beta = b.value
        self.fft = tuple(f)

    def direct_atom_energy_ptr(self):
        return self.lib.sponge_pme_direct_atom_energy(self.h)

    def reciprocal_force(self, crd, charge, frc, step=0, virial=None, potential=None):
        _check(self.lib.sponge_pme_reciprocal_force(
            self.h, _ptr(crd), *self.box.args(), _ptr(charge), _ptr(frc), int(virial is not None),
            int(potential is not None), _ptr(virial), _ptr(potential), int(step)),
            "sponge_pme_reciprocal_force")

    def excluded_force(self, crd, charge, excl, frc, atom_energy=None, atom_virial=None):
        es, el, en = excl
        _check(self.lib.sponge_pme_excluded_force(
            self.h, self.N, None, None, _ptr(crd), *self.box.args(), _ptr(charge), _ptr(es),
            _ptr(el), _ptr(en), _ptr(frc), int(atom_energy is not None), _ptr(atom_energy),
            int(atom_virial is not None), _ptr(atom_virial)), "sponge_pme_excluded_force")

    def step_print(self):
        v = [C.c_float() for _ in range(5)]
        _check(self.lib.sponge_pme_step_print(self.h, *[C.byref(x) for x in v]), "pme step_print")
        return dict(PM=v[0].value, direct=v[1].value, reciprocal=v[2].value,
                    self_energy=v[3].value, correction=v[4].value)

    def update_box(self, box):
        """Particle_Mesh::Update_Box for a new Box (same grid)."""
        _check(self.lib.sponge_pme_update_box(self.h, *box.args()), "sponge_pme_update_box")
        self.box = box

    def grids(self, Q=None, FBCFQ=None):
        _check(self.lib.sponge_pme_grids(self.h, _ptr(Q), _ptr(FBCFQ)), "sponge_pme_grids")

    def clear(self):
        if self.h:
            self.lib.sponge_pme_clear(self.h)
            self.h = None

    __del__ = clear


class LJSoftCoreOptions(C.Structure):
    _fields_ = [("lambda_", C.c_float), ("alpha", C.c_float), ("p", C.c_float),
                ("sigma", C.c_float), ("sigma_min", C.c_float)]


class LJSoftCore:
    """LJ_SOFT_CORE (SPONGE/Lennard_Jones_force/LJ_soft_core.h)."""

    def __init__(self, atom_numbers, AA, AB, BA, BB, typeA, typeB, cutoff, stream=None,
                 **opts):
        self.lib = load()
        o = LJSoftCoreOptions()
        self.lib.sponge_ljsc_default_options(C.byref(o))
        for k, v in opts.items():
            setattr(o, "lambda_" if k == "lambda" else k, v)
        tA = np.ascontiguousarray(typeA, np.int32)
        tB = np.ascontiguousarray(typeB, np.int32)
        TA = int(round((np.sqrt(8 * len(AA) + 1) - 1) / 2))
        TB = int(round((np.sqrt(8 * len(BA) + 1) - 1) / 2))
        arrs = [_fp(x) for x in (AA, AB, BA, BB)]
        h = C.c_void_p()
        ip = C.POINTER(C.c_int)
        _check(self.lib.sponge_ljsc_initial(C.byref(h), int(atom_numbers), TA, TB,
                                            *[p for _, p in arrs], tA.ctypes.data_as(ip),
                                            tB.ctypes.data_as(ip), float(cutoff), C.byref(o),
                                            _stream_ptr(stream)), "sponge_ljsc_initial")
        self.h = h
        self.N = int(atom_numbers)

    def direct_force(self, nl, crd, charge, frc, box, beta, atom_energy=None, atom_virial=None,
                     direct_energy=None):
        _check(self.lib.sponge_ljsc_direct_force(
            self.h, nl.h, _ptr(crd), _ptr(charge), _ptr(frc), *box.args(), float(beta),
            int(atom_energy is not None), _ptr(atom_energy), int(atom_virial is not None),
            _ptr(atom_virial), _ptr(direct_energy)), "sponge_ljsc_direct_force")

    def du_dlambda(self, nl, crd, charge, box, beta, charge_BA=None):
        a, b = C.c_float(), C.c_float()
        _check(self.lib.sponge_ljsc_du_dlambda(self.h, nl.h, _ptr(crd), _ptr(charge),
                                               _ptr(charge_BA), *box.args(), float(beta),
                                               C.byref(a), C.byref(b)), "sponge_ljsc_du_dlambda")
        return a.value, b.value

    def long_range_factors(self):
        a, b = C.c_float(), C.c_float()
        _check(self.lib.sponge_ljsc_long_range_factors(self.h, C.byref(a), C.byref(b)),
               "sponge_ljsc_long_range_factors")
        return a.value, b.value

    def step_print(self):
        a = C.c_float()
        _check(self.lib.sponge_ljsc_step_print(self.h, C.byref(a)), "sponge_ljsc_step_print")
        return a.value

    def clear(self):
        if self.h:
            self.lib.sponge_ljsc_clear(self.h)
            self.h = None

    __del__ = clear


class Batch:
    """Batched driver over MDSystem instances (one graph launch advances all)."""

    def __init__(self, systems):
        self.lib = load()
        self.systems = list(systems)
        arr = (_P * len(self.systems))(*[m.h for m in self.systems])
        h = C.c_void_p()
        _check(self.lib.sponge_batch_initial(C.byref(h), len(self.systems), arr),
               "sponge_batch_initial")
        self.h = h

    def autotune(self, reps=5):
        t = (C.c_int * (3 * len(self.systems)))()
        _check(self.lib.sponge_batch_autotune(self.h, int(reps), t), "sponge_batch_autotune")
        return [tuple(t[3 * i:3 * i + 3]) for i in range(len(self.systems))]

    def run(self, steps, steps_per_launch=8):
        _check(self.lib.sponge_batch_run(self.h, int(steps), int(steps_per_launch)),
               "sponge_batch_run")

    def synchronize(self):
        _check(self.lib.sponge_batch_synchronize(self.h), "sponge_batch_synchronize")

    def clear(self):
        if self.h:
            self.lib.sponge_batch_clear(self.h)
            self.h = None

    __del__ = clear


class RsurOptions(C.Structure):
    _fields_ = [("feature_cutoff", C.c_float), ("max_frames", C.c_int), ("seed", C.c_int)]


class ReciprocalSurrogate:
    """Learned reciprocal-force surrogate (no counterpart in the reference;
    it stands beside Particle_Mesh's reciprocal force, PM_force.cpp:1070-1169)."""

    N_FEAT = 48

    def __init__(self, atom_numbers, beta, stream=None, **opts):
        self.lib = load()
        o = RsurOptions()
        self.lib.sponge_rsur_default_options(C.byref(o))
        for k, v in opts.items():
            setattr(o, k, v)
        self.opt = o
        h = C.c_void_p()
        _check(self.lib.sponge_rsur_initial(C.byref(h), int(atom_numbers), float(beta), C.byref(o),
                                            _stream_ptr(stream)), "sponge_rsur_initial")
        self.h = h
        self.N = int(atom_numbers)
        self.model_size = self.lib.sponge_rsur_model_size()

    def features(self, nl, crd, charge, box, excl=None):
        es, el, en = excl if excl is not None else (None, None, None)
        _check(self.lib.sponge_rsur_features(self.h, nl.h, _ptr(crd), *box.args(), _ptr(charge),
                                             _ptr(es), _ptr(el), _ptr(en)), "sponge_rsur_features")

    def export_features(self, feat=None, basis=None):
        _check(self.lib.sponge_rsur_export_features(self.h, _ptr(feat), _ptr(basis)),
               "sponge_rsur_export_features")

    def predict(self, frc):
        _check(self.lib.sponge_rsur_predict(self.h, _ptr(frc)), "sponge_rsur_predict")

    def get_model(self):
        m = np.zeros(self.model_size, np.float32)
        _check(self.lib.sponge_rsur_get_model(self.h, m.ctypes.data_as(_FP)), "sponge_rsur_get_model")
        return m

    def set_model(self, model):
        m = np.ascontiguousarray(model, np.float32)
        assert m.size == self.model_size
        _check(self.lib.sponge_rsur_set_model(self.h, m.ctypes.data_as(_FP)), "sponge_rsur_set_model")

    def add_frame(self, rec_frc):
        _check(self.lib.sponge_rsur_add_frame(self.h, _ptr(rec_frc)), "sponge_rsur_add_frame")

    def frame_count(self):
        n = C.c_int()
        _check(self.lib.sponge_rsur_frame_count(self.h, C.byref(n)), "sponge_rsur_frame_count")
        return n.value

    def fit_normalization(self, n_frames):
        _check(self.lib.sponge_rsur_fit_normalization(self.h, int(n_frames)),
               "sponge_rsur_fit_normalization")

    def gradient(self, frame):
        g = np.zeros(self.model_size - 2 * self.N_FEAT, np.float32)
        loss = C.c_double()
        _check(self.lib.sponge_rsur_gradient(self.h, int(frame), g.ctypes.data_as(_FP),
                                             C.byref(loss)), "sponge_rsur_gradient")
        return loss.value, g

    def train(self, n_train_frames, epochs, lr):
        out = (C.c_double * max(1, epochs))()
        _check(self.lib.sponge_rsur_train(self.h, int(n_train_frames), int(epochs), float(lr), out),
               "sponge_rsur_train")
        return list(out)[:epochs]

    def evaluate(self, begin, end, lag=0):
        """Relative RMS errors over stored frames [begin, end): model,
        baseline and, with lag > 0, model / baseline anchored on frame f - lag
        and the stale force of frame f - lag."""
        out = (C.c_double * 5)()
        _check(self.lib.sponge_rsur_evaluate(self.h, int(begin), int(end), int(lag), out),
               "sponge_rsur_evaluate")
        keys = ("model", "baseline", "model_anchored", "baseline_anchored", "stale")
        return dict(zip(keys, list(out)))

    def clear(self):
        if self.h:
            self.lib.sponge_rsur_clear(self.h)
            self.h = None

    __del__ = clear


class MDSystem:
    """NVE step driver over a System (sponge_amd.system.System)."""

    ENERGY_KEYS = ("potential", "kinetic", "LJ_short", "LJ_long", "PM", "PM_direct",
                   "PM_reciprocal", "PM_self", "PM_correction", "atom_energy_sum")

    def __init__(self, s, stream=None, nl_opts=None, pme_opts=None):
        self.lib = load()
        self.s = s
        N = s.n_atoms
        no = NLOptions()
        self.lib.sponge_nl_default_options(C.byref(no))
        for k, v in (nl_opts or {}).items():
            setattr(no, k, v)
        po = PMEOptions()
        self.lib.sponge_pme_default_options(C.byref(po))
        if s.fft is not None:
            po.fftx, po.ffty, po.fftz = s.fft
        po.direct_tolerance = s.tolerance
        for k, v in (pme_opts or {}).items():
            setattr(po, k, v)
        ip = C.POINTER(C.c_int)
        arr = lambda a, t=np.float32: np.ascontiguousarray(a, t)
        self._keep = [arr(s.crd), arr(s.vel), arr(s.mass), arr(s.charge), arr(s.lj_A), arr(s.lj_B),
                      arr(s.lj_type, np.int32), arr(s.excl_start, np.int32),
                      arr(s.excl_num, np.int32),
                      arr(s.excl_list if len(s.excl_list) else np.zeros(1), np.int32),
                      arr(np.arange(s.water_start, s.water_start + 3 * s.n_waters), np.int32),
                      arr(s.box)]
        k = self._keep
        h = C.c_void_p()
        _check(self.lib.sponge_md_initial(
            C.byref(h), N, k[0].ctypes.data_as(_FP), k[1].ctypes.data_as(_FP),
            k[2].ctypes.data_as(_FP), k[3].ctypes.data_as(_FP), s.n_types,
            k[4].ctypes.data_as(_FP), k[5].ctypes.data_as(_FP), k[6].ctypes.data_as(ip),
            k[7].ctypes.data_as(ip), k[8].ctypes.data_as(ip), k[9].ctypes.data_as(ip), s.n_waters,
            k[10].ctypes.data_as(ip), k[11].ctypes.data_as(_FP), float(s.cutoff), float(s.skin),
            float(s.dt_ps), C.byref(no), C.byref(po), _stream_ptr(stream)), "sponge_md_initial")
        self.h = h
        self.N = N

    def run(self, steps, energy_interval=0, use_graph=False):
        _check(self.lib.sponge_md_run(self.h, int(steps), int(energy_interval), int(use_graph)),
               "sponge_md_run")

    def energies(self):
        out = (C.c_double * 10)()
        _check(self.lib.sponge_md_energies(self.h, out), "sponge_md_energies")
        return dict(zip(self.ENERGY_KEYS, list(out)))

    def state(self):
        crd = np.zeros((self.N, 3), np.float32)
        vel = np.zeros((self.N, 3), np.float32)
        frc = np.zeros((self.N, 3), np.float32)
        _check(self.lib.sponge_md_get_state(self.h, crd.ctypes.data_as(_FP), vel.ctypes.data_as(_FP),
                                            frc.ctypes.data_as(_FP)), "sponge_md_get_state")
        return crd, vel, frc

    def set_state(self, crd=None, vel=None):
        c = None if crd is None else np.ascontiguousarray(crd, np.float32)
        v = None if vel is None else np.ascontiguousarray(vel, np.float32)
        _check(self.lib.sponge_md_set_state(self.h, None if c is None else c.ctypes.data_as(_FP),
                                            None if v is None else v.ctypes.data_as(_FP)),
               "sponge_md_set_state")

    def equilibrate(self, steps, temperature=300.0, every=10):
        """Velocity rescaling to `temperature` every `every` steps (set-up of
        a synthetic start; not part of any timed region)."""
        from .system import KB
        m = np.asarray(self.s.mass, np.float64)
        # rigid 3-site waters have 6 degrees of freedom each, ions 3
        dof = 6 * self.s.n_waters + 3 * (self.N - 3 * self.s.n_waters) - 3
        for _ in range(max(1, steps // every)):
            self.run(every)
            _, vel, _ = self.state()
            ek = 0.5 * (m[:, None] * vel.astype(np.float64) ** 2).sum()
            t = 2 * ek / (dof * KB)
            if t > 0:
                self.set_state(vel=vel * np.float32(np.sqrt(temperature / t)))

    def autotune(self, reps=5):
        """Tunes the PME FFT tiling on the current coordinates; returns it."""
        t = (C.c_int * 3)()
        _check(self.lib.sponge_md_autotune(self.h, int(reps), t), "sponge_md_autotune")
        return tuple(t)

    def set_reciprocal_overlap(self, mode):
        """1: the reciprocal half of a plain PME step on a second stream
        beside the pair work; 0: one stream; -1: from the atom count."""
        _check(self.lib.sponge_md_set_reciprocal_overlap(self.h, int(mode)),
               "sponge_md_set_reciprocal_overlap")

    def set_soft_core(self, sc):
        """Lambda window: the step's LJ + PME direct term is the soft-core
        module sc (an LJSoftCore over this system's atoms; the system's
        charges are the window's); None returns to the plain LJ module."""
        _check(self.lib.sponge_md_set_soft_core(self.h, sc.h if sc is not None else None),
               "sponge_md_set_soft_core")
        self._sc = sc  # the module must outlive its use by the driver

    def set_surrogate(self, rsur, period=4, check_interval=8, bound=0.01):
        _check(self.lib.sponge_md_set_surrogate(self.h, rsur.h if rsur is not None else None,
                                                int(period), int(check_interval), float(bound)),
               "sponge_md_set_surrogate")

    def set_surrogate_mode(self, mode):
        """1: impulse schedule (default), 0: anchored."""
        _check(self.lib.sponge_md_set_surrogate_mode(self.h, int(mode)),
               "sponge_md_set_surrogate_mode")

    def surrogate_stats(self):
        out = (C.c_longlong * 4)()
        err = C.c_double()
        _check(self.lib.sponge_md_surrogate_stats(self.h, out, C.byref(err)),
               "sponge_md_surrogate_stats")
        return dict(surrogate_steps=out[0], checks=out[1], fallbacks=out[2],
                    switched_off=bool(out[3]), last_error=err.value)

    def collect_frames(self, rsur, n_frames, stride):
        _check(self.lib.sponge_md_collect_frames(self.h, rsur.h, int(n_frames), int(stride)),
               "sponge_md_collect_frames")

    def nl_options(self):
        """Options of the driver's current neighbor list (an overflow
        re-initialises it with grown limits)."""
        nl, _, _ = self.handles()
        o = NLOptions()
        _check(self.lib.sponge_nl_get_options(nl, C.byref(o)), "sponge_nl_get_options")
        return {k: getattr(o, k) for k, _ in NLOptions._fields_}

    def set_list_prune(self, delta):
        """Inner rows within cutoff + delta for the pair kernel (0: off)."""
        _check(self.lib.sponge_md_set_list_prune(self.h, float(delta)), "sponge_md_set_list_prune")

    def request_list_update(self):
        """The next step (alone or batched) rebuilds the list."""
        _check(self.lib.sponge_md_request_list_update(self.h), "sponge_md_request_list_update")

    def pme_grid(self):
        """The PME grid of the driver's reciprocal module."""
        _, _, pm = self.handles()
        b = C.c_float()
        f = (C.c_int * 3)()
        _check(self.lib.sponge_pme_info(pm, C.byref(b), f), "sponge_pme_info")
        return tuple(f)

    def handles(self):
        nl, lj, pm = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(self.lib.sponge_md_handles(self.h, C.byref(nl), C.byref(lj), C.byref(pm)),
               "sponge_md_handles")
        return nl, lj, pm

    def clear(self):
        if self.h:
            self.lib.sponge_md_clear(self.h)
            self.h = None

    __del__ = clear


NOALLOC_MAX_ATOMS = 262144
NOALLOC_MAX_TYPES = 64


def lj_direct_force_list_noalloc(n_types, lj_A, lj_B, lj_type, cutoff, counts, serial,
                                 max_numbers, crd, charge, frc, box, beta, atom_energy=None,
                                 atom_virial=None, lj_energy=None, direct_energy=None,
                                 stream=None):
    """sponge_lj_pme_direct_force_list_noalloc: the direct force over a
    reference-format half list with no module handle and no device
    allocation (every argument array is a caller-owned device tensor)."""
    lib = load()
    _check(lib.sponge_lj_pme_direct_force_list_noalloc(
        int(crd.shape[0]), int(n_types), _ptr(lj_A), _ptr(lj_B), _ptr(lj_type), float(cutoff),
        _ptr(counts), _ptr(serial), int(max_numbers), _ptr(crd), _ptr(charge), _ptr(frc),
        *box.args(), float(beta), int(atom_energy is not None), _ptr(atom_energy),
        int(atom_virial is not None), _ptr(atom_virial), _ptr(lj_energy), _ptr(direct_energy),
        _stream_ptr(stream)), "sponge_lj_pme_direct_force_list_noalloc")
