"""Readers of the reference's native input files and mdin, for the parts of
the system this path uses.

Formats (as the reference's loaders read them):
  coordinate_in_file  "N [start_time]", N lines "x y z", then the box lengths
                      and angles (SPONGE/xponge/load/native/md_core.hpp:106-190)
  charge_in_file      "N", N charges already in SPONGE units (x18.2223) (:57-103)
  mass_in_file        "N", N masses (:9-55)
  LJ_in_file          "N T", the T(T+1)/2 lower-triangle A then B coefficients,
                      scaled on load by 12 and 6, then N type indices
                      (SPONGE/xponge/load/native/lj.hpp:8-69)
  exclude_in_file     "N total", per atom its count then that many partner
                      ids; the count total must equal `total` (md_core.hpp:258-310)
  mdin                a title line, then "key = value" lines; `#` starts a
                      comment; TOML-like "[table]" headers prefix later keys
                      with "table_" (SPONGE/control.cpp:50-94)
Every reader raises ValueError on a malformed file, where the reference
raises its BadFileFormat error.

Writers of the outputs the reference's tools read:
  frc.dat   float32 VECTOR[N] per frame, appended (MD_core/output.hpp:1691-1706)
  mdout     a header of the printed keys, then one row per print step, every
            field "%15s" followed by a space (a newline after the last)
            (SPONGE/utils/control/print.hpp:143-205)
"""
from __future__ import annotations

import os

import numpy as np

from .system import System


def _tokens(path):
    with open(path) as f:
        return f.read().split()


def read_coordinate(path):
    """(crd [N, 3] float32, box lengths [3] float32, box angles [3] float32)."""
    with open(path) as f:
        head = f.readline().split()
        if not head:
            raise ValueError(f"{path}: empty coordinate file")
        n = int(head[0])
        rest = f.read().split()
    if len(rest) < 3 * n + 6:
        raise ValueError(f"{path}: expected {3 * n} coordinates and a box")
    v = np.array(rest[:3 * n + 6], np.float32)
    return v[:3 * n].reshape(n, 3), v[3 * n:3 * n + 3], v[3 * n + 3:3 * n + 6]


def read_per_atom(path, dtype=np.float32):
    t = _tokens(path)
    n = int(t[0])
    if len(t) < 1 + n:
        raise ValueError(f"{path}: expected {n} values")
    return np.array(t[1:1 + n], dtype)


def read_lj(path):
    """(types [N] int32, A [T(T+1)/2] float32 x12, B x6, T)."""
    t = _tokens(path)
    n, ntype = int(t[0]), int(t[1])
    npair = ntype * (ntype + 1) // 2
    if len(t) < 2 + 2 * npair + n:
        raise ValueError(f"{path}: the format of LJ_in_file is not right")
    A = np.array(t[2:2 + npair], np.float32) * np.float32(12.0)
    B = np.array(t[2 + npair:2 + 2 * npair], np.float32) * np.float32(6.0)
    types = np.array(t[2 + 2 * npair:2 + 2 * npair + n], np.int32)
    if types.min() < 0 or types.max() >= ntype:
        raise ValueError(f"{path}: LJ type out of range")
    return types, A, B, ntype


def read_exclude(path):
    """Exclusion CSR (start [N], num [N], list [E]) as listed per atom."""
    t = _tokens(path)
    n, total = int(t[0]), int(t[1])
    start = np.zeros(n, np.int32)
    num = np.zeros(n, np.int32)
    lst = []
    k = 2
    for i in range(n):
        if k >= len(t):
            raise ValueError(f"{path}: the format of exclude_in_file is not right")
        c = int(t[k])
        k += 1
        start[i] = len(lst)
        num[i] = c
        lst.extend(int(x) for x in t[k:k + c])
        k += c
    if len(lst) != total:
        raise ValueError(f"{path}: excluded_atom_numbers is not right")
    return start, num, np.array(lst, np.int32)


def read_mdin(path):
    """{key: value string} of an mdin file (keys lower case)."""
    out = {}
    table = ""
    with open(path) as f:
        lines = f.read().splitlines()
    for line in lines[1:]:
        line = line.split("#", 1)[0].strip()
        if not line:
            continue
        if line.startswith("[") and line.endswith("]"):
            table = line[1:-1].strip() + "_"
            continue
        if "=" not in line:
            continue
        k, v = line.split("=", 1)
        out[(table + k.strip()).lower()] = v.strip().strip('"')
    return out


def load_system(directory, prefix="sys_flexible", cutoff=10.0, skin=2.0, name=None,
                tolerance=1e-5, coordinate=None):
    """A System from native input files `<prefix>_{coordinate,charge,mass,LJ,
    exclude}.txt` in `directory` (velocities zero, no waters constrained);
    `coordinate` names another coordinate file."""
    p = lambda kind: os.path.join(directory, f"{prefix}_{kind}.txt")
    crd, box, angles = read_coordinate(coordinate or p("coordinate"))
    if not np.allclose(angles, 90.0):
        raise ValueError("only orthorhombic boxes are supported")
    n = crd.shape[0]
    charge = read_per_atom(p("charge"))
    mass = read_per_atom(p("mass")) if os.path.exists(p("mass")) else np.full(n, 20.0, np.float32)
    types, A, B, _ = read_lj(p("LJ"))
    es, en, el = read_exclude(p("exclude"))
    for arr, what in ((charge, "charge"), (mass, "mass"), (types, "LJ")):
        if len(arr) != n:
            raise ValueError(f"'atom_numbers' is different in the {what} file")
    if len(es) != n:
        raise ValueError("'atom_numbers' is different in the exclude file")
    return System(name=name or os.path.basename(os.path.normpath(directory)), box=box.copy(),
                  crd=crd, vel=np.zeros_like(crd), charge=charge, mass=mass, lj_type=types,
                  lj_A=A, lj_B=B, excl_start=es, excl_num=en, excl_list=el, n_waters=0,
                  cutoff=float(cutoff), skin=float(skin), fft=None, tolerance=tolerance)


def append_frc_frame(path, frc):
    """Appends one frame of forces (N x 3) to a frc.dat trajectory."""
    a = np.ascontiguousarray(frc, np.float32).reshape(-1)
    with open(path, "ab") as f:
        a.tofile(f)


class MdoutWriter:
    """mdout in the reference's layout: `keys` as the header, one row per
    write; values are formatted with `fmt` (a string is written as is)."""

    def __init__(self, path, keys, fmt="%.2f"):
        self.path = path
        self.keys = list(keys)
        self.fmt = fmt
        with open(path, "w") as f:
            f.write(self._line(self.keys))

    @staticmethod
    def _line(fields):
        return "".join("%15s%s" % (x, "\n" if i == len(fields) - 1 else " ")
                       for i, x in enumerate(fields))

    def write(self, values):
        row = []
        for k in self.keys:
            v = values[k]
            row.append(v if isinstance(v, str) else
                       ("%d" % v if isinstance(v, (int, np.integer)) else self.fmt % v))
        with open(self.path, "a") as f:
            f.write(self._line(row))


def read_mdout(path):
    """{key: [values]} of an mdout written in the reference's layout."""
    with open(path) as f:
        lines = [l.split() for l in f.read().splitlines() if l.strip()]
    keys = lines[0]
    out = {k: [] for k in keys}
    for row in lines[1:]:
        for k, v in zip(keys, row):
            out[k].append(float(v))
    return out
