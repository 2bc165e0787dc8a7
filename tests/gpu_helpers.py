"""Shared set-up of the GPU parity tests: systems on the device, module
instances through the C ABI, and comparison metrics."""
from __future__ import annotations

import numpy as np


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / max((b ** 2).sum(), 1e-300)))


def rel(a, b):
    return abs(float(a) - float(b)) / max(abs(float(b)), 1e-300)


def check_total(name, gpu, ref32, ref64, tol=1e-6):
    """Energy / virial totals: the GPU value must lie within `tol` (relative)
    of the FP64 oracle, or be no further from it than 1.5x the FP32
    restatement of the reference is (the restatement's own rounding sets
    the floor of any FP32 comparison).  Returns the three deviations."""
    gpu, ref32, ref64 = np.asarray(gpu, np.float64), np.asarray(ref32, np.float64), \
        np.asarray(ref64, np.float64)
    scale = max(float(np.abs(ref64).max()), 1e-300)
    d_gpu = float(np.abs(gpu - ref64).max()) / scale
    d_ref = float(np.abs(ref32 - ref64).max()) / scale
    d_gr = float(np.abs(gpu - ref32).max()) / scale
    print(f"check_total {name}: |gpu-fp64| {d_gpu:.2e} |ref32-fp64| {d_ref:.2e}")
    assert d_gpu <= max(tol, 1.5 * d_ref), \
        f"{name}: |gpu-fp64| {d_gpu:.3e}, |ref32-fp64| {d_ref:.3e}, |gpu-ref32| {d_gr:.3e}"
    assert d_gr <= tol + d_ref + d_gpu
    return d_gpu, d_ref, d_gr


class DevSystem:
    """Arrays of a System as device tensors in the reference layout."""

    def __init__(self, s, dev, crd=None):
        import torch
        self.s = s
        self.N = s.n_atoms
        c = s.crd if crd is None else crd
        self.crd = torch.tensor(np.ascontiguousarray(c, np.float32), device=dev)
        self.charge = torch.tensor(s.charge, device=dev)
        self.es = torch.tensor(s.excl_start, device=dev)
        self.en = torch.tensor(s.excl_num, device=dev)
        el = s.excl_list if len(s.excl_list) else np.zeros(1, np.int32)
        self.el = torch.tensor(el, device=dev)
        self.excl = (self.es, self.el, self.en)
        self.dev = dev

    def zeros(self, *shape, dtype=None):
        import torch
        return torch.zeros(*shape, dtype=dtype or torch.float32, device=self.dev)


def export_pairs(nl, N, dev, max_numbers=1200):
    """Reference-format half list of the GPU neighbor list as sorted pair keys."""
    import torch
    from oracle import oracle as O
    counts = torch.zeros(N, dtype=torch.int32, device=dev)
    serial = torch.zeros(N * max_numbers, dtype=torch.int32, device=dev)
    # the list's stream is not torch's: the zero fill must be done first
    torch.cuda.synchronize()
    nl.export_half_list(counts, serial, max_numbers)
    torch.cuda.synchronize()
    c = counts.cpu().numpy()
    s = serial.view(N, max_numbers).cpu().numpy()
    return O.pair_array(c, s), c, s
