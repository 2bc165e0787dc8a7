"""Triclinic (general lower-triangular) cells for the neighbor list and the
LJ + PME direct pair kernel.

The 1,944-atom water box is mapped into a sheared cell (box vectors (L, 0, 0),
(3, L, 0), (2, -1.5, L), fractional coordinates kept).  On the CPU the
oracle's reference cell algorithm (with the sheared-image stencil of
neighbor_list.cpp:96-181) is pinned against its brute-force pair search, and
its FP32 pair-kernel restatement against an FP64 numpy restatement of the
same pairs with the general minimum image (SPONGE/utils/boundary.h
Get_Displacement).  On the GPU the list must equal the brute-force set bit for
bit, and the pair kernel (energy path and force-only path) must meet the
force, energy and virial bars against them.  The PME reciprocal part takes
the same cells: the oracle's general-cell form is pinned by the invariance of
the lattice sum under a change of basis (an orthorhombic lattice written with
sheared cell vectors, errors shrinking with the grid as PME's own), and the GPU
must match the oracle on the sheared box.
"""
import dataclasses

import numpy as np
import pytest
from scipy.special import erfc

from oracle import oracle as O

TILT = (3.0, 2.0, -1.5)  # a21, a31, a32 in A


def _sheared():
    from sponge_amd import system as S
    s = S.config(1)
    L = s.box.astype(np.float64)
    H = np.array([[L[0], 0, 0], [TILT[0], L[1], 0], [TILT[1], TILT[2], L[2]]])
    crd = ((s.crd.astype(np.float64) / L) @ H).astype(np.float32)
    cell = np.array([H[0, 0], H[1, 0], H[1, 1], H[2, 0], H[2, 1], H[2, 2]], np.float32)
    Hf = np.array([[cell[0], 0, 0], [cell[1], cell[2], 0], [cell[3], cell[4], cell[5]]], np.float64)
    R = np.linalg.inv(Hf)
    rcell = np.array([R[0, 0], R[1, 0], R[1, 1], R[2, 0], R[2, 1], R[2, 2]], np.float32)

    class Sheared(S.System):
        def cell(self):
            return cell

        def rcell(self):
            return rcell

    fields = {f.name: getattr(s, f.name) for f in dataclasses.fields(s)}
    fields["crd"] = crd
    return Sheared(**fields), cell, rcell


@pytest.fixture(scope="module")
def tri():
    s, cell, rcell = _sheared()
    beta = O.get_beta(s.cutoff, s.tolerance)
    c, ser, flags = O.nl_build(s)
    assert flags == 0
    return s, cell, rcell, beta, c, ser


def _disp32(s, cell, rcell, ii, jj):
    """Get_Displacement's general form in FP32 (left to right, no FMA), as
    the reference and the kernels evaluate it: [P, 3] float32."""
    c32, r32 = cell.astype(np.float32), rcell.astype(np.float32)
    x32 = s.crd.astype(np.float32)
    dx, dy, dz = (x32[jj, k] - x32[ii, k] for k in range(3))
    half = np.float32(0.5)
    sx = np.floor(dx * r32[0] + dy * r32[1] + dz * r32[3] + half)
    sy = np.floor(dy * r32[2] + dz * r32[4] + half)
    sz = np.floor(dz * r32[5] + half)
    return np.stack([dx - (sx * c32[0] + sy * c32[1] + sz * c32[3]),
                     dy - (sy * c32[2] + sz * c32[4]), dz - sz * c32[5]], 1)


def _direct64(s, cell, rcell, beta, counts, serial):
    """FP64 LJ + direct Coulomb over the half-list pairs within the cutoff:
    forces, E_LJ, E_C and the virial as the FP64 oracle sums them.  The pair
    set is the FP32 one (sqrtf(r^2) < cutoff on the FP32 displacement), so
    that pairs within rounding of the cutoff count as in the reference."""
    N, M = serial.shape
    mask = np.arange(M)[None, :] < counts[:, None]
    ii = np.broadcast_to(np.arange(N)[:, None], (N, M))[mask]
    jj = serial[mask]
    x = s.crd.astype(np.float64)
    H = np.array([[cell[0], 0, 0], [cell[1], cell[2], 0], [cell[3], cell[4], cell[5]]], np.float64)
    R = np.array([[rcell[0], 0, 0], [rcell[1], rcell[2], 0], [rcell[3], rcell[4], rcell[5]]],
                 np.float64)
    d = x[jj] - x[ii]
    d -= np.floor(d @ R + 0.5) @ H
    o = _disp32(s, cell, rcell, ii, jj)
    keep = np.sqrt((o[:, 0] * o[:, 0] + o[:, 1] * o[:, 1]) + o[:, 2] * o[:, 2]) < np.float32(s.cutoff)
    ii, jj, d = ii[keep], jj[keep], d[keep]
    r2 = (d * d).sum(1)
    r = np.sqrt(r2)
    hi, lo = np.maximum(s.lj_type[ii], s.lj_type[jj]), np.minimum(s.lj_type[ii], s.lj_type[jj])
    tp = hi * (hi + 1) // 2 + lo
    A, B = s.lj_A[tp].astype(np.float64), s.lj_B[tp].astype(np.float64)
    qq = s.charge[ii].astype(np.float64) * s.charge[jj]
    r6 = 1.0 / r2 ** 3
    br = beta * r
    fr = (B - A * r6) * r6 / r2 - qq / (r2 * r) * (br * 2 / np.sqrt(np.pi) * np.exp(-br * br) + erfc(br))
    fl = fr[:, None] * d
    frc = np.zeros((N, 3))
    np.add.at(frc, ii, fl)
    np.add.at(frc, jj, -fl)
    vir = -np.array([(fl[:, 0] * d[:, 0]).sum(), (fl[:, 0] * d[:, 1] + fl[:, 1] * d[:, 0]).sum(),
                     (fl[:, 1] * d[:, 1]).sum(), (fl[:, 0] * d[:, 2] + fl[:, 2] * d[:, 0]).sum(),
                     (fl[:, 1] * d[:, 2] + fl[:, 2] * d[:, 1]).sum(), (fl[:, 2] * d[:, 2]).sum()])
    return dict(frc=frc, lj=float(((A / 12 * r6 - B / 6) * r6).sum()),
                coulomb=float((qq * erfc(br) / r).sum()), virial=vir)


def _rel_rms(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def test_oracle_cell_list_equals_bruteforce(tri):
    s, cell, rcell, beta, c, ser = tri
    cb, sb, total = O.nl_bruteforce(s)
    want = O.pair_array(cb, sb)
    assert total == len(want) and len(want) > 0
    assert np.array_equal(O.pair_array(c, ser), want)
    # the shear changes the set: some pairs are reached only through sheared images
    from sponge_amd import system as S
    s0 = S.config(1)
    c0, ser0, _ = O.nl_build(s0)
    assert not np.array_equal(O.pair_array(c0, ser0), want)


def test_oracle_direct_against_fp64(tri):
    s, cell, rcell, beta, c, ser = tri
    ref = O.lj_direct(s, c, ser, beta)
    r64 = _direct64(s, cell, rcell, beta, c, ser)
    assert _rel_rms(ref["frc"], r64["frc"]) < 1e-5
    assert abs(ref["lj_energy"].sum(dtype=np.float64) - r64["lj"]) < 1e-5 * abs(r64["lj"])
    assert abs(ref["direct_energy"].sum(dtype=np.float64) - r64["coulomb"]) < 1e-5 * abs(r64["coulomb"])
    v = ref["atom_virial"].sum(0, dtype=np.float64)
    assert np.abs(v - r64["virial"]).max() < 1e-5 * np.abs(r64["virial"]).max()


@pytest.mark.gpu
def test_gpu_list_and_pair_kernel_triclinic(dev, tri):
    import torch
    from gpu_helpers import DevSystem, check_total, export_pairs
    from sponge_amd import abi
    s, cell, rcell, beta, c, ser = tri
    print(f"case: {s.name} sheared by {TILT}, {s.n_atoms} atoms")
    d = DevSystem(s, dev)
    box = abi.Box.from_cell(cell, rcell)
    nl = abi.NeighborList(d.N, s.cutoff, s.skin, box)
    nl.update(d.crd, update=abi.NL_FORCED, excl=d.excl)
    got, _, _ = export_pairs(nl, d.N, dev)
    cb, sb, _ = O.nl_bruteforce(s)
    assert np.array_equal(got, O.pair_array(cb, sb))
    n, dims = nl.info()
    assert dims == tuple(O.nl_grid_dims(cell, s.cutoff, s.skin))

    ref = O.lj_direct(s, c, ser, beta)
    r64 = _direct64(s, cell, rcell, beta, c, ser)
    lj = abi.LennardJones(d.N, s.n_types, s.lj_A, s.lj_B, s.lj_type, s.cutoff)
    frc, e, v, edir = d.zeros(d.N, 3), d.zeros(d.N), d.zeros(d.N, 6), d.zeros(d.N)
    lj.direct_force(nl, d.crd, d.charge, frc, box, beta, e, v, edir)
    torch.cuda.synchronize()
    lj_short, _ = lj.step_print()
    f = frc.cpu().numpy()
    assert _rel_rms(f, ref["frc"]) < 1e-5
    assert _rel_rms(f, r64["frc"]) < 1e-5
    check_total("E_LJ", lj_short, ref["lj_energy"].sum(dtype=np.float64), r64["lj"])
    check_total("E_coulomb_direct", edir.cpu().numpy().sum(dtype=np.float64),
                ref["direct_energy"].sum(dtype=np.float64), r64["coulomb"])
    # virial total: the reference-parity bar (1e-6 of the largest component
    # against the reference's FP32 path).  Against FP64 this total sits at the
    # FP32 arithmetic noise of the sheared box: the reference's own FP32
    # restatement is 7.9e-7 away and an FP32 evaluation of its pair formula in
    # numpy 1.4e-6, so the FP64 bound here is 2e-6
    vg = v.cpu().numpy().sum(0, dtype=np.float64)
    v32 = ref["atom_virial"].sum(0, dtype=np.float64)
    scale = float(np.abs(r64["virial"]).max())
    d_ref = float(np.abs(vg - v32).max()) / scale
    d_64 = float(np.abs(vg - r64["virial"]).max()) / scale
    d_32_64 = float(np.abs(v32 - r64["virial"]).max()) / scale
    print(f"virial: |gpu-ref32| {d_ref:.2e}, |gpu-fp64| {d_64:.2e}, |ref32-fp64| {d_32_64:.2e}")
    assert d_ref < 1e-6
    assert d_64 < 2e-6
    # force-only call: the fast arithmetic on the general-cell displacement
    f2 = d.zeros(d.N, 3)
    lj.direct_force(nl, d.crd, d.charge, f2, box, beta)
    torch.cuda.synchronize()
    assert _rel_rms(f2.cpu().numpy(), ref["frc"]) < 1e-5


def _with_cell(s, cell):
    """System s (same coordinates) described with another cell basis."""
    from sponge_amd import system as S
    H = np.array([[cell[0], 0, 0], [cell[1], cell[2], 0], [cell[3], cell[4], cell[5]]], np.float64)
    R = np.linalg.inv(H)
    rcell = np.array([R[0, 0], R[1, 0], R[1, 1], R[2, 0], R[2, 1], R[2, 2]], np.float32)

    class Cell(S.System):
        def cell(self):
            return cell

        def rcell(self):
            return rcell

    return Cell(**{f.name: getattr(s, f.name) for f in dataclasses.fields(s)}), rcell


def test_oracle_pme_basis_invariance():
    """The same orthorhombic lattice with cell vectors (L,0,0), (L,L,0),
    (0,0,L): reciprocal energy, forces and virial agree with the orthorhombic
    description within PME's discretisation error, which shrinks from a 48^3
    to a 64^3 grid as order-4 interpolation does (a transposed or truncated
    rcell would differ at O(1))."""
    from sponge_amd import system as S
    s = S.config(1)
    L = s.box.astype(np.float32)
    beta = O.get_beta(s.cutoff, s.tolerance)
    ortho, _ = _with_cell(s, np.array([L[0], 0, L[1], 0, 0, L[2]], np.float32))
    shear, _ = _with_cell(s, np.array([L[0], L[0], L[1], 0, 0, L[2]], np.float32))
    errs = []
    for n in (48, 64):
        a = O.PME(ortho, (n, n, n), beta).reciprocal()
        b = O.PME(shear, (n, n, n), beta).reciprocal()
        de = abs(a["reciprocal"] - b["reciprocal"]) / abs(a["reciprocal"])
        df = _rel_rms(b["frc"], a["frc"])
        dv = float(np.abs(a["virial"].astype(np.float64) - b["virial"]).max() / np.abs(a["virial"]).max())
        print(f"grid {n}^3: energy {de:.2e}, forces {df:.2e}, virial {dv:.2e}")
        errs.append((de, df, dv))
    assert errs[1][0] < 1e-4 and errs[1][1] < 1e-3 and errs[1][2] < 1e-3
    assert all(errs[1][k] < 0.6 * errs[0][k] for k in range(3))


@pytest.mark.gpu
def test_gpu_pme_reciprocal_triclinic(dev, tri):
    import torch
    from gpu_helpers import DevSystem
    from sponge_amd import abi
    s, cell, rcell, beta, c, ser = tri
    d = DevSystem(s, dev)
    pm = abi.ParticleMesh(d.N, abi.Box.from_cell(cell, rcell), s.cutoff)
    print(f"case: {s.name} sheared by {TILT}, PME grid {pm.fft}")
    # the grid follows the cell vector lengths (the reference's box_length,
    # grid_spacing 1 A by default): |b| and |c| here, not a22 and a33
    c32 = cell.astype(np.float32)
    lb = np.sqrt(c32[1] * c32[1] + c32[2] * c32[2])
    lc = np.sqrt(c32[3] * c32[3] + c32[4] * c32[4] + c32[5] * c32[5])
    assert pm.fft == (O.get_fft_parameter(float(c32[0])), O.get_fft_parameter(float(lb)),
                      O.get_fft_parameter(float(lc)))
    ref = O.PME(s, pm.fft, beta).reciprocal()
    frc = d.zeros(d.N, 3)
    vir = d.zeros(6)
    pot = d.zeros(1)
    pm.reciprocal_force(d.crd, d.charge, frc, virial=vir, potential=pot)
    torch.cuda.synchronize()
    assert _rel_rms(frc.cpu().numpy(), ref["frc"]) < 1e-5
    e = pm.step_print()
    assert abs(e["reciprocal"] - ref["reciprocal"]) < 1e-6 * abs(ref["reciprocal"])
    assert abs(float(pot.cpu()[0]) - ref["potential"]) < 1e-6 * abs(ref["potential"])
    v, rv = vir.cpu().numpy().astype(np.float64), ref["virial"].astype(np.float64)
    dv = float(np.abs(v - rv).max() / np.abs(rv).max())
    print(f"E_reciprocal {e['reciprocal']:.4f} (oracle {ref['reciprocal']:.4f}), virial {dv:.2e}")
    assert dv < 1e-6
