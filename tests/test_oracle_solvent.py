"""Pins the oracle's solvent_LJ restatement (SPONGE/Lennard_Jones_force/
solvent_LJ.cpp:24-147): on a list fresh from a build, the plain rows below
the tail plus the solvent rows give exactly the plain kernel's pairs; on a
stale list the pair set is the reference's (each water's atoms against its
oxygen's half list), checked against an FP64 sum over that set spelled out
in numpy, and it differs from the plain kernel's."""
import numpy as np
import pytest

from oracle import oracle as O
from sponge_amd import system as S


@pytest.fixture(scope="module")
def tail_box():
    base = S.build_water_ion_box("tail", (8, 8, 8), 25.0, n_sites=500, n_na=10, n_cl=10,
                                 cutoff=8.0, seed=11)
    return S.solvent_tail(base)


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def _split(s, c, ser, beta, crd=None):
    n_solv = 3 * s.n_waters
    g = O.lj_direct(s, c, ser, beta, crd=crd, solvent_n=n_solv)
    w = O.solvent_lj(s, c, ser, beta, s.water_start, 3, crd=crd)
    return {k: g[k] + w[k] for k in g}, w


def test_solvent_tail_layout(tail_box):
    s = tail_box
    assert s.water_start == 20 and s.n_atoms == 20 + 3 * s.n_waters
    rs = s.residue_starts()
    assert len(rs) == 20 + s.n_waters and rs[20] == 20 and rs[21] == 23
    # the water exclusions moved with the atoms: O -> H1, H2
    o = s.water_start
    a = s.excl_start[o]
    assert s.excl_num[o] == 2 and list(s.excl_list[a:a + 2]) == [o + 1, o + 2]
    assert abs(float(s.mass[o]) - 16.0) < 1e-6 and abs(float(s.mass[o + 1]) - 1.008) < 1e-6


def test_fresh_list_equals_plain_kernel(tail_box):
    s = tail_box
    beta = O.get_beta(s.cutoff, s.tolerance)
    c, ser, flags = O.nl_build(s)
    assert flags == 0
    plain = O.lj_direct(s, c, ser, beta)
    both, solv = _split(s, c, ser, beta)
    r64 = O.direct64_list(s, beta, c, ser)
    assert _rel(both["frc"], plain["frc"]) < 1e-6
    # totals: the split sums the same pairs in another FP32 order, so it is
    # as close to the FP64 sum as the plain kernel is
    pairs = [("lj_energy", "lj"), ("direct_energy", "coulomb")]
    for k, k64 in pairs:
        d_both = abs(both[k].sum(dtype=np.float64) - r64[k64])
        d_plain = abs(plain[k].sum(dtype=np.float64) - r64[k64])
        assert d_both <= max(1e-6 * abs(r64[k64]), 2 * d_plain), (k, d_both, d_plain)
    d_both = np.abs(both["atom_virial"].sum(0, dtype=np.float64) - r64["virial"]).max()
    d_plain = np.abs(plain["atom_virial"].sum(0, dtype=np.float64) - r64["virial"]).max()
    assert d_both <= max(1e-6 * np.abs(r64["virial"]).max(), 2 * d_plain)
    # the solvent rows put a water's energies and virial on its oxygen
    w = np.arange(s.water_start, s.n_atoms).reshape(-1, 3)
    assert not solv["atom_energy"][w[:, 1:]].any() and not solv["atom_virial"][w[:, 1:]].any()
    assert np.count_nonzero(solv["atom_energy"][w[:, 0]]) > 0.9 * len(w)
    # and nothing on the atoms below the tail
    assert not solv["atom_energy"][:s.water_start].any()


def _solvent_pair_set(s, c, ser):
    """Half list of the reference's pair set: rows below the tail keep their
    own partners; every atom of a water takes its oxygen's partners."""
    N, w0 = s.n_atoms, s.water_start
    rows = []
    for i in range(N):
        if i < w0:
            rows.append(ser[i, :c[i]])
        else:
            o = w0 + 3 * ((i - w0) // 3)
            rows.append(ser[o, :c[o]])
    m = max(len(r) for r in rows)
    counts = np.array([len(r) for r in rows], np.int32)
    serial = np.zeros((N, m), np.int32)
    for i, r in enumerate(rows):
        serial[i, :len(r)] = r
    return counts, serial


def test_stale_list_is_the_reference_pair_set(tail_box):
    # a skin below the O-H distance (the list radius 8.5 A): an atom within
    # the cutoff of a hydrogen can lie beyond its oxygen's list, so the
    # reference's set differs from the plain one even on a fresh list; the
    # atoms then move by ~0.3 A without a rebuild
    import dataclasses
    s = dataclasses.replace(tail_box, skin=0.5)
    beta = O.get_beta(s.cutoff, s.tolerance)
    c, ser, _ = O.nl_build(s)
    rng = np.random.default_rng(5)
    moved = (s.crd + rng.normal(scale=0.2, size=s.crd.shape)).astype(np.float32)
    both, _ = _split(s, c, ser, beta, crd=moved)
    cs, ss = _solvent_pair_set(s, c, ser)
    ref64 = O.direct64_list(s, beta, cs, ss, crd=moved)
    plain = O.lj_direct(s, c, ser, beta, crd=moved)
    e_set = _rel(both["frc"], ref64["frc"])
    e_plain = _rel(both["frc"], plain["frc"])
    print(f"solvent split vs FP64 over the reference pair set {e_set:.2e}; vs plain {e_plain:.2e}")
    assert e_set < 1e-5
    # the hydrogens' own lists differ from their oxygen's here
    assert e_plain > 10 * e_set
    tot = both["lj_energy"].sum(dtype=np.float64) + both["direct_energy"].sum(dtype=np.float64)
    assert abs(tot - (ref64["lj"] + ref64["coulomb"])) < 1e-5 * abs(ref64["lj"] + ref64["coulomb"])
