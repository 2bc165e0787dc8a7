"""The reference's own known-answer fixture for this path: alanine dipeptide
in CHARMM TIP3P water (2129 atoms, cutoff 12 A) from its GROMACS comparison
(benchmarks/comparison/tests/gromacs/tests/comp_alanine_dipeptide.py), with
SPONGE's LJ_short compared to GROMACS' lj_sr (tolerance 25 kcal/mol) and
SPONGE's PM (direct + reciprocal + self + exclusion correction) to
coulomb_sr + coulomb_recip (tolerance 2 kcal/mol), the reference test's own
bars.  Here: the native-format readers and the oracle (CPU)."""
import json
import os

import numpy as np
import pytest

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gromacs_alanine_dipeptide")


def energies():
    with open(os.path.join(FIX, "energies.json")) as f:
        return json.load(f)


def fixture_system(it):
    from oracle import oracle as O
    from sponge_amd import native
    mdin = native.read_mdin(os.path.join(FIX, "sponge.mdin"))
    s = native.load_system(FIX, cutoff=float(mdin["cutoff"]), name=f"alanine_iter{it}",
                           coordinate=os.path.join(FIX, f"iter{it}_coordinate.txt"))
    # PM grid from the box and the default 1 A spacing (PM_force.cpp:383-390)
    s.fft = tuple(O.get_fft_parameter(float(L) / 1.0) for L in s.box)
    return s


def test_native_readers():
    from sponge_amd import native
    s = fixture_system(0)
    assert s.n_atoms == 2129 and s.n_types == 10
    assert np.allclose(s.box, 70.0)
    assert int(s.excl_num.sum()) == 2216 and len(s.excl_list) == 2216
    assert abs(float(s.charge.astype(np.float64).sum())) < 1e-3  # a neutral system
    # A and B were scaled by 12 and 6 on load, as Native_Load_LJ does
    t = native._tokens(os.path.join(FIX, "sys_flexible_LJ.txt"))
    assert s.lj_A[0] == np.float32(float(t[2])) * np.float32(12.0)
    mdin = native.read_mdin(os.path.join(FIX, "sponge.mdin"))
    assert mdin["cutoff"] == "12.0" and mdin["mode"] == "nve"
    # every listed exclusion is j > i and ascending, as the list filter assumes
    for i in range(s.n_atoms):
        row = s.excl_list[s.excl_start[i]:s.excl_start[i] + s.excl_num[i]]
        assert np.all(row > i) and np.all(np.diff(row) > 0)


def oracle_terms(s):
    from oracle import oracle as O
    beta = O.get_beta(s.cutoff, s.tolerance)
    c, ser, flags = O.nl_build(s)
    assert flags == 0
    d = O.lj_direct(s, c, ser, beta, need_energy=True, need_virial=False)
    ex = O.pme_excluded(s, beta, need_virial=False)
    rec = O.PME(s, s.fft, beta).reciprocal(need_virial=False)
    lj = float(d["lj_energy"].astype(np.float64).sum())
    pm = (float(d["direct_energy"].astype(np.float64).sum()) + rec["reciprocal"] +
          rec["self_energy"] + float(ex["correction_energy"].astype(np.float64).sum()))
    return lj, pm


@pytest.mark.parametrize("it", [0, 1, 2])
def test_oracle_against_gromacs(it):
    s = fixture_system(it)
    ref = energies()
    e = [x for x in ref["entries"] if x["iteration"] == it][0]
    tol = ref["tolerance_kcal_mol"]
    lj, pm = oracle_terms(s)
    d_lj = lj - e["lj_sr"]
    d_pm = pm - (e["coulomb_sr"] + e["coulomb_recip"])
    print(f"iter{it}: LJ_short {lj:.3f} (GROMACS lj_sr {e['lj_sr']:.3f}, d {d_lj:+.3f}); "
          f"PM {pm:.3f} (coulomb_sr + recip {e['coulomb_sr'] + e['coulomb_recip']:.3f}, "
          f"d {d_pm:+.3f})")
    assert abs(d_lj) <= tol["lj_short_vs_lj_sr"]
    assert abs(d_pm) <= tol["pm_vs_coulomb_sr_plus_recip"]


def test_output_writers(tmp_path):
    from sponge_amd import native
    frc = np.arange(12, dtype=np.float32).reshape(4, 3)
    p = tmp_path / "frc.dat"
    native.append_frc_frame(p, frc)
    native.append_frc_frame(p, 2 * frc)
    raw = np.fromfile(p, dtype=np.float32)
    # the reference's extractor takes the last frame of width 3N
    assert raw.size == 24 and np.array_equal(raw[-12:].reshape(4, 3), 2 * frc)
    m = tmp_path / "mdout.txt"
    w = native.MdoutWriter(m, ["step", "potential", "LJ", "PM"])
    w.write({"step": 0, "potential": -3976.7394, "LJ": 1237.96, "PM": -7437.5})
    w.write({"step": 10, "potential": -3970.1, "LJ": 1230.0, "PM": -7430.25})
    lines = open(m).read().splitlines()
    assert lines[0] == "%15s %15s %15s %15s" % ("step", "potential", "LJ", "PM")
    assert lines[1] == "%15s %15s %15s %15s" % ("0", "-3976.74", "1237.96", "-7437.50")
    back = native.read_mdout(m)
    assert back["step"] == [0.0, 10.0] and back["PM"] == [-7437.5, -7430.25]
