"""The reference's softcore point-energy case as a 133,266-atom known-answer
test of the plain path: no atom is perturbed there, so SPONGE's LJ_soft_short
and LJ_soft_long are the plain LJ short-range energy and long-range
correction, and its PM terms plain PME (cutoff 8 A, box 126.8 x 105.0 x
100.7 A, automatic ~1 A grid).  Every term within the reference test's
absolute tolerance of 2 kcal/mol (benchmarks/validation/misc/tests/
test_softcore.py), for the oracle (CPU) and the GPU path (which must also
match the oracle to 2e-6 relative)."""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def reference():
    with open(os.path.join(GOLD, "softcore_133k.json")) as f:
        return json.load(f)


def softcore_system():
    from oracle import oracle as O
    from sponge_amd.system import System
    z = np.load(os.path.join(GOLD, "softcore_133k.npz"))
    assert np.allclose(z["angles"], 90.0)
    crd = z["crd"]
    s = System(name="softcore_133k", box=z["box"].copy(), crd=crd, vel=np.zeros_like(crd),
               charge=z["charge"], mass=np.ones(len(crd), np.float32), lj_type=z["lj_type"],
               lj_A=z["lj_A"], lj_B=z["lj_B"], excl_start=z["excl_start"], excl_num=z["excl_num"],
               excl_list=z["excl_list"], n_waters=0, cutoff=reference()["cutoff"], skin=2.0)
    s.fft = tuple(O.get_fft_parameter(float(L) / 1.0) for L in s.box)
    return s


def oracle_terms(s):
    from oracle import oracle as O
    beta = O.get_beta(s.cutoff, s.tolerance)
    c, ser, flags = O.nl_build(s)
    assert flags == 0
    d = O.lj_direct(s, c, ser, beta, need_energy=True, need_virial=False)
    ex = O.pme_excluded(s, beta, need_virial=False)
    rec = O.PME(s, s.fft, beta).reciprocal(need_virial=False)
    volume = float(np.float32(s.box[0] * s.box[1]) * s.box[2])
    t = dict(LJ_soft_short=float(d["lj_energy"].astype(np.float64).sum()),
             LJ_soft_long=float(O.lj_long_range_factor(s)) / volume,
             PM_direct=float(d["direct_energy"].astype(np.float64).sum()),
             PM_reciprocal=rec["reciprocal"], PM_self=rec["self_energy"],
             PM_correction=float(ex["correction_energy"].astype(np.float64).sum()))
    t["PM"] = t["PM_direct"] + t["PM_reciprocal"] + t["PM_self"] + t["PM_correction"]
    return t


@pytest.fixture(scope="module")
def fixture():
    s = softcore_system()
    return s, oracle_terms(s)


def test_oracle_against_reference_energies(fixture):
    s, t = fixture
    ref = reference()
    assert s.n_atoms == 133266 and s.fft == (128, 108, 108)
    for k, v in ref["energies"].items():
        print(f"{k}: oracle {t[k]:.3f} reference {v:.3f} d {t[k] - v:+.3f}")
    for k, v in ref["energies"].items():
        assert abs(t[k] - v) <= ref["abs_tolerance_kcal_mol"], k


@pytest.mark.gpu
def test_gpu_against_reference_energies(dev, fixture):
    from sponge_amd import abi
    s, t = fixture
    ref = reference()
    md = abi.MDSystem(s)
    md.run(1, energy_interval=1)
    e = md.energies()
    got = dict(LJ_soft_short=e["LJ_short"], LJ_soft_long=e["LJ_long"], PM=e["PM"],
               PM_direct=e["PM_direct"], PM_reciprocal=e["PM_reciprocal"], PM_self=e["PM_self"],
               PM_correction=e["PM_correction"])
    for k, v in ref["energies"].items():
        print(f"case: {k} GPU {got[k]:.3f} oracle {t[k]:.3f} reference {v:.3f}")
    for k, v in ref["energies"].items():
        assert abs(got[k] - v) <= ref["abs_tolerance_kcal_mol"], k
        assert abs(got[k] - t[k]) <= max(2e-6 * abs(t[k]), 0.02), k
