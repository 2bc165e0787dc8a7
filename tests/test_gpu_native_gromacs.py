"""The reference's GROMACS known-answer fixture (alanine dipeptide in CHARMM
TIP3P water, 2129 atoms, cutoff 12 A) through the GPU path: one energy step
of the NVE driver from the native input files, LJ_short and PM against the
oracle (FP32 sums: 2e-6 relative) and against GROMACS with the reference
comparison's own tolerances (25 and 2 kcal/mol)."""
import pytest

from test_native_gromacs import energies, fixture_system, oracle_terms

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("it", [0, 1, 2])
def test_gpu_against_gromacs_and_oracle(dev, it):
    from sponge_amd import abi
    s = fixture_system(it)
    md = abi.MDSystem(s)
    md.run(1, energy_interval=1)
    e = md.energies()
    lj_o, pm_o = oracle_terms(s)
    ref = energies()
    g = [x for x in ref["entries"] if x["iteration"] == it][0]
    tol = ref["tolerance_kcal_mol"]
    print(f"case: iter{it} GPU LJ_short {e['LJ_short']:.4f} (oracle {lj_o:.4f}, GROMACS "
          f"{g['lj_sr']:.4f}); GPU PM {e['PM']:.4f} (oracle {pm_o:.4f}, GROMACS "
          f"{g['coulomb_sr'] + g['coulomb_recip']:.4f})")
    assert abs(e["LJ_short"] - lj_o) <= 2e-6 * abs(lj_o)
    assert abs(e["PM"] - pm_o) <= 2e-6 * abs(pm_o)
    assert abs(e["LJ_short"] - g["lj_sr"]) <= tol["lj_short_vs_lj_sr"]
    assert abs(e["PM"] - (g["coulomb_sr"] + g["coulomb_recip"])) <= tol["pm_vs_coulomb_sr_plus_recip"]
