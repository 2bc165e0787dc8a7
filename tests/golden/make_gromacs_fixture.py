"""Copies the reference's GROMACS comparison fixture for alanine dipeptide in
CHARMM TIP3P water (data files its own comparison test reads) into
tests/golden/gromacs_alanine_dipeptide/:
  benchmarks/comparison/reference/gromacs/sponge_inputs/alanine_dipeptide_charmm_tip3p/
      iter{0,1,2}/sys_flexible_coordinate.txt      (per iteration)
      iter0/sys_flexible_{charge,LJ,exclude,mass}.txt (identical in every iteration)
  benchmarks/comparison/tests/gromacs/statics/alanine_dipeptide_charmm_tip3p/sponge.mdin
  benchmarks/comparison/reference/gromacs/reference.json -> energies.json (the
      nonbonded terms and the perturbation of each iteration)
and the tolerances of the reference's comparison
(benchmarks/comparison/tests/gromacs/tests/comp_alanine_dipeptide.py:106-115).
Run in a container that has the reference at /root/reference."""
import json
import os
import shutil

SRC = "/root/reference/benchmarks/comparison"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gromacs_alanine_dipeptide")
CASE = "alanine_dipeptide_charmm_tip3p"


def main():
    inputs = os.path.join(SRC, "reference/gromacs/sponge_inputs", CASE)
    os.makedirs(DST, exist_ok=True)
    for kind in ("charge", "LJ", "exclude", "mass"):
        shutil.copyfile(os.path.join(inputs, "iter0", f"sys_flexible_{kind}.txt"),
                        os.path.join(DST, f"sys_flexible_{kind}.txt"))
    for it in range(3):
        shutil.copyfile(os.path.join(inputs, f"iter{it}", "sys_flexible_coordinate.txt"),
                        os.path.join(DST, f"iter{it}_coordinate.txt"))
    shutil.copyfile(os.path.join(SRC, "tests/gromacs/statics", CASE, "sponge.mdin"),
                    os.path.join(DST, "sponge.mdin"))
    with open(os.path.join(SRC, "reference/gromacs/reference.json")) as f:
        ref = json.load(f)
    keep = ("iteration", "perturbation", "natom", "lj_sr", "coulomb_sr", "coulomb_recip")
    out = {"source": "GROMACS " + ref["gromacs_version"], "energy_unit": ref["energy_unit"],
           "tolerance_kcal_mol": {"lj_short_vs_lj_sr": 25.0, "pm_vs_coulomb_sr_plus_recip": 2.0},
           "entries": [{k: e[k] for k in keep} for e in ref["entries"] if e["case_name"] == CASE]}
    with open(os.path.join(DST, "energies.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
