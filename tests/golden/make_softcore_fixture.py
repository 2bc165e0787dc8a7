"""Packs the reference's softcore point-energy fixture
(benchmarks/validation/misc/statics/softcore/: 133,266 atoms, cutoff 8 A,
orthorhombic box; its SPONGE baseline energies in
reference/energy_reference_np1.json, compared by
benchmarks/validation/misc/tests/test_softcore.py with an absolute tolerance
of 2 kcal/mol per term) into tests/golden/softcore_133k.npz + .json.  No atom
of that system is perturbed (identical A and B LJ types and tables), so its
LJ_soft terms are the plain LJ short-range energy and long-range correction,
and its PM terms are plain PME.  Stored: coordinates and box (as the
coordinate file holds them), charges, A-state LJ types and the A / B
coefficient tables scaled on load by 12 and 6 (Native_Load_LJ_Soft_Core,
SPONGE/xponge/load/native/lj_soft.hpp), the exclusion CSR.  Run in a
container that has the reference at /root/reference."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from sponge_amd import native  # noqa: E402

SRC = "/root/reference/benchmarks/validation/misc/statics/softcore"


def main():
    crd, box, angles = native.read_coordinate(os.path.join(SRC, "softcore_coordinate.txt"))
    charge = native.read_per_atom(os.path.join(SRC, "softcore_charge.txt"))
    es, en, el = native.read_exclude(os.path.join(SRC, "softcore_exclude.txt"))
    t = native._tokens(os.path.join(SRC, "softcore_LJ_soft_core.txt"))
    n, ta, tb = int(t[0]), int(t[1]), int(t[2])
    pa, pb = ta * (ta + 1) // 2, tb * (tb + 1) // 2
    k = 3
    AA = np.array(t[k:k + pa], np.float32) * np.float32(12.0)
    k += pa
    AB = np.array(t[k:k + pa], np.float32) * np.float32(6.0)
    k += pa
    BA = np.array(t[k:k + pb], np.float32) * np.float32(12.0)
    k += pb
    BB = np.array(t[k:k + pb], np.float32) * np.float32(6.0)
    k += pb
    types = np.array(t[k:k + 2 * n], np.int32).reshape(n, 2)
    assert np.array_equal(types[:, 0], types[:, 1]) and np.array_equal(AA, BA) and np.array_equal(AB, BB)
    np.savez_compressed(os.path.join(HERE, "softcore_133k.npz"), crd=crd, box=box, angles=angles,
                        charge=charge, lj_type=types[:, 0], lj_A=AA, lj_B=AB, excl_start=es,
                        excl_num=en, excl_list=el)
    with open(os.path.join(SRC, "reference", "energy_reference_np1.json")) as f:
        ref = json.load(f)
    mdin = native.read_mdin(os.path.join(SRC, "mdin.spg.in"))
    out = {"source": "SPONGE single-process baseline of the softcore validation case",
           "cutoff": float(mdin["cutoff"]), "lambda_lj": float(mdin["lambda_lj"]),
           "abs_tolerance_kcal_mol": 2.0,
           "energies": {k: ref["energies"][k] for k in
                        ("LJ_soft_short", "LJ_soft_long", "PM", "PM_direct", "PM_reciprocal",
                         "PM_self", "PM_correction")}}
    with open(os.path.join(HERE, "softcore_133k.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
