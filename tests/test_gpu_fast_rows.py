"""The step driver's force-only pair rows at every length class: dilute water
boxes (L = 50 A, rc = 10 A, so the pair kernel stages partners at their images
nearest the cell centre) whose rows hold ~35, ~160, ~260 and ~400 list
entries -- only a partial pass, one full pass of 128 entries, two full passes
and three full passes before the partial one.  The force of the first step
(list built from the step's coordinates) is the oracle's force of those
coordinates (LJ + PME direct with the reference arithmetic, the excluded-pair
correction and PME) within 1e-5 relative RMS."""
import numpy as np
import pytest

from gpu_helpers import rel_rms

pytestmark = pytest.mark.gpu

L_BOX = 50.0


@pytest.mark.parametrize("n_waters", [200, 900, 1500, 2300])
def test_fast_rows_against_oracle(dev, n_waters):
    from oracle import oracle as O
    from sponge_amd import abi
    from sponge_amd import system as S
    m = int(np.ceil(n_waters ** (1.0 / 3.0))) + 2
    s = S.build_water_ion_box(f"dilute{n_waters}", (m, m, m), L_BOX, n_sites=n_waters,
                              cutoff=10.0, skin=2.0, fft=None, seed=20260311 + n_waters)
    density = s.n_atoms / L_BOX ** 3
    rows = density * 4.0 / 3.0 * np.pi * (s.cutoff + s.skin) ** 3
    print(f"case: {n_waters} waters in {L_BOX} A, ~{rows:.0f} entries per row", flush=True)
    md = abi.MDSystem(s)
    crd0, _, _ = md.state()
    md.run(1)
    _, _, f = md.state()
    beta = O.get_beta(s.cutoff, s.tolerance)
    c, ser, flags = O.nl_build(s, crd=crd0)
    assert flags == 0
    ref = O.lj_direct(s, c, ser, beta, crd=crd0, need_energy=False,
                      need_virial=False)["frc"].astype(np.float64)
    ref += O.pme_excluded(s, beta, crd=crd0, need_energy=False, need_virial=False)["frc"]
    ref += O.PME(s, md.pme_grid(), beta).reciprocal(crd=crd0, need_energy=False,
                                                     need_virial=False)["frc"]
    e = rel_rms(f, ref)
    print(f"case: {n_waters} waters: force rel RMS {e:.2e}", flush=True)
    assert e < 1e-5
    md.clear()
