"""GPU parity of the soft-core LJ module (LJ_SOFT_CORE) through the C ABI:
forces, LJ / direct-Coulomb energies, virial and dU/dlambda against the
oracle's restatement at lambda = 0, 0.25, 0.6, 1 with a dummy B state, and
the reference's softcore case (133k atoms, lambda_lj = 0.2) in soft-core mode
against SPONGE's LJ_soft_short / LJ_soft_long within its 2 kcal/mol bar."""
import numpy as np
import pytest

from gpu_helpers import DevSystem, rel_rms
from test_oracle_softcore import _dummy_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def base():
    from oracle import oracle as O
    from sponge_amd import system as S
    s = S.small_water_box(n_side=5, spacing=3.1, cutoff=6.0, n_na=2, n_cl=2, seed=5)
    beta = O.get_beta(s.cutoff, s.tolerance)
    c, ser, flags = O.nl_build(s)
    assert flags == 0
    return s, beta, c, ser


@pytest.mark.parametrize("lam", [0.0, 0.25, 0.6, 1.0])
def test_softcore_against_oracle(dev, base, lam):
    import torch
    from oracle import oracle as O
    from sponge_amd import abi
    s, beta, c, ser = base
    sc = dict(_dummy_state(s, np.array([0, 1, 2, 5, s.n_atoms - 1]), qscale=-0.5), **{"lambda": lam})
    ref = O.ljsc_direct(s, c, ser, beta, sc)
    d = DevSystem(s, dev)
    box = abi.Box(s.box)
    nl = abi.NeighborList(d.N, s.cutoff, s.skin, box)
    nl.update(d.crd, update=abi.NL_FORCED, excl=d.excl)
    m = abi.LJSoftCore(d.N, sc["AA"], sc["AB"], sc["BA"], sc["BB"], sc["typeA"], sc["typeB"],
                       s.cutoff, alpha=sc["alpha"], p=sc["p"], sigma=3.0, sigma_min=0.0,
                       **{"lambda": lam})
    frc, e, v, edir = d.zeros(d.N, 3), d.zeros(d.N), d.zeros(d.N, 6), d.zeros(d.N)
    torch.cuda.synchronize()
    m.direct_force(nl, d.crd, d.charge, frc, box, beta, e, v, edir)
    lj_sum = m.step_print()
    qba = torch.tensor(sc["charge_BA"], device=dev)
    du_lj, du_cf = m.du_dlambda(nl, d.crd, d.charge, box, beta, qba)
    torch.cuda.synchronize()
    _, lr_ti = m.long_range_factors()
    volume = float(np.float32(s.box[0] * s.box[1]) * s.box[2])
    e_lj_ref = float(ref["lj_energy"].astype(np.float64).sum())
    e_cf_ref = float(ref["direct_energy"].astype(np.float64).sum())
    print(f"case: lambda {lam} LJ {lj_sum:.5f} (oracle {e_lj_ref:.5f}) dU/dl LJ {du_lj:.5f} "
          f"(oracle {ref['du_lj'] + lr_ti / volume:.5f}) direct {du_cf:.5f} (oracle "
          f"{ref['du_direct']:.5f})")
    # energy totals are sums of FP32 terms of both signs: the bound scales
    # with the sum of the per-atom magnitudes
    scale = lambda a: float(np.abs(np.asarray(a, np.float64)).sum())
    assert rel_rms(frc.cpu().numpy(), ref["frc"]) < 1e-5
    assert abs(lj_sum - e_lj_ref) <= 2e-6 * scale(ref["lj_energy"])
    assert abs(float(edir.cpu().numpy().astype(np.float64).sum()) - e_cf_ref) <= \
        2e-6 * scale(ref["direct_energy"])
    assert abs(e.cpu().numpy().astype(np.float64).sum() -
               ref["atom_energy"].astype(np.float64).sum()) <= 2e-6 * scale(ref["atom_energy"])
    assert np.allclose(v.cpu().numpy().astype(np.float64).sum(0),
                       ref["atom_virial"].astype(np.float64).sum(0), rtol=1e-5, atol=1e-3)
    assert abs(du_lj - (ref["du_lj"] + lr_ti / volume)) <= 1e-5 * max(1.0, abs(du_lj))
    assert abs(du_cf - ref["du_direct"]) <= 1e-5 * max(1.0, abs(du_cf))


def test_softcore_reference_case(dev):
    import torch
    from sponge_amd import abi
    from test_softcore_fixture import reference, softcore_system
    s = softcore_system()
    ref = reference()
    d = DevSystem(s, dev)
    box = abi.Box(s.box)
    nl = abi.NeighborList(d.N, s.cutoff, s.skin, box)
    nl.update(d.crd, update=abi.NL_FORCED, excl=d.excl)
    m = abi.LJSoftCore(d.N, s.lj_A, s.lj_B, s.lj_A, s.lj_B, s.lj_type, s.lj_type, s.cutoff,
                       **{"lambda": ref["lambda_lj"]})
    frc, e = d.zeros(d.N, 3), d.zeros(d.N)
    m.direct_force(nl, d.crd, d.charge, frc, box, 0.0 + abi.load().sponge_amd_get_beta(s.cutoff, s.tolerance), e)
    short = m.step_print()
    factor, _ = m.long_range_factors()
    long_ = factor / box.volume
    torch.cuda.synchronize()
    print(f"case: softcore 133k LJ_soft_short {short:.3f} (SPONGE {ref['energies']['LJ_soft_short']}), "
          f"LJ_soft_long {long_:.3f} (SPONGE {ref['energies']['LJ_soft_long']})")
    assert abs(short - ref["energies"]["LJ_soft_short"]) <= ref["abs_tolerance_kcal_mol"]
    assert abs(long_ - ref["energies"]["LJ_soft_long"]) <= ref["abs_tolerance_kcal_mol"]
