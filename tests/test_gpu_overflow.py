"""Neighbor-list overflow inside the NVE step driver (the reference's
NEIGHBOR_LIST::Check_Overflow, SPONGE/neighbor_list/neighbor_list.cpp:
1027-1103, every check_overflow_interval steps): with
throw_error_when_overflow the run stops with SPONGE's Overflow error (1009);
by default the list is re-initialised with 1.1x the limit that overflowed
until it holds every pair, and the forces are then complete."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle_force(s, crd):
    from oracle import oracle as O
    beta = O.get_beta(s.cutoff, s.tolerance)
    c, ser, _ = O.nl_build(s, crd=crd)
    f = O.lj_direct(s, c, ser, beta, crd=crd, need_energy=False,
                    need_virial=False)["frc"].astype(np.float64)
    f += O.pme_excluded(s, beta, crd=crd, need_energy=False, need_virial=False)["frc"]
    f += O.PME(s, s.fft, beta).reciprocal(crd=crd, need_energy=False, need_virial=False)["frc"]
    return f


def test_overflow_raises_when_asked(dev):
    from sponge_amd import abi
    from sponge_amd import system as S
    s = S.config(1)  # ~420 partners per atom in both directions
    md = abi.MDSystem(s, nl_opts={"max_neighbor_numbers": 120, "check_overflow_interval": 5,
                                  "throw_error_when_overflow": 1})
    md.run(4)  # before the first check
    with pytest.raises(abi.SpongeError) as err:
        md.run(1)
    assert err.value.code == 1009


@pytest.mark.parametrize("use_graph", [False, True])
def test_overflow_grows_the_list(dev, use_graph):
    from sponge_amd import abi
    from sponge_amd import system as S
    s = S.config(1)
    md = abi.MDSystem(s, nl_opts={"max_neighbor_numbers": 120, "check_overflow_interval": 5})
    md.run(150, use_graph=use_graph)
    grown = md.nl_options()["max_neighbor_numbers"]
    print(f"case: graph={use_graph}: max_neighbor_numbers 120 -> {grown}")
    assert grown > 120
    # one more step on the grown list: the force of that step is complete
    crd, _, _ = md.state()
    md.run(1)
    _, _, frc = md.state()
    ref = _oracle_force(s, crd)
    err = float(np.sqrt(((frc - ref) ** 2).sum() / (ref ** 2).sum()))
    print(f"case: graph={use_graph}: force rel RMS after growth {err:.2e}")
    assert err < 1e-5
