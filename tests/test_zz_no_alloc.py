"""The allocation-free direct-force entry point
(sponge_lj_pme_direct_force_list_noalloc) against the handle-based
sponge_lj_pme_direct_force_list on the same reference-format half lists:
two calls in a row on one GPU (the second on a smaller system, so stale
contents of the shared working array would show), forces, per-atom energies
and virials; and the size limit."""
import numpy as np
import pytest

from gpu_helpers import DevSystem, rel_rms

pytestmark = pytest.mark.gpu


def _handle_and_noalloc(dev, s):
    import torch
    from oracle import oracle as O
    from sponge_amd import abi
    d = DevSystem(s, dev)
    N = d.N
    beta = O.get_beta(s.cutoff, s.tolerance)
    c, ser, flags = O.nl_build(s)
    assert flags == 0
    M = ser.shape[1]
    counts = torch.tensor(np.ascontiguousarray(c, np.int32), device=dev)
    serial = torch.tensor(np.ascontiguousarray(ser, np.int32).reshape(-1), device=dev)
    box = abi.Box(s.box)
    out = []
    # the module entry point
    lj = abi.LennardJones(N, s.n_types, s.lj_A, s.lj_B, s.lj_type, s.cutoff)
    f, e, v, de = d.zeros(N, 3), d.zeros(N), d.zeros(N, 6), d.zeros(N)
    torch.cuda.synchronize()
    lj.direct_force_list(counts, serial, M, d.crd, d.charge, f, box, beta, atom_energy=e,
                         atom_virial=v, direct_energy=de)
    torch.cuda.synchronize()
    ljs, _ = lj.step_print()
    out.append(dict(frc=f.cpu().numpy(), e=e.cpu().numpy(), v=v.cpu().numpy(),
                    de=de.cpu().numpy(), lj_short=ljs))
    lj.clear()
    # the allocation-free entry point on the caller's device tables
    A = torch.tensor(s.lj_A, device=dev)
    B = torch.tensor(s.lj_B, device=dev)
    T = torch.tensor(np.ascontiguousarray(s.lj_type, np.int32), device=dev)
    f2, e2, v2, de2, le2 = d.zeros(N, 3), d.zeros(N), d.zeros(N, 6), d.zeros(N), d.zeros(N)
    torch.cuda.synchronize()
    abi.lj_direct_force_list_noalloc(s.n_types, A, B, T, s.cutoff, counts, serial, M, d.crd,
                                     d.charge, f2, box, beta, atom_energy=e2, atom_virial=v2,
                                     lj_energy=le2, direct_energy=de2)
    torch.cuda.synchronize()
    out.append(dict(frc=f2.cpu().numpy(), e=e2.cpu().numpy(), v=v2.cpu().numpy(),
                    de=de2.cpu().numpy(), lj_short=float(le2.double().sum().cpu())))
    return out


def _compare(name, ref, got):
    ef = rel_rms(got["frc"], ref["frc"])
    ee = rel_rms(got["e"], ref["e"])
    ev = rel_rms(got["v"], ref["v"])
    ed = rel_rms(got["de"], ref["de"])
    el = abs(got["lj_short"] - ref["lj_short"]) / abs(ref["lj_short"])
    print(f"case: {name}: force {ef:.2e}, atom energy {ee:.2e}, virial {ev:.2e}, "
          f"direct energy {ed:.2e}, LJ_short {el:.2e}")
    # the same kernels: only the order of the float atomics differs
    assert ef < 1e-6 and ee < 1e-6 and ev < 1e-6 and ed < 1e-6 and el < 1e-6


def test_noalloc_matches_module_entry_point(dev):
    from sponge_amd import system as S
    big = S.config(1)                      # 1,944 atoms
    small = S.water_box(400, cutoff=8.0)   # 1,200 atoms
    ref, got = _handle_and_noalloc(dev, big)
    _compare("1944 atoms", ref, got)
    ref, got = _handle_and_noalloc(dev, small)
    _compare("1200 atoms after 1944", ref, got)


def test_noalloc_rejects_sizes_beyond_the_limit(dev):
    import torch
    from sponge_amd import abi
    from sponge_amd import system as S
    s = S.water_box(400, cutoff=8.0)
    n = abi.NOALLOC_MAX_ATOMS + 1
    crd = torch.zeros(n, 3, device=dev)
    one = torch.zeros(4, device=dev)
    ione = torch.zeros(4, dtype=torch.int32, device=dev)
    with pytest.raises(abi.SpongeError) as err:
        abi.lj_direct_force_list_noalloc(1, one, one, ione, 8.0, ione, ione, 1, crd, one, one,
                                         abi.Box(s.box), 0.3)
    assert err.value.code == 1009  # SPONGE_AMD_ERROR_OVERFLOW
    with pytest.raises(abi.SpongeError) as err:
        small = torch.zeros(8, 3, device=dev)
        abi.lj_direct_force_list_noalloc(abi.NOALLOC_MAX_TYPES + 1, one, one, ione, 8.0, ione,
                                         ione, 1, small, one, one, abi.Box(s.box), 0.3)
    assert err.value.code == 1009
