"""GPU tests of the batched driver (one dispatch per kernel over every
system) and the FFT autotuner: the 48-box batched workload against the oracle
(every box's neighbor-list set bit for bit, forces within 1e-5); every system
of a batch matches its own single-system run; graph launches of several steps
follow the single runs; autotuned FFT tilings leave the reciprocal forces
unchanged."""
import ctypes as C

import numpy as np
import pytest

from gpu_helpers import DevSystem, rel_rms

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def boxes():
    from sponge_amd import system as S
    return [S.water_box(n, seed=40 + n) for n in (500, 777, 1300)]


def _half_list(md, N, dev, M=1200):
    import torch
    from oracle import oracle as O
    from sponge_amd import abi
    nl_h, _, _ = md.handles()
    counts = torch.zeros(N, dtype=torch.int32, device=dev)
    serial = torch.zeros(N * M, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    rc = abi.load().sponge_nl_export_half_list(nl_h, abi._ptr(counts), abi._ptr(serial), M)
    assert rc == 0
    torch.cuda.synchronize()
    return O.pair_array(counts.cpu().numpy(), serial.view(N, M).cpu().numpy())


def test_batch_matches_single_runs(dev, boxes):
    from sponge_amd import abi
    singles = [abi.MDSystem(s) for s in boxes]
    batched = [abi.MDSystem(s) for s in boxes]
    b = abi.Batch(batched)
    # first step: the lists are built (eagerly), forces of the initial state
    for m in singles:
        m.run(1)
    b.run(1)
    b.synchronize()
    for s, m1, m2 in zip(boxes, singles, batched):
        c1, _, f1 = m1.state()
        c2, _, f2 = m2.state()
        e = rel_rms(f2, f1)
        print(f"case: {s.name} first-step force rel rms {e:.2e}")
        assert e < 1e-5
        assert np.array_equal(_half_list(m1, s.n_atoms, dev), _half_list(m2, s.n_atoms, dev))
    # graph launches of 8 steps (2 launches and a remainder of 3 steps)
    for m in singles:
        m.run(19, use_graph=True)
    b.run(19, steps_per_launch=8)
    b.synchronize()
    for s, m1, m2 in zip(boxes, singles, batched):
        c1, v1, f1 = m1.state()
        c2, v2, f2 = m2.state()
        dmax = float(np.abs(c1.astype(np.float64) - c2).max())
        print(f"case: {s.name} after 20 steps max |dx| {dmax:.2e} A, force rel rms "
              f"{rel_rms(f2, f1):.2e}")
        # the runs differ only by the order of float atomics
        assert dmax < 1e-3
        assert rel_rms(f2, f1) < 1e-3
    b.clear()


def test_autotune_keeps_forces(dev, boxes):
    import torch
    from sponge_amd import abi
    s = boxes[2]
    d = DevSystem(s, dev)
    box = abi.Box(s.box)
    pm = abi.ParticleMesh(d.N, box, s.cutoff)
    f0 = d.zeros(d.N, 3)
    pm.reciprocal_force(d.crd, d.charge, f0)
    chosen = (C.c_int * 3)()
    rc = abi.load().sponge_pme_autotune(pm.h, abi._ptr(d.crd), abi._ptr(d.charge), 3, chosen)
    assert rc == 0
    print(f"case: grid {pm.fft} tiling {tuple(chosen)}")
    assert all(c in (2, 4, 8, 16) for c in chosen)
    f1 = d.zeros(d.N, 3)
    pm.reciprocal_force(d.crd, d.charge, f1)
    torch.cuda.synchronize()
    assert rel_rms(f1.cpu().numpy(), f0.cpu().numpy()) < 1e-5
    # tilings are validated
    bad = (C.c_int * 3)(3, 4, 4)
    assert abi.load().sponge_pme_set_fft_tiling(pm.h, bad) == 1006


def test_batch_autotune_and_graph(dev, boxes):
    from sponge_amd import abi
    ms = [abi.MDSystem(s) for s in boxes]
    b = abi.Batch(ms)
    b.run(1)
    tiles = b.autotune(3)
    print(f"case: tilings {tiles}")
    assert len(tiles) == len(boxes)
    b.run(32, steps_per_launch=16)
    b.synchronize()
    for m in ms:
        m.run(1, energy_interval=1)
        e = m.energies()
        assert np.isfinite(e["potential"]) and e["kinetic"] > 0
    b.clear()


def test_batch_48_boxes_against_oracle(dev):
    """The batched workload: 48 water boxes of 500-6000 waters (log-uniform,
    seeded), cutoff 9 A, each with its own ~1 A PME grid.  After a few batched
    steps every list is rebuilt by a batched step; that step's list of every
    box is the oracle's pair set of the step's coordinates, exactly, and its
    force is the oracle's within 1e-5 relative RMS."""
    from oracle import oracle as O
    from sponge_amd import abi
    from sponge_amd import system as S
    boxes = S.batched_boxes(48)
    print(f"case: 48 boxes, {sum(s.n_atoms for s in boxes)} atoms", flush=True)
    ms = [abi.MDSystem(s) for s in boxes]
    b = abi.Batch(ms)
    b.run(6, steps_per_launch=4)   # first step eager, then batched launches
    b.synchronize()
    x0 = [m.state()[0] for m in ms]
    for m in ms:
        m.request_list_update()
    b.run(1, steps_per_launch=1)
    b.synchronize()
    worst = 0.0
    for k, (s, m) in enumerate(zip(boxes, ms)):
        _, _, f = m.state()
        got = _half_list(m, s.n_atoms, dev)
        c, ser, flags = O.nl_build(s, crd=x0[k])
        assert flags == 0
        want = O.pair_array(c, ser)
        assert np.array_equal(got, want), (k, s.name, len(got), len(want))
        beta = O.get_beta(s.cutoff, s.tolerance)
        ref = O.lj_direct(s, c, ser, beta, crd=x0[k], need_energy=False,
                          need_virial=False)["frc"].astype(np.float64)
        ref += O.pme_excluded(s, beta, crd=x0[k], need_energy=False, need_virial=False)["frc"]
        ref += O.PME(s, m.pme_grid(), beta).reciprocal(crd=x0[k], need_energy=False,
                                                      need_virial=False)["frc"]
        e = rel_rms(f, ref)
        worst = max(worst, e)
        assert e < 1e-5, (k, s.name, e)
    print(f"48 boxes: lists bit-exact, worst force rel RMS {worst:.2e}", flush=True)
    b.clear()
