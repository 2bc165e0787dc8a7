"""NVE step driver on the GPU: SETTLE keeps the rigid-water geometry, the
total energy is conserved, and a HIP-graph replay matches eager stepping."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def water_geometry_error(s, crd):
    """Largest deviation (A) of O-H1, O-H2, H1-H2 from the initial geometry."""
    nw = s.n_waters
    c = crd[:3 * nw].reshape(nw, 3, 3).astype(np.float64)
    c0 = s.crd[:3 * nw].reshape(nw, 3, 3).astype(np.float64)
    L = s.box.astype(np.float64)

    def d(a, i, j):
        x = a[:, i] - a[:, j]
        x -= np.round(x / L) * L
        return np.sqrt((x * x).sum(-1))

    err = 0.0
    for i, j in ((0, 1), (0, 2), (1, 2)):
        err = max(err, float(np.abs(d(c, i, j) - d(c0, i, j)).max()))
    return err


def test_settle_and_energy_conservation(dev):
    from sponge_amd import abi
    from sponge_amd import system as S
    s = S.config(1)
    md = abi.MDSystem(s)
    md.run(1, energy_interval=1)
    e_first = md.energies()
    totals = []
    for block in range(10):
        md.run(20, energy_interval=20)
        e = md.energies()
        totals.append(e["potential"] + e["kinetic"])
    crd, vel, frc = md.state()
    assert np.isfinite(crd).all() and np.isfinite(vel).all()
    assert water_geometry_error(s, crd) < 2e-4
    # leap-frog + SETTLE: v(n+1/2) = (x(n+1) - x(n)) / dt with both ends on the
    # constraint manifold, so each bond's relative velocity is orthogonal to
    # the bond at the midpoint configuration
    nw = s.n_waters
    dt = s.dt_ps * 20.455
    c = crd[:3 * nw].reshape(nw, 3, 3).astype(np.float64)
    v = vel[:3 * nw].reshape(nw, 3, 3).astype(np.float64)
    c_prev = c - dt * v
    for i, j in ((0, 1), (0, 2), (1, 2)):
        r_mid = 0.5 * ((c[:, i] - c[:, j]) + (c_prev[:, i] - c_prev[:, j]))
        r_mid -= np.round(r_mid / s.box) * s.box
        dv = v[:, i] - v[:, j]
        proj = np.abs((r_mid * dv).sum(-1)) / np.sqrt((r_mid * r_mid).sum(-1))
        assert proj.max() < 2e-3 * np.abs(v).max(), proj.max()
    totals = np.array(totals)
    kin = e_first["kinetic"]
    # leap-frog with half-step kinetic energy: fluctuation, no runaway drift
    # (the lattice start still relaxes here; the equilibrated 95k-atom test
    # in test_gpu_fullsize.py holds the tight bound)
    assert np.abs(totals - totals[0]).max() < 0.05 * abs(kin), totals


def test_graph_replay_matches_eager(dev):
    from sponge_amd import abi
    from sponge_amd import system as S
    s = S.config(1)
    a = abi.MDSystem(s)
    b = abi.MDSystem(s)
    a.run(12, energy_interval=0, use_graph=False)
    b.run(12, energy_interval=0, use_graph=True)
    ca, va, _ = a.state()
    cb, vb, _ = b.state()
    # float atomics make the summation order run-dependent: agreement, not
    # bit equality
    assert np.abs(ca - cb).max() < 1e-3
    assert np.abs(va - vb).max() < 1e-2 * np.abs(va).max()


@pytest.mark.parametrize("use_graph", [False, True])
def test_reciprocal_overlap_matches_one_stream(dev, use_graph):
    """The reciprocal half on its own stream (forced on for this small box)
    gives the one-stream trajectory up to the order of the float additions;
    switching the mode between runs rebuilds the captured step."""
    from sponge_amd import abi
    from sponge_amd import system as S
    s = S.config(1)
    a = abi.MDSystem(s)
    b = abi.MDSystem(s)
    a.set_reciprocal_overlap(0)
    b.set_reciprocal_overlap(1)
    a.run(12, energy_interval=0, use_graph=use_graph)
    b.run(12, energy_interval=0, use_graph=use_graph)
    ca, va, fa = a.state()
    cb, vb, fb = b.state()
    print(f"case: graph={use_graph} max |dx| {np.abs(ca - cb).max():.2e}")
    assert np.abs(ca - cb).max() < 1e-3
    assert np.abs(va - vb).max() < 1e-2 * np.abs(va).max()
    # energies on an energy step (the reciprocal energy sum is accumulated
    # on the side stream) and a switch back to one stream
    a.run(1, energy_interval=1)
    b.run(1, energy_interval=1)
    ea, eb = a.energies(), b.energies()
    assert abs(ea["potential"] - eb["potential"]) < 1e-3 * abs(ea["potential"])
    b.set_reciprocal_overlap(0)
    b.run(8, energy_interval=0, use_graph=use_graph)
    assert np.isfinite(b.state()[0]).all()
    with pytest.raises(abi.SpongeError):
        b.set_reciprocal_overlap(2)


def test_reciprocal_overlap_with_impulse_schedule(dev):
    """PME update_interval = 2 (the reference's impulse schedule: the
    reciprocal force every second step, doubled): the two-stream step skips
    the reciprocal half on the other steps exactly as the one-stream step."""
    from sponge_amd import abi
    from sponge_amd import system as S
    s = S.config(1)
    a = abi.MDSystem(s, pme_opts={"update_interval": 2})
    b = abi.MDSystem(s, pme_opts={"update_interval": 2})
    a.set_reciprocal_overlap(0)
    b.set_reciprocal_overlap(1)
    a.run(11, energy_interval=0)
    b.run(11, energy_interval=0)
    ca, va, _ = a.state()
    cb, vb, _ = b.state()
    print(f"case: impulse schedule max |dx| {np.abs(ca - cb).max():.2e}")
    assert np.abs(ca - cb).max() < 1e-3
    assert np.abs(va - vb).max() < 1e-2 * np.abs(va).max()


def test_overlap_first_force_matches_one_stream(dev):
    """First force-only step from the same coordinates: on two streams the
    PME gather assigns each atom's force (the pair kernel's per-row force
    found through the list's rank map, plus the reciprocal and the
    side-stream excluded-pair forces); on one stream the finalize pass
    assigns the pair force, the excluded-pair correction and the gather add
    to it.  The totals differ only in the order of the float additions."""
    from sponge_amd import abi
    from sponge_amd import system as S
    s = S.config(1)
    a = abi.MDSystem(s)
    b = abi.MDSystem(s)
    a.set_reciprocal_overlap(0)
    b.set_reciprocal_overlap(1)
    a.run(1, energy_interval=0)
    b.run(1, energy_interval=0)
    fa = a.state()[2].astype(np.float64)
    fb = b.state()[2].astype(np.float64)
    e = np.sqrt(((fb - fa) ** 2).sum() / (fa ** 2).sum())
    print(f"case: first step force rel RMS two streams vs one {e:.2e}, "
          f"max |df| {np.abs(fb - fa).max():.2e}", flush=True)
    # measured 3.0e-7 (a wrong row -> atom map gives order 1)
    assert e < 2e-6
    a.clear()
    b.clear()


def test_kernel_timers_inside_graph(dev):
    """Timers recorded during capture become graph event nodes that are
    re-armed on every replay: one timed pair-kernel launch per step."""
    from sponge_amd import abi
    from sponge_amd import system as S
    s = S.config(1)
    md = abi.MDSystem(s)
    md.run(2, use_graph=False)
    abi.set_kernel_timing_filter(["lj_pairs"])
    try:
        abi.set_kernel_timing(True)
        md.run(1, use_graph=True)
        abi.reset_kernel_timing()
        md.run(7, use_graph=True)
        ms, n = abi.kernel_timing("lj_pairs")
        assert n == 7 and ms > 0.0, (ms, n)
        # a filtered-out timer records nothing
        assert abi.kernel_timing("pme_spread")[1] == 0
    finally:
        abi.set_kernel_timing(False)
        abi.set_kernel_timing_filter(None)
        abi.reset_kernel_timing()


@pytest.mark.parametrize("delta", [0.0, 0.5, 1.0])
def test_pruned_rows_keep_every_pair(dev, delta):
    """The pair kernel of the step driver reads inner rows pruned to cutoff +
    delta (pruned again once an atom moved delta / 2): after 200 steps the
    force of a step is the oracle's force of its coordinates (no pair within
    the cutoff missing), for several buffers and for the full rows."""
    from oracle import oracle as O
    from sponge_amd import abi
    from sponge_amd import system as S
    s = S.config(2)
    md = abi.MDSystem(s)
    md.set_list_prune(delta)
    md.run(200, use_graph=True)
    crd, _, _ = md.state()
    md.run(1)
    _, _, frc = md.state()
    beta = O.get_beta(s.cutoff, s.tolerance)
    c, ser, _ = O.nl_build(s, crd=crd)
    ref = O.lj_direct(s, c, ser, beta, crd=crd, need_energy=False,
                      need_virial=False)["frc"].astype(np.float64)
    ref += O.pme_excluded(s, beta, crd=crd, need_energy=False, need_virial=False)["frc"]
    ref += O.PME(s, s.fft, beta).reciprocal(crd=crd, need_energy=False, need_virial=False)["frc"]
    err = float(np.sqrt(((frc - ref) ** 2).sum() / (ref ** 2).sum()))
    print(f"case: prune delta {delta}: force rel RMS {err:.2e}")
    assert err < 1e-5
