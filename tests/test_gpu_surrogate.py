"""GPU tests of the learned reciprocal-force surrogate through the C ABI:
features and basis against the FP64 oracle (also after the atoms moved on
an unchanged list), the BF16 MFMA forward pass and loss gradient against the
oracle's network, training on frames from the plain engine (the ablation:
trained network vs the same prediction without it), and the learned
multiple-time-step schedule in the MD driver (checks, fallbacks, graphs)."""
import numpy as np
import pytest

from gpu_helpers import DevSystem, rel_rms

pytestmark = pytest.mark.gpu

FEAT_TOL = 2e-5    # per feature column, relative to the column's RMS (FP32 sums)
BASIS_TOL = 1e-5   # relative RMS of each basis vector


@pytest.fixture(scope="module")
def cfg1():
    from sponge_amd import system as S
    return S.config(1)


def _moved(s, seed, scale=0.25, cap=0.45):
    rng = np.random.default_rng(seed)
    d = np.clip(rng.normal(scale=scale, size=s.crd.shape), -cap, cap)
    return (s.crd.astype(np.float64) + d).astype(np.float32)


def _setup(dev, s, crd_list, crd_feat, max_frames=4):
    import torch
    from oracle import oracle as O
    from sponge_amd import abi
    d = DevSystem(s, dev, crd_list)
    box = abi.Box(s.box)
    nl = abi.NeighborList(d.N, s.cutoff, s.skin, box)
    nl.update(d.crd, step=0, update=abi.NL_FORCED, excl=d.excl)
    if crd_feat is not None:
        d.crd.copy_(torch.tensor(crd_feat, device=dev))
        # within skin / 2 of the build: the conditional update keeps the list
        nl.update(d.crd, step=1, update=abi.NL_CONDITIONAL, excl=d.excl)
    beta = O.get_beta(s.cutoff, s.tolerance)
    rs = abi.ReciprocalSurrogate(d.N, beta, feature_cutoff=s.cutoff, max_frames=max_frames)
    return d, box, nl, rs, beta


def _export(rs, d):
    import torch
    feat = d.zeros(d.N, 48)
    basis = d.zeros(d.N, 9)
    torch.cuda.synchronize()
    rs.export_features(feat, basis)
    torch.cuda.synchronize()
    return feat.cpu().numpy().astype(np.float64), basis.cpu().numpy().reshape(d.N, 3, 3)


@pytest.mark.parametrize("case", ["at_build", "moved"])
def test_features_against_oracle(dev, cfg1, case):
    from oracle import surrogate as R
    s = cfg1
    crd0 = _moved(s, 11, 0.2)
    crd = crd0
    if case == "moved":
        # every atom moved by less than skin / 2 since the build
        rng = np.random.default_rng(12)
        step = np.clip(rng.normal(scale=0.25, size=crd0.shape), -0.45, 0.45)
        crd = (crd0.astype(np.float64) + step).astype(np.float32)
    d, box, nl, rs, beta = _setup(dev, s, crd0, None if case == "at_build" else crd)
    rs.features(nl, d.crd, d.charge, box, d.excl)
    feat, basis = _export(rs, d)
    rf, rb = R.features(crd, s.charge, s.box, s.cutoff, beta, s.excl_start, s.excl_num, s.excl_list)
    col_rms = np.sqrt((rf ** 2).mean(0))
    err = np.abs(feat - rf).max(0) / col_rms
    print(f"case: {case} worst feature column error {err.max():.2e} (column {err.argmax()})")
    assert np.all(err < FEAT_TOL), err
    for m in range(3):
        e = rel_rms(basis[:, m], rb[:, m])
        print(f"basis U_{m}: rel rms {e:.2e}")
        assert e < BASIS_TOL


def _random_model(seed, feat):
    from oracle import surrogate as R
    rng = np.random.default_rng(seed)
    p = R.init_params(seed)
    u = R.unpack(p)
    u["W3"] = rng.normal(scale=0.3, size=u["W3"].shape)
    u["b1"] = rng.normal(scale=0.1, size=u["b1"].shape)
    u["b2"] = rng.normal(scale=0.1, size=u["b2"].shape)
    u["b3"] = np.array([0.9, 1.1, 0.2])
    params = R.pack(u).astype(np.float32)
    mu = feat.mean(0).astype(np.float32)
    sd = feat.std(0)
    isd = np.where(sd > 0, 1.0 / np.maximum(sd, 1e-30), 1.0).astype(np.float32)
    return params, mu, isd


def test_mlp_forward_against_oracle(dev, cfg1):
    import torch
    from oracle import surrogate as R
    s = cfg1
    d, box, nl, rs, beta = _setup(dev, s, _moved(s, 21, 0.2), None)
    rs.features(nl, d.crd, d.charge, box, d.excl)
    feat, basis = _export(rs, d)
    params, mu, isd = _random_model(4, feat)
    rs.set_model(np.concatenate([params, mu, isd]))
    frc = d.zeros(d.N, 3)
    torch.cuda.synchronize()
    rs.predict(frc)
    torch.cuda.synchronize()
    got = frc.cpu().numpy()
    p64 = params.astype(np.float64)
    want_bf = R.predict(p64, mu.astype(np.float64), isd.astype(np.float64), feat, basis, s.charge,
                        bf16=True)
    want = R.predict(p64, mu.astype(np.float64), isd.astype(np.float64), feat, basis, s.charge)
    e_bf, e64 = rel_rms(got, want_bf), rel_rms(got, want)
    print(f"forward: rel rms vs bf16-rounded oracle {e_bf:.2e}, vs FP64 oracle {e64:.2e}")
    # the oracle with BF16-rounded operands differs from the MFMA path only in
    # FP32 accumulation order: a wrong fragment layout in any layer would give
    # errors of order one
    assert e_bf < 1e-4
    assert e64 < 2e-2
    # a second call accumulates (frc +=)
    rs.predict(frc)
    torch.cuda.synchronize()
    assert rel_rms(frc.cpu().numpy(), 2 * got) < 1e-6
    assert np.array_equal(rs.get_model(), np.concatenate([params, mu, isd]))


def test_gradient_against_oracle(dev, cfg1):
    import torch
    from oracle import surrogate as R
    from sponge_amd import abi
    s = cfg1
    crd = _moved(s, 31, 0.2)
    d, box, nl, rs, beta = _setup(dev, s, crd, None)
    pm = abi.ParticleMesh(d.N, box, s.cutoff, fftx=s.fft[0], ffty=s.fft[1], fftz=s.fft[2])
    rec = d.zeros(d.N, 3)
    pm.reciprocal_force(d.crd, d.charge, rec)
    rs.features(nl, d.crd, d.charge, box, d.excl)
    rs.add_frame(rec)
    torch.cuda.synchronize()
    assert rs.frame_count() == 1
    feat, basis = _export(rs, d)
    params, mu, isd = _random_model(9, feat)
    rs.set_model(np.concatenate([params, mu, isd]))
    loss, g = rs.gradient(0)
    r = rec.cpu().numpy().astype(np.float64)
    target = r - r.mean(0)
    p64, mu64, isd64 = params.astype(np.float64), mu.astype(np.float64), isd.astype(np.float64)
    L, g64 = R.loss_and_grad(p64, mu64, isd64, feat, basis, s.charge, target)
    F_bf = R.predict(p64, mu64, isd64, feat, basis, s.charge, bf16=True)
    L_bf = ((F_bf - target) ** 2).sum() / (3 * d.N)
    print(f"loss gpu {loss:.6e} bf16-oracle {L_bf:.6e} fp64 {L:.6e}")
    assert abs(loss - L_bf) < 1e-3 * L_bf
    o = 0
    for name, shape in R._SHAPES:
        n = int(np.prod(shape))
        e = rel_rms(g[o:o + n], g64[o:o + n])
        print(f"grad {name}: rel rms {e:.2e}")
        assert e < 1e-2, name
        o += n


@pytest.fixture(scope="module")
def trained(cfg1):
    """Frames from the plain engine (equilibrated 300 K water): 48 training
    frames 5 steps apart, then 32 consecutive held-out frames; the network
    is trained on the first and scored on the second, alone and anchored on
    the exact force 1-3 steps earlier (the learned multiple-time-step form)."""
    import torch
    from oracle import oracle as O
    from sponge_amd import abi
    s = cfg1
    md = abi.MDSystem(s)
    md.equilibrate(100, temperature=300.0, every=1)
    md.equilibrate(200, temperature=300.0, every=10)
    beta = O.get_beta(s.cutoff, s.tolerance)
    rs = abi.ReciprocalSurrogate(s.n_atoms, beta, feature_cutoff=s.cutoff, max_frames=80)
    md.collect_frames(rs, 48, 5)
    md.collect_frames(rs, 32, 1)
    assert rs.frame_count() == 80
    rs.fit_normalization(48)
    before = {lag: rs.evaluate(51, 80, lag) for lag in (1, 2, 3)}
    losses = rs.train(48, 40, 2e-3)
    after = {lag: rs.evaluate(51, 80, lag) for lag in (1, 2, 3)}
    torch.cuda.synchronize()
    return md, rs, before, after, losses


def test_training_beats_the_method_without_the_network(trained):
    md, rs, before, after, losses = trained
    print(f"loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    for lag in (1, 2, 3):
        b, a = before[lag], after[lag]
        print(f"held-out rel RMSE, {lag} step(s) after the exact force: stale {a['stale']:.4f}; "
              f"baseline {a['baseline']:.4f} (anchored {a['baseline_anchored']:.4f}); "
              f"network untrained {b['model']:.4f}, trained {a['model']:.4f} "
              f"(anchored {a['model_anchored']:.4f})")
        # the untrained network is exactly the physics baseline
        assert abs(b["model"] - b["baseline"]) < 1e-6
        assert abs(b["model_anchored"] - b["baseline_anchored"]) < 1e-6
        # training improves on the baseline, and anchored it improves on the
        # stale force (what the reference's impulse schedule applies)
        assert a["model"] < a["baseline"]
        assert a["model_anchored"] < a["stale"]
    assert losses[-1] < losses[0]


def test_md_learned_schedule(dev, trained):
    md, rs, _, _, _ = trained
    md.set_surrogate(rs, period=4, check_interval=2, bound=1e9)
    st0 = md.surrogate_stats()
    md.run(64, use_graph=False)
    st1 = md.surrogate_stats()
    # 3 of every 4 steps; checks on the first surrogate step of every 2nd period
    assert st1["surrogate_steps"] - st0["surrogate_steps"] == 48
    assert st1["checks"] - st0["checks"] == 8
    assert st1["fallbacks"] == st0["fallbacks"]
    err = st1["last_error"]
    print(f"last checked step: rel RMS error {err:.4f}")
    assert 0 < err < 0.5
    assert not st1["switched_off"]
    # a bound below the error: the first check falls back to PME and switches
    # the surrogate off; every later step is plain PME
    md.set_surrogate(rs, period=4, check_interval=2, bound=1e-9)
    md.run(64, use_graph=False)
    st2 = md.surrogate_stats()
    assert st2["checks"] - st1["checks"] == 1 and st2["fallbacks"] - st1["fallbacks"] == 1
    # the checked step comes first or after one unchecked period, depending on
    # the phase of the step counter: at most 3 unchecked steps before it
    assert 1 <= st2["surrogate_steps"] - st1["surrogate_steps"] <= 4 and st2["switched_off"]
    md.run(32, use_graph=False)
    st2b = md.surrogate_stats()
    assert st2b["surrogate_steps"] == st2["surrogate_steps"] and st2b["checks"] == st2["checks"]
    st2 = st2b
    # graph replay of whole 8-step cycles
    md.set_surrogate(rs, period=4, check_interval=2, bound=1e9)
    assert not md.surrogate_stats()["switched_off"]
    md.run(64, use_graph=True)
    st3 = md.surrogate_stats()
    assert st3["surrogate_steps"] - st2["surrogate_steps"] == 48
    assert st3["checks"] - st2["checks"] == 8 and st3["fallbacks"] == st2["fallbacks"]
    # under graph replay a failed check switches the surrogate off after the
    # cycle that held it: up to 3 unchecked eager steps until the step counter
    # reaches a cycle boundary, then one captured cycle of 6 surrogate steps,
    # one of them checked
    md.set_surrogate(rs, period=4, check_interval=2, bound=1e-9)
    md.run(64, use_graph=True)
    st4 = md.surrogate_stats()
    assert 1 <= st4["surrogate_steps"] - st3["surrogate_steps"] <= 9
    assert st4["checks"] - st3["checks"] == 1 and st4["fallbacks"] - st3["fallbacks"] == 1
    assert st4["switched_off"]
    md.run(32, use_graph=True)
    assert md.surrogate_stats()["surrogate_steps"] == st4["surrogate_steps"]
    md.run(1, energy_interval=1)
    e = md.energies()
    assert np.isfinite(e["potential"]) and np.isfinite(e["kinetic"])
    md.set_surrogate(None)


def test_impulse_schedule_conserves_energy(dev, trained):
    """The surrogate in the impulse schedule (its force every step, the kick
    period x (F_PME - F_surrogate) on the PME steps) keeps the total energy
    of the equilibrated box as the plain step does; the anchored schedule
    (not a conservative force) heats it far more."""
    md, rs, _, _, _ = trained
    crd0, vel0, _ = md.state()

    def drift(mode):
        md.set_state(crd0, vel0)
        if mode is None:
            md.set_surrogate(None)
        else:
            md.set_surrogate_mode(mode)
            md.set_surrogate(rs, period=4, check_interval=8, bound=1e9)
        e = []
        for k in range(9):
            md.run(1, energy_interval=1)
            en = md.energies()
            e.append(en["potential"] + en["kinetic"])
            md.run(63, use_graph=True)
        return float(np.polyfit(np.arange(9) * 64, np.asarray(e), 1)[0]) * 1000.0

    d_plain, d_impulse, d_anchored = drift(None), drift(1), drift(0)
    md.set_surrogate(None)
    md.set_surrogate_mode(1)
    print(f"drift kcal/mol per 1000 steps: plain {d_plain:.2f}, impulse {d_impulse:.2f}, "
          f"anchored {d_anchored:.2f}")
    # a slope fitted to 9 energies over 576 steps carries a few kcal/mol per
    # 1000 steps of noise (the bench measures 10,000 steps): the bars are
    # 10 kcal/mol per 1000 steps, or 5x the plain step's, against the
    # anchored schedule's hundreds
    assert abs(d_impulse) < max(5.0 * abs(d_plain), 10.0)
    assert abs(d_anchored) > 10.0 * abs(d_impulse)
