"""The GPU's NVE step (leap-frog + SETTLE, SPONGE/MD_core/nve.hpp:3-16 and
SPONGE/constrain/settle.cpp:100-258 in the order of SPONGE/main.cpp:
1705-1777) against the oracle: one step from the GPU's own force (the
integrator alone), and ten steps with the oracle's forces at every step (the
whole MD step), on the 1,944-atom water box."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _mimg(d, box):
    return d - np.round(d / box) * box


def _oracle_force(s, crd):
    beta = O.get_beta(s.cutoff, s.tolerance)
    c, ser, _ = O.nl_build(s, crd=crd)
    f = O.lj_direct(s, c, ser, beta, crd=crd, need_energy=False,
                    need_virial=False)["frc"].astype(np.float64)
    f += O.pme_excluded(s, beta, crd=crd, need_energy=False, need_virial=False)["frc"]
    f += O.PME(s, s.fft, beta).reciprocal(crd=crd, need_energy=False, need_virial=False)["frc"]
    return f


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def test_one_step_against_oracle(dev):
    from sponge_amd import abi
    from sponge_amd import system as S
    s = S.config(1)
    md = abi.MDSystem(s)
    md.run(1)
    crd1, vel1, frc0 = md.state()
    # the force the step used
    f_ref = _oracle_force(s, s.crd)
    ef = _rel(frc0, f_ref)
    # the integrator alone, from the GPU's own force
    c64, v64, _ = O.nve_shake64(s, s.crd, s.vel, frc0.astype(np.float64))
    c32, v32 = O.nve_settle(s, s.crd, s.vel, frc0)
    dx64 = np.abs(_mimg(crd1 - c64, s.box)).max()
    dx32 = np.abs(_mimg(crd1 - c32, s.box)).max()
    ev64, ev32 = _rel(vel1, v64), _rel(vel1, v32)
    print(f"case: cfg1 one step: force rel {ef:.2e}; crd max |dx| vs FP64 {dx64:.2e}, "
          f"vs FP32 restatement {dx32:.2e}; vel rel vs FP64 {ev64:.2e}, vs FP32 {ev32:.2e}")
    assert ef < 1e-5
    # FP32 positions near 27 A (ulp 1.9e-6 A)
    assert dx64 < 6e-6 and dx32 < 6e-6
    # velocities: the GPU forms each water's centre-of-mass frame from
    # relative vectors, the FP32 restatement through the absolute position of
    # A as the reference does (up to half an ulp of a coordinate / dt per water)
    assert ev64 < 2e-5
    assert ev32 < 1e-4
    # the constraint correction is a large part of the step: the comparison
    # sees it, not only leap-frog
    dt = s.dt_ps * O.TIME_CONVERSION
    v_free = s.vel.astype(np.float64) + dt * frc0.astype(np.float64) / s.mass[:, None]
    assert _rel(v64, v_free) > 0.1


def test_ten_steps_against_oracle(dev):
    from sponge_amd import abi
    from sponge_amd import system as S
    s = S.config(1)
    md = abi.MDSystem(s)
    md.run(10)
    crd_g, vel_g, _ = md.state()
    crd = s.crd.astype(np.float32)
    vel = s.vel.astype(np.float64)
    for step in range(10):
        f = _oracle_force(s, crd)
        c64, vel, _ = O.nve_shake64(s, crd, vel, f)
        crd = c64.astype(np.float32)
    dx = np.abs(_mimg(crd_g - crd, s.box)).max()
    ev = _rel(vel_g, vel)
    print(f"case: cfg1 ten steps: crd max |dx| {dx:.2e} A, vel rel RMS {ev:.2e}")
    assert dx < 4e-5
    assert ev < 1e-4
