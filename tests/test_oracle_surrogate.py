"""Pins of the surrogate's FP64 oracle (CPU): hand-computed two- and
three-atom features, the reciprocal pair field against a numerical
derivative, Chebyshev polynomials against numpy's, bfloat16 rounding on known
values, and the analytic loss gradient against central finite differences."""
import math

import numpy as np
import pytest

from oracle import surrogate as R

BOX = np.array([40.0, 40.0, 40.0])
BETA = 0.3123
RF = 10.0
NO_EXCL = (np.zeros(3, np.int32), np.zeros(3, np.int32), np.zeros(1, np.int32))


def test_dphi_matches_numerical_derivative():
    for r in (0.9, 1.5, 3.0, 7.5, 9.9):
        h = 1e-5
        phi = lambda x: math.erf(BETA * x) / x
        num = (phi(r + h) - phi(r - h)) / (2 * h)
        assert abs(R.dphi(r, BETA) - num) < 1e-9 * max(1.0, abs(num))
        # small r: phi' ~ -(4 / (3 sqrt(pi))) beta^3 r
    r = 1e-3
    assert abs(R.dphi(r, BETA) + 4 / (3 * math.sqrt(math.pi)) * BETA ** 3 * r) < 1e-9


def test_chebyshev_against_numpy():
    x = np.linspace(-1, 1, 17)
    t = R.chebyshev(x, R.N_RADIAL)
    for k in range(R.N_RADIAL):
        c = np.zeros(k + 1)
        c[k] = 1
        assert np.allclose(t[k], np.polynomial.chebyshev.chebval(x, c), atol=1e-12)


def test_two_atoms_by_hand():
    # atom 1 at distance r along x from atom 0, charge q1
    r, q0, q1 = 3.7, -0.8, 0.45
    crd = np.array([[5.0, 6.0, 7.0], [5.0 + r, 6.0, 7.0], [30.0, 30.0, 30.0]])
    q = np.array([q0, q1, 0.0])
    feat, basis = R.features(crd, q, BOX, RF, BETA, *NO_EXCL, atoms=[0])
    u = r / RF
    fc = (1 - u * u) ** 2
    x = 2 * u - 1
    w = q1 * fc
    T = [math.cos(k * math.acos(x)) for k in range(R.N_RADIAL)]
    assert np.allclose(feat[0, :R.N_RADIAL], [w * t for t in T], rtol=1e-12, atol=1e-14)
    # P_a = w T_a xhat: P_a . P_b = w^2 T_a T_b;  Q = w (xx^T - I/3): |Q|^2 = 2/3 w^2
    ang = [w * w * T[a] * T[b] for a in range(5) for b in range(a, 5)] + [2.0 / 3.0 * w * w]
    assert np.allclose(feat[0, R.N_RADIAL:], ang, rtol=1e-12, atol=1e-14)
    br = BETA * r
    dphi = (2 / math.sqrt(math.pi) * br * math.exp(-br * br) - math.erf(br)) / (r * r)
    assert np.allclose(basis[0, 0], [q1 * fc * dphi, 0, 0], rtol=1e-12)
    assert np.allclose(basis[0, 1], 0)
    assert np.allclose(basis[0, 2], [w, 0, 0], rtol=1e-12)
    # beyond the cutoff: nothing
    crd[1, 0] = 5.0 + RF + 0.01
    feat, basis = R.features(crd, q, BOX, RF, BETA, *NO_EXCL, atoms=[0])
    assert np.all(feat == 0) and np.all(basis == 0)


def test_three_atoms_exclusion_and_image():
    # atom 1 excluded from atom 0 (reference half CSR: j > i); atom 2 reached
    # through the periodic image
    crd = np.array([[1.0, 1.0, 1.0], [1.0, 1.96, 1.0], [39.0, 1.0, 1.0]])
    q = np.array([-0.834, 0.417, 0.9])
    ex_start = np.array([0, 1, 1], np.int32)
    ex_num = np.array([1, 0, 0], np.int32)
    ex_list = np.array([1], np.int32)
    feat, basis = R.features(crd, q, BOX, RF, BETA, ex_start, ex_num, ex_list)
    # atom 0: only atom 2 at d = (-2, 0, 0) on the list
    u = 2.0 / RF
    fc = (1 - u * u) ** 2
    assert math.isclose(feat[0, 0], 0.9 * fc, rel_tol=1e-12)
    assert np.allclose(basis[0, 2], [-0.9 * fc, 0, 0], rtol=1e-12)
    # exclusion field on 0 from 1 (d = +0.96 y) and on 1 from 0 (opposite sign pattern)
    r = 0.96
    assert np.allclose(basis[0, 1], [0, 0.417 * R.dphi(r, BETA), 0], rtol=1e-12)
    assert np.allclose(basis[1, 1], [0, -(-0.834) * R.dphi(r, BETA), 0], rtol=1e-12)
    # atom 1's list partner is atom 2 only
    d = np.array([-2.0, -0.96, 0.0])
    rr = np.linalg.norm(d)
    uu = rr / RF
    assert math.isclose(feat[1, 0], 0.9 * (1 - uu * uu) ** 2, rel_tol=1e-12)


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 1.0 + 2 ** -7 + 2 ** -8],
                    np.float32)
    # halfway 1 + 2^-8 rounds to even (1.0); 1 + 2^-8 + 2^-9 rounds up to 1 + 2^-7;
    # 1 + 2^-7 + 2^-8 is halfway between odd 1 + 2^-7 and even 1 + 2^-6: up
    want = np.array([1.0, 1.0, 1.0 + 2 ** -7, -2.5, 1.0 + 2 ** -6])
    assert np.array_equal(R.bf16_round(vals), want)


def test_initial_params_are_the_physics_baseline():
    p = R.init_params(3)
    feat = np.random.default_rng(0).normal(size=(7, R.N_FEAT))
    c, _ = R.forward(p, np.zeros(R.N_FEAT), np.ones(R.N_FEAT), feat)
    assert np.allclose(c, [[1.0, 1.0, 0.0]] * 7)


def test_loss_gradient_finite_differences():
    rng = np.random.default_rng(5)
    n = 9
    p = R.init_params(1) + 0.05 * rng.normal(size=R.N_PARAMS)
    feat = rng.normal(size=(n, R.N_FEAT))
    mu = rng.normal(size=R.N_FEAT) * 0.1
    isd = 1.0 + 0.2 * rng.random(R.N_FEAT)
    basis = rng.normal(size=(n, 3, 3))
    q = rng.normal(size=n)
    target = rng.normal(size=(n, 3))
    L, g = R.loss_and_grad(p, mu, isd, feat, basis, q, target)
    idx = rng.choice(R.N_PARAMS, 60, replace=False)
    # every block of the parameter vector is represented
    idx = np.concatenate([idx, [0, 3072, 3072 + 64, 7231, 7295, 7488, 7490]])
    h = 1e-6
    for i in idx:
        e = np.zeros(R.N_PARAMS)
        e[i] = h
        lp, _ = R.loss_and_grad(p + e, mu, isd, feat, basis, q, target)
        lm, _ = R.loss_and_grad(p - e, mu, isd, feat, basis, q, target)
        num = (lp - lm) / (2 * h)
        assert abs(num - g[i]) < 1e-6 * max(1.0, abs(num)), (i, num, g[i])
    assert R.N_PARAMS == 7491 and R.N_MODEL == 7491 + 96


@pytest.mark.parametrize("seed", [0, 1])
def test_rotation_equivariance(seed):
    # features are invariant and the basis rotates with the environment
    rng = np.random.default_rng(seed)
    n = 12
    crd = 20.0 + rng.normal(scale=2.5, size=(n, 3))
    q = rng.normal(size=n)
    big = np.array([400.0, 400.0, 400.0])
    f1, b1 = R.features(crd, q, big, RF, BETA, *(np.zeros(n, np.int32),) * 2, np.zeros(1, np.int32))
    a = rng.normal(size=(3, 3))
    rot, _ = np.linalg.qr(a)
    crd2 = (crd - 20.0) @ rot.T + 20.0
    f2, b2 = R.features(crd2, q, big, RF, BETA, *(np.zeros(n, np.int32),) * 2, np.zeros(1, np.int32))
    assert np.allclose(f1, f2, rtol=1e-9, atol=1e-12)
    assert np.allclose(b1 @ rot.T, b2, rtol=1e-9, atol=1e-12)
