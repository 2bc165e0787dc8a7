"""ORACLE -- TEST INFRASTRUCTURE ONLY.

FP64 restatement of the learned reciprocal-force surrogate: the per-atom
charge-environment features, the equivariant basis vectors, the MLP
48 -> 64 -> 64 -> 3 (SiLU) and its loss gradient.  This form has no
counterpart in the reference (it is a project-added form beside the plain
PM_force path); the restatement is pinned by hand-checkable two- and
three-atom cases, by closed forms of the basis (the exclusion field is the
reference's own reciprocal pair force between excluded atoms, the
derivative of erf(beta r) / r, SPONGE/PM_force/PM_force.cpp:1745-1817 uses
the same kernel) and by central finite differences of the loss.  Only
tests/ and the smoke check import it.

Features of atom i (sums over the partners j with r_ij < Rf that are not
excluded from i; d = r_j - r_i as a minimum image, r = |d|, u = r / Rf):
  fc = (1 - u^2)^2,  w_j = q_j fc,  x = 2u - 1,  T_k = Chebyshev T_k(x)
  R_k  = sum_j w_j T_k(x)                        k = 0..31   (radial)
  P_a  = sum_j w_j T_a(x) d / r                  a = 0..4    (vectors)
  S    = sum_j w_j d d^T / r^2,  Q = S - (R_0 / 3) I
  A    = [P_a . P_b for 0 <= a <= b <= 4] (15) + [|Q|_F^2] (1)  (angular)
Basis (equivariant vectors; the predicted force is q_i sum_m c_m U_m):
  U_0 = sum_j q_j fc phi'(r) d / r     (list partners, smooth cutoff)
  U_1 = sum_{j excluded} q_j phi'(r) d / r   (no cutoff)
  U_2 = P_0
with phi(r) = erf(beta r) / r, the reciprocal-space pair potential.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf as _erf

N_RADIAL = 32
N_VEC = 5
N_ANGULAR = 16
N_FEAT = N_RADIAL + N_ANGULAR
N_HIDDEN = 64
N_OUT = 3

# parameter vector layout: W1[64][48], b1[64], W2[64][64], b2[64], W3[3][64], b3[3]
_SHAPES = (("W1", (N_HIDDEN, N_FEAT)), ("b1", (N_HIDDEN,)), ("W2", (N_HIDDEN, N_HIDDEN)),
           ("b2", (N_HIDDEN,)), ("W3", (N_OUT, N_HIDDEN)), ("b3", (N_OUT,)))
N_PARAMS = sum(int(np.prod(s)) for _, s in _SHAPES)
# model blob = params + feature mean (48) + feature inverse std (48)
N_MODEL = N_PARAMS + 2 * N_FEAT


def unpack(params):
    out, o = {}, 0
    for name, shape in _SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(params[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in _SHAPES])


def dphi(r, beta):
    """d/dr [erf(beta r) / r]."""
    br = beta * np.asarray(r, np.float64)
    return (2.0 / math.sqrt(math.pi) * br * np.exp(-br * br) - _erf(br)) / (r * r)


def chebyshev(x, n):
    """T_0..T_{n-1} at x (array) by the three-term recurrence."""
    t = np.empty((n,) + np.shape(x))
    t[0] = 1.0
    if n > 1:
        t[1] = x
    for k in range(2, n):
        t[k] = 2.0 * x * t[k - 1] - t[k - 2]
    return t


def _excluded_sets(n, excl_start, excl_num, excl_list):
    ex = [set() for _ in range(n)]
    for i in range(n):
        for e in range(excl_num[i]):
            j = int(excl_list[excl_start[i] + e])
            ex[i].add(j)
            ex[j].add(i)
    return ex


def features(crd, charge, box, rf, beta, excl_start, excl_num, excl_list, atoms=None):
    """Features [n, 48] and basis [n, 3, 3] (U_m rows) of `atoms` (default
    all), FP64, brute force over every partner."""
    crd = np.asarray(crd, np.float64)
    q = np.asarray(charge, np.float64)
    box = np.asarray(box, np.float64)
    n_all = crd.shape[0]
    atoms = np.arange(n_all) if atoms is None else np.asarray(atoms)
    ex = _excluded_sets(n_all, excl_start, excl_num, excl_list)
    feat = np.zeros((len(atoms), N_FEAT))
    basis = np.zeros((len(atoms), 3, 3))
    for row, i in enumerate(atoms):
        d = crd - crd[i]
        d -= np.floor(d / box + 0.5) * box
        r2 = (d * d).sum(1)
        mask = (r2 < rf * rf) & (r2 > 0)
        mask[i] = False
        exm = np.zeros(n_all, bool)
        if ex[i]:
            exm[list(ex[i])] = True
        sel = mask & ~exm
        dj, rj, qj = d[sel], np.sqrt(r2[sel]), q[sel]
        u = rj / rf
        fc = (1.0 - u * u) ** 2
        w = qj * fc
        t = chebyshev(2.0 * u - 1.0, N_RADIAL)
        feat[row, :N_RADIAL] = (w[None, :] * t).sum(1)
        rhat = dj / rj[:, None]
        P = np.stack([((w * t[a])[:, None] * rhat).sum(0) for a in range(N_VEC)])
        S = (w[:, None, None] * rhat[:, :, None] * rhat[:, None, :]).sum(0)
        Q = S - feat[row, 0] / 3.0 * np.eye(3)
        ang = [P[a] @ P[b] for a in range(N_VEC) for b in range(a, N_VEC)]
        ang.append((Q * Q).sum())
        feat[row, N_RADIAL:] = ang
        basis[row, 0] = ((qj * fc * dphi(rj, beta))[:, None] * rhat).sum(0)
        basis[row, 2] = P[0]
        if ex[i]:
            js = np.array(sorted(ex[i]))
            de = d[js]
            re = np.sqrt((de * de).sum(1))
            basis[row, 1] = ((q[js] * dphi(re, beta))[:, None] * de / re[:, None]).sum(0)
    return feat, basis


def bf16_round(a):
    """Round-to-nearest-even of float32 values to bfloat16, returned as float64."""
    a32 = np.ascontiguousarray(a, np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    rounded = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, mu, isd, feat, bf16=False):
    """Coefficients c [n, 3] and the intermediates (x, z1, h1, z2, h2).
    bf16=True rounds the matmul operands (inputs, weights, activations) to
    bfloat16 as the MFMA kernels do."""
    p = unpack(params)
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    x = rnd((np.asarray(feat, np.float64) - mu) * isd)
    z1 = x @ rnd(p["W1"]).T + p["b1"]
    h1 = rnd(silu(z1))
    z2 = h1 @ rnd(p["W2"]).T + p["b2"]
    h2 = rnd(silu(z2))
    c = h2 @ rnd(p["W3"]).T + p["b3"]
    return c, (x, z1, h1, z2, h2)


def predict(params, mu, isd, feat, basis, charge, bf16=False):
    c, _ = forward(params, mu, isd, feat, bf16)
    return np.asarray(charge, np.float64)[:, None] * np.einsum("nm,nmk->nk", c, basis)


def loss_and_grad(params, mu, isd, feat, basis, charge, target):
    """L = mean over atoms and components of (F_pred - target)^2 and dL/dparams."""
    p = unpack(params)
    q = np.asarray(charge, np.float64)
    c, (x, z1, h1, z2, h2) = forward(params, mu, isd, feat)
    F = q[:, None] * np.einsum("nm,nmk->nk", c, basis)
    res = F - target
    n = feat.shape[0]
    L = (res * res).sum() / (3 * n)
    dc = (2.0 / (3 * n)) * q[:, None] * np.einsum("nk,nmk->nm", res, basis)
    g = {"W3": dc.T @ h2, "b3": dc.sum(0)}
    dz2 = (dc @ p["W3"]) * dsilu(z2)
    g["W2"], g["b2"] = dz2.T @ h1, dz2.sum(0)
    dz1 = (dz2 @ p["W2"]) * dsilu(z1)
    g["W1"], g["b1"] = dz1.T @ x, dz1.sum(0)
    return L, pack(g)


def init_params(seed):
    """Initial parameters: scaled normal weights, the output layer zero with
    biases (1, 1, 0) -- the physics baseline F = q (U_0 + U_1)."""
    rng = np.random.default_rng(seed)
    p = {"W1": rng.standard_normal((N_HIDDEN, N_FEAT)) / math.sqrt(N_FEAT),
         "b1": np.zeros(N_HIDDEN),
         "W2": rng.standard_normal((N_HIDDEN, N_HIDDEN)) / math.sqrt(N_HIDDEN),
         "b2": np.zeros(N_HIDDEN),
         "W3": np.zeros((N_OUT, N_HIDDEN)),
         "b3": np.array([1.0, 1.0, 0.0])}
    return pack(p)


def rel_rmse(pred, ref):
    """Relative RMS error sqrt(sum |pred - ref|^2 / sum |ref|^2)."""
    pred = np.asarray(pred, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.sqrt(((pred - ref) ** 2).sum() / (ref * ref).sum()))
