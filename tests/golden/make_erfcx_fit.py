"""Generator of the erfcx polynomial used by the pair kernel's force-only path.

erfcx(x) = exp(x^2) erfc(x) is approximated on 0 <= x <= 4.5 as t P(t) with
t = 1 / (1 + 0.4 x) and P of degree 7.  The coefficients minimise the largest
relative error: Lawson's algorithm (iteratively reweighted least squares, the
weights raised in proportion to each sample's error) over 4000 Chebyshev-
spaced samples, in FP64 against scipy.special.erfcx.

Run:  python tests/golden/make_erfcx_fit.py
prints the coefficients (highest power first, the order the kernel's Horner
evaluation consumes them) and the largest relative error of the fit in FP64.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erfcx

X_MAX = 4.5
T_SCALE = 0.4
DEGREE = 7


def samples(n=4000):
    k = np.arange(n)
    return 0.5 * X_MAX * (1.0 - np.cos(np.pi * (k + 0.5) / n))


def design(x):
    t = 1.0 / (1.0 + T_SCALE * x)
    return np.stack([t ** (p + 1) for p in range(DEGREE + 1)], 1)  # t P(t): powers 1..8


def fit(iterations=400):
    """Coefficients c[p] of t^(p+1), p = 0..DEGREE (lowest first)."""
    x = samples()
    y = erfcx(x)
    A = design(x) / y[:, None]      # relative error: (A c - y) / y
    b = np.ones_like(x)
    w = np.full(len(x), 1.0 / len(x))
    c = None
    for _ in range(iterations):
        sw = np.sqrt(w)
        c, *_ = np.linalg.lstsq(A * sw[:, None], b * sw, rcond=None)
        err = np.abs(A @ c - b)
        w = w * err
        w /= w.sum()
    return c


def max_rel_error(c_low_first, x=None):
    x = samples(20000) if x is None else x
    approx = design(x) @ np.asarray(c_low_first, np.float64)
    return float(np.max(np.abs(approx / erfcx(x) - 1.0)))


if __name__ == "__main__":
    c = fit()
    print("coefficients, highest power of t first (Horner order of the kernel):")
    for v in c[::-1]:
        print(f"    {v:.9e}")
    print(f"max relative error on [0, {X_MAX}]: {max_rel_error(c):.3e}")
