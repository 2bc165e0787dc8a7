"""The erfcx polynomial of the pair kernel's force-only path
(sponge_amd/csrc/common.h, erfcx_poly_t) against scipy.special.erfcx, and its
generator (tests/golden/make_erfcx_fit.py) reproducing a fit of the same
accuracy."""
import os
import re
import sys

import numpy as np
from scipy.special import erfcx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

import make_erfcx_fit as G  # noqa: E402


def committed_coefficients():
    """The literals of erfcx_poly_t in Horner order (highest power first), and
    the t scale kErfcxT of t = 1 / (1 + kErfcxT x)."""
    src = open(os.path.join(ROOT, "sponge_amd", "csrc", "common.h")).read()
    body = src[src.index("float erfcx_poly_t(float t)"):]
    body = body[:body.index("return p * t;")]
    vals = re.findall(r"([-+]?\d\.\d+e[-+]\d+)f", body)
    assert len(vals) == G.DEGREE + 1, vals
    scale = re.search(r"constexpr float kErfcxT = ([0-9.]+)f;", src)
    assert scale and float(scale.group(1)) == G.T_SCALE
    return [float(v) for v in vals]


def test_committed_constants_fp64():
    c_high = committed_coefficients()
    err = G.max_rel_error(c_high[::-1])
    print(f"FP64 max relative error {err:.3e}")
    assert err < 3.0e-8


def test_committed_constants_fp32_horner():
    c_high = np.asarray(committed_coefficients(), np.float32)
    x = G.samples(20000).astype(np.float32)
    t = (np.float32(1) / (np.float32(1) + np.float32(G.T_SCALE) * x)).astype(np.float32)
    p = np.full_like(t, c_high[0])
    for c in c_high[1:]:
        p = (p * t + c).astype(np.float32)
    approx = (p * t).astype(np.float32).astype(np.float64)
    err = float(np.max(np.abs(approx / erfcx(x.astype(np.float64)) - 1.0)))
    print(f"FP32 Horner max relative error {err:.3e}")
    assert err < 4.0e-7


def test_generator_reproduces_the_fit():
    c = G.fit(iterations=200)
    assert G.max_rel_error(c) < 3.0e-8
    # the generator lands next to the committed coefficients
    committed = np.asarray(committed_coefficients()[::-1])
    assert np.max(np.abs(c - committed) / np.abs(committed)) < 1e-3
