"""Writes tests/golden/fft_golden.npz: seeded real 3D arrays and their
unnormalised real-to-complex transforms from numpy.fft (the library routine
the oracle's own mixed-radix FFT is pinned against)."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPES = [(6, 5, 9), (8, 12, 10), (7, 4, 14), (3, 3, 27), (5, 2, 96)]


def main():
    rng = np.random.default_rng(20260217)
    out = {}
    for k, shp in enumerate(SHAPES):
        x = rng.normal(size=shp)
        out[f"x{k}"] = x
        out[f"X{k}"] = np.fft.rfftn(x)
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


if __name__ == "__main__":
    main()
