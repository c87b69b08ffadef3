"""Writes the small golden vectors the oracle is pinned against.

numpy's FFT (pocketfft) is an independent library routine; its outputs on
tiny inputs, including radix-3/5/7 lengths, are stored as fixtures so the
CPU suite does not depend on regenerating them.  Run from the repo root:
    python tests/golden/make_golden.py
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20261202)
    d = {}
    for n in (6, 9, 12, 15, 16, 21, 24, 35, 48, 64, 96, 125, 128):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        d[f"fft_in_{n}"] = x
        d[f"fft_out_{n}"] = np.fft.fft(x)
    for shape in ((4, 6, 8), (3, 5, 12), (6, 9, 10), (8, 12, 18)):
        a = rng.standard_normal(shape)
        tag = "x".join(map(str, shape))
        d[f"rfftn_in_{tag}"] = a
        d[f"rfftn_out_{tag}"] = np.fft.rfftn(a)
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **d)


if __name__ == "__main__":
    main()
