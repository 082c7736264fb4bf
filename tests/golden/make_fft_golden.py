"""Writes the FFT golden vectors used to pin the oracle's (and the engine's)
3-D FFT: numpy.fft.fftn / ifftn (unnormalised inverse) of seeded inputs on
non-cubic shapes, including radix-3 lengths.  numpy is an independent
library, not the reference; the reference's FFT is cuFFT, which cannot run
here.  Run: python tests/golden/make_fft_golden.py"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SHAPES = [(8, 6, 4), (16, 4, 8), (48, 2, 4), (4, 24, 6)]


def main():
    out = {}
    for s, shape in enumerate(SHAPES):
        rng = np.random.default_rng(1000 + s)
        x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        key = "x".join(str(d) for d in shape)
        out["in_" + key] = x
        out["fwd_" + key] = np.fft.fftn(x)
        out["inv_" + key] = np.fft.ifftn(x) * x.size
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


if __name__ == "__main__":
    main()
