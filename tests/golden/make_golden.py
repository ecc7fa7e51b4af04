"""Writes tests/golden/r2r_scipy.npz: seeded inputs and scipy.fft outputs for
the four FFTW r2r kinds the reference uses (REDFT10/REDFT01/RODFT10/RODFT01,
unnormalised; scipy's dct/dst types 2 and 3 with norm=None are the same
definitions).  FFTW itself is not available in this image; scipy 1.15 is.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy
import scipy.fft as sf

HERE = os.path.dirname(os.path.abspath(__file__))
KINDS = {0: ("dct", 2), 1: ("dct", 3), 2: ("dst", 2), 3: ("dst", 3)}


def apply1(x, kind, axis=-1):
    fn, t = KINDS[kind]
    return getattr(sf, fn)(x, type=t, axis=axis, norm=None)


def main():
    rng = np.random.default_rng(20260101)
    out = {"scipy_version": np.array(scipy.__version__)}
    for n in (4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512):
        x = rng.standard_normal(n)
        out[f"x_{n}"] = x
        for kind in range(4):
            out[f"y_{n}_{kind}"] = apply1(x, kind)
    # 3D mixed kinds (x, y, z) on a 6x8x4 grid, applied z, y, x
    x3 = rng.standard_normal((6, 8, 4))
    out["x3"] = x3
    for kx, ky, kz in ((0, 0, 0), (1, 1, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2),
                       (3, 1, 1), (1, 3, 1), (1, 1, 3)):
        y = apply1(apply1(apply1(x3, kz, 2), ky, 1), kx, 0)
        out[f"y3_{kx}{ky}{kz}"] = y
    np.savez_compressed(os.path.join(HERE, "r2r_scipy.npz"), **out)


if __name__ == "__main__":
    main()
