"""Closed forms and independent transforms for pinning the oracle.

The r2r transforms here are scipy.fft's unnormalised DCT/DST (types 2 and 3,
norm=None), which equal FFTW's REDFT10/REDFT01/RODFT10/RODFT01 (the committed
fixture tests/golden/r2r_scipy.npz and tests/test_oracle.py pin that
equality).  Grid functions are sampled where the DCT-II puts its points,
x_n = (n + 1/2) dx, so cos(q_a x_n) with q_a = pi a / L is the a-th DCT mode.
"""
import numpy as np
import scipy.fft as sf

REDFT10, REDFT01, RODFT10, RODFT01 = 0, 1, 2, 3


def spacing(p):
    return 2 * np.pi / (p.q0 * p.nw)


def coords(p):
    """Sample positions (x, y, z) of each axis, shapes broadcastable to the grid."""
    h = spacing(p)
    x = (np.arange(p.nx) + 0.5) * h
    y = (np.arange(p.ny) + 0.5) * h
    z = (np.arange(p.nz) + 0.5) * h
    return x[:, None, None], y[None, :, None], z[None, None, :]


def wavenumbers(p):
    h = spacing(p)
    return [np.pi * np.arange(n) / (n * h) for n in (p.nx, p.ny, p.nz)]


def mode(p, a, b, c):
    """cos(q_a x) cos(q_b y) cos(q_c z) and its analytic gradient."""
    x, y, z = coords(p)
    qx, qy, qz = (w[i] for w, i in zip(wavenumbers(p), (a, b, c)))
    cx, cy, cz = np.cos(qx * x), np.cos(qy * y), np.cos(qz * z)
    sx, sy, sz = np.sin(qx * x), np.sin(qy * y), np.sin(qz * z)
    f = cx * cy * cz
    g = [-qx * sx * cy * cz, -qy * cx * sy * cz, -qz * cx * cy * sz]
    return f, g


def mq2(p, a, b, c):
    qx, qy, qz = (w[i] for w, i in zip(wavenumbers(p), (a, b, c)))
    return qx * qx + qy * qy + qz * qz


def alin(p, m2):
    """L(q) = eps - alpha (|q|^2 - q0^2)^2."""
    return p.eps - p.alpha * (m2 - p.q0 ** 2) ** 2


def r2r(x, kinds):
    """Separable unnormalised 3D r2r with FFTW kinds per axis (x, y, z)."""
    y = np.asarray(x, dtype=np.float64)
    for ax, k in enumerate(kinds):
        if k == REDFT10:
            y = sf.dct(y, type=2, axis=ax)
        elif k == REDFT01:
            y = sf.dct(y, type=3, axis=ax)
        elif k == RODFT10:
            y = sf.dst(y, type=2, axis=ax)
        else:
            y = sf.dst(y, type=3, axis=ax)
    return y


def ct(x):
    return r2r(x, (REDFT10,) * 3)


def ict(x):
    return r2r(x, (REDFT01,) * 3)


ST = ((RODFT10, REDFT10, REDFT10), (REDFT10, RODFT10, REDFT10), (REDFT10, REDFT10, RODFT10))
IST = ((RODFT01, REDFT01, REDFT01), (REDFT01, RODFT01, REDFT01), (REDFT01, REDFT01, RODFT01))


def scale(p):
    return 0.125 / (p.nx * p.ny * p.nz)


def mq2_grid(p):
    qx, qy, qz = wavenumbers(p)
    return qx[:, None, None] ** 2 + qy[None, :, None] ** 2 + qz[None, None, :] ** 2
