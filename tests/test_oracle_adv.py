"""CPU pinning of the advective oracle (adv/cosAdvTrackMult.cpp:827-1369).

The reference's velocity projection reads Sy[i+1,j-1,k], Sz[i+1,j,k-1], ...
(:1103-1274).  These shifted reads are what make the three DST/DCT force
components refer to one physical wavevector; the tests below fail for any
transposed or mis-shifted read:
  * planar layers psi(z) give u == 0 and the plain-model trajectory;
  * the spectral divergence q . u_hat of the projected velocity vanishes on
    all modes the reference does not truncate;
  * a single force mode gives (I - q q^T/|q|^2) F / (nu |q|^2).
"""
import numpy as np
import pytest

from conftest import rel_l2


def adv_params(smadft, shape, **kw):
    p = smadft.Params.default(smadft.MODEL_ADV, shape[0])
    p.nx, p.ny, p.nz = shape
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_planar_layers_no_flow(smadft, oracle):
    shape = (16, 8, 32)
    p = adv_params(smadft, shape)
    dz = 2 * np.pi / (p.q0 * p.nw)
    z = np.arange(shape[2]) * dz
    psi0 = np.broadcast_to(1.2 * np.cos(p.q0 * z) * np.exp(-((z - z.mean()) / 8) ** 2), shape).copy()
    a = oracle.AdvOracle(p, psi0)
    a.step(15)
    assert np.abs(a.vel).max() < 1e-12 * np.abs(a.grad[2]).max()
    pl = oracle.PlainOracle(p, psi0)
    pl.step(15)
    assert rel_l2(a.psi, pl.psi) < 1e-12
    # grad psi of planar layers is along z only
    assert np.abs(a.grad[0]).max() < 1e-14 and np.abs(a.grad[1]).max() < 1e-14


def spectral_velocity(oracle, vel):
    # forward transforms of u: STx, STy, STz (round trip gives 8 N^3 u_hat)
    ux = oracle.r2r3d(vel[0], (oracle.RODFT10, oracle.REDFT10, oracle.REDFT10))
    uy = oracle.r2r3d(vel[1], (oracle.REDFT10, oracle.RODFT10, oracle.REDFT10))
    uz = oracle.r2r3d(vel[2], (oracle.REDFT10, oracle.REDFT10, oracle.RODFT10))
    return ux, uy, uz


def test_projected_velocity_is_divergence_free(smadft, oracle):
    shape = (16, 16, 16)
    p = adv_params(smadft, shape)
    psi0 = smadft.ic_curved(p, 1.328, 1e-2, seed=2)
    a = oracle.AdvOracle(p, psi0)
    a.step(1)
    ux, uy, uz = spectral_velocity(oracle, a.vel)
    dx = 2 * np.pi / (p.q0 * p.nw)
    q = [np.pi * np.arange(n) / (n * dx) for n in shape]
    nx, ny, nz = shape
    div = np.zeros((nx - 1, ny - 2, nz - 2))
    mag = np.zeros_like(div)
    # DCT mode (i, j, k), i in 1..nx-1, j in 1..ny-2, k in 1..nz-2
    for i in range(1, nx):
        for j in range(1, ny - 1):
            for k in range(1, nz - 1):
                t = (q[0][i] * ux[i - 1, j, k] + q[1][j] * uy[i, j - 1, k] + q[2][k] * uz[i, j, k - 1])
                div[i - 1, j - 1, k - 1] = t
                mag[i - 1, j - 1, k - 1] = abs(q[0][i] * ux[i - 1, j, k]) + abs(q[1][j] * uy[i, j - 1, k]) + abs(q[2][k] * uz[i, j, k - 1])
    assert mag.max() > 1e-6
    assert np.abs(div).max() < 1e-12 * mag.max()


def test_single_force_mode_stokes_solution(smadft, oracle):
    shape = (8, 8, 16)
    p = adv_params(smadft, shape, nu=1.7)
    rng = np.random.default_rng(11)
    F = rng.standard_normal(3)
    a_, b_, c_ = 3, 2, 5  # DCT wavevector indices of the physical mode
    nx, ny, nz = shape
    # force components of wavevector (qx(a), qy(b), qz(c)) in their own bases
    Sx = np.zeros(shape); Sy = np.zeros(shape); Sz = np.zeros(shape)
    Sx[a_ - 1, b_, c_] = F[0]
    Sy[a_, b_ - 1, c_] = F[1]
    Sz[a_, b_, c_ - 1] = F[2]
    ux, uy, uz = oracle.adv_stokes(p, Sx, Sy, Sz)
    dx = 2 * np.pi / (p.q0 * p.nw)
    qv = np.array([np.pi * a_ / (nx * dx), np.pi * b_ / (ny * dx), np.pi * c_ / (nz * dx)])
    q2 = qv @ qv
    scale = 0.125 / (nx * ny * nz)
    u = scale * (F - qv * (qv @ F) / q2) / (p.nu * q2)
    got = np.array([ux[a_ - 1, b_, c_], uy[a_, b_ - 1, c_], uz[a_, b_, c_ - 1]])
    assert np.allclose(got, u, rtol=1e-13, atol=0)
    # no other mode is excited
    assert np.count_nonzero(ux) == 1 and np.count_nonzero(uy) == 1 and np.count_nonzero(uz) == 1
