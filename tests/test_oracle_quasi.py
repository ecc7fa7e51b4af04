"""CPU pinning of the quasi-incompressible oracle (quasi/quasi.cpp:990-2251).

  * kp = 0 gives a uniform density rho0 and divv == 0;
  * the mean pressure is zero (p_hat[0,0,0] = 0, :1577-1581);
  * quasi-incompressibility: the spectral divergence of the velocity equals
    the imposed divv on every mode the reference does not truncate -- this
    pins the shifted reads of S and of divv in the solenoidal and
    irrotational parts (:1599-2109);
  * before the switch step the velocity is solenoidal;
  * planar layers psi(z) keep ux = uy = 0.
"""
import numpy as np
import pytest


def qparams(smadft, shape, **kw):
    p = smadft.Params.default(smadft.MODEL_QUASI, shape[0])
    p.nx, p.ny, p.nz = shape
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_default_quasi_is_stable_at_reference_dt(smadft, oracle):
    p = qparams(smadft, (16, 16, 16), div_switch=4)
    q = oracle.QuasiOracle(p, smadft.ic_curved(p, 1.34164, 1e-3, seed=3))
    q.step(30)
    assert np.isfinite(q.psi).all() and np.abs(q.psi).max() < 2.0


def test_kp_zero_uniform_density(smadft, oracle):
    p = qparams(smadft, (16, 16, 16), kp=0.0, rho0=1.0, rho_m=0.5, div_switch=3)
    q = oracle.QuasiOracle(p, smadft.ic_curved(p, 1.34164, 1e-3, seed=3))
    q.step(8)
    assert np.abs(q.rho - 1.0).max() < 1e-13
    assert np.abs(q.divv).max() < 1e-9  # unscaled REDFT10 of roundoff


def test_mean_pressure_zero(smadft, oracle):
    p = qparams(smadft, (16, 8, 16), div_switch=2)
    q = oracle.QuasiOracle(p, smadft.ic_curved(p, 1.34164, 1e-3, seed=3))
    q.step(5)
    assert abs(q.p.mean()) < 1e-12 * np.abs(q.p).max()


def spectral_div(oracle, p, vel):
    ux = oracle.r2r3d(vel[0], (2, 0, 0))
    uy = oracle.r2r3d(vel[1], (0, 2, 0))
    uz = oracle.r2r3d(vel[2], (0, 0, 2))
    dx = 2 * np.pi / (p.q0 * p.nw)
    nx, ny, nz = p.nx, p.ny, p.nz
    qx = np.pi * np.arange(nx) / (nx * dx)
    qy = np.pi * np.arange(ny) / (ny * dx)
    qz = np.pi * np.arange(nz) / (nz * dx)
    sl = np.s_[1:nx, 1:ny - 1, 1:nz - 1]
    div = (qx[1:, None, None] * ux[0:nx - 1, 1:ny - 1, 1:nz - 1]
           + qy[None, 1:ny - 1, None] * uy[1:nx, 0:ny - 2, 1:nz - 1]
           + qz[None, None, 1:nz - 1] * uz[1:nx, 1:ny - 1, 0:nz - 2])
    return div, sl


def test_velocity_divergence_equals_divv(smadft, oracle):
    p = qparams(smadft, (16, 16, 16), div_switch=2)
    q = oracle.QuasiOracle(p, smadft.ic_curved(p, 1.34164, 1e-3, seed=3))
    q.step(6)
    div, sl = spectral_div(oracle, p, q.vel)
    n3 = 8 * p.nx * p.ny * p.nz
    scale = 1.0 / n3
    # ST(vel) = 8N^3 u_hat and q . u_hat = s divv_hat (divv_hat unscaled)
    target = n3 * scale * q.divv[sl]
    assert np.abs(target).max() > 1e-6
    assert np.abs(div - target).max() < 1e-9 * np.abs(target).max()


def test_velocity_solenoidal_before_switch(smadft, oracle):
    p = qparams(smadft, (16, 16, 16), div_switch=100)
    q = oracle.QuasiOracle(p, smadft.ic_curved(p, 1.34164, 1e-3, seed=3))
    q.step(4)
    div, _ = spectral_div(oracle, p, q.vel)
    ux = oracle.r2r3d(q.vel[0], (2, 0, 0))
    assert np.abs(ux).max() > 1e-6
    assert np.abs(div).max() < 1e-10 * np.abs(ux).max() * np.pi


def test_planar_layers_flow_along_normal_only(smadft, oracle):
    p = qparams(smadft, (8, 8, 32), div_switch=2)
    dz = 2 * np.pi / (p.q0 * p.nw)
    z = np.arange(p.nz) * dz
    psi0 = np.broadcast_to(1.34 * np.cos(z) * np.exp(-((z - z.mean()) / 6) ** 2), (p.nx, p.ny, p.nz)).copy()
    q = oracle.QuasiOracle(p, psi0)
    q.step(6)
    assert np.abs(q.vel[2]).max() > 1e-8
    assert np.abs(q.vel[0]).max() < 1e-12 * np.abs(q.vel[2]).max() + 1e-15
    assert np.abs(q.vel[1]).max() < 1e-12 * np.abs(q.vel[2]).max() + 1e-15
    assert np.abs(q.psi - q.psi[:1, :1, :]).max() < 1e-12
