"""Pins the oracle's conserved plain variant (no-adv/cosNoAdvConserv.cpp:647-697,
:780-812): aLin and Nq carry |q|^2.  Closed forms: single linear modes decay
by the CN factor of |q|^2 L; the mean of psi (the q = 0 mode) is conserved
exactly; the first step from the pre-loop equals a scipy restatement of the
scheme (Euler in N on the first step)."""
import numpy as np
import pytest
import scipy.fft as sf

from test_oracle import cos_mode


def cparams(smadft, shape, **kw):
    p = smadft.Params.default(smadft.MODEL_PLAIN_CONSERVED, shape[0])
    p.nx, p.ny, p.nz = shape
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def q2_grid(p):
    dx = 2 * np.pi / (p.q0 * p.nw)
    q = [np.pi * np.arange(n) / (n * dx) for n in (p.nx, p.ny, p.nz)]
    return q[0][:, None, None] ** 2 + q[1][None, :, None] ** 2 + q[2][None, None, :] ** 2


def test_defaults_follow_the_reference(smadft):
    p = smadft.Params.default(smadft.MODEL_PLAIN_CONSERVED, 64)
    assert p.nw == 16 and p.dt == 0.0005


@pytest.mark.parametrize("mode", [(1, 2, 3), (5, 0, 7), (3, 3, 3)])
def test_conserved_linear_mode_decay(smadft, oracle, mode):
    shape = (16, 8, 24)
    p = cparams(smadft, shape, beta=0.0, gamma=0.0, eps=-0.3, dt=0.01)
    psi0 = cos_mode(shape, *mode)
    o = oracle.PlainOracle(p, psi0)
    nst = 30
    o.step(nst)
    dx = 2 * np.pi / (p.q0 * p.nw)
    q2 = sum((np.pi * m / (n * dx)) ** 2 for m, n in zip(mode, shape))
    lin = q2 * (p.eps - p.alpha * (q2 - p.q0 ** 2) ** 2)
    g = (1 + p.dt / 2 * lin) / (1 - p.dt / 2 * lin)
    err = np.linalg.norm(o.psi - g ** nst * psi0) / np.linalg.norm(psi0)
    assert err < 1e-13


def test_conserved_mean_is_exact(smadft, oracle):
    shape = (16, 16, 16)
    p = cparams(smadft, shape)
    psi0 = 0.3 + smadft.ic_curved(p, 1.328, 1e-3, seed=2)
    o = oracle.PlainOracle(p, psi0)
    o.step(40)
    assert abs(o.psi.mean() - psi0.mean()) < 1e-13
    assert np.linalg.norm(o.psi - psi0) > 1e-4  # it did evolve


def test_conserved_first_step_scipy(smadft, oracle):
    shape = (16, 12, 20)
    p = cparams(smadft, shape, dt=0.002)
    psi0 = 0.2 + smadft.ic_curved(p, 1.328, 1e-2, seed=5)
    o = oracle.PlainOracle(p, psi0)
    o.step(1)
    s = 1.0 / (8 * np.prod(shape))
    q2 = q2_grid(p)
    L = q2 * (p.eps - p.alpha * (q2 - p.q0 ** 2) ** 2)
    psiq = sf.dctn(psi0, type=2)
    nq = q2 * sf.dctn(p.beta * psi0 ** 3 - p.gamma * psi0 ** 5, type=2)
    # first step: Nq_past = Nq (pre-loop), so 3 Nq - Nq_past = 2 Nq
    new = s * ((1 + p.dt / 2 * L) * psiq + p.dt / 2 * (3 * nq - nq)) / (1 - p.dt / 2 * L)
    ref = sf.dctn(new, type=3)
    assert np.max(np.abs(o.psi - ref)) < 1e-12 * np.max(np.abs(ref))


# ---------------------------------------------------------------- advective --

def acparams(smadft, shape, model=None, **kw):
    p = smadft.Params.default(smadft.MODEL_ADV_CONSERVED if model is None else model, shape[0])
    p.nx, p.ny, p.nz = shape
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_adv_conserved_defaults(smadft):
    p = smadft.Params.default(smadft.MODEL_ADV_CONSERVED, 64)
    assert p.nw == 16 and p.dt == 0.0005 and p.nu > 0


def test_adv_conserved_planar_layers_follow_the_conserved_plain_step(smadft, oracle):
    """Planar layers have no flow, so the conserved advective step is the
    conserved plain step, except for the pre-loop history, which the
    advective variant leaves without |q|^2 (adv/cosAdvConSingle.cpp:755-790)."""
    shape = (16, 8, 32)
    p = acparams(smadft, shape, dt=0.002)
    dz = 2 * np.pi / (p.q0 * p.nw)
    z = np.arange(shape[2]) * dz
    psi0 = np.broadcast_to(1.2 * np.cos(p.q0 * z) * np.exp(-((z - z.mean()) / 8) ** 2), shape).copy()
    a = oracle.AdvOracle(p, psi0)
    a.step(12)
    assert np.abs(a.vel).max() < 1e-12
    pc = acparams(smadft, shape, model=smadft.MODEL_PLAIN_CONSERVED, dt=0.002, nw=p.nw)
    pl = oracle.PlainOracle(pc, psi0)
    pl.nq[...] = sf.dctn(p.beta * psi0 ** 3 - p.gamma * psi0 ** 5, type=2)  # unscaled history
    pl.step(12)
    assert np.max(np.abs(a.psi - pl.psi)) < 1e-12 * np.max(np.abs(psi0))


def test_adv_conserved_first_step_scipy(smadft, oracle):
    """One step from the pre-loop, restated with scipy transforms and the
    flow and gradient of the non-conserved advective step from the same psi
    (the force does not change between the variants)."""
    shape = (16, 16, 16)
    pc = acparams(smadft, shape)
    psi0 = smadft.ic_curved(pc, 1.328, 1e-2, seed=2)
    a = oracle.AdvOracle(pc, psi0)
    a.step(1)
    pn = acparams(smadft, shape, model=smadft.MODEL_ADV, dt=pc.dt, nw=pc.nw, eps=pc.eps)
    b = oracle.AdvOracle(pn, psi0)
    b.step(1)
    assert np.allclose(a.vel, b.vel, rtol=0, atol=1e-14 * np.abs(b.vel).max())
    assert np.abs(b.vel).max() > 1e-8
    s = 1.0 / (8 * np.prod(shape))
    q2 = q2_grid(pc)
    L = q2 * (pc.eps - pc.alpha * (q2 - pc.q0 ** 2) ** 2)
    th = pc.beta * psi0 ** 3 - pc.gamma * psi0 ** 5
    adv = (b.vel * b.grad).sum(axis=0)
    nq = q2 * sf.dctn(th, type=2) - sf.dctn(adv, type=2)
    nq_past = sf.dctn(th, type=2)
    psiq = s * sf.dctn(psi0, type=2)
    new = ((1 + pc.dt / 2 * L) * psiq + pc.dt / 2 * s * (3 * nq - nq_past)) / (1 - pc.dt / 2 * L)
    ref = sf.dctn(new, type=3)
    assert np.max(np.abs(a.psi - ref)) < 1e-12 * np.max(np.abs(ref))
