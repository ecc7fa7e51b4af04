"""GPU parity of the advective-model step (adv/cosAdvTrackMult.cpp:827-1369)
against the oracle."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def adv_params(smadft, shape, prec=64, **kw):
    p = smadft.Params.config(2, shape[0])  # advective, nw = 8, dt = 0.01
    p.nx, p.ny, p.nz = shape
    p.precision = prec
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("shape,nsteps", [((32, 32, 32), 40), ((16, 32, 64), 25), ((384, 16, 32), 20),
                                          ((16, 384, 32), 20), ((32, 16, 384), 20), ((130, 16, 32), 15),
                                          ((16, 484, 16), 10), ((16, 16, 578), 10)])
def test_adv_fp64_matches_oracle(smadft, oracle, shape, nsteps):
    p = adv_params(smadft, shape)
    psi0 = smadft.ic_curved(p, 1.328, 1e-3, seed=2)
    o = oracle.AdvOracle(p, psi0)
    o.step(nsteps)
    with smadft.Stepper(p, psi0) as s:
        s.step(nsteps)
        psi = s.fetch(smadft.FIELD_PSI)
        vel = [s.fetch(f) for f in (smadft.FIELD_VEL_X, smadft.FIELD_VEL_Y, smadft.FIELD_VEL_Z)]
        grad = [s.fetch(f) for f in (smadft.FIELD_GRAD_X, smadft.FIELD_GRAD_Y, smadft.FIELD_GRAD_Z)]
    assert rel_l2(psi, o.psi) < 1e-10
    for c in range(3):
        assert rel_l2(grad[c], o.grad[c]) < 1e-10
        assert rel_l2(vel[c], o.vel[c]) < 1e-9
    assert np.abs(o.vel).max() > 1e-8  # the flow is not trivially zero


@pytest.mark.parametrize("shape", [(32, 32, 32), (32, 16, 384)])
def test_adv_fp32_mixed_error(smadft, oracle, shape):
    p = adv_params(smadft, shape, prec=32)
    psi0 = smadft.ic_curved(p, 1.328, 1e-3, seed=2)
    o = oracle.AdvOracle(p, psi0)
    o.step(40)
    with smadft.Stepper(p, psi0) as s:
        s.step(40)
        psi = s.fetch()
    assert rel_l2(psi, o.psi) < 1e-5


def test_adv_l1(smadft, oracle):
    p = adv_params(smadft, (16, 16, 32))
    psi0 = smadft.ic_curved(p, 1.328, 1e-3, seed=2)
    o = oracle.AdvOracle(p, psi0)
    o.step(9)
    before = o.psi.copy()
    o.step(1)
    ref = oracle.l1_sums(o.psi, before)
    with smadft.Stepper(p, psi0) as s:
        s.step(9)
        s.step(1)
        red = s.reduce(smadft.RED_L1)
    assert red[0] == pytest.approx(ref[0], rel=1e-8)
    assert red[1] == pytest.approx(ref[1], rel=1e-12)
