"""GPU parity of the conserved plain variant (no-adv/cosNoAdvConserv.cpp)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def setup(smadft, n=32, prec=64):
    p = smadft.Params.default(smadft.MODEL_PLAIN_CONSERVED, n)
    p.precision = prec
    psi0 = 0.1 + smadft.ic_curved(p, 1.328, 1e-3, seed=3)
    return p, psi0


@pytest.mark.parametrize("shape", [(32, 32, 32), (16, 32, 64)])
def test_conserved_fp64_matches_oracle(smadft, oracle, shape):
    p, _ = setup(smadft)
    p.nx, p.ny, p.nz = shape
    psi0 = 0.1 + smadft.ic_curved(p, 1.328, 1e-3, seed=3)
    o = oracle.PlainOracle(p, psi0)
    o.step(60)
    with smadft.Stepper(p, psi0) as s:
        s.step(60)
        psi = s.fetch()
        l1 = s.reduce(smadft.RED_L1)
    assert rel_l2(psi, o.psi) < 1e-10
    assert abs(psi.mean() - psi0.mean()) < 1e-13
    assert l1[0] > 0


def test_conserved_fp32_mixed(smadft, oracle):
    p, psi0 = setup(smadft, prec=32)
    o = oracle.PlainOracle(p, psi0)
    o.step(60)
    with smadft.Stepper(p, psi0) as s:
        s.step(60)
        psi = s.fetch()
    assert rel_l2(psi, o.psi) < 1e-5


def test_conserved_has_no_adjoint_or_surface(smadft):
    p, psi0 = setup(smadft, n=16)
    with smadft.Stepper(p, psi0) as s:
        for call in (lambda: s.forward_tape(4, 2), s.surface):
            with pytest.raises(smadft.SmadError) as e:
                call()
            assert e.value.code == smadft.EINVAL


@pytest.mark.parametrize("shape,nsteps", [((32, 32, 32), 30), ((16, 32, 64), 20)])
def test_adv_conserved_fp64_matches_oracle(smadft, oracle, shape, nsteps):
    p = smadft.Params.default(smadft.MODEL_ADV_CONSERVED, shape[0])
    p.nx, p.ny, p.nz = shape
    psi0 = smadft.ic_curved(p, 1.328, 1e-3, seed=2)
    o = oracle.AdvOracle(p, psi0)
    o.step(nsteps)
    with smadft.Stepper(p, psi0) as s:
        s.step(nsteps)
        psi = s.fetch()
        vel = [s.fetch(f) for f in (smadft.FIELD_VEL_X, smadft.FIELD_VEL_Y, smadft.FIELD_VEL_Z)]
    assert rel_l2(psi, o.psi) < 1e-10
    for c in range(3):
        assert rel_l2(vel[c], o.vel[c]) < 1e-9
    assert np.abs(o.vel).max() > 1e-8


def test_adv_conserved_fp32_mixed(smadft, oracle):
    p = smadft.Params.default(smadft.MODEL_ADV_CONSERVED, 32)
    p.precision = 32
    psi0 = smadft.ic_curved(p, 1.328, 1e-3, seed=2)
    o = oracle.AdvOracle(p, psi0)
    o.step(30)
    with smadft.Stepper(p, psi0) as s:
        s.step(30)
        psi = s.fetch()
    assert rel_l2(psi, o.psi) < 1e-5
