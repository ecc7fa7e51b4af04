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
