"""GPU parity of the plain-model step (no-adv/cosNoAdvTrackMult.cpp:729-805)
against the oracle, plus closed-form and fault-reporting checks."""
import numpy as np
import pytest

from conftest import rel_l2
from test_oracle import cn_factor, cos_mode

pytestmark = pytest.mark.gpu


def params(smadft, shape, prec=64, **kw):
    p = smadft.Params.default(0, shape[0])
    p.nx, p.ny, p.nz = shape
    p.precision = prec
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("shape,nsteps", [((32, 32, 32), 100), ((16, 32, 64), 60), ((64, 64, 64), 400),
                                          ((1024, 16, 32), 30), ((32, 16, 1024), 30), ((16, 16, 512), 30),
                                          ((384, 16, 32), 30), ((32, 384, 16), 30), ((16, 32, 384), 30),
                                          ((578, 16, 16), 20), ((16, 484, 16), 20), ((16, 16, 130), 30),
                                          ((130, 32, 578), 10)])
def test_plain_fp64_matches_oracle(smadft, oracle, shape, nsteps):
    p = params(smadft, shape)
    psi0 = smadft.ic_slab(p, 0.6, 1e-3, seed=0)
    o = oracle.PlainOracle(p, psi0)
    o.step(nsteps)
    with smadft.Stepper(p, psi0) as s:
        s.step(nsteps)
        psi = s.fetch(smadft.FIELD_PSI)
        nhat = s.fetch(smadft.FIELD_N_HAT)
    assert rel_l2(psi, o.psi) < 1e-10
    assert rel_l2(nhat, o.nq) < 1e-10


def test_plain_fp64_1000_steps_32(smadft, oracle):
    p = params(smadft, (32, 32, 32))
    psi0 = smadft.ic_curved(p, 0.6, 1e-3, seed=1)
    o = oracle.PlainOracle(p, psi0)
    o.step(1000)
    with smadft.Stepper(p, psi0) as s:
        s.step(1000)
        psi = s.fetch()
    assert rel_l2(psi, o.psi) < 1e-10


@pytest.mark.parametrize("shape", [(1024, 16, 32), (16, 32, 1024), (384, 16, 32), (16, 32, 384)])
def test_plain_fp32_mixed_1024_lines(smadft, oracle, shape):
    p = params(smadft, shape, prec=32)
    psi0 = smadft.ic_slab(p, 1.328, 1e-3, seed=0)
    o = oracle.PlainOracle(p, psi0)
    o.step(30)
    with smadft.Stepper(p, psi0) as s:
        s.step(30)
        psi = s.fetch()
    assert rel_l2(psi, o.psi) < 1e-5


def test_plain_fp32_mixed_error(smadft, oracle):
    # reference amplitude 1.328 (no-adv/cosNoAdvTrackSing.cpp:389): the
    # smectic persists, so the relative error is taken on an O(1) field
    p = params(smadft, (64, 64, 64), prec=32)
    psi0 = smadft.ic_slab(p, 1.328, 1e-3, seed=0)
    o = oracle.PlainOracle(p, psi0)
    o.step(300)
    with smadft.Stepper(p, psi0) as s:
        s.step(300)
        psi = s.fetch()
    assert rel_l2(psi, o.psi) < 1e-5


def test_plain_linear_mode_gpu(smadft):
    shape = (32, 16, 64)
    p = params(smadft, shape, beta=0.0, gamma=0.0, eps=-0.3, dt=0.2)
    mode = (3, 2, 5)
    psi0 = cos_mode(shape, *mode)
    with smadft.Stepper(p, psi0) as s:
        s.step(20)
        psi = s.fetch()
    g = cn_factor(p, *mode)
    assert rel_l2(psi, g ** 20 * psi0) < 1e-12


def test_l1_matches_oracle(smadft, oracle):
    p = params(smadft, (32, 32, 32))
    psi0 = smadft.ic_slab(p, 0.6, 1e-3, seed=0)
    o = oracle.PlainOracle(p, psi0)
    o.step(99)
    before = o.psi.copy()
    o.step(1)
    ref = oracle.l1_sums(o.psi, before)
    with smadft.Stepper(p, psi0) as s:
        s.step(99)
        s.step(1)
        red = s.reduce(smadft.RED_L1)
    assert red[0] == pytest.approx(ref[0], rel=1e-9)
    assert red[1] == pytest.approx(ref[1], rel=1e-12)
    assert red[2] == pytest.approx(ref[0] / (p.dt * ref[1]), rel=1e-9)


def test_nonfinite_is_reported(smadft):
    p = params(smadft, (16, 16, 16))
    psi0 = smadft.ic_slab(p, 0.6, 1e-3, seed=0)
    with smadft.Stepper(p, psi0) as s:
        bad = psi0.copy()
        bad[3, 4, 5] = np.nan
        s.set_field(smadft.FIELD_PSI, bad)
        s.step(2)
        with pytest.raises(smadft.SmadError) as e:
            s.reduce(smadft.RED_L1)
        assert e.value.code == smadft.ENONFINITE
        with pytest.raises(smadft.SmadError) as e:
            s.fetch()
        assert e.value.code == smadft.ENONFINITE


def test_restart_equals_continuous(smadft):
    # set_field(PSI) re-runs the pre-loop (first step Euler again), as the
    # reference's load == 1 restart does; compare with the oracle doing the same
    p = params(smadft, (32, 16, 32))
    psi0 = smadft.ic_slab(p, 0.6, 1e-3, seed=4)
    with smadft.Stepper(p, psi0) as s:
        s.step(10)
        mid = s.fetch()
        s.set_field(smadft.FIELD_PSI, mid)
        s.step(10)
        gpu = s.fetch()
    import oracle as orc
    o = orc.PlainOracle(p, mid)
    o.step(10)
    assert rel_l2(gpu, o.psi) < 1e-10
