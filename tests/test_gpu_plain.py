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


def test_history_fields_match_oracle(smadft, oracle):
    """The step carries w = REDFT01_y REDFT01_x(psi_hat) and the combined
    history H = C1 psi_hat - (dt/2) s N_hat_past; psi_hat and N_hat are
    recovered from them on fetch."""
    p = params(smadft, (32, 16, 64))
    psi0 = smadft.ic_curved(p, 1.328, 1e-3, seed=3)
    o = oracle.PlainOracle(p, psi0)
    o.step(40)
    with smadft.Stepper(p, psi0) as s:
        ph0 = s.fetch(smadft.FIELD_PSI_HAT)  # pre-loop: exact
        nq0 = s.fetch(smadft.FIELD_N_HAT)
        s.step(40)
        ph = s.fetch(smadft.FIELD_PSI_HAT)
        nq = s.fetch(smadft.FIELD_N_HAT)
    scale = 0.125 / psi0.size
    kinds = (oracle.REDFT10,) * 3
    assert rel_l2(ph0, scale * oracle.r2r3d(psi0, kinds)) < 1e-13
    assert rel_l2(ph, scale * oracle.r2r3d(o.psi, kinds)) < 1e-10
    assert rel_l2(nq, o.nq) < 1e-10
    o0 = oracle.PlainOracle(p, psi0)
    assert rel_l2(nq0, o0.nq) < 1e-13


@pytest.mark.parametrize("prec", [64, 32])
def test_set_params_mid_run_keeps_psi_and_history(smadft, oracle, prec):
    """smad_set_params on a running context re-expresses the carried history
    for the new coefficients: the run continues from the same psi and AB2
    history, as the oracle does when its parameters change between steps."""
    pa = params(smadft, (32, 32, 16), prec=prec)
    pb = params(smadft, (32, 32, 16), prec=prec, eps=-0.65, alpha=1.1, beta=2.1, dt=0.05)
    psi0 = smadft.ic_curved(pa, 1.328, 1e-3, seed=5)
    o = oracle.PlainOracle(pa, psi0)
    o.step(30)
    nq_a = o.nq.copy()
    o.op = oracle.oparams(pb)
    o.step(30)
    with smadft.Stepper(pa, psi0) as s:
        s.step(30)
        s.set_params(pb)
        # the history is still the nonlinearity formed with the old beta and
        # gamma (recovered from H, no longer from psi_prev)
        nq = s.fetch(smadft.FIELD_N_HAT)
        s.step(30)
        psi = s.fetch()
    assert rel_l2(psi, o.psi) < (1e-10 if prec == 64 else 1e-5)
    if prec == 64:
        assert rel_l2(nq, nq_a) < 1e-8


def test_set_psi_hat_and_n_hat_restart(smadft):
    """The classic state (psi_hat, N_hat) set on a fresh context continues
    the run: setting psi_hat keeps the history, setting N_hat keeps psi_hat."""
    p = params(smadft, (16, 32, 32))
    psi0 = smadft.ic_curved(p, 1.328, 1e-3, seed=6)
    with smadft.Stepper(p, psi0) as a:
        a.step(25)
        ph, nq = a.fetch(smadft.FIELD_PSI_HAT), a.fetch(smadft.FIELD_N_HAT)
        a.step(25)
        want = a.fetch()
    for order in ((smadft.FIELD_PSI_HAT, smadft.FIELD_N_HAT), (smadft.FIELD_N_HAT, smadft.FIELD_PSI_HAT)):
        with smadft.Stepper(p, np.zeros(p.shape)) as b:
            for f in order:
                b.set_field(f, ph if f == smadft.FIELD_PSI_HAT else nq)
            b.step(25)
            assert rel_l2(b.fetch(), want) < 1e-12
