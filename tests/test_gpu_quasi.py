"""GPU parity of the quasi-incompressible step (quasi/quasi.cpp:1099-2251)
against the oracle, across the divv switch."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def qparams(smadft, shape, prec=64, **kw):
    p = smadft.Params.config(3, shape[0])  # quasi, rho0 = 0.9, kp = 0.1/1.34164
    p.nx, p.ny, p.nz = shape
    p.precision = prec
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("shape,nsteps,switch", [((32, 32, 32), 14, 5), ((16, 32, 64), 9, 3),
                                                 ((384, 16, 32), 6, 2), ((32, 16, 384), 6, 2),
                                                 ((578, 16, 16), 4, 2), ((16, 484, 16), 4, 2), ((16, 16, 130), 5, 2)])
def test_quasi_fp64_matches_oracle(smadft, oracle, shape, nsteps, switch):
    p = qparams(smadft, shape, div_switch=switch)
    psi0 = smadft.ic_curved(p, 1.34164, 1e-3, seed=3)
    q = oracle.QuasiOracle(p, psi0)
    q.step(nsteps)
    with smadft.Stepper(p, psi0) as s:
        s.step(nsteps)
        psi = s.fetch(smadft.FIELD_PSI)
        fields = {f: s.fetch(f) for f in (smadft.FIELD_VEL_X, smadft.FIELD_VEL_Y, smadft.FIELD_VEL_Z,
                                          smadft.FIELD_PRESSURE, smadft.FIELD_RHO, smadft.FIELD_ENERGY)}
        l1 = s.reduce(smadft.RED_L1)
        mass = s.reduce(smadft.RED_MASS)[0]
        energy = s.reduce(smadft.RED_ENERGY)[0]
    assert rel_l2(psi, q.psi) < 1e-10
    assert rel_l2(fields[smadft.FIELD_VEL_X], q.vel[0]) < 1e-9
    assert rel_l2(fields[smadft.FIELD_VEL_Y], q.vel[1]) < 1e-9
    assert rel_l2(fields[smadft.FIELD_VEL_Z], q.vel[2]) < 1e-9
    assert rel_l2(fields[smadft.FIELD_PRESSURE], q.p) < 1e-9
    assert rel_l2(fields[smadft.FIELD_RHO], q.rho) < 1e-12
    assert rel_l2(fields[smadft.FIELD_ENERGY], q.energy) < 1e-10
    assert l1[0] == pytest.approx(q.sums[0], rel=1e-8)
    assert l1[1] == pytest.approx(q.sums[1], rel=1e-10)
    assert mass == pytest.approx(q.sums[2], rel=1e-12)
    assert energy == pytest.approx(q.sums[3], rel=1e-9)


@pytest.mark.parametrize("shape", [(32, 32, 32), (32, 16, 384)])
def test_quasi_fp32_mixed_error(smadft, oracle, shape):
    p = qparams(smadft, shape, prec=32, div_switch=5)
    psi0 = smadft.ic_curved(p, 1.34164, 1e-3, seed=3)
    q = oracle.QuasiOracle(p, psi0)
    q.step(14)
    with smadft.Stepper(p, psi0) as s:
        s.step(14)
        psi = s.fetch()
        pr = s.fetch(smadft.FIELD_PRESSURE)
    assert rel_l2(psi, q.psi) < 1e-5
    # the pressure comes from S = mu grad psi + e grad rho, i.e. up to fourth
    # derivatives of FP32 psi; its FP32 error is reported, bounded loosely
    assert rel_l2(pr, q.p) < 1e-3


def test_quasi_fp32_velocity_and_pressure_after_the_switch(smadft, oracle):
    """In FP32 storage the density's change over a step is formed from the
    transforms of psi_hat^n - psi_hat^{n-1} (k_q_drho_z), so (rho - rho_old)/dt
    in divv (dt = 1e-3) no longer divides the FP32 transforms' rounding of rho
    by dt: the velocity and the pressure follow the FP64 oracle closely over
    hundreds of steps after the switch."""
    p = smadft.Params.config(3, 32)
    p.precision = 32
    p.div_switch = 50
    psi0 = smadft.config_ic(3, p)
    q = oracle.QuasiOracle(p, psi0)
    q.step(300)
    with smadft.Stepper(p, psi0) as s:
        s.step(300)
        got = [s.fetch(f) for f in (smadft.FIELD_VEL_X, smadft.FIELD_VEL_Y, smadft.FIELD_VEL_Z,
                                    smadft.FIELD_PRESSURE, smadft.FIELD_PSI)]
    errs = [rel_l2(g, r) for g, r in zip(got, list(q.vel) + [q.p, q.psi])]
    print(f"quasi FP32, config 3 32^3, 300 steps (switch at 50): ux uy uz p psi {errs}", flush=True)
    # measured: u 1.2e-5 - 5.4e-5, p 1.1e-5, psi 1.0e-6 (the difference of two
    # FP32 densities gave u and p errors near 1e-3 at this horizon)
    assert max(errs[:3]) < 2e-4 and errs[3] < 1e-4 and errs[4] < 1e-5, errs
