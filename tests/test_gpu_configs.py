"""The benchmark configurations 2-4 with their own initial conditions and
parameters, on the GPU against the oracle (configs 0 and 1 are the plain
model: tests/test_gpu_long.py runs config 0 at full size, the bench config 1).

  * config 2: the advective model from two seeded bumped interfaces;
  * config 3: the quasi model from a seeded tilted planar slab, across the
    divv switch and the coupling term;
  * config 4: the adjoint of K = 200 config-1 steps with checkpoints every 20
    for the loss |psi_K - psi_target|^2, psi_target from the same initial
    condition at eps = -0.65, alpha = 1.1; against the oracle adjoint and
    central differences (h = 1e-6)."""
import dataclasses

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [64, 32])
def test_config2_two_bumps_matches_oracle(smadft, oracle, prec):
    p = smadft.Params.config(2, 32)
    p.precision = prec
    psi0 = smadft.config_ic(2, p)
    o = oracle.AdvOracle(p, psi0)
    o.step(100)
    with smadft.Stepper(p, psi0) as s:
        s.step(100)
        psi = s.fetch()
        vel = [s.fetch(f) for f in (smadft.FIELD_VEL_X, smadft.FIELD_VEL_Y, smadft.FIELD_VEL_Z)]
    e_psi = rel_l2(psi, o.psi)
    e_u = max(rel_l2(vel[a], o.vel[a]) for a in range(3))
    print(f"config 2, 32^3, 100 steps, fp{prec}: psi {e_psi:.2e}, u {e_u:.2e}", flush=True)
    assert np.abs(o.vel).max() > 1e-4  # the bumps drive a flow
    if prec == 64:
        assert e_psi < 1e-10 and e_u < 1e-9
    else:
        assert e_psi < 1e-5 and e_u < 1e-3


@pytest.mark.parametrize("prec", [64, 32])
def test_config3_tilted_slab_matches_oracle(smadft, oracle, prec):
    p = smadft.Params.config(3, 32)
    p.precision = prec
    p.div_switch = 10
    psi0 = smadft.config_ic(3, p)
    q = oracle.QuasiOracle(p, psi0)
    q.step(30)
    with smadft.Stepper(p, psi0) as s:
        s.step(30)
        got = {"psi": s.fetch(), "rho": s.fetch(smadft.FIELD_RHO), "p": s.fetch(smadft.FIELD_PRESSURE),
               "ux": s.fetch(smadft.FIELD_VEL_X), "uz": s.fetch(smadft.FIELD_VEL_Z)}
    ref = {"psi": q.psi, "rho": q.rho, "p": q.p, "ux": q.vel[0], "uz": q.vel[2]}
    errs = {k: rel_l2(got[k], ref[k]) for k in got}
    print(f"config 3, 32^3, 30 steps (switch at 10), fp{prec}: {errs}", flush=True)
    if prec == 64:
        assert all(e < 1e-9 for e in errs.values()), errs
    else:
        assert errs["psi"] < 1e-5 and errs["rho"] < 1e-5, errs


def config4_case(smadft, oracle, n, K=200):
    p = smadft.Params.config(4, n)
    psi0 = smadft.config_ic(4, p)
    pt = dataclasses.replace(p, eps=-0.65, alpha=1.1)
    target = oracle.plain_forward(pt, psi0, K)
    return p, psi0, target


def gpu_adjoint(smadft, p, psi0, target, K, ck, budget=0.0):
    with smadft.Stepper(p, psi0) as s:
        s.set_tape_budget(budget)
        s.forward_tape(K, ck)
        psik = s.fetch()
        d0, dp = s.adjoint(2 * (psik - target))
    return psik, d0, dp


def test_config4_adjoint_matches_oracle_and_fd(smadft, oracle):
    K, ck = 200, 20
    p, psi0, target = config4_case(smadft, oracle, 32, K)
    psik_o = oracle.plain_forward(p, psi0, K)
    _, d0_o, dp_o = oracle.plain_adjoint(p, psi0, K, 2 * (psik_o - target))
    psik, d0, dp = gpu_adjoint(smadft, p, psi0, target, K, ck)
    assert rel_l2(psik, psik_o) < 1e-11
    assert rel_l2(d0, d0_o) < 1e-9
    assert np.allclose(dp, dp_o, rtol=1e-9, atol=1e-9 * np.abs(dp_o).max())
    h = 1e-6

    def J(q, x0):
        with smadft.Stepper(q, x0) as s:
            s.step(K)
            return float(np.sum((s.fetch() - target) ** 2))

    for i, name in enumerate(("eps", "alpha", "beta", "gamma")):
        v = getattr(p, name)
        fd = (J(dataclasses.replace(p, **{name: v + h}), psi0) - J(dataclasses.replace(p, **{name: v - h}), psi0)) / (2 * h)
        print(f"config 4 dJ/d{name}: adjoint {dp[i]:.9e}, FD {fd:.9e}", flush=True)
        assert dp[i] == pytest.approx(fd, rel=1e-6, abs=1e-9 * np.abs(dp).max()), name
    # along a unit direction, h = 1e-3: a step of 1e-6 perturbs psi0 by ~1e-8
    # of its norm, where the loss's own rounding shows at ~1e-6 relative
    v = np.random.default_rng(4).standard_normal(psi0.shape)
    v /= np.linalg.norm(v)
    hv = 1e-3
    fd = (J(p, psi0 + hv * v) - J(p, psi0 - hv * v)) / (2 * hv)
    print(f"config 4 dJ/dpsi0 . v: adjoint {float(np.sum(d0 * v)):.9e}, FD {fd:.9e}", flush=True)
    assert float(np.sum(d0 * v)) == pytest.approx(fd, rel=1e-6, abs=0)


def test_config4_adjoint_128_matches_oracle(smadft, oracle):
    """Configuration 4 at 128^3 in FP64 (the oracle adjoint holds all 201
    states on the host, ~6.5 GB), checkpointed with strict recomputation and
    with every segment taped (the gradients are identical)."""
    K, ck = 200, 20
    p, psi0, target = config4_case(smadft, oracle, 128, K)
    print("config 4 at 128^3: target ready", flush=True)
    psik, d0, dp = gpu_adjoint(smadft, p, psi0, target, K, ck)
    _, d0t, dpt = gpu_adjoint(smadft, p, psi0, target, K, ck, budget=64.0)
    assert np.array_equal(d0, d0t) and np.array_equal(dp, dpt)
    _, d0_o, dp_o = oracle.plain_adjoint(p, psi0, K, 2 * (psik - target))
    print("config 4 at 128^3: oracle adjoint done", flush=True)
    assert rel_l2(d0, d0_o) < 1e-9
    assert np.allclose(dp, dp_o, rtol=1e-9, atol=1e-9 * np.abs(dp_o).max())
