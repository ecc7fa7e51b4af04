"""The benchmark configurations at their full sizes on the GPU: config 1's
plain step and config 2's advective step at 256^3, config 3's
quasi-incompressible step at 384^3 and config 4's adjoint at 256^3.

The oracle finishes a few plain and advective steps at 256^3 in seconds, so
those are compared with it directly. Everything else at full size is checked
through properties that do not depend on the size:

  * r2r transforms: REDFT01 o REDFT10 = RODFT01 o RODFT10 = 2N per axis, and
    linearity;
  * the plain step with beta = gamma = 0: every cosine mode decays by its
    Crank-Nicolson factor (1 + dt L/2)/(1 - dt L/2), independently of the
    others (no-adv/cosNoAdvTrackMult.cpp:778-800 with N = 0);
  * the quasi step: a run saved and restored through the full state continues
    bit-identically, and FP32-mixed stays close to FP64;
  * the adjoint: its gradients against central differences of the GPU's own
    loss (h = 1e-6 in the parameters, 1e-3 along a unit direction of psi0)."""
import dataclasses

import numpy as np
import pytest

from conftest import rel_l2
from test_oracle import cn_factor, cos_mode

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("n", [256, 384])
def test_r2r_round_trip_and_linearity_full_size(smadft, n, prec):
    rng = np.random.default_rng(n + prec)
    shape = (n, n, n)
    x = rng.standard_normal(shape)
    fwd = (smadft.REDFT10, smadft.RODFT10, smadft.REDFT10)
    inv = (smadft.REDFT01, smadft.RODFT01, smadft.REDFT01)
    print(f"r2r round trip {shape} fp{prec}", flush=True)
    X = smadft.r2r_3d(x, fwd, prec)
    back = smadft.r2r_3d(X, inv, prec)
    tol = 1e-13 if prec == 64 else 2e-6
    assert rel_l2(back, 8.0 * n ** 3 * x) < tol
    y = rng.standard_normal(shape)
    Y = smadft.r2r_3d(y, fwd, prec)
    Z = smadft.r2r_3d(2.0 * x - 0.5 * y, fwd, prec)
    assert rel_l2(Z, 2.0 * X - 0.5 * Y) < tol


def test_plain_config1_256_matches_oracle(smadft, oracle):
    p = smadft.Params.config(1, 256)
    psi0 = smadft.config_ic(1, p)
    print("plain config 1 256^3, 3 steps vs oracle", flush=True)
    o = oracle.PlainOracle(p, psi0)
    o.step(3)
    with smadft.Stepper(p, psi0) as s:
        s.step(3)
        psi = s.fetch()
        nhat = s.fetch(smadft.FIELD_N_HAT)
    assert rel_l2(psi, o.psi) < 1e-12
    assert rel_l2(nhat, o.nq) < 1e-12


@pytest.mark.parametrize("prec", [64, 32])
def test_plain_linear_modes_256(smadft, prec):
    """beta = gamma = 0 on config 1's grid: a sum of cosine modes, each
    decaying by its own factor G(q) over 50 steps."""
    p = smadft.Params.config(1, 256)
    p = dataclasses.replace(p, beta=0.0, gamma=0.0, precision=prec)
    shape = (256, 256, 256)
    modes = [((0, 0, 0), 0.3), ((8, 0, 8), 1.0), ((5, 7, 3), -0.7), ((40, 1, 200), 0.5), ((255, 128, 17), 0.2)]
    psi0 = sum(a * cos_mode(shape, *m) for m, a in modes)
    K = 50
    want = sum(a * cn_factor(p, *m) ** K * cos_mode(shape, *m) for m, a in modes)
    print(f"plain linear modes 256^3 fp{prec}", flush=True)
    with smadft.Stepper(p, psi0) as s:
        s.step(K)
        psi = s.fetch()
    assert rel_l2(psi, want) < (1e-12 if prec == 64 else 1e-5)


def test_adv_config2_256_matches_oracle(smadft, oracle):
    p = smadft.Params.config(2, 256)
    psi0 = smadft.config_ic(2, p)
    print("adv config 2 256^3, 2 steps vs oracle", flush=True)
    o = oracle.AdvOracle(p, psi0)
    o.step(2)
    with smadft.Stepper(p, psi0) as s:
        s.step(2)
        psi = s.fetch()
        vel = [s.fetch(f) for f in (smadft.FIELD_VEL_X, smadft.FIELD_VEL_Y, smadft.FIELD_VEL_Z)]
    assert np.abs(o.vel).max() > 1e-6  # the bumps drive a flow
    assert rel_l2(psi, o.psi) < 1e-12
    assert max(rel_l2(vel[a], o.vel[a]) for a in range(3)) < 1e-10


def test_quasi_config3_384_restart_and_precision(smadft):
    """Config 3 on its 384^3 grid with the divv switch at step 2, so that the
    post-switch step and its coupling term run: saved after 3 steps and
    restored into a fresh context, it continues bit-identically to 6
    uninterrupted steps; FP32-mixed stays within 1e-5 of FP64 in psi and rho."""
    p = smadft.Params.config(3, 384)
    p.div_switch = 2
    psi0 = smadft.config_ic(3, p)
    K, M = 3, 3
    fields = (smadft.FIELD_PSI, smadft.FIELD_RHO, smadft.FIELD_PRESSURE, smadft.FIELD_VEL_X)
    print("quasi config 3 384^3 fp64, continuous", flush=True)
    with smadft.Stepper(p, psi0) as a:
        a.step(K + M)
        want = [a.fetch(f) for f in fields]
    print("quasi config 3 384^3 fp64, save and restore", flush=True)
    with smadft.Stepper(p, psi0) as b:
        b.step(K)
        state = b.save_state()
    with smadft.Stepper(p, np.zeros(p.shape)) as c:
        c.load_state(state)
        del state
        c.step(M)
        for f, w in zip(fields, want):
            assert np.array_equal(c.fetch(f), w), f
    print("quasi config 3 384^3 fp32", flush=True)
    with smadft.Stepper(dataclasses.replace(p, precision=smadft.FP32_MIXED), psi0) as d:
        d.step(K + M)
        psi32, rho32 = d.fetch(smadft.FIELD_PSI), d.fetch(smadft.FIELD_RHO)
    assert rel_l2(psi32, want[0]) < 1e-5
    assert rel_l2(rho32, want[1]) < 1e-5


def test_config4_adjoint_256_against_central_differences(smadft):
    """Config 4 at 256^3: K = 200 steps checkpointed every 20 (strict
    recomputation), the loss |psi_K - psi_target|^2 with psi_target from
    eps = -0.65, alpha = 1.1; dJ/deps, dJ/dalpha (h = 1e-6) and dJ/dpsi0
    along a random unit direction (h = 1e-3) against central differences of
    the GPU loss.
    The FP32-mixed adjoint with the same cotangent stays within 2e-5 of the
    FP64 one."""
    K, ck, h = 200, 20, 1e-6
    p = smadft.Params.config(4, 256)
    psi0 = smadft.config_ic(4, p)
    print("config 4 256^3: target", flush=True)
    with smadft.Stepper(dataclasses.replace(p, eps=-0.65, alpha=1.1), psi0) as t:
        t.step(K)
        target = t.fetch()

    def J(q, x0):
        with smadft.Stepper(q, x0) as s:
            s.step(K)
            return float(np.sum((s.fetch() - target) ** 2))

    print("config 4 256^3: adjoint fp64", flush=True)
    with smadft.Stepper(p, psi0) as s:
        s.set_tape_budget(0.0)
        s.forward_tape(K, ck)
        g = 2 * (s.fetch() - target)
        d0, dp = s.adjoint(g)
    for i, name in enumerate(("eps", "alpha")):
        v = getattr(p, name)
        fd = (J(dataclasses.replace(p, **{name: v + h}), psi0) - J(dataclasses.replace(p, **{name: v - h}), psi0)) / (2 * h)
        print(f"config 4 256^3 dJ/d{name}: adjoint {dp[i]:.9e}, FD {fd:.9e}", flush=True)
        assert dp[i] == pytest.approx(fd, rel=1e-6, abs=0), name
    # along a unit direction, h = 1e-3: a step of 1e-6 perturbs psi0 (norm ~1700)
    # by ~6e-10 of itself, below what the loss's own rounding resolves
    v = np.random.default_rng(4).standard_normal(psi0.shape)
    v /= np.linalg.norm(v)
    hv = 1e-3
    fd = (J(p, psi0 + hv * v) - J(p, psi0 - hv * v)) / (2 * hv)
    print(f"config 4 256^3 dJ/dpsi0 . v: adjoint {float(np.sum(d0 * v)):.9e}, FD {fd:.9e}", flush=True)
    assert float(np.sum(d0 * v)) == pytest.approx(fd, rel=1e-6, abs=0)
    print("config 4 256^3: adjoint fp32", flush=True)
    with smadft.Stepper(dataclasses.replace(p, precision=smadft.FP32_MIXED), psi0) as s:
        s.set_tape_budget(0.0)
        s.forward_tape(K, ck)
        d0f, dpf = s.adjoint(g)
    assert rel_l2(d0f, d0) < 2e-5
    assert np.allclose(dpf, dp, rtol=2e-5, atol=0.0)
