"""GPU parity of the discrete adjoint of the advective step (the linearised
psi^3/psi^5 terms, the force mu grad psi, the Stokes projection and u .
grad psi transposed) against the FP64 oracle adjoint, which is pinned by
central differences and dot products (tests/test_oracle_adv.py)."""
import dataclasses

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def setup(smadft, shape, prec=64):
    p = smadft.Params.config(2, shape[0])
    p.nx, p.ny, p.nz = shape
    p.precision = prec
    psi0 = smadft.ic_curved(p, 1.328, 1e-2, seed=2)
    tgt = np.random.default_rng(3).standard_normal(p.shape) * 0.1
    return p, psi0, tgt


@pytest.mark.parametrize("shape,K", [((16, 16, 16), 6), ((16, 32, 32), 10), ((32, 16, 64), 4)])
def test_adv_adjoint_fp64_matches_oracle(smadft, oracle, shape, K):
    p, psi0, tgt = setup(smadft, shape)
    psik_o = oracle.adv_forward(p, psi0, K)
    g = 2 * (psik_o - tgt)
    _, d0_o, dp_o = oracle.adv_adjoint(p, psi0, K, g)
    print(f"adv adjoint {shape} K={K}", flush=True)
    with smadft.Stepper(p, psi0) as s:
        s.forward_tape(K, K)
        psik = s.fetch()
        d0, dp = s.adjoint(g)
        after = s.fetch()
    assert rel_l2(psik, psik_o) < 1e-11
    assert rel_l2(after, psik_o) < 1e-11  # context left at step K
    assert rel_l2(d0, d0_o) < 1e-9
    assert np.allclose(dp, dp_o, rtol=1e-9, atol=1e-9 * np.abs(dp_o).max())


def test_adv_adjoint_central_differences_on_gpu(smadft):
    p, psi0, tgt = setup(smadft, (16, 16, 32))
    K = 8
    with smadft.Stepper(p, psi0) as s:
        s.forward_tape(K, K)
        d0, dp = s.adjoint(2 * (s.fetch() - tgt))
    h = 1e-6

    def J(q, x0):
        with smadft.Stepper(q, x0) as s:
            s.step(K)
            return float(np.sum((s.fetch() - tgt) ** 2))

    for i, name in enumerate(("eps", "alpha", "beta", "gamma")):
        v = getattr(p, name)
        fd = (J(dataclasses.replace(p, **{name: v + h}), psi0) - J(dataclasses.replace(p, **{name: v - h}), psi0)) / (2 * h)
        print(f"adv dJ/d{name}: adjoint {dp[i]:.9e}, FD {fd:.9e}", flush=True)
        assert dp[i] == pytest.approx(fd, rel=1e-6, abs=0), name
    u = d0 / np.linalg.norm(d0)
    fd = (J(p, psi0 + h * u) - J(p, psi0 - h * u)) / (2 * h)
    print(f"adv dJ/dpsi0 . u: adjoint {float(np.sum(d0 * u)):.9e}, FD {fd:.9e}", flush=True)
    assert float(np.sum(d0 * u)) == pytest.approx(fd, rel=1e-6, abs=0)


def test_adv_adjoint_fp32_close(smadft, oracle):
    p, psi0, tgt = setup(smadft, (16, 32, 32), prec=32)
    K = 10
    g = 2 * (oracle.adv_forward(p, psi0, K) - tgt)
    _, d0_o, dp_o = oracle.adv_adjoint(p, psi0, K, g)
    with smadft.Stepper(p, psi0) as s:
        s.forward_tape(K, K)
        d0, dp = s.adjoint(g)
    e0 = rel_l2(d0, d0_o)
    ep = np.max(np.abs(dp - dp_o) / np.abs(dp_o))
    print(f"adv adjoint FP32: dpsi0 {e0:.2e}, dtheta {ep:.2e}", flush=True)
    assert e0 < 5e-6 and ep < 2e-5


def test_adv_adjoint_needs_a_fresh_start(smadft):
    p, psi0, tgt = setup(smadft, (16, 16, 16))
    with smadft.Stepper(p, psi0) as s:
        s.step(2)
        with pytest.raises(smadft.SmadError) as e:
            s.forward_tape(4, 4)
        assert e.value.code == smadft.ESTATE


@pytest.mark.parametrize("prec", [64, 32])
def test_adv_adjoint_velocity_tape_is_exact(smadft, prec):
    """The velocity taped during the forward run (within the tape budget) and
    the velocity recomputed in the backward sweep are the same numbers, so the
    gradients are bit-identical whatever the budget; a budget of part of the
    run tapes the last steps only."""
    p, psi0, tgt = setup(smadft, (16, 32, 32), prec=prec)
    K = 6
    field = 16 * 32 * 32 * (8 if prec == 64 else 4)
    out = []
    for gib in (0.0, 2.5 * 3 * field / 2 ** 30, 100.0):
        with smadft.Stepper(p, psi0) as s:
            s.set_tape_budget(gib)
            s.forward_tape(K, K)
            psik = s.fetch()
            ux = s.fetch(smadft.FIELD_VEL_X)
            d0, dp = s.adjoint(2 * (psik - tgt))
            out.append((psik, ux, d0, dp))
    for o in out[1:]:
        for a, b in zip(o, out[0]):
            assert np.array_equal(a, b)
