"""GPU parity of the checkpointed discrete adjoint of the plain step against
the FP64 oracle adjoint (which is itself pinned by central differences)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def setup(smadft, shape=(16, 32, 32), prec=64):
    p = smadft.Params.default(0, shape[0])
    p.nx, p.ny, p.nz = shape
    p.precision = prec
    psi0 = smadft.ic_slab(p, 1.328, 1e-2, seed=3)
    tgt = np.random.default_rng(0).standard_normal(psi0.shape) * 0.1
    return p, psi0, tgt


@pytest.mark.parametrize("K,ck", [(24, 5), (10, 10), (1, 1), (13, 4)])
def test_adjoint_fp64_matches_oracle(smadft, oracle, K, ck):
    p, psi0, tgt = setup(smadft)
    psik_o = oracle.plain_forward(p, psi0, K)
    g = 2 * (psik_o - tgt)
    _, d0_o, dp_o = oracle.plain_adjoint(p, psi0, K, g)
    with smadft.Stepper(p, psi0) as s:
        s.forward_tape(K, ck)
        psik = s.fetch()
        d0, dp = s.adjoint(g)
        after = s.fetch()
    assert rel_l2(psik, psik_o) < 1e-11
    assert rel_l2(after, psik_o) < 1e-11  # context left at step K
    assert rel_l2(d0, d0_o) < 1e-9
    for q in range(4):
        assert dp[q] == pytest.approx(dp_o[q], rel=1e-9, abs=1e-9 * np.abs(dp_o).max())


@pytest.mark.parametrize("gib", [0.0026, 1.0])
def test_adjoint_tape_budget_bit_identical(smadft, gib):
    # taping more segments replaces recomputation by stored copies of the
    # same values: forward state and gradients must not change by one bit
    p, psi0, tgt = setup(smadft)
    K, ck = 24, 5
    out = {}
    for b in (0.0, gib):
        with smadft.Stepper(p, psi0) as s:
            s.set_tape_budget(b)
            s.forward_tape(K, ck)
            psik = s.fetch()
            d0, dp = s.adjoint(2 * (psik - tgt))
            out[b] = (psik, d0, dp, s.fetch())
    for a, b in zip(out[0.0], out[gib]):
        assert np.array_equal(a, b)


def test_adjoint_fp32_mixed_close_to_oracle(smadft, oracle):
    p, psi0, tgt = setup(smadft, prec=32)
    K = 20
    g = 2 * (oracle.plain_forward(p, psi0, K) - tgt)
    _, d0_o, dp_o = oracle.plain_adjoint(p, psi0, K, g)
    with smadft.Stepper(p, psi0) as s:
        s.forward_tape(K, 5)
        d0, dp = s.adjoint(g)
    # FP32 storage, FP64 CN coefficients and lambda recurrence: measured
    # 1.3e-6 (dpsi0) and <= 9.2e-7 (theta) on this case
    assert rel_l2(d0, d0_o) < 3e-6
    assert np.max(np.abs(dp - dp_o) / np.abs(dp_o)) < 3e-6


def test_adjoint_state_errors(smadft):
    p, psi0, tgt = setup(smadft, shape=(16, 16, 16))
    with smadft.Stepper(p, psi0) as s:
        with pytest.raises(smadft.SmadError) as e:
            s.adjoint(tgt)
        assert e.value.code == smadft.ESTATE
        s.step(3)
        with pytest.raises(smadft.SmadError) as e:
            s.forward_tape(5, 2)
        assert e.value.code == smadft.ESTATE
        with pytest.raises(smadft.SmadError) as e:
            s.set_tape_budget(-1.0)
        assert e.value.code == smadft.EINVAL


def test_autograd_rollout_matches_oracle_and_fd(smadft, oracle):
    import importlib
    import torch
    ag = importlib.import_module("sma-dft_amd.autograd")
    p, psi0, tgt = setup(smadft, shape=(16, 16, 32))
    K = 16
    theta = torch.tensor([p.eps, p.alpha, p.beta, p.gamma], dtype=torch.float64, requires_grad=True)
    x0 = torch.tensor(psi0, requires_grad=True)
    t = torch.tensor(tgt)
    loss = ((ag.rollout(x0, theta, p, K, 4) - t) ** 2).sum()
    loss.backward()
    g = 2 * (oracle.plain_forward(p, psi0, K) - tgt)
    _, d0_o, dp_o = oracle.plain_adjoint(p, psi0, K, g)
    assert rel_l2(x0.grad.numpy(), d0_o) < 1e-9
    assert np.allclose(theta.grad.numpy(), dp_o, rtol=1e-9)
    # central differences of the GPU loss in the parameters, h = 1e-6
    import dataclasses
    h = 1e-6

    def J(**kw):
        q = dataclasses.replace(p, **kw)
        with smadft.Stepper(q, psi0) as s:
            s.step(K)
            return float(np.sum((s.fetch() - tgt) ** 2))

    for i, name in enumerate(["eps", "alpha", "beta", "gamma"]):
        v = getattr(p, name)
        fd = (J(**{name: v + h}) - J(**{name: v - h})) / (2 * h)
        print(f"autograd dJ/d{name}: {theta.grad[i].item():.9e}, FD {fd:.9e}", flush=True)
        assert theta.grad[i].item() == pytest.approx(fd, rel=1e-6, abs=0), name


def test_adjoint_reports_nonfinite_gradients(smadft):
    """A NaN in the cotangent gives SMAD_ENONFINITE (as smad_fetch does for a
    non-finite field), not a silent SMAD_OK."""
    p, psi0, tgt = setup(smadft, shape=(16, 16, 16))
    g = tgt.copy()
    g[3, 4, 5] = np.nan
    with smadft.Stepper(p, psi0) as s:
        s.forward_tape(4, 2)
        with pytest.raises(smadft.SmadError) as e:
            s.adjoint(g)
        assert e.value.code == smadft.ENONFINITE


def test_device_buffer_adjoint_matches_host(smadft):
    """smad_forward_tape_dev / smad_adjoint_dev on torch device tensors give
    the host entry points' results bit for bit; smad_set_params switches the
    coefficients of a live context."""
    import torch
    p, psi0, tgt = setup(smadft, shape=(16, 32, 32))
    K, ck = 12, 5
    with smadft.Stepper(p, psi0) as s:
        s.forward_tape(K, ck)
        psik_h = s.fetch()
        g = 2 * (psik_h - tgt)
        d0_h, dp_h = s.adjoint(g)
    q = dataclasses_replace(p, eps=-0.5)
    with smadft.Stepper(q, np.zeros(p.shape)) as s:
        s.set_params(p)
        x = torch.tensor(psi0, device="cuda:0")
        out = torch.empty_like(x)
        torch.cuda.synchronize()
        s.forward_tape_dev(x.data_ptr(), K, ck, out.data_ptr())
        assert np.array_equal(out.cpu().numpy(), psik_h)
        gd = torch.tensor(g, device="cuda:0")
        d0 = torch.empty_like(gd)
        torch.cuda.synchronize()
        dp = s.adjoint_dev(gd.data_ptr(), d0.data_ptr())
    assert np.array_equal(d0.cpu().numpy(), d0_h)
    assert np.array_equal(dp, dp_h)


def dataclasses_replace(p, **kw):
    import dataclasses
    return dataclasses.replace(p, **kw)


def test_autograd_on_device_tensors(smadft, oracle):
    """rollout on cuda tensors (device buffers, pooled contexts) twice in a
    row with different theta: gradients against the oracle each time."""
    import importlib
    import torch
    ag = importlib.import_module("sma-dft_amd.autograd")
    p, psi0, tgt = setup(smadft, shape=(16, 16, 32))
    K = 10
    for eps in (p.eps, -0.6):
        q = dataclasses_replace(p, eps=eps)
        theta = torch.tensor([q.eps, q.alpha, q.beta, q.gamma], dtype=torch.float64, device="cuda:0",
                             requires_grad=True)
        x0 = torch.tensor(psi0, device="cuda:0", requires_grad=True)
        t = torch.tensor(tgt, device="cuda:0")
        loss = ((ag.rollout(x0, theta, p, K, 4) - t) ** 2).sum()
        loss.backward()
        g = 2 * (oracle.plain_forward(q, psi0, K) - tgt)
        _, d0_o, dp_o = oracle.plain_adjoint(q, psi0, K, g)
        assert x0.grad.is_cuda
        assert rel_l2(x0.grad.cpu().numpy(), d0_o) < 1e-9
        assert np.allclose(theta.grad.cpu().numpy(), dp_o, rtol=1e-9)
