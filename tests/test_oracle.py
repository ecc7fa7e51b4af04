"""CPU tests of the oracle: transforms against scipy fixtures, closed forms and
brute force; the plain step against the CN amplification factor and the
uniform-field AB2 recurrence."""
import os

import numpy as np
import pytest

from conftest import rel_l2

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "r2r_scipy.npz"))
SIZES = (4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("kind", range(4))
@pytest.mark.parametrize("naive", (False, True))
def test_r2r1d_scipy_fixtures(oracle, n, kind, naive):
    if naive and n > 128:
        pytest.skip("brute force only on small lengths")
    y = oracle.r2r1d(GOLD[f"x_{n}"], kind, naive=naive)
    ref = GOLD[f"y_{n}_{kind}"]
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref)) * np.log2(n)


@pytest.mark.parametrize("combo", ["000", "111", "200", "020", "002", "311", "131", "113"])
def test_r2r3d_scipy_fixtures(oracle, combo):
    kinds = tuple(int(c) for c in combo)
    y = oracle.r2r3d(GOLD["x3"], kinds)
    assert rel_l2(y, GOLD[f"y3_{combo}"]) < 1e-14


@pytest.mark.parametrize("n", (8, 16, 48, 128))
def test_dct2_known_answer(oracle, n):
    # REDFT10 of cos(pi m (2j+1)/2N) is N delta_km for m > 0, 2N for m = 0
    j = np.arange(n)
    for m in (0, 1, n // 3, n - 1):
        y = oracle.r2r1d(np.cos(np.pi * m * (2 * j + 1) / (2 * n)), oracle.REDFT10)
        ref = np.zeros(n)
        ref[m] = 2 * n if m == 0 else n
        assert np.max(np.abs(y - ref)) < 1e-12 * n


@pytest.mark.parametrize("n", (8, 24, 64))
def test_dst2_known_answer(oracle, n):
    # RODFT10 of sin(pi (m+1)(2j+1)/2N) is N delta_km (k < N-1), 2N at m = N-1
    j = np.arange(n)
    for m in (0, 3, n - 1):
        y = oracle.r2r1d(np.sin(np.pi * (m + 1) * (2 * j + 1) / (2 * n)), oracle.RODFT10)
        ref = np.zeros(n)
        ref[m] = 2 * n if m == n - 1 else n
        assert np.max(np.abs(y - ref)) < 1e-12 * n


@pytest.mark.parametrize("n", (16, 96, 256))
def test_round_trips_and_transposes(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n)
    yv = rng.standard_normal(n)
    for f, b in ((0, 1), (2, 3)):
        back = oracle.r2r1d(oracle.r2r1d(x, f), b)
        assert np.max(np.abs(back - 2 * n * x)) < 1e-12 * n
    # <REDFT10 x, y> = <x, REDFT01(diag(2,1,..,1) y)>
    d = yv.copy(); d[0] *= 2
    assert abs(oracle.r2r1d(x, 0) @ yv - x @ oracle.r2r1d(d, 1)) < 1e-11 * n
    # <RODFT10 x, y> = <x, RODFT01(diag(1,..,1,2) y)>
    d = yv.copy(); d[-1] *= 2
    assert abs(oracle.r2r1d(x, 2) @ yv - x @ oracle.r2r1d(d, 3)) < 1e-11 * n


def _params(smadft, shape, **kw):
    p = smadft.Params.default(0, shape[0])
    p.nx, p.ny, p.nz = shape
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def cos_mode(shape, a, b, c):
    nx, ny, nz = shape
    cx = np.cos(np.pi * a * (2 * np.arange(nx) + 1) / (2 * nx))
    cy = np.cos(np.pi * b * (2 * np.arange(ny) + 1) / (2 * ny))
    cz = np.cos(np.pi * c * (2 * np.arange(nz) + 1) / (2 * nz))
    return cx[:, None, None] * cy[None, :, None] * cz[None, None, :]


def cn_factor(p, a, b, c):
    # Crank-Nicolson amplification of psi_t = L psi, L = eps - alpha (|q|^2 - q0^2)^2,
    # q = pi (a, b, c) / (N dx), dx = 2 pi / (q0 nw)
    dx = 2 * np.pi / (p.q0 * p.nw)
    q2 = sum((np.pi * m / (n * dx)) ** 2 for m, n in ((a, p.nx), (b, p.ny), (c, p.nz)))
    lin = p.eps - p.alpha * (q2 - p.q0 ** 2) ** 2
    return (1 + p.dt / 2 * lin) / (1 - p.dt / 2 * lin)


@pytest.mark.parametrize("mode", [(0, 0, 0), (1, 2, 3), (5, 0, 7), (3, 3, 3)])
def test_plain_linear_mode_decay(smadft, oracle, mode):
    shape = (16, 8, 24)
    p = _params(smadft, shape, beta=0.0, gamma=0.0, eps=-0.3, dt=0.2)
    psi0 = cos_mode(shape, *mode)
    o = oracle.PlainOracle(p, psi0)
    nst = 30
    o.step(nst)
    g = cn_factor(p, *mode)
    # error measured against the initial amplitude: roundoff of early steps
    # stays at ~1e-16 |psi0| while a stiff mode decays far below it
    err = np.linalg.norm(o.psi - g ** nst * psi0) / np.linalg.norm(psi0)
    assert err < 1e-13
    if abs(g) ** nst > 1e-3:
        assert rel_l2(o.psi, g ** nst * psi0) < 1e-11


def test_plain_uniform_recurrence(smadft, oracle):
    shape = (8, 16, 8)
    p = _params(smadft, shape, eps=0.3, dt=0.05)
    u = 0.37
    o = oracle.PlainOracle(p, np.full(shape, u))
    lin0 = p.eps - p.alpha * p.q0 ** 4
    nl = lambda v: p.beta * v ** 3 - p.gamma * v ** 5
    c1, c2 = 1 + p.dt / 2 * lin0, 1 - p.dt / 2 * lin0
    prev_n = nl(u)  # pre-loop history: first step is Euler in N
    for _ in range(25):
        o.step(1)
        cur_n = nl(u)
        u = (c1 * u + p.dt / 2 * (3 * cur_n - prev_n)) / c2
        prev_n = cur_n
        assert np.max(np.abs(o.psi - u)) < 1e-12


def test_l1_sums(oracle):
    rng = np.random.default_rng(3)
    a, b = rng.standard_normal(1000), rng.standard_normal(1000)
    s = oracle.l1_sums(a, b)
    assert np.allclose(s, [np.abs(a - b).sum(), np.abs(a).sum()], rtol=1e-13)


def _adj_setup(smadft):
    p = _params(smadft, (16, 8, 16), dt=0.1)
    psi0 = smadft.ic_slab(p, 1.328, 1e-2, seed=3)
    tgt = np.random.default_rng(0).standard_normal(psi0.shape) * 0.1
    return p, psi0, tgt


def test_plain_adjoint_central_fd(smadft, oracle):
    import copy
    p, psi0, tgt = _adj_setup(smadft)
    K = 12

    def J(pp, x0):
        return float(np.sum((oracle.plain_forward(pp, x0, K) - tgt) ** 2))

    g = 2 * (oracle.plain_forward(p, psi0, K) - tgt)
    psik, d0, dp = oracle.plain_adjoint(p, psi0, K, g)
    h = 1e-6
    for idx, name in enumerate(["eps", "alpha", "beta", "gamma"]):
        pp = copy.copy(p); setattr(pp, name, getattr(p, name) + h)
        pm = copy.copy(p); setattr(pm, name, getattr(p, name) - h)
        fd = (J(pp, psi0) - J(pm, psi0)) / (2 * h)
        assert dp[idx] == pytest.approx(fd, rel=1e-6), name
    v = np.random.default_rng(1).standard_normal(psi0.shape)
    fd = (J(p, psi0 + h * v) - J(p, psi0 - h * v)) / (2 * h)
    assert float(np.sum(d0 * v)) == pytest.approx(fd, rel=1e-6)


def test_plain_adjoint_linear_dot_product(smadft, oracle):
    # beta = gamma = 0: the K-step map is linear, M psi0 = psi_K, and the
    # adjoint must satisfy <M x, y> = <x, M^T y> to roundoff
    p, _, _ = _adj_setup(smadft)
    p.beta = 0.0
    p.gamma = 0.0
    rng = np.random.default_rng(5)
    x = rng.standard_normal((p.nx, p.ny, p.nz))
    y = rng.standard_normal(x.shape)
    K = 7
    mx = oracle.plain_forward(p, x, K)
    _, mty, _ = oracle.plain_adjoint(p, x, K, y)
    assert float(np.sum(mx * y)) == pytest.approx(float(np.sum(x * mty)), rel=1e-12)
