"""CPU pinning of the advective oracle (adv/cosAdvTrackMult.cpp:827-1369).

The reference's velocity projection reads Sy[i+1,j-1,k], Sz[i+1,j,k-1], ...
(:1103-1274).  These shifted reads are what make the three DST/DCT force
components refer to one physical wavevector; the tests below fail for any
transposed or mis-shifted read:
  * planar layers psi(z) give u == 0 and the plain-model trajectory;
  * the spectral divergence q . u_hat of the projected velocity vanishes on
    all modes the reference does not truncate;
  * a single force mode gives (I - q q^T/|q|^2) F / (nu |q|^2).
"""
import numpy as np
import pytest

from conftest import rel_l2


def adv_params(smadft, shape, **kw):
    p = smadft.Params.config(2, shape[0])  # advective, nw = 8, dt = 0.01
    p.nx, p.ny, p.nz = shape
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_planar_layers_no_flow(smadft, oracle):
    shape = (16, 8, 32)
    p = adv_params(smadft, shape)
    dz = 2 * np.pi / (p.q0 * p.nw)
    z = np.arange(shape[2]) * dz
    psi0 = np.broadcast_to(1.2 * np.cos(p.q0 * z) * np.exp(-((z - z.mean()) / 8) ** 2), shape).copy()
    a = oracle.AdvOracle(p, psi0)
    a.step(15)
    assert np.abs(a.vel).max() < 1e-12 * np.abs(a.grad[2]).max()
    pl = oracle.PlainOracle(p, psi0)
    pl.step(15)
    assert rel_l2(a.psi, pl.psi) < 1e-12
    # grad psi of planar layers is along z only
    assert np.abs(a.grad[0]).max() < 1e-14 and np.abs(a.grad[1]).max() < 1e-14


def spectral_velocity(oracle, vel):
    # forward transforms of u: STx, STy, STz (round trip gives 8 N^3 u_hat)
    ux = oracle.r2r3d(vel[0], (oracle.RODFT10, oracle.REDFT10, oracle.REDFT10))
    uy = oracle.r2r3d(vel[1], (oracle.REDFT10, oracle.RODFT10, oracle.REDFT10))
    uz = oracle.r2r3d(vel[2], (oracle.REDFT10, oracle.REDFT10, oracle.RODFT10))
    return ux, uy, uz


def test_projected_velocity_is_divergence_free(smadft, oracle):
    shape = (16, 16, 16)
    p = adv_params(smadft, shape)
    psi0 = smadft.ic_curved(p, 1.328, 1e-2, seed=2)
    a = oracle.AdvOracle(p, psi0)
    a.step(1)
    ux, uy, uz = spectral_velocity(oracle, a.vel)
    dx = 2 * np.pi / (p.q0 * p.nw)
    q = [np.pi * np.arange(n) / (n * dx) for n in shape]
    nx, ny, nz = shape
    div = np.zeros((nx - 1, ny - 2, nz - 2))
    mag = np.zeros_like(div)
    # DCT mode (i, j, k), i in 1..nx-1, j in 1..ny-2, k in 1..nz-2
    for i in range(1, nx):
        for j in range(1, ny - 1):
            for k in range(1, nz - 1):
                t = (q[0][i] * ux[i - 1, j, k] + q[1][j] * uy[i, j - 1, k] + q[2][k] * uz[i, j, k - 1])
                div[i - 1, j - 1, k - 1] = t
                mag[i - 1, j - 1, k - 1] = abs(q[0][i] * ux[i - 1, j, k]) + abs(q[1][j] * uy[i, j - 1, k]) + abs(q[2][k] * uz[i, j, k - 1])
    assert mag.max() > 1e-6
    assert np.abs(div).max() < 1e-12 * mag.max()


def test_single_force_mode_stokes_solution(smadft, oracle):
    shape = (8, 8, 16)
    p = adv_params(smadft, shape, nu=1.7)
    rng = np.random.default_rng(11)
    F = rng.standard_normal(3)
    a_, b_, c_ = 3, 2, 5  # DCT wavevector indices of the physical mode
    nx, ny, nz = shape
    # force components of wavevector (qx(a), qy(b), qz(c)) in their own bases
    Sx = np.zeros(shape); Sy = np.zeros(shape); Sz = np.zeros(shape)
    Sx[a_ - 1, b_, c_] = F[0]
    Sy[a_, b_ - 1, c_] = F[1]
    Sz[a_, b_, c_ - 1] = F[2]
    ux, uy, uz = oracle.adv_stokes(p, Sx, Sy, Sz)
    dx = 2 * np.pi / (p.q0 * p.nw)
    qv = np.array([np.pi * a_ / (nx * dx), np.pi * b_ / (ny * dx), np.pi * c_ / (nz * dx)])
    q2 = qv @ qv
    scale = 0.125 / (nx * ny * nz)
    u = scale * (F - qv * (qv @ F) / q2) / (p.nu * q2)
    got = np.array([ux[a_ - 1, b_, c_], uy[a_, b_ - 1, c_], uz[a_, b_, c_ - 1]])
    assert np.allclose(got, u, rtol=1e-13, atol=0)
    # no other mode is excited
    assert np.count_nonzero(ux) == 1 and np.count_nonzero(uy) == 1 and np.count_nonzero(uz) == 1


# ---------------------------------------------------------------------------
# Closed-form pins of the gradient, mu, the force and the advection term
# (adv/cosAdvTrackMult.cpp:866-1303).  Everything but the Stokes projection
# (pinned above by the single force mode and the divergence test) is
# computed here from closed forms and scipy's transforms.

import forms  # noqa: E402


@pytest.mark.parametrize("modes", [
    [(0, 0, 0, 0.3), (3, 0, 0, 0.5), (15, 0, 0, 0.2)],   # constant + mid + top x mode
    [(0, 0, 0, -0.4), (0, 5, 0, 0.7), (0, 7, 0, 0.1)],   # y, incl. the top mode of ny = 8
    [(0, 0, 0, 0.2), (0, 0, 9, 0.6), (0, 0, 31, 0.3)],   # z, incl. the top mode
    [(2, 3, 4, 0.8), (15, 7, 31, 0.25)],                # products of the three axes
])
def test_gradient_of_cosine_modes_is_the_analytic_derivative(smadft, oracle, modes):
    """d/dx cos(q_a x) = -q_a sin(q_a x) exactly: the DST-shift gradient maps
    DCT mode a to DST mode a-1 with the factor -qs(a-1) = -q_a.  The
    constant mode must give no gradient, which fails if the top DST mode
    (which has no DCT partner, :877-902) were fed from a wrapped index."""
    p = adv_params(smadft, (16, 8, 32))
    psi0 = np.zeros(p.shape)
    grad = [np.zeros(p.shape) for _ in range(3)]
    for a, b, c, amp in modes:
        f, g = forms.mode(p, a, b, c)
        psi0 += amp * f
        for ax in range(3):
            grad[ax] += amp * g[ax]
    o = oracle.AdvOracle(p, psi0)
    o.step(1)  # the gradient of the step is that of psi^0
    gmax = max(np.abs(g).max() for g in grad)
    for ax in range(3):
        assert np.abs(o.grad[ax] - grad[ax]).max() < 1e-12 * max(gmax, 1.0), ax


def first_step_closed_form(p, amps_modes, advection_sign=-1.0, mu_sign=1.0):
    """psi^1, u and S of the first advective step for psi^0 = sum of cosine
    modes, from closed forms: Sp2 = -L(q) psi per mode (mu's linear part),
    mu = Sp2 - beta psi^3 + gamma psi^5, S = mu grad psi, the projection of
    the (pinned) oracle Stokes solve, N = beta psi^3 - gamma psi^5 - u . grad
    psi and the CN/AB2 update with the pre-loop history
    N_hat^{-1} = CT(beta psi^3 - gamma psi^5) (first step Euler in N)."""
    import oracle as orc
    psi = np.zeros(p.shape)
    grad = [np.zeros(p.shape) for _ in range(3)]
    sp2 = np.zeros(p.shape)
    for amp, (a, b, c) in amps_modes:
        f, g = forms.mode(p, a, b, c)
        psi += amp * f
        sp2 += -forms.alin(p, forms.mq2(p, a, b, c)) * amp * f
        for ax in range(3):
            grad[ax] += amp * g[ax]
    mu = mu_sign * (sp2 - p.beta * psi ** 3 + p.gamma * psi ** 5)
    S = [mu * g for g in grad]
    Sh = [forms.r2r(S[ax], forms.ST[ax]) for ax in range(3)]
    uh = orc.adv_stokes(p, *Sh)
    u = [forms.r2r(uh[ax], forms.IST[ax]) for ax in range(3)]
    nl0 = p.beta * psi ** 3 - p.gamma * psi ** 5
    nr = nl0 + advection_sign * sum(u[ax] * grad[ax] for ax in range(3))
    s = forms.scale(p)
    L = forms.alin(p, forms.mq2_grid(p))
    c1, c2 = 1 + p.dt / 2 * L, 1 - p.dt / 2 * L
    ph = s * forms.ct(psi)
    ph1 = (c1 * ph + p.dt / 2 * s * (3 * forms.ct(nr) - forms.ct(nl0))) / c2
    return forms.ict(ph1), u, S


def test_first_step_closed_form_mu_force_and_advection_sign(smadft, oracle):
    p = adv_params(smadft, (16, 16, 16), dt=0.02)
    am = [(1.1, (2, 1, 3)), (0.45, (1, 3, 2)), (-0.3, (3, 2, 1))]
    psi0 = sum(a * forms.mode(p, *m)[0] for a, m in am)
    o = oracle.AdvOracle(p, psi0)
    o.step(1)
    psi1, u, S = first_step_closed_form(p, am)
    umax = max(np.abs(v).max() for v in u)
    assert umax > 1e-6
    for ax in range(3):
        assert np.abs(o.vel[ax] - u[ax]).max() < 1e-11 * umax, ax
    tol = 1e-13 * np.abs(psi1).max()
    assert np.abs(o.psi - psi1).max() < tol
    # the checks are sharp: the wrong advection sign, or mu with its
    # nonlinear terms' signs reversed, moves psi^1 far beyond the tolerance
    flipped, _, _ = first_step_closed_form(p, am, advection_sign=+1.0)
    assert np.abs(flipped - psi1).max() > 1e4 * tol
    _, u_wrong, _ = first_step_closed_form(p, am, mu_sign=-1.0)
    assert max(np.abs(u_wrong[ax] - u[ax]).max() for ax in range(3)) > 1.0 * umax
    # mu's local part g(psi) = -beta psi^3 + gamma psi^5 adds g grad psi =
    # grad G(psi) to S, a gradient the projection removes: only Sp2 (the
    # nonlocal part) drives the flow, so the local terms' signs are not
    # observable through u (and the reference's choice cannot be pinned
    # there); the projection must show exactly that
    import oracle as orc
    psi = sum(a * forms.mode(p, *m)[0] for a, m in am)
    grad = [sum(a * forms.mode(p, *m)[1][ax] for a, m in am) for ax in range(3)]
    G = [(p.beta * psi ** 3) * g for g in grad]
    uG = orc.adv_stokes(p, *[forms.r2r(G[ax], forms.ST[ax]) for ax in range(3)])
    assert max(np.abs(v).max() for v in uG) < 1e-13 * umax


def test_stokes_flow_dissipates(smadft, oracle):
    """The flow does positive work against the force, sum u . S > 0
    (u = P S/(nu |q|^2)): advection lowers the free energy, as the model's
    -u . grad psi with the force mu grad psi must."""
    p = adv_params(smadft, (16, 16, 16))
    am = [(1.2, (2, 1, 3)), (0.5, (1, 3, 2))]
    psi0 = sum(a * forms.mode(p, *m)[0] for a, m in am)
    o = oracle.AdvOracle(p, psi0)
    o.step(1)
    _, _, S = first_step_closed_form(p, am)
    work = sum(float(np.sum(o.vel[ax] * S[ax])) for ax in range(3))
    scale_ = sum(float(np.sum(np.abs(o.vel[ax] * S[ax]))) for ax in range(3))
    assert work > 0.2 * scale_


# ---------------------------------------------------------------------------
# The advective adjoint (orc_adv_adjoint): transposes by dot products, the
# gradients by central differences.

@pytest.mark.parametrize("kinds", [(0, 0, 0), (1, 1, 1), (2, 0, 0), (1, 3, 1), (3, 2, 1)])
def test_transform_transposes(oracle, kinds):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((8, 16, 12))
    y = rng.standard_normal((8, 16, 12))
    lhs = np.sum(oracle.r2r3d(x, kinds) * y)
    rhs = np.sum(x * oracle.r2r3d_T(y, kinds))
    assert lhs == pytest.approx(rhs, rel=1e-13)


def test_projection_transpose(smadft, oracle):
    """<P S, U> = <S, P^T U> for the projection with its shifted reads and
    truncations (adv/cosAdvTrackMult.cpp:1103-1274)."""
    p = adv_params(smadft, (8, 12, 16))
    rng = np.random.default_rng(2)
    S = [rng.standard_normal(p.shape) for _ in range(3)]
    U = [rng.standard_normal(p.shape) for _ in range(3)]
    PS = oracle.adv_stokes(p, *S)
    PtU = oracle.adv_stokes_T(p, *U)
    lhs = sum(np.sum(a * b) for a, b in zip(PS, U))
    rhs = sum(np.sum(a * b) for a, b in zip(S, PtU))
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_adv_adjoint_central_differences(smadft, oracle):
    """dJ/d(eps, alpha, beta, gamma) and dJ/dpsi0 along a random direction
    of J = |psi_K - target|^2 after K advective steps, against central
    differences (h = 1e-6) of the forward oracle."""
    p = adv_params(smadft, (16, 16, 16), dt=0.01)
    psi0 = smadft.ic_curved(p, 1.328, 1e-2, seed=2)
    K = 6
    tgt = np.random.default_rng(3).standard_normal(p.shape) * 0.1
    psik = oracle.adv_forward(p, psi0, K)
    _, d0, dp = oracle.adv_adjoint(p, psi0, K, 2 * (psik - tgt))
    import dataclasses
    h = 1e-6

    def J(q, x0):
        return float(np.sum((oracle.adv_forward(q, x0, K) - tgt) ** 2))

    for i, name in enumerate(("eps", "alpha", "beta", "gamma")):
        v = getattr(p, name)
        fd = (J(dataclasses.replace(p, **{name: v + h}), psi0) - J(dataclasses.replace(p, **{name: v - h}), psi0)) / (2 * h)
        assert dp[i] == pytest.approx(fd, rel=1e-6), name
    # dJ/dpsi0: along a random unit direction the derivative is ~1e-5 of J,
    # so h = 1e-6 would be roundoff-limited; h = 1e-4 (central differences
    # converge there to 1e-8), and h = 1e-6 along the gradient itself
    v = np.random.default_rng(4).standard_normal(p.shape)
    v /= np.linalg.norm(v)
    fd = (J(p, psi0 + 1e-4 * v) - J(p, psi0 - 1e-4 * v)) / 2e-4
    assert float(np.sum(d0 * v)) == pytest.approx(fd, rel=1e-6)
    u = d0 / np.linalg.norm(d0)
    fd = (J(p, psi0 + h * u) - J(p, psi0 - h * u)) / (2 * h)
    assert float(np.sum(d0 * u)) == pytest.approx(fd, rel=1e-6)
    # the flow matters: the advective gradient differs from the plain model's
    _, d0p, _ = oracle.plain_adjoint(p, psi0, K, 2 * (oracle.plain_forward(p, psi0, K) - tgt))
    assert np.linalg.norm(d0 - d0p) > 1e-6 * np.linalg.norm(d0)
