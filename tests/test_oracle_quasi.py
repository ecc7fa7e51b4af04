"""CPU pinning of the quasi-incompressible oracle (quasi/quasi.cpp:990-2251).

  * kp = 0 gives a uniform density rho0 and divv == 0;
  * the mean pressure is zero (p_hat[0,0,0] = 0, :1577-1581);
  * quasi-incompressibility: the spectral divergence of the velocity equals
    the imposed divv on every mode the reference does not truncate -- this
    pins the shifted reads of S and of divv in the solenoidal and
    irrotational parts (:1599-2109);
  * before the switch step the velocity is solenoidal;
  * planar layers psi(z) keep ux = uy = 0.
"""
import numpy as np
import pytest


def qparams(smadft, shape, **kw):
    p = smadft.Params.config(3, shape[0])  # quasi, rho0 = 0.9, kp = 0.1/1.34164
    p.nx, p.ny, p.nz = shape
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_default_quasi_is_stable_at_reference_dt(smadft, oracle):
    p = qparams(smadft, (16, 16, 16), div_switch=4)
    q = oracle.QuasiOracle(p, smadft.ic_curved(p, 1.34164, 1e-3, seed=3))
    q.step(30)
    assert np.isfinite(q.psi).all() and np.abs(q.psi).max() < 2.0


def test_kp_zero_uniform_density(smadft, oracle):
    p = qparams(smadft, (16, 16, 16), kp=0.0, rho0=1.0, rho_m=0.5, div_switch=3)
    q = oracle.QuasiOracle(p, smadft.ic_curved(p, 1.34164, 1e-3, seed=3))
    q.step(8)
    assert np.abs(q.rho - 1.0).max() < 1e-13
    assert np.abs(q.divv).max() < 1e-9  # unscaled REDFT10 of roundoff


def test_mean_pressure_zero(smadft, oracle):
    p = qparams(smadft, (16, 8, 16), div_switch=2)
    q = oracle.QuasiOracle(p, smadft.ic_curved(p, 1.34164, 1e-3, seed=3))
    q.step(5)
    assert abs(q.p.mean()) < 1e-12 * np.abs(q.p).max()


def spectral_div(oracle, p, vel):
    ux = oracle.r2r3d(vel[0], (2, 0, 0))
    uy = oracle.r2r3d(vel[1], (0, 2, 0))
    uz = oracle.r2r3d(vel[2], (0, 0, 2))
    dx = 2 * np.pi / (p.q0 * p.nw)
    nx, ny, nz = p.nx, p.ny, p.nz
    qx = np.pi * np.arange(nx) / (nx * dx)
    qy = np.pi * np.arange(ny) / (ny * dx)
    qz = np.pi * np.arange(nz) / (nz * dx)
    sl = np.s_[1:nx, 1:ny - 1, 1:nz - 1]
    div = (qx[1:, None, None] * ux[0:nx - 1, 1:ny - 1, 1:nz - 1]
           + qy[None, 1:ny - 1, None] * uy[1:nx, 0:ny - 2, 1:nz - 1]
           + qz[None, None, 1:nz - 1] * uz[1:nx, 1:ny - 1, 0:nz - 2])
    return div, sl


def test_velocity_divergence_equals_divv(smadft, oracle):
    p = qparams(smadft, (16, 16, 16), div_switch=2)
    q = oracle.QuasiOracle(p, smadft.ic_curved(p, 1.34164, 1e-3, seed=3))
    q.step(6)
    div, sl = spectral_div(oracle, p, q.vel)
    n3 = 8 * p.nx * p.ny * p.nz
    scale = 1.0 / n3
    # ST(vel) = 8N^3 u_hat and q . u_hat = s divv_hat (divv_hat unscaled)
    target = n3 * scale * q.divv[sl]
    assert np.abs(target).max() > 1e-6
    assert np.abs(div - target).max() < 1e-9 * np.abs(target).max()


def test_velocity_solenoidal_before_switch(smadft, oracle):
    p = qparams(smadft, (16, 16, 16), div_switch=100)
    q = oracle.QuasiOracle(p, smadft.ic_curved(p, 1.34164, 1e-3, seed=3))
    q.step(4)
    div, _ = spectral_div(oracle, p, q.vel)
    ux = oracle.r2r3d(q.vel[0], (2, 0, 0))
    assert np.abs(ux).max() > 1e-6
    assert np.abs(div).max() < 1e-10 * np.abs(ux).max() * np.pi


def test_planar_layers_flow_along_normal_only(smadft, oracle):
    p = qparams(smadft, (8, 8, 32), div_switch=2)
    dz = 2 * np.pi / (p.q0 * p.nw)
    z = np.arange(p.nz) * dz
    psi0 = np.broadcast_to(1.34 * np.cos(z) * np.exp(-((z - z.mean()) / 6) ** 2), (p.nx, p.ny, p.nz)).copy()
    q = oracle.QuasiOracle(p, psi0)
    q.step(6)
    assert np.abs(q.vel[2]).max() > 1e-8
    assert np.abs(q.vel[0]).max() < 1e-12 * np.abs(q.vel[2]).max() + 1e-15
    assert np.abs(q.vel[1]).max() < 1e-12 * np.abs(q.vel[2]).max() + 1e-15
    assert np.abs(q.psi - q.psi[:1, :1, :]).max() < 1e-12


# ---------------------------------------------------------------------------
# Closed-form pins of the quasi step's pieces (quasi/quasi.cpp:1121-2182).
# The reference writes mu = lap + rho (f - eps psi - beta psi^3 + gamma psi^5)
# with f = alpha (q0^2 + grad^2) psi and lap = rho grad^2 f (before the
# switch) or grad^2(rho f) (after it): mu = rho delta F/delta psi of the
# Swift-Hohenberg energy for q0 = 1, the reference's value, which these
# tests use.

import forms  # noqa: E402


def scalar_bdf2(c0, K, rhs, lin, dt, rho_m):
    """BDF2/AB2 recurrence of one spectral amplitude: C1 = 2 - dt rho_m L,
    C2 = 1.5 - dt rho_m L, psi_old^0 = psi^0, Nl_old = Nl at the first step."""
    c1, c2 = 2 - dt * rho_m * lin, 1.5 - dt * rho_m * lin
    c_prev = c = c0
    nl_prev = None
    out = [c0]
    for n in range(1, K + 1):
        nl = rhs(c)
        if n == 1:
            nl_prev = nl
        c_new = (c1 * c - 0.5 * c_prev + dt / 2 * (3 * nl - nl_prev)) / c2
        c_prev, c, nl_prev = c, c_new, nl
        out.append(c)
    return out


def test_uniform_psi_follows_the_scalar_bdf2_recurrence(smadft, oracle):
    """Uniform psi = c: grad psi = 0, rho = kp q0 |c| + rho0, f = alpha q0^2 c,
    mu = rho (f - eps c - beta c^3 + gamma c^5), u = 0, Nl = -mu; mass and
    energy are n dx^3 rho(c) and n dx^3 e(c) with
    e = -eps c^2/2 + f^2/2 - beta c^4/4 + gamma c^6/6.  Pins C1, C2 (with
    rho_m and L(0) = eps - alpha q0^4), mu's nonlinear part and its rho
    factor, the density law without gradients, e, mass and energy."""
    p = qparams(smadft, (8, 8, 8), kp=0.3, rho0=0.9, rho_m=0.6, dt=0.01, div_switch=3)
    c0, K = 0.8, 8
    q = oracle.QuasiOracle(p, np.full(p.shape, c0))
    q.step(K)
    rho = lambda c: p.kp * p.q0 * abs(c) + p.rho0
    f = lambda c: p.alpha * p.q0 ** 2 * c
    mu = lambda c: rho(c) * (f(c) - p.eps * c - p.beta * c ** 3 + p.gamma * c ** 5)
    e = lambda c: -0.5 * p.eps * c * c + 0.5 * f(c) ** 2 - 0.25 * p.beta * c ** 4 + p.gamma / 6 * c ** 6
    L0 = p.eps - p.alpha * p.q0 ** 4
    cs = scalar_bdf2(c0, K, lambda c: -mu(c), L0, p.dt, p.rho_m)
    assert abs(cs[K] - c0) > 1e-3  # the state moved
    assert np.abs(q.psi - cs[K]).max() < 1e-12
    dv = forms.spacing(p) ** 3 * q.psi.size
    assert q.sums[2] == pytest.approx(dv * rho(cs[K - 1]), rel=1e-12)   # from psi^{K-1}
    assert q.sums[3] == pytest.approx(dv * e(cs[K - 1]), rel=1e-12)


def test_linear_single_mode_follows_the_bdf2_recurrence(smadft, oracle):
    """beta = gamma = kp = 0: rho = rho0 and mu = -rho0 L(q) psi for one
    cosine mode, in both the lap branches; mu grad psi is then a gradient,
    so u = 0 and Nl = rho0 L(q) psi: the mode's amplitude follows the scalar
    recurrence across the switch step (pins C1, C2 at q != 0, f and both
    lap branches)."""
    p = qparams(smadft, (16, 16, 16), beta=0.0, gamma=0.0, kp=0.0, rho0=0.9, rho_m=0.6, dt=0.01,
                div_switch=2)
    m, _ = forms.mode(p, 2, 1, 3)
    K, A0 = 6, 0.5
    q = oracle.QuasiOracle(p, A0 * m)
    q.step(K)
    L = forms.alin(p, forms.mq2(p, 2, 1, 3))
    amps = scalar_bdf2(A0, K, lambda a: p.rho0 * L * a, L, p.dt, p.rho_m)
    assert abs(amps[K] - A0) > 1e-3
    assert np.abs(q.psi - amps[K] * m).max() < 1e-12 * A0
    assert np.abs(q.vel).max() < 1e-12


def test_lap_branches_agree_for_uniform_density(smadft, oracle):
    """kp = 0 makes rho uniform, where grad^2(rho f) (after the switch) equals
    rho grad^2 f (before it): a run switched at step 0 and one never switched
    must coincide (the divv of the first switched step is uniform and only
    reaches the mean mode, which the projection does not use)."""
    kw = dict(kp=0.0, rho0=0.9, rho_m=0.6)
    pa = qparams(smadft, (16, 16, 16), div_switch=0, **kw)
    pb = qparams(smadft, (16, 16, 16), div_switch=1000, **kw)
    psi0 = smadft.ic_curved(pa, 1.34164, 1e-3, seed=3)
    a, b = oracle.QuasiOracle(pa, psi0), oracle.QuasiOracle(pb, psi0)
    a.step(5)
    b.step(5)
    assert np.abs(a.psi - psi0).max() > 1e-4
    assert np.abs(a.psi - b.psi).max() < 1e-12 * np.abs(b.psi).max()


def test_density_law_and_gaussian_filter(smadft, oracle):
    """rho = iCT(s CT(kp sqrt(q0^2 psi^2 + |grad psi|^2) + rho0) exp(-2 sigma^2 |q|^2)),
    with psi and grad psi in closed form and scipy's transforms."""
    p = qparams(smadft, (16, 16, 16), kp=0.3, rho0=0.9, q0=1.3)  # q0 != 1 pins its factor
    am = [(1.0, (2, 1, 3)), (0.4, (1, 2, 1)), (0.2, (0, 0, 0))]
    psi = np.zeros(p.shape)
    grad = [np.zeros(p.shape) for _ in range(3)]
    for amp, m in am:
        f, g = forms.mode(p, *m)
        psi += amp * f
        for ax in range(3):
            grad[ax] += amp * g[ax]
    q = oracle.QuasiOracle(p, psi)
    q.step(1)
    raw = p.kp * np.sqrt(p.q0 ** 2 * psi ** 2 + sum(g * g for g in grad)) + p.rho0

    def filt(expo):
        return forms.ict(forms.scale(p) * forms.ct(raw) * np.exp(expo * forms.mq2_grid(p)))

    sig2 = p.filter_sigma ** 2
    rho = filt(-2 * sig2)
    tol = 1e-13 * np.abs(rho).max()
    assert np.abs(q.rho - rho).max() < tol
    # the variants' exponents (-sigma^2/2, -8 sigma^2) are far off
    assert np.abs(filt(-0.5 * sig2) - rho).max() > 1e6 * tol
    assert np.abs(filt(-8 * sig2) - rho).max() > 1e6 * tol


def test_energy_density_and_mu_of_cosine_modes(smadft, oracle):
    """e = -eps psi^2/2 + f^2/2 - beta psi^4/4 + gamma psi^6/6 with
    f = alpha (q0^2 - |q|^2) psi per mode; mu = rho (grad^2 f + f - eps psi
    - beta psi^3 + gamma psi^5) before the switch and
    grad^2(rho f) + rho (f - ...) after it."""
    for div_switch in (100, 0):
        p = qparams(smadft, (16, 16, 16), kp=0.3, rho0=0.9, div_switch=div_switch)
        am = [(0.9, (2, 1, 3)), (0.3, (0, 0, 0)), (-0.2, (1, 2, 2))]
        psi = np.zeros(p.shape)
        f = np.zeros(p.shape)
        lapf = np.zeros(p.shape)
        for amp, m in am:
            fm, _ = forms.mode(p, *m)
            m2 = forms.mq2(p, *m)
            psi += amp * fm
            f += p.alpha * (p.q0 ** 2 - m2) * amp * fm
            lapf += -m2 * p.alpha * (p.q0 ** 2 - m2) * amp * fm
        q = oracle.QuasiOracle(p, psi)
        q.step(1)
        e = -0.5 * p.eps * psi ** 2 + 0.5 * f ** 2 - 0.25 * p.beta * psi ** 4 + p.gamma / 6 * psi ** 6
        assert np.abs(q.energy - e).max() < 1e-13 * np.abs(e).max()
        rho = q.rho
        assert np.abs(rho - p.rho0).max() > 1e-3  # the density varies
        loc = f - p.eps * psi - p.beta * psi ** 3 + p.gamma * psi ** 5
        if div_switch:
            mu = rho * (lapf + loc)
        else:
            s = forms.scale(p)
            mu = forms.ict(-s * forms.mq2_grid(p) * forms.ct(rho * f)) + rho * loc
        assert np.abs(q.mu - mu).max() < 1e-12 * np.abs(mu).max(), div_switch


def test_projection_of_single_force_and_divergence_modes(smadft, oracle):
    """nu grad^2 u + (lambda + nu) grad(div u) - grad p + S = 0 with
    div u = divv: for one wavevector q, p_hat = -(q . S)/|q|^2 +
    (lambda + 2 nu) divv_hat and u_hat = (I - q q^T/|q|^2) S/(nu |q|^2) +
    q divv_hat/|q|^2 (times s: S_hat and divv_hat are unscaled), with S_x,
    S_y, S_z and u in their DST bases one index below the DCT mode on their
    own axis."""
    p = qparams(smadft, (8, 8, 16), nu=1.3, lambda_=-0.4)
    rng = np.random.default_rng(4)
    F = rng.standard_normal(3)
    dv = 0.7
    a, b, c = 3, 2, 5
    Sx, Sy, Sz, D = (np.zeros(p.shape) for _ in range(4))
    Sx[a - 1, b, c], Sy[a, b - 1, c], Sz[a, b, c - 1] = F
    D[a, b, c] = dv
    ph, ux, uy, uz = oracle.quasi_project(p, Sx, Sy, Sz, D)
    qv = np.array([w[i] for w, i in zip(forms.wavenumbers(p), (a, b, c))])
    q2 = qv @ qv
    s = forms.scale(p)
    u = s * ((F - qv * (qv @ F) / q2) / (p.nu * q2) + qv * dv / q2)
    pr = s * (-(qv @ F) / q2 + (p.lambda_ + 2 * p.nu) * dv)
    got = np.array([ux[a - 1, b, c], uy[a, b - 1, c], uz[a, b, c - 1]])
    assert np.allclose(got, u, rtol=1e-13, atol=0)
    assert ph[a, b, c] == pytest.approx(pr, rel=1e-13)
    for arr in (ph, ux, uy, uz):
        assert np.count_nonzero(arr) == 1


def test_divv_coupling_and_force_with_a_manufactured_grad_rho(smadft, oracle):
    """After the switch, with the previous step's filtered rho_hat, rho_old
    and u replaced by manufactured fields before one step:
      divv = -((rho - rho_old)/dt + u . grad rho)/((rho + rho_old)/2),
      S = mu grad psi + e grad rho,
      Nl = -mu - u' . grad psi + (p + rho e) kp |grad rho|/rho^2,
    where grad rho is the closed-form gradient of the manufactured previous
    density (a single DCT mode), rho this step's density, u the manufactured
    (previous) velocity and u' this step's."""
    p = qparams(smadft, (16, 16, 16), kp=0.2, rho0=0.9, div_switch=1)
    q = oracle.QuasiOracle(p, smadft.ic_curved(p, 1.34164, 1e-3, seed=3))
    q.step(2)  # nloop = 3 next: divv and the coupling term are both active
    s = forms.scale(p)
    a, b, c, R = 2, 1, 3, 0.01
    q.rhoq[...] = 0.0
    q.rhoq[a, b, c] = R  # previous filtered rho = iCT(rhoq) = 8 R mode(a,b,c)
    m, gm = forms.mode(p, a, b, c)
    grho = [8 * R * g for g in gm]
    rho_old = p.rho0 + 0.02 * forms.mode(p, 1, 2, 1)[0] + 0.01 * forms.mode(p, 0, 3, 0)[0]
    q.rho_old[...] = rho_old
    x, y, z = forms.coords(p)
    u_prev = [0.3 * np.sin(0.4 * y + 0.2) * np.cos(0.3 * z) + 0 * x,
              0.2 * np.cos(0.5 * x) * np.sin(0.6 * z) + 0 * y,
              -0.25 * np.sin(0.35 * x + 0.1) * np.cos(0.45 * y) + 0 * z]
    for ax in range(3):
        q.vel[ax][...] = u_prev[ax]
    q.step(1)
    rho = q.rho
    divv = -((rho - rho_old) / p.dt + sum(u_prev[ax] * grho[ax] for ax in range(3))) / (0.5 * (rho + rho_old))
    got = s * forms.ict(q.divv)
    assert np.abs(got - divv).max() < 1e-10 * np.abs(divv).max()
    # the force and the projection of this step
    gpsi = q.grad_psi
    S = [q.mu * gpsi[ax] + q.energy * grho[ax] for ax in range(3)]
    Sh = [forms.r2r(S[ax], forms.ST[ax]) for ax in range(3)]
    _, uxh, uyh, uzh = oracle.quasi_project(p, *Sh, q.divv)
    u = [forms.r2r(h, forms.IST[ax]) for ax, h in enumerate((uxh, uyh, uzh))]
    umax = max(np.abs(v).max() for v in u)
    for ax in range(3):
        assert np.abs(q.vel[ax] - u[ax]).max() < 1e-10 * umax, ax
    # Nl and its coupling term
    nl = s * forms.ict(q.nl)
    coupling = (q.p + rho * q.energy) * p.kp * np.sqrt(sum(g * g for g in grho)) / rho ** 2
    rest = nl + q.mu + sum(q.vel[ax] * gpsi[ax] for ax in range(3))
    assert np.abs(coupling).max() > 1e-6 * np.abs(nl).max()
    assert np.abs(rest - coupling).max() < 1e-10 * np.abs(coupling).max()


def test_load_start_takes_rho_old_from_the_first_step(smadft, oracle):
    """The reference's load == 1 start (quasi/quasi.cpp:1365-1369): divv and
    grad^2(rho f) from the first step with rho_old = rho there, so the first
    divv has no (rho - rho_old)/dt term.  With kp = 0 (uniform rho) the
    trajectory equals the one switched at step 0, whose first divv is the
    uniform -2/dt of rho_old = 0 (it only reaches the mean mode)."""
    kw = dict(kp=0.0, rho0=0.9, rho_m=0.6)
    pl = qparams(smadft, (16, 16, 16), div_switch=200, load=1, **kw)
    ps = qparams(smadft, (16, 16, 16), div_switch=0, **kw)
    psi0 = smadft.ic_curved(pl, 1.34164, 1e-3, seed=3)
    a, b = oracle.QuasiOracle(pl, psi0), oracle.QuasiOracle(ps, psi0)
    a.step(1)
    b.step(1)
    n3 = 8 * pl.nx * pl.ny * pl.nz
    assert abs(a.divv[0, 0, 0]) < 1e-9 * n3
    assert b.divv[0, 0, 0] == pytest.approx(-2.0 / pl.dt * n3, rel=1e-12)
    a.step(4)
    b.step(4)
    assert np.abs(a.psi - b.psi).max() < 1e-12 * np.abs(b.psi).max()
    # with a varying density the load start differs from a fresh start
    pk = qparams(smadft, (16, 16, 16), div_switch=200, load=1, kp=0.2)
    pf = qparams(smadft, (16, 16, 16), div_switch=200, load=0, kp=0.2)
    c, d = oracle.QuasiOracle(pk, psi0), oracle.QuasiOracle(pf, psi0)
    c.step(3)
    d.step(3)
    assert np.abs(c.psi - d.psi).max() > 1e-9
