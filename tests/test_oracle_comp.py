"""CPU pinning of the compressible oracle (compressible/compressible.cpp:
1114-2700, one slab).

  * uniform psi reduces the step to a scalar recurrence derived here from the
    equations (amp = q0 |psi|, f = alpha q0^2 psi, no flow, the density reset
    to rho0 + kp amp before the switch and frozen after it, the penalty term
    once amp moves away from the frozen density, C1/C2 with rho_m, the L1,
    mass and energy sums);
  * planar layers at the grid wavenumber q0 have amp = q0 A exactly and a
    force that is a gradient, so the solenoidal velocity vanishes and the
    flow stays one-dimensional; the same layers along x, y and z give the
    same profiles (axis-specific shifts and wavenumbers);
  * the velocity of a random force: the solenoidal part is divergence-free
    on every untruncated mode, the irrotational part is the gradient of one
    potential built from G and q . S (with the reference's k = 0 filter of
    the z component, :2350);
  * the relaxation density is lifted by rho_up planes along z (:1606-1628).
"""
import numpy as np
import pytest


def cparams(smadft, shape, **kw):
    p = smadft.Params.default(smadft.MODEL_COMPRESSIBLE, shape[0])
    p.nx, p.ny, p.nz = shape
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_defaults_follow_the_program(smadft):
    p = smadft.Params.default(smadft.MODEL_COMPRESSIBLE, 32)
    amp = np.sqrt((3 * 2.0 + np.sqrt(9 * 4.0 + 40 * p.eps * 1.0)) / 5)
    assert (p.dt, p.nw, p.rho0, p.zeta, p.rho_up, p.div_switch) == (0.001, 8.0, 0.01, 0.001, 16, 200)
    assert p.lambda_ == p.nu
    assert p.kp == pytest.approx((1 - 0.01) / amp, rel=1e-15)
    assert p.kp * amp + p.rho0 == pytest.approx(1.0, rel=1e-14)  # rho_s = 1
    assert p.rho_m == pytest.approx(0.5, rel=1e-14)


def scalar_run(p, c, nsteps):
    """Uniform psi = c: every gradient, the flow and the divergence vanish,
    so each quantity is its DC value (amp and f exact through the filter
    and transforms).  Returns per-step (psi, rho, sums)."""
    q02 = p.q0 ** 2
    al0 = p.eps - p.alpha * q02 ** 2          # aLin at |q| = 0
    c1 = 2.0 - p.dt * p.rho_m * al0
    c2 = 1.5 - p.dt * p.rho_m * al0
    psi, psi_old, rho, nl_old = c, c, 0.0, None
    out = []
    for n in range(1, nsteps + 1):
        amp = p.q0 * abs(psi)
        f = p.alpha * q02 * psi
        if not (n > p.div_switch or p.load == 1):
            rho = p.rho0 + p.kp * amp           # relaxation density (rho_up = 0)
        mu = rho * (f - p.eps * psi - p.beta * psi ** 3 + p.gamma * psi ** 5)
        # continuity: no flow, no divergence, zero AB2 history -> rho unchanged
        nl = -mu + (p.kp * p.zeta / (amp + 0.001)) * (rho - p.rho0 - p.kp * amp) * psi
        if n == 1:
            nl_old = nl
        new = (c1 * psi - 0.5 * psi_old + p.dt / 2 * (3 * nl - nl_old)) / c2
        psi_old, psi, nl_old = psi, new, nl
        e = -0.5 * p.eps * psi ** 2 + 0.5 * f ** 2 - 0.25 * p.beta * psi ** 4 + p.gamma / 6 * psi ** 6
        out.append((psi, rho, abs(psi - psi_old), rho, rho * e + 0.5 * p.zeta * (rho - p.rho0 - p.kp * amp) ** 2))
    return out


@pytest.mark.parametrize("q0", [1.0, 0.9])
def test_uniform_psi_scalar_recurrence(smadft, oracle, q0):
    # a large zeta and dt so that the penalty after the switch moves psi visibly
    p = cparams(smadft, (8, 8, 8), q0=q0, rho_up=0, div_switch=3, zeta=5.0, dt=0.05)
    c = 0.45
    o = oracle.CompOracle(p, np.full(p.shape, c))
    want = scalar_run(p, c, 8)
    n = p.nx * p.ny * p.nz
    dx = 2 * np.pi / (p.q0 * p.nw)
    pen_seen = 0.0
    for s in range(8):
        o.step(1)
        psi_w, rho_w, da, mass, en = want[s]
        assert np.abs(o.psi - psi_w).max() < 1e-12, s
        assert np.abs(o.rho - rho_w).max() < 1e-12, s
        assert np.abs(o.vel).max() < 1e-12 and np.abs(o.divv).max() < 1e-10
        assert np.abs(o.amp - q0 * abs(want[s - 1][0] if s else c)).max() < 1e-12
        assert o.sums[0] == pytest.approx(n * da, rel=1e-9, abs=1e-12)
        assert o.sums[2] == pytest.approx(n * mass * dx ** 3, rel=1e-12)
        assert o.sums[3] == pytest.approx(n * en * dx ** 3, rel=1e-10)
        pen_seen = max(pen_seen, abs(rho_w - p.rho0 - p.kp * q0 * abs(psi_w)))
    assert pen_seen > 1e-3  # the frozen density departs from rho0 + kp amp


def layers(p, axis, amp):
    """amp cos(q0 (m + 1/2) d) along `axis`: a DCT-II basis vector at |q| = q0."""
    dx = 2 * np.pi / (p.q0 * p.nw)
    n = p.shape[axis]
    prof = amp * np.cos(p.q0 * (np.arange(n) + 0.5) * dx)
    sh = [1, 1, 1]
    sh[axis] = n
    return np.broadcast_to(prof.reshape(sh), p.shape).copy()


def test_layers_amp_exact_and_flow_irrotational(smadft, oracle):
    p = cparams(smadft, (8, 8, 32), rho_up=0, div_switch=2, nu=1.0, dt=0.01)
    A = 1.1
    o = oracle.CompOracle(p, layers(p, 2, A))
    o.step(1)
    assert np.abs(o.amp - p.q0 * A).max() < 1e-12           # sqrt(q0^2 psi^2 + psi_z^2) = q0 A
    assert np.abs(o.f).max() < 1e-12                          # alpha (q0^2 - |q|^2) at |q| = q0
    # a modulated profile has low wavenumbers that pass the filter
    psi0 = layers(p, 2, A) * (1 + 0.3 * np.cos(2 * np.pi * (np.arange(32) + 0.5) / 32))
    o = oracle.CompOracle(p, psi0)
    o.step(4)  # across the switch: rho now carried by continuity
    assert np.abs(o.vsol).max() < 1e-12 * max(1.0, np.abs(o.virr).max())
    assert np.abs(o.virr[2]).max() > 1e-8                     # a real compressive flow
    assert np.abs(o.vel[0]).max() < 1e-14 and np.abs(o.vel[1]).max() < 1e-14
    for fld in (o.psi, o.rho, o.vel[2]):                      # still a function of z alone
        assert np.abs(fld - fld[:1, :1, :]).max() < 1e-12 * max(1.0, np.abs(fld).max())


def test_layers_same_along_every_axis(smadft, oracle):
    """Layers along x, y and z on the transposed grids give the same profile
    of psi, rho and the normal velocity (rho_up = 0 keeps the model isotropic)."""
    profs = []
    for axis in range(3):
        shape = [8, 8, 8]
        shape[axis] = 32
        p = cparams(smadft, tuple(shape), rho_up=0, div_switch=2, dt=0.01)
        psi0 = layers(p, axis, 1.0) * (1 + 0.2 * np.cos(2 * np.pi * (np.indices(p.shape)[axis] + 0.5) / 32))
        o = oracle.CompOracle(p, psi0)
        o.step(5)
        take = [0, 0, 0]
        take[axis] = slice(None)
        profs.append([o.psi[tuple(take)], o.rho[tuple(take)], o.vel[axis][tuple(take)]])
    for q in range(3):
        ref = profs[2][q]
        assert np.abs(ref).max() > 1e-6
        for axis in range(2):
            assert np.abs(profs[axis][q] - ref).max() < 1e-12 * np.abs(ref).max(), (axis, q)


def grids(p):
    dx = 2 * np.pi / (p.q0 * p.nw)
    q = [np.pi * np.arange(n) / (n * dx) for n in p.shape]
    s = [np.pi * (np.arange(n) + 1) / (n * dx) for n in p.shape]
    return q, s


def test_velocity_projection_structure(smadft, oracle):
    p = cparams(smadft, (8, 12, 16), nu=0.7, lambda_=1.3)
    rng = np.random.default_rng(5)
    S = [rng.standard_normal(p.shape) for _ in range(3)]
    G = rng.standard_normal(p.shape)
    vs, vr = oracle.comp_velocity(p, *S, G)
    nx, ny, nz = p.shape
    (qx, qy, qz), (sx, sy, sz) = grids(p)
    sc = 0.125 / (nx * ny * nz)
    # solenoidal: the staggered divergence at P = (i, j, k) vanishes wherever
    # no truncated component enters (j < ny-1, k < nz-1)
    div = np.zeros(p.shape)
    div[1:] += qx[1:, None, None] * vs[0][:-1]
    div[:, 1:] += qy[None, 1:, None] * vs[1][:, :-1]
    div[:, :, 1:] += qz[None, None, 1:] * vs[2][:, :, :-1]
    assert np.abs(div[:, :ny - 1, :nz - 1]).max() < 1e-12 * np.abs(vs).max() * qx.max()
    # irrotational: virr_a at the staggered point = q_a(P) Phi(P) with
    # Phi(P) = s (G(P) + D(P) g(P)/|q(P)|^2)/((lambda + 2 nu)|q(P)|^2),
    # D(P) the staggered divergence of S at P, g the Gaussian filter
    Sp = [np.pad(a, ((1, 0), (1, 0), (1, 0))) for a in S]   # Sp[a][i+1] = S[a][i], zero at -1
    D = (qx[:, None, None] * Sp[0][:-1, 1:, 1:] + qy[None, :, None] * Sp[1][1:, :-1, 1:]
         + qz[None, None, :] * Sp[2][1:, 1:, :-1])
    m2 = qx[:, None, None] ** 2 + qy[None, :, None] ** 2 + qz[None, None, :] ** 2
    g = np.exp(-2 * p.filter_sigma ** 2 * m2)
    with np.errstate(divide="ignore", invalid="ignore"):
        phi = sc * (G + D * g / m2) / ((p.lambda_ + 2 * p.nu) * m2)
    want_x = qx[1:, None, None] * phi[1:]
    assert np.abs(vr[0][:-1] - want_x).max() < 1e-12 * np.abs(want_x).max()
    assert np.abs(vr[0][-1]).max() == 0
    want_y = qy[None, 1:, None] * phi[:, 1:]
    assert np.abs(vr[1][:, :-1, :-1] - want_y[:, :, :-1]).max() < 1e-12 * np.abs(want_y).max()
    assert np.abs(vr[1][:, -1]).max() == 0 and np.abs(vr[1][:, :, -1]).max() == 0
    want_z = qz[None, None, 1:] * phi[:, :, 1:]
    quirk = np.zeros(p.shape, dtype=bool)
    quirk[1:, 1:, 0] = True
    ok = ~quirk[:, :-1, :-1]
    assert np.abs((vr[2][:, :-1, :-1] - want_z[:, :-1, :])[ok]).max() < 1e-12 * np.abs(want_z).max()
    # at k = 0, i >= 1, j >= 1 the filter of the q . S term takes the
    # y-staggered |q|^2 = qx^2 + sy^2 + qz(0)^2 (CM1y at :2350)
    i, j = np.meshgrid(np.arange(1, nx), np.arange(1, ny - 1), indexing="ij")
    mz = qx[i] ** 2 + qy[j] ** 2 + sz[0] ** 2
    my = qx[i] ** 2 + sy[j] ** 2
    Dz = D[i, j, 1]
    quirk_want = sc * sz[0] * (G[i, j, 1] + Dz * np.exp(-2 * p.filter_sigma ** 2 * my) / mz) / (
        (p.lambda_ + 2 * p.nu) * mz)
    assert np.abs(vr[2][i, j, 0] - quirk_want).max() < 1e-12 * np.abs(quirk_want).max()
    assert np.abs(vr[2][i, j, 0] - want_z[i, j, 0]).max() > 1e-6 * np.abs(quirk_want).max()


@pytest.mark.parametrize("up", [0, 3])
def test_relaxation_density_lift(smadft, oracle, up):
    # dt tiny: after one step rho equals the relaxation density to O(dt)
    p = cparams(smadft, (8, 8, 16), rho_up=up, div_switch=5, dt=1e-10)
    psi0 = smadft.ic_curved(p, 1.2, 1e-2, seed=2)
    o = oracle.CompOracle(p, psi0)
    o.step(1)
    rel = p.rho0 + p.kp * o.amp
    want = rel.copy()
    if up:
        want[:, :, up:] = rel[:, :, :-up]
        want[:, :, :up] = 1.0
    assert np.abs(o.rho - want).max() < 1e-8
    assert np.abs(o.amp).max() > 0.5


def test_load_start_uses_the_given_density(smadft, oracle):
    """load == 1: no relaxation; the given density enters lap(rho f), the
    stress and the penalty from the first step (:1542)."""
    p = cparams(smadft, (8, 8, 8), load=1, dt=0.01)
    psi0 = smadft.ic_curved(p, 1.2, 1e-2, seed=2)
    rho_in = 0.5 + 0.1 * np.cos(2 * np.pi * (np.indices(p.shape)[0] + 0.5) / 8)
    a = oracle.CompOracle(p, psi0, rho_in)
    a.step(1)
    b = oracle.CompOracle(p, psi0, np.full(p.shape, 0.5))
    b.step(1)
    assert np.abs(a.psi - b.psi).max() > 1e-6
    assert np.abs(a.rho - rho_in).max() < 0.05          # carried, not reset to rho0 + kp amp


def test_fc_torus_reference_geometry(smadft):
    p = cparams(smadft, (578, 484, 8))
    amp = np.sqrt((3 * 2.0 + np.sqrt(9 * 4.0 + 40 * p.eps)) / 5)
    psi = smadft.ic_fc_torus(p, amp, 6.0, 6.0, 5.2 / 6.2)
    dz = 2 * np.pi / (p.q0 * p.nw)
    # on the ring (ds = 121) at k = 0 the value is amp; the ring passes
    # through j = 242 + 121 at x_scale i = 242
    i = int(round(242 / (5.2 / 6.2)))
    xs = (5.2 / 6.2) * i - 242
    j = 242 + int(round(np.sqrt(121 ** 2 - xs ** 2)))
    ds = np.hypot(xs, j - 242)
    assert psi[i, j, 0] == pytest.approx(amp * np.cos(p.q0 * dz * abs(ds - 121)), rel=1e-12)
    assert psi[i, j, 7] == 0.0                               # above b_e + 1
    assert psi[0, 0, 0] == 0.0                               # far off the ring


def _real(oracle, hat, p):
    """Real field of an unscaled REDFT10^3 spectrum."""
    return oracle.r2r3d(hat, (1, 1, 1)) * (0.125 / (p.nx * p.ny * p.nz))


def _free_energies(oracle, p, psi, rho):
    """F = sum rho (f^2/2 - eps psi^2/2 - beta psi^4/4 + gamma psi^6/6) and
    F_pen = sum zeta/2 (rho - rho0 - kp amp)^2 on the grid, with
    f = alpha (q0^2 + lap) psi and amp = sqrt(q0^2 psi^2 + |grad psi|^2)
    from spectral derivatives (unfiltered)."""
    sc = 0.125 / (p.nx * p.ny * p.nz)
    (qx, qy, qz), (sx, sy, sz) = grids(p)
    ph = oracle.r2r3d(psi, (0, 0, 0)) * sc
    m2 = qx[:, None, None] ** 2 + qy[None, :, None] ** 2 + qz[None, None, :] ** 2
    f = oracle.r2r3d(p.alpha * (p.q0 ** 2 - m2) * ph, (1, 1, 1))
    g2 = np.zeros(p.shape)
    for ax, (kinds, s) in enumerate((((3, 1, 1), sx), ((1, 3, 1), sy), ((1, 1, 3), sz))):
        sh = np.zeros(p.shape)
        src = [slice(None)] * 3
        dst = [slice(None)] * 3
        src[ax], dst[ax] = slice(1, None), slice(0, -1)
        shape = [1, 1, 1]
        shape[ax] = -1
        sh[tuple(dst)] = -s[:-1].reshape(shape) * ph[tuple(src)]
        g2 += oracle.r2r3d(sh, kinds) ** 2
    amp = np.sqrt(p.q0 ** 2 * psi ** 2 + g2)
    F = np.sum(rho * (0.5 * f ** 2 - 0.5 * p.eps * psi ** 2 - 0.25 * p.beta * psi ** 4 + p.gamma / 6 * psi ** 6))
    Fp = np.sum(0.5 * p.zeta * (rho - p.rho0 - p.kp * amp) ** 2)
    return F, Fp


def _nl_first_step(oracle, p, psi, rho):
    o = oracle.CompOracle(p, psi, rho)
    o.step(1)
    return _real(oracle, o.nl, p)


def _gradient_flow_params(smadft, zeta):
    # load == 1 keeps the given density; nu, lambda huge make u . grad psi
    # negligible; no filter, so amp is the field whose penalty F_pen holds
    return cparams(smadft, (8, 8, 8), load=1, nu=1e12, lambda_=1e12, filter_sigma=0.0, zeta=zeta, q0=1.0,
                   alpha=1.0, dt=1e-3)


def _smooth_fields(p, seed):
    rng = np.random.default_rng(seed)
    x = (np.indices(p.shape) + 0.5) / np.array(p.shape)[:, None, None, None]
    psi = 1.0 + 0.25 * np.cos(np.pi * x[0]) * np.cos(2 * np.pi * x[1]) + 0.15 * np.cos(np.pi * x[2])
    rho = 0.6 + 0.2 * np.cos(np.pi * x[0] + 0.3) * np.cos(np.pi * x[2])
    v = rng.standard_normal(p.shape)
    return psi, rho, v


def test_nl_is_minus_the_free_energy_gradient(smadft, oracle):
    """With no flow and zeta = 0, Nl = -mu = -dF/dpsi: mu = grad^2(rho f) +
    rho (f - eps psi - beta psi^3 + gamma psi^5) is the variational
    derivative of F for alpha = q0 = 1 (:1653-1655, :2581-2582)."""
    p = _gradient_flow_params(smadft, 0.0)
    psi, rho, v = _smooth_fields(p, 11)
    nl = _nl_first_step(oracle, p, psi, rho)
    h = 1e-5  # central-difference error ~1e-9 relative
    dF = (_free_energies(oracle, p, psi + h * v, rho)[0] - _free_energies(oracle, p, psi - h * v, rho)[0]) / (2 * h)
    assert abs(np.sum(nl * v) + dF) < 1e-8 * abs(dF)


def test_penalty_terms_are_minus_the_penalty_gradient(smadft, oracle):
    """The zeta terms of Nl, (kp zeta/(amp + 0.001)) w psi - div(kp zeta w
    grad psi/(amp + 0.001)) with w = rho - rho0 - kp amp (:1665-1668,
    :2583-2585), are -dF_pen/dpsi up to the 0.001 regularisation."""
    p0 = _gradient_flow_params(smadft, 0.0)
    p1 = _gradient_flow_params(smadft, 1.0)
    psi, rho, v = _smooth_fields(p0, 12)
    d = _nl_first_step(oracle, p1, psi, rho) - _nl_first_step(oracle, p0, psi, rho)
    h = 1e-5
    dF = (_free_energies(oracle, p1, psi + h * v, rho)[1] - _free_energies(oracle, p1, psi - h * v, rho)[1]) / (2 * h)
    assert abs(dF) > 1e-3
    assert abs(np.sum(d * v) + dF) < 3e-3 * abs(dF)


def _div_rho_u(oracle, p, rho, u):
    """Spectral divergence of rho u (u_a in the DST basis of axis a)."""
    sc = 0.125 / (p.nx * p.ny * p.nz)
    (qx, qy, qz), _ = grids(p)
    qs = (qx, qy, qz)
    out = np.zeros(p.shape)
    for ax in range(3):
        kinds = [0, 0, 0]
        kinds[ax] = 2                      # RODFT10 on the own axis
        hat = oracle.r2r3d(rho * u[ax], tuple(kinds)) * sc
        d = np.zeros(p.shape)
        src = [slice(None)] * 3
        dst = [slice(None)] * 3
        src[ax], dst[ax] = slice(0, -1), slice(1, None)
        shape = [1, 1, 1]
        shape[ax] = -1
        d[tuple(dst)] = qs[ax][1:].reshape(shape) * hat[tuple(src)]
        out += oracle.r2r3d(d, (1, 1, 1))
    return out


def test_continuity_is_minus_div_rho_u_with_ab2(smadft, oracle):
    """R = -u . grad rho - rho divv equals -div(rho u) for smooth fields
    (product rule), and rho advances by dt (3 R_n - R_{n-1})/2, so over many
    steps by about dt sum R (:2559-2572)."""
    p = cparams(smadft, (16, 16, 16), load=1, nu=0.05, lambda_=0.05, dt=2e-3, zeta=0.05)
    x = (np.indices(p.shape) + 0.5) / 16.0
    psi = 1.2 * np.cos(np.pi * x[0]) * np.cos(np.pi * x[1]) + 0.3 * np.cos(np.pi * x[2])
    rho = 0.7 + 0.2 * np.cos(np.pi * x[0]) * np.cos(np.pi * x[2])
    o = oracle.CompOracle(p, psi, rho)
    rho_start = o.rho.copy()
    rsum = np.zeros(p.shape)
    for n in range(12):
        before = o.rho.copy()
        o.step(1)
        r = o.frho.copy()
        want = -_div_rho_u(oracle, p, before, o.vel)
        assert np.abs(r).max() > 1e-4
        assert np.abs(r - want).max() < 0.02 * np.abs(want).max(), n
        rsum += r
    drho = o.rho - rho_start
    assert np.abs(drho - p.dt * rsum).max() < 0.1 * np.abs(drho).max()


@pytest.mark.parametrize("mode_axis", [0, 2])
def test_density_mode_relaxes_at_the_bulk_rate(smadft, oracle, mode_axis):
    """Uniform psi and a small density mode: the force (f^2/2) grad rho is the
    gradient of rho f^2/2, which the filtered rho e0 term of G cancels, so the
    flow is driven by the pressure fgrad = zeta (rho^2 - rho_eq^2)/2 alone:
    u = grad lap^-1 fgrad/(lambda + 2 nu), div u = fgrad/(lambda + 2 nu), and
    the mode decays under AB2 at the rate a = zeta rho^2/(lambda + 2 nu) for
    every wavenumber (:1641-1651, :1826-1834, :1940-2410, :2555-2573).  The
    first step is exact; later psi picks up O(delta) variations from
    grad^2(rho f), which kp = 0 keeps out of rho_eq."""
    p = cparams(smadft, (8, 8, 8), load=1, nu=0.3, lambda_=0.2, zeta=2.0, dt=0.01, kp=0.0)
    rbar, d0 = 0.8, 1e-5
    x = (np.indices(p.shape)[mode_axis] + 0.5) / 8.0
    mode = np.cos(np.pi * x)
    o = oracle.CompOracle(p, np.full(p.shape, 0.7), rbar + d0 * mode)
    a = p.zeta * rbar ** 2 / (p.lambda_ + 2 * p.nu)
    coef = lambda fld: np.sum((fld - fld.mean()) * mode) / np.sum(mode * mode)
    d, r_old = d0, 0.0
    for n in range(40):
        o.step(1)
        if n == 0:
            assert coef(o.frho) == pytest.approx(-a * d0, rel=1e-8)
            assert np.abs(o.vsol).max() < 1e-12 * np.abs(o.virr).max()
        r = -a * d
        d = d + p.dt * (1.5 * r - 0.5 * r_old)
        r_old = r
        assert abs(coef(o.rho) - d) < 0.01 * (d0 - d), n
    assert d < 0.6 * d0  # decayed visibly


def test_amp_is_the_gaussian_filtered_amplitude(smadft, oracle):
    """Each cosine mode of amp is that of sqrt(q0^2 psi^2 + |grad psi|^2)
    times exp(-2 sigma^2 |q|^2) (:1477-1515): the ratio of the spectra is the
    Gaussian, the mean is kept."""
    p = cparams(smadft, (8, 8, 16), q0=1.1, filter_sigma=0.8)
    psi = smadft.ic_curved(p, 1.2, 1e-2, seed=4)
    o = oracle.CompOracle(p, psi)
    o.step(1)
    (qx, qy, qz), (sx, sy, sz) = grids(p)
    sc = 0.125 / (p.nx * p.ny * p.nz)
    ph = oracle.r2r3d(psi, (0, 0, 0)) * sc
    g2 = np.zeros(p.shape)
    for ax, (kinds, s) in enumerate((((3, 1, 1), sx), ((1, 3, 1), sy), ((1, 1, 3), sz))):
        sh = np.zeros(p.shape)
        src = [slice(None)] * 3
        dst = [slice(None)] * 3
        src[ax], dst[ax] = slice(1, None), slice(0, -1)
        shape = [1, 1, 1]
        shape[ax] = -1
        sh[tuple(dst)] = -s[:-1].reshape(shape) * ph[tuple(src)]
        g2 += oracle.r2r3d(sh, kinds) ** 2
    raw = oracle.r2r3d(np.sqrt(p.q0 ** 2 * psi ** 2 + g2), (0, 0, 0))
    got = oracle.r2r3d(o.amp, (0, 0, 0))
    m2 = qx[:, None, None] ** 2 + qy[None, :, None] ** 2 + qz[None, None, :] ** 2
    big = np.abs(raw) > 1e-6 * np.abs(raw).max()
    assert big.sum() > 100
    assert np.abs(got[big] / raw[big] - np.exp(-2 * p.filter_sigma ** 2 * m2[big])).max() < 1e-9
