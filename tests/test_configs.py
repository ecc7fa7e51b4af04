"""The benchmark configurations: presets, seeded initial conditions, and the
time steps the configurations specify for the advective (0.05) and quasi
(0.02) models, which the oracle shows to diverge (the presets take 0.01 and
the reference's 1e-3; at 256^3 the advective step also diverges at 0.02,
where the bumped interfaces first drive |u| ~ 0.5)."""
import numpy as np
import pytest


def test_presets(smadft):
    want = {0: (0, 128, 8.0, 0.1), 1: (0, 256, 8.0, 0.1), 2: (1, 256, 8.0, 0.01), 3: (2, 384, 8.0, 0.001),
            4: (0, 256, 8.0, 0.1)}
    for c, (model, n, nw, dt) in want.items():
        p = smadft.Params.config(c)
        assert (p.model, p.nx, p.ny, p.nz, p.nw, p.dt) == (model, n, n, n, nw, dt)
        assert (p.eps, p.alpha, p.beta, p.gamma, p.q0) == (-0.7, 1.0, 2.0, 1.0, 1.0)
    p = smadft.Params.config(3)
    assert p.rho0 == 0.9 and p.kp == pytest.approx(0.1 / 1.34164)
    # rho_smectic = kp A + rho0 = 1.0 at the amplitude 1.34164, rho_m = rho_s/2
    assert p.kp * 1.34164 + p.rho0 == pytest.approx(1.0) and p.rho_m == pytest.approx(0.5)
    assert smadft.Params.config(2).nu == 1.0
    with pytest.raises(smadft.SmadError):
        smadft.Params.config(5)


def test_program_defaults(smadft):
    """smad_default_params: the reference programs' own constants."""
    a = smadft.Params.default(smadft.MODEL_ADV, 64)
    assert (a.nw, a.dt) == (16.0, 0.0005)  # adv/cosAdvTrackMult.cpp:181-182
    q = smadft.Params.default(smadft.MODEL_QUASI, 64)
    assert (q.nw, q.dt, q.rho0, q.kp, q.rho_m, q.div_switch) == (8.0, 0.001, 1.0, 0.0, 0.5, 200)  # quasi.cpp:204-252
    assert q.lambda_ == pytest.approx(-2.0 / 3.0)


def test_two_bumps_ic(smadft):
    p = smadft.Params.config(2, 64)
    a = smadft.ic_two_bumps(p, 1.328, 0.0, seed=2)
    b = smadft.ic_two_bumps(p, 1.328, 0.0, seed=2)
    c = smadft.ic_two_bumps(p, 1.328, 0.0, seed=9)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    h = 2 * np.pi / (p.q0 * p.nw)
    x = np.arange(p.nx) * h
    inside = np.abs(a).max(axis=2) > 0  # (i, j) columns holding layers
    # each (y, z) line of the slab: [lo, hi] widens around the bumps only
    width = inside.sum(axis=0)
    flat = int(np.sum(np.abs(x - p.nx * h / 2) < p.nx * h / 4))
    assert width.min() >= flat - 1
    assert width.max() >= flat + 2  # a bump bulges out by at least a layer spacing
    i_mid = p.nx // 2  # the slab's middle holds flat layers amp cos(q0 k dz)
    assert np.allclose(a[i_mid, 5, :], 1.328 * np.cos(p.q0 * np.arange(p.nz) * h), atol=1e-15)
    assert np.array_equal(smadft.config_ic(2, p), smadft.ic_two_bumps(p, 1.328, 1e-3, seed=2))


def test_tilted_slab_ic(smadft):
    p = smadft.Params.config(3, 32)
    tilts = [smadft.ic_tilted_slab(p, 1.34164, 15.0, 0.0, seed=s)[1] for s in range(40)]
    assert all(0 <= t <= 15.0 for t in tilts)
    assert max(tilts) > 10 and min(tilts) < 5  # spread over the range
    flat, t0 = smadft.ic_tilted_slab(p, 1.34164, 0.0, 0.0, seed=3)
    assert t0 == 0.0
    assert np.allclose(flat, smadft.ic_slab(p, 1.34164, 0.0, seed=0), atol=1e-15)
    psi, t = smadft.ic_tilted_slab(p, 1.34164, 15.0, 1e-3, seed=3)
    assert np.array_equal(smadft.config_ic(3, p), psi)
    # the layers' normal makes the angle t with z: the dominant spectral
    # direction of the slab interior
    h = 2 * np.pi / (p.q0 * p.nw)
    inner = psi[p.nx // 4 + 2: 3 * p.nx // 4 - 2]
    g = np.gradient(inner, h)
    gz = np.abs(g[2]).mean()
    gxy = np.hypot(np.abs(g[0]).mean(), np.abs(g[1]).mean())
    assert np.degrees(np.arctan2(gxy, gz)) == pytest.approx(t, abs=3.0)


def test_specified_time_steps_diverge(smadft, oracle):
    """Config 2's dt = 0.05 and config 3's dt = 0.02 give NaN within 50 steps
    on their own initial conditions (32^3, the configurations' spacing),
    while the presets' 0.01 and 1e-3 stay bounded."""
    p = smadft.Params.config(2, 32)
    psi0 = smadft.config_ic(2, p)
    for dt, ok in ((0.05, False), (0.01, True)):
        p.dt = dt
        o = oracle.AdvOracle(p, psi0)
        o.step(50)
        assert bool(np.isfinite(o.psi).all() and np.abs(o.psi).max() < 2) == ok, dt
    q = smadft.Params.config(3, 32)
    psi0 = smadft.config_ic(3, q)
    for dt, ok in ((0.02, False), (0.001, True)):
        q.dt = dt
        o = oracle.QuasiOracle(q, psi0)
        o.step(50)
        assert bool(np.isfinite(o.psi).all() and np.abs(o.psi).max() < 2) == ok, dt
