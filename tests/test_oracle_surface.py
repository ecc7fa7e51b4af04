"""Pins the oracle's surface diagnostics (no-adv/cosNoAdvTrackMult.cpp:787-796,
853-1022; the |psi|-envelope scan of no-adv/cosNoAdvTrackBump.cpp:929-1031)
against closed forms: single cosine modes for the auxiliary spectral fields,
spherical and planar level sets for the curvatures, hand-built columns for
the scans and an independent vectorised formulation of the envelope index."""
import numpy as np
import pytest


def plain_params(smadft, n):
    p = smadft.Params.default(smadft.MODEL_PLAIN, n)
    return p


def test_aux_fields_single_mode(smadft, oracle):
    p = plain_params(smadft, 16)
    p.nx, p.ny, p.nz = 16, 12, 20
    h = 2 * np.pi / (p.q0 * p.nw)
    L = np.array([p.nx, p.ny, p.nz]) * h
    m = (3, 2, 5)
    idx = np.meshgrid(*[np.arange(n) + 0.5 for n in (p.nx, p.ny, p.nz)], indexing="ij")
    q = [np.pi * mi / Li for mi, Li in zip(m, L)]
    psi = 0.8 * np.cos(q[0] * idx[0] * h) * np.cos(q[1] * idx[1] * h) * np.cos(q[2] * idx[2] * h)
    dxx, dyy, dzz, dt = oracle.surface_aux(p, psi)
    assert np.allclose(dxx, -q[0] ** 2 * psi, atol=1e-12)
    assert np.allclose(dyy, -q[1] ** 2 * psi, atol=1e-12)
    assert np.allclose(dzz, -q[2] ** 2 * psi, atol=1e-12)
    q2 = sum(qi ** 2 for qi in q)
    alin = p.eps - p.alpha * (q2 - p.q0 ** 2) ** 2
    assert np.allclose(dt, alin * psi + p.beta * psi ** 3 - p.gamma * psi ** 5, atol=1e-12)


def radial_case(n=64, R=20.0, w=8.0, c=(31.3, 30.6, 14.2)):
    """psi = tanh((r - R)/w) (grid units), with its analytic second
    derivatives; level sets are spheres."""
    x, y, z = np.meshgrid(*[np.arange(n, dtype=float)] * 3, indexing="ij")
    X, Y, Z = x - c[0], y - c[1], z - c[2]
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    t = np.tanh((r - R) / w)
    f1 = (1 - t ** 2) / w
    f2 = -2 * t * (1 - t ** 2) / w ** 2
    d2 = [f2 * (A / r) ** 2 + f1 * (1 / r - A ** 2 / r ** 3) for A in (X, Y, Z)]
    return t, d2, (X, Y, Z, r), f1


@pytest.mark.parametrize("c", [(31.3, 30.6, 14.2), (22.7, 39.1, 12.4)])
def test_curvatures_of_spheres(oracle, c):
    """The level set through a point at distance r from the centre is a
    sphere: 2H = 2/r, K = 1/r^2.  Off the column axes every term of both
    formulas (the mixed derivatives included) contributes."""
    R = 20.0
    psi, (dxx, dyy, dzz), (X, Y, Z, r), f1 = radial_case(R=R, c=c)
    out = oracle.surface_track(psi, dxx, dyy, dzz, np.ones_like(psi), 1.0)
    s = out["surf"]
    ii, jj = np.nonzero(s >= 0)
    assert len(ii) > 1000
    k2 = s[ii, jj].astype(int)
    rp = r[ii, jj, k2]
    # the crossing sits next to the sphere
    assert np.all(np.abs(rp - R) < 1.0)
    # columns away from the sphere's rim, where the gradient turns horizontal
    Xp, Yp = X[ii, jj, k2], Y[ii, jj, k2]
    inner = (Xp ** 2 + Yp ** 2) < (0.7 * R) ** 2
    # the tested points exercise the mixed terms: n_x n_y up to ~0.2
    assert np.max(np.abs(Xp * Yp)[inner] / rp[inner] ** 2) > 0.2
    h = out["curv_h"][ii, jj][inner]
    kk = out["curv_k"][ii, jj][inner]
    assert np.max(np.abs(h * rp[inner] / 2 - 1)) < 1e-2
    assert np.max(np.abs(kk * rp[inner] ** 2 - 1)) < 2e-2
    # vel = dTpsi / |grad psi| with dTpsi = 1 and |grad psi| = f'(r) up to O(h^2)
    v = out["vel"][ii, jj][inner]
    assert np.max(np.abs(v * f1[ii, jj, k2][inner] - 1)) < 1e-2


def test_tilted_plane_has_no_curvature(oracle):
    n = 40
    x, y, z = np.meshgrid(*[np.arange(n, dtype=float)] * 3, indexing="ij")
    th = 0.3
    s = np.sin(th) * (x - 20) + np.cos(th) * (z - 20)
    w = 4.0
    psi = np.tanh(s / w)
    t = psi
    f2 = -2 * t * (1 - t ** 2) / w ** 2
    dxx, dzz = f2 * np.sin(th) ** 2, f2 * np.cos(th) ** 2
    out = oracle.surface_track(psi, dxx, np.zeros_like(psi), dzz, np.ones_like(psi), 1.0)
    ii, jj = np.nonzero(out["surf"] >= 0)
    assert len(ii) == n * n
    # curvature scale 1/w: the residual is the O(h^2) finite-difference error
    assert np.max(np.abs(out["curv_h"])) < 2e-2 / w
    assert np.max(np.abs(out["curv_k"])) < 1e-6 / w ** 2


def test_scan_rules(oracle):
    nx, ny, nz = 3, 3, 16
    psi = np.full((nx, ny, nz), 0.3)
    # (0,0): arm at k=12, crossing at k=9 with |psi(9)| < |psi(10)|  -> 9
    psi[0, 0, 12] = 0.9
    psi[0, 0, 10] = 0.2
    psi[0, 0, 9] = -0.1
    # (0,1): crossing at k=9 but |psi(9)| > |psi(10)|                 -> 10
    psi[0, 1, 12] = 0.9
    psi[0, 1, 10] = 0.05
    psi[0, 1, 9] = -0.3
    # (0,2): never above 0.7                                           -> -1
    psi[0, 2, 5] = -0.5
    # (1,0): negative above the arming point is ignored                -> 3
    psi[1, 0, 14] = -0.4
    psi[1, 0, 8] = 0.8
    psi[1, 0, 3] = -0.2
    psi[1, 0, 4] = 0.25
    # (1,1): armed, never negative                                     -> -1
    psi[1, 1, 7] = 0.75
    # (1,2): only the first crossing counts                            -> 6
    psi[1, 2, 11] = 0.71
    psi[1, 2, 6] = -0.2
    psi[1, 2, 7] = 0.4
    psi[1, 2, 2] = -0.6
    z = np.zeros_like(psi)
    out = oracle.surface_track(psi, z, z, z, z, 0.5)
    assert out["surf"][0, 0] == 9
    assert out["surf"][0, 1] == 10
    assert out["surf"][0, 2] == -1
    assert out["surf"][1, 0] == 3
    assert out["surf"][1, 1] == -1
    assert out["surf"][1, 2] == 6
    for key in ("vel", "curv_h", "curv_k"):
        assert out[key][0, 2] == 0 and out[key][1, 1] == 0


# ---------------------------------------------------------------------------
# |psi|-envelope scan (no-adv/cosNoAdvTrackBump.cpp:929-1031): the first
# |psi| > 0.6 below the top, a walk up the flat index to |psi| <= 0.1, the
# next point taken when smaller in magnitude; the derivatives are taken at the
# |psi| > 0.6 point unless that swap happened.


def columns_grid(profile, nxy=3):
    """A grid whose columns all equal `profile`: the x and y differences vanish
    and |grad psi| = |d psi/dz|."""
    return np.broadcast_to(np.asarray(profile, float), (nxy, nxy, len(profile))).copy()


def dz_at(profile, k, h):
    """The z difference at k with the one-sided top and bottom rule."""
    f, n = np.asarray(profile, float), len(profile)
    if 0 < k < n - 1:
        return (f[k + 1] - f[k - 1]) / (2 * h)
    return (f[1] - f[0]) / h if k == 0 else (f[-1] - f[-2]) / h


@pytest.mark.parametrize("profile,surf,k_at", [
    # walk 12 -> 14 (0.05 <= 0.1); |0.05| > |0.02|: the next point, derivatives there
    ({12: 0.9, 13: 0.3, 14: 0.05, 15: 0.02}, 15, 15),
    # no swap (|0.05| < |0.08|): surface at 14, derivatives at the 0.6 point 12
    ({12: 0.9, 13: 0.3, 14: 0.05, 15: 0.08}, 14, 12),
    # negative values count by magnitude; the first from the top wins
    ({4: 0.95, 10: -0.7, 11: -0.2, 12: 0.05, 13: 0.0}, 13, 13),
    # |psi| > 0.1 up to the top: the walk runs into the next column's k = 0
    # (flat index), where psi = 0 stops it; no swap (0 > |psi(k=1)| is false)
    ({1: 0.04, 12: 0.9, 13: 0.5, 14: 0.5, 15: 0.5}, 16, 12),
])
def test_envelope_scan_rules(oracle, profile, surf, k_at):
    nz, h = 16, 0.5
    prof = np.zeros(nz)
    for k, v in profile.items():
        prof[k] = v
    psi = columns_grid(prof)
    dtp = np.broadcast_to(np.arange(nz, dtype=float) + 1, psi.shape).copy()  # marks the point used
    z = np.zeros_like(psi)
    out = oracle.surface_track(psi, z, z, z, dtp, h, envelope=True)
    assert np.all(out["surf"] == surf)
    g = abs(dz_at(prof, k_at, h))
    assert np.allclose(out["vel"], (k_at + 1) / g, rtol=1e-14)


def test_envelope_without_large_values(oracle):
    psi = columns_grid(np.full(12, 0.55))
    z = np.zeros_like(psi)
    out = oracle.surface_track(psi, z, z, z, z, 1.0, envelope=True)
    assert np.all(out["surf"] == -1)
    assert not out["vel"].any() and not out["curv_h"].any() and not out["curv_k"].any()


def test_envelope_surface_index_vectorised(oracle):
    """The surface index against an independent formulation: numpy masks over
    the flat |psi| array instead of the per-column loop."""
    rng = np.random.default_rng(7)
    nx, ny, nz = 6, 5, 24
    # smooth columns: a random envelope height, layers below it, noise above
    top = rng.integers(6, 20, size=(nx, ny))
    k = np.arange(nz)
    psi = np.where(k[None, None, :] <= top[..., None],
                   np.cos(0.9 * (k[None, None, :] - top[..., None])) * rng.uniform(0.7, 1.0, (nx, ny, 1)),
                   rng.uniform(-0.3, 0.3, (nx, ny, nz)) * (rng.random((nx, ny, nz)) < 0.5))
    z = np.zeros_like(psi)
    out = oracle.surface_track(psi, z, z, z, np.ones_like(psi), 1.0, envelope=True)
    a = np.abs(psi).ravel()
    n = a.size
    expect = np.full((nx, ny), -1.0)
    for i in range(nx):
        for j in range(ny):
            c = i * ny + j
            big = np.nonzero(a[c * nz:(c + 1) * nz] > 0.6)[0]
            if big.size == 0:
                continue
            p0 = c * nz + big.max()
            tail = np.append(a[p0:], 0.0)  # zero past the end of the grid
            p = p0 + int(np.argmax(tail <= 0.1))
            if p - p0 > nz:
                continue
            nxt = a[p + 1] if p + 1 < n else 0.0
            cur = a[p] if p < n else 0.0
            expect[i, j] = (p + (1 if cur > nxt else 0)) - c * nz
    assert (expect >= 0).sum() > 20
    assert np.array_equal(out["surf"], expect)


def test_envelope_curvatures_of_spherical_level_sets(oracle):
    """psi = (1 - tanh((r - R)/w))/2: every level set is a sphere, so at any
    point with a nonzero gradient (curv_h/2)^2 = curv_k, 2H = -2/r for this
    inward-increasing field, and the point lies between the 0.1 and 0.6
    levels."""
    n, R, w, c = 64, 18.0, 4.0, (31.3, 30.6, 12.2)
    x, y, zz = np.meshgrid(*[np.arange(n, dtype=float)] * 3, indexing="ij")
    X, Y, Z = x - c[0], y - c[1], zz - c[2]
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    t = np.tanh((r - R) / w)
    psi = 0.5 * (1 - t)
    f1 = -0.5 * (1 - t ** 2) / w
    f2 = t * (1 - t ** 2) / w ** 2
    d2 = [f2 * (A / r) ** 2 + f1 * (1 / r - A ** 2 / r ** 3) for A in (X, Y, Z)]
    out = oracle.surface_track(psi, *d2, np.ones_like(psi), 1.0, envelope=True)
    ii, jj = np.nonzero(out["surf"] >= 0)
    k2 = out["surf"][ii, jj].astype(int)
    inner = (X[ii, jj, 0] ** 2 + Y[ii, jj, 0] ** 2) < (0.6 * R) ** 2
    assert inner.sum() > 300
    # the surface index sits just above the 0.1 level
    assert np.all(psi[ii, jj, k2][inner] <= 0.1 + 1e-12)
    h = out["curv_h"][ii, jj][inner]
    kk = out["curv_k"][ii, jj][inner]
    assert np.all(h < 0)
    assert np.max(np.abs((h / 2) ** 2 / kk - 1)) < 3e-2
    # 1/r between the radii of the 0.6 level and two grid steps past the 0.1
    # level (the walk's end and the swap to the next point)
    rmin, rmax = R + w * np.arctanh(1 - 2 * 0.6) - 1.0, R + w * np.arctanh(1 - 2 * 0.1) + 2.5
    assert np.all(-h / 2 > 1 / rmax) and np.all(-h / 2 < 1 / rmin)


def test_gradient_array_scan_of_spheres(oracle):
    """The advective program's scan reads gradient arrays instead of
    differencing psi (adv/cosAdvTrackMult.cpp:1431-1435): fed the analytic
    gradient of spherical level sets it gives 2H = 2/r and K = 1/r^2."""
    R, c = 20.0, (31.3, 30.6, 14.2)
    psi, (dxx, dyy, dzz), (X, Y, Z, r), f1 = radial_case(R=R, c=c)
    grad = [f1 * A / r for A in (X, Y, Z)]
    out = oracle.surface_track_grad(psi, grad, dxx, dyy, dzz, np.ones_like(psi), 1.0)
    s = out["surf"]
    ii, jj = np.nonzero(s >= 0)
    assert len(ii) > 1000
    rp = r[ii, jj, s[ii, jj].astype(int)]
    assert np.allclose(out["curv_h"][ii, jj], 2 / rp, rtol=1e-2)
    assert np.allclose(out["curv_k"][ii, jj], 1 / rp ** 2, rtol=2e-2)
    # vel = dTpsi/|grad psi| with the given gradient
    assert np.allclose(out["vel"][ii, jj], 1 / np.abs(f1[ii, jj, s[ii, jj].astype(int)]), rtol=1e-12)


def test_adv_surface_is_the_scan_of_closed_form_fields(smadft, oracle):
    """orc_adv_surface (adv/cosAdvTrackMult.cpp:1311-1347, 1412-1500) equals
    the gradient-array scan of closed forms: for psi a sum of cosine modes,
    psi_xx = sum -qx^2 a_m m, grad psi analytic, and dTpsi = d psi/dt of the
    advective model = sum L(q_m) a_m m + beta psi^3 - gamma psi^5 - u . grad
    psi with u from the (pinned) Stokes solve of the closed-form force."""
    import forms
    from test_oracle_adv import first_step_closed_form
    p = smadft.Params.default(smadft.MODEL_ADV, 16)
    p.nx, p.ny, p.nz = 16, 16, 32
    p.nw, p.dt = 8.0, 0.02
    am = [(1.2, (0, 0, 7)), (0.35, (2, 1, 3)), (0.25, (1, 3, 5))]
    psi = np.zeros(p.shape)
    grad = [np.zeros(p.shape) for _ in range(3)]
    d2 = [np.zeros(p.shape) for _ in range(3)]
    dtp = np.zeros(p.shape)
    for amp, m in am:
        f, g = forms.mode(p, *m)
        q = [w[i] for w, i in zip(forms.wavenumbers(p), m)]
        psi += amp * f
        dtp += forms.alin(p, forms.mq2(p, *m)) * amp * f
        for ax in range(3):
            grad[ax] += amp * g[ax]
            d2[ax] += -q[ax] ** 2 * amp * f
    _, u, _ = first_step_closed_form(p, am)
    dtp += p.beta * psi ** 3 - p.gamma * psi ** 5 - sum(u[ax] * grad[ax] for ax in range(3))
    h = forms.spacing(p)
    ref = oracle.surface_track_grad(psi, grad, *d2, dtp, h)
    got = oracle.adv_surface(p, psi)
    assert (ref["surf"] >= 0).sum() > p.nx * p.ny // 2
    assert np.array_equal(got["surf"], ref["surf"])
    for key in ("vel", "curv_h", "curv_k"):
        assert np.abs(got[key] - ref[key]).max() < 1e-10 * np.abs(ref[key]).max(), key
