"""Full restart state (smad_state_fields / smad_set_steps_taken) and the
quasi model's load == 1 start (quasi/quasi.cpp:1365-1369) on the GPU.

A run interrupted after K steps, saved through smad_fetch and restored into
a fresh context with smad_set_field, continues bit-identically to the
uninterrupted run; the reference's own restart keeps psi only and re-runs
the pre-loop (no-adv/cosNoAdvTrackMult.cpp:517-537), which smad_set_field
(PSI) still does."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def params(smadft, model, shape, prec):
    p = smadft.Params.default(model, shape[0])
    p.nx, p.ny, p.nz = shape
    p.nw = 8.0
    p.precision = prec
    if model == smadft.MODEL_ADV:
        p.dt = 0.02
    if model == smadft.MODEL_QUASI:
        p.kp, p.rho0 = 0.1 / 1.34164, 0.9
        p.rho_m = (p.kp * 1.34164 + p.rho0) / 2
        p.div_switch = 4
    return p


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("model", [0, 1, 2])
def test_full_state_restart_is_bit_identical(smadft, model, prec):
    p = params(smadft, model, (16, 32, 32), prec)
    psi0 = smadft.ic_curved(p, 1.34164 if model == 2 else 1.328, 1e-3, seed=7)
    K, M = 5, 4  # quasi: the switch (4) and the coupling term (from step 6) fall after the restart
    print(f"restart model={model} prec={prec}", flush=True)
    with smadft.Stepper(p, psi0) as a:
        a.step(K + M)
        want = a.fetch()
        want_l1 = a.reduce()[:2]
    with smadft.Stepper(p, psi0) as b:
        b.step(K)
        state = b.save_state()
    assert state[0] == K
    with smadft.Stepper(p, np.zeros(p.shape)) as c:
        c.load_state(state)
        assert c.steps_taken == K
        c.step(M)
        got = c.fetch()
        got_l1 = c.reduce()[:2]
    assert np.array_equal(got, want)
    assert np.array_equal(got_l1, want_l1)


def test_psi_only_restart_reruns_the_preloop(smadft):
    """smad_set_field(PSI) is the reference's restart: the AB2 history is
    rebuilt from psi (first step Euler in N), so it departs from the
    uninterrupted run at roundoff-times-dt level, not bit-identically."""
    p = params(smadft, 0, (16, 32, 32), 64)
    psi0 = smadft.ic_curved(p, 1.328, 1e-3, seed=7)
    with smadft.Stepper(p, psi0) as a:
        a.step(8)
        want = a.fetch()
    with smadft.Stepper(p, psi0) as b:
        b.step(4)
        mid = b.fetch()
        b.set_field(smadft.FIELD_PSI, mid)
        b.step(4)
        got = b.fetch()
    err = rel_l2(got, want)
    assert 0 < err < 1e-3


@pytest.mark.parametrize("load", [0, 1])
def test_quasi_load_start_matches_oracle(smadft, oracle, load):
    p = params(smadft, 2, (16, 32, 32), 64)
    p.div_switch = 200
    p.load = load
    psi0 = smadft.ic_curved(p, 1.34164, 1e-3, seed=3)
    print(f"quasi load={load}", flush=True)
    with smadft.Stepper(p, psi0) as s:
        s.step(6)
        psi = s.fetch()
        dv = s.fetch(smadft.FIELD_DIVV_HAT)
    q = oracle.QuasiOracle(p, psi0)
    q.step(6)
    assert rel_l2(psi, q.psi) < 1e-10
    if load:
        assert rel_l2(dv, q.divv) < 1e-9
    else:
        assert np.abs(dv).max() == 0.0


def test_quasi_save_block_outputs_match_oracle(smadft, oracle):
    """vsol, the projected force, S, divv (quasi/quasi.cpp:2379-2487) from the
    spectral fields the step leaves behind, against the oracle's force and
    projection of the same step."""
    import forms
    p = params(smadft, 2, (16, 32, 32), 64)
    p.div_switch = 2
    psi0 = smadft.ic_curved(p, 1.34164, 1e-3, seed=3)
    with smadft.Stepper(p, psi0) as s:
        s.step(5)
        out = {f: s.fetch(f) for f in range(smadft.FIELD_VSOL_X, smadft.FIELD_DIVV + 1)}
        grho = [s.fetch(f) for f in (smadft.FIELD_GRAD_RHO_X, smadft.FIELD_GRAD_RHO_Y, smadft.FIELD_GRAD_RHO_Z)]
    q = oracle.QuasiOracle(p, psi0)
    q.step(4)
    rhoq_prev = q.rhoq.copy()
    q.step(1)
    # this step's force with the oracle's fields: grad rho from the previous filtered rho
    sc = forms.scale(p)
    gr = [forms.r2r(v, k) for v, k in zip(_shift_grad(p, rhoq_prev), forms.IST)]
    S = [q.mu * q.grad_psi[ax] + q.energy * gr[ax] for ax in range(3)]
    for ax in range(3):
        assert rel_l2(out[smadft.FIELD_FORCE_X + ax], S[ax]) < 1e-10, ax
    Sh = [forms.r2r(S[ax], forms.ST[ax]) for ax in range(3)]
    zero = np.zeros(p.shape)
    _, *us = oracle.quasi_project(p, *Sh, zero)  # solenoidal part: no divv term
    qx, qy, qz = forms.wavenumbers(p)
    h = forms.spacing(p)
    qs = [np.pi * (np.arange(n) + 1) / (n * h) for n in p.shape]
    mq2 = [qs[0][:, None, None] ** 2 + qy[None, :, None] ** 2 + qz[None, None, :] ** 2,
           qx[:, None, None] ** 2 + qs[1][None, :, None] ** 2 + qz[None, None, :] ** 2,
           qx[:, None, None] ** 2 + qy[None, :, None] ** 2 + qs[2][None, None, :] ** 2]
    for ax in range(3):
        vsol = forms.r2r(us[ax], forms.IST[ax])
        assert rel_l2(out[smadft.FIELD_VSOL_X + ax], vsol) < 1e-9, ax
        # f = s P S_hat = nu |q|^2 vsol_hat in the component's own basis
        fproj = forms.r2r(p.nu * mq2[ax] * us[ax], forms.IST[ax])
        assert rel_l2(out[smadft.FIELD_FPROJ_X + ax], fproj) < 1e-9, ax
    divv = sc * forms.ict(q.divv)
    assert rel_l2(out[smadft.FIELD_DIVV], divv) < 1e-9
    assert grho[0].shape == p.shape


def _shift_grad(p, ah):
    """DST-shift gradient of a scaled spectral field (spectral outputs)."""
    import forms
    qx, qy, qz = forms.wavenumbers(p)
    h = forms.spacing(p)
    sx = np.pi * (np.arange(p.nx) + 1) / (p.nx * h)
    sy = np.pi * (np.arange(p.ny) + 1) / (p.ny * h)
    sz = np.pi * (np.arange(p.nz) + 1) / (p.nz * h)
    gx = np.zeros(p.shape); gy = np.zeros(p.shape); gz = np.zeros(p.shape)
    gx[:-1] = -sx[:-1, None, None] * ah[1:]
    gy[:, :-1] = -sy[None, :-1, None] * ah[:, 1:]
    gz[:, :, :-1] = -sz[None, None, :-1] * ah[:, :, 1:]
    return gx, gy, gz
