"""GPU parity of the compressible step (compressible/compressible.cpp:
1114-2700) against the oracle: across the relaxation switch, with the z
lift of the relaxation density, from a loaded density (load == 1), in FP32
mixed precision, and through a full-state restart."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def cparams(smadft, shape, prec=64, **kw):
    p = smadft.Params.default(smadft.MODEL_COMPRESSIBLE, shape[0])
    p.nx, p.ny, p.nz = shape
    p.precision = prec
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def gpu_fields(smadft, s):
    F = smadft
    out = {nm: s.fetch(f) for nm, f in (("psi", F.FIELD_PSI), ("rho", F.FIELD_RHO), ("frho", F.FIELD_F_RHO),
                                        ("amp", F.FIELD_AMP), ("divv", F.FIELD_DIVV))}
    out["vel"] = np.stack([s.fetch(f) for f in (F.FIELD_VEL_X, F.FIELD_VEL_Y, F.FIELD_VEL_Z)])
    out["virr"] = np.stack([s.fetch(f) for f in (F.FIELD_VIRR_X, F.FIELD_VIRR_Y, F.FIELD_VIRR_Z)])
    out["grad_rho"] = np.stack([s.fetch(f) for f in (F.FIELD_GRAD_RHO_X, F.FIELD_GRAD_RHO_Y, F.FIELD_GRAD_RHO_Z)])
    out["grad_psi"] = np.stack([s.fetch(f) for f in (F.FIELD_GRAD_X, F.FIELD_GRAD_Y, F.FIELD_GRAD_Z)])
    out["l1"] = s.reduce(F.RED_L1)
    out["mass"] = s.reduce(F.RED_MASS)[0]
    out["energy"] = s.reduce(F.RED_ENERGY)[0]
    return out


def check_fp64(g, o, tol=1e-9):
    assert rel_l2(g["psi"], o.psi) < tol * 0.1
    for nm in ("rho", "frho", "amp", "divv", "vel", "virr", "grad_rho", "grad_psi"):
        want = getattr(o, nm)
        assert np.abs(want).max() > 0, nm
        assert rel_l2(g[nm], want) < tol, nm
    assert g["l1"][0] == pytest.approx(o.sums[0], rel=1e-8)
    assert g["l1"][1] == pytest.approx(o.sums[1], rel=1e-10)
    assert g["mass"] == pytest.approx(o.sums[2], rel=1e-11)
    assert g["energy"] == pytest.approx(o.sums[3], rel=1e-9)


@pytest.mark.parametrize("shape,nsteps,switch,up", [((32, 32, 32), 10, 4, 3), ((16, 32, 64), 8, 3, 16),
                                                    ((384, 16, 32), 5, 2, 2), ((32, 16, 384), 5, 2, 16),
                                                    ((130, 16, 32), 5, 2, 3), ((16, 484, 16), 4, 2, 2),
                                                    ((32, 16, 578), 4, 2, 16), ((578, 16, 130), 3, 1, 16)])
def test_comp_fp64_matches_oracle(smadft, oracle, shape, nsteps, switch, up):
    p = cparams(smadft, shape, div_switch=switch, rho_up=up)
    psi0 = smadft.ic_curved(p, 1.4142, 1e-3, seed=3)
    print(f"comp fp64 {shape} steps={nsteps} switch={switch} up={up}", flush=True)
    o = oracle.CompOracle(p, psi0)
    o.step(nsteps)
    with smadft.Stepper(p, psi0) as s:
        s.step(nsteps)
        g = gpu_fields(smadft, s)
    check_fp64(g, o)


def test_comp_fc_torus_matches_oracle(smadft, oracle):
    """The program's own initial condition (the FC torus) with its rho_up = 16."""
    p = cparams(smadft, (64, 64, 64), div_switch=3)
    psi0 = smadft.ic_fc_torus(p, None, 40.0, 40.0)
    print("comp fc torus 64^3", flush=True)
    o = oracle.CompOracle(p, psi0)
    o.step(6)
    with smadft.Stepper(p, psi0) as s:
        s.step(6)
        g = gpu_fields(smadft, s)
    check_fp64(g, o)


def test_comp_load_start_matches_oracle(smadft, oracle):
    """load == 1: no relaxation; the loaded density is the state from the
    first step (compressible.cpp:1542, :874-924)."""
    p = cparams(smadft, (32, 32, 32), load=1)
    psi0 = smadft.ic_curved(p, 1.4142, 1e-3, seed=5)
    x = (np.indices(p.shape)[2] + 0.5) / p.nz
    rho_in = 0.55 + 0.4 * np.tanh(4 * (x - 0.5))
    print("comp load start", flush=True)
    o = oracle.CompOracle(p, psi0, rho_in)
    o.step(6)
    with smadft.Stepper(p, psi0) as s:
        s.set_field(smadft.FIELD_RHO, rho_in)
        s.step(6)
        g = gpu_fields(smadft, s)
    check_fp64(g, o)


def test_comp_fp32_mixed_close(smadft, oracle):
    p = cparams(smadft, (32, 32, 32), prec=32, div_switch=4, rho_up=3)
    psi0 = smadft.ic_curved(p, 1.4142, 1e-3, seed=3)
    print("comp fp32", flush=True)
    o = oracle.CompOracle(p, psi0)
    o.step(10)
    with smadft.Stepper(p, psi0) as s:
        s.step(10)
        psi = s.fetch()
        rho = s.fetch(smadft.FIELD_RHO)
        uz = s.fetch(smadft.FIELD_VEL_Z)
    assert rel_l2(psi, o.psi) < 1e-5
    assert rel_l2(rho, o.rho) < 1e-5
    assert rel_l2(uz, o.vel[2]) < 1e-3


@pytest.mark.parametrize("prec", [64, 32])
def test_comp_full_state_restart_is_bit_identical(smadft, prec):
    p = cparams(smadft, (16, 32, 32), prec=prec, div_switch=4, rho_up=3)
    psi0 = smadft.ic_curved(p, 1.4142, 1e-3, seed=7)
    K, M = 3, 4  # the switch falls after the restart
    print(f"comp restart prec={prec}", flush=True)
    with smadft.Stepper(p, psi0) as a:
        a.step(K + M)
        want = a.fetch()
        want_rho = a.fetch(smadft.FIELD_RHO)
        want_l1 = a.reduce()[:2]
    with smadft.Stepper(p, psi0) as b:
        b.step(K)
        state = b.save_state()
    with smadft.Stepper(p, np.zeros(p.shape)) as c:
        c.load_state(state)
        c.step(M)
        got = c.fetch()
        got_rho = c.fetch(smadft.FIELD_RHO)
        got_l1 = c.reduce()[:2]
    assert np.array_equal(got, want)
    assert np.array_equal(got_rho, want_rho)
    assert np.array_equal(got_l1, want_l1)


def test_comp_reference_grid_runs(smadft):
    """The program's own grid, 578 x 484 x 130, FP64, from its FC torus
    (x scaled by 5.2/6.2, centre 242, radius 121): a few steps across the
    switch stay finite, and the density stays within [rho0, 1] up to the
    continuity's small excursions."""
    p = cparams(smadft, (578, 484, 130), div_switch=2)
    amp = np.sqrt((3 * p.beta + np.sqrt(9 * p.beta ** 2 + 40 * p.eps * p.gamma)) / 5)
    psi0 = smadft.ic_fc_torus(p, amp, 124.0, 124.0, 5.2 / 6.2)
    print("comp reference grid 578x484x130", flush=True)
    with smadft.Stepper(p, psi0) as s:
        s.step(4)
        l1 = s.reduce(smadft.RED_L1)
        rho = s.fetch(smadft.FIELD_RHO)
        psi = s.fetch()
    assert np.isfinite(l1).all() and l1[0] > 0
    assert np.isfinite(psi).all() and np.abs(psi).max() < 2
    assert rho.min() > 0.0 and rho.max() < 1.1
