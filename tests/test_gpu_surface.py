"""GPU surface diagnostics (smad_surface, smad_surface_envelope) against the
oracle's restatement of no-adv/cosNoAdvTrackMult.cpp:787-796, 853-1022 and of
the |psi|-envelope scan of no-adv/cosNoAdvTrackBump.cpp:929-1031 on the same
psi."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_state(smadft, n, prec, steps):
    p = smadft.Params.default(smadft.MODEL_PLAIN, n)
    p.precision = prec
    psi0 = smadft.ic_curved(p, 1.328, 1e-3, seed=4)
    s = smadft.Stepper(p, psi0)
    s.step(steps)
    return p, s


@pytest.mark.parametrize("envelope", [False, True])
@pytest.mark.parametrize("n,steps", [(32, 30), (64, 60)])
def test_surface_fp64_matches_oracle(smadft, oracle, n, steps, envelope):
    p, s = run_state(smadft, n, 64, steps)
    with s:
        psi = s.fetch()
        got = s.surface(envelope=envelope)
    ref = oracle.surface(p, psi, envelope=envelope)
    assert (ref["surf"] >= 0).sum() > n * n // 4  # the curved interface is found
    assert np.array_equal(got["surf"], ref["surf"])
    for key in ("vel", "curv_h", "curv_k"):
        scale = np.abs(ref[key]).max()
        assert np.max(np.abs(got[key] - ref[key])) < 1e-8 * scale, key


@pytest.mark.parametrize("envelope", [False, True])
def test_surface_fp32_mixed_close(smadft, oracle, envelope):
    p, s = run_state(smadft, 64, 32, 60)
    with s:
        psi = s.fetch()
        got = s.surface(envelope=envelope)
    ref = oracle.surface(p, psi, envelope=envelope)
    same = got["surf"] == ref["surf"]
    assert same.mean() > 0.99
    for key in ("vel", "curv_h", "curv_k"):
        scale = np.abs(ref[key][same]).max()
        assert np.max(np.abs(got[key][same] - ref[key][same])) < 1e-3 * scale, key


def adv_state(smadft, n, prec, steps):
    p = smadft.Params.default(smadft.MODEL_ADV, n)
    p.nw, p.dt, p.nu = 8.0, 0.02, 1.0
    p.precision = prec
    s = smadft.Stepper(p, smadft.ic_curved(p, 1.328, 1e-3, seed=4))
    s.step(steps)
    return p, s


@pytest.mark.parametrize("n,steps", [(32, 20), (64, 10)])
def test_surface_adv_matches_oracle(smadft, oracle, n, steps):
    """The advective save block's tracker (adv/cosAdvTrackMult.cpp:1311-1347,
    1412-1500) on the GPU against the oracle for the same psi; afterwards
    the GRAD and VEL fields are those of the current psi."""
    p, s = adv_state(smadft, n, 64, steps)
    with s:
        psi = s.fetch()
        got = s.surface()
        vel = [s.fetch(f) for f in (smadft.FIELD_VEL_X, smadft.FIELD_VEL_Y, smadft.FIELD_VEL_Z)]
        psi_again = s.fetch()
        s.step(2)
        psi2 = s.fetch()
    _, plain_run = adv_state(smadft, n, 64, steps + 2)  # the same run without the surface call
    with plain_run:
        want2 = plain_run.fetch()
    ref = oracle.adv_surface(p, psi)
    assert (ref["surf"] >= 0).sum() > n * n // 4
    assert np.array_equal(got["surf"], ref["surf"])
    for key in ("vel", "curv_h", "curv_k"):
        assert np.max(np.abs(got[key] - ref[key])) < 1e-8 * np.abs(ref[key]).max(), key
    o = oracle.AdvOracle(p, psi)
    o.step(1)
    for ax in range(3):
        assert np.max(np.abs(vel[ax] - o.vel[ax])) < 1e-9 * np.abs(o.vel).max()
    assert np.max(np.abs(psi - psi_again)) < 1e-13 * np.abs(psi).max()
    assert np.array_equal(psi2, want2)  # stepping is unaffected


def test_surface_adv_fp32_mixed_close(smadft, oracle):
    p, s = adv_state(smadft, 64, 32, 10)
    with s:
        psi = s.fetch()
        got = s.surface()
    ref = oracle.adv_surface(p, psi)
    same = got["surf"] == ref["surf"]
    assert same.mean() > 0.99
    for key in ("vel", "curv_h", "curv_k"):
        scale = np.abs(ref[key][same]).max()
        assert np.max(np.abs(got[key][same] - ref[key][same])) < 1e-3 * scale, key


def test_surface_not_for_quasi_and_conserved_adv(smadft):
    for model in (smadft.MODEL_QUASI, smadft.MODEL_ADV_CONSERVED):
        p = smadft.Params.default(model, 16)
        with smadft.Stepper(p, smadft.ic_slab(p, 1.328, 0.0, seed=0)) as s:
            for envelope in (False, True):
                with pytest.raises(smadft.SmadError) as e:
                    s.surface(envelope=envelope)
                assert e.value.code == smadft.EINVAL
