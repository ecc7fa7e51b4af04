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


def test_surface_is_plain_only(smadft):
    p = smadft.Params.default(smadft.MODEL_ADV, 16)
    with smadft.Stepper(p, smadft.ic_slab(p, 1.328, 0.0, seed=0)) as s:
        for envelope in (False, True):
            with pytest.raises(smadft.SmadError) as e:
                s.surface(envelope=envelope)
            assert e.value.code == smadft.EINVAL
