"""Full-size plain-model parity at 128^3 over 1000 steps: configuration 0
(the reference's 128^3 slab relaxation, nw = 8, eps = -0.7, dt = 0.1,
amplitude 0.6) and the reference's amplitude 1.328, where the smectic
persists; GPU vs the FP64 oracle (about a minute of host time each, printed
in stages)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("amp", [0.6, 1.328])
def test_plain_128_1000_steps(smadft, oracle, amp):
    p = smadft.Params.config(0)
    psi0 = smadft.ic_slab(p, amp, 1e-3, seed=0)
    if amp == 0.6:
        assert np.array_equal(psi0, smadft.config_ic(0, p))
    o = oracle.PlainOracle(p, psi0)
    for done in range(250, 1001, 250):
        o.step(250)
        print(f"oracle amp {amp}: {done} steps", flush=True)
    with smadft.Stepper(p, psi0) as s:
        s.step(1000)
        psi = s.fetch()
        l1 = s.l1()
    err = rel_l2(psi, o.psi)
    print(f"amp {amp}: |psi| {np.linalg.norm(o.psi):.3e}, rel L2 GPU vs oracle {err:.3e}, L1 {l1:.3e}", flush=True)
    assert err < 1e-10
