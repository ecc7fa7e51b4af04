"""Full-size plain-model parity at 128^3 over 1000 steps (the reference's
128^3 slab relaxation, nw = 8, eps = -0.7, dt = 0.1), GPU vs the FP64 oracle.
The oracle needs about a minute of host time, so this runs only with
SMAD_LONG=1 in the environment."""
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.environ.get("SMAD_LONG"), reason="set SMAD_LONG=1 (long oracle run)")]


@pytest.mark.parametrize("amp", [0.6, 1.328])
def test_plain_128_1000_steps(smadft, oracle, amp):
    p = smadft.Params.default(smadft.MODEL_PLAIN, 128)
    psi0 = smadft.ic_slab(p, amp, 1e-3, seed=0)
    o = oracle.PlainOracle(p, psi0)
    o.step(1000)
    with smadft.Stepper(p, psi0) as s:
        s.step(1000)
        psi = s.fetch()
        l1 = s.l1()
    err = rel_l2(psi, o.psi)
    print(f"amp {amp}: |psi| {np.linalg.norm(o.psi):.3e}, rel L2 GPU vs oracle {err:.3e}, L1 {l1:.3e}", flush=True)
    assert err < 1e-10
