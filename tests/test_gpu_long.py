"""Parity over 1000 steps: the plain model at full size (128^3,
configuration 0: the reference's slab relaxation, nw = 8, eps = -0.7,
dt = 0.1, amplitude 0.6, and the reference's amplitude 1.328, where the
smectic persists), and the advective, quasi and compressible models at 32^3
across their switches; GPU vs the FP64 oracle (up to a minute of host time
each, printed in stages)."""
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


def _run_long(smadft, oracle_obj, p, psi0, label, fields):
    for done in range(250, 1001, 250):
        oracle_obj.step(250)
        print(f"oracle {label}: {done} steps", flush=True)
    with smadft.Stepper(p, psi0) as s:
        s.step(1000)
        got = {nm: s.fetch(f) for nm, f in fields.items()}
    return got


def test_adv_1000_steps_config2_ic(smadft, oracle):
    """The advective step over 1000 steps on config 2's two-bump IC (32^3,
    its dt = 0.01): psi and u within 1e-10 / 1e-9 of the oracle."""
    p = smadft.Params.config(2, 32)
    psi0 = smadft.config_ic(2, p)
    o = oracle.AdvOracle(p, psi0)
    got = _run_long(smadft, o, p, psi0, "advective", {"psi": smadft.FIELD_PSI, "uz": smadft.FIELD_VEL_Z})
    e_psi, e_u = rel_l2(got["psi"], o.psi), rel_l2(got["uz"], o.vel[2])
    print(f"advective 1000 steps: psi {e_psi:.3e}, uz {e_u:.3e}", flush=True)
    assert e_psi < 1e-10 and e_u < 1e-9


def test_quasi_1000_steps_across_the_switch(smadft, oracle):
    """The quasi step over 1000 steps on config 3's tilted slab (32^3), its
    divv switch at step 200 included."""
    p = smadft.Params.config(3, 32)
    psi0 = smadft.config_ic(3, p)
    o = oracle.QuasiOracle(p, psi0)
    got = _run_long(smadft, o, p, psi0, "quasi", {"psi": smadft.FIELD_PSI, "rho": smadft.FIELD_RHO,
                                                 "ux": smadft.FIELD_VEL_X})
    e = {k: rel_l2(got[k], getattr(o, "vel")[0] if k == "ux" else getattr(o, k)) for k in got}
    print(f"quasi 1000 steps: {e}", flush=True)
    assert e["psi"] < 1e-10 and e["rho"] < 1e-10 and e["ux"] < 1e-8


def test_comp_1000_steps_fc_torus(smadft, oracle):
    """The compressible step over 1000 steps from the FC torus (32^3,
    rho_up = 3), its relaxation switch at step 200 included."""
    p = smadft.Params.default(smadft.MODEL_COMPRESSIBLE, 32)
    p.rho_up = 3
    psi0 = smadft.ic_fc_torus(p, None, 30.0, 30.0)
    o = oracle.CompOracle(p, psi0)
    got = _run_long(smadft, o, p, psi0, "compressible", {"psi": smadft.FIELD_PSI, "rho": smadft.FIELD_RHO,
                                                        "uz": smadft.FIELD_VEL_Z})
    e = {"psi": rel_l2(got["psi"], o.psi), "rho": rel_l2(got["rho"], o.rho), "uz": rel_l2(got["uz"], o.vel[2])}
    print(f"compressible 1000 steps: {e}", flush=True)
    assert e["psi"] < 1e-10 and e["rho"] < 1e-10 and e["uz"] < 1e-8
