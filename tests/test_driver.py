"""smadft_run, the host driver with the reference programs' loop control
(no-adv/cosNoAdvTrackMult.cpp:692-850): argument handling on the CPU, and on
the GPU the L1.dat / psi.dat / psi.bin it writes against the library driven
directly through the ABI."""
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sma-dft_amd", "smadft_run")


def run(args, **kw):
    if not os.path.exists(EXE):
        pytest.fail("smadft_run is not built (make -C sma-dft_amd)")
    return subprocess.run([EXE] + args, capture_output=True, text=True, timeout=300, **kw)


@pytest.mark.parametrize("args", [[], ["nosuch"], ["adv", "1", "0.7"], ["quasi", "1"],
                                  ["plain", "--bogus"], ["plain", "--n"], ["plain", "--steps", "x"],
                                  ["plain", "--steps", "0"], ["plain", "--tracker", "top"]])
def test_usage_errors(args):
    r = run(args)
    assert r.returncode == 2
    assert "usage" in r.stderr


def read_bin(path):
    with open(path, "rb") as f:
        assert f.read(8) == b"SMADPSI1"
        nx, ny, nz = struct.unpack("<3i", f.read(12))
        (steps,) = struct.unpack("<q", f.read(8))
        data = np.frombuffer(f.read(), dtype=np.float64)
    assert data.size == nx * ny * nz
    return (nx, ny, nz), steps, data.reshape(nx, ny, nz)


@pytest.mark.gpu
def test_plain_run_matches_library(smadft, tmp_path):
    out = tmp_path / "run"
    r = run(["plain", "--n", "32", "--steps", "300", "--save-every", "2", "--ic", "slab",
             "--amp", "1.328", "--noise", "1e-3", "--seed", "3", "--out", str(out)])
    assert r.returncode == 0, r.stderr
    l1 = np.loadtxt(out / "L1.dat")
    assert l1.shape == (3,)
    dims, steps, psi = read_bin(out / "psi.bin")
    assert dims == (32, 32, 32) and steps == 300
    txt = np.loadtxt(out / "psi.dat")
    assert np.allclose(txt, psi.ravel(), rtol=1e-5, atol=1e-6)  # 6 significant digits
    p = smadft.Params.default(smadft.MODEL_PLAIN, 32)
    psi0 = smadft.ic_slab(p, 1.328, 1e-3, seed=3)
    with smadft.Stepper(p, psi0) as s:
        ref = []
        for _ in range(3):
            s.step(100)
            ref.append(s.l1())
        ref_psi = s.fetch()
    assert np.allclose(l1, ref, rtol=1e-5)
    assert rel_l2(psi, ref_psi) < 1e-13


@pytest.mark.gpu
def test_plain_restart_and_l1_stop(smadft, tmp_path):
    out = tmp_path / "run"
    r = run(["plain", "--n", "16", "--steps", "200", "--ic", "slab", "--amp", "1.328",
             "--out", str(out)])
    assert r.returncode == 0, r.stderr
    _, steps, psi200 = read_bin(out / "psi.bin")
    assert steps == 200
    # restart re-runs the pre-loop from psi, as the reference's load path
    r = run(["plain", "--n", "16", "--steps", "100", "--restart", str(out / "psi.bin"),
             "--out", str(tmp_path / "cont")])
    assert r.returncode == 0, r.stderr
    _, steps, psi300 = read_bin(tmp_path / "cont" / "psi.bin")
    assert steps == 300
    p = smadft.Params.default(smadft.MODEL_PLAIN, 16)
    with smadft.Stepper(p, psi200.copy()) as s:
        s.step(100)
        assert rel_l2(psi300, s.fetch()) < 1e-13
    # L1-stop: runs until L1 <= limit, last line of L1.dat below it
    r = run(["plain", "--n", "16", "--l1", "1e-3", "--max-steps", "20000", "--ic", "slab",
             "--amp", "1.328", "--out", str(tmp_path / "stop")])
    assert r.returncode == 0, r.stderr
    l1 = np.atleast_1d(np.loadtxt(tmp_path / "stop" / "L1.dat"))
    assert l1[-1] <= 1e-3 and (l1[:-1] > 1e-3).all()


@pytest.mark.gpu
@pytest.mark.parametrize("tracker", ["crossing", "envelope"])
def test_plain_save_block_surface_files(smadft, tmp_path, tracker):
    """Save blocks write psiMid.dat and the surface files of the state one
    step before the block's end, as the reference's save block does."""
    out = tmp_path / "run"
    n = 32
    r = run(["plain", "--n", str(n), "--steps", "40", "--l1-every", "10", "--save-every", "2",
             "--ic", "curved", "--amp", "1.328", "--seed", "4", "--tracker", tracker, "--out", str(out)])
    assert r.returncode == 0, r.stderr
    files = {k: np.loadtxt(out / f"{k}.dat").reshape(2, -1)
             for k in ("surfPsi", "velSurf", "curvH", "curvK", "psiMid")}
    p = smadft.Params.default(smadft.MODEL_PLAIN, n)
    psi0 = smadft.ic_curved(p, 1.328, 1e-3, seed=4)
    with smadft.Stepper(p, psi0) as s:
        for blk, nsteps in enumerate((19, 20)):
            s.step(nsteps)
            sf = s.surface(envelope=tracker == "envelope")
            psi = s.fetch()
            assert np.array_equal(files["surfPsi"][blk], sf["surf"].ravel())
            assert (sf["surf"] >= 0).any()
            for key, name in (("vel", "velSurf"), ("curv_h", "curvH"), ("curv_k", "curvK")):
                assert np.allclose(files[name][blk], sf[key].ravel(), rtol=1e-5, atol=1e-6 * np.abs(sf[key]).max())
            assert np.allclose(files["psiMid"][blk], psi[:, n // 2, :].ravel(), rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("argv,model,l1_every,n", [
    (["adv", "1", "700", "8", "0"], 1, 100, 16),
    (["quasi", "1", "700", "8"], 2, 50, 16),
    (["conserved", "70", "8", "16"], 3, 100, 16),
    (["advcons", "1", "70", "8", "16", "0"], 4, 100, 16),
])
def test_model_modes_match_library(smadft, tmp_path, argv, model, l1_every, n):
    """The positional argv of each reference main and its default cadence:
    two L1 blocks, L1.dat against the library driven directly."""
    out = tmp_path / "run"
    extra = [] if "--n" in argv or model in (3, 4) else ["--n", str(n)]
    r = run(argv + extra + ["--steps", str(2 * l1_every), "--ic", "slab", "--amp", "1.328",
                            "--seed", "2", "--out", str(out)])
    assert r.returncode == 0, r.stderr
    l1 = np.loadtxt(out / "L1.dat")
    assert l1.shape == (2,)
    p = smadft.Params.default(model, n)
    if model in (1, 2, 4):
        p.nu = 1.0
        p.lambda_ = -2.0 / 3.0
    p.eps = (-0.01 if model in (3, 4) else -0.001) * float(argv[2 if model in (1, 2, 4) else 1])
    psi0 = smadft.ic_slab(p, 1.328, 1e-3, seed=2)
    with smadft.Stepper(p, psi0) as s:
        ref = []
        for _ in range(2):
            s.step(l1_every)
            ref.append(s.l1())
    assert np.allclose(l1, ref, rtol=1e-5)
    if model == 2:
        assert np.loadtxt(out / "mass.dat").shape == (2,)
