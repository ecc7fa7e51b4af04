"""CPU tests of the C ABI library: it loads, exports every function the
header declares, validates arguments, and its host-only IC builders are
deterministic and match their defining formulas."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "smadft", "smadft.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(smad_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_what_python_binds(smadft):
    assert sorted(smadft.EXPORTS) == declared_functions()


def test_library_exports_every_declared_symbol(smadft):
    L = smadft.lib()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_params_struct_layout(smadft, tmp_path):
    # model..nz ints (16 B), 13 doubles (104 B), 3 ints (12 B), zeta at 136, rho_up at 144,
    # padded to 152 B; the offsets the C compiler gives the header's struct
    assert ctypes.sizeof(smadft.CParams) == 152
    names = [n for n, _ in smadft.CParams._fields_]
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"smadft/smadft.h\"\n"
                   "int main(void){printf(\"%zu\\n\", sizeof(smad_params));\n" +
                   "".join(f'printf("%zu\\n", offsetof(smad_params, {"lambda" if n == "lambda_" else n}));\n'
                           for n in names) + "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    assert got[0] == ctypes.sizeof(smadft.CParams)
    assert got[1:] == [getattr(smadft.CParams, n).offset for n in names]


def test_init_rejects_bad_sizes(smadft):
    p = smadft.Params.default(0, 32)
    p.nx = 20
    with pytest.raises(smadft.SmadError) as e:
        smadft.Stepper(p, np.zeros(p.shape))
    assert e.value.code == smadft.EINVAL
    p = smadft.Params.default(0, 32)
    p.precision = 16
    with pytest.raises(smadft.SmadError) as e:
        smadft.Stepper(p, np.zeros(p.shape))
    assert e.value.code == smadft.EINVAL
    with pytest.raises(smadft.SmadError):
        smadft.r2r_3d(np.zeros((20, 16, 16)), (0, 0, 0))


def test_noise_is_deterministic_and_bounded(smadft):
    p = smadft.Params.default(0, 16)
    a = smadft.add_noise(p, np.zeros(p.shape), 1e-3, seed=7)
    b = smadft.add_noise(p, np.zeros(p.shape), 1e-3, seed=7)
    c = smadft.add_noise(p, np.zeros(p.shape), 1e-3, seed=8)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.abs(a).max() <= 1e-3 and abs(a.mean()) < 1e-4
    assert a.std() == pytest.approx(1e-3 / np.sqrt(3), rel=0.05)


def test_slab_ic_formula(smadft):
    p = smadft.Params.default(0, 32)
    psi = smadft.ic_slab(p, amp=0.6, noise=0.0)
    dx = 2 * np.pi / (p.q0 * p.nw)
    lx = p.nx * dx
    x = np.arange(p.nx) * dx
    k = np.arange(p.nz)
    inside = np.abs(x - lx / 2) < lx / 4
    ref = np.where(inside[:, None, None], 0.6 * np.cos(p.q0 * k * dx)[None, None, :], 0.0)
    ref = np.broadcast_to(ref, p.shape)
    assert np.max(np.abs(psi - ref)) < 1e-15


def test_focal_conic_ic_formula(smadft):
    # no-adv/cosNoAdvTrackSing.cpp:389-449, checked point by point
    p = smadft.Params.default(0, 32)
    a_e = b_e = 10.0
    psi = smadft.ic_focal_conic(p, amp=1.328, a_e=a_e, b_e=b_e)
    dz = 2 * np.pi / (p.q0 * p.nw)
    mid = p.nx / 2
    for (i, j, k) in [(16, 16, 3), (0, 0, 2), (5, 16, 9), (16, 5, 12), (30, 30, 1), (10, 20, 4)]:
        ds = np.hypot(i - mid, j - mid)
        v = 0.0
        if k < b_e + 1:
            if ds < mid:
                if np.sqrt(((ds - mid) / a_e) ** 2 + (k / b_e) ** 2) <= 1:
                    v = 1.328 * np.cos(p.q0 * dz * np.sqrt(((b_e / a_e) * (ds - mid)) ** 2 + k * k))
            elif abs(k) < b_e:
                v = 1.328 * np.cos(p.q0 * k * dz)
        assert psi[i, j, k] == pytest.approx(v, abs=1e-14)


def test_noalloc_rejects_bad_arguments_before_any_device_call(smadft):
    """smad_r2r_3d_noalloc validates on the host: no device is touched for
    a rejected call (this runs without a GPU)."""
    fake = 1 << 20  # aligned, never dereferenced
    with pytest.raises(smadft.SmadError) as e:  # FP32 staging holds at most 2^24 points
        smadft.r2r_3d_noalloc(fake, fake, (512, 256, 256), (0, 0, 0), smadft.FP32_MIXED)
    assert e.value.code == smadft.EINVAL and "scratch area" in str(e.value)
    with pytest.raises(smadft.SmadError) as e:
        smadft.r2r_3d_noalloc(fake, fake, (20, 16, 16), (0, 0, 0))
    assert e.value.code == smadft.EINVAL
    with pytest.raises(smadft.SmadError) as e:
        smadft.r2r_3d_noalloc(fake + 8, fake, (16, 16, 16), (0, 0, 0))
    assert e.value.code == smadft.EINVAL and "aligned" in str(e.value)
    with pytest.raises(smadft.SmadError) as e:
        smadft.r2r_3d_noalloc(fake, fake, (16, 16, 16), (0, 4, 0))
    assert e.value.code == smadft.EINVAL


def test_stream_copy_rejects_unaligned_arguments(smadft):
    L = smadft.lib()
    for dst, src, nb in ((0, 0, 16), (16, 32, 15), (17, 32, 16), (16, 33, 16)):
        assert L.smad_stream_copy(dst or None, src or None, nb, None) == smadft.EINVAL


def test_odd_radix_lengths_are_fp64_only(smadft):
    """578, 484 and 130 points (the compressible program's grid) run in FP64;
    FP32 is rejected before any device work."""
    for shape in ((578, 16, 16), (16, 484, 16), (16, 16, 130)):
        with pytest.raises(smadft.SmadError) as e:
            smadft.r2r_3d(np.zeros(shape), (0, 0, 0), smadft.FP32_MIXED)
        assert e.value.code == smadft.EINVAL
        p = smadft.Params.default(smadft.MODEL_COMPRESSIBLE, 16)
        p.nx, p.ny, p.nz = shape
        p.precision = 32
        with pytest.raises(smadft.SmadError) as e:
            smadft.Stepper(p, np.zeros(p.shape))
        assert e.value.code == smadft.EINVAL
