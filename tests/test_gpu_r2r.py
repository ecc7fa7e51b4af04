"""GPU parity of the hand-written r2r transforms against the oracle."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

SHAPES = [(16, 16, 16), (32, 16, 64), (64, 32, 16), (16, 128, 32), (384, 16, 16), (16, 16, 256),
          (1024, 16, 16), (16, 16, 1024), (16, 32, 512)]
COMBOS = [(0, 0, 0), (1, 1, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2), (3, 1, 1), (1, 3, 1), (1, 1, 3)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("kinds", COMBOS)
def test_r2r3d_fp64(smadft, oracle, shape, kinds):
    rng = np.random.default_rng(hash((shape, kinds)) % 2**32)
    x = rng.standard_normal(shape)
    ref = oracle.r2r3d(x, kinds)
    y = smadft.r2r_3d(x, kinds, smadft.FP64)
    assert rel_l2(y, ref) < 1e-14


@pytest.mark.parametrize("shape", [(32, 16, 64), (16, 384, 16), (128, 32, 32), (16, 1024, 32)])
@pytest.mark.parametrize("kinds", COMBOS)
def test_r2r3d_fp32(smadft, oracle, shape, kinds):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(shape)
    ref = oracle.r2r3d(x, kinds)
    y = smadft.r2r_3d(x, kinds, smadft.FP32_MIXED)
    assert rel_l2(y, ref) < 3e-7


# the compressible program's grid lengths (578 = 2 17^2, 484 = 4 11^2,
# 130 = 2 5 13: radices 17, 11, 13, 5), FP64 only
ODD_SHAPES = [(130, 16, 16), (16, 484, 16), (16, 16, 578), (16, 16, 130), (578, 16, 16), (16, 130, 484)]


@pytest.mark.parametrize("shape", ODD_SHAPES)
@pytest.mark.parametrize("kinds", COMBOS)
def test_r2r3d_fp64_odd_radices(smadft, oracle, shape, kinds):
    rng = np.random.default_rng(hash((shape, kinds)) % 2**32)
    x = rng.standard_normal(shape)
    ref = oracle.r2r3d(x, kinds)
    y = smadft.r2r_3d(x, kinds, smadft.FP64)
    assert rel_l2(y, ref) < 1e-14


def test_round_trip_scale(smadft):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((64, 64, 64))
    y = smadft.r2r_3d(smadft.r2r_3d(x, (0, 0, 0)), (1, 1, 1))
    assert rel_l2(y, 8 * 64 ** 3 * x) < 1e-14
    y = smadft.r2r_3d(smadft.r2r_3d(x, (2, 2, 2)), (3, 3, 3))
    assert rel_l2(y, 8 * 64 ** 3 * x) < 1e-14


@pytest.mark.parametrize("nbytes", [16, 4096 + 16, 3 * (1 << 20) + 48])
def test_stream_copy_is_exact(smadft, nbytes):
    """smad_stream_copy (the copy-ceiling kernel) copies every 16-byte group,
    including a partial last block, and nothing beyond."""
    import torch
    src = torch.arange(nbytes // 8 + 4, dtype=torch.float64, device="cuda:0")
    dst = torch.full_like(src, -1.0)
    smadft.stream_copy(dst.data_ptr(), src.data_ptr(), nbytes, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    k = nbytes // 8
    assert torch.equal(dst[:k], src[:k])
    assert torch.all(dst[k:] == -1.0)
