"""smad_r2r_3d_noalloc (allocates nothing; device buffers in and out) against
smad_r2r_3d (host buffers, context-owned device memory) on the same inputs.
Both run the same kernels with the same twiddle values, so the results are
bit-identical; two calls in a row, the second on a smaller grid, check that
the shared scratch array carries nothing from one call into the next."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

KINDS = [(0, 0, 0), (1, 1, 1), (2, 0, 3), (3, 2, 1)]


def run_noalloc(smadft, x, kinds, precision, inplace=False):
    dev = torch.device("cuda:0")
    xin = torch.from_numpy(x).to(dev)
    out = xin if inplace else torch.empty_like(xin)
    stream = torch.cuda.current_stream(dev)
    smadft.r2r_3d_noalloc(xin.data_ptr(), out.data_ptr(), x.shape, kinds, precision, stream.cuda_stream)
    stream.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("precision", [64, 32])
def test_noalloc_matches_r2r_3d_two_calls(smadft, oracle, precision):
    rng = np.random.default_rng(precision)
    for shape in ((64, 32, 128), (16, 32, 16)):  # second call smaller than the first
        for kinds in KINDS:
            print(f"noalloc precision={precision} shape={shape} kinds={kinds}", flush=True)
            x = rng.standard_normal(shape)
            got = run_noalloc(smadft, x, kinds, precision)
            ref = smadft.r2r_3d(x, kinds, precision)
            assert np.array_equal(got, ref), (shape, kinds)
            orc = oracle.r2r3d(x, kinds)
            err = np.linalg.norm(got - orc) / np.linalg.norm(orc)
            assert err < (1e-14 if precision == 64 else 3e-7), err


def test_noalloc_in_place_and_largest_fp32_grid(smadft):
    rng = np.random.default_rng(5)
    x = rng.standard_normal((32, 64, 32))
    got = run_noalloc(smadft, x, (1, 0, 2), 64, inplace=True)
    assert np.array_equal(got, smadft.r2r_3d(x, (1, 0, 2), 64))
    # the largest FP32 grid the staging area holds (2^24 points), then a small one
    x = rng.standard_normal((256, 256, 256))
    print("noalloc FP32 256^3", flush=True)
    got = run_noalloc(smadft, x, (0, 0, 0), 32)
    assert np.array_equal(got, smadft.r2r_3d(x, (0, 0, 0), 32))
    x = rng.standard_normal((16, 16, 16))
    got = run_noalloc(smadft, x, (3, 3, 3), 32)
    assert np.array_equal(got, smadft.r2r_3d(x, (3, 3, 3), 32))
