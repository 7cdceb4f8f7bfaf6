"""The Gaussian pre-filter (images/filters/GaussianFilter2D.java): the oracle
against the reference's golden filtered images (GaussianFilter2DTest.java,
tolerances 4e-5 for 3 px / 1e-4 for 5 px), and the HIP kernels against the
oracle, value for value."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN

CASES = [("g3_2p0", 2.0, 3, 4e-5, (0,)), ("g5_0p48", 0.48, 5, 1e-4, (0, 1))]


def _parts(img):
    return {"top": img[:16], "bottom": img[-16:], "left": img[:, :16], "right": img[:, -16:],
            "centre": img[224:288, 224:288]}


@pytest.mark.parametrize("key,sigma,K,tol,images", CASES)
def test_oracle_filter_matches_reference_golden(rankine_pair, key, sigma, K, tol, images):
    g = np.load(os.path.join(GOLDEN, "gauss_filter_golden.npz"))
    for k in images:
        img = rankine_pair[k].astype(np.float32)
        out = O.gaussian_filter2d(img, np.float32(sigma), K)
        for name, part in _parts(out).items():
            ref = g[f"{key}_{k}_{name}"]
            assert np.abs(part - ref).max() <= tol, (key, k, name)


def test_oracle_convolution_kat():
    """simpleConvolutionTest (GaussianFilter2DTest.java:47-60): [1.5, 2.5, 3.5]
    convolved with [1, 2] gives [1.5, 5.5, 8.5]; and the normalised kernel."""
    np.testing.assert_allclose(O.convolve1d([1.5, 2.5, 3.5], [1.0, 2.0]), [1.5, 5.5, 8.5],
                               atol=1e-6)
    k = O.gauss_kernel(0.8, 3)
    assert abs(k.sum() - 1) < 1e-6 and k[0] == k[2]


@pytest.mark.gpu
@pytest.mark.parametrize("K,sigma", [(3, 2.0), (5, 0.48), (7, 1.3), (4, 1.0), (1, 1.0)])
def test_gpu_filter_equals_oracle(rankine_pair, K, sigma):
    import torch

    import qlpiv

    ctx = qlpiv.Context(0)
    a, b = rankine_pair
    imgs = np.stack([a, b])
    got = ctx.gaussian_filter(torch.from_numpy(imgs).cuda(), sigma, K)
    gotf = ctx.gaussian_filter(torch.from_numpy(imgs.astype(np.float32)).cuda(), sigma, K)
    ctx.synchronize()
    got, gotf = got.cpu().numpy(), gotf.cpu().numpy()
    for k in range(2):
        ref = O.gaussian_filter2d(imgs[k].astype(np.float32), np.float32(sigma), K)
        assert np.array_equal(got[k], ref) and np.array_equal(gotf[k], ref), (K, sigma, k)
    zero = ctx.gaussian_filter(torch.from_numpy(imgs).cuda(), 0.0, 3)
    ctx.synchronize()
    assert np.array_equal(zero.cpu().numpy(), imgs.astype(np.float32))
    ctx.close()
