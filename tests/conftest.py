"""Shared test setup: markers, import paths and fixtures."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "quicklabpiv-ng_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) -- run with -m gpu")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


def load_json(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def rankine_pair():
    d = np.load(os.path.join(GOLDEN, "rankine_pair.npz"))
    return d["a"], d["b"]


@pytest.fixture(scope="session")
def golden_planes():
    return np.load(os.path.join(GOLDEN, "gpu_fft_golden.npz"))["planes"]


def gaussian_plane(dim, uX, sigmaX, uY, sigmaY, gain=1000.0):
    """GaussianTestHelper.createMatrixWithGaussian2DAssymmetric :42-53 with
    gaussian2DAssymmetric :27-30 (double evaluation, float result)."""
    f32 = np.float32
    out = np.zeros((dim, dim), np.float32)
    sx, sy = f32(sigmaX), f32(sigmaY)
    for i in range(dim):
        for j in range(dim):
            x = f32(j - dim // 2)
            y = f32(i - dim // 2)
            xu = float(f32(x - f32(uX)))
            yu = float(f32(y - f32(uY)))
            v = (float(f32(gain) * f32(1.0)) / (2.0 * np.pi * float(sx) * float(sy)) *
                 np.exp(-(xu ** 2 / (2.0 * float(sx) ** 2) + yu ** 2 / (2.0 * float(sy) ** 2))))
            out[i, j] = f32(v)
    return out
