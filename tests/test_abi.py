"""CPU-side checks of the C ABI library and the host logic (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def header_symbols():
    src = open(os.path.join(ROOT, "include", "qlpiv.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(qlp_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from qlpiv import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(lib, s), s
    # the Python binding declares a signature for each of them
    assert set(syms) == set(_lib.SIGNATURES)


def test_version_and_error_without_device():
    """Without a visible device the library reports QLP_ERR_DEVICE instead of
    falling back to anything."""
    import torch

    from qlpiv import _lib

    lib = _lib.load()
    assert b"gfx950" in lib.qlp_version()
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    rc = lib.qlp_ctx_create(0, 0, ctypes.byref(h))
    assert rc == _lib.QLP_ERR_DEVICE
    assert lib.qlp_last_error()


def test_null_context_is_an_argument_error():
    from qlpiv import _lib

    lib = _lib.load()
    assert lib.qlp_xcorr_batch(None, 1, 32, None, None, None, None, 0, None) == _lib.QLP_ERR_ARG
    assert b"context" in lib.qlp_last_error()


@pytest.mark.parametrize("H,W,N,count", [(1024, 1024, 32, 63 * 63), (2048, 2048, 64, 63 * 63),
                                         (2048, 2048, 32, 127 * 127), (2048, 2048, 16, 255 * 255),
                                         (512, 512, 16, 63 * 63), (1024, 2048, 16, 127 * 255),
                                         (2048, 2048, 128, 31 * 31)])
def test_ia_counts(H, W, N, count):
    """IA counts at 50 % overlap: (2W/N - 1)(2H/N - 1)
    (AdaptiveInterAreaStrategySuperPosition.java:79-80,161-162)."""
    from qlpiv.geometry import level_grid

    tops, lefts = level_grid(H, W, N, 0.5)
    assert tops.size == count
    assert tops.min() >= 0 and tops.max() + N <= H and lefts.max() + N <= W
    assert np.all(np.diff(lefts[0]) == N // 2)


def test_multipass_grid_levels():
    from qlpiv.geometry import multipass_grids

    g = multipass_grids(2048, 2048, 64, 16, 0.5)
    assert [t for t, _, _ in g] == [64, 32, 16]
    assert [a.size for _, a, _ in g] == [3969, 16129, 65025]


def test_synthetic_pair_is_seeded_and_dense():
    from qlpiv.synth import particle_pair

    a1, b1, _ = particle_pair(128, 128, seed=5)
    a2, b2, _ = particle_pair(128, 128, seed=5)
    a3, _, _ = particle_pair(128, 128, seed=6)
    assert np.array_equal(a1, a2) and np.array_equal(b1, b2) and not np.array_equal(a1, a3)
    assert a1.dtype == np.uint8 and 10 < a1.mean() < 60


def test_composite_last_level_interpolator_start_step():
    """CompositeLastLevelInterpolator (first-levels interpolator only): the
    reference's inverted isLastStep (:38) applies it on the last level alone,
    or on every level with applyMainInterpolatorAtLastStep, and never before
    the job's interpolatorStartStep."""
    from qlpiv import composite_last_level_start_step as f

    assert f(3, 0, False) == 2 and f(3, 0, True) == 0
    assert f(3, 1, True) == 1 and f(3, 2, False) == 2 and f(1, 0, False) == 0
    # the levels the composite interpolates, by the reference's own rule
    for n in (1, 2, 3, 4):
        for start in range(n + 1):
            for apply in (False, True):
                levels = [s for s in range(n)
                          if s >= start and ((s == n - 1) or apply)]
                assert levels == [s for s in range(n) if s >= f(n, start, apply)]
