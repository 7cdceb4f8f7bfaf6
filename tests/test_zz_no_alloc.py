"""The allocation-free correlation entry point (qlp_xcorr_batch_noalloc_dev)
against qlp_xcorr_batch_dev, through the C ABI: planes value-identical for
every IA size and both arithmetic modes, on two calls in a row (the second
with fewer IAs, reusing the static scratch), including 128 px batches larger
than the 64 scratch slots."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("N,counts", [(16, (700, 90)), (32, (400, 77)), (64, (150, 33)),
                                      (128, (150, 20))])
@pytest.mark.parametrize("mode", [0, 1])
def test_noalloc_equals_allocating_entry_point(N, counts, mode):
    import torch

    import qlpiv
    from qlpiv import _lib

    lib = _lib.load()
    ctx = qlpiv.Context(0)
    rng = np.random.default_rng(N * 10 + mode)
    for n in counts:
        tA = torch.from_numpy(rng.integers(0, 256, (n, N, N), dtype=np.uint8)).cuda()
        tB = torch.from_numpy(rng.integers(0, 256, (n, N, N), dtype=np.uint8)).cuda()
        mA = tA.reshape(n, -1).amax(1).clamp(min=16).float().contiguous()
        mB = tB.reshape(n, -1).amax(1).clamp(min=16).float().contiguous()
        D = 2 * N - 1
        ref = torch.empty((n, D, D), dtype=torch.float32, device="cuda")
        got = torch.full((n, D, D), float("nan"), dtype=torch.float32, device="cuda")
        say(f"noalloc N={N} mode={mode} n={n}")
        p = lambda t: ctypes.c_void_p(t.data_ptr())
        qlpiv.check(lib.qlp_xcorr_batch_dev(ctx.h, n, N, p(tA), p(tB), p(mA), p(mB), mode, p(ref)))
        qlpiv.check(lib.qlp_xcorr_batch_noalloc_dev(ctx.h, n, N, p(tA), p(tB), p(mA), p(mB), mode,
                                                    p(got)))
        ctx.synchronize()
        assert torch.equal(got, ref), (N, mode, n)
    ctx.close()


def test_noalloc_rejects_bad_sizes():
    import qlpiv
    from qlpiv import _lib

    lib = _lib.load()
    ctx = qlpiv.Context(0)
    rc = lib.qlp_xcorr_batch_noalloc_dev(ctx.h, 4, 256, None, None, None, None, 0, None)
    assert rc == _lib.QLP_ERR_UNSUPPORTED or rc == _lib.QLP_ERR_ARG
    assert lib.qlp_xcorr_batch_noalloc_dev(ctx.h, -1, 32, None, None, None, None, 0, None) != 0
    ctx.close()
