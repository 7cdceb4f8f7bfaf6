"""IA 128 (2N = 256) on the GPU against the oracle, through the C ABI.

The 256 x 256 complex tile does not fit in LDS; the kernel keeps it in an HBM
workspace slot per work-group and runs each 1D pass on batches of lines in LDS.
Bars as for the other sizes: planes value-identical to the Java arithmetic,
integer peaks bit-exact, floors and neighbourhoods equal.
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

N = 128
D = 2 * N - 1


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import qlpiv

    c = qlpiv.Context(0)
    yield c
    c.close()


def tiles(rng, n):
    """Particle-like tiles, plus the edge cases: all zero (max clamped to 16),
    saturated, and one bright pixel in a corner of A and the far corner of B."""
    A = np.where(rng.random((n, N, N)) < 0.05, rng.integers(100, 256, (n, N, N)),
                 rng.integers(0, 8, (n, N, N))).astype(np.uint8)
    B = (np.roll(A, (3, -5), axis=(1, 2)) // 2 + rng.integers(0, 20, (n, N, N)) // 2).astype(np.uint8)
    A[0] = 0
    B[0] = 0
    A[1] = 255
    B[1] = 255
    A[2] = 0
    B[2] = 0
    A[2, 0, 0] = 200
    B[2, N - 1, N - 1] = 200
    mA = np.maximum(A.reshape(n, -1).max(1), 16).astype(np.float32)
    mB = np.maximum(B.reshape(n, -1).max(1), 16).astype(np.float32)
    return A, B, mA, mB


@pytest.mark.parametrize("mode,n", [(0, 20), (1, 300)])
def test_ia128_planes_identical_to_java_arithmetic(ctx, mode, n):
    """n = 300 exceeds the 256 workspace slots, so work-groups loop over IAs."""
    import qlpiv

    A, B, mA, mB = tiles(np.random.default_rng(128 + mode), n)
    qmode = qlpiv.QLP_NORM_A_MAX if mode == 0 else qlpiv.QLP_NORM_PER_TILE
    say(f"IA 128 planes mode={mode} n={n}")
    got = ctx.xcorr(A, B, mA, mB, qmode)
    ref = O.xcorr_batch_u8(A, B, mA, mB, mode)
    diff = got != ref
    assert not diff.any(), (int(diff.sum()), float(np.abs(got - ref).max()))


def test_ia128_device_entry_and_empty_batch(ctx):
    import torch

    import qlpiv

    A, B, mA, mB = tiles(np.random.default_rng(5), 6)
    say("IA 128 device entry")
    dev = ctx.xcorr(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(),
                    torch.from_numpy(mA).cuda(), torch.from_numpy(mB).cuda(),
                    qlpiv.QLP_NORM_PER_TILE)
    ctx.synchronize()
    assert np.array_equal(dev.cpu().numpy(), O.xcorr_batch_u8(A, B, mA, mB, 1))
    empty = ctx.xcorr(np.zeros((0, N, N), np.uint8), np.zeros((0, N, N), np.uint8),
                      np.zeros(0, np.float32), np.zeros(0, np.float32), qlpiv.QLP_NORM_PER_TILE)
    assert empty.shape == (0, D, D)


@pytest.mark.parametrize("warp", [0, 1])
@pytest.mark.parametrize("mode", [0, 1])
def test_ia128_fused_pass_from_images(ctx, warp, mode):
    """Tiles cut from 384^2 images (shifted clip or micro warp of B, windows
    partly outside the frame), planes, FromCenter peaks, floors and the 7-cell
    neighbourhoods against the oracle."""
    import torch

    import qlpiv
    from qlpiv.geometry import level_grid
    from qlpiv.synth import particle_pair

    a, b, _ = particle_pair(384, 384, seed=77, shift=(2.6, -1.4), core_radius=90.0, u_max=5.0)
    tops, lefts = level_grid(384, 384, N, 0.5)
    top = tops.ravel().astype(np.int32)
    left = lefts.ravel().astype(np.int32)
    n = top.size
    rng = np.random.default_rng(31 + warp + 2 * mode)
    du = rng.uniform(-6, 6, n).astype(np.float32)
    dv = rng.uniform(-6, 6, n).astype(np.float32)
    qmode = qlpiv.QLP_NORM_A_MAX if mode == 0 else qlpiv.QLP_NORM_PER_TILE
    say(f"IA 128 fused pass warp={warp} mode={mode} n={n}")
    pl, pk, mf, nb = ctx.xcorr_images(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(),
                                      torch.from_numpy(top).cuda(), torch.from_numpy(left).cuda(), N,
                                      torch.from_numpy(du).cuda(), torch.from_numpy(dv).cuda(),
                                      warp_mode=warp, mode=qmode, planes=True, nbr_half=3)
    ctx.synchronize()
    pl, pk, mf, nb = (x.cpu().numpy() for x in (pl, pk, mf, nb))
    tA = np.empty((n, N, N), np.uint8)
    tB = np.empty((n, N, N), np.uint8)
    mA = np.empty(n, np.float32)
    mB = np.empty(n, np.float32)
    for k in range(n):
        tA[k], mA[k] = O.clip_tile(a, top[k], left[k], N)
        if warp:
            tB[k] = O.warp_micro(b, top[k], left[k], N, du[k], dv[k])
            mB[k] = 255.0
        else:
            t0 = int(np.floor(np.float32(top[k]) + du[k] + np.float32(0.5)))
            l0 = int(np.floor(np.float32(left[k]) + dv[k] + np.float32(0.5)))
            tB[k], mB[k] = O.clip_tile(b, t0, l0, N)
    ref = O.xcorr_batch_u8(tA, tB, mA, mB, mode)
    assert np.array_equal(pl, ref)
    for k in range(n):
        rp = O.find_max_from_center(ref[k])
        assert tuple(pk[k]) == tuple(rp[:3]) and mf[k] == rp[3], k
        pi, pj = int(rp[0]), int(rp[1])
        for d in range(-3, 4):
            assert nb[k, 0, d + 3] == (ref[k, pi + d, pj] if 0 <= pi + d < D else 0.0)
            assert nb[k, 1, d + 3] == (ref[k, pi, pj + d] if 0 <= pj + d < D else 0.0)


def test_ia128_mfma_form_is_rejected(ctx):
    import torch

    import qlpiv

    z = torch.zeros((1, N, N), dtype=torch.uint8, device="cuda")
    m = torch.full((1,), 16.0, dtype=torch.float32, device="cuda")
    with pytest.raises(qlpiv.QlpError):
        ctx.xcorr(z, z, m, m, qlpiv.QLP_NORM_PER_TILE, form=qlpiv._lib.QLP_XCORR_MFMA_F16)
