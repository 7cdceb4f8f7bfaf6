"""The MFMA (matrix-core) forms of the correlation against an FP64 correlation.

The MFMA forms compute the same plane as the reference's FFT job (linear
cross-correlation of the normalised tiles, zero displacement at [N-1][N-1]) by
DFT-basis products in fp16/bf16 with fp32 accumulation; they are not
value-identical to the Java arithmetic, so they are held to stated error bounds
against an FP64 correlation of the same normalised tiles:
  max |plane - fp64| / max |fp64|   <= TOL[form]
"""
import sys

import numpy as np
import pytest

import oracle as O

F16, BF16, F16S, BF16S = 1, 2, 3, 4
NAMES = {F16: "f16 split", BF16: "bf16 split", F16S: "f16", BF16S: "bf16"}
# error bounds relative to the plane maximum (measured on MI355X: 6.9e-7,
# 9.4e-6, 6.9e-4 and 6.4e-3 at most over these cases)
TOL = {F16: 2e-6, BF16: 3e-5, F16S: 2e-3, BF16S: 2e-2}
# 64 px, longer sums (M = 128): measured 2.4e-6, 1.2e-5, 7.5e-4 and 6.4e-3
TOL64 = {F16: 7e-6, BF16: 4e-5, F16S: 2.5e-3, BF16S: 2e-2}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def normalise(tA, tB, mA, mB):
    """x / max * 16 + 1 in float32, per tile (Matrix.copyMatrixToArrayAndNormalizeAndOffset)."""
    a = (tA.astype(np.float32) / mA[:, None, None].astype(np.float32)) * np.float32(16) + np.float32(1)
    g = (tB.astype(np.float32) / mB[:, None, None].astype(np.float32)) * np.float32(16) + np.float32(1)
    return a.astype(np.float32), g.astype(np.float32)


def corr_fp64(a, g):
    """plane[p][q] = sum_x a(x) g(x + (p - (N-1), q - (N-1))) in FP64 (batched FFT)."""
    n, N, _ = a.shape
    M = 2 * N
    A = np.fft.fft2(a.astype(np.float64), s=(M, M))
    G = np.fft.fft2(g.astype(np.float64), s=(M, M))
    r = np.fft.ifft2(np.conj(A) * G).real
    idx = (np.arange(M - 1) - (N - 1)) % M
    return r[:, idx][:, :, idx]


def test_fp64_reference_is_the_direct_sum():
    """The FFT helper equals the oracle's brute-force direct sum."""
    rng = np.random.default_rng(5)
    for N in (2, 4, 8, 16):
        tA = rng.integers(0, 256, (3, N, N)).astype(np.uint8)
        tB = rng.integers(0, 256, (3, N, N)).astype(np.uint8)
        mA = np.maximum(16, tA.reshape(3, -1).max(1)).astype(np.float32)
        mB = np.maximum(16, tB.reshape(3, -1).max(1)).astype(np.float32)
        a, g = normalise(tA, tB, mA, mB)
        got = corr_fp64(a, g)
        for k in range(3):
            ref = O.xcorr_direct64(a[k, ::-1, ::-1], g[k])
            assert np.abs(got[k] - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.fixture(scope="module")
def ctx():
    import qlpiv

    c = qlpiv.Context(0)
    yield c
    c.close()


def _tiles(rng, n, N, kind):
    if kind == "random":
        tA = rng.integers(0, 256, (n, N, N)).astype(np.uint8)
        tB = rng.integers(0, 256, (n, N, N)).astype(np.uint8)
    elif kind == "shifted":
        # a particle image and its shifted copy (a correlation with a clear peak)
        base = rng.integers(0, 40, (n, N + 8, N + 8)).astype(np.int32)
        ys, xs = rng.integers(0, N + 8, (2, n, 12))
        for k in range(n):
            base[k, ys[k], xs[k]] = 255
        tA = base[:, 4:4 + N, 4:4 + N].astype(np.uint8)
        dy, dx = rng.integers(-3, 4, 2)
        tB = base[:, 4 + dy:4 + dy + N, 4 + dx:4 + dx + N].astype(np.uint8)
    elif kind == "zeros":
        tA = np.zeros((n, N, N), np.uint8)
        tB = np.zeros((n, N, N), np.uint8)
    else:  # saturated
        tA = np.full((n, N, N), 255, np.uint8)
        tB = np.full((n, N, N), 255, np.uint8)
    mA = np.maximum(16, tA.reshape(n, -1).max(1)).astype(np.float32)
    mB = np.maximum(16, tB.reshape(n, -1).max(1)).astype(np.float32)
    return tA, tB, mA, mB


@pytest.mark.gpu
@pytest.mark.parametrize("N", [16, 32, 64])
@pytest.mark.parametrize("form", [F16, BF16, F16S, BF16S])
def test_mfma_planes_against_fp64(ctx, N, form):
    import torch

    rng = np.random.default_rng(100 * N + form)
    for kind, n in (("random", 257), ("shifted", 130), ("zeros", 5), ("saturated", 7)):
        tA, tB, mA, mB = _tiles(rng, n, N, kind)
        for mode in (0, 1):
            say(f"mfma {NAMES[form]} N={N} {kind} n={n} mode={mode}")
            out = ctx.xcorr(torch.from_numpy(tA).cuda(), torch.from_numpy(tB).cuda(),
                            torch.from_numpy(mA).cuda(), torch.from_numpy(mB).cuda(), mode=mode,
                            form=form)
            ctx.synchronize()
            got = out.cpu().numpy()
            a, g = normalise(tA, tB, mA, mA if mode == 1 else mB)
            ref = corr_fp64(a, g)
            err = np.abs(got - ref).max(axis=(1, 2)) / np.abs(ref).max(axis=(1, 2))
            say(f"   max rel err {err.max():.3e}  rms {np.sqrt((err ** 2).mean()):.3e}")
            assert np.isfinite(got).all()
            assert err.max() <= (TOL64 if N == 64 else TOL)[form], (kind, mode, err.max())


@pytest.mark.gpu
def test_mfma_small_batches_and_errors(ctx):
    """Batches that leave waves of a work-group idle, n = 0, and the argument
    checks (IA size, form)."""
    import torch

    import qlpiv

    rng = np.random.default_rng(9)
    for n in (1, 2, 3, 4, 5, 8191):
        tA, tB, mA, mB = _tiles(rng, n, 32, "random")
        say(f"mfma small batch n={n}")
        out = ctx.xcorr(torch.from_numpy(tA).cuda(), torch.from_numpy(tB).cuda(),
                        torch.from_numpy(mA).cuda(), torch.from_numpy(mB).cuda(), form=F16)
        ctx.synchronize()
        a, g = normalise(tA, tB, mA, mB)
        ref = corr_fp64(a, g)
        err = np.abs(out.cpu().numpy() - ref).max(axis=(1, 2)) / np.abs(ref).max(axis=(1, 2))
        assert err.max() <= TOL[F16]
    d = torch.zeros(1, dtype=torch.uint8, device="cuda")
    f = torch.zeros(1, dtype=torch.float32, device="cuda")
    out = torch.zeros(1, dtype=torch.float32, device="cuda")
    assert ctx.lib.qlp_xcorr_form_dev(ctx.h, F16, 0, 32, None, None, None, None, 0, None) == 0
    with pytest.raises(qlpiv.QlpError):
        ctx.xcorr(torch.zeros((1, 128, 128), dtype=torch.uint8, device="cuda"),
                  torch.zeros((1, 128, 128), dtype=torch.uint8, device="cuda"), f, f, form=F16)
    with pytest.raises(qlpiv.QlpError):
        ctx.xcorr(torch.zeros((1, 32, 32), dtype=torch.uint8, device="cuda"),
                  torch.zeros((1, 32, 32), dtype=torch.uint8, device="cuda"), f, f, form=9)
    del d, out


@pytest.mark.gpu
@pytest.mark.parametrize("warp", [0, 1])
def test_mfma64_fused_pass_matches_its_planes(ctx, warp):
    """The 64 px matrix-core pass from images: planes equal those of the same
    form from the clipped / warped tiles, and its fused FromCenter peaks,
    floors and neighbourhoods equal the peak search on those planes."""
    import torch

    import qlpiv
    from qlpiv.geometry import level_grid
    from qlpiv.synth import particle_pair

    N, D = 64, 127
    a, b, _ = particle_pair(384, 384, seed=91, shift=(2.3, -1.6), core_radius=90.0, u_max=5.0)
    tops, lefts = level_grid(384, 384, N, 0.5)
    top = torch.from_numpy(tops.ravel().astype(np.int32)).cuda()
    left = torch.from_numpy(lefts.ravel().astype(np.int32)).cuda()
    n = top.numel()
    rng = np.random.default_rng(5 + warp)
    du = torch.from_numpy(rng.uniform(-4, 4, n).astype(np.float32)).cuda()
    dv = torch.from_numpy(rng.uniform(-4, 4, n).astype(np.float32)).cuda()
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    say(f"mfma64 fused pass warp={warp} n={n}")
    pl = torch.empty((n, D, D), dtype=torch.float32, device="cuda")
    pk = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    mf = torch.empty(n, dtype=torch.float32, device="cuda")
    nb = torch.empty((n, 2, 7), dtype=torch.float32, device="cuda")
    from qlpiv import _ptr
    qlpiv.check(ctx.lib.qlp_xcorr_images_form_dev(ctx.h, F16, n, N, qlpiv.QLP_NORM_PER_TILE, _ptr(da),
                                                  _ptr(db), 384, 384, _ptr(top), _ptr(left),
                                                  _ptr(du), _ptr(dv), warp, _ptr(pl), _ptr(pk),
                                                  _ptr(mf), _ptr(nb), 3))
    tA, mA = ctx.clip_tiles(da, top, left, N)
    if warp:
        tB, mB = ctx.warp_micro_tiles(db, top, left, du, dv, N)
    else:
        tbt = torch.floor(top.float() + du + 0.5).int()
        tbl = torch.floor(left.float() + dv + 0.5).int()
        tB, mB = ctx.clip_tiles(db, tbt, tbl, N)
    ref = ctx.xcorr(tA, tB, mA, mB, qlpiv.QLP_NORM_PER_TILE, form=F16)
    rpk, rmf = ctx.find_peaks(ref)
    ctx.synchronize()
    assert torch.equal(pl, ref)
    assert torch.equal(pk, rpk[:, 0, :]) and torch.equal(mf, rmf)
    pl_, pk_, nb_ = ref.cpu().numpy(), pk.cpu().numpy(), nb.cpu().numpy()
    for k in range(n):
        pi, pj = int(pk_[k, 0]), int(pk_[k, 1])
        for d in range(-3, 4):
            assert nb_[k, 0, d + 3] == (pl_[k, pi + d, pj] if 0 <= pi + d < D else 0.0)
            assert nb_[k, 1, d + 3] == (pl_[k, pi, pj + d] if 0 <= pj + d < D else 0.0)
