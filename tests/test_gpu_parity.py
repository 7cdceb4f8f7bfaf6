"""GPU parity of the HIP path against the CPU oracle (the C restatement of the
reference's Java code), through the C ABI.

Bars: correlation planes value-identical to the Java arithmetic (float ==),
integer peaks bit-exact, sub-pixel results within 1e-4 px, tiles bit-exact.
"""
import os
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

SUBPIX_TOL = 1e-4  # px, sub-pixel displacement vs the reference's Java path


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import qlpiv

    c = qlpiv.Context(0)
    yield c
    c.close()


def sparse_tiles(rng, n, N, density=0.06):
    """Particle-like tiles: mostly dark with bright spots, a few saturated."""
    t = np.where(rng.random((n, N, N)) < density, rng.integers(100, 256, (n, N, N)),
                 rng.integers(0, 8, (n, N, N)))
    return t.astype(np.uint8)


def tile_max(t):
    return np.maximum(t.reshape(len(t), -1).max(1), 16).astype(np.float32)


@pytest.mark.parametrize("N", [16, 32, 64])
@pytest.mark.parametrize("mode", [0, 1])
def test_xcorr_planes_identical_to_java_arithmetic(ctx, N, mode):
    import qlpiv

    rng = np.random.default_rng(1000 + N + mode)
    n = 48
    A = sparse_tiles(rng, n, N)
    B = np.roll(A, (2, -3), axis=(1, 2)) // 2 + sparse_tiles(rng, n, N, 0.01) // 2
    mA, mB = tile_max(A), tile_max(B)
    qmode = qlpiv.QLP_NORM_A_MAX if mode == 0 else qlpiv.QLP_NORM_PER_TILE
    say(f"xcorr parity N={N} mode={mode}")
    got = ctx.xcorr(A, B, mA, mB, qmode)
    ref = O.xcorr_batch_u8(A, B, mA, mB, mode)
    assert got.shape == ref.shape
    diff = got != ref
    assert not diff.any(), (int(diff.sum()), float(np.abs(got - ref).max()))


def test_xcorr_edge_tiles(ctx):
    """All-zero tiles (max clamped to 16), saturated tiles, a lone pixel in a
    corner, and the empty batch."""
    import qlpiv

    N = 32
    A = np.zeros((4, N, N), np.uint8)
    B = np.zeros((4, N, N), np.uint8)
    A[1] = 255
    B[1] = 255
    A[2, 0, 0] = 200
    B[2, N - 1, N - 1] = 200
    A[3, 5:9, 7:11] = 250
    B[3, 6:10, 9:13] = 250
    mA, mB = tile_max(A), tile_max(B)
    for mode, qm in ((0, qlpiv.QLP_NORM_A_MAX), (1, qlpiv.QLP_NORM_PER_TILE)):
        got = ctx.xcorr(A, B, mA, mB, qm)
        ref = O.xcorr_batch_u8(A, B, mA, mB, mode)
        assert np.array_equal(got, ref)
    empty = ctx.xcorr(np.zeros((0, N, N), np.uint8), np.zeros((0, N, N), np.uint8),
                      np.zeros(0, np.float32), np.zeros(0, np.float32), qlpiv.QLP_NORM_PER_TILE)
    assert empty.shape == (0, 2 * N - 1, 2 * N - 1)


def test_golden_planes_on_gpu(ctx, rankine_pair, golden_planes):
    """The reference's device self-test on the GPU path: 1024 planes of 31x31
    within 1e-2 of the stored ones (GpuFFTBenchmarkJob.java:316-334) and equal to
    the oracle value for value."""
    import qlpiv

    a, b = rankine_pair
    N, n = 16, 32
    tA = np.ascontiguousarray(a.reshape(n, N, n, N).transpose(0, 2, 1, 3).reshape(-1, N, N))
    tB = np.ascontiguousarray(b.reshape(n, N, n, N).transpose(0, 2, 1, 3).reshape(-1, N, N))
    mA, mB = tile_max(tA), tile_max(tB)
    say("golden planes on GPU")
    got = ctx.xcorr(tA, tB, mA, mB, qlpiv.QLP_NORM_PER_TILE)
    assert np.abs(got - golden_planes).max() <= 1e-2
    assert np.array_equal(got, O.xcorr_batch_u8(tA, tB, mA, mB, 1))


def test_device_entry_matches_host_entry(ctx):
    import torch

    import qlpiv

    rng = np.random.default_rng(7)
    A = sparse_tiles(rng, 100, 32)
    B = sparse_tiles(rng, 100, 32)
    mA, mB = tile_max(A), tile_max(B)
    host = ctx.xcorr(A, B, mA, mB, qlpiv.QLP_NORM_PER_TILE)
    dev = ctx.xcorr(torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(),
                    torch.from_numpy(mA).cuda(), torch.from_numpy(mB).cuda(),
                    qlpiv.QLP_NORM_PER_TILE)
    ctx.synchronize()
    assert np.array_equal(dev.cpu().numpy(), host)


# --------------------------------------------------------------------------
# peaks
# --------------------------------------------------------------------------

def planes_for_peaks(rng, n, D):
    p = rng.random((n, D, D)).astype(np.float32) * 100
    # ties, zeros, negatives, maxima on row 0 / col 0 and on the scan centre
    p[0] = np.round(p[0] / 25)
    p[1, 0, :] = 1e4
    p[1, :, 0] = 1e4
    p[2] = -1.0
    p[3] = 0.0
    p[4, D // 2 + 1, D // 2 + 1] = 500
    p[5, 3, 4] = p[5, D - 2, D - 3] = 700  # equal maxima, different distances
    p[6, 7, 7] = p[6, D - 7 + 1, D - 7 + 1] = 700  # equal maxima, equal distances
    p[7, min(10, D - 1), min(12, D - 1)] = -0.0
    return p


@pytest.mark.parametrize("D", [31, 63, 127])
def test_from_center_bit_exact(ctx, D):
    import qlpiv

    rng = np.random.default_rng(D)
    planes = planes_for_peaks(rng, 64, D)
    say(f"from-center D={D}")
    peaks, mf = ctx.find_peaks(planes, qlpiv.QLP_PEAK_FROM_CENTER, 1)
    for k in range(len(planes)):
        ref = O.find_max_from_center(planes[k])
        assert tuple(peaks[k, 0]) == tuple(ref[:3]), k
        assert mf[k] == ref[3], k


@pytest.mark.parametrize("D,P,kw", [(31, 3, 3), (63, 3, 3), (63, 4, 5), (127, 5, 3), (10, 4, 3),
                                     (63, 16, 3), (127, 3, 1), (31, 1, 3)])
def test_multipeaks_bit_exact(ctx, D, P, kw):
    import qlpiv

    rng = np.random.default_rng(D * 10 + P)
    planes = planes_for_peaks(rng, 48, D)
    planes[10] = np.round(planes[10] / 10)  # many equal block maxima
    say(f"multipeaks D={D} P={P} k={kw}")
    peaks, mf = ctx.find_peaks(planes, qlpiv.QLP_PEAK_MULTI, P, kw)
    for k in range(len(planes)):
        ref, found, rmf = O.find_max_multipeaks(planes[k], P, kw)
        assert np.array_equal(peaks[k], ref), (k, peaks[k], ref)
        assert mf[k] == rmf


@pytest.mark.parametrize("W", [8, 16, 32, 64])
@pytest.mark.parametrize("lst", ["lanes", "lds"])
def test_multipeaks_every_segment_width_and_list_form(ctx, W, lst):
    """Both forms of the search's bounded list (in the lanes of a plane's
    segment, or in LDS) at every segment width, against the oracle: P below,
    equal to and above the width (the LDS list then), kernel widths 1, 2, 3."""
    import qlpiv

    saved = {k: os.environ.get(k) for k in ("QLPIV_MULTIPEAK_LANES", "QLPIV_MULTIPEAK_LIST",
                                            "QLPIV_MULTIPEAK_FORM")}
    os.environ["QLPIV_MULTIPEAK_LANES"] = str(W)
    os.environ["QLPIV_MULTIPEAK_LIST"] = lst
    os.environ["QLPIV_MULTIPEAK_FORM"] = "segment"
    try:
        for D, P, kw in [(31, 3, 3), (31, 8, 2), (63, 16, 3), (63, 1, 3), (15, 5, 1)]:
            rng = np.random.default_rng(W + D + P)
            planes = planes_for_peaks(rng, 40, D)
            planes[7] = np.round(planes[7] / 10)  # many equal block maxima
            say(f"multipeaks W={W} list={lst} D={D} P={P} k={kw}")
            peaks, mf = ctx.find_peaks(planes, qlpiv.QLP_PEAK_MULTI, P, kw)
            for k in range(len(planes)):
                ref, found, rmf = O.find_max_multipeaks(planes[k], P, kw)
                assert np.array_equal(peaks[k], ref), (D, P, kw, k)
                assert mf[k] == rmf, (D, P, kw, k)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("G", [1, 4, 16, 64])
def test_multipeaks_replay_every_plane_count(ctx, G):
    """The two-kernel form (kernel width 3, P <= 4: block maxima by one
    wavefront per plane, then one lane per plane replaying the list) at several
    planes-per-work-group counts of the replay, against the oracle: the C2
    plane sizes and odd ones, P = 1..4, batches that leave lanes of the last
    work-group idle, planes with ties, zeros, negatives (blocks without an
    update) and equal maxima."""
    import qlpiv

    saved = {k: os.environ.get(k) for k in ("QLPIV_MULTIPEAK_PLANES", "QLPIV_MULTIPEAK_FORM")}
    os.environ["QLPIV_MULTIPEAK_PLANES"] = str(G)
    os.environ["QLPIV_MULTIPEAK_FORM"] = "entries"
    try:
        for D, P, n in [(31, 3, 48), (63, 3, 37), (127, 3, 21), (63, 4, 40), (31, 1, 33),
                        (127, 2, 9), (15, 4, 50), (10, 3, 19), (5, 2, 12)]:
            rng = np.random.default_rng(G * 1000 + D * 10 + P)
            planes = planes_for_peaks(rng, max(n, 8), D) if D >= 10 else \
                rng.random((n, D, D)).astype(np.float32) * 10
            planes = planes[:n]
            if D >= 10:
                planes[7] = np.round(planes[7] / 10)  # many equal block maxima
            say(f"multipeaks replay G={G} D={D} P={P} n={n}")
            peaks, mf = ctx.find_peaks(planes, qlpiv.QLP_PEAK_MULTI, P, 3)
            for k in range(n):
                ref, found, rmf = O.find_max_multipeaks(planes[k], P, 3)
                assert np.array_equal(peaks[k], ref), (D, P, k, peaks[k], ref)
                assert mf[k] == rmf, (D, P, k)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def test_multipeaks_replay_on_correlation_planes(ctx):
    """The entries form (block maxima kernel, then the replay) against the
    segment form and the oracle on correlation planes of particle images at the
    three C2 sizes (the planes the step's search sees: one sharp peak over a
    particle-pair background)."""
    import qlpiv
    from qlpiv.geometry import level_grid
    from qlpiv.synth import particle_pair

    for N in (16, 32, 64):
        a, b, _ = particle_pair(384, 384, seed=N, shift=(2.3, -1.6))
        tops, lefts = level_grid(384, 384, N, 0.5)
        top, left = tops.ravel().astype(np.int32), lefts.ravel().astype(np.int32)
        n = top.size
        u = np.full(n, 2.0, np.float32)
        v = np.full(n, -2.0, np.float32)
        _, _, pl = O.level_batch(a, b, top, left, u, v, N, 1, mode=1, peak_kind=1, P=3, k=3,
                                 planes=True)
        say(f"multipeaks replay on correlation planes N={N} n={n}")
        saved = os.environ.get("QLPIV_MULTIPEAK_FORM")
        try:
            os.environ["QLPIV_MULTIPEAK_FORM"] = "segment"
            speaks, smf = ctx.find_peaks(pl, qlpiv.QLP_PEAK_MULTI, 3, 3)
            os.environ["QLPIV_MULTIPEAK_FORM"] = "entries"
            peaks, mf = ctx.find_peaks(pl, qlpiv.QLP_PEAK_MULTI, 3, 3)
        finally:
            if saved is None:
                os.environ.pop("QLPIV_MULTIPEAK_FORM", None)
            else:
                os.environ["QLPIV_MULTIPEAK_FORM"] = saved
        assert np.array_equal(peaks, speaks) and np.array_equal(mf, smf)
        for k in range(0, n, 7):
            ref, found, rmf = O.find_max_multipeaks(pl[k], 3, 3)
            assert np.array_equal(peaks[k], ref), (N, k)
            assert mf[k] == rmf, (N, k)


def test_multipeaks_reference_kat(ctx):
    import json
    import os

    import qlpiv
    from conftest import GOLDEN

    kat = json.load(open(os.path.join(GOLDEN, "multipeak_kat.json")))
    plane = np.array(kat["plane"], np.float32)[None]
    peaks, _ = ctx.find_peaks(plane, qlpiv.QLP_PEAK_MULTI, kat["numberOfPeaks"], kat["kernelWidth"])
    for p, e in enumerate(kat["expect"]):
        if e[0] is not None:
            assert (peaks[0, p, 0], peaks[0, p, 1]) == (e[0], e[1])
        assert abs(peaks[0, p, 2] - e[2]) <= kat["tol"]


# --------------------------------------------------------------------------
# sub-pixel
# --------------------------------------------------------------------------

def gaussian_peak_planes(rng, n, D):
    """Planes with one Gaussian peak at a random sub-pixel position plus floor."""
    yy, xx = np.mgrid[0:D, 0:D]
    planes = np.empty((n, D, D), np.float32)
    peaks = np.empty((n, 1, 3), np.float32)
    for k in range(n):
        ci, cj = rng.uniform(4, D - 5, 2)
        s = rng.uniform(0.7, 2.0)
        a = rng.uniform(1e3, 1e5)
        p = a * np.exp(-((yy - ci) ** 2 + (xx - cj) ** 2) / (2 * s * s)) + rng.uniform(1, 50)
        p += rng.normal(0, 0.002 * a, p.shape)
        planes[k] = p.astype(np.float32)
        i, j = np.unravel_index(np.argmax(planes[k]), (D, D))
        peaks[k, 0] = (i, j, planes[k, i, j])
    return planes, peaks


@pytest.mark.parametrize("kind,pixels", [("gauss", 3), ("gauss", 5), ("gauss", 7), ("guo", 3),
                                         ("guo", 5), ("guo", 7), ("guo", 9), ("guo", 15),
                                         ("poly", 3)])
def test_subpixel_vs_oracle(ctx, kind, pixels):
    import qlpiv

    rng = np.random.default_rng(pixels + len(kind))
    D = 63
    planes, peaks = gaussian_peak_planes(rng, 200, D)
    # border and origin cases keep the integer peak
    peaks[0, 0, :2] = (0, 0)
    peaks[1, 0, :2] = (1, 30)
    mf = planes.reshape(len(planes), -1).min(1).astype(np.float32)
    mf[2] = -1.0  # negative floor is added back before the fit
    qk = {"gauss": qlpiv.QLP_SUBPIX_GAUSS1D, "guo": qlpiv.QLP_SUBPIX_GUO,
          "poly": qlpiv.QLP_SUBPIX_POLY}[kind]
    say(f"subpixel {kind} {pixels}")
    res = ctx.subpixel(planes, mf, peaks, qk, pixels, 20)
    for k in range(len(planes)):
        if kind == "gauss":
            ref = O.gauss1d(planes[k], pixels, mf[k], peaks[k, 0])
        elif kind == "guo":
            ref = O.guo(planes[k], pixels, 20, mf[k], peaks[k, 0])
        else:
            ref = O.poly(planes[k], peaks[k, 0])
        assert res[k, 0, 3] == ref[3], (k, res[k, 0], ref)
        # wide windows can take the log of a non-positive sample: both sides
        # then carry NaN through the fit
        for c in (0, 1):
            assert (np.isnan(res[k, 0, c]) and np.isnan(ref[c])) or \
                abs(res[k, 0, c] - ref[c]) <= SUBPIX_TOL, (k, res[k, 0], ref)
        if kind == "guo" and not np.isnan(ref[2]):
            assert abs(res[k, 0, 2] - ref[2]) <= 1e-4 * abs(ref[2]) + 1e-3


# --------------------------------------------------------------------------
# tiles
# --------------------------------------------------------------------------

def test_clip_and_warp_tiles_bit_exact(ctx):
    import torch

    rng = np.random.default_rng(9)
    H, W, N = 200, 260, 32
    img = rng.integers(0, 256, (H, W)).astype(np.uint8)
    n = 60
    top = rng.integers(-20, H - N + 20, n).astype(np.int32)
    left = rng.integers(-20, W - N + 20, n).astype(np.int32)
    u = rng.uniform(-7.5, 7.5, n).astype(np.float32)
    v = rng.uniform(-7.5, 7.5, n).astype(np.float32)
    u[:6] = [0.0, 1.0, -2.0, 0.5, -0.5, 3.25]
    dimg = torch.from_numpy(img).cuda()
    say("clip/warp tiles")
    tiles, mx = ctx.clip_tiles(dimg, torch.from_numpy(top).cuda(), torch.from_numpy(left).cuda(), N)
    wt, wmx = ctx.warp_micro_tiles(dimg, torch.from_numpy(top).cuda(), torch.from_numpy(left).cuda(),
                                   torch.from_numpy(u).cuda(), torch.from_numpy(v).cuda(), N)
    ctx.synchronize()
    tiles, mx, wt, wmx = tiles.cpu().numpy(), mx.cpu().numpy(), wt.cpu().numpy(), wmx.cpu().numpy()
    for k in range(n):
        rt, rm = O.clip_tile(img, top[k], left[k], N)
        assert np.array_equal(tiles[k], rt) and mx[k] == rm, k
        assert np.array_equal(wt[k], O.warp_micro(img, top[k], left[k], N, u[k], v[k])), k
        assert wmx[k] == 255.0


# --------------------------------------------------------------------------
# the job interface and error behaviour
# --------------------------------------------------------------------------

def test_job_interface_and_errors(ctx):
    import qlpiv
    from qlpiv import CrossCorrelationJob, JobResultEnum

    rng = np.random.default_rng(4)
    A = sparse_tiles(rng, 5, 16)
    B = sparse_tiles(rng, 5, 16)
    job = CrossCorrelationJob(ctx, qlpiv.QLP_NORM_PER_TILE)
    job.set_input_parameters(JobResultEnum.JOB_RESULT_CLIPPED_TILES_A, [(t, float(max(16, t.max()))) for t in A])
    job.set_input_parameters(JobResultEnum.JOB_RESULT_CLIPPED_TILES_B, [(t, float(max(16, t.max()))) for t in B])
    job.analyze()
    job.compute()
    planes = job.get_job_result(JobResultEnum.JOB_RESULT_CROSS_MATRICES)
    ref = O.xcorr_batch_u8(A, B, tile_max(A), tile_max(B), 1)
    assert len(planes) == 5 and all(np.array_equal(p, r) for p, r in zip(planes, ref))
    job.set_input_parameters(JobResultEnum.JOB_RESULT_CLIPPED_TILES_B, [(t, 16.0) for t in B[:4]])
    with pytest.raises(RuntimeError):
        job.analyze()
    with pytest.raises(qlpiv.QlpError) as e:
        ctx.xcorr(np.zeros((2, 256, 256), np.uint8), np.zeros((2, 256, 256), np.uint8),
                  np.ones(2, np.float32), np.ones(2, np.float32))
    assert e.value.code == qlpiv._lib.QLP_ERR_ARG
    with pytest.raises(qlpiv.QlpError) as e:
        ctx.xcorr(np.zeros((2, 24, 24), np.uint8), np.zeros((2, 24, 24), np.uint8),
                  np.ones(2, np.float32), np.ones(2, np.float32))
    assert e.value.code == qlpiv._lib.QLP_ERR_ARG


@pytest.mark.parametrize("D", [31, 63, 127])
def test_simple_bit_exact(ctx, D):
    """FindMaximumSimple: row 0 and column 0 included, first maximum in row-major
    order, (D/2, D/2, 0) without a positive value, NaN never taken."""
    import qlpiv

    rng = np.random.default_rng(D + 1)
    planes = planes_for_peaks(rng, 64, D)
    planes[8] = np.nan
    planes[9, 1, 1] = np.nan
    planes[10, 0, 0] = 1e5
    say(f"simple D={D}")
    peaks, mf = ctx.find_peaks(planes, qlpiv.QLP_PEAK_SIMPLE, 1)
    for k in range(len(planes)):
        ref = O.find_max_simple(planes[k])
        assert tuple(peaks[k, 0]) == tuple(ref[:3]), k
        assert mf[k] == ref[3], k
