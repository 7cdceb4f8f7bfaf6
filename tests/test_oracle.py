"""Pin the CPU oracle (the C restatement of the reference's Java path) against the
reference's golden fixtures, its unit-test KATs and what the mathematics fixes.

These run on the CPU only; they are the reason the GPU parity tests may trust
the oracle.
"""
import math

import numpy as np
import pytest

import oracle as O
from conftest import gaussian_plane, load_json


# --------------------------------------------------------------------------
# FFT building blocks
# --------------------------------------------------------------------------

@pytest.mark.parametrize("m", range(1, 9))
def test_perfect_shuffle_is_bit_reversal(m):
    n = 1 << m
    out = O.perfect_shuffle(np.arange(n, dtype=np.float32)).astype(int)
    br = [int(format(i, "0%db" % m)[::-1], 2) for i in range(n)] if m else [0]
    assert list(out) == br


@pytest.mark.parametrize("M", [8, 16, 32, 64, 128, 256])
def test_sine_table_robust_to_last_ulp_of_double_sin(M):
    """The reference builds its twiddles as (float)FastMath.sin(double).  FastMath.sin
    and libm's sin may differ in the last double ulp; the float table is the same
    for both whenever no entry lies within a few double ulps of a float rounding
    boundary.  Checked here for every table size used."""
    ws = O.sine_table(M)
    for k in range(M):
        d = math.sin(2.0 * math.pi * k / M)
        assert ws[k] == np.float32(d)
        if d == 0.0:
            continue
        f = np.float32(d)
        # distance of d to the nearest midpoint between adjacent floats
        lo = np.nextafter(f, np.float32(-np.inf))
        hi = np.nextafter(f, np.float32(np.inf))
        mid_lo = (float(f) + float(lo)) / 2
        mid_hi = (float(f) + float(hi)) / 2
        margin = min(abs(d - mid_lo), abs(d - mid_hi))
        assert margin > 16 * np.spacing(abs(d)), (M, k, margin)


def test_sin_pi_is_not_special_cased():
    ws = O.sine_table(64)
    assert ws[32] != 0.0 and abs(ws[32]) < 1e-15  # sin(pi) in double, cast to float


# --------------------------------------------------------------------------
# Cross-correlation
# --------------------------------------------------------------------------

def test_fastreal_kat_against_direct_sum():
    """FastRealFFTXCorrTests.java:31-51,195-199: un-normalised 4x4 tiles, FFT result
    equals the direct sum within 1e-3."""
    kat = load_json("fastreal_kat.json")
    F = np.array(kat["F"], np.float32)
    G = np.array(kat["G"], np.float32)
    got = O.xcorr_floats(F, G)
    ref = O.xcorr_direct64(F[::-1, ::-1], G)
    np.testing.assert_allclose(got, ref, atol=kat["tol"], rtol=0)


@pytest.mark.parametrize("N", [2, 4, 8, 16])
@pytest.mark.parametrize("mode", [0, 1])
def test_xcorr_restatement_matches_direct_sum(N, mode):
    rng = np.random.default_rng(100 + N + 7 * mode)
    A = rng.integers(0, 256, (N, N)).astype(np.float32)
    B = rng.integers(0, 256, (N, N)).astype(np.float32)
    mA, mB = max(16.0, A.max()), max(16.0, B.max())
    if mode == 0:
        got = O.xcorr_javacpu(A, B, mA)
        F, G = O.normalise_pair(A, B, mA, mA)
    else:
        got = O.xcorr_javagpu(A, B, mA, mB)
        F, G = O.normalise_pair(A, B, mA, mB)
    ref = O.xcorr_direct64(F, G)
    assert np.max(np.abs(got - ref)) <= 2e-6 * np.max(ref)
    # the argmax of the plane is the brute-force one
    assert np.unravel_index(np.argmax(got), got.shape) == np.unravel_index(np.argmax(ref), ref.shape)


def test_fp64_fft_helper_is_the_direct_sum():
    """xcorr_fft64 equals the brute-force sum: whole planes for small N, and 40
    random cells (borders included) of a 128 px plane."""
    rng = np.random.default_rng(17)
    for N in (2, 4, 8, 16):
        F = rng.uniform(1, 17, (N, N)).astype(np.float32)
        G = rng.uniform(1, 17, (N, N)).astype(np.float32)
        ref = O.xcorr_direct64(F, G)
        assert np.abs(O.xcorr_fft64(F, G) - ref).max() <= 1e-9 * ref.max()
    N = 128
    Fm = rng.uniform(1, 17, (N, N))
    G = rng.uniform(1, 17, (N, N))
    got = O.xcorr_fft64(Fm, G)
    F = Fm[::-1, ::-1]
    cells = [(0, 0), (2 * N - 2, 2 * N - 2), (N - 1, N - 1), (0, 2 * N - 2)]
    cells += [tuple(c) for c in rng.integers(0, 2 * N - 1, (36, 2))]
    for pi, pj in cells:
        i, j = pi - (N - 1), pj - (N - 1)
        n0, n1 = max(0, -i), min(N, N - i)
        m0, m1 = max(0, -j), min(N, N - j)
        ref = np.sum(F[n0:n1, m0:m1] * G[n0 + i:n1 + i, m0 + j:m1 + j])
        assert abs(got[pi, pj] - ref) <= 1e-9 * abs(ref), (pi, pj)


@pytest.mark.parametrize("N", [32, 64, 128])
@pytest.mark.parametrize("mode", [0, 1])
def test_xcorr_restatement_matches_fp64_fft_at_large_sizes(N, mode):
    """The IA sizes the direct sum is too slow for, up to 128 (M = 256 twiddles):
    against an FP64 FFT of the same normalised tiles (itself pinned to the direct
    sum below).  A wrong twiddle or stage at these sizes moves the plane by far
    more than float rounding (measured: <= 4.3e-7 of the maximum)."""
    rng = np.random.default_rng(200 + N + mode)
    A = rng.integers(0, 256, (N, N)).astype(np.float32)
    B = rng.integers(0, 256, (N, N)).astype(np.float32)
    mA, mB = max(16.0, A.max()), max(16.0, B.max())
    if mode == 0:
        got = O.xcorr_javacpu(A, B, mA)
        F, G = O.normalise_pair(A, B, mA, mA)
    else:
        got = O.xcorr_javagpu(A, B, mA, mB)
        F, G = O.normalise_pair(A, B, mA, mB)
    ref = O.xcorr_fft64(F, G)
    assert np.max(np.abs(got - ref)) <= 2e-6 * np.max(ref)
    assert np.unravel_index(np.argmax(got), got.shape) == np.unravel_index(np.argmax(ref), ref.shape)


def test_xcorr_modes_differ_only_by_normalisation_and_inverse_twiddles():
    """With A's max used for G too, the GPU-job arithmetic and the CPU path compute
    the same plane up to float rounding of the inverse twiddles."""
    rng = np.random.default_rng(5)
    A = rng.integers(0, 256, (32, 32)).astype(np.float32)
    B = rng.integers(0, 256, (32, 32)).astype(np.float32)
    c = O.xcorr_javacpu(A, B, 200.0)
    g = O.xcorr_javagpu(A, B, 200.0, 200.0)
    assert np.max(np.abs(c - g)) <= 1e-6 * np.max(c)


def test_shift_of_ten_pixels_lands_at_expected_peak():
    """ImageCrossCorrelationRealFFTParStdJobTests.java:232-306 recreated with a
    synthetic image: B's window sits 10 px further into the image than A's, so
    the content moved by -10 px and the plane maximum sits at (dim/2 - 10,
    dim/2 - 10), as the reference's test asserts."""
    rng = np.random.default_rng(3)
    img = np.where(rng.random((128, 128)) < 0.05, rng.integers(150, 256, (128, 128)), 0)
    img = img.astype(np.uint8)  # sparse particles, as in a PIV image
    N = 32
    A = img[40:40 + N, 40:40 + N].astype(np.float32)
    B = img[50:50 + N, 50:50 + N].astype(np.float32)
    p = O.xcorr_javagpu(A, B, A.max(), B.max())
    i, j = np.unravel_index(np.argmax(p), p.shape)
    assert (i, j) == (N - 1 - 10, N - 1 - 10)


def _rankine_tiles(rankine_pair, N=16):
    a, b = rankine_pair
    n = 512 // N
    tA = a.reshape(n, N, n, N).transpose(0, 2, 1, 3).reshape(-1, N, N)
    tB = b.reshape(n, N, n, N).transpose(0, 2, 1, 3).reshape(-1, N, N)
    mA = np.maximum(tA.reshape(len(tA), -1).max(1), 16).astype(np.float32)
    mB = np.maximum(tB.reshape(len(tB), -1).max(1), 16).astype(np.float32)
    return np.ascontiguousarray(tA), np.ascontiguousarray(tB), mA, mB


def test_golden_gpu_planes(rankine_pair, golden_planes):
    """The reference's own device self-test (GpuFFTBenchmarkJob.java:161-245,316-334):
    16 x 16 IAs without overlap over the 512^2 Rankine pair, every element within
    1e-2 of the stored planes, with per-tile normalisation."""
    tA, tB, mA, mB = _rankine_tiles(rankine_pair)
    planes = O.xcorr_batch_u8(tA, tB, mA, mB, 1)
    assert planes.shape == golden_planes.shape == (1024, 31, 31)
    err = np.abs(planes - golden_planes)
    assert err.max() <= 1e-2, err.max()
    am = planes.reshape(1024, -1).argmax(1)
    ag = golden_planes.reshape(1024, -1).argmax(1)
    assert np.all(am == ag)
    # the CPU-path normalisation does not reproduce the fixture (F5 of the survey)
    cpu = O.xcorr_batch_u8(tA[:64], tB[:64], mA[:64], mA[:64], 0)
    assert np.abs(cpu - golden_planes[:64]).max() > 1.0


# --------------------------------------------------------------------------
# Peak search
# --------------------------------------------------------------------------

def from_center_brute(plane):
    """Sequential semantics of FindMaximumFromCenter restated as a key order:
    largest value over rows/cols 1..D-1, then smallest distance to
    (D/2+1, D/2+1), then earliest visit; (0,0) with value 0 when nothing > 0."""
    D = plane.shape[0]
    h = D // 2
    c = h + 1
    visit = {}
    idx = 0

    def mark(i, j):
        nonlocal idx
        if (i, j) not in visit:
            visit[(i, j)] = idx
        idx += 1
    mark(c, c)
    for i in range(h):
        mark(h - i, c)
        mark(h + 1 + i, c)
    for j in range(h):
        mark(c, h - j)
        mark(c, h + 1 + j)
    for i in range(h):
        for j in range(h):
            mark(h - i, h - j)
            mark(h + 1 + i, h - j)
            mark(h - i, h + 1 + j)
            mark(h + 1 + i, h + 1 + j)
    best = None
    for (i, j), v in visit.items():
        val = plane[i, j]
        if not (val >= 0):
            continue
        key = (-val, (i - c) ** 2 + (j - c) ** 2, v)
        if best is None or key < best[0]:
            best = (key, i, j, val)
    mn = min(plane[i, j] for (i, j) in visit)
    if best is None:
        return (0, 0, 0.0, mn)
    return (best[1], best[2], best[3], mn)


@pytest.mark.parametrize("seed", range(8))
def test_from_center_vs_brute_force(seed):
    rng = np.random.default_rng(seed)
    D = [7, 15, 31, 63][seed % 4]
    plane = rng.integers(0, 6, (D, D)).astype(np.float32)  # many ties
    if seed % 2:
        plane[0, :] = 99.0  # row 0 / col 0 are never visited
        plane[:, 0] = 99.0
    got = O.find_max_from_center(plane)
    exp = from_center_brute(plane)
    assert (int(got[0]), int(got[1]), float(got[2])) == (exp[0], exp[1], float(exp[2]))
    assert got[3] == exp[3]


def test_from_center_all_negative_keeps_origin():
    plane = -np.ones((15, 15), np.float32)
    got = O.find_max_from_center(plane)
    assert tuple(got[:3]) == (0.0, 0.0, 0.0) and got[3] == -1.0


def test_multipeak_kat():
    kat = load_json("multipeak_kat.json")
    plane = np.array(kat["plane"], np.float32)
    peaks, found, _ = O.find_max_multipeaks(plane, kat["numberOfPeaks"], kat["kernelWidth"])
    for p, e in enumerate(kat["expect"]):
        if e[0] is not None:
            assert (peaks[p, 0], peaks[p, 1]) == (e[0], e[1])
        assert abs(peaks[p, 2] - e[2]) <= kat["tol"]


@pytest.mark.parametrize("case", load_json("list_kats.json"), ids=lambda c: c["source"].split()[0])
def test_list_kats(case):
    lst, _ = O.list_script(case["maxSize"], case["kernelWidth"], case["ops"])
    assert [list(map(float, e)) for e in lst] == [list(map(float, e)) for e in case["expect"]]


def test_list_against_set_semantics_fuzz():
    """Random insert sequences: the list always holds at most P entries, ascending,
    and no two entries within the suppression distance."""
    rng = np.random.default_rng(11)
    for _ in range(200):
        P, k = int(rng.integers(1, 6)), int(rng.integers(2, 6))
        ops = [[0, float(rng.integers(0, 50)), int(rng.integers(0, 20)), int(rng.integers(0, 20))]
               for _ in range(int(rng.integers(1, 40)))]
        lst, _ = O.list_script(P, k, ops)
        assert len(lst) <= P
        vals = [e[0] for e in lst]
        assert vals == sorted(vals)
        for a in range(len(lst)):
            for b in range(a + 1, len(lst)):
                d2 = (lst[a][1] - lst[b][1]) ** 2 + (lst[a][2] - lst[b][2]) ** 2
                assert d2 >= (k - 1) * k


# --------------------------------------------------------------------------
# Sub-pixel fits
# --------------------------------------------------------------------------

GAUSS_CASES = load_json("gauss1d_cases.json")


@pytest.mark.parametrize("case", GAUSS_CASES, ids=lambda c: c["file"][28:-9] + "." + c["test"])
def test_reference_gaussian_cases(case):
    """The 567 Gaussian1DInterpolatorAdvanced* cases: noise-free Gaussians, the fit
    must recover the centre within the reference's tolerance (8e-3 px)."""
    dim = case["dim"]
    plane = gaussian_plane(dim, case["uX"], case["sigmaX"], case["uY"], case["sigmaY"])
    peak = np.array([dim // 2, dim // 2, 1000.0], np.float32)
    if case["kind"] == "guo":
        res = O.guo(plane, case["pixels"], 20, 0.0, peak)
    else:
        res = O.gauss1d(plane, case["pixels"], 0.0, peak)
    assert abs(res[0] - (case["uY"] + dim // 2)) <= case["tol"]
    assert abs(res[1] - (case["uX"] + dim // 2)) <= case["tol"]


def test_gauss3_closed_form_exact():
    """Three samples of an exact Gaussian: the interpolant recovers the centre to
    double precision (the LM fitter converges to the same interpolant)."""
    for mu in (-0.45, -0.2, 0.0, 0.13, 0.49):
        for s in (0.6, 1.0, 2.5):
            x = np.arange(29, 34, dtype=np.float64)
            y = (5e4 * np.exp(-(x - 31 - mu) ** 2 / (2 * s * s))).astype(np.float32)
            plane = np.ones((63, 63), np.float32)
            plane[29:34, 31] = y
            plane[31, 29:34] = y
            r = O.gauss1d(plane, 3, 0.0, np.array([31, 31, y[2]], np.float32))
            assert abs(r[0] - (31 + mu)) < 2e-5 and abs(r[1] - (31 + mu)) < 2e-5


def test_guo_exact_on_noise_free_gaussian_and_amplitude():
    for pixels in (3, 5, 7):
        x = np.arange(0, 63, dtype=np.float64)
        mu, s, a = 30.3, 1.3, 7e4
        y = (a * np.exp(-(x - mu) ** 2 / (2 * s * s))).astype(np.float32)
        plane = np.ones((63, 63), np.float32)
        plane[:, 30] = y
        plane[30, :] = y
        r = O.guo(plane, pixels, 20, 0.0, np.array([30, 30, y[30]], np.float32))
        assert abs(r[0] - mu) < 1e-4 and abs(r[1] - mu) < 1e-4
        assert abs(r[2] - a) / a < 1e-4  # the peak value is replaced by the amplitude


def test_guo_rejects_far_fit_and_border():
    plane = np.ones((31, 31), np.float32)
    r = O.guo(plane, 3, 20, 0.0, np.array([0, 15, 1.0], np.float32))
    assert r[3] == 4.0 and r[0] == 0.0  # border: integer peak kept


def test_poly_closed_form():
    y = np.array([np.exp(-(1.3) ** 2 / 2), np.exp(-(0.3) ** 2 / 2), np.exp(-(0.7) ** 2 / 2)]) * 1e4
    plane = np.ones((9, 9), np.float32)
    plane[3:6, 4] = y
    plane[4, 3:6] = y
    r = O.poly(plane, np.array([4, 4, y[1]], np.float32))
    assert abs(r[0] - 4.3) < 1e-4 and abs(r[1] - 4.3) < 1e-4


# --------------------------------------------------------------------------
# Tiles
# --------------------------------------------------------------------------

def test_clip_tile_zero_fill_and_max_clamp():
    rng = np.random.default_rng(1)
    img = rng.integers(0, 12, (40, 50)).astype(np.uint8)
    t, mx = O.clip_tile(img, -3, 45, 16)
    ref = np.zeros((16, 16), np.uint8)
    ref[3:, :5] = img[0:13, 45:50]
    assert np.array_equal(t, ref) and mx == 16.0
    img[5, 5] = 200
    t, mx = O.clip_tile(img, 0, 0, 16)
    assert np.array_equal(t, img[:16, :16]) and mx == 200.0


def test_float_clip_max_semantics():
    """MatrixFloat tiles: zero fill outside the frame, maximum over the tile
    with no >= 16 clamp and 0 -> 1; integer-valued frames give the 8-bit clip's
    tiles."""
    rng = np.random.default_rng(3)
    img = (rng.random((40, 50)) * 3.0).astype(np.float32)
    t, mx = O.clip_tile_f32(img, -3, 45, 16)
    ref = np.zeros((16, 16), np.float32)
    ref[3:, :5] = img[0:13, 45:50]
    assert np.array_equal(t, ref) and mx == np.float32(img[0:13, 45:50].max())
    t, mx = O.clip_tile_f32(img, 100, 100, 16)  # wholly outside: zeros, maximum 1
    assert not t.any() and mx == 1.0
    t, mx = O.clip_tile_f32(-img, 0, 0, 16)  # no value above 0: maximum 1
    assert mx == 1.0
    img8 = rng.integers(0, 256, (40, 50)).astype(np.uint8)
    for top, left in [(0, 0), (-5, 40), (30, -7)]:
        a, ma = O.clip_tile(img8, top, left, 16)
        b, mb = O.clip_tile_f32(img8.astype(np.float32), top, left, 16)
        assert np.array_equal(a.astype(np.float32), b)
        assert mb == max(float(a.max()), 1.0) and ma == max(float(a.max()), 16.0)


def test_float_warp_micro_reads_float_frames():
    """The float-frame Micro warp: on integer-valued frames it equals the 8-bit
    warp (every side); with integer displacements it is the shifted frame
    rounded half up; on a constant frame c it is round(c) everywhere inside;
    its maximum is 65535."""
    rng = np.random.default_rng(4)
    img8 = rng.integers(0, 256, (80, 90)).astype(np.uint8)
    f = img8.astype(np.float32)
    for side in range(4):
        for top, left, du, dv in [(10, 12, 1.3, -2.7), (-4, 70, -0.4, 3.9), (60, -3, 2.5, 0.5)]:
            a = O.warp_micro(img8, top, left, 16, du, dv, side)
            b, m = O.warp_micro_f32(f, top, left, 16, du, dv, side)
            assert np.array_equal(a.astype(np.float32), b) and m == 65535.0
    g = (rng.random((80, 90)) * 300.0).astype(np.float32)
    b, _ = O.warp_micro_f32(g, 20, 30, 16, 3.0, -2.0)
    assert np.array_equal(b, np.floor(g[23:39, 28:44] + np.float32(0.5)))
    c = np.full((80, 90), 37.625, np.float32)
    b, _ = O.warp_micro_f32(c, 20, 30, 16, 1.7, -0.3)
    assert np.all(b == 38.0)


def test_u16_clip_and_warp_follow_the_8bit_rules():
    """10/12/16-bit frames: 8-bit values widened give the 8-bit clip and warp
    (every side); the clip maximum keeps the >= 16 clamp with 16-bit values;
    integer displacements shift the frame; the warp maximum is 2^bits - 1."""
    rng = np.random.default_rng(6)
    img8 = rng.integers(0, 256, (70, 80)).astype(np.uint8)
    w = img8.astype(np.uint16)
    for side in range(4):
        for top, left, du, dv in [(10, 12, 1.3, -2.7), (-4, 60, -0.4, 3.9)]:
            a = O.warp_micro(img8, top, left, 16, du, dv, side)
            for bits in (10, 12, 16):
                b, m = O.warp_micro_u16(w, bits, top, left, 16, du, dv, side)
                assert np.array_equal(a.astype(np.float32), b) and m == (1 << bits) - 1
    for top, left in [(0, 0), (-5, 70), (60, -7)]:
        a, ma = O.clip_tile(img8, top, left, 16)
        b, mb = O.clip_tile_u16(w, top, left, 16)
        assert np.array_equal(a.astype(np.float32), b) and ma == mb
    big = rng.integers(0, 65536, (70, 80)).astype(np.uint16)
    t, m = O.clip_tile_u16(big, 3, 4, 16)
    assert np.array_equal(t, big[3:19, 4:20].astype(np.float32)) and m == float(big[3:19, 4:20].max())
    t, m = O.clip_tile_u16(np.full((70, 80), 9, np.uint16), 3, 4, 16)
    assert m == 16.0
    t, _ = O.warp_micro_u16(big, 16, 20, 30, 16, 3.0, -2.0)
    assert np.array_equal(t, big[23:39, 28:44].astype(np.float32))


def test_float_full_and_mini_warps_follow_the_8bit_ones():
    """Full and Mini warps over float and 16-bit frames: integer-valued frames
    give the 8-bit warps' samples on every side."""
    rng = np.random.default_rng(8)
    img8 = rng.integers(0, 256, (64, 80)).astype(np.uint8)
    fu = rng.uniform(-3, 3, (64, 80)).astype(np.float32)
    fv = rng.uniform(-3, 3, (64, 80)).astype(np.float32)
    nI, nJ, N = 3, 4, 16
    top = (np.arange(nI)[:, None] * 16 + np.zeros((1, nJ), int)).astype(np.int32)
    left = (np.zeros((nI, 1), int) + np.arange(nJ)[None, :] * 16).astype(np.int32)
    u = rng.uniform(-2, 2, (nI, nJ)).astype(np.float32)
    v = rng.uniform(-2, 2, (nI, nJ)).astype(np.float32)
    for side in range(4):
        a = O.warp_image(img8, fu, fv, side).astype(np.float32)
        assert np.array_equal(a, O.warp_image_f32(img8.astype(np.float32), fu, fv, side))
        for ti, tj in [(0, 0), (1, 2), (2, 3)]:
            m8 = O.warp_mini(img8, top, left, N, u, v, ti, tj, side).astype(np.float32)
            assert np.array_equal(m8, O.warp_mini_frame(img8.astype(np.float32), top, left, N, u, v,
                                                         ti, tj, side))
            assert np.array_equal(m8, O.warp_mini_frame(img8.astype(np.uint16), top, left, N, u, v,
                                                         ti, tj, side))


def test_warp_micro_invariants():
    rng = np.random.default_rng(2)
    img = rng.integers(0, 256, (96, 96)).astype(np.uint8)
    # zero displacement: identity; integer displacement: pure shift
    assert np.array_equal(O.warp_micro(img, 20, 30, 16, 0.0, 0.0), img[20:36, 30:46])
    assert np.array_equal(O.warp_micro(img, 20, 30, 16, 3.0, -2.0), img[23:39, 28:44])
    # constant fraction: separable bilinear with round-half-up, away from borders
    du, dv = 2.25, 1.5
    got = O.warp_micro(img, 20, 30, 16, du, dv).astype(np.float64)
    f = img.astype(np.float64)
    # frac 0.25 rounds to p = i (delta +0.25), frac 0.5 rounds up (delta -0.5)
    ref = np.zeros((16, 16))
    for i in range(16):
        for j in range(16):
            si, sj = 22 + i + 0.25, 31 + j + 0.5
            pi, pj = int(np.floor(si + 0.5)), int(np.floor(sj + 0.5))
            di, dj = si - pi, sj - pj
            oi, oj = (1 if di >= 0 else -1), (1 if dj >= 0 else -1)
            wi, wj = abs(di), abs(dj)
            v = ((1 - wi) * (1 - wj) * f[pi, pj] + (1 - wi) * wj * f[pi, pj + oj] +
                 wi * (1 - wj) * f[pi + oi, pj] + wi * wj * f[pi + oi, pj + oj])
            ref[i, j] = np.floor(v + 0.5)
    assert np.max(np.abs(got - ref)) <= 1.0  # float vs double rounding at x.5 only
    assert np.mean(got == ref) > 0.98


# --------------------------------------------------------------------------
# FindMaximumSimple and the per-level batch
# --------------------------------------------------------------------------

def simple_brute(plane):
    """FindMaximumSimple.findMaximum :19-48 stated on the array: the first cell in
    row-major order holding the largest value above 0 ((D/2, D/2, 0) without
    one); minFloor the smallest non-NaN value (Float.MAX_VALUE without one)."""
    flat = plane.ravel()
    pos = np.where(flat > 0, flat, -np.inf)
    if not np.any(flat > 0):
        pi, pj, pv = plane.shape[0] // 2, plane.shape[1] // 2, 0.0
    else:
        idx = int(np.argmax(pos))  # first occurrence of the maximum
        pi, pj, pv = idx // plane.shape[1], idx % plane.shape[1], flat[idx]
    finite = flat[~np.isnan(flat)]
    mn = finite.min() if finite.size else np.float32(3.4028235e38)
    return pi, pj, np.float32(pv), np.float32(mn)


def test_find_max_simple_brute_force():
    rng = np.random.default_rng(5)
    planes = []
    for D in (3, 10, 31, 63):
        p = rng.random((12, D, D)).astype(np.float32) * 100
        p[0] = np.round(p[0] / 30)            # ties: first in row-major order wins
        p[1] = -1.0                           # nothing positive
        p[2] = 0.0
        p[3, 0, 0] = 1e4                      # row 0 / column 0 are visited
        p[4, D - 1, D - 1] = 1e4
        p[5, :, :] = np.nan
        p[6, 1, 1] = np.nan
        p[7, D // 2, 0] = p[7, D // 2, D - 1] = 500
        p[8] = -np.abs(p[8])
        p[8, D - 1, 0] = 1e-30               # a tiny positive maximum
        p[9, 2 % D, 1] = -0.0
        planes += list(p)
    for p in planes:
        r = O.find_max_simple(p)
        pi, pj, pv, mn = simple_brute(p)
        assert (r[0], r[1], r[2]) == (pi, pj, pv)
        assert r[3] == mn


def test_level_batch_composes_the_restated_steps():
    """or_level_batch equals the per-IA oracle functions called one by one, for
    every warp, mode, peak finder and fit (the full-size GPU tests use it)."""
    from qlpiv.geometry import level_grid
    from qlpiv.synth import particle_pair

    a, b, _ = particle_pair(128, 128, seed=3, shift=(1.2, -2.1))
    N = 16
    tops, lefts = level_grid(128, 128, N, 0.5)
    top, left = tops.ravel()[::5].astype(np.int32), lefts.ravel()[::5].astype(np.int32)
    n = top.size
    rng = np.random.default_rng(0)
    u = rng.uniform(-3, 3, n).astype(np.float32)
    v = rng.uniform(-3, 3, n).astype(np.float32)
    D = 2 * N - 1
    for warp in (0, 1, 2):
        for mode in (0, 1):
            for peak_kind, P in ((0, 1), (1, 3), (2, 1)):
                for fit in (-1, 0, 1, 2):
                    out, cnt, pl = O.level_batch(a, b, top, left, u, v, N, warp, mode, peak_kind,
                                                 P, 3, fit, 3, 20, planes=True)
                    for t in range(n):
                        tA, mA = O.clip_tile(a, top[t], left[t], N)
                        if warp == 2:
                            tB, mB = O.warp_micro(b, top[t], left[t], N, u[t], v[t]), 255.0
                        elif warp == 1:
                            tt = int(np.floor(np.float32(top[t]) + u[t] + np.float32(0.5)))
                            ll = int(np.floor(np.float32(left[t]) + v[t] + np.float32(0.5)))
                            tB, mB = O.clip_tile(b, tt, ll, N)
                        else:
                            tB, mB = O.clip_tile(b, top[t], left[t], N)
                        A, B = tA.astype(np.float32), tB.astype(np.float32)
                        ref = O.xcorr_javagpu(A, B, mA, mB) if mode == 0 else O.xcorr_javacpu(A, B, mA)
                        assert np.array_equal(pl[t], ref)
                        if peak_kind == 1:
                            pk, found, mf = O.find_max_multipeaks(ref, P, 3)
                        else:
                            r = O.find_max_simple(ref) if peak_kind == 2 else O.find_max_from_center(ref)
                            pk, found, mf = r[None, :3], 1, r[3]
                        assert cnt[t] == max(found, 1)
                        for q in range(out.shape[1]):
                            if q >= found:
                                assert not out[t, q].any()
                                continue
                            if fit == 0:
                                e = O.gauss1d(ref, 3, mf, pk[q])
                            elif fit == 1:
                                e = O.guo(ref, 3, 20, mf, pk[q])
                            elif fit == 2:
                                e = O.poly(ref, pk[q])
                            else:
                                e = np.array([*pk[q], 4.0], np.float32)
                            assert np.array_equal(out[t, q], e), (warp, mode, peak_kind, fit, t, q)
