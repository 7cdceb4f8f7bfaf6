"""Pins for the oracle's sub-pixel fits beyond the noise-free cases.

* Guo's SVD fallback (Gaussian1DHongweiGuoInterpolator.estimateGaussianParameters
  :113-122, commons-math3 SingularValueDecomposition solver): the oracle's
  pseudo-inverse solve equals the minimum-norm least-squares solution (numpy)
  on well-conditioned, small-valued and exactly rank-deficient normal matrices,
  and a Guo fit whose normal matrix the LU rejects still recovers the Gaussian.
* Guo's iteration (initMatrices :47-70, updateMatrices :72-102) on noisy 5-15
  point data: one iteration is the weighted least squares with the data weights
  y^2 (numpy lstsq), and the converged coefficients satisfy the fixed point of
  the re-weighting, X^T diag(yhat^2) (ln y - X c) = 0 with yhat = exp(X c).  A
  rule that kept the data weights fails the fixed point.
* Gaussian1D 5/7-point fits (Gaussian1DInterpolator.java:78-94, commons-math3
  GaussianCurveFitter): the oracle's optimum equals scipy.optimize.least_squares
  on noisy samples within 1e-6 px, through the interpolator itself too.
"""
import numpy as np
import pytest
from scipy.optimize import least_squares

import oracle as O

EPS = np.finfo(np.float64).eps


def commons_pinv(A):
    """The pseudo-inverse with commons-math3 3.6.1's SVD threshold
    max(m * s_max * 2^-52, sqrt(Double.MIN_NORMAL)), m = 3."""
    u, s, vt = np.linalg.svd(A)
    tol = max(3 * s[0] * EPS, np.sqrt(np.finfo(np.float64).tiny))
    inv = np.where(s > tol, 1.0 / np.where(s > tol, s, 1.0), 0.0)
    return (vt.T * inv) @ u.T


def normal_matrix(x, w):
    X = np.stack([np.ones_like(x), x, x * x], 1)
    return X.T @ (w[:, None] * X)


# ---------------------------------------------------------------------------
# pseudo-inverse solve (the SVD fallback)
# ---------------------------------------------------------------------------

def test_pinv_equals_solve_on_well_conditioned_matrices():
    rng = np.random.default_rng(1)
    for _ in range(200):
        n = int(rng.integers(3, 16))
        x = np.arange(n, dtype=np.float64) + float(rng.integers(0, 40))
        A = normal_matrix(x, rng.uniform(0.5, 2.0, n))
        b = rng.normal(size=3) * np.abs(A).max()
        ref = np.linalg.solve(A, b)
        got = O.pinv_solve3(A, b)
        assert np.allclose(got, ref, rtol=1e-6, atol=1e-9 * np.abs(ref).max()), (got, ref)


def test_pinv_small_valued_matrix_reaches_fallback_and_solves():
    """Weights of 1e-16: every entry of the normal matrix is below the LU's
    absolute singularity threshold 1e-11, so Guo takes the SVD path; the matrix
    is still full rank and the pseudo-inverse is the inverse."""
    x = np.arange(5, dtype=np.float64)
    A = normal_matrix(x, np.full(5, 1e-16))
    assert O.lu_singular3(A)
    b = A @ np.array([0.7, -1.3, 0.25])
    got = O.pinv_solve3(A, b)
    assert np.allclose(got, [0.7, -1.3, 0.25], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("vectors", [
    [(1.0, 2.0, 3.0)],
    [(1.0, 1.0, 1.0)],
    [(1.0, 2.0, 3.0), (2.0, 0.0, 1.0)],
    [(0.0, 1.0, 4.0), (1.0, -1.0, 2.0)],
])
def test_pinv_rank_deficient_is_minimum_norm(vectors):
    """Exactly rank-1 / rank-2 integer matrices: consistent and inconsistent
    right-hand sides give the minimum-norm least-squares solution A+ b."""
    A = sum(np.outer(v, v) for v in np.array(vectors))
    assert O.lu_singular3(A)
    rng = np.random.default_rng(len(vectors))
    for _ in range(20):
        for b in (A @ rng.normal(size=3), rng.normal(size=3)):
            ref = commons_pinv(A) @ b
            ref2 = np.linalg.lstsq(A, b, rcond=1e-12)[0]
            assert np.allclose(ref, ref2, atol=1e-12)
            got = O.pinv_solve3(A, b)
            assert np.allclose(got, ref, rtol=1e-9, atol=1e-12), (got, ref)


def test_pinv_zero_matrix():
    assert np.array_equal(O.pinv_solve3(np.zeros((3, 3)), np.ones(3)), np.zeros(3))


def test_guo_plane_through_svd_fallback_recovers_gaussian():
    """A peak of amplitude 1e-7 (normal-matrix entries far below 1e-11): the fit
    goes through the pseudo-inverse on every iteration and still recovers the
    noise-free Gaussian's centre, as the LU path does at full amplitude."""
    for pixels in (3, 5, 7):
        xs = np.arange(28, 28 + pixels, dtype=np.float32)
        mu, s = 28 + pixels // 2 + 0.31, 1.4
        for amp in (1e-7, 7e4):
            ys = (amp * np.exp(-(xs.astype(np.float64) - mu) ** 2 / (2 * s * s))).astype(np.float32)
            w = (ys * ys).astype(np.float64)
            assert O.lu_singular3(normal_matrix(xs.astype(np.float64), w)) == (amp < 1)
            c = O.guo_coefs(xs, ys, 20)
            assert abs(-c[1] / (2 * c[2]) - mu) < 1e-4, (pixels, amp)


# ---------------------------------------------------------------------------
# Guo's iteratively re-weighted least squares on noisy data
# ---------------------------------------------------------------------------

def noisy_gaussian(rng, n, noise):
    x0 = float(rng.integers(0, 30))
    xs = (np.arange(n) + x0).astype(np.float32)
    mu = x0 + n // 2 + rng.uniform(-0.5, 0.5)
    s = rng.uniform(0.8, 2.5)
    y = 5e4 * np.exp(-(xs.astype(np.float64) - mu) ** 2 / (2 * s * s))
    y *= 1.0 + noise * rng.normal(size=n)
    return xs, np.maximum(y, 50.0).astype(np.float32)


def design(xs):
    x = xs.astype(np.float64)
    x2 = (xs * xs).astype(np.float64)  # the reference's float product
    return np.stack([np.ones_like(x), x, x2], 1)


@pytest.mark.parametrize("n", [5, 7, 9, 15])
def test_guo_first_iteration_is_data_weighted_least_squares(n):
    rng = np.random.default_rng(n)
    for _ in range(40):
        xs, ys = noisy_gaussian(rng, n, 0.05)
        X = design(xs)
        w = (ys * ys).astype(np.float64)  # float product, as the reference
        sw = np.sqrt(w)
        ref = np.linalg.lstsq(sw[:, None] * X, sw * np.log(ys.astype(np.float64)), rcond=None)[0]
        got = O.guo_coefs(xs, ys, 1)
        assert abs(-got[1] / (2 * got[2]) - (-ref[1] / (2 * ref[2]))) < 1e-7
        assert np.allclose(got, ref, rtol=1e-6, atol=1e-9 * np.abs(ref).max())


@pytest.mark.parametrize("n", [5, 7, 11, 15])
def test_guo_converges_to_reweighting_fixed_point(n):
    """After the reference's 20 iterations the re-weighting has converged (the
    centre moves < 1e-8 px by 60 iterations) and the normal equations with the
    fitted weights hold to double rounding (measured: <= 3.3e-13 of the summed
    term magnitudes); the data-weighted solution misses them by >= 4e-7."""
    rng = np.random.default_rng(100 + n)
    for _ in range(40):
        xs, ys = noisy_gaussian(rng, n, 0.08)
        X = design(xs)
        lny = np.log(ys.astype(np.float64))

        def residual(c):
            yh = np.exp(X @ c)
            t = (X * (yh * yh * (lny - X @ c))[:, None]).sum(0)
            return np.max(np.abs(t) / np.abs(X * (yh * yh * lny)[:, None]).sum(0))

        c20, c60 = O.guo_coefs(xs, ys, 20), O.guo_coefs(xs, ys, 60)
        assert abs(-c20[1] / (2 * c20[2]) + c60[1] / (2 * c60[2])) < 1e-8
        assert residual(c20) <= 1e-11
        assert residual(O.guo_coefs(xs, ys, 1)) > 1e-8


def test_guo_three_samples_iterations_reproduce_the_first():
    """With 3 samples the weighted normal equations are square: the first solve
    is the interpolant, and the 20 iterations of the reference reproduce it up
    to FP64 rounding (the device stops after the first successful solve for
    3-pixel fits).  The fitted centre agrees to 1e-5 px, the amplitude to a
    relative 1e-6; a 5-sample fit does move between its iterations (the
    shortcut is specific to 3 samples)."""
    rng = np.random.default_rng(33)
    for _ in range(300):
        xs, ys = noisy_gaussian(rng, 3, 0.05)
        c1, c20 = O.guo_coefs(xs, ys, 1), O.guo_coefs(xs, ys, 20)
        m1, m20 = -c1[1] / (2 * c1[2]), -c20[1] / (2 * c20[2])
        assert abs(np.float32(m1) - np.float32(m20)) <= 1e-5
        a1 = np.exp(c1[0] - c1[1] ** 2 / (4 * c1[2]))
        a20 = np.exp(c20[0] - c20[1] ** 2 / (4 * c20[2]))
        assert abs(a1 - a20) <= 1e-6 * abs(a20)
    moved = 0
    for _ in range(100):
        xs, ys = noisy_gaussian(rng, 5, 0.08)
        c1, c20 = O.guo_coefs(xs, ys, 1), O.guo_coefs(xs, ys, 20)
        moved += abs(c1[1] / (2 * c1[2]) - c20[1] / (2 * c20[2])) > 1e-6
    assert moved > 50


def test_guo_interpolator_uses_the_iterated_coefficients():
    """or_guo's mean and amplitude come from the coefficients above."""
    rng = np.random.default_rng(7)
    for pixels in (5, 7):
        plane = np.full((63, 63), 40.0, np.float32)
        xs, ys = noisy_gaussian(rng, pixels, 0.05)
        c = pixels // 2
        xs = np.arange(31 - c, 32 + c, dtype=np.float32)
        plane[31 - c:32 + c, 31] = ys
        plane[31, 31 - c:32 + c] = ys
        r = O.guo(plane, pixels, 20, 0.0, np.array([31, 31, ys[c]], np.float32))
        k = O.guo_coefs(xs, ys, 20)
        mean = np.float32(-k[1] / (2 * k[2]))
        assert r[0] == mean and r[1] == mean
        assert r[2] == np.float32(np.exp(k[0] - k[1] * k[1] / (4 * k[2])))


# ---------------------------------------------------------------------------
# Gaussian1D least-squares fit (5 / 7 points)
# ---------------------------------------------------------------------------

def scipy_gauss(x, y):
    def res(p):
        return p[0] * np.exp(-(x - p[1]) ** 2 / (2 * p[2] ** 2)) - y
    best = None
    c = len(x) // 2
    for b0 in (x[c], x[c] - 0.4, x[c] + 0.4):
        for s0 in (0.8, 1.5, 2.5):
            r = least_squares(res, [y.max(), b0, s0], method="lm", xtol=1e-15, ftol=1e-15,
                              gtol=1e-15, max_nfev=20000)
            if best is None or r.cost < best.cost:
                best = r
    return best.x, best.cost


@pytest.mark.parametrize("n", [5, 7])
def test_gauss_lsq_equals_scipy_optimum_on_noisy_samples(n):
    rng = np.random.default_rng(200 + n)
    for _ in range(60):
        x = np.arange(n, dtype=np.float64) + 20.0
        mu = 20 + n // 2 + rng.uniform(-0.5, 0.5)
        s = rng.uniform(0.7, 2.5)
        y = 3e4 * np.exp(-(x - mu) ** 2 / (2 * s * s)) * (1 + 0.05 * rng.normal(size=n)) + 30
        ref, ref_cost = scipy_gauss(x, y)
        got = O.gauss_lsq(x, y)
        cost = 0.5 * np.sum((got[0] * np.exp(-(x - got[1]) ** 2 / (2 * got[2] ** 2)) - y) ** 2)
        assert cost <= ref_cost * (1 + 1e-9) + 1e-12
        assert abs(got[1] - ref[1]) <= 1e-6, (got, ref)


@pytest.mark.parametrize("pixels", [5, 7])
def test_gauss1d_interpolator_noisy_planes_vs_scipy(pixels):
    rng = np.random.default_rng(300 + pixels)
    c = pixels // 2
    for _ in range(30):
        plane = np.full((63, 63), 25.0, np.float32)
        mu_i, mu_j = 31 + rng.uniform(-0.5, 0.5), 31 + rng.uniform(-0.5, 0.5)
        s = rng.uniform(0.8, 2.0)
        idx = np.arange(31 - c, 32 + c, dtype=np.float64)
        col = 4e4 * np.exp(-(idx - mu_i) ** 2 / (2 * s * s)) * (1 + 0.03 * rng.normal(size=pixels))
        row = 4e4 * np.exp(-(idx - mu_j) ** 2 / (2 * s * s)) * (1 + 0.03 * rng.normal(size=pixels))
        plane[31 - c:32 + c, 31] = col.astype(np.float32)
        plane[31, 31 - c:32 + c] = row.astype(np.float32)
        plane[31, 31] = max(plane[31, 31], plane[30, 31], plane[31, 30]) + 1.0
        r = O.gauss1d(plane, pixels, 0.0, np.array([31, 31, plane[31, 31]], np.float32))
        ri, _ = scipy_gauss(idx, plane[31 - c:32 + c, 31].astype(np.float64))
        rj, _ = scipy_gauss(idx, plane[31, 31 - c:32 + c].astype(np.float64))
        assert r[3] == 3.0
        assert abs(r[0] - ri[1]) <= 1e-5 and abs(r[1] - rj[1]) <= 1e-5
