"""Pin the oracle's restatements of the velocity inheritance strategies
(iareas/AdaptiveInterVelocityInheritanceStrategy{Area,DistanceGeneric,BiCubicSpline}.java).

Area and Distance are pinned by the reference's own unit tests, restated as
data: AdaptiveInterAreaVelocityInheritanceByAreaCompleteTests (every child of an
8 x 8 level, its contributing parents and overlap areas) and
AdaptiveInterAreaVelocityInheritanceByDistanceGenericTests (the contributing
parents of every child).  BiCubicSpline follows commons-math3 3.6.1's
PiecewiseBicubicSplineInterpolator (Akima splines over 5 x 5 windows), a
library absent here: it is pinned by what the mathematics fixes (Akima
reproduces quadratics on uniform knots, so biquadratic fields come back
exactly; interpolation at the knots; clamping outside the knot range; the
reference's zero result below 5 knots).  The choice of the 5-point window
away from the knots is parity unpinned.
"""
import numpy as np
import pytest

import oracle as O
from qlpiv.geometry import level_grid

# the parent displacements both reference tests set (stepTiles1Array[i][j])
PARENT_UV = np.array([[(-11, -11), (-11, 1), (-11, 11)],
                      [(1, -11), (1, 1), (1, 11)],
                      [(11, -11), (11, 1), (11, 11)]], np.float32)


def test_area_complete_kat():
    """AdaptiveInterAreaVelocityInheritanceByAreaCompleteTests :31-210: super
    position, 171 x 171 image, margins 32, overlap factor 1/3, 64 -> 32."""
    of = np.float32(1.0) / np.float32(3.0)
    m = (32, 32, 32, 32)
    pt, pl = level_grid(171, 171, 64, of, m, level0=True)
    ct, cl = level_grid(171, 171, 32, of, m, level0=False)
    assert pt.shape == (3, 3) and ct.shape == (8, 8)
    u, v, rc = O.inherit("area", (pt, pl, 64), (ct, cl, 32), PARENT_UV[..., 0], PARENT_UV[..., 1])
    assert rc == 0
    idx = [[0, 1], [0, 1], [0, 1, 2], [0, 1, 2], [0, 1, 2], [0, 1, 2], [1, 2], [1, 2]]
    areas = [[32, 11], [32, 22], [32, 32, 10], [32, 32, 21], [21, 32, 32], [11, 32, 32],
             [21, 32], [10, 32]]
    for i in range(8):
        for j in range(8):
            tiles, ar = [], []
            for a, ai in zip(idx[i], areas[i]):
                for b, bj in zip(idx[j], areas[j]):
                    tiles.append((a, b))
                    ar.append(np.float32(ai * bj))
            total = np.float32(0)
            for x in ar:
                total = np.float32(total + x)
            eu = ev = np.float32(0)
            for (a, b), x in zip(tiles, ar):
                eu = np.float32(eu + np.float32(x / total) * PARENT_UV[a, b, 0])
                ev = np.float32(ev + np.float32(x / total) * PARENT_UV[a, b, 1])
            assert abs(u[i, j] - eu) <= 1e-4 and abs(v[i, j] - ev) <= 1e-4, (i, j)


def _distance_expected(ctop, cleft, ptop, pleft, parents):
    """computeDisplacements of the reference test (:200-235): inverse-distance
    weights between the child centre and each listed parent centre."""
    cc = (np.float32(ctop + 16), np.float32(cleft + 16))
    d = []
    for a, b in parents:
        pc = (np.float32(ptop[a] + 32), np.float32(pleft[b] + 32))
        d.append(np.float32(np.sqrt(float(cc[1] - pc[1]) ** 2 + float(cc[0] - pc[0]) ** 2)))
    tot = np.float32(0)
    zero = -1
    for k, x in enumerate(d):
        if x > 0:
            tot = np.float32(tot + np.float32(1) / x)
        else:
            zero = k
    eu = ev = np.float32(0)
    for k, (a, b) in enumerate(parents):
        w = (np.float32(np.float32(1) / d[k]) / tot) if zero < 0 else np.float32(k == zero)
        eu = np.float32(eu + np.float32(w * PARENT_UV[a, b, 0]))
        ev = np.float32(ev + np.float32(w * PARENT_UV[a, b, 1]))
    return eu, ev


def test_distance_generic_kat():
    """AdaptiveInterAreaVelocityInheritanceByDistanceGenericTests :24-198: mixed
    super position, parents of 64 px at 16/37/59, children of 32 px at
    16/27/37/48/59/69/80/91; the test's table of contributing parents."""
    p1 = np.array([16, 37, 59])
    c1 = np.array([16, 27, 37, 48, 59, 69, 80, 91])
    pt, pl = np.meshgrid(p1, p1, indexing="ij")
    ct, cl = np.meshgrid(c1, c1, indexing="ij")
    u, v, rc = O.inherit("distance", (pt, pl, 64), (ct, cl, 32), PARENT_UV[..., 0],
                         PARENT_UV[..., 1])
    assert rc == 0
    sol = [[0, -1], [0, -1], [0, 1], [0, 1], [1, 2], [1, 2], [2, -1], [2, -1]]
    for i in range(8):
        for j in range(8):
            parents = [(a, b) for a in sol[i] if a >= 0 for b in sol[j] if b >= 0]
            eu, ev = _distance_expected(c1[i], c1[j], p1, p1, parents)
            assert abs(u[i, j] - eu) <= 1e-6 and abs(v[i, j] - ev) <= 1e-6, (i, j, u[i, j], eu)


@pytest.mark.parametrize("kind", ["area", "distance"])
def test_stencil_strategies_on_superposition_levels(kind):
    """C2-like levels (2048^2, 64 -> 32 -> 16, 50 %): the walk never hits the
    reference's IllegalStateException, weights sum to one (a constant field
    stays constant) and rounding gives integers."""
    grids = [(64, *level_grid(2048, 2048, 64, 0.5, level0=True)),
             (32, *level_grid(2048, 2048, 32, 0.5, level0=False)),
             (16, *level_grid(2048, 2048, 16, 0.5, level0=False))]
    for (ps, pt, pl), (cs, ct, cl) in zip(grids[:-1], grids[1:]):
        pu = np.full(pt.shape, 2.5, np.float32)
        pv = np.full(pt.shape, -1.25, np.float32)
        u, v, rc = O.inherit(kind, (pt, pl, ps), (ct, cl, cs), pu, pv)
        assert rc == 0
        assert np.abs(u - 2.5).max() <= 1e-6 and np.abs(v + 1.25).max() <= 1e-6
        ur, vr, rc = O.inherit(kind, (pt, pl, ps), (ct, cl, cs), pu, pv, round_it=True)
        assert rc == 0 and (ur == 3.0).all() and (vr == -1.0).all()


def _centres(tops, size):
    return tops.astype(np.float64) + size / 2.0 - 0.5


def test_bicubic_reproduces_biquadratic_fields():
    pt, pl = level_grid(1024, 1024, 64, 0.5, level0=True)
    ct, cl = level_grid(1024, 1024, 32, 0.5, level0=False)
    py, px = _centres(pt, 64), _centres(pl, 64)
    f = lambda y, x: 1e-5 * (y - 400) ** 2 - 2e-5 * (x - 500) ** 2 + 3e-3 * x - 2e-3 * y + 1.5
    g = lambda y, x: -2e-5 * (y - 300) ** 2 + 1e-5 * (x - 600) ** 2 - 1e-3 * x + 0.25
    pu = f(py, px).astype(np.float32)
    pv = g(py, px).astype(np.float32)
    u, v, _ = O.inherit("bicubic", (pt, pl, 64), (ct, cl, 32), pu, pv)
    cy = np.clip(_centres(ct, 32), py.min(), py.max())
    cx = np.clip(_centres(cl, 32), px.min(), px.max())
    # the samples are float32: compare with the biquadratic through them
    # within float32 rounding of the inputs and the output
    assert np.abs(u - f(cy, cx)).max() < 2e-5
    assert np.abs(v - g(cy, cx)).max() < 2e-5


def test_bicubic_interpolates_at_knots_and_clamps():
    rng = np.random.default_rng(3)
    xs = np.arange(7) * 16.0 + 31.5
    ys = np.arange(6) * 16.0 + 31.5
    f = rng.normal(0, 2, (6, 7))
    for i in range(6):
        for j in range(7):
            # exact where the piece starts at the knot; the last knot is the far
            # end of the last piece (Horner over the full interval)
            got = O.piecewise_bicubic(ys, xs, f, ys[i], xs[j])
            assert abs(got - f[i, j]) <= 1e-13, (i, j)
    # clamping: a child centre outside the parent centres takes the edge value
    pt = (np.arange(6) * 16)[:, None].repeat(7, 1) + 0
    pl = (np.arange(7) * 16)[None, :].repeat(6, 0) + 0
    ct = np.array([[-40, 200]])
    cl = np.array([[-40, 300]])
    fu = f.astype(np.float32)
    u, _, _ = O.inherit("bicubic", (pt, pl, 32), (ct, cl, 16), fu, fu)
    assert u[0, 0] == fu[0, 0] and abs(u[0, 1] - fu[5, 6]) <= 1e-6


def test_bicubic_below_five_knots_keeps_zero():
    pt, pl = np.meshgrid(np.arange(4) * 16, np.arange(9) * 16, indexing="ij")
    ct, cl = np.meshgrid(np.arange(8) * 8, np.arange(18) * 8, indexing="ij")
    pu = np.ones(pt.shape, np.float32)
    u, v, _ = O.inherit("bicubic", (pt, pl, 32), (ct, cl, 16), pu, pu)
    assert not u.any() and not v.any()
