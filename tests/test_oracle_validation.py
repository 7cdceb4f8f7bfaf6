"""Pin the oracle's restatement of the vector validation job
(jobs/validator/VectorValidatorJob.java and the validators/replacements it
drives).  The reference has no unit test for this job, so the pins are
independent specifications (the spiral as a geometric ring walk), hand-worked
known answers that separate the reference's index arithmetic from the
textbook median, and invariants (linear fields are reproduced exactly by the
bilinear replacement, smooth fields are left untouched).
"""
import numpy as np
import pytest

import oracle as O

MED, DIFF = 0, 1
R_NONE, R_SEC, R_BIL = 0, 1, 2


def grid(nI, nJ, step=8, top0=3, left0=5):
    top = (top0 + step * np.arange(nI))[:, None].repeat(nJ, 1).astype(np.int32)
    left = (left0 + step * np.arange(nJ))[None, :].repeat(nI, 0).astype(np.int32)
    return top, left


def ring_walk(nI, nJ):
    """The visit order described geometrically: rings around the centre cell
    (odd size) or the centre pair (even size), each walked clockwise from its
    top-left corner (right along the top row, down the right column, left
    along the bottom row, up the left column), cells outside the grid skipped."""
    ci0 = nI // 2 if nI % 2 else nI // 2 - 1
    cj0 = nJ // 2 if nJ % 2 else nJ // 2 - 1
    hI = 0 if nI % 2 else 1
    hJ = 0 if nJ % 2 else 1
    out, seen = [], set()
    r = 0
    while len(out) < nI * nJ:
        t, b = ci0 - r, ci0 + hI + r
        lft, rgt = cj0 - r, cj0 + hJ + r
        cells = [(t, j) for j in range(lft, rgt + 1)]
        cells += [(i, rgt) for i in range(t + 1, b + 1)]
        if b > t:
            cells += [(b, j) for j in range(rgt - 1, lft - 1, -1)]
            cells += [(i, lft) for i in range(b - 1, t, -1)]
        for i, j in cells:
            if 0 <= i < nI and 0 <= j < nJ and (i, j) not in seen:
                seen.add((i, j))
                out.append(i * nJ + j)
        r += 1
    return np.array(out, np.int32)


SHAPES = [(a, b) for a in range(1, 10) for b in range(1, 10)] + [(63, 63), (127, 127), (31, 7),
                                                                 (8, 127), (255, 255)]


@pytest.mark.parametrize("nI,nJ", SHAPES)
def test_spiral_is_the_clockwise_ring_walk(nI, nJ):
    order = O.spiral_order(nI, nJ)
    assert order is not None, "a tile was visited twice"
    np.testing.assert_array_equal(order, ring_walk(nI, nJ))


def test_uniform_field_is_untouched():
    top, left = grid(6, 7)
    u = np.full((6, 7), 1.25, np.float32)
    v = np.full((6, 7), -2.5, np.float32)
    st = [(MED, 2.0, 0.1, R_BIL, 0)]
    u2, v2, inv, stats = O.validate(top, left, u, v, st)
    np.testing.assert_array_equal(u2, u)
    np.testing.assert_array_equal(v2, v)
    assert not inv.any()
    assert list(stats) == [1, 0, 0, 0]


def _centre_case(c):
    """3x3 grid: eight neighbours u = 0..7 (row-major), v = 0, centre u = c."""
    top, left = grid(3, 3)
    u = np.array([[0, 1, 2], [3, c, 4], [5, 6, 7]], np.float32)
    v = np.zeros((3, 3), np.float32)
    return top, left, u, v


def test_median_known_answer_eight_neighbours():
    """Hand-worked: median of the norm-sorted neighbours = (3 + 4)/2 = 3.5;
    residuals (3.5-k)^2 sorted = .25 .25 2.25 2.25 6.25 6.25 12.25 12.25.
    With 8 neighbours the reference takes rMed = sqrt(rIs[4]) = 2.5 (the
    textbook median would be sqrt((2.25+6.25)/2) = 2.06).  Threshold 2,
    eps 0.1: the centre is valid iff |3.5 - c| < 2 * 2.6 = 5.2.  c = 8.2 is
    valid with the reference's rMed and invalid with the textbook one."""
    st = [(MED, 2.0, 0.1, R_NONE, 0)]
    for c, expect_invalid in [(8.2, False), (8.69, False), (8.71, True), (-1.69, False),
                              (-1.71, True)]:
        top, left, u, v = _centre_case(c)
        _, _, inv, _ = O.validate(top, left, u, v, st)
        assert bool(inv[1, 1]) == expect_invalid, c


def test_median_known_answer_five_neighbours():
    """Tile (0, 1) of a 2x3 grid has 5 neighbours: u = 0, 10, 1, 2, 3 at
    (0,0) (0,2) (1,0) (1,1) (1,2).  Norm order 0 1 2 3 10 -> median 2;
    residuals 4 1 0 1 64 -> sorted 0 1 1 4 64.  With 5 neighbours the reference
    takes rMed = sqrt((rIs[2] + rIs[3]) / 2) = sqrt(2.5) (textbook: rIs[2] = 1).
    Valid iff |2 - c| < 2 * (1.5811 + 0.1) = 3.3623."""
    top, left = grid(2, 3)
    st = [(MED, 2.0, 0.1, R_NONE, 0)]
    for c, expect_invalid in [(5.3, False), (5.37, True), (-1.36, False), (-1.37, True)]:
        u = np.array([[0, c, 10], [1, 2, 3]], np.float32)
        v = np.zeros((2, 3), np.float32)
        _, _, inv, _ = O.validate(top, left, u, v, st)
        assert bool(inv[0, 1]) == expect_invalid, c


def test_median_single_neighbour_uses_unit_rmed():
    """One neighbour: rMed = 1, so valid iff |du| < 2 * 1.1."""
    top, left = grid(1, 2)
    st = [(MED, 2.0, 0.1, R_NONE, 0)]
    for d, expect in [(2.19, False), (2.21, True)]:
        u = np.array([[0.0, d]], np.float32)
        v = np.zeros((1, 2), np.float32)
        _, _, inv, _ = O.validate(top, left, u, v, st)
        assert bool(inv[0, 0]) == expect and bool(inv[0, 1]) == expect


def test_single_tile_is_always_valid():
    top, left = grid(1, 1)
    u = np.array([[1e6]], np.float32)
    _, _, inv, stats = O.validate(top, left, u, u.copy(), [(MED, 2.0, 0.1, R_BIL, 0)])
    assert not inv.any() and stats[2] == 0


def _linear(nI, nJ):
    i, j = np.meshgrid(np.arange(nI), np.arange(nJ), indexing="ij")
    u = (0.5 * i + 0.25 * j + 1.0).astype(np.float32)
    v = (-0.25 * i + 0.75 * j - 2.0).astype(np.float32)
    return u, v


@pytest.mark.parametrize("pos", [(3, 4), (0, 4), (3, 0), (6, 4), (3, 8), (1, 1)])
def test_bilinear_reproduces_a_linear_field(pos):
    """An outlier in a linear field, replaced by the bilinear estimate, gets the
    field's value back exactly (dyadic values: every weighted mean is exact)."""
    nI, nJ = 7, 9
    top, left = grid(nI, nJ)
    u, v = _linear(nI, nJ)
    uo, vo = u.copy(), v.copy()
    uo[pos] += 40.0
    vo[pos] -= 30.0
    u2, v2, inv, stats = O.validate(top, left, uo, vo, [(MED, 2.0, 0.1, R_BIL, 0)])
    assert not inv.any()
    np.testing.assert_array_equal(u2, u)
    np.testing.assert_array_equal(v2, v)
    assert stats[1] == 1


def test_bilinear_corner_takes_the_last_row_with_a_valid_neighbour():
    """Corner (0, 0): no left/right or top/bottom pair and no diagonal pair, so
    the fallback copies a valid neighbour.  The reference's break leaves only
    the inner loop, so the last window row with a valid neighbour wins: the
    tile below (1, 0), not the one to the right (0, 1)."""
    top, left = grid(4, 4)
    u = np.zeros((4, 4), np.float32)
    v = np.zeros((4, 4), np.float32)
    u[0, 1], u[1, 0], u[1, 1] = 0.5, 0.25, 0.375
    u[0, 0] = 50.0
    u2, _, inv, _ = O.validate(top, left, u, v, [(MED, 2.0, 0.1, R_BIL, 0)])
    assert not inv[0, 0]
    assert u2[0, 0] == np.float32(0.25)


def test_secondary_peak_replaces_main_peak_increment():
    """u = u - u_peak0 + u_peak1: an outlier whose second peak holds the true
    increment is restored exactly; peak counts below the index leave it to the
    next stage."""
    nI, nJ = 5, 5
    top, left = grid(nI, nJ)
    base = np.full((nI, nJ), 1.5, np.float32)
    inc = np.zeros((nI, nJ, 3, 2), np.float32)
    inc[..., 0, 0] = 0.25
    u = base + inc[..., 0, 0]
    v = np.full((nI, nJ), -0.5, np.float32)
    inc[2, 2, 0] = (9.0, 7.0)
    inc[2, 2, 1] = (0.25, 0.0)
    u[2, 2] = 1.5 + 9.0
    v[2, 2] = -0.5 + 7.0
    stages = [(MED, 2.0, 0.1, R_SEC, 1), (MED, 2.0, 0.1, R_SEC, 2), (MED, 2.0, 0.1, R_BIL, 0)]
    u2, v2, inv, stats = O.validate(top, left, u, v, stages, peak_uv=inc)
    assert not inv.any()
    assert u2[2, 2] == np.float32(1.75) and v2[2, 2] == np.float32(-0.5)
    # only one peak found: the secondary stages skip, bilinear replaces
    cnt = np.full((nI, nJ), 1, np.int32)
    u3, v3, inv3, _ = O.validate(top, left, u, v, stages, peak_uv=inc, peak_count=cnt)
    assert not inv3.any()
    assert u3[2, 2] == np.float32(1.75) and v3[2, 2] == np.float32(-0.5)


def test_unreplaceable_vector_becomes_zero():
    """Median validation without replacement: the outlier stays invalid and is
    set to (0, 0) at the end of the job."""
    top, left = grid(5, 5)
    u = np.full((5, 5), 2.0, np.float32)
    v = np.full((5, 5), 1.0, np.float32)
    u[2, 3] = 30.0
    u2, v2, inv, stats = O.validate(top, left, u, v, [(MED, 2.0, 0.1, R_NONE, 0)])
    assert inv[2, 3] and inv.sum() == 1
    assert u2[2, 3] == 0.0 and v2[2, 3] == 0.0
    assert list(stats) == [1, 0, 1, 0]


@pytest.mark.parametrize("rep", [R_NONE, R_BIL])
def test_difference_validator_leaves_everything_invalid(rep):
    """Every tile starts invalid and the difference validator only updates a
    tile from a valid neighbour, so no tile ever becomes valid and the whole
    field is zeroed (the reference's behaviour with this validator)."""
    top, left = grid(6, 5)
    rng = np.random.default_rng(5)
    u = rng.normal(0, 0.1, (6, 5)).astype(np.float32)
    v = rng.normal(0, 0.1, (6, 5)).astype(np.float32)
    u2, v2, inv, stats = O.validate(top, left, u, v, [(DIFF, 1.0, 0.0, rep, 0)])
    assert inv.all() and not u2.any() and not v2.any()


def test_iterations_follow_should_do_another_attempt():
    """Two adjacent outliers whose bilinear estimates use each other: the first
    sweep corrects both, the second revalidates them unchanged and stops."""
    top, left = grid(7, 7)
    u, v = _linear(7, 7)
    uo, vo = u.copy(), v.copy()
    uo[3, 3] += 25.0
    uo[3, 4] -= 25.0
    _, _, _, s1 = O.validate(top, left, uo, vo, [(MED, 2.0, 0.1, R_BIL, 0)],
                             max_iterations=0, until_no_more=False)
    assert s1[0] == 1
    _, _, _, s2 = O.validate(top, left, uo, vo, [(MED, 2.0, 0.1, R_BIL, 0)],
                             max_iterations=5, until_no_more=False)
    assert s2[0] >= 2 and s2[1] >= s1[1]
    _, _, _, s3 = O.validate(top, left, uo, vo, [(MED, 2.0, 0.1, R_BIL, 0)],
                             until_no_more=True)
    assert s3[0] == s2[0]


def test_locked_tiles_keep_their_verdict():
    """1 x 3 grid u = (0, 1, 10), v = 0; spiral order: middle, right, left.
    Initializer: middle has median 5 and rMed = sqrt(rIs[1]) = 5, distance
    4/5.1 -> valid (locked); right: |1 - 10|/1.1 -> invalid; left:
    1/1.1 -> valid (locked).  The right tile's second peak gives
    10 - 10 + (-1) = -1, valid next to 1 (2/1.1 < 2).  The second sweep would
    find the middle invalid against (0, -1) (median -0.5, rMed 0.5, distance
    1.5/0.6 = 2.5) but it is locked, so the job ends with (0, 1, -1), nothing
    invalid, two sweeps and one correction."""
    top, left = grid(1, 3)
    u = np.array([[0.0, 1.0, 10.0]], np.float32)
    v = np.zeros((1, 3), np.float32)
    inc = np.zeros((1, 3, 2, 2), np.float32)
    inc[0, 2, 0] = (10.0, 0.0)
    inc[0, 2, 1] = (-1.0, 0.0)
    stages = [(MED, 2.0, 0.1, R_SEC, 1)]
    u2, v2, inv, stats = O.validate(top, left, u, v, stages, peak_uv=inc, max_iterations=5)
    np.testing.assert_array_equal(u2, [[0.0, 1.0, -1.0]])
    assert not inv.any()
    assert list(stats) == [2, 1, 0, 0]


def test_norm_ties_keep_insertion_order():
    """Neighbours in window order: (0,1) (2,0) (-1,0) (0,-2) (1,0) (0,2) (0,-1)
    (-2,0); norms 1 and 4 tie in fours.  The stable sort keeps window order
    within a tie, so the median is ((0,-1) + (2,0))/2 = (1, -0.5); the sorted
    residuals are .25 1.25 1.25 3.25 3.25 4.25 7.25 9.25, rMed = sqrt(3.25).
    The centre (4.7, -0.5) lies 3.7 / 1.9028 = 1.94 from it (valid), (4.9,
    -0.5) 2.05 (invalid)."""
    top, left = grid(3, 3)
    nb = [(0, 1), (2, 0), (-1, 0), (0, -2), (1, 0), (0, 2), (0, -1), (-2, 0)]
    cells = [(0, 0), (0, 1), (0, 2), (1, 0), (1, 2), (2, 0), (2, 1), (2, 2)]
    for c, expect in [((4.7, -0.5), False), ((4.9, -0.5), True)]:
        u = np.zeros((3, 3), np.float32)
        v = np.zeros((3, 3), np.float32)
        for (i, j), (a, b) in zip(cells, nb):
            u[i, j], v[i, j] = a, b
        u[1, 1], v[1, 1] = c
        _, _, inv, _ = O.validate(top, left, u, v, [(MED, 2.0, 0.1, R_NONE, 0)])
        assert bool(inv[1, 1]) == expect, c


def test_bilinear_classification_weights():
    """Quadratic field u = c*i^2, v = c*j^2 (c = 1/32) with an outlier at (4, 4).
    The horizontal pair gives v = c(j^2 + 1) (weight 2 for v, 1 for u), the
    vertical pair u = c(i^2 + 1) (weight 2 for u, 1 for v), the four corners
    c(i^2 + 1), c(j^2 + 1) (weight 3), so the replacement is
    u = c*i^2 + 5c/6, v = c*j^2 + 5c/6.  A second stage with a very large
    threshold accepts the replacement whatever its value."""
    c = 1.0 / 32
    top, left = grid(9, 9)
    i, j = np.meshgrid(np.arange(9), np.arange(9), indexing="ij")
    u = (c * i * i).astype(np.float32)
    v = (c * j * j).astype(np.float32)
    u[4, 4] += 20.0
    stages = [(MED, 2.0, 0.1, R_NONE, 0), (MED, 1e30, 0.1, R_BIL, 0)]
    u2, v2, inv, stats = O.validate(top, left, u, v, stages)
    assert not inv.any() and stats[1] == 1
    assert abs(u2[4, 4] - (c * 16 + 5 * c / 6)) < 1e-6
    assert abs(v2[4, 4] - (c * 16 + 5 * c / 6)) < 1e-6


def test_each_stage_restarts_from_the_main_peak():
    """Tile.restoreDisplacement before every replacement: with a wrong first
    secondary peak, the second is applied to the main-peak displacement
    (1.5 + 9 - 9 + 0.25), not on top of the first replacement."""
    nI, nJ = 5, 5
    top, left = grid(nI, nJ)
    inc = np.zeros((nI, nJ, 3, 2), np.float32)
    inc[..., 0, 0] = 0.25
    u = np.full((nI, nJ), 1.75, np.float32)
    v = np.full((nI, nJ), -0.5, np.float32)
    inc[2, 2, 0] = (9.0, 7.0)
    inc[2, 2, 1] = (5.0, 5.0)
    inc[2, 2, 2] = (0.25, 0.0)
    u[2, 2] = 1.5 + 9.0
    v[2, 2] = -0.5 + 7.0
    stages = [(MED, 2.0, 0.1, R_SEC, 1), (MED, 2.0, 0.1, R_SEC, 2)]
    u2, v2, inv, stats = O.validate(top, left, u, v, stages, peak_uv=inc)
    assert not inv.any()
    assert u2[2, 2] == np.float32(1.75) and v2[2, 2] == np.float32(-0.5)


def test_difference_validator_after_a_secondary_peak():
    """A secondary-peak replacement marks its tile valid, which the difference
    validator then keeps (it has no valid neighbour to compare with); later
    tiles compare against it with the squared threshold.  1 x 2 grid, visit
    order left, right: left (5, 0) -> 5 - 5 + 1 = (1, 0), valid; right (2, 1)
    is 1^2 + 1^2 = 2 < 1.5^2 from it -> valid, unchanged."""
    top, left = grid(1, 2)
    u = np.array([[5.0, 2.0]], np.float32)
    v = np.array([[0.0, 1.0]], np.float32)
    inc = np.zeros((1, 2, 2, 2), np.float32)
    inc[0, 0, 0] = (5.0, 0.0)
    inc[0, 0, 1] = (1.0, 0.0)
    inc[0, 1, 1] = (5.0, 5.0)
    u2, v2, inv, stats = O.validate(top, left, u, v, [(DIFF, 1.5, 0.0, R_SEC, 1)], peak_uv=inc)
    assert not inv.any()
    np.testing.assert_array_equal(u2, [[1.0, 2.0]])
    np.testing.assert_array_equal(v2, [[0.0, 1.0]])


def test_difference_validator_consults_only_three_neighbours():
    """1 x 3 grid visited middle, right, left.  Middle: second peak -> 1,
    valid.  Right (1.5): its left neighbour (the middle) is within 1.5 ->
    valid.  Left: its upper-left, upper and left neighbours do not exist, so
    after its second-peak replacement (20) nothing can invalidate it, although
    its right neighbour is far away."""
    top, left = grid(1, 3)
    u = np.array([[10.0, 5.0, 1.5]], np.float32)
    v = np.zeros((1, 3), np.float32)
    inc = np.zeros((1, 3, 2, 2), np.float32)
    inc[0, :, 0, 0] = u[0]
    inc[0, 0, 1, 0] = 20.0
    inc[0, 1, 1, 0] = 1.0
    inc[0, 2, 1, 0] = 1.5
    u2, v2, inv, _ = O.validate(top, left, u, v, [(DIFF, 1.5, 0.0, R_SEC, 1)], peak_uv=inc)
    assert not inv.any()
    np.testing.assert_array_equal(u2, [[20.0, 1.0, 1.5]])
