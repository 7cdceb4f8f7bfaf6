"""The oracle's warp orders (BiLinearImageWarpingModeEnum: SecondImage,
FirstImage, BothImages) for the Micro, Mini and Full-image warps, pinned by what
the geometry fixes: the sampled position of every side on a linear ramp, where
the bilinear sample is exact (BothImages: the two frames are sampled u apart,
centred as the reference's derivation in warpAndClipImage :96-117 states),
FirstImage of u being SecondImage of -u, integer fields being shifts, and a
uniform grid making the Mini field the Micro one."""
from fractions import Fraction as Fr
import math

import numpy as np
import pytest

import oracle as O
from qlpiv.geometry import level_grid

H, W = 40, 60
RAMP = (np.arange(H)[:, None] + 2 * np.arange(W)[None, :] + 10).astype(np.uint8)
DISPS = [(2.25, -1.75), (-3.5, 0.75), (0.25, 1.5), (-0.75, -2.25), (4.0, -3.0)]


def ramp_value(y, x):
    """Round-half-up of the ramp at a position inside the frame."""
    return math.floor(y + 2 * x + 10 + Fr(1, 2))


def frac(a):
    return a - math.floor(a)


@pytest.mark.parametrize("u,v", DISPS)
def test_micro_sides_sample_their_positions_on_a_ramp(u, v):
    top, left, N = 12, 20, 8
    fu, fv = Fr(u), Fr(v)
    pos = {  # the position each side samples for pixel (0, 0) of the window
        O.SIDE_B_SECOND: (fu, fv),
        O.SIDE_A_FIRST: (-fu, -fv),
        O.SIDE_B_BOTH: (fu / 2 + frac(fu) / 2, fv / 2 + frac(fv) / 2),
        O.SIDE_A_BOTH: (-fu / 2 + frac(fu) / 2, -fv / 2 + frac(fv) / 2),
    }
    # BothImages: B sampled exactly u after A, as the reference intends
    assert pos[O.SIDE_B_BOTH][0] - pos[O.SIDE_A_BOTH][0] == fu
    for side, (dy, dx) in pos.items():
        tile = O.warp_micro(RAMP, top, left, N, u, v, side)
        exp = np.array([[ramp_value(top + i + dy, left + j + dx) for j in range(N)] for i in range(N)])
        assert np.array_equal(tile.astype(int), exp), side


@pytest.mark.parametrize("u,v", DISPS)
def test_micro_first_image_is_second_image_of_the_negated_displacement(u, v):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (H, W)).astype(np.uint8)
    for top, left in ((0, 0), (12, 20), (H - 8, W - 8), (-3, 55)):
        a = O.warp_micro(img, top, left, 8, u, v, O.SIDE_A_FIRST)
        b = O.warp_micro(img, top, left, 8, -u, -v, O.SIDE_B_SECOND)
        assert np.array_equal(a, b), (top, left)


def test_micro_both_images_halves_differ_from_plain_orders():
    """A non-trivial displacement moves each BothImages side by half."""
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (H, W)).astype(np.uint8)
    a = O.warp_micro(img, 12, 20, 8, 3.0, 2.0, O.SIDE_A_BOTH)
    b = O.warp_micro(img, 12, 20, 8, 3.0, 2.0, O.SIDE_B_BOTH)
    # integer u: A at -u/2 + 0, B at +u/2 (1.5 and 1 px away from the window)
    assert np.array_equal(a, O.warp_micro(img, 12, 20, 8, -1.5, -1.0, O.SIDE_B_SECOND))
    assert np.array_equal(b, O.warp_micro(img, 12, 20, 8, 1.5, 1.0, O.SIDE_B_SECOND))


@pytest.mark.parametrize("side", [O.SIDE_B_SECOND, O.SIDE_A_FIRST, O.SIDE_A_BOTH, O.SIDE_B_BOTH])
def test_mini_on_a_uniform_grid_is_the_micro_warp(side):
    """Every tile displaced alike: the Mini field is the constant fractional
    part everywhere, so each side's Mini tile is its Micro tile."""
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (128, 128)).astype(np.uint8)
    top, left = level_grid(128, 128, 16, 0.5, level0=False)
    u = np.full(top.shape, 2.25, np.float32)
    v = np.full(top.shape, -1.75, np.float32)
    for ti, tj in ((0, 0), (3, 4), (top.shape[0] - 1, top.shape[1] - 1)):
        m = O.warp_mini(img, top, left, 16, u, v, ti, tj, side)
        r = O.warp_micro(img, int(top[ti, tj]), int(left[ti, tj]), 16, 2.25, -1.75, side)
        assert np.array_equal(m, r), (side, ti, tj)


def test_full_image_sides_with_integer_fields_are_shifts():
    rng = np.random.default_rng(6)
    img = rng.integers(0, 256, (H, W)).astype(np.uint8)
    F, G = 3.0, -2.0
    fu = np.full((H, W), F, np.float32)
    fv = np.full((H, W), G, np.float32)
    inner = (slice(8, H - 8), slice(8, W - 8))
    second = O.warp_image(img, fu, fv, O.SIDE_B_SECOND)
    first = O.warp_image(img, fu, fv, O.SIDE_A_FIRST)
    assert np.array_equal(second[inner], img[11:H - 5, 6:W - 10])
    assert np.array_equal(first[inner], img[5:H - 11, 10:W - 6])
    # BothImages halves the field: 2F moves A by -F and B by +F
    a = O.warp_image(img, 2 * fu, 2 * fv, O.SIDE_A_BOTH)
    b = O.warp_image(img, 2 * fu, 2 * fv, O.SIDE_B_BOTH)
    assert np.array_equal(a, first) and np.array_equal(b, second)
