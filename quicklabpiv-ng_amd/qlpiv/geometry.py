"""Interrogation-area placement of the reference's adaptive super-position
strategy (iareas/AdaptiveInterAreaStrategySuperPosition.java:31-197), in the
reference's float arithmetic (np.float32 where Java uses float)."""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32


def _round_half_up(x):
    return int(math.floor(float(f32(x) + f32(0.5))))


def _count(extent, margins, tile, inc):
    # (int)(0.01f+((float)(width - margins) - (tileWidth - widthIncrement)) / widthIncrement)
    return int(f32(0.01) + (f32(extent - margins) - (f32(tile) - inc)) / inc)


def level_grid(H, W, tile, overlap_factor, margins=(0, 0, 0, 0), level0=True):
    """Tile tops/lefts of one adaptive level.

    overlap_factor = 1 - overlap (0.5 for 50 % overlap; the reference's
    PIVInputParameters.getOverlapFactor, PIVConfigurationConverterVisitor.java:146).
    margins = (top, bottom, left, right).  Returns (tops, lefts) as int arrays of
    shape (nI, nJ)."""
    mt, mb, ml, mr = margins
    of = f32(overlap_factor)
    hinc = of * f32(tile)
    winc = of * f32(tile)
    nJ = _count(W, ml + mr, tile, winc)
    nI = _count(H, mt + mb, tile, hinc)
    remI = H - int(math.ceil(float((nI - 1) * hinc + f32(tile)))) - (mt + mb)
    remJ = W - int(math.ceil(float((nJ - 1) * winc + f32(tile)))) - (ml + mr)
    top0 = mt + (remI // 2)
    left0 = ml + (remJ // 2)
    tops = np.zeros((nI, nJ), np.int32)
    lefts = np.zeros((nI, nJ), np.int32)
    if level0:
        for i in range(nI):
            for j in range(nJ):
                tops[i, j] = top0 + _round_half_up(f32(i) * hinc)
                lefts[i, j] = left0 + _round_half_up(f32(j) * winc)
    else:
        tp = f32(top0)
        for i in range(nI):
            lp = f32(left0)
            for j in range(nJ):
                tops[i, j] = _round_half_up(tp)
                lefts[i, j] = _round_half_up(lp)
                lp = f32(lp + winc)
            tp = f32(tp + hinc)
    return tops, lefts


def multipass_grids(H, W, ia_start, ia_end, overlap_factor, margins=(0, 0, 0, 0)):
    """All levels from ia_start down to ia_end (halving), as a list of
    (tile, tops, lefts)."""
    steps = ia_start // ia_end
    levels = int(round(math.log2(steps))) + 1
    if 2 ** (levels - 1) != steps or ia_start % ia_end:
        raise ValueError("IA start must be a power-of-two multiple of IA end")
    out = []
    tile = ia_start
    for lv in range(levels):
        tops, lefts = level_grid(H, W, tile, overlap_factor, margins, level0=(lv == 0))
        out.append((tile, tops, lefts))
        tile //= 2
    return out
