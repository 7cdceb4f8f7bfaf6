"""The warp orders (SecondImage, FirstImage, BothImages) of the Micro, Mini and
Full-image warps on the GPU against the oracle: warped and clipped tiles and
frames bit for bit, and every level of the multi-pass plan in each order."""
import sys

import numpy as np
import pytest

import oracle as O
from qlpiv.geometry import level_grid

pytestmark = pytest.mark.gpu

SECOND, FIRST, BOTH = 0, 1, 2
ORDERS = {"second": SECOND, "first": FIRST, "both": BOTH}
A_SIDE = {FIRST: O.SIDE_A_FIRST, BOTH: O.SIDE_A_BOTH}
B_SIDE = {SECOND: O.SIDE_B_SECOND, BOTH: O.SIDE_B_BOTH}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import qlpiv

    c = qlpiv.Context(0)
    yield c
    c.close()


def field(top, left, N, seed, pairs):
    rng = np.random.default_rng(seed)
    cy = top + N / 2.0
    cx = left + N / 2.0
    u = np.stack([(1.3 + np.sin(cy / 40.0) * 2.5 + 0.01 * cx + rng.normal(0, 0.3, top.shape))
                  for _ in range(pairs)]).astype(np.float32)
    v = np.stack([(-0.8 + np.cos(cx / 35.0) * 2.0 - 0.01 * cy + rng.normal(0, 0.3, top.shape))
                  for _ in range(pairs)]).astype(np.float32)
    return u, v


def ref_tiles(mode, order, a, b, top, left, N, u, v, i, j):
    """Oracle A and B tiles and maxima of IA (i, j) at a level > 0."""
    t, l0 = int(top[i, j]), int(left[i, j])

    def warp(img, side):
        if mode == "micro":
            return O.warp_micro(img, t, l0, N, u[i, j], v[i, j], side), 255.0
        return O.warp_mini(img, top, left, N, u, v, i, j, side), 255.0

    tA, mA = O.clip_tile(a, t, l0, N) if order == SECOND else warp(a, A_SIDE[order])
    tB, mB = O.clip_tile(b, t, l0, N) if order == FIRST else warp(b, B_SIDE[order])
    return tA, mA, tB, mB


GRIDS = [("256x256/16", 256, 256, 16, (0, 0, 0, 0)), ("300x220/32", 300, 220, 32, (3, 5, 7, 2))]


@pytest.mark.parametrize("mode", ["micro", "mini"])
@pytest.mark.parametrize("order", list(ORDERS))
@pytest.mark.parametrize("grid", GRIDS, ids=[g[0] for g in GRIDS])
def test_warp_clip_tiles_equal_oracle(ctx, mode, order, grid):
    import torch

    import qlpiv

    name, H, W, N, m = grid
    o = ORDERS[order]
    rng = np.random.default_rng(H + W + o)
    a = rng.integers(0, 256, (2, H, W)).astype(np.uint8)
    b = rng.integers(0, 256, (2, H, W)).astype(np.uint8)
    top, left = level_grid(H, W, N, 0.5, m, level0=False)
    u, v = field(top, left, N, 30 + o, 2)
    wm = qlpiv._lib.QLP_WARP_MICRO if mode == "micro" else qlpiv._lib.QLP_WARP_MINI
    say(f"warp_clip_tiles {mode} {order} {name}")
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    tA, mA, tB, mB = ctx.warp_clip_tiles(dev(a), dev(b), dev(top.astype(np.int32)),
                                         dev(left.astype(np.int32)), dev(u), dev(v), N, wm, o)
    ctx.synchronize()
    tA, mA, tB, mB = (x.cpu().numpy() for x in (tA, mA, tB, mB))
    nI, nJ = top.shape
    for k in range(2):
        for i in range(nI):
            for j in range(nJ):
                ra, rma, rb, rmb = ref_tiles(mode, o, a[k], b[k], top, left, N, u[k], v[k], i, j)
                t = i * nJ + j
                assert np.array_equal(tA[k, t], ra) and mA[k, t] == rma, (k, i, j)
                assert np.array_equal(tB[k, t], rb) and mB[k, t] == rmb, (k, i, j)


@pytest.mark.parametrize("order", list(ORDERS))
def test_warp_images_equal_oracle(ctx, order):
    import torch

    H, W, N = 300, 220, 32
    o = ORDERS[order]
    rng = np.random.default_rng(77 + o)
    a = rng.integers(0, 256, (2, H, W)).astype(np.uint8)
    b = rng.integers(0, 256, (2, H, W)).astype(np.uint8)
    top, left = level_grid(H, W, N, 0.5, (3, 5, 7, 2), level0=False)
    u, v = field(top, left, N, 40 + o, 2)
    say(f"warp_images {order}")
    outA, outB = ctx.warp_images(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), top, left,
                                 torch.from_numpy(u).cuda(), torch.from_numpy(v).cuda(), N, o)
    ctx.synchronize()
    assert (outA is None) == (o == SECOND) and (outB is None) == (o == FIRST)
    for k in range(2):
        fu, fv = O.full_field(top, left, N, u[k], v[k], H, W)
        if outA is not None:
            assert np.array_equal(outA[k].cpu().numpy(), O.warp_image(a[k], fu, fv, A_SIDE[o])), k
        if outB is not None:
            assert np.array_equal(outB[k].cpu().numpy(), O.warp_image(b[k], fu, fv, B_SIDE[o])), k


CASES = [(m, o) for m in ("micro", "mini", "full") for o in ("first", "both")] + [("micro", "second")]


@pytest.mark.parametrize("mode,order", CASES, ids=[f"{m}-{o}" for m, o in CASES])
def test_multipass_warp_orders_against_oracle(ctx, mode, order):
    """Every level of the device loop equals the oracle run from the previous
    level's device output: integer peaks bit-exact, displacements within 1e-4 px."""
    import torch

    import qlpiv
    from qlpiv.synth import particle_pair

    o = ORDERS[order]
    wm = {"micro": qlpiv._lib.QLP_WARP_MICRO, "mini": qlpiv._lib.QLP_WARP_MINI,
          "full": qlpiv._lib.QLP_WARP_FULL}[mode]
    Hh = Ww = 256
    a, b, _ = particle_pair(Hh, Ww, seed=47 + o, shift=(2.7, -3.4), core_radius=60.0, u_max=4.0)
    plan = qlpiv.PivPlan(ctx, Hh, Ww, 1, ia_start=64, ia_end=16, warp_mode=wm, warp_order=o,
                         subpix_kind=qlpiv.QLP_SUBPIX_GUO)
    say(f"multipass warp={mode} order={order}")
    plan.run(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    prev = None
    for lv, L in enumerate(plan.levels):
        u, vv, pk = (x[0] for x in plan.result(lv))
        N, nI, nJ = L["ia"], L["nI"], L["nJ"]
        uin = np.zeros((nI, nJ), np.float32)
        vin = np.zeros((nI, nJ), np.float32)
        if lv > 0:
            P, pu, pv = prev
            for i in range(nI):
                for j in range(nJ):
                    if i // 2 < P["nI"] and j // 2 < P["nJ"]:
                        uin[i, j] = pu[i // 2, j // 2]
                        vin[i, j] = pv[i // 2, j // 2]
            if mode == "full":
                fu, fv = O.full_field(L["top"], L["left"], N, uin, vin, Hh, Ww)
                wa = O.warp_image(a, fu, fv, A_SIDE[o]) if o != SECOND else None
                wb = O.warp_image(b, fu, fv, B_SIDE[o]) if o != FIRST else None
        c = np.float32((2 * N - 1) // 2)
        for i in range(nI):
            for j in range(nJ):
                t, l0 = int(L["top"][i, j]), int(L["left"][i, j])
                if lv == 0:
                    tA, mA = O.clip_tile(a, t, l0, N)
                    tB, mB = O.clip_tile(b, t, l0, N)
                elif mode == "full":
                    tA, mA = (O.clip_tile(wa, t, l0, N)[0], 255.0) if wa is not None else O.clip_tile(a, t, l0, N)
                    tB, mB = (O.clip_tile(wb, t, l0, N)[0], 255.0) if wb is not None else O.clip_tile(b, t, l0, N)
                else:
                    tA, mA, tB, mB = ref_tiles(mode, o, a, b, L["top"], L["left"], N, uin, vin, i, j)
                plane = O.xcorr_javagpu(tA.astype(np.float32), tB.astype(np.float32), mA, mB)
                rp = O.find_max_from_center(plane)
                rs = O.guo(plane, 3, 20, rp[3], rp[:3])
                assert int(pk[i, j, 3]) == int(rs[3]), (lv, i, j)
                eu = np.float32(uin[i, j] + np.float32(rs[0] - c))
                ev = np.float32(vin[i, j] + np.float32(rs[1] - c))
                assert abs(u[i, j] - eu) <= 1e-4 and abs(vv[i, j] - ev) <= 1e-4, (lv, i, j)
        prev = (L, u, vv)
    plan.close()


def test_warp_order_errors(ctx):
    import qlpiv

    with pytest.raises(qlpiv.QlpError):
        qlpiv.PivPlan(ctx, 256, 256, 1, ia_start=64, ia_end=16, warp_order=3)
