"""GPU tests of the Mini and Full-image warps (qlp_warp_mini_tiles_dev,
qlp_warp_image_dev, and the plan's warp modes) against the oracle, bit for bit
on the warped pixels; the multi-pass plan with each mode level by level."""
import sys

import numpy as np
import pytest

import oracle as O
from qlpiv.geometry import level_grid

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import qlpiv

    c = qlpiv.Context(0)
    yield c
    c.close()


def _field(top, left, N, seed):
    rng = np.random.default_rng(seed)
    cy = top + N / 2.0
    cx = left + N / 2.0
    u = (1.3 + np.sin(cy / 40.0) * 2.0 + 0.01 * cx + rng.normal(0, 0.2, top.shape)).astype(np.float32)
    v = (-0.8 + np.cos(cx / 35.0) * 1.5 - 0.01 * cy + rng.normal(0, 0.2, top.shape)).astype(np.float32)
    return u, v


GRIDS = [("256x256/16", 256, 256, 16, (0, 0, 0, 0)), ("300x220/32", 300, 220, 32, (3, 5, 7, 2)),
         ("96x64/32 (3x2 grid)", 96, 64, 32, (0, 0, 0, 0)), ("128x512/16", 128, 512, 16, (0, 0, 0, 0))]


@pytest.mark.parametrize("case", GRIDS, ids=[g[0] for g in GRIDS])
def test_mini_warp_equals_oracle(ctx, case):
    import torch

    name, H, W, N, m = case
    rng = np.random.default_rng(H + W)
    imgs = rng.integers(0, 256, (2, H, W)).astype(np.uint8)
    top, left = level_grid(H, W, N, 0.5, m, level0=False)
    uv = [_field(top, left, N, 10 + k) for k in range(2)]
    u = np.stack([a for a, _ in uv])
    v = np.stack([b for _, b in uv])
    say(f"mini warp {name}: grid {top.shape}")
    tiles, mx = ctx.warp_mini_tiles(torch.from_numpy(imgs).cuda(),
                                    torch.from_numpy(np.ascontiguousarray(top, np.int32)).cuda(),
                                    torch.from_numpy(np.ascontiguousarray(left, np.int32)).cuda(),
                                    torch.from_numpy(u).cuda(), torch.from_numpy(v).cuda(), N)
    ctx.synchronize()
    tiles = tiles.cpu().numpy()
    assert (mx.cpu().numpy() == 255.0).all()
    nI, nJ = top.shape
    for k in range(2):
        for ti in range(nI):
            for tj in range(nJ):
                ref = O.warp_mini(imgs[k], top, left, N, u[k], v[k], ti, tj)
                assert np.array_equal(tiles[k, ti * nJ + tj], ref), (k, ti, tj)


@pytest.mark.parametrize("case", GRIDS, ids=[g[0] for g in GRIDS])
def test_full_warp_equals_oracle(ctx, case):
    import torch

    name, H, W, N, m = case
    rng = np.random.default_rng(7 * H + W)
    imgs = rng.integers(0, 256, (2, H, W)).astype(np.uint8)
    top, left = level_grid(H, W, N, 0.5, m, level0=False)
    uv = [_field(top, left, N, 20 + k) for k in range(2)]
    u = np.stack([a for a, _ in uv])
    v = np.stack([b for _, b in uv])
    say(f"full warp {name}: grid {top.shape}")
    out = ctx.warp_image(torch.from_numpy(imgs).cuda(), top, left, torch.from_numpy(u).cuda(),
                         torch.from_numpy(v).cuda(), N)
    ctx.synchronize()
    out = out.cpu().numpy()
    for k in range(2):
        fu, fv = O.full_field(top, left, N, u[k], v[k], H, W)
        ref = O.warp_image(imgs[k], fu, fv)
        assert np.array_equal(out[k], ref), k


def test_spline_warps_reject_single_row_grids(ctx):
    import torch

    import qlpiv

    img = torch.zeros((1, 32, 200), dtype=torch.uint8, device="cuda")
    top, left = level_grid(32, 200, 32, 0.5, level0=False)
    assert top.shape[0] == 1
    z = torch.zeros(top.shape, dtype=torch.float32, device="cuda")
    with pytest.raises(qlpiv.QlpError):
        ctx.warp_mini_tiles(img, torch.from_numpy(np.ascontiguousarray(top, np.int32)).cuda(),
                            torch.from_numpy(np.ascontiguousarray(left, np.int32)).cuda(), z, z, 32)
    with pytest.raises(qlpiv.QlpError):
        ctx.warp_image(img, top, left, z, z, 32)


@pytest.mark.parametrize("mode", ["mini", "full"])
def test_multipass_with_spline_warp_against_oracle(ctx, mode):
    """Every level of the device loop with the Mini or Full warp equals the
    oracle run from the previous level's device output: integer peaks
    bit-exact, displacements within 1e-4 px."""
    import torch

    import qlpiv
    from qlpiv.synth import particle_pair

    wm = {"mini": qlpiv._lib.QLP_WARP_MINI, "full": qlpiv._lib.QLP_WARP_FULL}[mode]
    a, b, _ = particle_pair(256, 256, seed=43, shift=(1.7, -2.2), core_radius=60.0, u_max=4.0)
    plan = qlpiv.PivPlan(ctx, 256, 256, 1, ia_start=64, ia_end=16, warp_mode=wm,
                         subpix_kind=qlpiv.QLP_SUBPIX_GUO)
    say(f"multipass warp={mode}")
    plan.run(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    prev = None
    for lv, L in enumerate(plan.levels):
        u, v, pk = (x[0] for x in plan.result(lv))
        N, nI, nJ = L["ia"], L["nI"], L["nJ"]
        if lv == 0:
            uin = np.zeros((nI, nJ), np.float32)
            vin = np.zeros((nI, nJ), np.float32)
        else:
            P, pu, pv = prev
            uin = np.zeros((nI, nJ), np.float32)
            vin = np.zeros((nI, nJ), np.float32)
            for i in range(nI):
                for j in range(nJ):
                    if i // 2 < P["nI"] and j // 2 < P["nJ"]:
                        uin[i, j] = pu[i // 2, j // 2]
                        vin[i, j] = pv[i // 2, j // 2]
            if mode == "full":
                fu, fv = O.full_field(L["top"], L["left"], N, uin, vin, 256, 256)
                wb = O.warp_image(b, fu, fv)
        c = np.float32((2 * N - 1) // 2)
        for i in range(nI):
            for j in range(nJ):
                t, l0 = int(L["top"][i, j]), int(L["left"][i, j])
                tA, mA = O.clip_tile(a, t, l0, N)
                if lv == 0:
                    tB, mB = O.clip_tile(b, t, l0, N)
                elif mode == "mini":
                    tB, mB = O.warp_mini(b, L["top"], L["left"], N, uin, vin, i, j), 255.0
                else:
                    tB, mB = O.clip_tile(wb, t, l0, N)[0], 255.0
                plane = O.xcorr_javagpu(tA.astype(np.float32), tB.astype(np.float32), mA, mB)
                rp = O.find_max_from_center(plane)
                rs = O.guo(plane, 3, 20, rp[3], rp[:3])
                assert int(pk[i, j, 3]) == int(rs[3]), (lv, i, j)
                eu = np.float32(uin[i, j] + np.float32(rs[0] - c))
                ev = np.float32(vin[i, j] + np.float32(rs[1] - c))
                assert abs(u[i, j] - eu) <= 1e-4 and abs(v[i, j] - ev) <= 1e-4, (lv, i, j)
        prev = (L, u, v)
    plan.close()
