"""Float tiles, float frames (frames filtered before warping) and 10/12/16-bit
frames on the GPU.

GaussianFilter2D.applyFilter(IImage, IImage) (images/filters/GaussianFilter2D
.java:116-126, JavaImageFilterJob.compute :50-53) turns an 8-bit frame into an
ImageFloat; its tiles are MatrixFloat (ImageFloat.clipImageMatrix :278-384,
maximum over the tile, 0 -> 1) and its Micro-warped tiles carry the rounded
samples with the ImageFloat maximum 65535.  Every stage is checked against the
oracle on the same float frames: tiles, maxima and planes bit for bit, the
multi-pass displacements within 1e-4 px (the Guo fit's FP64 log/exp).
"""
import sys

import numpy as np
import pytest

import oracle as O

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


@pytest.fixture(scope="module")
def fpair(ctx):
    """A synthetic 256^2 pair filtered by the library's Gaussian (sigma 1.2, 5
    taps): device float frames and their host copies."""
    import torch

    from qlpiv.synth import particle_pair

    a, b, _ = particle_pair(256, 256, seed=41, shift=(1.6, -2.4), core_radius=60.0, u_max=4.0)
    da = ctx.gaussian_filter(torch.from_numpy(a).cuda(), 1.2, 5)
    db = ctx.gaussian_filter(torch.from_numpy(b).cuda(), 1.2, 5)
    ctx.synchronize()  # the library runs on its own stream
    ha, hb = da.cpu().numpy(), db.cpu().numpy()
    assert not np.array_equal(ha, np.round(ha))  # genuinely fractional frames
    return da, db, ha, hb


@pytest.mark.parametrize("N", [16, 32, 64, 128])
@pytest.mark.parametrize("mode", [0, 1])
def test_xcorr_f32_tiles_equal_oracle(ctx, fpair, N, mode):
    import torch

    import qlpiv

    _, _, ha, hb = fpair
    rng = np.random.default_rng(N + mode)
    n = 12 if N < 128 else 4
    tA = np.empty((n, N, N), np.float32)
    tB = np.empty((n, N, N), np.float32)
    mA = np.empty(n, np.float32)
    mB = np.empty(n, np.float32)
    for k in range(n):
        t, l0 = (int(x) for x in rng.integers(-N // 2, 256 - N // 2, 2))
        tA[k], mA[k] = O.clip_tile_f32(ha, t, l0, N)
        tB[k], mB[k] = O.clip_tile_f32(hb, t + 2, l0 - 1, N)
    tA[0] = 0.0  # an all-zero tile: maximum 1
    mA[0] = 1.0
    say(f"xcorr_f32 N={N} mode={mode}")
    host = ctx.xcorr_f32(tA, tB, mA, mB, mode)
    dev = ctx.xcorr_f32(torch.from_numpy(tA).cuda(), torch.from_numpy(tB).cuda(),
                        torch.from_numpy(mA).cuda(), torch.from_numpy(mB).cuda(), mode)
    ctx.synchronize()
    dev = dev.cpu().numpy()
    for k in range(n):
        if mode == qlpiv.QLP_NORM_A_MAX:
            ref = O.xcorr_javacpu(tA[k], tB[k], mA[k])
        else:
            ref = O.xcorr_javagpu(tA[k], tB[k], mA[k], mB[k])
        assert np.array_equal(host[k], ref), k
        assert np.array_equal(dev[k], ref), k


@pytest.mark.parametrize("side", [0, 1, 2, 3])
def test_float_clip_and_micro_warp_equal_oracle(ctx, fpair, side):
    import torch

    da, _, ha, _ = fpair
    N = 32
    rng = np.random.default_rng(5 + side)
    n = 40
    top = rng.integers(-20, 240, n).astype(np.int32)
    left = rng.integers(-20, 240, n).astype(np.int32)
    u = rng.uniform(-6, 6, n).astype(np.float32)
    v = rng.uniform(-6, 6, n).astype(np.float32)
    dt, dl = torch.from_numpy(top).cuda(), torch.from_numpy(left).cuda()
    du, dv = torch.from_numpy(u).cuda(), torch.from_numpy(v).cuda()
    say(f"float clip / micro warp side={side}")
    outs = [ctx.clip_tiles_f32(da, dt, dl, N), ctx.clip_tiles_f32(da, dt, dl, N, du, dv),
            ctx.warp_micro_tiles_f32(da, dt, dl, du, dv, N, side)]
    ctx.synchronize()
    (ct, cm), (st, sm), (wt, wm) = ([x.cpu().numpy() for x in o] for o in outs)
    for k in range(n):
        rt, rm = O.clip_tile_f32(ha, top[k], left[k], N)
        assert np.array_equal(ct[k], rt) and cm[k] == np.float32(rm), k
        tb = int(np.floor(np.float32(top[k]) + u[k] + np.float32(0.5)))
        lb = int(np.floor(np.float32(left[k]) + v[k] + np.float32(0.5)))
        rt, rm = O.clip_tile_f32(ha, tb, lb, N)
        assert np.array_equal(st[k], rt) and sm[k] == np.float32(rm), k
        rt, rm = O.warp_micro_f32(ha, top[k], left[k], N, u[k], v[k], side)
        assert np.array_equal(wt[k], rt) and wm[k] == np.float32(rm), k


def direct_inherit(pu, nI, nJ, round_it=False):
    """Direct inheritance: child (i, j) takes parent (i // 2, j // 2); children
    past the parent grid start from 0."""
    out = np.zeros((nI, nJ), np.float32)
    pI, pJ = pu.shape
    for i in range(nI):
        for j in range(nJ):
            if i // 2 < pI and j // 2 < pJ:
                x = pu[i // 2, j // 2]
                out[i, j] = np.float32(np.floor(x + np.float32(0.5))) if round_it else x
    return out


@pytest.mark.parametrize("warp,order", [(1, 0), (1, 2), (1, 1), (0, 0), (2, 0), (2, 2), (2, 1),
                                        (3, 0), (3, 1), (3, 2)])
def test_float_frame_plan_against_oracle(ctx, fpair, warp, order):
    """Every level of a float-frame plan (64 -> 16 px, Direct inheritance,
    FromCenter, Guo 3 px) equals the oracle restatement from the same inherited
    displacements: planes bit-exact on a sample, fit status equal, displacements
    within 1e-4 px."""
    import qlpiv

    da, db, ha, hb = fpair
    plan = qlpiv.PivPlan(ctx, 256, 256, 1, ia_start=64, ia_end=16, overlap_factor=0.5,
                         warp_mode=warp, warp_order=order, subpix_kind=qlpiv.QLP_SUBPIX_GUO,
                         subpix_pixels=3, subpix_iterations=20, keep_planes=2,
                         frame_type=qlpiv._lib.QLP_FRAME_F32)
    say(f"float-frame plan warp={warp} order={order}")
    plan.run(da.unsqueeze(0).contiguous(), db.unsqueeze(0).contiguous())
    prev_u = prev_v = None
    sides_a = {1: O.SIDE_A_FIRST, 2: O.SIDE_A_BOTH}
    sides_b = {0: O.SIDE_B_SECOND, 2: O.SIDE_B_BOTH}
    for lv, L in enumerate(plan.levels):
        u, v, pk = (x[0] for x in plan.result(lv))
        planes = plane_copy(plan, lv)
        N, nI, nJ = L["ia"], L["nI"], L["nJ"]
        if lv == 0:
            uin = np.zeros((nI, nJ), np.float32)
            vin = np.zeros((nI, nJ), np.float32)
        else:
            uin = direct_inherit(prev_u, nI, nJ, round_it=(warp == 0))
            vin = direct_inherit(prev_v, nI, nJ, round_it=(warp == 0))
        D = 2 * N - 1
        c = np.float32(D // 2)
        if lv > 0 and warp == 3:
            # Full warp: whole frames warped, tiles cut from them with the source
            # frame's maximum
            fu, fv = O.full_field(L["top"], L["left"], N, uin, vin, 256, 256)
            wa = O.warp_image_f32(ha, fu, fv, sides_a[order]) if order in sides_a else None
            wb = O.warp_image_f32(hb, fu, fv, sides_b[order]) if order in sides_b else None
        for i in range(nI):
            for j in range(nJ):
                t, l0 = int(L["top"][i, j]), int(L["left"][i, j])
                if lv > 0 and warp == 3 and wa is not None:
                    tA, mA = wa[t:t + N, l0:l0 + N].copy(), max(float(ha.max()), 1.0)
                elif lv > 0 and warp == 2 and order in sides_a:
                    tA, mA = O.warp_mini_frame(ha, L["top"], L["left"], N, uin, vin, i, j,
                                               sides_a[order]), 65535.0
                elif lv > 0 and warp and order in sides_a:
                    tA, mA = O.warp_micro_f32(ha, t, l0, N, uin[i, j], vin[i, j], sides_a[order])
                else:
                    tA, mA = O.clip_tile_f32(ha, t, l0, N)
                if lv > 0 and warp == 3 and wb is not None:
                    tB, mB = wb[t:t + N, l0:l0 + N].copy(), max(float(hb.max()), 1.0)
                elif lv > 0 and warp == 2 and order in sides_b:
                    tB, mB = O.warp_mini_frame(hb, L["top"], L["left"], N, uin, vin, i, j,
                                               sides_b[order]), 65535.0
                elif lv > 0 and warp and order in sides_b:
                    tB, mB = O.warp_micro_f32(hb, t, l0, N, uin[i, j], vin[i, j], sides_b[order])
                elif lv > 0 and not warp:
                    tb = int(np.floor(np.float32(t) + uin[i, j] + np.float32(0.5)))
                    lb = int(np.floor(np.float32(l0) + vin[i, j] + np.float32(0.5)))
                    tB, mB = O.clip_tile_f32(hb, tb, lb, N)
                else:
                    tB, mB = O.clip_tile_f32(hb, t, l0, N)
                plane = O.xcorr_javagpu(tA, tB, mA, mB)
                if (i * nJ + j) % 5 == 0:
                    assert np.array_equal(planes[i, j], plane), (lv, i, j)
                rp = O.find_max_from_center(plane)
                rs = O.guo(plane, 3, 20, rp[3], rp[:3])
                eu = np.float32(uin[i, j] + np.float32(rs[0] - c))
                ev = np.float32(vin[i, j] + np.float32(rs[1] - c))
                assert int(pk[i, j, 3]) == int(rs[3]), (lv, i, j)
                assert abs(u[i, j] - eu) <= 1e-4 and abs(v[i, j] - ev) <= 1e-4, (lv, i, j)
        say(f"level {lv}: N={N} {nI}x{nJ} equal")
        prev_u, prev_v = u, v
    plan.close()


@pytest.mark.parametrize("bits,warp,order", [(12, 1, 0), (16, 1, 2), (10, 0, 0), (12, 2, 2)])
def test_u16_frame_plan_against_oracle(ctx, bits, warp, order):
    """10/12/16-bit frames (ImageInt16): clipped tiles take the MatrixInt16
    maximum (>= 16), Micro-warped tiles the bit-depth maximum; every level
    equals the oracle from the same inherited displacements."""
    import torch

    import qlpiv
    from qlpiv.synth import particle_pair

    a8, b8, _ = particle_pair(192, 192, seed=50 + bits, shift=(2.3, -1.2), core_radius=50.0,
                              u_max=3.0)
    scale = (1 << bits) - 1
    rng = np.random.default_rng(bits)
    ha = (a8.astype(np.float64) / 255.0 * scale + rng.uniform(0, 3, a8.shape)).clip(0, scale)
    hb = (b8.astype(np.float64) / 255.0 * scale + rng.uniform(0, 3, b8.shape)).clip(0, scale)
    ha, hb = ha.astype(np.uint16), hb.astype(np.uint16)
    ha[:40, :40] = 3  # a dim corner: maxima below 16 are clamped
    hb[:40, :40] = 5
    da = torch.from_numpy(ha.view(np.int16)).cuda().unsqueeze(0).contiguous()
    db = torch.from_numpy(hb.view(np.int16)).cuda().unsqueeze(0).contiguous()
    plan = qlpiv.PivPlan(ctx, 192, 192, 1, ia_start=64, ia_end=16, overlap_factor=0.5,
                         warp_mode=warp, warp_order=order, subpix_kind=qlpiv.QLP_SUBPIX_GUO,
                         subpix_pixels=3, subpix_iterations=20, keep_planes=2,
                         frame_type=qlpiv._lib.QLP_FRAME_U16, frame_bits=bits)
    say(f"u16-frame plan bits={bits} warp={warp} order={order}")
    plan.run(da, db)
    sides_a = {1: O.SIDE_A_FIRST, 2: O.SIDE_A_BOTH}
    sides_b = {0: O.SIDE_B_SECOND, 2: O.SIDE_B_BOTH}
    prev_u = prev_v = None
    for lv, L in enumerate(plan.levels):
        u, v, pk = (x[0] for x in plan.result(lv))
        planes = plane_copy(plan, lv)
        N, nI, nJ = L["ia"], L["nI"], L["nJ"]
        if lv == 0:
            uin = np.zeros((nI, nJ), np.float32)
            vin = np.zeros((nI, nJ), np.float32)
        else:
            uin = direct_inherit(prev_u, nI, nJ, round_it=(warp == 0))
            vin = direct_inherit(prev_v, nI, nJ, round_it=(warp == 0))
        c = np.float32((2 * N - 1) // 2)
        for i in range(nI):
            for j in range(nJ):
                t, l0 = int(L["top"][i, j]), int(L["left"][i, j])
                if lv > 0 and warp == 2 and order in sides_a:
                    tA, mA = O.warp_mini_frame(ha, L["top"], L["left"], N, uin, vin, i, j,
                                               sides_a[order]), float((1 << bits) - 1)
                elif lv > 0 and warp and order in sides_a:
                    tA, mA = O.warp_micro_u16(ha, bits, t, l0, N, uin[i, j], vin[i, j], sides_a[order])
                else:
                    tA, mA = O.clip_tile_u16(ha, t, l0, N)
                if lv > 0 and warp == 2 and order in sides_b:
                    tB, mB = O.warp_mini_frame(hb, L["top"], L["left"], N, uin, vin, i, j,
                                               sides_b[order]), float((1 << bits) - 1)
                elif lv > 0 and warp and order in sides_b:
                    tB, mB = O.warp_micro_u16(hb, bits, t, l0, N, uin[i, j], vin[i, j], sides_b[order])
                elif lv > 0 and not warp:
                    tb = int(np.floor(np.float32(t) + uin[i, j] + np.float32(0.5)))
                    lb = int(np.floor(np.float32(l0) + vin[i, j] + np.float32(0.5)))
                    tB, mB = O.clip_tile_u16(hb, tb, lb, N)
                else:
                    tB, mB = O.clip_tile_u16(hb, t, l0, N)
                plane = O.xcorr_javagpu(tA, tB, mA, mB)
                if (i * nJ + j) % 4 == 0:
                    assert np.array_equal(planes[i, j], plane), (lv, i, j)
                rp = O.find_max_from_center(plane)
                rs = O.guo(plane, 3, 20, rp[3], rp[:3])
                eu = np.float32(uin[i, j] + np.float32(rs[0] - c))
                ev = np.float32(vin[i, j] + np.float32(rs[1] - c))
                assert int(pk[i, j, 3]) == int(rs[3]), (lv, i, j)
                assert abs(u[i, j] - eu) <= 1e-4 and abs(v[i, j] - ev) <= 1e-4, (lv, i, j)
        say(f"level {lv}: N={N} {nI}x{nJ} equal")
        prev_u, prev_v = u, v
    plan.close()


def plane_copy(plan, lv):
    import ctypes

    import qlpiv

    L = plan.levels[lv]
    u, v, pk, pl = (ctypes.c_void_p() for _ in range(4))
    plan.ctx.synchronize()
    qlpiv.check(plan.ctx.lib.qlp_piv_result_dev(plan.h, lv, ctypes.byref(u), ctypes.byref(v),
                                                ctypes.byref(pk), ctypes.byref(pl)))
    D = 2 * L["ia"] - 1
    n = L["nI"] * L["nJ"]
    return qlpiv._fetch(pl, n * D * D, np.float32).reshape(L["nI"], L["nJ"], D, D)


def test_float_frame_plan_errors(ctx):
    import torch

    import qlpiv

    with pytest.raises(qlpiv.QlpError):
        qlpiv.PivPlan(ctx, 128, 128, 1, ia_start=32, ia_end=16, warp_mode=qlpiv._lib.QLP_WARP_FULL,
                      frame_type=qlpiv._lib.QLP_FRAME_U16, frame_bits=12)
    with pytest.raises(qlpiv.QlpError):
        qlpiv.PivPlan(ctx, 128, 128, 1, ia_start=32, ia_end=16, xcorr_form=1,
                      frame_type=qlpiv._lib.QLP_FRAME_F32)
    with pytest.raises(qlpiv.QlpError):
        qlpiv.PivPlan(ctx, 128, 128, 1, ia_start=32, ia_end=16,
                      frame_type=qlpiv._lib.QLP_FRAME_U16, frame_bits=14)
    plan = qlpiv.PivPlan(ctx, 128, 128, 1, ia_start=32, ia_end=16,
                         frame_type=qlpiv._lib.QLP_FRAME_F32)
    img8 = torch.zeros((1, 128, 128), dtype=torch.uint8, device="cuda")
    with pytest.raises(qlpiv.QlpError):
        qlpiv.check(ctx.lib.qlp_piv_run_dev(plan.h, img8.data_ptr(), img8.data_ptr()))
    plan.close()
