"""GPU tests of the fused correlation pass and of the device-resident multi-pass
loop, against the separate kernels and the CPU oracle."""
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
def pair256():
    from qlpiv.synth import particle_pair

    a, b, field = particle_pair(256, 256, seed=77, shift=(1.3, -0.7), core_radius=60.0, u_max=4.0)
    return a, b, field


@pytest.mark.parametrize("N", [16, 32, 64])
@pytest.mark.parametrize("warp", [0, 1])
@pytest.mark.parametrize("mode", [0, 1])
def test_fused_pass_equals_separate_kernels(ctx, pair256, N, warp, mode):
    """clip + (micro-warp | shifted clip) + xcorr + FromCenter fused in one kernel
    gives the planes, peaks and floors of the separate kernels, which equal the
    oracle (tested elsewhere)."""
    import torch

    import qlpiv
    from qlpiv.geometry import level_grid

    a, b, _ = pair256
    tops, lefts = level_grid(256, 256, N, 0.5)
    top = tops.ravel().astype(np.int32)
    left = lefts.ravel().astype(np.int32)
    n = top.size
    rng = np.random.default_rng(N + warp)
    du = rng.uniform(-5, 5, n).astype(np.float32)
    dv = rng.uniform(-5, 5, n).astype(np.float32)
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    dt, dl = torch.from_numpy(top).cuda(), torch.from_numpy(left).cuda()
    ddu, ddv = torch.from_numpy(du).cuda(), torch.from_numpy(dv).cuda()
    qmode = qlpiv.QLP_NORM_A_MAX if mode == 0 else qlpiv.QLP_NORM_PER_TILE
    say(f"fused pass N={N} warp={warp} mode={mode}")
    pl, pk, mf, nb = ctx.xcorr_images(da, db, dt, dl, N, ddu, ddv, warp_mode=warp, mode=qmode,
                                      planes=True, nbr_half=3)
    ctx.synchronize()
    pl, pk, mf, nb = (x.cpu().numpy() for x in (pl, pk, mf, nb))
    # the same pass from oracle tiles
    tA = np.empty((n, N, N), np.uint8)
    tB = np.empty((n, N, N), np.uint8)
    mA = np.empty(n, np.float32)
    mB = np.empty(n, np.float32)
    for k in range(n):
        tA[k], mA[k] = O.clip_tile(a, top[k], left[k], N)
        if warp:
            tB[k] = O.warp_micro(b, top[k], left[k], N, du[k], dv[k])
            mB[k] = 255.0
        else:
            t0 = int(np.floor(np.float32(top[k]) + du[k] + np.float32(0.5)))
            l0 = int(np.floor(np.float32(left[k]) + dv[k] + np.float32(0.5)))
            tB[k], mB[k] = O.clip_tile(b, t0, l0, N)
    ref = O.xcorr_batch_u8(tA, tB, mA, mB, mode)
    assert np.array_equal(pl, ref)
    D = 2 * N - 1
    for k in range(n):
        rp = O.find_max_from_center(ref[k])
        assert tuple(pk[k]) == tuple(rp[:3]) and mf[k] == rp[3], k
        pi, pj = int(rp[0]), int(rp[1])
        for d in range(-3, 4):
            if 0 <= pi + d < D:
                assert nb[k, 0, d + 3] == ref[k, pi + d, pj]
            if 0 <= pj + d < D:
                assert nb[k, 1, d + 3] == ref[k, pi, pj + d]


def direct_inherit(pu, nI, nJ, round_it=False):
    out = np.zeros((nI, nJ), np.float32)
    pI, pJ = pu.shape
    for i in range(nI):
        for j in range(nJ):
            if i // 2 < pI and j // 2 < pJ:
                x = pu[i // 2, j // 2]
                out[i, j] = np.float32(np.floor(x + np.float32(0.5))) if round_it else x
    return out


@pytest.mark.parametrize("warp,ia_start", [(1, 64), (0, 64), (1, 128)])
def test_multipass_levels_against_oracle(ctx, pair256, warp, ia_start):
    """Every level of the device loop equals the oracle restatement of that level
    run from the same inherited displacements: integer peaks bit-exact, the
    updated displacements within 1e-4 px (the Guo fit's FP64 log/exp)."""
    import torch

    import qlpiv

    a, b, _ = pair256
    plan = qlpiv.PivPlan(ctx, 256, 256, 1, ia_start=ia_start, ia_end=16, overlap_factor=0.5,
                         warp_mode=warp, subpix_kind=qlpiv.QLP_SUBPIX_GUO, subpix_pixels=3,
                         subpix_iterations=20)
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    say(f"multipass warp={warp} ia_start={ia_start}")
    plan.run(da, db)
    prev_u = prev_v = None
    for lv, L in enumerate(plan.levels):
        u, v, pk = plan.result(lv)
        u, v, pk = u[0], v[0], pk[0]
        N, nI, nJ = L["ia"], L["nI"], L["nJ"]
        if lv == 0:
            uin = np.zeros((nI, nJ), np.float32)
            vin = np.zeros((nI, nJ), np.float32)
        else:
            uin = direct_inherit(prev_u, nI, nJ, round_it=(warp == 0))
            vin = direct_inherit(prev_v, nI, nJ, round_it=(warp == 0))
        D = 2 * N - 1
        c = np.float32(D // 2)
        for i in range(nI):
            for j in range(nJ):
                t, l0 = int(L["top"][i, j]), int(L["left"][i, j])
                tA, mA = O.clip_tile(a, t, l0, N)
                if lv == 0:
                    tB, mB = O.clip_tile(b, t, l0, N)
                elif warp:
                    tB, mB = O.warp_micro(b, t, l0, N, uin[i, j], vin[i, j]), 255.0
                else:
                    tb = int(np.floor(np.float32(t) + uin[i, j] + np.float32(0.5)))
                    lb = int(np.floor(np.float32(l0) + vin[i, j] + np.float32(0.5)))
                    tB, mB = O.clip_tile(b, tb, lb, N)
                plane = O.xcorr_javagpu(tA.astype(np.float32), tB.astype(np.float32), mA, mB)
                rp = O.find_max_from_center(plane)
                rs = O.guo(plane, 3, 20, rp[3], rp[:3])
                eu = np.float32(uin[i, j] + np.float32(rs[0] - c))
                ev = np.float32(vin[i, j] + np.float32(rs[1] - c))
                assert int(pk[i, j, 3]) == int(rs[3]), (lv, i, j)
                assert abs(u[i, j] - eu) <= 1e-4 and abs(v[i, j] - ev) <= 1e-4, (lv, i, j)
        prev_u, prev_v = u, v
    plan.close()


def test_multipass_recovers_the_synthetic_field(ctx):
    """End to end on a 512^2 Rankine pair with a uniform shift: the last level
    (16 px IAs) matches the generating field away from the borders."""
    import torch

    import qlpiv
    from qlpiv.synth import particle_pair

    a, b, field = particle_pair(512, 512, seed=11, shift=(3.3, 3.3), core_radius=128.0, u_max=6.0)
    plan = qlpiv.PivPlan(ctx, 512, 512, 1, ia_start=64, ia_end=16, warp_mode=1,
                         subpix_kind=qlpiv.QLP_SUBPIX_GUO)
    plan.run(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    u, v, _ = plan.result(-1)
    L = plan.levels[-1]
    N = L["ia"]
    cy = L["top"] + N / 2.0 - 0.5
    cx = L["left"] + N / 2.0 - 0.5
    fx, fy = field(cx, cy)
    inner = (slice(4, -4), slice(4, -4))
    err = np.hypot(u[0][inner] - fy[inner], v[0][inner] - fx[inner])
    say(f"field error: median {np.median(err):.3f} px, p95 {np.percentile(err, 95):.3f} px")
    assert np.median(err) < 0.15 and np.percentile(err, 95) < 0.6
    plan.close()


def test_plan_errors(ctx):
    import qlpiv

    with pytest.raises(qlpiv.QlpError):
        qlpiv.PivPlan(ctx, 256, 256, 1, ia_start=64, ia_end=24)
    with pytest.raises(qlpiv.QlpError):
        qlpiv.PivPlan(ctx, 256, 256, 1, overlap_factor=0.95)
    small = qlpiv.Context(0, hbm_budget=1 << 16)
    with pytest.raises(qlpiv.QlpError) as e:
        qlpiv.PivPlan(small, 2048, 2048, 4)
    assert e.value.code == qlpiv._lib.QLP_ERR_BUDGET
    small.close()


PLAN_CASES = [
    # name, plan options, oracle options
    ("simple-gauss5-micro",
     dict(warp_mode=1, peak_kind=2, subpix_kind=0, subpix_pixels=5),
     dict(warp_mode=1, norm_mode=0, peak_kind=2, fit_kind=0, pixels=5)),
    ("fromcenter-guo-start1-micro",
     dict(warp_mode=1, subpix_kind=1, subpix_pixels=3, subpix_iterations=20,
          interpolator_start_step=1),
     dict(warp_mode=1, norm_mode=0, peak_kind=0, fit_kind=1, start_step=1)),
    ("multi-poly-start2-shift-amax",
     dict(warp_mode=0, norm_mode=1, peak_kind=1, n_peaks=3, peak_kernel_width=3, subpix_kind=2,
          interpolator_start_step=2),
     dict(warp_mode=0, norm_mode=1, peak_kind=1, n_peaks=3, fit_kind=2, start_step=2)),
    ("simple-guo7-start1-shift",
     dict(warp_mode=0, peak_kind=2, subpix_kind=1, subpix_pixels=7, subpix_iterations=20,
          interpolator_start_step=1, keep_planes=2),
     dict(warp_mode=0, norm_mode=0, peak_kind=2, fit_kind=1, pixels=7, start_step=1)),
]


@pytest.mark.parametrize("case", PLAN_CASES, ids=[c[0] for c in PLAN_CASES])
def test_plan_finders_fits_and_start_step(ctx, pair256, case):
    """The device loop with each peak finder (Simple, FromCenter, MultiPeaks),
    each fit and an interpolator start step above 0 (the levels before it keep
    the integer peak, CompositeLastLevelInterpolator / LocalPIVOpenCLGpuJob
    :373), level by level against the oracle."""
    import torch

    import qlpiv
    from plan_oracle import check_plan_levels

    name, popt, oopt = case
    a, b, _ = pair256
    plan = qlpiv.PivPlan(ctx, 256, 256, 1, ia_start=64, ia_end=16, overlap_factor=0.5, **popt)
    say(f"plan {name}")
    plan.run(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    check_plan_levels(plan, a, b, log=say, **oopt)
    plan.close()


@pytest.mark.parametrize("pixels", [3, 5, 7])
@pytest.mark.parametrize("form", ["exact", "mfma_f16", "mfma_bf16"])
def test_multipeaks_entries_in_plans_equal_segment_form(ctx, form, pixels):
    """MultiPeaks in a multi-pass plan, with the block maxima written by the
    correlation passes (exact: from the LDS plane; matrix-core forms: from the
    plane each pass has just stored) and replayed one lane per plane, gives at
    every level the u, v, peak records, validation and stored planes of the same
    plan with the segment form searching the stored planes
    (QLPIV_MULTIPEAK_FORM=segment): both replay the reference's list on the same
    planes; Guo fits of 3, 5 and 7 pixels follow."""
    import os

    import torch

    import qlpiv
    from qlpiv.synth import particle_pair

    arith = {"exact": qlpiv.QLP_XCORR_EXACT, "mfma_f16": qlpiv.QLP_XCORR_MFMA_F16,
             "mfma_bf16": qlpiv.QLP_XCORR_MFMA_BF16}[form]
    a, b, _ = particle_pair(512, 384, seed=41, shift=(2.6, -1.2), core_radius=80.0, u_max=5.0)
    cfg = dict(ia_start=64, ia_end=16, warp_mode=1, peak_kind=qlpiv.QLP_PEAK_MULTI, n_peaks=3,
               peak_kernel_width=3, xcorr_form=arith, subpix_kind=qlpiv.QLP_SUBPIX_GUO,
               subpix_pixels=pixels, validation=qlpiv._lib.multipeak_median_stages(3, 2.0, 0.1))
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    out = {}
    saved = os.environ.get("QLPIV_MULTIPEAK_FORM")
    try:
        for name in ("entries", "segment"):
            if name == "segment":
                os.environ["QLPIV_MULTIPEAK_FORM"] = "segment"
            else:
                os.environ.pop("QLPIV_MULTIPEAK_FORM", None)
            say(f"multipeaks plan {form} guo {pixels} px {name}")
            plan = qlpiv.PivPlan(ctx, 512, 384, 1, **cfg)
            plan.run(da, db)
            out[name] = [plan.result(lv) for lv in range(len(plan.levels))]
            out[name + "_valid"] = [plan.validation(lv) for lv in range(len(plan.levels))]
            out[name + "_planes"] = [plan.planes(lv) for lv in range(len(plan.levels))]
            plan.close()
    finally:
        if saved is None:
            os.environ.pop("QLPIV_MULTIPEAK_FORM", None)
        else:
            os.environ["QLPIV_MULTIPEAK_FORM"] = saved
    for lv, (r1, r2) in enumerate(zip(out["entries"], out["segment"])):
        for k in range(3):
            assert np.array_equal(r1[k], r2[k]), (form, lv, k)
        v1, v2 = out["entries_valid"][lv], out["segment_valid"][lv]
        assert np.array_equal(v1[0], v2[0]) and np.array_equal(v1[1], v2[1]), (form, lv)
        # the stored planes too (the 32 px matrix-core pass stores them band by
        # band from LDS when it writes the entries)
        assert np.array_equal(out["entries_planes"][lv], out["segment_planes"][lv]), (form, lv)


def test_multipeaks_plan_without_room_for_the_entries(ctx):
    """A context budget that holds a MultiPeaks plan's buffers but not the
    entries workspace (qlp_piv_plan_estimate counts it): the plan is created and
    runs on the segment form, with the results of the same plan on a context
    with room for the workspace."""
    import torch

    import qlpiv
    from qlpiv.synth import particle_pair

    a, b, _ = particle_pair(512, 384, seed=43, shift=(1.6, 2.2), core_radius=80.0, u_max=5.0)
    cfg = dict(ia_start=64, ia_end=16, warp_mode=1, peak_kind=qlpiv.QLP_PEAK_MULTI, n_peaks=3,
               peak_kernel_width=3, validation=qlpiv._lib.multipeak_median_stages(3, 2.0, 0.1))
    est = qlpiv.plan_estimate(512, 384, **cfg)
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    res = {}
    small = qlpiv.Context(0, hbm_budget=est - 1)
    try:
        for name, c in (("tight", small), ("room", ctx)):
            say(f"multipeaks plan budget {name} (estimate {est} bytes)")
            plan = qlpiv.PivPlan(c, 512, 384, 1, **cfg)
            plan.run(da, db)
            res[name] = [plan.result(lv) for lv in range(len(plan.levels))]
            plan.close()
    finally:
        small.close()
    for lv, (r1, r2) in enumerate(zip(res["tight"], res["room"])):
        for k in range(3):
            assert np.array_equal(r1[k], r2[k]), (lv, k)
