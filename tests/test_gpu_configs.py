"""The reference's CPU-path configuration at its full size on the GPU.

One synthetic 1024 x 1024 pair (Gaussian particles d = 2.5 px at 0.05 ppp,
I0 ~ U(150, 255), noise sigma 2; frame B advected by a Rankine vortex R = 128 px,
U = 6 px), 32 x 32 IAs at 50 % overlap (63 x 63 = 3,969 IAs), the Java CPU
path's arithmetic (both tiles normalised by A's maximum, inverse twiddles),
FromCenter integer peak and 3-point Gaussian 1D-1D fit, single pass, no warp:
planes value-identical to the oracle's restatement of the Java path, every
integer peak bit-exact, displacements within 1e-4 px.
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

SUBPIX_TOL = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_reference_cpu_path_configuration_full_size():
    import torch

    import qlpiv
    from qlpiv.geometry import level_grid
    from qlpiv.synth import particle_pair

    H = W = 1024
    N = 32
    a, b, _ = particle_pair(H, W, seed=1000, shift=(0.0, 0.0), core_radius=128.0, u_max=6.0)
    tops, lefts = level_grid(H, W, N, 0.5)
    assert tops.shape == (63, 63)
    top = tops.ravel().astype(np.int32)
    left = lefts.ravel().astype(np.int32)
    n = top.size
    tA = np.empty((n, N, N), np.uint8)
    tB = np.empty((n, N, N), np.uint8)
    mA = np.empty(n, np.float32)
    mB = np.empty(n, np.float32)
    for k in range(n):
        tA[k], mA[k] = O.clip_tile(a, top[k], left[k], N)
        tB[k], mB[k] = O.clip_tile(b, top[k], left[k], N)
    ctx = qlpiv.Context(0)
    say(f"C1 planes: {n} IAs")
    planes = ctx.xcorr(tA, tB, mA, mB, qlpiv.QLP_NORM_A_MAX)
    ref = O.xcorr_batch_u8(tA, tB, mA, mB, 0)
    assert np.array_equal(planes, ref)
    peaks, mf = ctx.find_peaks(planes)
    say("C1 plan: single pass, Gaussian 3 px")
    plan = qlpiv.PivPlan(ctx, H, W, 1, ia_start=N, ia_end=N, overlap_factor=0.5, warp_mode=0,
                         norm_mode=qlpiv.QLP_NORM_A_MAX, subpix_kind=qlpiv.QLP_SUBPIX_GAUSS1D,
                         subpix_pixels=3)
    plan.run(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    u, v, sub = (x[0].reshape(n, -1) for x in plan.result(0))
    u, v = u.ravel(), v.ravel()
    c = np.float32(N - 1)
    worst = 0.0
    for k in range(n):
        rp = O.find_max_from_center(ref[k])
        assert tuple(peaks[k, 0]) == tuple(rp[:3]) and mf[k] == rp[3], k
        rs = O.gauss1d(ref[k], 3, rp[3], rp[:3])
        assert int(sub[k, 3]) == int(rs[3]), k
        eu, ev = np.float32(rs[0] - c), np.float32(rs[1] - c)
        worst = max(worst, abs(float(u[k] - eu)), abs(float(v[k] - ev)))
    say(f"C1 worst displacement difference {worst:.2e} px")
    assert worst <= SUBPIX_TOL
    plan.close()
    ctx.close()


def test_c2_full_size_every_level_against_oracle():
    """The plain FP32 form's configuration at its full size: the bench's 2048 x
    2048 pair (uniform 3.3 px shift plus a Rankine vortex R = 128 px, U = 6 px)
    through the multi-pass loop 64 -> 32 -> 16 px at 50 % overlap (3,969 +
    16,129 + 65,025 IAs), micro warp of B, per-tile-normalised correlation
    storing every plane, MultiPeaks 3 peaks / 3 x 3, Guo 3 px / 20 iterations on
    every peak, and the multi-peak normalised median validation (2.0, 0.1) with
    secondary-peak and bilinear replacement.  Every level against the oracle
    from the device's inherited field: all 85,123 planes value-identical, fit
    status equal and fits within 1e-4 px on every peak, the validated field,
    flags and statistics bit-exact."""
    import torch

    import qlpiv
    from plan_oracle import check_plan_levels
    from qlpiv.synth import particle_pair

    H = W = 2048
    a, b, _ = particle_pair(H, W, seed=1100, shift=(3.3, 3.3), core_radius=128.0, u_max=6.0)
    stages = qlpiv._lib.multipeak_median_stages(3, 2.0, 0.1)
    ctx = qlpiv.Context(0)
    plan = qlpiv.PivPlan(ctx, H, W, 1, ia_start=64, ia_end=16, overlap_factor=0.5,
                         warp_mode=qlpiv._lib.QLP_WARP_MICRO, norm_mode=qlpiv.QLP_NORM_PER_TILE,
                         subpix_kind=qlpiv.QLP_SUBPIX_GUO, subpix_pixels=3, subpix_iterations=20,
                         keep_planes=2, peak_kind=qlpiv.QLP_PEAK_MULTI, n_peaks=3,
                         peak_kernel_width=3, validation=stages, valid_max_iterations=0)
    assert [L["nI"] * L["nJ"] for L in plan.levels] == [3969, 16129, 65025]
    say("C2 full size: 85,123 IAs on the device, then the oracle level by level")
    plan.run(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    check_plan_levels(plan, a, b, warp_mode=1, norm_mode=0, peak_kind=1, n_peaks=3,
                      kernel_width=3, fit_kind=1, pixels=3, iterations=20, stages=stages,
                      max_iterations=0, log=say)
    plan.close()
    ctx.close()
