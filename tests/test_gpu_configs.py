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
