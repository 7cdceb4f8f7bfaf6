"""GPU tests of the vector validation (qlp_validate_vectors[_dev]) and of the
multi-pass loop with MultiPeaks and validation, against the CPU oracle.

The validation's integer and float work is reproduced operation for
operation, so fields, flags and statistics must be bit-identical to the
oracle's sequential spiral-order restatement."""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

MED, DIFF = 0, 1
R_NONE, R_SEC, R_BIL = 0, 1, 2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import qlpiv

    c = qlpiv.Context(0)
    yield c
    c.close()


def grid(nI, nJ, step=8, top0=3, left0=5):
    top = (top0 + step * np.arange(nI))[:, None].repeat(nJ, 1).astype(np.int32)
    left = (left0 + step * np.arange(nJ))[None, :].repeat(nI, 0).astype(np.int32)
    return top, left


def random_case(nI, nJ, P, seed, outliers=0.1, nan=False):
    """A smooth field plus planted outliers; peak increments where a planted
    outlier's true increment sits in a random secondary peak."""
    rng = np.random.default_rng(seed)
    i, j = np.meshgrid(np.arange(nI), np.arange(nJ), indexing="ij")
    base_u = (2.0 * np.sin(i / 7.0) + 0.3 * j / max(nJ, 1)).astype(np.float32)
    base_v = (1.5 * np.cos(j / 5.0) - 0.2 * i / max(nI, 1)).astype(np.float32)
    inc = rng.normal(0, 0.3, (nI, nJ, P, 2)).astype(np.float32)
    u = (base_u + inc[..., 0, 0]).astype(np.float32)
    v = (base_v + inc[..., 0, 1]).astype(np.float32)
    bad = rng.random((nI, nJ)) < outliers
    jump = rng.uniform(4, 12, (nI, nJ, 2)).astype(np.float32) * rng.choice([-1, 1], (nI, nJ, 2))
    inc[bad, 0] += jump[bad]
    u[bad] += jump[bad, 0]
    v[bad] += jump[bad, 1]
    if P > 1:
        k = rng.integers(1, P, (nI, nJ))
        for (a, b) in zip(*np.nonzero(bad)):
            inc[a, b, k[a, b]] = inc[a, b, 0] - jump[a, b]  # the true increment
    cnt = rng.integers(1, P + 1, (nI, nJ)).astype(np.int32)
    if nan:
        m = rng.random((nI, nJ)) < 0.02
        u[m] = np.nan
    return u, v, inc, cnt


def same(a, b):
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)],
                                                                      b[~np.isnan(b)])


CONFIGS = {
    "multipeak3": ([(MED, 2.0, 0.1, R_SEC, 1), (MED, 2.0, 0.1, R_SEC, 2),
                    (MED, 2.0, 0.1, R_BIL, 0)], 3),
    "median_bilinear": ([(MED, 2.0, 0.1, R_BIL, 0)], 1),
    "median_only": ([(MED, 2.0, 0.1, R_NONE, 0)], 1),
    "difference_bilinear": ([(DIFF, 1.0, 0.0, R_BIL, 0)], 1),
    "difference_secondary": ([(DIFF, 1.5, 0.0, R_SEC, 1), (MED, 3.0, 0.2, R_BIL, 0)], 2),
}

SHAPES = [(1, 1), (1, 7), (7, 1), (2, 2), (5, 8), (8, 5), (31, 31), (63, 63), (127, 127)]


@pytest.mark.parametrize("cfg", list(CONFIGS))
@pytest.mark.parametrize("nI,nJ", SHAPES)
def test_validation_equals_oracle(ctx, cfg, nI, nJ):
    import torch

    stages, P = CONFIGS[cfg]
    pairs = 3
    top, left = grid(nI, nJ)
    cases = [random_case(nI, nJ, P, seed=1000 * nI + nJ + 17 * k) for k in range(pairs)]
    for iters, until in [(0, False), (4, False), (0, True)]:
        say(f"validation {cfg} {nI}x{nJ} pairs={pairs} iters={iters} until={until}")
        u = torch.from_numpy(np.stack([c[0] for c in cases])).cuda()
        v = torch.from_numpy(np.stack([c[1] for c in cases])).cuda()
        inc = torch.from_numpy(np.stack([c[2] for c in cases])).cuda()
        cnt = torch.from_numpy(np.stack([c[3] for c in cases])).cuda()
        u, v, inv, st = ctx.validate_vectors(top, left, u, v, stages, inc, cnt, iters, until)
        ctx.synchronize()
        u, v, inv, st = (x.cpu().numpy() for x in (u, v, inv, st))
        for k, (cu, cv, ci, cc) in enumerate(cases):
            ru, rv, rinv, rst = O.validate(top, left, cu, cv, stages, ci, cc, iters, until)
            assert np.array_equal(u[k], ru) and np.array_equal(v[k], rv), (k, iters, until)
            assert np.array_equal(inv[k].astype(bool), rinv.astype(bool)), (k, iters, until)
            assert list(st[k]) == list(rst), (k, list(st[k]), list(rst))


@pytest.mark.parametrize("frac", [0.5, 1.0])
def test_validation_dense_outliers(ctx, frac):
    """Most or all vectors invalid: long chains of dependent visits."""
    import torch

    stages, P = CONFIGS["multipeak3"]
    nI, nJ = 40, 57
    top, left = grid(nI, nJ)
    u0, v0, inc, cnt = random_case(nI, nJ, P, seed=9, outliers=frac)
    say(f"validation dense frac={frac}")
    u, v, inv, st = ctx.validate_vectors(top, left, torch.from_numpy(u0).cuda(),
                                         torch.from_numpy(v0).cuda(), stages,
                                         torch.from_numpy(inc).cuda(),
                                         torch.from_numpy(cnt).cuda(), 3, False)
    ctx.synchronize()
    ru, rv, rinv, rst = O.validate(top, left, u0, v0, stages, inc, cnt, 3, False)
    assert np.array_equal(u.cpu().numpy(), ru) and np.array_equal(v.cpu().numpy(), rv)
    assert np.array_equal(inv.cpu().numpy().astype(bool), rinv.astype(bool))
    assert list(st.cpu().numpy()[0]) == list(rst)


def test_validation_with_nan_displacements(ctx):
    import torch

    stages, P = CONFIGS["multipeak3"]
    nI, nJ = 33, 30
    top, left = grid(nI, nJ)
    u0, v0, inc, cnt = random_case(nI, nJ, P, seed=4, nan=True)
    say("validation NaN field")
    u, v, inv, st = ctx.validate_vectors(top, left, torch.from_numpy(u0).cuda(),
                                         torch.from_numpy(v0).cuda(), stages,
                                         torch.from_numpy(inc).cuda(),
                                         torch.from_numpy(cnt).cuda(), 2, False)
    ctx.synchronize()
    ru, rv, rinv, rst = O.validate(top, left, u0, v0, stages, inc, cnt, 2, False)
    assert same(u.cpu().numpy(), ru) and same(v.cpu().numpy(), rv)
    assert np.array_equal(inv.cpu().numpy().astype(bool), rinv.astype(bool))
    assert list(st.cpu().numpy()[0]) == list(rst)


def test_validation_host_entry_and_job(ctx):
    """The host-pointer entry point and the VectorValidatorJob interface give
    the device entry point's results; bad configurations raise."""
    import qlpiv

    stages, P = CONFIGS["multipeak3"]
    top, left = grid(20, 24)
    u0, v0, inc, cnt = random_case(20, 24, P, seed=3)
    say("validation host entry")
    u, v, inv, st = ctx.validate_vectors(top, left, u0, v0, stages, inc, cnt, 1, False)
    ru, rv, rinv, rst = O.validate(top, left, u0, v0, stages, inc, cnt, 1, False)
    assert np.array_equal(u, ru) and np.array_equal(v, rv)
    assert np.array_equal(inv, rinv.astype(bool)) and list(st[0]) == list(rst)
    job = qlpiv.VectorValidatorJob(ctx, stages, max_iterations=1)
    job.set_input_parameters(qlpiv.JobResultEnum.JOB_RESULT_VALIDATOR_DATA,
                             dict(top=top, left=left, u=u0, v=v0, peak_uv=inc, peak_count=cnt))
    job.analyze()
    job.compute()
    out = job.get_job_result(qlpiv.JobResultEnum.JOB_RESULT_VALIDATED_VECTORS)
    assert np.array_equal(out["u"], ru) and np.array_equal(out["v"], rv)
    with pytest.raises(qlpiv.QlpError):
        ctx.validate_vectors(top, left, u0, v0, [(5, 2.0, 0.1, R_NONE, 0)])
    with pytest.raises(qlpiv.QlpError):
        ctx.validate_vectors(top, left, u0, v0, [(MED, 2.0, 0.1, R_SEC, 0)], inc)
    with pytest.raises(RuntimeError):
        qlpiv.VectorValidatorJob(ctx, None).analyze()


def direct_inherit(pu, nI, nJ):
    out = np.zeros((nI, nJ), np.float32)
    pI, pJ = pu.shape
    for i in range(nI):
        for j in range(nJ):
            if i // 2 < pI and j // 2 < pJ:
                out[i, j] = pu[i // 2, j // 2]
    return out


def test_multipass_multipeak_validation_against_oracle(ctx):
    """The device loop with MultiPeaks (P = 3, k = 3), Guo fits of every peak
    and the multi-peak normalised median validation after every level.  Each
    stage is checked against the oracle from the previous stage's device
    output: planes bit-exact, peaks bit-exact, fits within 1e-4 px (FP64
    log/exp), and the validated field bit-exact from the device's fits."""
    import torch

    import qlpiv
    from qlpiv.synth import particle_pair

    a, b, _ = particle_pair(256, 256, seed=31, shift=(2.2, -1.4), core_radius=60.0, u_max=5.0)
    stages = qlpiv._lib.multipeak_median_stages(3, 2.0, 0.1)
    plan = qlpiv.PivPlan(ctx, 256, 256, 1, ia_start=64, ia_end=16, warp_mode=1,
                         subpix_kind=qlpiv.QLP_SUBPIX_GUO, subpix_pixels=3, subpix_iterations=20,
                         peak_kind=qlpiv.QLP_PEAK_MULTI, n_peaks=3, peak_kernel_width=3,
                         validation=stages, keep_planes=2)
    say("multipass multipeak + validation")
    plan.run(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    prev_u = prev_v = None
    total_invalid = 0
    for lv, L in enumerate(plan.levels):
        u, v, sub = (x[0] for x in plan.result(lv))
        inv, st = plan.validation(lv)
        N, nI, nJ = L["ia"], L["nI"], L["nJ"]
        D = 2 * N - 1
        c = np.float32(D // 2)
        uin = np.zeros((nI, nJ), np.float32) if lv == 0 else direct_inherit(prev_u, nI, nJ)
        vin = np.zeros((nI, nJ), np.float32) if lv == 0 else direct_inherit(prev_v, nI, nJ)
        # planes of this level from the device, checked against the oracle on a sample
        planes = plan_planes(plan, lv)
        cnt = np.zeros((nI, nJ), np.int32)
        inc = np.zeros((nI, nJ, 3, 2), np.float32)
        for i in range(nI):
            for j in range(nJ):
                t, l0 = int(L["top"][i, j]), int(L["left"][i, j])
                if (i * nJ + j) % 7 == 0:
                    tA, mA = O.clip_tile(a, t, l0, N)
                    if lv == 0:
                        tB, mB = O.clip_tile(b, t, l0, N)
                    else:
                        tB, mB = O.warp_micro(b, t, l0, N, uin[i, j], vin[i, j]), 255.0
                    ref = O.xcorr_javagpu(tA.astype(np.float32), tB.astype(np.float32), mA, mB)
                    assert np.array_equal(planes[i, j], ref), (lv, i, j)
                pk, found, mf = O.find_max_multipeaks(planes[i, j], 3, 3)
                cnt[i, j] = max(found, 1)
                for k in range(3):
                    if k < found:
                        rs = O.guo(planes[i, j], 3, 20, mf, pk[k])
                        assert int(sub[i, j, k, 3]) == int(rs[3]), (lv, i, j, k)
                        assert abs(sub[i, j, k, 0] - rs[0]) <= 1e-4, (lv, i, j, k)
                        assert abs(sub[i, j, k, 1] - rs[1]) <= 1e-4, (lv, i, j, k)
                    inc[i, j, k] = (np.float32(sub[i, j, k, 0] - c), np.float32(sub[i, j, k, 1] - c))
        # the validation from the device's own fits
        upre = (uin + inc[..., 0, 0]).astype(np.float32)
        vpre = (vin + inc[..., 0, 1]).astype(np.float32)
        ru, rv, rinv, rst = O.validate(L["top"], L["left"], upre, vpre, stages, inc, cnt, 0, False)
        assert np.array_equal(u, ru) and np.array_equal(v, rv), lv
        assert np.array_equal(inv[0], rinv.astype(bool)) and list(st[0]) == list(rst), lv
        say(f"level {lv}: N={N} {nI}x{nJ} corrected={rst[1]} invalid={rst[2]}")
        total_invalid += int(rst[2])
        prev_u, prev_v = u, v
    plan.close()


def plan_planes(plan, lv):
    """Host copy of a level's stored planes (nI x nJ x D x D)."""
    import ctypes

    import qlpiv

    L = plan.levels[lv]
    u, v, pk, pl = (ctypes.c_void_p() for _ in range(4))
    plan.ctx.synchronize()
    qlpiv.check(plan.ctx.lib.qlp_piv_result_dev(plan.h, lv, ctypes.byref(u), ctypes.byref(v),
                                                ctypes.byref(pk), ctypes.byref(pl)))
    D = 2 * L["ia"] - 1
    out = np.empty(L["nI"] * L["nJ"] * D * D, np.float32)
    qlpiv._hip_copy_to_host(out, pl.value)
    return out.reshape(L["nI"], L["nJ"], D, D)
