"""GPU tests of the serving form (qlp_server): requests of mixed frame sizes,
IA sizes, level counts and correlation forms run concurrently on several slots
under an HBM budget, and each result equals the same request run alone."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _frames():
    from qlpiv.synth import particle_pair

    out = {}
    for k, (H, W) in enumerate(((256, 256), (384, 512), (512, 256), (512, 512))):
        out[(H, W)] = particle_pair(H, W, seed=70 + k, shift=(1.3, -0.8), core_radius=60.0,
                                    u_max=4.0)[:2]
    return out


def _requests():
    import qlpiv

    reqs = []
    shapes = ((256, 256), (384, 512), (512, 256), (512, 512))
    configs = [dict(ia_start=64, ia_end=16), dict(ia_start=32, ia_end=32),
               dict(ia_start=64, ia_end=32, subpix_kind=qlpiv.QLP_SUBPIX_GAUSS1D),
               dict(ia_start=32, ia_end=16, xcorr_form=qlpiv.QLP_XCORR_MFMA_F16),
               dict(ia_start=64, ia_end=16, peak_kind=qlpiv.QLP_PEAK_MULTI, n_peaks=3,
                    validation=qlpiv._lib.multipeak_median_stages(3, 2.0, 0.1))]
    for k in range(12):
        reqs.append((shapes[k % 4], configs[k % 5]))
    return reqs


def _alone(ctx, a, b, cfg):
    import torch

    import qlpiv

    plan = qlpiv.PivPlan(ctx, a.shape[0], a.shape[1], 1, **cfg)
    plan.run(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
    u, v, _ = plan.result(len(plan.levels) - 1)
    plan.close()
    return u[0], v[0]


def test_results_in_the_mix_equal_alone():
    import qlpiv

    frames = _frames()
    reqs = _requests()
    ctx = qlpiv.Context(0)
    alone = []
    for (shape, cfg) in reqs:
        a, b = frames[shape]
        alone.append(_alone(ctx, a, b, cfg))
    ctx.close()
    say("alone results computed")
    srv = qlpiv.Server(0, hbm_budget=4 << 30, slots=4)
    tickets = [srv.submit(*frames[shape], **cfg) for (shape, cfg) in reqs]
    for k, t in enumerate(tickets):
        assert srv.wait(t, 60000), k
        u, v, ia, times = srv.result(t)
        ua, va = alone[k]
        say(f"request {k}: {reqs[k][0]} ia {ia} latency {times['done_ms'] - times['submit_ms']:.2f} ms")
        assert u.shape == ua.shape and np.array_equal(u, ua) and np.array_equal(v, va), k
        assert times["submit_ms"] <= times["admit_ms"] <= times["done_ms"]
    st = srv.stats()
    say("stats", st)
    assert st["completed"] == len(reqs) and st["failed"] == 0 and st["pending"] == 0
    assert st["hbm_peak"] <= 4 << 30
    srv.close()


def test_tight_budget_admission_and_eviction():
    """A budget that holds about one request's buffers: requests wait for memory,
    idle plans are evicted, every request still completes with the alone result."""
    import qlpiv

    frames = _frames()
    cfgs = [dict(ia_start=64, ia_end=16, keep_planes=2), dict(ia_start=32, ia_end=16, keep_planes=2)]
    shapes = [(512, 512), (384, 512)]
    need = max(qlpiv.plan_estimate(H, W, **c) + 2 * H * W for (H, W) in shapes for c in cfgs)
    budget = int(need * 1.2)
    ctx = qlpiv.Context(0)
    jobs = [(shapes[k % 2], cfgs[(k // 2) % 2]) for k in range(8)]
    alone = [_alone(ctx, *frames[s], c) for (s, c) in jobs]
    ctx.close()
    srv = qlpiv.Server(0, hbm_budget=budget, slots=3)
    tickets = [srv.submit(*frames[s], **c) for (s, c) in jobs]
    for k, t in enumerate(tickets):
        assert srv.wait(t, 60000), k
        u, v, _, _ = srv.result(t)
        assert np.array_equal(u, alone[k][0]) and np.array_equal(v, alone[k][1]), k
    st = srv.stats()
    say("stats", st, "budget", budget)
    assert st["completed"] == len(jobs) and st["hbm_peak"] <= budget and st["evictions"] > 0
    srv.close()


def test_device_frames_and_errors():
    import torch

    import qlpiv

    frames = _frames()
    a, b = frames[(256, 256)]
    srv = qlpiv.Server(0, hbm_budget=1 << 30, slots=2)
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    torch.cuda.synchronize()
    t = srv.submit(da, db, ia_start=32, ia_end=16)
    assert srv.wait(t, 60000)
    u, v, ia, _ = srv.result(t)
    ctx = qlpiv.Context(0)
    ua, va = _alone(ctx, a, b, dict(ia_start=32, ia_end=16))
    ctx.close()
    assert ia == 16 and np.array_equal(u, ua) and np.array_equal(v, va)
    # a configuration the library rejects, and a request larger than the budget
    with pytest.raises(qlpiv.QlpError):
        srv.submit(a, b, ia_start=48, ia_end=16)
    small = qlpiv.Server(0, hbm_budget=1 << 20, slots=1)
    with pytest.raises(qlpiv.QlpError) as e:
        small.submit(a, b, ia_start=64, ia_end=16, keep_planes=2)
    assert e.value.code == qlpiv._lib.QLP_ERR_BUDGET
    small.close()
    srv.close()


def test_c5_full_size_mix_equals_alone():
    """The serving configuration at its sizes: frames 512^2, 1024^2, 1024x2048
    and 2048^2, final IA 16 / 32 / 64 with 1-3 levels, micro warp, FromCenter,
    Guo 3 px, a 64 GB budget; 36 requests submitted in bursts while earlier
    ones run (continuous batching), in the exact and the MFMA form.  Every
    result equals the same request run alone through a plan, bit for bit; two
    of the plans are held level by level to the oracle."""
    import time

    import torch

    import qlpiv
    from plan_oracle import check_plan_levels
    from qlpiv.synth import particle_pair

    sizes = [(512, 512), (1024, 1024), (1024, 2048), (2048, 2048)]
    frames = {}
    for k, (H, W) in enumerate(sizes):
        a, b, _ = particle_pair(H, W, seed=1500 + k, shift=(3.3, 3.3), core_radius=40.0 + 50 * k,
                                u_max=2.0 + 2 * k)
        frames[(H, W)] = (a, b)
    rng = np.random.default_rng(1600)
    jobs = []
    for k in range(36):
        H, W = sizes[int(rng.integers(4))]
        ia_end = int(rng.choice([16, 32, 64]))
        levels = int(rng.integers(1, {16: 3, 32: 2, 64: 1}[ia_end] + 1))
        form = qlpiv.QLP_XCORR_MFMA_F16 if k % 3 == 2 else qlpiv.QLP_XCORR_EXACT
        jobs.append(((H, W), dict(ia_start=ia_end << (levels - 1), ia_end=ia_end, warp_mode=1,
                                  subpix_kind=qlpiv.QLP_SUBPIX_GUO, subpix_pixels=3,
                                  subpix_iterations=20, xcorr_form=form)))
    ctx = qlpiv.Context(0)
    alone = {}
    for shape, cfg in jobs:
        key = (shape, tuple(sorted(cfg.items())))
        if key not in alone:
            alone[key] = _alone(ctx, *frames[shape], cfg)
    say(f"C5: {len(alone)} distinct requests run alone")
    # two of the plans, level by level against the oracle
    for shape, cfg in (((2048, 2048), dict(ia_start=64, ia_end=16)), ((1024, 2048), dict(ia_start=32, ia_end=16))):
        a, b = frames[shape]
        plan = qlpiv.PivPlan(ctx, shape[0], shape[1], 1, warp_mode=1, subpix_kind=1, subpix_pixels=3,
                             subpix_iterations=20, **cfg)
        plan.run(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
        check_plan_levels(plan, a, b, warp_mode=1, norm_mode=0, peak_kind=0, fit_kind=1, log=say)
        plan.close()
    ctx.close()
    pinned = {s: tuple(torch.from_numpy(x).pin_memory().numpy() for x in frames[s]) for s in sizes}
    srv = qlpiv.Server(0, hbm_budget=64 << 30, slots=2)
    tickets = []
    for burst in range(4):
        for shape, cfg in jobs[burst * 9:(burst + 1) * 9]:
            tickets.append(srv.submit(*pinned[shape], **cfg))
        time.sleep(0.002)  # the next burst arrives while these run
    lat = []
    for k, t in enumerate(tickets):
        assert srv.wait(t, 120000), k
        u, v, ia, times = srv.result(t)
        shape, cfg = jobs[k]
        ua, va = alone[(shape, tuple(sorted(cfg.items())))]
        assert ia == cfg["ia_end"]
        assert np.array_equal(u, ua) and np.array_equal(v, va), (k, shape, cfg)
        lat.append(times["done_ms"] - times["submit_ms"])
    st = srv.stats()
    say(f"C5: {len(jobs)} requests equal alone; latency p50 {np.percentile(lat, 50):.2f} ms, "
        f"p99 {np.percentile(lat, 99):.2f} ms; stats {st}")
    assert st["completed"] == len(jobs) and st["failed"] == 0 and st["hbm_peak"] <= 64 << 30
    srv.close()
