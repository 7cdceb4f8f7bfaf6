"""Mixed-size batches in one graph launch (qlp_batch) against one-level plans.

Every pair's IAs must give the same values through the batch (descriptor clip,
tiles, tuned kernel shape, graph replay) as through a one-level plan of that
pair alone: displacements and sub-pixel records equal (float ==).  The plans
themselves are held to the oracle in test_gpu_pipeline.py and
test_gpu_ia128.py.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PAIRS = [(256, 256, 16), (192, 320, 32), (256, 256, 64), (384, 384, 128), (200, 240, 16),
         (300, 260, 32), (128, 512, 64)]
CFG = dict(overlap_factor=0.5, subpix_kind=1, subpix_pixels=3, subpix_iterations=20)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import qlpiv

    c = qlpiv.Context(0)
    yield c
    c.close()


def frames(seed):
    from qlpiv.synth import particle_pair

    out = []
    for k, (H, W, _) in enumerate(PAIRS):
        a, b, _ = particle_pair(H, W, seed=seed + k, shift=(1.5 + 0.3 * k, -2.0 + 0.2 * k),
                                core_radius=min(H, W) / 3.0, u_max=3.0 + k % 3)
        out.append((a, b))
    return out


def plan_results(ctx, fr, **extra):
    import torch

    import qlpiv

    res = []
    for (H, W, ia), (a, b) in zip(PAIRS, fr):
        plan = qlpiv.PivPlan(ctx, H, W, 1, ia_start=ia, ia_end=ia, warp_mode=0, **CFG, **extra)
        plan.run(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
        u, v, pk = plan.result(0)
        res.append((u[0], v[0], pk[0]))
        plan.close()
    return res


def check_equal(batch, ref):
    for k in range(len(PAIRS)):
        u, v, sub = batch.result(k)
        ru, rv, rs = ref[k]
        assert np.array_equal(u, ru) and np.array_equal(v, rv), k
        assert np.array_equal(sub, rs), k


CASES = [("tuned", None, -1), ("lds serial", 2, 0), ("half concurrent", 3, 1),
         ("lane concurrent", 1, 1), ("wide serial", 4, 0)]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_batch_equals_one_level_plans(ctx, case):
    import torch

    import qlpiv

    name, shape, schedule = case
    fr = frames(500)
    ref = plan_results(ctx, fr)
    batch = qlpiv.Batch(ctx, PAIRS, **CFG)
    assert batch.total_ias == sum(r[0].size for r in ref)
    if shape is not None:
        if shape == qlpiv._lib.QLP_SHAPE_LANE:
            batch.fix(16, shape, schedule)
        else:
            batch.fix(0, shape, schedule)
    dA = [torch.from_numpy(a).cuda() for a, _ in fr]
    dB = [torch.from_numpy(b).cuda() for _, b in fr]
    say(f"batch {name}: {batch.total_ias} IAs")
    batch.run(dA, dB)
    say(f"report {batch.report()}")
    check_equal(batch, ref)
    batch.close()


def test_batch_graph_replays_with_new_frames(ctx):
    """Three frame sets through one instantiated graph, with eager torch work
    and synchronisations between the replays: every replay equals the plans."""
    import torch

    import qlpiv

    batch = qlpiv.Batch(ctx, PAIRS, **CFG)
    for rep, seed in enumerate((600, 700, 600)):
        fr = frames(seed)
        ref = plan_results(ctx, fr)
        dA = [torch.from_numpy(a).cuda() for a, _ in fr]
        dB = [torch.from_numpy(b).cuda() for _, b in fr]
        say(f"replay {rep} seed {seed}")
        batch.run(dA, dB)
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        check_equal(batch, ref)
        del junk
    batch.close()


def test_batch_mfma_form_equals_plans(ctx):
    """The MFMA form in the 16 and 32 px classes (64 and 128 stay exact)."""
    import torch

    import qlpiv

    fr = frames(800)
    ref = plan_results(ctx, fr, xcorr_form=qlpiv.QLP_XCORR_MFMA_F16)
    batch = qlpiv.Batch(ctx, PAIRS, xcorr_form=qlpiv.QLP_XCORR_MFMA_F16, **CFG)
    dA = [torch.from_numpy(a).cuda() for a, _ in fr]
    dB = [torch.from_numpy(b).cuda() for _, b in fr]
    say("batch mfma form")
    batch.run(dA, dB)
    check_equal(batch, ref)
    batch.close()


def test_batch_errors(ctx):
    import qlpiv

    with pytest.raises(qlpiv.QlpError):
        qlpiv.Batch(ctx, [(256, 256, 24)], **CFG)
    with pytest.raises(qlpiv.QlpError):
        qlpiv.Batch(ctx, [], **CFG)
    with pytest.raises(qlpiv.QlpError):
        qlpiv.Batch(ctx, [(8, 8, 16)], **CFG)
    b = qlpiv.Batch(ctx, [(64, 64, 16)], **CFG)
    with pytest.raises(qlpiv.QlpError):
        b.fix(32, qlpiv._lib.QLP_SHAPE_LANE)
    with pytest.raises(qlpiv.QlpError):
        b.fix(0, 7)
    b.close()


ML_PAIRS = [(256, 256, 16, 3), (192, 320, 32, 2), (256, 256, 64, 1), (384, 384, 32, 3),
            (200, 240, 16, 2), (300, 260, 64, 2), (384, 384, 128, 1)]


@pytest.mark.parametrize("warp", [0, 1])
@pytest.mark.parametrize("form", [0, 1])
def test_multilevel_batch_equals_plans(ctx, warp, form):
    """Pairs with 1-3 adaptive levels (Direct inheritance, rounded shift or micro
    warp of B between levels, Guo fits from the second level on) in one batch:
    every pair's last level equals a multi-pass plan of that pair alone, bit
    for bit, in the exact and the MFMA form."""
    import torch

    import qlpiv

    cfg = dict(CFG, warp_mode=warp, interpolator_start_step=1, xcorr_form=form)
    fr = []
    for k, (H, W, _, _) in enumerate(ML_PAIRS):
        from qlpiv.synth import particle_pair

        a, b, _ = particle_pair(H, W, seed=900 + k, shift=(2.5, -1.5), core_radius=min(H, W) / 3.0,
                                u_max=3.0)
        fr.append((a, b))
    ref = []
    for (H, W, ia, L), (a, b) in zip(ML_PAIRS, fr):
        plan = qlpiv.PivPlan(ctx, H, W, 1, ia_start=ia << (L - 1), ia_end=ia, **cfg)
        plan.run(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda())
        u, v, pk = plan.result(-1)
        ref.append((u[0], v[0], pk[0]))
        plan.close()
    batch = qlpiv.Batch(ctx, ML_PAIRS, **cfg)
    dA = [torch.from_numpy(a).cuda() for a, _ in fr]
    dB = [torch.from_numpy(b).cuda() for _, b in fr]
    say(f"multi-level batch warp={warp} form={form}: {batch.total_ias} IAs")
    for rep in range(2):
        batch.run(dA, dB)
        for k in range(len(ML_PAIRS)):
            u, v, sub = batch.result(k)
            ru, rv, rs = ref[k]
            assert np.array_equal(u, ru) and np.array_equal(v, rv), (rep, k)
            assert np.array_equal(sub, rs), (rep, k)
    batch.close()


def test_c3_full_size_against_plans_and_oracle(ctx):
    """The batching configuration at its full size: 64 pairs, 16 each at 512^2,
    1024^2, 1024x2048 and 2048^2, IA size per pair from {16, 32, 64, 128}
    (769,664 IAs), one pass (clip, per-tile-normalised exact correlation,
    FromCenter, Guo 3 px / 20 iterations) in one graph launch.  Every pair's
    field and fit records equal its one-level plan's (float ==); every 37th IA
    of every pair, 128 px pairs included, against the oracle: fit status equal
    and fits within 1e-4 px."""
    import torch

    import oracle as O
    import qlpiv
    from qlpiv.synth import batch_frames, batch_specs

    specs = batch_specs(1700)
    say("C3: generating 64 frame pairs")
    fr = batch_frames(specs)
    pairs = [(h, w, ia) for (h, w, ia, _, _, _) in specs]
    batch = qlpiv.Batch(ctx, pairs, **CFG)
    assert batch.total_ias == 769664
    dA = [torch.from_numpy(a).cuda() for a, _ in fr]
    dB = [torch.from_numpy(b).cuda() for _, b in fr]
    say(f"C3: batch of {batch.total_ias} IAs")
    batch.run(dA, dB)
    say(f"C3 report {batch.report()}")
    worst = 0.0
    for k, ((H, W, ia), (a, b)) in enumerate(zip(pairs, fr)):
        u, v, sub = batch.result(k)
        plan = qlpiv.PivPlan(ctx, H, W, 1, ia_start=ia, ia_end=ia, warp_mode=0, **CFG)
        plan.run(dA[k], dB[k])
        pu, pv, pk = plan.result(0)
        assert np.array_equal(u, pu[0]) and np.array_equal(v, pv[0]), k
        assert np.array_equal(sub, pk[0]), k
        L = plan.levels[0]
        idx = np.arange(0, L["nI"] * L["nJ"], 37)
        out, _, _ = O.level_batch(a, b, L["top"].ravel()[idx], L["left"].ravel()[idx], None, None,
                                  ia, 0, 0, 0, 1, 3, 1, 3, 20)
        s = sub.reshape(-1, 4)[idx]
        assert np.array_equal(s[:, 3], out[:, 0, 3]), k
        d = float(np.abs(s[:, :2] - out[:, 0, :2]).max())
        assert d <= 1e-4, (k, d)
        worst = max(worst, d)
        plan.close()
    say(f"C3: 64 pairs equal their plans; oracle sample worst fit difference {worst:.2e} px")
    batch.close()
