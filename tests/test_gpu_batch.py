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
