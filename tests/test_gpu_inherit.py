"""GPU tests of the velocity inheritance strategies (qlp_inherit_velocity_dev and
the multi-pass plan's inherit_mode) against the oracle, bit for bit."""
import sys

import numpy as np
import pytest

import oracle as O
from qlpiv.geometry import level_grid

pytestmark = pytest.mark.gpu

KINDS = {"direct": 0, "area": 1, "distance": 2, "bicubic": 3}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import qlpiv

    c = qlpiv.Context(0)
    yield c
    c.close()


def direct_inherit(pu, nI, nJ, round_it=False):
    out = np.zeros((nI, nJ), np.float32)
    pI, pJ = pu.shape
    for i in range(nI):
        for j in range(nJ):
            if i // 2 < pI and j // 2 < pJ:
                x = pu[i // 2, j // 2]
                out[i, j] = np.float32(np.floor(x + np.float32(0.5))) if round_it else x
    return out


def oracle_inherit(kind, parent, child, pu, pv, round_it):
    if kind == "direct":
        ct = child[0]
        return (direct_inherit(pu, *ct.shape, round_it), direct_inherit(pv, *ct.shape, round_it))
    u, v, rc = O.inherit(kind, parent, child, pu, pv, round_it)
    assert rc == 0
    return u, v


def geometries():
    of = np.float32(1.0) / np.float32(3.0)
    m = (32, 32, 32, 32)
    yield "area-kat", (64, *level_grid(171, 171, 64, of, m, True)), (32, *level_grid(171, 171, 32, of, m, False))
    p1 = np.array([16, 37, 59])
    c1 = np.array([16, 27, 37, 48, 59, 69, 80, 91])
    yield "mixed-kat", (64, *np.meshgrid(p1, p1, indexing="ij")), (32, *np.meshgrid(c1, c1, indexing="ij"))
    g = [(64, *level_grid(2048, 2048, 64, 0.5, level0=True)),
         (32, *level_grid(2048, 2048, 32, 0.5, level0=False)),
         (16, *level_grid(2048, 2048, 16, 0.5, level0=False))]
    yield "c2-64-32", g[0], g[1]
    yield "c2-32-16", g[1], g[2]
    yield "odd-600x450", (64, *level_grid(600, 450, 64, 0.5, (3, 5, 7, 2), True)), \
        (32, *level_grid(600, 450, 32, 0.5, (3, 5, 7, 2), False))


@pytest.mark.parametrize("kind", list(KINDS))
@pytest.mark.parametrize("round_it", [False, True])
def test_inherit_equals_oracle(ctx, kind, round_it):
    import torch

    import qlpiv

    for name, (ps, pt, pl), (cs, ct, cl) in geometries():
        rng = np.random.default_rng(pt.size + 7 * ct.size)
        pairs = 2
        i, j = np.meshgrid(np.arange(pt.shape[0]), np.arange(pt.shape[1]), indexing="ij")
        pu = np.stack([(np.sin(i / 3.0 + k) * 4 + rng.normal(0, 0.3, i.shape)) for k in range(pairs)]).astype(np.float32)
        pv = np.stack([(np.cos(j / 4.0 - k) * 3 + rng.normal(0, 0.3, i.shape)) for k in range(pairs)]).astype(np.float32)
        say(f"inherit {kind} round={round_it} {name}")
        u, v = qlpiv.inherit_velocity(ctx, KINDS[kind], (pt, pl, ps), (ct, cl, cs),
                                      torch.from_numpy(pu).cuda(), torch.from_numpy(pv).cuda(),
                                      round_it)
        ctx.synchronize()
        u, v = u.cpu().numpy(), v.cpu().numpy()
        for k in range(pairs):
            eu, ev = oracle_inherit(kind, (pt, pl, ps), (ct, cl, cs), pu[k], pv[k], round_it)
            assert np.array_equal(u[k], eu) and np.array_equal(v[k], ev), (name, k)


@pytest.mark.parametrize("kind", ["area", "distance", "bicubic"])
def test_multipass_with_inheritance_against_oracle(ctx, kind):
    """Every level of the device loop with the chosen inheritance equals the
    oracle run from the previous level's device output: integer peaks
    bit-exact, displacements within 1e-4 px."""
    import torch

    import qlpiv
    from qlpiv.synth import particle_pair

    a, b, _ = particle_pair(256, 256, seed=41, shift=(1.7, -2.2), core_radius=60.0, u_max=4.0)
    plan = qlpiv.PivPlan(ctx, 256, 256, 1, ia_start=64, ia_end=16, warp_mode=1,
                         subpix_kind=qlpiv.QLP_SUBPIX_GUO, inherit_mode=KINDS[kind])
    say(f"multipass inherit={kind}")
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
            uin, vin = oracle_inherit(kind, (P["top"], P["left"], P["ia"]),
                                      (L["top"], L["left"], N), pu, pv, False)
        c = np.float32((2 * N - 1) // 2)
        for i in range(nI):
            for j in range(nJ):
                t, l0 = int(L["top"][i, j]), int(L["left"][i, j])
                tA, mA = O.clip_tile(a, t, l0, N)
                if lv == 0:
                    tB, mB = O.clip_tile(b, t, l0, N)
                else:
                    tB, mB = O.warp_micro(b, t, l0, N, uin[i, j], vin[i, j]), 255.0
                plane = O.xcorr_javagpu(tA.astype(np.float32), tB.astype(np.float32), mA, mB)
                rp = O.find_max_from_center(plane)
                rs = O.guo(plane, 3, 20, rp[3], rp[:3])
                assert int(pk[i, j, 3]) == int(rs[3]), (lv, i, j)
                eu = np.float32(uin[i, j] + np.float32(rs[0] - c))
                ev = np.float32(vin[i, j] + np.float32(rs[1] - c))
                assert abs(u[i, j] - eu) <= 1e-4 and abs(v[i, j] - ev) <= 1e-4, (lv, i, j)
        prev = (L, u, v)
    plan.close()
