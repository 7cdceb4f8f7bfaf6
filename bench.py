"""Benchmark of the plain FP32 form on one MI355X: the C2 frame-pair configuration.

Workload (one step): one synthetic 2048 x 2048 8-bit pair (0.05 particles/px,
d = 2.5 px, I0 ~ U(150, 255), noise sigma 2; frame B = frame A's particles moved
by a uniform 3.3 px shift plus a Rankine vortex R = 128 px / U = 6 px) through
the adaptive multi-pass loop 64 -> 32 -> 16 px at 50 % overlap (3,969 + 16,129
+ 65,025 = 85,123 IAs): per level the Direct velocity inheritance, clip of A,
bilinear micro-warp of B (SecondImage), FFT cross-correlation with the
reference GPU job's arithmetic (per-tile normalisation), storing every plane
(the drop-in returns them), MultiPeaks integer search (3 peaks, 3 x 3 blocks),
Guo 3 px / 20 iteration fits of every peak, the displacement update and the
vector validation of the reference UI's multi-peak normalised median model
(threshold 2.0, epsilon0 0.1: median + secondary peak 1, median + secondary
peak 2, median + bilinear; one sweep).  The frames live in HBM before the
timed region.

Prints one JSON line (the driver's contract).  value is the metric's figure:
the 16,129 IAs of the 32 x 32 level over that level's device time in the step
(correlation pass, peaks and fits, validation; the plan's HIP events over the
timed steps); the all-level rate and per-level times sit in config.  The
roofline entry is the 32 x 32 level's correlation pass.  The CPU baseline is
the oracle's C restatement of the reference's Java CPU path on this host over
the same 32 x 32 level (one thread, and IA-parallel over the host threads),
plus the single-pass 1024 x 1024 CPU-path configuration.  Unless --no-legs,
the line also carries bounded runs of the batching configuration (64 mixed
pairs in one graph launch) and of the serving configuration (a Poisson stream
at 400 req/s through the request loop, against the plain FP32 plan one
request at a time).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "quicklabpiv-ng_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
H = W = 2048


def entry_bytes(N):
    """MultiPeaks entries an exact pass writes per IA (kernel width 3): 8 bytes
    per 3x3 block of the (2N-1)^2 plane in the reference's visit order, padded to
    chunks of 16 entries, 4 bytes per chunk maximum and 4 for the plane minimum."""
    D = 2 * N - 1
    slices = (D + 2) // 3
    top = slices - slices // 2
    stride = (4 * top * top + 15) // 16 * 16
    return stride * 8 + stride // 16 * 4 + 4


def pass_bytes(n_ia, N):
    """Algorithmic HBM bytes of one plane-writing correlation pass of the C2
    step: the A and B tile pixels (2 N^2), the (2N-1)^2 float32 plane, the
    8-byte inherited displacement and the MultiPeaks entries it writes for the
    peak search (entry_bytes)."""
    return n_ia * (2 * N * N + 4 * (2 * N - 1) ** 2 + 8 + entry_bytes(N))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Host threads for the IA-parallel variant: the cores this process may run
    on, capped by OMP_NUM_THREADS (the box's CPU share for one GPU)."""
    n = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(cap))) if cap and cap.isdigit() else n


def _median_time(fn, runs):
    fn()  # warm-up (caches, page faults, lazy library load)
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def _level_parallel(O, threads, a, b, top, left, u, v, N, **kw):
    """One level's IAs split into `threads` contiguous chunks, each chunk one
    oracle call on its own thread (ctypes releases the GIL during the call)."""
    from concurrent.futures import ThreadPoolExecutor

    n = top.size
    cuts = np.linspace(0, n, threads + 1).astype(int)
    with ThreadPoolExecutor(threads) as ex:
        parts = list(ex.map(lambda k: O.level_batch(a, b, top[cuts[k]:cuts[k + 1]],
                                                    left[cuts[k]:cuts[k + 1]],
                                                    u[cuts[k]:cuts[k + 1]], v[cuts[k]:cuts[k + 1]],
                                                    N, **kw),
                            range(threads)))
    return np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts])


def cpu_baseline(a, b, lv32, stages, runs=3):
    """The reference's CPU path restated in C by the oracle (built -O2
    -ffp-contract=off; not JVM timing), on this host's cores:
      * value: the 32 x 32 level of the same C2 pair (16,129 IAs from the
        device's inherited field): clip A, micro-warp B (SecondImage), the Java
        CPU correlation job's arithmetic (CrossCorrelationSoftRealFFTJob),
        MultiPeaks 3 x 3 / 3, Guo 3 px / 20 iterations on every peak, the
        multi-peak normalised median validation; one thread;
      * all_cores: the same split by IA over the host threads (restatement,
        IA-parallel -- not a reference capability: the Java CPU job runs one
        IA after another);
      * c1: the reference CPU-path configuration (one 1024 x 1024 pair, 63 x 63
        IAs of 32 x 32, FromCenter, Gaussian 3-point, single pass, no warp),
        one thread and all threads.
    Each figure is the median of `runs` timed runs after one warm-up run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    from qlpiv.geometry import level_grid
    from qlpiv.synth import particle_pair

    threads = cpu_threads()
    top, left, u, v, N, nI, nJ = lv32
    n = top.size
    P = 3
    c = np.float32(N - 1)
    lvkw = dict(warp=2, mode=1, peak_kind=1, P=P, k=3, fit_kind=1, pixels=3, iterations=20)

    def validate(out, cnt):
        inc = (out[..., :2] - c).astype(np.float32)
        up = (u + inc[:, 0, 0]).astype(np.float32).reshape(nI, nJ)
        vp = (v + inc[:, 0, 1]).astype(np.float32).reshape(nI, nJ)
        O.validate(top.reshape(nI, nJ), left.reshape(nI, nJ), up, vp, stages,
                   inc.reshape(nI, nJ, P, 2), cnt.reshape(nI, nJ), 0, False)

    def one():
        out, cnt, _ = O.level_batch(a, b, top, left, u, v, N, **lvkw)
        validate(out, cnt)

    def allc():
        out, cnt = _level_parallel(O, threads, a, b, top, left, u, v, N, **lvkw)
        validate(out, cnt)

    print(f"[bench] CPU baseline: 32x32 level, 1 and {threads} threads", file=sys.stderr, flush=True)
    t1, ts1 = _median_time(one, runs)
    tm, tsm = _median_time(allc, runs)

    a1, b1, _ = particle_pair(1024, 1024, seed=1000, shift=(0.0, 0.0), core_radius=128.0, u_max=6.0)
    ct, cl = level_grid(1024, 1024, 32, 0.5)
    ct = np.ascontiguousarray(ct.ravel(), np.int32)
    cl = np.ascontiguousarray(cl.ravel(), np.int32)
    z = np.zeros(ct.size, np.float32)
    c1kw = dict(warp=0, mode=1, peak_kind=0, P=1, k=3, fit_kind=0, pixels=3, iterations=20)
    print("[bench] CPU baseline: C1 pair", file=sys.stderr, flush=True)
    c1_1, _ = _median_time(lambda: O.level_batch(a1, b1, ct, cl, z, z, 32, **c1kw), max(runs, 5))
    c1_m, _ = _median_time(lambda: _level_parallel(O, threads, a1, b1, ct, cl, z, z, 32, **c1kw),
                           max(runs, 5))
    return {"value": round(n / t1, 1), "unit": "IAs/s", "cores": 1, "kind": "port",
            "cpu_model": cpu_model(), "host_threads_available": len(os.sched_getaffinity(0)),
            "sample": f"the {n} IAs of the 32x32 level of the same 2048x2048 pair (inherited "
                      f"field from the device's 64x64 level): clip, micro-warp, Java CPU xcorr, "
                      f"MultiPeaks 3x3/3, Guo 3px/20it per peak, multi-peak median validation; "
                      f"C restatement of the reference CPU path (-O2 -ffp-contract=off, not JVM "
                      f"timing), 1 thread, median of {runs} runs after a warm-up "
                      f"({', '.join(f'{x:.2f}' for x in ts1)} s)",
            "level_latency_ms": round(t1 * 1e3, 2),
            "all_cores": {"label": "restatement, IA-parallel -- not a reference capability",
                          "value": round(n / tm, 1), "unit": "IAs/s", "cores": threads,
                          "level_latency_ms": round(tm * 1e3, 2),
                          "runs_s": [round(x, 3) for x in tsm]},
            "c1": {"workload": "C1: one 1024x1024 pair, 63x63 IAs of 32x32 at 50% overlap, Java "
                               "CPU xcorr (A-max normalisation), FromCenter, Gaussian 3-point, "
                               "single pass, no warp",
                   "ias": int(ct.size),
                   "one_thread": {"ias_per_s": round(ct.size / c1_1, 1),
                                  "pair_latency_ms": round(c1_1 * 1e3, 2)},
                   "all_cores": {"ias_per_s": round(ct.size / c1_m, 1),
                                 "pair_latency_ms": round(c1_m * 1e3, 2), "cores": threads,
                                 "label": "restatement, IA-parallel -- not a reference capability"},
                   "runs": max(runs, 5)}}


def time_plan(ctx, plan, da, db, steps, warmup):
    """Run warmup untimed steps, then time exactly `steps` steps: the host clock
    around the whole run, HIP events per step on the library's stream and the
    plan's own per-level events read after each step.  Returns (wall seconds,
    per-step ms, per-level mean ms as levels x (pass, peaks+fit, validation))."""
    import torch

    s = ctx.torch_stream()
    plan.set_timing(True)
    for _ in range(warmup):
        plan.run(da, db)
    ctx.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(steps)]
    per_level = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record(s)
        plan.run(da, db)
        e1.record(s)
        per_level.append(plan.timing())  # waits for this step
    ctx.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    plan.set_timing(False)
    step_ms = np.array([e0.elapsed_time(e1) for e0, e1 in evs])
    return dt, step_ms, np.mean(np.array(per_level), axis=0)


def precision_report(total32, total64):
    """The precision configuration: independent IAs cut from synthetic 2048 x 2048
    pairs (50 % overlap grids; 16,129 IAs of 32 x 32 or 3,969 of 64 x 64 per
    pair), correlated in every arithmetic form and compared with an FP64 FFT
    correlation of the same normalised tiles (torch float64 on the device).
    Per form: plane error relative to the plane maximum (max and RMS over IAs),
    FromCenter integer-peak agreement with the FP64 plane (rounded to float32
    for the same peak finder) on IAs whose FP64 peak ratio P1/P2 of 3 x 3 local
    maxima exceeds 1.2, and the 3-point Gaussian sub-pixel difference on those
    IAs where both fits are accepted."""
    import torch
    import torch.nn.functional as F

    import qlpiv
    from qlpiv.geometry import level_grid
    from qlpiv.synth import particle_pair

    dev = torch.device("cuda:0")
    ctx = qlpiv.Context(0)
    forms = {"exact_fp32": qlpiv.QLP_XCORR_EXACT, "mfma_f16_split": qlpiv.QLP_XCORR_MFMA_F16,
             "mfma_bf16_split": qlpiv.QLP_XCORR_MFMA_BF16, "mfma_f16": qlpiv.QLP_XCORR_MFMA_F16_SINGLE,
             "mfma_bf16": qlpiv.QLP_XCORR_MFMA_BF16_SINGLE}
    report = {}
    for N, total in ((32, total32), (64, total64)):
        M, D = 2 * N, 2 * N - 1
        tops, lefts = level_grid(H, W, N, 0.5, level0=True)
        top_all = torch.from_numpy(np.ascontiguousarray(tops.ravel(), np.int32)).to(dev)
        left_all = torch.from_numpy(np.ascontiguousarray(lefts.ravel(), np.int32)).to(dev)
        acc = {k: {"err_max": 0.0, "err_sq": 0.0, "n": 0, "agree": 0, "n_ppr": 0,
                   "sub_max": 0.0, "sub_sq": 0.0, "n_sub": 0} for k in forms}
        done, pair = 0, 0
        while done < total:
            n = min(total - done, top_all.numel())
            print(f"[precision] N={N} pair {pair}: {n} IAs", file=sys.stderr, flush=True)
            a, b, _ = particle_pair(H, W, seed=1400 + pair, shift=(3.3, 3.3), core_radius=128.0,
                                    u_max=6.0)
            da, db = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)
            tA, mA = ctx.clip_tiles(da, top_all[:n].contiguous(), left_all[:n].contiguous(), N)
            tB, mB = ctx.clip_tiles(db, top_all[:n].contiguous(), left_all[:n].contiguous(), N)
            ctx.synchronize()
            # FP64 reference of the normalised tiles (float32 normalisation as the job's)
            an = (tA.float() / mA[:, None, None]) * 16.0 + 1.0
            gn = (tB.float() / mB[:, None, None]) * 16.0 + 1.0
            ref = torch.empty((n, D, D), dtype=torch.float64, device=dev)
            idx = (torch.arange(D, device=dev) - (N - 1)) % M
            for s in range(0, n, 8192):
                e = min(n, s + 8192)
                A = torch.fft.fft2(an[s:e].double(), s=(M, M))
                G = torch.fft.fft2(gn[s:e].double(), s=(M, M))
                r = torch.fft.ifft2(torch.conj(A) * G).real
                ref[s:e] = r[:, idx][:, :, idx]
            ref_max = ref.abs().amax(dim=(1, 2))
            # peak ratio of the FP64 plane: the two largest 3 x 3 local maxima
            pooled = F.max_pool2d(ref[:, None], 3, stride=1, padding=1)[:, 0]
            loc = torch.where(ref == pooled, ref, torch.full_like(ref, -float("inf")))
            top2 = torch.topk(loc.reshape(n, -1), 2, dim=1).values
            ppr = torch.where(top2[:, 1] > 0, top2[:, 0] / top2[:, 1],
                              torch.full_like(top2[:, 0], float("inf")))
            sel = ppr > 1.2
            ref32 = ref.float().contiguous()
            torch.cuda.synchronize()  # torch's stream wrote ref32; the library's reads it
            pk_r, mf_r = ctx.find_peaks(ref32)
            sub_r = ctx.subpixel(ref32, mf_r, pk_r, kind=qlpiv.QLP_SUBPIX_GAUSS1D, pixels=3)
            for name, form in forms.items():
                planes = ctx.xcorr(tA, tB, mA, mB, form=form)
                pk, mf = ctx.find_peaks(planes)
                sub = ctx.subpixel(planes, mf, pk, kind=qlpiv.QLP_SUBPIX_GAUSS1D, pixels=3)
                ctx.synchronize()
                err = (planes.double() - ref).abs().amax(dim=(1, 2)) / ref_max
                st = acc[name]
                st["err_max"] = max(st["err_max"], float(err.max()))
                st["err_sq"] += float((err ** 2).sum())
                st["n"] += n
                agree = (pk[:, 0, :2] == pk_r[:, 0, :2]).all(dim=1)
                st["agree"] += int((agree & sel).sum())
                st["n_ppr"] += int(sel.sum())
                ok = sel & agree & (sub[:, 0, 3] == 3) & (sub_r[:, 0, 3] == 3)
                d = (sub[:, 0, :2] - sub_r[:, 0, :2]).abs().amax(dim=1)[ok]
                if d.numel():
                    st["sub_max"] = max(st["sub_max"], float(d.max()))
                    st["sub_sq"] += float((d.double() ** 2).sum())
                    st["n_sub"] += int(d.numel())
                del planes
            done += n
            pair += 1
        report[f"{N}x{N}"] = {
            "ias": total, "pairs": pair,
            "forms": {k: {"plane_err_max_rel": st["err_max"],
                          "plane_err_rms_rel": (st["err_sq"] / st["n"]) ** 0.5,
                          "int_peak_agreement_ppr_gt_1.2": st["agree"] / max(st["n_ppr"], 1),
                          "ias_ppr_gt_1.2": st["n_ppr"],
                          "subpixel_err_max_px": st["sub_max"],
                          "subpixel_err_rms_px": (st["sub_sq"] / max(st["n_sub"], 1)) ** 0.5,
                          "fits_compared": st["n_sub"]}
                      for k, st in acc.items() if st["n"]}}
    ctx.close()
    print(json.dumps({"metric": "precision of the correlation forms vs an FP64 correlation",
                      "config": "independent IAs cut from synthetic 2048x2048 pairs",
                      "report": report}), flush=True)


SERVE_SIZES = ((512, 512), (1024, 1024), (1024, 2048), (2048, 2048))  # (H, W)


def serve_schedule(n, rate, seed):
    """Poisson arrivals at `rate` req/s; per request a frame size (uniform over
    SERVE_SIZES), a final IA size from {16, 32, 64} and 1-3 levels (IA sizes
    halving, the first at most 64 px)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for k in range(n):
        size = int(rng.integers(len(SERVE_SIZES)))
        ia_end = int(rng.choice([16, 32, 64]))
        levels = int(rng.integers(1, {16: 3, 32: 2, 64: 1}[ia_end] + 1))
        out.append((float(t[k]), size, ia_end * 2 ** (levels - 1), ia_end))
    return out


def serve_frames(qlpiv, seed=1500):
    """One synthetic pair per serving frame size, device-resident (requests are
    served from frames already in HBM, for the server and the plain form alike)."""
    import torch

    from qlpiv.synth import particle_pair

    frames = []
    for k, (h, w) in enumerate(SERVE_SIZES):
        rng = np.random.default_rng(k + 10)
        a, b, _ = particle_pair(h, w, seed=seed + 10 * k, shift=(3.3, 3.3),
                                core_radius=float(rng.uniform(32, 256)),
                                u_max=float(rng.uniform(2, 10)))
        frames.append((torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()))
    torch.cuda.synchronize()
    return frames


def serve_common(qlpiv, form):
    return dict(overlap_factor=0.5, warp_mode=qlpiv._lib.QLP_WARP_MICRO,
                subpix_kind=qlpiv.QLP_SUBPIX_GUO, subpix_pixels=3, subpix_iterations=20,
                xcorr_form=form)


def latency_stats(arrive, done, ias, wall):
    lat = (np.array(done) - np.array(arrive)) * 1e3
    return {"requests": len(lat), "ias": int(ias), "wall_s": round(wall, 4),
            "ias_per_s": round(ias / wall, 1), "req_per_s": round(len(lat) / wall, 2),
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                           "p99": round(float(np.percentile(lat, 99)), 3),
                           "max": round(float(lat.max()), 3)}}


def serve_server(qlpiv, frames, schedule, form, saturate):
    """The request loop (continuous batching) over the schedule: submitted at
    their arrival times, or all at once when saturate."""
    srv = qlpiv.Server(0, hbm_budget=64 << 30, slots=2)
    common = serve_common(qlpiv, form)
    # warm-up: every (size, levels) combination (engine graphs, level grids), then
    # a burst of the stream's first requests (steps large enough for the engine
    # to time its schedule)
    warm = [srv.submit(*frames[s], ia_start=a, ia_end=e, **common)
            for s in range(len(SERVE_SIZES))
            for (a, e) in ((64, 16), (32, 16), (16, 16), (64, 32), (32, 32), (64, 64))]
    for t in warm:
        srv.wait(t)
        srv.result(t)
    warm = [srv.submit(*frames[s], ia_start=a, ia_end=e, **common) for (_, s, a, e) in schedule[:200]]
    for t in warm:
        srv.wait(t)
        srv.result(t)
    cfgs = {(a, e): qlpiv.piv_config(ia_start=a, ia_end=e, **common)
            for (_, _, a, e) in schedule}
    tickets = []
    t0 = time.perf_counter()
    for (ta, s, a, e) in schedule:
        if not saturate:
            while time.perf_counter() - t0 < ta:
                pass
        tickets.append(srv.submit(*frames[s], cfg=cfgs[(a, e)]))
    t_submit = time.perf_counter() - t0
    arrive, done, ias = [], [], 0
    first = last = None
    for t in tickets:
        srv.wait(t)
        _, _, _, tm = srv.result(t)
        arrive.append(tm["submit_ms"] * 1e-3)
        done.append(tm["done_ms"] * 1e-3)
        ias += tm["ias"]
        first = tm["submit_ms"] if first is None else min(first, tm["submit_ms"])
        last = tm["done_ms"] if last is None else max(last, tm["done_ms"])
    st = srv.stats()
    srv.close()
    out = latency_stats(arrive, done, ias, (last - first) * 1e-3)
    out["hbm_peak_bytes"] = st["hbm_peak"]
    out["failed"] = st["failed"]
    out["submit_loop_s"] = round(t_submit, 4)
    return out


def serve_plain(qlpiv, frames, schedule, saturate):
    """The plain FP32 GPU form on the same stream: the drop-in multi-pass plan,
    one request at a time in arrival order, each enqueued when it has arrived
    with its last level's field copied to pinned host memory behind it on the
    stream (no host wait between requests: the copies overlap the next
    request's kernels as the server's do).  A request is done when the stream
    passes the event recorded after its copy."""
    import torch

    ctx = qlpiv.Context(0)
    stream = ctx.torch_stream()
    plans = {}
    common = serve_common(qlpiv, qlpiv.QLP_XCORR_EXACT)
    hosts = []
    for (_, s, a, e) in schedule:
        if (s, a, e) not in plans:
            H, W = SERVE_SIZES[s]
            plans[(s, a, e)] = qlpiv.PivPlan(ctx, H, W, 1, ia_start=a, ia_end=e, **common)
            plans[(s, a, e)].run(*frames[s])
            plans[(s, a, e)].result(-1)
        L = plans[(s, a, e)].levels[-1]
        hosts.append(tuple(torch.empty((1, L["nI"], L["nJ"]), dtype=torch.float32,
                                       pin_memory=True).numpy() for _ in range(2)))
    evs = [torch.cuda.Event(enable_timing=True) for _ in schedule]
    start = torch.cuda.Event(enable_timing=True)
    ctx.synchronize()
    ias = 0
    t0 = time.perf_counter()
    start.record(stream)
    for k, (ta, s, a, e) in enumerate(schedule):
        if not saturate:
            while time.perf_counter() - t0 < ta:
                pass
        p = plans[(s, a, e)]
        p.run(*frames[s])
        p.result_async(hosts[k][0], hosts[k][1], -1)
        evs[k].record(stream)
        ias += sum(L["nI"] * L["nJ"] for L in p.levels)
    evs[-1].synchronize()
    wall = time.perf_counter() - t0
    done = [start.elapsed_time(ev) * 1e-3 for ev in evs]
    arrive = [0.0 if saturate else ta for (ta, _, _, _) in schedule]
    for p in plans.values():
        p.close()
    ctx.close()
    return latency_stats(arrive, done, ias, wall)


def serve_report(n_offered, n_saturated, rate, reps):
    """The serving configuration: a Poisson stream of frame-pair requests of mixed
    sizes (512^2, 1024^2, 1024x2048, 2048^2; final IA 16/32/64 with 1-3 levels;
    micro warp, FromCenter, Guo 3 px) through the request loop with continuous
    batching and a 64 GB HBM budget, in the MFMA fp16 hi+lo and the exact FP32
    form, against the plain FP32 form (the drop-in plan, one request at a time).
    'offered': the first n_offered requests of the stream at their arrival
    times (latency under the 400 req/s load); 'saturated': n_saturated requests
    queued at once (the throughput limit), `reps` runs per form on fresh servers,
    the median run reported beside every run's rate.  Frames are device-resident
    for every form."""
    import qlpiv

    frames = serve_frames(qlpiv)
    schedule = serve_schedule(max(n_offered, n_saturated), rate, seed=1600)
    out = {}

    def median_run(runs):
        runs = sorted(runs, key=lambda r: r["ias_per_s"])
        med = dict(runs[len(runs) // 2])
        med["runs_ias_per_s"] = [r["ias_per_s"] for r in runs]
        return med

    forms = (("serving_mfma_f16_split", qlpiv.QLP_XCORR_MFMA_F16),
             ("serving_exact_fp32", qlpiv.QLP_XCORR_EXACT))
    for name, form in forms:
        print(f"[serve] {name} offered: {n_offered} requests", file=sys.stderr, flush=True)
        out.setdefault(name, {})["offered"] = serve_server(qlpiv, frames, schedule[:n_offered],
                                                           form, False)
        print(f"[serve] {name} saturated: {reps} x {n_saturated} requests", file=sys.stderr,
              flush=True)
        out[name]["saturated"] = median_run(
            [serve_server(qlpiv, frames, schedule[:n_saturated], form, True) for _ in range(reps)])
    print(f"[serve] plain FP32 form offered: {n_offered} requests", file=sys.stderr, flush=True)
    out["plain_fp32_one_at_a_time"] = {"offered": serve_plain(qlpiv, frames, schedule[:n_offered],
                                                              False)}
    print(f"[serve] plain FP32 form saturated: {reps} x {n_saturated} requests", file=sys.stderr,
          flush=True)
    out["plain_fp32_one_at_a_time"]["saturated"] = median_run(
        [serve_plain(qlpiv, frames, schedule[:n_saturated], True) for _ in range(reps)])
    plain = out["plain_fp32_one_at_a_time"]["saturated"]["ias_per_s"]
    return {"metric": "serving: throughput and latency of a Poisson stream of mixed frame-pair "
                      "requests",
            "config": {"arrival_rate_req_per_s": rate, "requests_offered": n_offered,
                       "requests_saturated": n_saturated, "saturated_runs": reps,
                       "sizes_HxW": SERVE_SIZES, "final_ia": [16, 32, 64], "levels": "1-3",
                       "hbm_budget_bytes": 64 << 30,
                       "per_request": "micro warp, FromCenter, Guo 3px/20it, direct inheritance; "
                                      "last level returned to the host; frames device-resident"},
            "forms": out,
            "speedup_serving_mfma_vs_plain_saturated": round(
                out["serving_mfma_f16_split"]["saturated"]["ias_per_s"] / plain, 3),
            "speedup_serving_exact_vs_plain_saturated": round(
                out["serving_exact_fp32"]["saturated"]["ias_per_s"] / plain, 3)}


def batch_report(steps, warmup, gen_workers=8):
    """The batching configuration: 64 frame pairs, 16 each at 512x512,
    1024x1024, 1024x2048 and 2048x2048, the IA size of each pair drawn from
    {16, 32, 64, 128} at 50 % overlap, per-pair Rankine vortex R ~ U(32, 256),
    U ~ U(2, 10).  One pass per pair (clip, exact correlation with per-tile
    normalisation, FromCenter, Guo 3 px / 20 iterations) for all 64 pairs in one
    graph launch (qlp_batch), against the same work as one one-level plan per
    pair run one after the other."""
    import torch

    import qlpiv
    from qlpiv.synth import batch_frames, batch_specs

    specs = batch_specs(1700)
    print(f"[batch] generating {len(specs)} frame pairs", file=sys.stderr, flush=True)
    got = batch_frames(specs, gen_workers)
    dA = [torch.from_numpy(a).cuda() for a, _ in got]
    dB = [torch.from_numpy(b).cuda() for _, b in got]
    del got
    cfg = dict(overlap_factor=0.5, subpix_kind=qlpiv.QLP_SUBPIX_GUO, subpix_pixels=3,
               subpix_iterations=20)
    ctx = qlpiv.Context(0)
    pairs = [(h, w, ia) for (h, w, ia, _, _, _) in specs]
    batch = qlpiv.Batch(ctx, pairs, **cfg)
    print(f"[batch] {batch.total_ias} IAs; tuning and building the graph", file=sys.stderr, flush=True)
    batch.run(dA, dB)
    ctx.synchronize()
    for _ in range(warmup):
        batch.run(dA, dB)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run(dA, dB)
    ctx.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    report = batch.report()
    # the same work without the batch: one one-level plan per pair, in sequence
    print("[batch] one plan per pair, in sequence", file=sys.stderr, flush=True)
    plans = [qlpiv.PivPlan(ctx, h, w, 1, ia_start=ia, ia_end=ia, warp_mode=0, **cfg)
             for (h, w, ia) in pairs]
    for p, a, b in zip(plans, dA, dB):
        p.run(a, b)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for p, a, b in zip(plans, dA, dB):
            p.run(a, b)
    ctx.synchronize()
    ms_seq = (time.perf_counter() - t0) / steps * 1e3
    # equal results (spot check of every pair's field)
    same = all(np.array_equal(batch.result(k)[0], plans[k].result(0)[0][0]) for k in range(len(pairs)))
    for p in plans:
        p.close()
    batch.close()
    ctx.close()
    mix = {str(n): sum(1 for p in pairs if p[2] == n) for n in (16, 32, 64, 128)}
    return {"metric": "batching: correlated IAs/sec of 64 mixed frame pairs in one graph launch",
                      "value": round(batch.total_ias / ms * 1e3, 1), "unit": "IAs/s",
                      "ms_per_batch": round(ms, 4), "steps": steps, "warmup": warmup,
                      "ias": batch.total_ias, "dtype": "f32", "data": "synthetic",
                      "config": {"workload": "64 pairs: 16 each at 512x512, 1024x1024, 1024x2048, "
                                             "2048x2048; IA per pair from {16,32,64,128}, 50% "
                                             "overlap; one pass: clip, exact correlation, "
                                             "FromCenter, Guo 3px/20it",
                                 "pairs_per_ia": mix},
                      "autotuner": report,
                      "sequential_plans": {"ms_per_batch": round(ms_seq, 4),
                                           "ias_per_s": round(batch.total_ias / ms_seq * 1e3, 1)},
                      "speedup_vs_sequential_plans": round(ms_seq / ms, 3),
                      "results_equal_to_plans": bool(same)}


def level_ms(plan, per_level):
    return {str(plan.levels[k]["ia"]): {"pass": round(float(per_level[k][0]), 4),
                                        "peaks_fit": round(float(per_level[k][1]), 4),
                                        "validation": round(float(per_level[k][2]), 4)}
            for k in range(len(plan.levels))}


def direct_inherit(pu, nI, nJ):
    """Direct inheritance: IA (i, j) takes the parent level's (i // 2, j // 2)
    where that exists, else 0."""
    out = np.zeros((nI, nJ), np.float32)
    ci, cj = min(nI, 2 * pu.shape[0]), min(nJ, 2 * pu.shape[1])
    out[:ci, :cj] = pu[np.arange(ci)[:, None] // 2, np.arange(cj)[None, :] // 2]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mfma", action="store_true", help="skip the matrix-core form")
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the bounded batching and serving legs of the default line")
    ap.add_argument("--leg-requests", type=int, nargs=2, default=(400, 2000),
                    metavar=("OFFERED", "SATURATED"),
                    help="requests of the offered-rate run and of each saturated run in the "
                         "default line's serving leg")
    ap.add_argument("--precision", action="store_true",
                    help="print the precision report of the correlation forms instead")
    ap.add_argument("--precision-ias", type=int, nargs=2, default=(200000, 50000),
                    metavar=("N32", "N64"), help="IAs of 32x32 and of 64x64 in the report")
    ap.add_argument("--serve", action="store_true",
                    help="print the full serving report (Poisson stream through the request loop)")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-runs", type=int, default=3,
                    help="saturated runs per form (the median is reported)")
    ap.add_argument("--serve-rate", type=float, default=400.0)
    ap.add_argument("--batch", action="store_true",
                    help="print the batching report (64 mixed pairs in one graph launch)")
    ap.add_argument("--gen-workers", type=int, default=8,
                    help="processes generating the batch frames (0: in this process, for "
                         "profilers that do not follow forks)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one MI355X per box); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.precision:
        precision_report(*args.precision_ias)
        return
    if args.serve:
        print(json.dumps(serve_report(args.serve_requests, args.serve_requests, args.serve_rate,
                                      args.serve_runs)), flush=True)
        return
    if args.batch:
        print(json.dumps(batch_report(args.steps, args.warmup, args.gen_workers)), flush=True)
        return

    import torch

    import qlpiv
    from qlpiv.synth import particle_pair

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda:0")
    print("[bench] generating the 2048x2048 pair", file=sys.stderr, flush=True)
    a, b, _field = particle_pair(H, W, seed=1100, shift=(3.3, 3.3), core_radius=128.0, u_max=6.0)
    da, db = torch.from_numpy(a).to(dev), torch.from_numpy(b).to(dev)

    ctx = qlpiv.Context(0)
    stages = qlpiv._lib.multipeak_median_stages(3, 2.0, 0.1)
    c2 = dict(ia_start=64, ia_end=16, overlap_factor=0.5, warp_mode=qlpiv._lib.QLP_WARP_MICRO,
              norm_mode=qlpiv.QLP_NORM_PER_TILE, subpix_kind=qlpiv.QLP_SUBPIX_GUO,
              subpix_pixels=3, subpix_iterations=20, keep_planes=2,
              peak_kind=qlpiv.QLP_PEAK_MULTI, n_peaks=3, peak_kernel_width=3,
              validation=stages, valid_max_iterations=0)
    plan = qlpiv.PivPlan(ctx, H, W, 1, **c2)
    n_levels = [L["nI"] * L["nJ"] for L in plan.levels]
    n_step = sum(n_levels)
    lv32 = [L["ia"] for L in plan.levels].index(32)
    n32 = n_levels[lv32]
    print(f"[bench] exact FP32 form: {args.warmup} warmup + {args.steps} timed steps of "
          f"{n_step} IAs", file=sys.stderr, flush=True)
    dt, step_ms, per_level = time_plan(ctx, plan, da, db, args.steps, args.warmup)
    lvl32_ms = float(per_level[lv32].sum())
    pass32_ms = float(per_level[lv32][0])
    achieved = pass_bytes(n32, 32) / (pass32_ms * 1e-3) / 1e9
    ue, ve, _ = plan.result(len(plan.levels) - 1)

    mfma = None
    if not args.no_mfma:
        print("[bench] MFMA form (fp16 hi+lo, all three levels)", file=sys.stderr, flush=True)
        plan_m = qlpiv.PivPlan(ctx, H, W, 1, xcorr_form=qlpiv.QLP_XCORR_MFMA_F16, **c2)
        dt_m, step_m, per_m = time_plan(ctx, plan_m, da, db, args.steps, args.warmup)
        um, vm, _ = plan_m.result(len(plan_m.levels) - 1)
        d = np.hypot(um - ue, vm - ve).ravel()
        plan_m.close()
        m32_ms = float(per_m[lv32].sum())
        mfma = {"form": "MFMA fp16 hi+lo operands, fp32 accumulation (all three levels)",
                "ias_per_s_at_32x32": round(n32 / (m32_ms * 1e-3), 1), "unit": "IAs/s",
                "speedup_vs_exact": round(lvl32_ms / m32_ms, 3),
                "pass_32x32_ms": round(float(per_m[lv32][0]), 4),
                "ms_per_step": round(dt_m / args.steps * 1e3, 3),
                "ias_per_s_all_levels": round(n_step * args.steps / dt_m, 1),
                "speedup_vs_exact_step": round(dt / dt_m, 3),
                "pair_latency_ms": {"p50": round(float(np.percentile(step_m, 50)), 3),
                                    "p99": round(float(np.percentile(step_m, 99)), 3)},
                "level_ms": level_ms(plan, per_m),
                "final_field_vs_exact_px": {"median": round(float(np.median(d)), 6),
                                            "p99": round(float(np.percentile(d, 99)), 6),
                                            "frac_within_0.01": round(float((d <= 0.01).mean()), 5)}}

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "xcorr_pass_c2_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch", {}).get("32")

    cpu = None
    if not args.no_cpu_baseline:
        L32 = plan.levels[lv32]
        u64, v64, _ = plan.result(lv32 - 1)
        ui = direct_inherit(u64[0], L32["nI"], L32["nJ"])
        vi = direct_inherit(v64[0], L32["nI"], L32["nJ"])
        cpu = cpu_baseline(a, b, (np.ascontiguousarray(L32["top"].ravel(), np.int32),
                                  np.ascontiguousarray(L32["left"].ravel(), np.int32),
                                  ui.ravel(), vi.ravel(), 32, L32["nI"], L32["nJ"]), stages)
    lvms = level_ms(plan, per_level)
    plan.close()
    ctx.close()
    del da, db

    legs = {}
    if not args.no_legs:
        print("[bench] batching leg", file=sys.stderr, flush=True)
        legs["batch"] = batch_report(10, 2, min(args.gen_workers, cpu_threads()))
        print("[bench] serving leg", file=sys.stderr, flush=True)
        legs["serving"] = serve_report(args.leg_requests[0], args.leg_requests[1], 400.0, 3)

    line = {
        "metric": "correlated IAs/sec at 32x32; p50/p99 pair latency; MFMA vs FP32 form",
        "value": round(n32 / (lvl32_ms * 1e-3), 1),
        "unit": "IAs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "C2 plain FP32 form: one 2048x2048 pair, multipass 64->32->16 at "
                               "50% overlap, micro warp of B, per-tile-normalised FFT xcorr "
                               "storing every plane, MultiPeaks 3 peaks/3x3, Guo 3px/20it on "
                               "every peak, multi-peak normalised median validation (2.0, 0.1) "
                               "with secondary-peak and bilinear replacement, 1 sweep",
                   "value_is": "the 16,129 IAs of the 32x32 level / that level's device time "
                               "(correlation pass + peaks and fits + validation, the plan's own "
                               "HIP events, mean over the timed steps)",
                   "ias_per_step": n_step, "ias_per_level": n_levels,
                   "ias_per_s_all_levels": round(n_step * args.steps / dt, 1),
                   "pair_latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 3),
                                       "p99": round(float(np.percentile(step_ms, 99)), 3)},
                   "level_ms": lvms},
        "mfma_form": mfma,
        "roofline": {"bound": "hbm",
                     "kernel": "xcorr_pass_kernel<6,false,true,false,512> (32x32 level, 512 lanes)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": "profiles/xcorr_pass_c2_traffic.json: rocprofv3 --pmc "
                                       "FETCH_SIZE and WRITE_SIZE of this kernel on this workload, "
                                       "scaled by a copy kernel's reading (not measured in this run)",
                     "avg_launch_ms": round(pass32_ms, 4),
                     "algorithmic_bytes_per_launch": pass_bytes(n32, 32)},
        "cpu_baseline": cpu,
    }
    line.update(legs)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
