"""Benchmark of the plain FP32 correlation form on one MI355X.

Workload (one step): the 32 x 32 level of the C2 frame-pair configuration --
synthetic 2048 x 2048 8-bit pairs (0.05 particles/px, d = 2.5 px, Rankine
vortex R = 128 px / U = 6 px plus a uniform 3.3 px shift), IAs of 32 x 32 at
50 % overlap (127 x 127 = 16,129 per pair), PAIRS pairs per step.  Per IA:
clip of frame A, bilinear micro-warp of frame B by the predictor displacement
of the coarser level, FFT cross-correlation (per-tile normalisation, the GPU
job's arithmetic), FromCenter integer peak, Guo 3 px / 20 iteration sub-pixel
fit.  Images, IA grid and predictor live in HBM before the timed region.

Prints one JSON line (the driver's contract), with the roofline of the
dominant kernel measured by HIP events on the library's stream and the
reference CPU path (the oracle's C restatement of the Java code) timed on a
bounded sample on this host.
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
N = 32
H = W = 2048


def ia_bytes(n_ia):
    """Algorithmic HBM bytes of the xcorr kernel per launch: u8 tiles A and B
    (2 N^2), their max values (8 B) and the (2N-1)^2 float32 plane."""
    return n_ia * (2 * N * N + 8 + 4 * (2 * N - 1) ** 2)


def cpu_baseline(tA, tB, mA, max_seconds=20.0):
    """The reference CPU path (CrossCorrelationSoftRealFFTJob -> FastRealFFTXCorr ->
    FindMaximumFromCenter -> Guo 3 px) as restated in C by the oracle, one
    thread, on a bounded sample of the same IAs."""
    import ctypes

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    L = O.lib()
    L.or_cpu_path_batch.argtypes = [ctypes.c_int, ctypes.c_int, O._u8p, O._u8p, O._f32p,
                                    ctypes.c_int, O._f32p]
    n = tA.shape[0]
    out = np.zeros((n, 4), np.float32)
    # warm-up on a few IAs, then time the sample
    L.or_cpu_path_batch(32, N, tA[:32], tB[:32], mA[:32], 1, out[:32])
    t0 = time.perf_counter()
    L.or_cpu_path_batch(n, N, tA, tB, mA, 1, out)
    dt = time.perf_counter() - t0
    return {"value": round(n / dt, 1), "unit": "IAs/s", "cores": 1, "kind": "port",
            "sample": f"{n} IAs of {N}x{N} from one 2048x2048 pair (Java CPU path restated in C, "
                      f"-O2 -ffp-contract=off, 1 thread), {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pairs", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one MI355X per box); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import qlpiv
    from qlpiv.geometry import level_grid
    from qlpiv.synth import particle_pair

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda:0")
    P = args.pairs
    print(f"[bench] generating {P} synthetic {H}x{W} pairs", file=sys.stderr, flush=True)
    pairs = [particle_pair(H, W, seed=1000 + 100 * 1 + k, shift=(3.3, 3.3)) for k in range(P)]
    tops, lefts = level_grid(H, W, N, 0.5)
    nI, nJ = tops.shape
    n_pair = tops.size
    # predictor displacement of the coarser level: the true field at IA centres
    cy = tops.ravel() + N / 2.0 - 0.5
    cx = lefts.ravel() + N / 2.0 - 0.5
    imgsA, imgsB, preds = [], [], []
    for a, b, field in pairs:
        dx, dy = field(cx, cy)
        imgsA.append(torch.from_numpy(a).to(dev))
        imgsB.append(torch.from_numpy(b).to(dev))
        preds.append((torch.from_numpy(dy.astype(np.float32)).to(dev),
                      torch.from_numpy(dx.astype(np.float32)).to(dev)))
    top_d = torch.from_numpy(tops.ravel().astype(np.int32)).to(dev)
    left_d = torch.from_numpy(lefts.ravel().astype(np.int32)).to(dev)
    n_step = n_pair * P

    ctx = qlpiv.Context(0)
    s = ctx.torch_stream()
    tilesA = torch.empty((P, n_pair, N, N), dtype=torch.uint8, device=dev)
    tilesB = torch.empty_like(tilesA)
    maxA = torch.empty((P, n_pair), dtype=torch.float32, device=dev)
    maxB = torch.empty_like(maxA)
    planes = torch.empty((n_step, 2 * N - 1, 2 * N - 1), dtype=torch.float32, device=dev)
    peaks = torch.empty((n_step, 1, 3), dtype=torch.float32, device=dev)
    mfloor = torch.empty(n_step, dtype=torch.float32, device=dev)
    sub = torch.empty((n_step, 1, 4), dtype=torch.float32, device=dev)
    lib = ctx.lib
    from qlpiv import _lib

    ev = []

    def step(timed):
        for p in range(P):
            _lib.check(lib.qlp_clip_tiles_dev(ctx.h, imgsA[p].data_ptr(), H, W, n_pair, N,
                                              top_d.data_ptr(), left_d.data_ptr(),
                                              tilesA[p].data_ptr(), maxA[p].data_ptr()))
            _lib.check(lib.qlp_warp_micro_tiles_dev(ctx.h, imgsB[p].data_ptr(), H, W, n_pair, N,
                                                    top_d.data_ptr(), left_d.data_ptr(),
                                                    preds[p][0].data_ptr(), preds[p][1].data_ptr(),
                                                    tilesB[p].data_ptr(), maxB[p].data_ptr()))
        if timed:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
        _lib.check(lib.qlp_xcorr_batch_dev(ctx.h, n_step, N, tilesA.data_ptr(), tilesB.data_ptr(),
                                           maxA.data_ptr(), maxB.data_ptr(),
                                           qlpiv.QLP_NORM_PER_TILE, planes.data_ptr()))
        if timed:
            e1.record(s)
            ev.append((e0, e1))
        _lib.check(lib.qlp_find_peaks_dev(ctx.h, qlpiv.QLP_PEAK_FROM_CENTER, n_step, 2 * N - 1,
                                          planes.data_ptr(), 1, 3, peaks.data_ptr(),
                                          mfloor.data_ptr()))
        _lib.check(lib.qlp_subpixel_dev(ctx.h, qlpiv.QLP_SUBPIX_GUO, 3, 20, n_step, 2 * N - 1, 1,
                                        planes.data_ptr(), mfloor.data_ptr(), peaks.data_ptr(),
                                        sub.data_ptr()))

    for _ in range(args.warmup):
        step(False)
    ctx.synchronize()
    torch.cuda.synchronize()
    print(f"[bench] warmup done; timing {args.steps} steps of {n_step} IAs", file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    ctx.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    xc_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    value = n_step * args.steps / dt
    achieved = ia_bytes(n_step) / (xc_ms * 1e-3) / 1e9

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "xcorr_n32_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        if t.get("n_ia") == n_step:
            traffic = t.get("hbm_bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] timing the CPU baseline sample", file=sys.stderr, flush=True)
        tA = tilesA[0].cpu().numpy()
        tB = tilesB[0].cpu().numpy()
        mA = maxA[0].cpu().numpy()
        cpu = cpu_baseline(np.ascontiguousarray(tA), np.ascontiguousarray(tB),
                           np.ascontiguousarray(mA))
    ctx.close()
    line = {
        "metric": "correlated IAs/sec at 32x32",
        "value": round(value, 1),
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
        "config": {"workload": f"plain FP32 form, 32x32 level of C2: {P} x 2048^2 pairs, "
                               f"{n_pair} IAs/pair at 50% overlap, clip+micro-warp+FFT xcorr "
                               f"(per-tile norm)+FromCenter+Guo 3px/20it",
                   "ia": N, "ias_per_step": n_step, "pairs_per_step": P},
        "roofline": {"bound": "hbm", "kernel": "xcorr_exact_kernel<6,false>",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "avg_launch_ms": round(xc_ms, 4),
                     "algorithmic_bytes_per_launch": ia_bytes(n_step)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
