"""Seeded synthetic PIV frame pairs.

8-bit frames of Gaussian particles: uniform positions at a given density
(particles per pixel), profile I0 * exp(-8 r^2 / d^2) with d = 2.5 px (the e^-2
diameter), I0 ~ U(150, 255), additive N(0, sigma^2) noise, clamp to [0, 255]
and round.  Frame B renders the same particle list moved once by the
displacement field evaluated at each particle's frame-A position (a uniform
shift plus a Rankine vortex centred in the image), so B carries no
interpolation bias.  Positions, intensities and noise use independent seeded
streams.
"""
from __future__ import annotations

import numpy as np


def rankine_displacement(x, y, W, H, core_radius=128.0, u_max=6.0):
    """Tangential Rankine vortex about the image centre: speed U r/R inside the
    core, U R/r outside.  Returns (dx, dy) in pixels (x = column, y = row)."""
    cx, cy = (W - 1) / 2.0, (H - 1) / 2.0
    rx, ry = x - cx, y - cy
    r = np.hypot(rx, ry)
    speed = np.where(r <= core_radius, u_max * r / core_radius,
                     u_max * core_radius / np.maximum(r, 1e-12))
    with np.errstate(invalid="ignore", divide="ignore"):
        ux = np.where(r > 0, -speed * ry / r, 0.0)
        uy = np.where(r > 0, speed * rx / r, 0.0)
    return ux, uy


def _render(H, W, xs, ys, i0, diameter, rng_noise, noise_sigma):
    rad = int(np.ceil(diameter)) + 1
    ix = np.floor(xs).astype(np.int64)
    iy = np.floor(ys).astype(np.int64)
    k = 8.0 / (diameter * diameter)
    # every (particle, offset) contribution in one bincount: per pixel the sum
    # runs over the offsets in order, particles in order within an offset
    idx, val = [], []
    for oy in range(-rad, rad + 2):
        for ox in range(-rad, rad + 2):
            px, py = ix + ox, iy + oy
            ok = (px >= 0) & (px < W) & (py >= 0) & (py < H)
            d2 = (px[ok] - xs[ok]) ** 2 + (py[ok] - ys[ok]) ** 2
            idx.append(py[ok] * W + px[ok])
            val.append(i0[ok] * np.exp(-k * d2))
    img = np.bincount(np.concatenate(idx), weights=np.concatenate(val),
                      minlength=H * W).reshape(H, W)
    if noise_sigma > 0:
        img += rng_noise.normal(0.0, noise_sigma, img.shape)
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def particle_pair(H, W, seed, density=0.05, diameter=2.5, noise_sigma=2.0, shift=(0.0, 0.0),
                  core_radius=128.0, u_max=6.0, vortex=True):
    """Return (frameA, frameB, field) with field(x, y) -> (dx, dy) the true
    displacement used for frame B.  shift = (dx, dy) uniform part."""
    rng_pos = np.random.default_rng([seed, 1])
    rng_int = np.random.default_rng([seed, 2])
    rng_noise = np.random.default_rng([seed, 3])
    npart = int(round(density * H * W))
    # particles may sit slightly outside the frame so that edges are populated
    margin = 16.0
    xs = rng_pos.uniform(-margin, W + margin, npart)
    ys = rng_pos.uniform(-margin, H + margin, npart)
    i0 = rng_int.uniform(150.0, 255.0, npart)

    def field(x, y):
        if vortex:
            ux, uy = rankine_displacement(x, y, W, H, core_radius, u_max)
        else:
            ux, uy = np.zeros_like(x), np.zeros_like(y)
        return ux + shift[0], uy + shift[1]

    dx, dy = field(xs, ys)
    a = _render(H, W, xs, ys, i0, diameter, rng_noise, noise_sigma)
    b = _render(H, W, xs + dx, ys + dy, i0, diameter, rng_noise, noise_sigma)
    return a, b, field


# The mixed batch workload: 64 pairs, 16 each at 512x512, 1024x1024, 1024x2048
# and 2048x2048 (H x W), the IA size of each pair drawn from {16, 32, 64, 128},
# per-pair Rankine vortex R ~ U(32, 256), U ~ U(2, 10).
BATCH_SIZES = [(512, 512), (1024, 1024), (1024, 2048), (2048, 2048)]


def batch_specs(seed=1700):
    """(H, W, ia, frame seed, core radius, max speed) of the 64 pairs."""
    rng = np.random.default_rng(seed)
    specs = []
    for (h, w) in BATCH_SIZES:
        for _ in range(16):
            specs.append((h, w, int(rng.choice([16, 32, 64, 128])), seed + len(specs),
                          float(rng.uniform(32, 256)), float(rng.uniform(2, 10))))
    return specs


def _spec_frames(spec):
    h, w, _ia, s, R, U = spec
    a, b, _ = particle_pair(h, w, seed=s, shift=(0.0, 0.0), core_radius=R, u_max=U)
    return a, b


def batch_frames(specs, workers=8):
    """The frame pairs of batch_specs, generated in a process pool (workers = 0:
    in this process, e.g. under a profiler that does not follow forks)."""
    if workers <= 0:
        return [_spec_frames(s) for s in specs]
    from multiprocessing import Pool

    with Pool(workers) as pool:
        return pool.map(_spec_frames, specs)
