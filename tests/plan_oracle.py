"""Level-by-level check of a device multi-pass plan against the oracle.

Each level is recomputed by the oracle (oracle.level_batch: clip, B tile by the
level's warp, the correlation job's arithmetic, the peak finder and the fit of
every peak) from the displacements the device level inherited, which are
rebuilt on the host from the previous level's device output (Direct
inheritance).  Checked per level:
  * stored planes value-identical (float ==) on every IA;
  * fit status equal and fitted positions within SUBPIX_TOL of the oracle's on
    every peak found (the integer peaks then agree: a different integer peak
    moves the fit by about a pixel);
  * the displacement update bit-exact from the device's own fits, and within
    SUBPIX_TOL of the oracle's;
  * with validation, the validated field, flags and statistics bit-exact
    against the oracle validator run on the device's own fits.
"""
import numpy as np

import oracle as O

SUBPIX_TOL = 1e-4


def direct_inherit(pu, nI, nJ, round_it=False):
    """AdaptiveInterVelocityInheritanceStrategyDirect: child (i, j) takes parent
    (i // 2, j // 2) when it exists, else 0; NoImageWarping rounds it."""
    out = np.zeros((nI, nJ), np.float32)
    pI, pJ = pu.shape
    ci, cj = min(nI, 2 * pI), min(nJ, 2 * pJ)
    src = pu[np.arange(ci)[:, None] // 2, np.arange(cj)[None, :] // 2]
    if round_it:
        src = np.floor(src + np.float32(0.5)).astype(np.float32)
    out[:ci, :cj] = src
    return out


def check_plan_levels(plan, a, b, *, warp_mode, norm_mode, peak_kind, n_peaks=1, kernel_width=3,
                      fit_kind=1, pixels=3, iterations=20, start_step=0, stages=None,
                      max_iterations=0, log=print):
    prev_u = prev_v = None
    for lv, L in enumerate(plan.levels):
        u, v, sub = (x[0] for x in plan.result(lv))
        N, nI, nJ = L["ia"], L["nI"], L["nJ"]
        n, D = nI * nJ, 2 * N - 1
        c = np.float32(D // 2)
        if lv == 0:
            uin = np.zeros((nI, nJ), np.float32)
            vin = np.zeros((nI, nJ), np.float32)
        else:
            uin = direct_inherit(prev_u, nI, nJ, round_it=(warp_mode == 0))
            vin = direct_inherit(prev_v, nI, nJ, round_it=(warp_mode == 0))
        warp = 0 if lv == 0 else (2 if warp_mode == 1 else 1)
        fit = fit_kind if lv >= start_step else -1
        gpl = plan.planes(lv)
        out, cnt, rpl = O.level_batch(a, b, L["top"], L["left"], uin, vin, N, warp, norm_mode,
                                      peak_kind, n_peaks, kernel_width, fit, pixels, iterations,
                                      planes=gpl is not None)
        if gpl is not None:
            gpl = gpl[0].reshape(n, D, D)
            bad = np.flatnonzero(~np.all(gpl.reshape(n, -1) == rpl.reshape(n, -1), axis=1))
            assert bad.size == 0, (lv, bad[:10])
            del gpl, rpl
        P = out.shape[1]
        sub = sub.reshape(n, P, 4)
        found = np.arange(P)[None, :] < cnt[:, None]
        assert np.array_equal(sub[..., 3][found], out[..., 3][found]), lv
        err = np.abs(sub[..., :2] - out[..., :2])[found]
        worst = float(err.max()) if err.size else 0.0
        assert worst <= SUBPIX_TOL, (lv, worst)
        inc = (sub[..., :2] - c).astype(np.float32)
        upre = (uin.ravel() + inc[:, 0, 0]).astype(np.float32).reshape(nI, nJ)
        vpre = (vin.ravel() + inc[:, 0, 1]).astype(np.float32).reshape(nI, nJ)
        eu = (uin.ravel() + (out[:, 0, 0] - c).astype(np.float32)).astype(np.float32)
        assert np.abs(upre.ravel() - eu).max() <= SUBPIX_TOL
        if stages:
            ru, rv, rinv, rst = O.validate(L["top"], L["left"], upre, vpre, stages,
                                           inc.reshape(nI, nJ, P, 2), cnt.reshape(nI, nJ),
                                           max_iterations, False)
            inv, st = plan.validation(lv)
            assert np.array_equal(u, ru) and np.array_equal(v, rv), lv
            assert np.array_equal(inv[0], rinv.astype(bool)) and list(st[0]) == list(rst), lv
            log(f"level {lv}: {N} px, {nI}x{nJ} IAs, worst fit difference {worst:.2e} px, "
                f"corrected {rst[1]}, left invalid {rst[2]}")
        else:
            assert np.array_equal(u, upre) and np.array_equal(v, vpre), lv
            log(f"level {lv}: {N} px, {nI}x{nJ} IAs, worst fit difference {worst:.2e} px")
        prev_u, prev_v = u, v
