"""CPU oracle for the PIV correlation path -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check in __graft_entry__.py and the cpu_baseline leg of
bench.py may import this module.  It wraps:

* ``_build/libqlpiv_oracle.so`` -- the C restatement of the reference's Java
  code (qlpiv_oracle.c; built by ``make -C oracle``);
* FP64 numpy oracles that follow the mathematics rather than the Java
  operation order (direct-sum and FFT correlation);
* readers for the reference's fixture formats (MATFloat).
"""
from __future__ import annotations

import ctypes
import os
import struct

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# QLPIV_ORACLE_LIB selects another build of the same source (the sanitizer
# build of tests/sanitize/).
_LIB_PATH = os.environ.get("QLPIV_ORACLE_LIB") or os.path.join(_HERE, "_build", "libqlpiv_oracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build():
    """Compile the C restatement (gcc; seconds)."""
    import subprocess

    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_int, c_float = ctypes.c_int, ctypes.c_float
        L.or_sine_table.argtypes = [c_int, _f32p]
        L.or_perfect_shuffle.argtypes = [_f32p, c_int]
        L.or_xcorr_javacpu.argtypes = [_f32p, _f32p, c_float, c_int, _f32p]
        L.or_xcorr_floats.argtypes = [_f32p, _f32p, c_int, _f32p]
        L.or_xcorr_javagpu.argtypes = [_f32p, _f32p, c_float, c_float, c_int, _f32p]
        L.or_xcorr_batch_u8.argtypes = [c_int, c_int, _u8p, _u8p, _f32p, _f32p, c_int, _f32p]
        L.or_find_max_from_center.argtypes = [_f32p, c_int, c_int, _f32p]
        L.or_find_max_simple.argtypes = [_f32p, c_int, c_int, _f32p]
        L.or_level_batch.argtypes = [_u8p, _u8p, c_int, c_int, c_int, c_int, _i32p, _i32p, _f32p,
                                     _f32p, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                     ctypes.c_void_p, _f32p, _i32p]
        L.or_list_script.argtypes = [c_int, c_int, c_int, _f32p, _i32p, _f32p]
        L.or_list_script.restype = c_int
        L.or_find_max_multipeaks.argtypes = [_f32p, c_int, c_int, c_int, c_int, _f32p, _f32p]
        L.or_find_max_multipeaks.restype = c_int
        L.or_gauss1d.argtypes = [_f32p, c_int, c_int, c_int, c_float, _f32p, _f32p]
        L.or_guo.argtypes = [_f32p, c_int, c_int, c_int, c_int, c_float, _f32p, _f32p]
        L.or_poly.argtypes = [_f32p, c_int, c_int, _f32p, _f32p]
        L.or_guo_coefs.argtypes = [c_int, c_int, _f32p, _f32p, _f64p]
        L.or_pinv_solve3.argtypes = [_f64p, _f64p, _f64p]
        L.or_lu_singular3.argtypes = [_f64p]
        L.or_lu_singular3.restype = c_int
        L.or_gauss_lsq.argtypes = [c_int, _f64p, _f64p, _f64p]
        L.or_clip_tile.argtypes = [_u8p, c_int, c_int, c_int, c_int, c_int, _u8p]
        L.or_clip_tile.restype = c_float
        L.or_clip_tile_f32.argtypes = [_f32p, c_int, c_int, c_int, c_int, c_int, _f32p]
        L.or_clip_tile_f32.restype = c_float
        L.or_warp_micro_side_f32.argtypes = [_f32p, c_int, c_int, c_int, c_int, c_int, c_float,
                                             c_float, c_int, _f32p]
        L.or_warp_micro_side_f32.restype = c_float
        _u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
        L.or_clip_tile_u16.argtypes = [_u16p, c_int, c_int, c_int, c_int, c_int, _f32p]
        L.or_clip_tile_u16.restype = c_float
        L.or_warp_micro_side_u16.argtypes = [_u16p, c_int, c_int, c_int, c_int, c_int, c_int,
                                             c_float, c_float, c_int, _f32p]
        L.or_warp_micro_side_u16.restype = c_float
        L.or_warp_micro.argtypes = [_u8p, c_int, c_int, c_int, c_int, c_int, c_float, c_float, _u8p]
        L.or_warp_micro_side.argtypes = [_u8p, c_int, c_int, c_int, c_int, c_int, c_float, c_float,
                                         c_int, _u8p]
        L.or_spiral_order.argtypes = [c_int, c_int, _i32p]
        L.or_spiral_order.restype = c_int
        L.or_validate.argtypes = [c_int, c_int, _i32p, _i32p, _f32p, _f32p, c_int, _f32p,
                                  ctypes.c_void_p, c_int, ctypes.c_void_p, c_int, c_int, c_int,
                                  _u8p, _i32p]
        for fn in ("or_inherit_area", "or_inherit_distance", "or_inherit_bicubic"):
            f = getattr(L, fn)
            f.argtypes = [c_int, c_int, c_int, c_int, _i32p, _i32p, _f32p, _f32p, c_int, c_int,
                          c_int, c_int, _i32p, _i32p, c_int, _f32p, _f32p]
        L.or_inherit_area.restype = c_int
        L.or_inherit_distance.restype = c_int
        L.or_inherit_bicubic.restype = None
        L.or_piecewise_bicubic.argtypes = [_f64p, c_int, _f64p, c_int, _f64p, ctypes.c_double,
                                           ctypes.c_double]
        L.or_piecewise_bicubic.restype = ctypes.c_double
        L.or_gauss_kernel.argtypes = [c_float, c_int, _f32p]
        L.or_gaussian_filter2d.argtypes = [_f32p, c_int, c_int, _f32p, c_int, _f32p]
        L.or_convolve1d.argtypes = [_f32p, c_int, _f32p, c_int, _f32p]
        L.or_natural_slopes.argtypes = [c_int, _f64p, _f64p, _f64p]
        L.or_spline_surface.argtypes = [c_int, _f64p, c_int, _f64p, _f64p, _f64p, _f64p, _f64p]
        L.or_spline_surface_value.argtypes = [c_int, _f64p, c_int, _f64p, _f64p, _f64p, _f64p,
                                              _f64p, ctypes.c_double, ctypes.c_double]
        L.or_spline_surface_value.restype = ctypes.c_double
        L.or_bilinear_value.argtypes = [c_int, _f64p, c_int, _f64p, _f64p, ctypes.c_double,
                                        ctypes.c_double]
        L.or_bilinear_value.restype = ctypes.c_double
        L.or_warp_mini.argtypes = [_u8p, c_int, c_int, c_int, c_int, _i32p, _i32p, c_int, _f32p,
                                   _f32p, c_int, c_int, _u8p]
        L.or_warp_mini_side.argtypes = [_u8p, c_int, c_int, c_int, c_int, _i32p, _i32p, c_int,
                                        _f32p, _f32p, c_int, c_int, c_int, _u8p]
        L.or_warp_mini_side_f32.argtypes = [_f32p, c_int, c_int, c_int, c_int, _i32p, _i32p, c_int,
                                            _f32p, _f32p, c_int, c_int, c_int, _f32p]
        L.or_warp_mini_side_u16.argtypes = [np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS"),
                                            c_int, c_int, c_int, c_int, _i32p, _i32p, c_int, _f32p,
                                            _f32p, c_int, c_int, c_int, _f32p]
        L.or_full_field.argtypes = [c_int, c_int, _i32p, _i32p, c_int, c_int, _f32p, _f32p, c_int,
                                    c_int, _f32p, _f32p]
        L.or_warp_image.argtypes = [_u8p, c_int, c_int, _f32p, _f32p, _u8p]
        L.or_warp_image_side.argtypes = [_u8p, c_int, c_int, _f32p, _f32p, c_int, _u8p]
        L.or_warp_image_side_f32.argtypes = [_f32p, c_int, c_int, _f32p, _f32p, c_int, _f32p]
        _lib = L
    return _lib


def inherit(kind, parent, child, pu, pv, round_it=False):
    """Velocity inheritance of one frame.  kind: "area", "distance" or
    "bicubic"; parent / child: (top, left, tile_size) with nI x nJ origin
    arrays; pu, pv: parent displacements (pnI x pnJ).  Returns (u, v, rc)."""
    ptop, pleft, psz = parent
    ctop, cleft, csz = child
    ptop = np.ascontiguousarray(ptop, np.int32)
    pleft = np.ascontiguousarray(pleft, np.int32)
    ctop = np.ascontiguousarray(ctop, np.int32)
    cleft = np.ascontiguousarray(cleft, np.int32)
    pnI, pnJ = ptop.shape
    cnI, cnJ = ctop.shape
    u = np.zeros((cnI, cnJ), np.float32)
    v = np.zeros((cnI, cnJ), np.float32)
    f = getattr(lib(), "or_inherit_" + kind)
    rc = f(pnI, pnJ, int(psz), int(psz), ptop, pleft, np.ascontiguousarray(pu, np.float32),
           np.ascontiguousarray(pv, np.float32), cnI, cnJ, int(csz), int(csz), ctop, cleft,
           int(bool(round_it)), u, v)
    return u, v, (0 if rc is None else rc)


def piecewise_bicubic(xval, yval, f, x, y):
    """commons-math3 PiecewiseBicubicSplineInterpolatingFunction.value(x, y)."""
    xval = np.ascontiguousarray(xval, np.float64)
    yval = np.ascontiguousarray(yval, np.float64)
    f = np.ascontiguousarray(f, np.float64)
    return lib().or_piecewise_bicubic(xval, len(xval), yval, len(yval), f, float(x), float(y))


def gauss_kernel(sigma, kernel_px):
    k = np.zeros(kernel_px, np.float32)
    lib().or_gauss_kernel(float(sigma), int(kernel_px), k)
    return k


def convolve1d(a, b):
    """DirectConvolution.convolve1D(a, |a|, b, |b|): the longer vector's length."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = np.zeros(max(a.size, b.size), np.float32)
    lib().or_convolve1d(a, a.size, b, b.size, out)
    return out


def gaussian_filter2d(img, sigma, kernel_px):
    """GaussianFilter2D.applyFilter on a float image (rows, then columns)."""
    img = np.ascontiguousarray(img, np.float32)
    out = np.zeros_like(img)
    k = gauss_kernel(sigma, kernel_px)
    lib().or_gaussian_filter2d(img, img.shape[0], img.shape[1], k, int(kernel_px), out)
    return out


class VStage(ctypes.Structure):
    """or_vstage: one validator + replacement stage."""

    _fields_ = [("validator", ctypes.c_int), ("threshold", ctypes.c_float),
                ("epsilon0", ctypes.c_float), ("replacement", ctypes.c_int),
                ("peak_index", ctypes.c_int)]


def spiral_order(nI, nJ):
    """Visit order of VectorValidatorJob.circularVisitor; None on a double visit."""
    order = np.full(nI * nJ, -1, np.int32)
    got = lib().or_spiral_order(nI, nJ, order)
    return None if got != nI * nJ else order


def validate(top, left, u, v, stages, peak_uv=None, peak_count=None, max_iterations=1,
             until_no_more=False, max_sweeps=256):
    """VectorValidatorJob.compute over one grid.

    top, left: nI x nJ tile origins; u, v: nI x nJ displacements; stages: list
    of (validator, threshold, epsilon0, replacement, peak_index); peak_uv:
    nI x nJ x P x 2 increments.  Returns (u, v, invalid, stats)."""
    top = np.ascontiguousarray(top, np.int32)
    nI, nJ = top.shape
    left = np.ascontiguousarray(left, np.int32)
    u = np.array(u, np.float32, copy=True, order="C")
    v = np.array(v, np.float32, copy=True, order="C")
    if peak_uv is None:
        peak_uv = np.zeros((nI, nJ, 1, 2), np.float32)
    peak_uv = np.ascontiguousarray(peak_uv, np.float32)
    P = peak_uv.shape[2]
    cnt = None if peak_count is None else np.ascontiguousarray(peak_count, np.int32)
    arr = (VStage * len(stages))(*[VStage(int(a), float(b), float(c), int(d), int(e))
                                   for a, b, c, d, e in stages])
    invalid = np.zeros((nI, nJ), np.uint8)
    stats = np.zeros(4, np.int32)
    lib().or_validate(nI, nJ, top, left, u, v, P, peak_uv,
                      None if cnt is None else cnt.ctypes.data, len(stages),
                      ctypes.cast(arr, ctypes.c_void_p), int(max_iterations), int(until_no_more),
                      int(max_sweeps), invalid, stats)
    return u, v, invalid, stats


# --------------------------------------------------------------------------
# Restatement wrappers
# --------------------------------------------------------------------------

def sine_table(n):
    out = np.zeros(n, np.float32)
    lib().or_sine_table(n, out)
    return out


def perfect_shuffle(a):
    a = np.ascontiguousarray(a, np.float32).copy()
    lib().or_perfect_shuffle(a, a.size)
    return a


def xcorr_javacpu(A, B, maxA):
    """FastRealFFTXCorr.computeXCorr(Matrix, Matrix) on float tiles."""
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    N = A.shape[0]
    out = np.zeros((2 * N - 1, 2 * N - 1), np.float32)
    lib().or_xcorr_javacpu(A, B, float(maxA), N, out)
    return out


def xcorr_floats(A, B):
    """FastRealFFTXCorr.computeXCorr(float[][], float[][]) (no normalisation)."""
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    N = A.shape[0]
    out = np.zeros((2 * N - 1, 2 * N - 1), np.float32)
    lib().or_xcorr_floats(A, B, N, out)
    return out


def xcorr_javagpu(A, B, maxA, maxB):
    """CrossCorrelationRealFFTParStdJob + kernel arithmetic (Java semantics)."""
    A = np.ascontiguousarray(A, np.float32)
    B = np.ascontiguousarray(B, np.float32)
    N = A.shape[0]
    out = np.zeros((2 * N - 1, 2 * N - 1), np.float32)
    lib().or_xcorr_javagpu(A, B, float(maxA), float(maxB), N, out)
    return out


def xcorr_batch_u8(tilesA, tilesB, maxA, maxB, mode):
    """mode 0: Java CPU path (A max for both); mode 1: GPU job (own maxima)."""
    tilesA = np.ascontiguousarray(tilesA, np.uint8)
    tilesB = np.ascontiguousarray(tilesB, np.uint8)
    n, N = tilesA.shape[0], tilesA.shape[1]
    out = np.zeros((n, 2 * N - 1, 2 * N - 1), np.float32)
    lib().or_xcorr_batch_u8(n, N, tilesA, tilesB, np.ascontiguousarray(maxA, np.float32),
                            np.ascontiguousarray(maxB, np.float32), int(mode), out)
    return out


def find_max_from_center(plane):
    plane = np.ascontiguousarray(plane, np.float32)
    out = np.zeros(4, np.float32)
    lib().or_find_max_from_center(plane, plane.shape[0], plane.shape[1], out)
    return out  # i, j, value, minFloor


def level_batch(a, b, top, left, u, v, N, warp, mode=0, peak_kind=0, P=1, k=3, fit_kind=1,
                pixels=3, iterations=20, planes=False):
    """One level of the per-frame loop for the IAs at (top, left) of the pair
    (a, b) from the inherited displacements (u, v) (or_level_batch).  Returns
    (out: n x P x 4 fits, count: n, planes: n x D x D or None)."""
    a = np.ascontiguousarray(a, np.uint8)
    b = np.ascontiguousarray(b, np.uint8)
    top = np.ascontiguousarray(np.ravel(top), np.int32)
    left = np.ascontiguousarray(np.ravel(left), np.int32)
    n = top.size
    u = np.ascontiguousarray(np.ravel(u), np.float32) if u is not None else np.zeros(n, np.float32)
    v = np.ascontiguousarray(np.ravel(v), np.float32) if v is not None else np.zeros(n, np.float32)
    np_ = P if peak_kind == 1 else 1
    out = np.zeros((n, np_, 4), np.float32)
    cnt = np.zeros(n, np.int32)
    D = 2 * N - 1
    pl = np.empty((n, D, D), np.float32) if planes else None
    lib().or_level_batch(a, b, a.shape[0], a.shape[1], n, N, top, left, u, v, int(warp), int(mode),
                         int(peak_kind), int(P), int(k), int(fit_kind), int(pixels),
                         int(iterations), None if pl is None else pl.ctypes.data, out, cnt)
    return out, cnt, pl


def find_max_simple(plane):
    plane = np.ascontiguousarray(plane, np.float32)
    out = np.zeros(4, np.float32)
    lib().or_find_max_simple(plane, plane.shape[0], plane.shape[1], out)
    return out  # i, j, value, minFloor


def list_script(max_size, kernel_width, ops):
    """ops: list of (opcode, value, i, j); 0=addInAscendOrder, 1=remove, 2=clear."""
    arr = np.ascontiguousarray(np.array(ops, np.float32).reshape(-1, 4))
    acc = np.zeros(len(ops), np.int32)
    out = np.zeros(3 * 70, np.float32)
    n = lib().or_list_script(max_size, kernel_width, len(ops), arr, acc, out)
    return [tuple(out[3 * k:3 * k + 3]) for k in range(n)], acc


def find_max_multipeaks(plane, n_peaks, kernel_width):
    plane = np.ascontiguousarray(plane, np.float32)
    out = np.zeros(3 * n_peaks, np.float32)
    mf = np.zeros(1, np.float32)
    found = lib().or_find_max_multipeaks(plane, plane.shape[0], plane.shape[1], n_peaks,
                                         kernel_width, out, mf)
    return out.reshape(n_peaks, 3), int(found), float(mf[0])


def gauss1d(plane, pixels, min_floor, peak):
    plane = np.ascontiguousarray(plane, np.float32)
    res = np.zeros(4, np.float32)
    lib().or_gauss1d(plane, plane.shape[0], plane.shape[1], pixels, float(min_floor),
                     np.ascontiguousarray(peak, np.float32), res)
    return res


def guo(plane, pixels, iterations, min_floor, peak):
    plane = np.ascontiguousarray(plane, np.float32)
    res = np.zeros(4, np.float32)
    lib().or_guo(plane, plane.shape[0], plane.shape[1], pixels, iterations, float(min_floor),
                 np.ascontiguousarray(peak, np.float32), res)
    return res


def guo_coefs(xs, ys, iterations):
    """Coefficients (a, b, c) of ln y = a + b x + c x^2 after 'iterations' solves
    of Guo's weighted least squares (float xs, ys as the reference holds them)."""
    xs = np.ascontiguousarray(xs, np.float32)
    ys = np.ascontiguousarray(ys, np.float32)
    out = np.zeros(3, np.float64)
    lib().or_guo_coefs(int(iterations), xs.size, xs, ys, out)
    return out


def pinv_solve3(A, b):
    out = np.zeros(3, np.float64)
    lib().or_pinv_solve3(np.ascontiguousarray(A, np.float64).ravel(),
                         np.ascontiguousarray(b, np.float64), out)
    return out


def lu_singular3(A):
    return bool(lib().or_lu_singular3(np.ascontiguousarray(A, np.float64).ravel()))


def gauss_lsq(x, y):
    """(a, b, c) of the least-squares Gaussian a exp(-(x-b)^2 / (2c^2))."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    out = np.zeros(3, np.float64)
    lib().or_gauss_lsq(x.size, x, y, out)
    return out


def poly(plane, peak):
    plane = np.ascontiguousarray(plane, np.float32)
    res = np.zeros(4, np.float32)
    lib().or_poly(plane, plane.shape[0], plane.shape[1], np.ascontiguousarray(peak, np.float32),
                  res)
    return res


def clip_tile(img, top, left, N):
    img = np.ascontiguousarray(img, np.uint8)
    tile = np.zeros((N, N), np.uint8)
    mx = lib().or_clip_tile(img, img.shape[0], img.shape[1], int(top), int(left), N, tile)
    return tile, mx


# sides of a tile warp: which frame is sampled in which order
# (BiLinearImageWarpingModeEnum): B in SecondImage, A in FirstImage, A and B in
# BothImages
SIDE_B_SECOND, SIDE_A_FIRST, SIDE_A_BOTH, SIDE_B_BOTH = 0, 1, 2, 3


def warp_micro(img, top, left, N, du, dv, side=SIDE_B_SECOND):
    img = np.ascontiguousarray(img, np.uint8)
    tile = np.zeros((N, N), np.uint8)
    lib().or_warp_micro_side(img, img.shape[0], img.shape[1], int(top), int(left), N, float(du),
                             float(dv), int(side), tile)
    return tile


def clip_tile_f32(img, top, left, N):
    """N x N window of a float frame, zero outside; returns (tile, max), the
    max over the tile with 0 -> 1 (no >= 16 clamp for float tiles)."""
    img = np.ascontiguousarray(img, np.float32)
    tile = np.zeros((N, N), np.float32)
    mx = lib().or_clip_tile_f32(img, img.shape[0], img.shape[1], int(top), int(left), N, tile)
    return tile, mx


def warp_micro_f32(img, top, left, N, du, dv, side=SIDE_B_SECOND):
    """Micro warp of a float frame's tile; returns (tile, 65535)."""
    img = np.ascontiguousarray(img, np.float32)
    tile = np.zeros((N, N), np.float32)
    mx = lib().or_warp_micro_side_f32(img, img.shape[0], img.shape[1], int(top), int(left), N,
                                      float(du), float(dv), int(side), tile)
    return tile, mx


def clip_tile_u16(img, top, left, N):
    """N x N window of a 10/12/16-bit frame as floats; (tile, max >= 16)."""
    img = np.ascontiguousarray(img, np.uint16)
    tile = np.zeros((N, N), np.float32)
    mx = lib().or_clip_tile_u16(img, img.shape[0], img.shape[1], int(top), int(left), N, tile)
    return tile, mx


def warp_micro_u16(img, bits, top, left, N, du, dv, side=SIDE_B_SECOND):
    """Micro warp of a 10/12/16-bit frame's tile: (float tile, 2^bits - 1)."""
    img = np.ascontiguousarray(img, np.uint16)
    tile = np.zeros((N, N), np.float32)
    mx = lib().or_warp_micro_side_u16(img, img.shape[0], img.shape[1], int(bits), int(top), int(left),
                                      N, float(du), float(dv), int(side), tile)
    return tile, mx


def natural_slopes(x, y):
    """Slopes at the knots of the natural cubic spline through (x, y)."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    s = np.zeros_like(x)
    lib().or_natural_slopes(x.size, x, y, s)
    return s


def spline_surface(xs, ys, f):
    """(fx, fy, fxy) knot slopes of the bicubic spline surface, f[i, j] at (xs[i], ys[j])."""
    xs = np.ascontiguousarray(xs, np.float64)
    ys = np.ascontiguousarray(ys, np.float64)
    f = np.ascontiguousarray(f, np.float64)
    fx, fy, fxy = (np.zeros_like(f) for _ in range(3))
    lib().or_spline_surface(xs.size, xs, ys.size, ys, f, fx, fy, fxy)
    return fx, fy, fxy


def spline_surface_value(xs, ys, f, x, y):
    xs = np.ascontiguousarray(xs, np.float64)
    ys = np.ascontiguousarray(ys, np.float64)
    f = np.ascontiguousarray(f, np.float64)
    fx, fy, fxy = spline_surface(xs, ys, f)
    return lib().or_spline_surface_value(xs.size, xs, ys.size, ys, f, fx, fy, fxy, float(x),
                                         float(y))


def bilinear_value(ys, xs, v, y, x):
    ys = np.ascontiguousarray(ys, np.float64)
    xs = np.ascontiguousarray(xs, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    return lib().or_bilinear_value(ys.size, ys, xs.size, xs, v, float(y), float(x))


def warp_mini(img, top, left, N, u, v, ti, tj, side=SIDE_B_SECOND):
    """Mini-warped tile (ti, tj) of a grid (top, left: nI x nJ origins, u, v: the
    grid's displacements) for a side (SIDE_*)."""
    img = np.ascontiguousarray(img, np.uint8)
    top = np.ascontiguousarray(top, np.int32)
    left = np.ascontiguousarray(left, np.int32)
    u = np.ascontiguousarray(u, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    tile = np.zeros((N, N), np.uint8)
    lib().or_warp_mini_side(img, img.shape[0], img.shape[1], top.shape[0], top.shape[1], top, left,
                            N, u, v, int(ti), int(tj), int(side), tile)
    return tile


def warp_mini_frame(img, top, left, N, u, v, ti, tj, side=SIDE_B_SECOND):
    """Mini-warped tile of a float (ImageFloat) or uint16 (ImageInt16) frame,
    as float samples."""
    top = np.ascontiguousarray(top, np.int32)
    left = np.ascontiguousarray(left, np.int32)
    u = np.ascontiguousarray(u, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    tile = np.zeros((N, N), np.float32)
    if img.dtype == np.uint16:
        img = np.ascontiguousarray(img)
        fn = lib().or_warp_mini_side_u16
    else:
        img = np.ascontiguousarray(img, np.float32)
        fn = lib().or_warp_mini_side_f32
    fn(img, img.shape[0], img.shape[1], top.shape[0], top.shape[1], top, left, N, u, v, int(ti),
       int(tj), int(side), tile)
    return tile


def full_field(top, left, N, u, v, H, W):
    top = np.ascontiguousarray(top, np.int32)
    left = np.ascontiguousarray(left, np.int32)
    u = np.ascontiguousarray(u, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    fu = np.zeros((H, W), np.float32)
    fv = np.zeros((H, W), np.float32)
    lib().or_full_field(top.shape[0], top.shape[1], top, left, N, N, u, v, H, W, fu, fv)
    return fu, fv


def warp_image(img, fu, fv, side=SIDE_B_SECOND):
    img = np.ascontiguousarray(img, np.uint8)
    out = np.zeros_like(img)
    lib().or_warp_image_side(img, img.shape[0], img.shape[1], np.ascontiguousarray(fu, np.float32),
                             np.ascontiguousarray(fv, np.float32), int(side), out)
    return out


def warp_image_f32(img, fu, fv, side=SIDE_B_SECOND):
    """Full warp of a float frame (ImageFloat): rounded samples as floats."""
    img = np.ascontiguousarray(img, np.float32)
    out = np.zeros_like(img)
    lib().or_warp_image_side_f32(img, img.shape[0], img.shape[1], np.ascontiguousarray(fu, np.float32),
                                 np.ascontiguousarray(fv, np.float32), int(side), out)
    return out


# --------------------------------------------------------------------------
# FP64 mathematics oracles
# --------------------------------------------------------------------------

def normalise_pair(A, B, maxF, maxG):
    """x / max * 16 + 1 in float32 as Java evaluates it; F mirrored."""
    A = np.asarray(A, np.float32)
    B = np.asarray(B, np.float32)
    F = (A[::-1, ::-1] / np.float32(maxF)) * np.float32(16.0) + np.float32(1.0)
    G = (B / np.float32(maxG)) * np.float32(16.0) + np.float32(1.0)
    return F.astype(np.float32), G.astype(np.float32)


def xcorr_direct64(Fm, G):
    """R[i][j] = sum_nm Fm[n][m] * G[n+i-(N-1)][m+j-(N-1)] where Fm is the
    mirrored, normalised tile -- the convolution the reference computes
    (CrossCorrelationTestHelper.localCrossCorrelation :32-73 for the
    unmirrored form).  FP64 brute force; for small N."""
    Fm = np.asarray(Fm, np.float64)
    G = np.asarray(G, np.float64)
    N = Fm.shape[0]
    F = Fm[::-1, ::-1]
    D = 2 * N - 1
    R = np.zeros((D, D))
    for i in range(-(N - 1), N):
        for j in range(-(N - 1), N):
            n0, n1 = max(0, -i), min(N, N - i)
            m0, m1 = max(0, -j), min(N, N - j)
            R[i + N - 1, j + N - 1] = np.sum(F[n0:n1, m0:m1] * G[n0 + i:n1 + i, m0 + j:m1 + j])
    return R


def xcorr_fft64(Fm, G):
    """Same plane through an FP64 FFT of the zero-padded 2N x 2N product."""
    Fm = np.asarray(Fm, np.float64)
    G = np.asarray(G, np.float64)
    N = Fm.shape[-1]
    M = 2 * N
    shp = Fm.shape[:-2] + (M, M)
    a = np.zeros(shp)
    b = np.zeros(shp)
    a[..., :N, :N] = Fm
    b[..., :N, :N] = G
    r = np.fft.ifft2(np.fft.fft2(a) * np.fft.fft2(b)).real
    return r[..., :M - 1, :M - 1]


# --------------------------------------------------------------------------
# Fixture formats
# --------------------------------------------------------------------------

def read_matfloat(path):
    """SimpleFloatMatrixImporterExporter format (exporter/SimpleFloatMatrixImporterExporter
    .java:263-314): 'MATFloat\\0', int16 count, names, then per matrix
    int16 sizeX, int16 sizeY, int16 0, big-endian float32 data."""
    with open(path, "rb") as f:
        d = f.read()
    assert d[:9] == b"MATFloat\x00"
    off = 9
    (n,) = struct.unpack(">h", d[off:off + 2])
    off += 2
    names = []
    for _ in range(n):
        (ln,) = struct.unpack(">h", d[off:off + 2])
        off += 2
        names.append(d[off:off + ln].decode())
        off += ln
    mats = []
    for _ in range(n):
        sx, sy, _z = struct.unpack(">hhh", d[off:off + 6])
        off += 6
        m = np.frombuffer(d, ">f4", sx * sy, off).reshape(sy, sx).astype(np.float32)
        off += 4 * sx * sy
        mats.append(m)
    return names, mats
