"""qlpiv -- MI355X-native PIV correlation hot path of QuickLabPIV-ng.

Host-side Python mirror of the reference's correlation plug-in interfaces over
the C ABI of libqlpiv.so (include/qlpiv.h):

* ``Context``            -- one HIP stream and the device tables (qlp_ctx)
* ``CrossCorrelationJob`` -- Job<List<Tile>, XCorrelationResults>
  (jobs/Job.java:14-67; CrossCorrelationRealFFTParStdJob / SoftRealFFTJob)
* ``find_maximum`` / ``interpolate`` -- IMaximumFinder / ICrossCorrelationInterpolator

Device buffers are torch tensors on ``cuda``; host buffers are numpy arrays.
Every call runs the HIP kernels: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from enum import Enum

import numpy as np

from . import _lib
from ._lib import (QLP_NORM_A_MAX, QLP_NORM_PER_TILE, QLP_PEAK_FROM_CENTER, QLP_PEAK_MULTI,
                   QLP_PEAK_SIMPLE,
                   QLP_SUBPIX_GAUSS1D, QLP_SUBPIX_GUO, QLP_SUBPIX_POLY, QLP_XCORR_EXACT,
                   QLP_XCORR_MFMA_BF16, QLP_XCORR_MFMA_BF16_SINGLE, QLP_XCORR_MFMA_F16,
                   QLP_XCORR_MFMA_F16_SINGLE, QlpError, build, check)

__all__ = [
    "Batch", "Context", "CrossCorrelationJob", "JobResultEnum", "PivPlan", "QlpError", "Server",
    "VelocityExporter", "composite_last_level_start_step", "read_image", "read_image16",
    "piv_config", "plan_estimate",
    "VectorValidatorJob", "build",
    "QLP_NORM_PER_TILE", "QLP_NORM_A_MAX", "QLP_PEAK_FROM_CENTER", "QLP_PEAK_MULTI",
    "QLP_PEAK_SIMPLE",
    "QLP_SUBPIX_GAUSS1D", "QLP_SUBPIX_GUO", "QLP_SUBPIX_POLY", "QLP_XCORR_EXACT",
    "QLP_XCORR_MFMA_F16", "QLP_XCORR_MFMA_BF16", "QLP_XCORR_MFMA_F16_SINGLE",
    "QLP_XCORR_MFMA_BF16_SINGLE",
]


def _ptr(x):
    """Raw pointer of a torch tensor (device) or numpy array (host); None -> NULL."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        assert x.flags["C_CONTIGUOUS"], "host buffers must be C-contiguous"
        return x.ctypes.data
    assert x.is_contiguous(), "device buffers must be contiguous"
    return x.data_ptr()


class Context:
    """qlp_ctx: a HIP stream on one device plus the twiddle tables."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        check(self.lib.qlp_ctx_create(int(device), int(hbm_budget), ctypes.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            self.lib.qlp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def stream_handle(self) -> int:
        return self.lib.qlp_ctx_stream(self.h)

    def torch_stream(self):
        """The context's stream as a torch stream (to order torch work with ours)."""
        import torch

        return torch.cuda.ExternalStream(self.stream_handle, device=self.device)

    def synchronize(self):
        check(self.lib.qlp_ctx_synchronize(self.h))

    # ---- correlation -----------------------------------------------------
    def xcorr(self, tilesA, tilesB, maxA, maxB=None, mode=QLP_NORM_PER_TILE, out=None, form=0):
        """Correlate n IA pairs; numpy inputs use the host entry point, torch
        CUDA tensors the device one (asynchronous on this context's stream).
        form: QLP_XCORR_EXACT (the reference's arithmetic) or one of the MFMA
        forms QLP_XCORR_MFMA_* (device tensors only)."""
        n, N = int(tilesA.shape[0]), int(tilesA.shape[-1])
        D = 2 * N - 1
        if form != QLP_XCORR_EXACT:
            if isinstance(tilesA, np.ndarray):
                raise TypeError("the MFMA forms take device tensors")
            import torch

            if out is None:
                out = torch.empty((n, D, D), dtype=torch.float32, device=tilesA.device)
            check(self.lib.qlp_xcorr_form_dev(self.h, int(form), n, N, _ptr(tilesA), _ptr(tilesB),
                                              _ptr(maxA), _ptr(maxB), int(mode), _ptr(out)))
            return out
        if isinstance(tilesA, np.ndarray):
            tilesA = np.ascontiguousarray(tilesA, np.uint8)
            tilesB = np.ascontiguousarray(tilesB, np.uint8)
            maxA = np.ascontiguousarray(maxA, np.float32)
            maxB = None if maxB is None else np.ascontiguousarray(maxB, np.float32)
            if out is None:
                out = np.empty((n, D, D), np.float32)
            check(self.lib.qlp_xcorr_batch(self.h, n, N, _ptr(tilesA), _ptr(tilesB), _ptr(maxA),
                                           _ptr(maxB), int(mode), _ptr(out)))
            return out
        import torch

        if out is None:
            out = torch.empty((n, D, D), dtype=torch.float32, device=tilesA.device)
        check(self.lib.qlp_xcorr_batch_dev(self.h, n, N, _ptr(tilesA), _ptr(tilesB), _ptr(maxA),
                                           _ptr(maxB), int(mode), _ptr(out)))
        return out

    def xcorr_noalloc(self, tilesA, tilesB, maxA, maxB=None, mode=QLP_NORM_PER_TILE, out=None):
        """xcorr on device tensors through qlp_xcorr_batch_noalloc_dev: the
        library allocates nothing (out must be given or is allocated here by
        torch, outside the library)."""
        import torch

        n, N = int(tilesA.shape[0]), int(tilesA.shape[-1])
        D = 2 * N - 1
        if out is None:
            out = torch.empty((n, D, D), dtype=torch.float32, device=tilesA.device)
        check(self.lib.qlp_xcorr_batch_noalloc_dev(self.h, n, N, _ptr(tilesA), _ptr(tilesB),
                                                   _ptr(maxA), _ptr(maxB), int(mode), _ptr(out)))
        return out

    def xcorr_f32(self, tilesA, tilesB, maxA, maxB=None, mode=QLP_NORM_PER_TILE, out=None):
        """Correlate n pairs of float tiles (MatrixFloat, e.g. cut from filtered
        frames) in the exact form; numpy inputs use the host entry point, torch
        CUDA tensors the device one."""
        n, N = int(tilesA.shape[0]), int(tilesA.shape[-1])
        D = 2 * N - 1
        if isinstance(tilesA, np.ndarray):
            tilesA = np.ascontiguousarray(tilesA, np.float32)
            tilesB = np.ascontiguousarray(tilesB, np.float32)
            maxA = np.ascontiguousarray(maxA, np.float32)
            maxB = None if maxB is None else np.ascontiguousarray(maxB, np.float32)
            if out is None:
                out = np.empty((n, D, D), np.float32)
            check(self.lib.qlp_xcorr_f32(self.h, n, N, _ptr(tilesA), _ptr(tilesB), _ptr(maxA),
                                         _ptr(maxB), int(mode), _ptr(out)))
            return out
        import torch

        if out is None:
            out = torch.empty((n, D, D), dtype=torch.float32, device=tilesA.device)
        check(self.lib.qlp_xcorr_f32_dev(self.h, n, N, _ptr(tilesA), _ptr(tilesB), _ptr(maxA),
                                         _ptr(maxB), int(mode), _ptr(out)))
        return out

    # ---- peaks -------------------------------------------------------------
    def find_peaks(self, planes, kind=QLP_PEAK_FROM_CENTER, n_peaks=1, kernel_width=3,
                   out=None, min_floor=None):
        n, D = int(planes.shape[0]), int(planes.shape[-1])
        if isinstance(planes, np.ndarray):
            planes = np.ascontiguousarray(planes, np.float32)
            peaks = np.empty((n, n_peaks, 3), np.float32)
            mf = np.empty(n, np.float32)
            check(self.lib.qlp_find_peaks(self.h, int(kind), n, D, _ptr(planes), int(n_peaks),
                                          int(kernel_width), _ptr(peaks), _ptr(mf)))
            return peaks, mf
        import torch

        peaks = out if out is not None else torch.empty((n, n_peaks, 3), dtype=torch.float32,
                                                        device=planes.device)
        mf = min_floor if min_floor is not None else torch.empty(n, dtype=torch.float32,
                                                                 device=planes.device)
        check(self.lib.qlp_find_peaks_dev(self.h, int(kind), n, D, _ptr(planes), int(n_peaks),
                                          int(kernel_width), _ptr(peaks), _ptr(mf)))
        return peaks, mf

    # ---- sub-pixel -----------------------------------------------------------
    def subpixel(self, planes, min_floor, peaks, kind=QLP_SUBPIX_GAUSS1D, pixels=3,
                 iterations=20, out=None):
        n, D = int(planes.shape[0]), int(planes.shape[-1])
        P = int(peaks.shape[1])
        if isinstance(planes, np.ndarray):
            res = np.empty((n, P, 4), np.float32)
            check(self.lib.qlp_subpixel(self.h, int(kind), int(pixels), int(iterations), n, D, P,
                                        _ptr(np.ascontiguousarray(planes, np.float32)),
                                        _ptr(np.ascontiguousarray(min_floor, np.float32)),
                                        _ptr(np.ascontiguousarray(peaks, np.float32)), _ptr(res)))
            return res
        import torch

        res = out if out is not None else torch.empty((n, P, 4), dtype=torch.float32,
                                                      device=planes.device)
        check(self.lib.qlp_subpixel_dev(self.h, int(kind), int(pixels), int(iterations), n, D, P,
                                        _ptr(planes), _ptr(min_floor), _ptr(peaks), _ptr(res)))
        return res

    # ---- tiles ---------------------------------------------------------------
    def clip_tiles(self, img, top, left, N):
        import torch

        n = int(top.shape[0])
        H, W = int(img.shape[0]), int(img.shape[1])
        tiles = torch.empty((n, N, N), dtype=torch.uint8, device=img.device)
        mx = torch.empty(n, dtype=torch.float32, device=img.device)
        check(self.lib.qlp_clip_tiles_dev(self.h, _ptr(img), H, W, n, N, _ptr(top), _ptr(left),
                                          _ptr(tiles), _ptr(mx)))
        return tiles, mx

    def xcorr_images(self, imgA, imgB, top, left, N, du=None, dv=None, warp_mode=1,
                     mode=QLP_NORM_PER_TILE, planes=True, nbr_half=0):
        """Fused pass straight from device images: returns (planes or None, peaks
        n x 3, min_floor n, nbr or None)."""
        import torch

        n = int(top.shape[0])
        H, W = int(imgA.shape[-2]), int(imgA.shape[-1])
        D = 2 * N - 1
        dev = imgA.device
        pl = torch.empty((n, D, D), dtype=torch.float32, device=dev) if planes else None
        pk = torch.empty((n, 3), dtype=torch.float32, device=dev)
        mf = torch.empty(n, dtype=torch.float32, device=dev)
        nb = (torch.empty((n, 2, 2 * nbr_half + 1), dtype=torch.float32, device=dev)
              if nbr_half else None)
        check(self.lib.qlp_xcorr_images_dev(self.h, n, N, int(mode), _ptr(imgA), _ptr(imgB), H, W,
                                            _ptr(top), _ptr(left), _ptr(du), _ptr(dv),
                                            int(warp_mode), _ptr(pl), _ptr(pk), _ptr(mf), _ptr(nb),
                                            int(nbr_half)))
        return pl, pk, mf, nb

    def warp_micro_tiles(self, img, top, left, u, v, N):
        import torch

        n = int(top.shape[0])
        H, W = int(img.shape[0]), int(img.shape[1])
        tiles = torch.empty((n, N, N), dtype=torch.uint8, device=img.device)
        mx = torch.empty(n, dtype=torch.float32, device=img.device)
        check(self.lib.qlp_warp_micro_tiles_dev(self.h, _ptr(img), H, W, n, N, _ptr(top),
                                                _ptr(left), _ptr(u), _ptr(v), _ptr(tiles),
                                                _ptr(mx)))
        return tiles, mx

    def clip_tiles_f32(self, img, top, left, N, u=None, v=None):
        """Float tiles of a device float frame at (top, left), or shifted by the
        rounded (u, v): (n x N x N tiles, maxima)."""
        import torch

        n = int(top.shape[0])
        H, W = int(img.shape[0]), int(img.shape[1])
        tiles = torch.empty((n, N, N), dtype=torch.float32, device=img.device)
        mx = torch.empty(n, dtype=torch.float32, device=img.device)
        check(self.lib.qlp_clip_tiles_f32_dev(self.h, _ptr(img), H, W, n, N, _ptr(top), _ptr(left),
                                              _ptr(u), _ptr(v), _ptr(tiles), _ptr(mx)))
        return tiles, mx

    def warp_micro_tiles_f32(self, img, top, left, u, v, N, side=_lib.QLP_SIDE_B_SECOND):
        """Micro warp of one side of n tiles of a device float frame:
        (n x N x N float tiles, maxima 65535)."""
        import torch

        n = int(top.shape[0])
        H, W = int(img.shape[0]), int(img.shape[1])
        tiles = torch.empty((n, N, N), dtype=torch.float32, device=img.device)
        mx = torch.empty(n, dtype=torch.float32, device=img.device)
        check(self.lib.qlp_warp_micro_tiles_f32_dev(self.h, _ptr(img), H, W, n, N, _ptr(top),
                                                    _ptr(left), _ptr(u), _ptr(v), int(side),
                                                    _ptr(tiles), _ptr(mx)))
        return tiles, mx

    def warp_mini_tiles(self, imgs, top, left, u, v, N):
        """Mini-warped B tiles of every tile of an nI x nJ grid for 'pairs' frames
        (imgs: pairs x H x W device uint8; top, left: device nI x nJ int32; u, v:
        device pairs x nI x nJ) -> (pairs x nI*nJ x N x N tiles, maxima)."""
        import torch

        imgs = imgs if imgs.dim() == 3 else imgs[None]
        pairs, H, W = (int(x) for x in imgs.shape)
        nI, nJ = int(top.shape[0]), int(top.shape[1])
        tiles = torch.empty((pairs, nI * nJ, N, N), dtype=torch.uint8, device=imgs.device)
        mx = torch.empty((pairs, nI * nJ), dtype=torch.float32, device=imgs.device)
        check(self.lib.qlp_warp_mini_tiles_dev(self.h, _ptr(imgs), H, W, pairs, nI, nJ, N,
                                               _ptr(top), _ptr(left), _ptr(u), _ptr(v),
                                               _ptr(tiles), _ptr(mx)))
        return tiles, mx

    def warp_image(self, imgs, top, left, u, v, N):
        """Full-image warp of 'pairs' second frames by the spline field of an
        nI x nJ grid (top, left: host numpy nI x nJ; u, v: device pairs x nI x nJ)."""
        import torch

        imgs = imgs if imgs.dim() == 3 else imgs[None]
        pairs, H, W = (int(x) for x in imgs.shape)
        top = np.ascontiguousarray(top, np.int32)
        left = np.ascontiguousarray(left, np.int32)
        nI, nJ = top.shape
        out = torch.empty_like(imgs)
        check(self.lib.qlp_warp_image_dev(self.h, _ptr(imgs), H, W, pairs, nI, nJ, N, _ptr(top),
                                          _ptr(left), _ptr(u), _ptr(v), _ptr(out)))
        return out

    def warp_clip_tiles(self, imgsA, imgsB, top, left, u, v, N, warp_mode=_lib.QLP_WARP_MICRO,
                        order=_lib.QLP_WARP_ORDER_SECOND):
        """warpAndClipImage of the Micro / Mini strategies at a level > 0: the A
        and B tiles of every IA of an nI x nJ grid for 'pairs' frame pairs in a
        warp order (imgs: pairs x H x W device uint8; top, left: device nI x nJ
        int32; u, v: device pairs x nI x nJ) -> (tilesA, maxA, tilesB, maxB)."""
        import torch

        imgsA = imgsA if imgsA.dim() == 3 else imgsA[None]
        imgsB = imgsB if imgsB.dim() == 3 else imgsB[None]
        pairs, H, W = (int(x) for x in imgsA.shape)
        nI, nJ = int(top.shape[0]), int(top.shape[1])
        dev = imgsA.device
        tA = torch.empty((pairs, nI * nJ, N, N), dtype=torch.uint8, device=dev)
        tB = torch.empty_like(tA)
        mA = torch.empty((pairs, nI * nJ), dtype=torch.float32, device=dev)
        mB = torch.empty_like(mA)
        check(self.lib.qlp_warp_clip_tiles_dev(self.h, int(warp_mode), int(order), _ptr(imgsA),
                                               _ptr(imgsB), H, W, pairs, nI, nJ, N, _ptr(top),
                                               _ptr(left), _ptr(u), _ptr(v), _ptr(tA), _ptr(mA),
                                               _ptr(tB), _ptr(mB)))
        return tA, mA, tB, mB

    def warp_images(self, imgsA, imgsB, top, left, u, v, N, order=_lib.QLP_WARP_ORDER_SECOND):
        """The frames the Full-image warp produces in a warp order: (warped A or
        None, warped B or None); top, left host numpy nI x nJ."""
        import torch

        imgsA = imgsA if imgsA.dim() == 3 else imgsA[None]
        imgsB = imgsB if imgsB.dim() == 3 else imgsB[None]
        pairs, H, W = (int(x) for x in imgsA.shape)
        top = np.ascontiguousarray(top, np.int32)
        left = np.ascontiguousarray(left, np.int32)
        nI, nJ = top.shape
        outA = None if order == _lib.QLP_WARP_ORDER_SECOND else torch.empty_like(imgsA)
        outB = None if order == _lib.QLP_WARP_ORDER_FIRST else torch.empty_like(imgsB)
        check(self.lib.qlp_warp_images_dev(self.h, int(order), _ptr(imgsA), _ptr(imgsB), H, W,
                                           pairs, nI, nJ, N, _ptr(top), _ptr(left), _ptr(u),
                                           _ptr(v), _ptr(outA), _ptr(outB)))
        return outA, outB

    # ---- image pre-filter ----------------------------------------------------
    def gaussian_filter(self, imgs, sigma, kernel_px):
        """GaussianFilter2D.applyFilter on device images (n x H x W or H x W,
        uint8 or float32 CUDA tensors); returns float32 images."""
        import torch

        shape = imgs.shape
        H, W = int(shape[-2]), int(shape[-1])
        n = int(imgs.numel()) // (H * W)
        out = torch.empty(shape, dtype=torch.float32, device=imgs.device)
        tmp = torch.empty(shape, dtype=torch.float32, device=imgs.device)
        check(self.lib.qlp_gaussian_filter_dev(self.h, n, H, W, _ptr(imgs),
                                               int(imgs.dtype == torch.float32), float(sigma),
                                               int(kernel_px), _ptr(tmp), _ptr(out)))
        return out

    # ---- vector validation ---------------------------------------------------
    def validate_vectors(self, top, left, u, v, stages, peak_uv=None, peak_count=None,
                         max_iterations=1, until_no_more=False):
        """VectorValidatorJob.compute over 'pairs' grids.

        top, left: nI x nJ tile origins (host).  u, v: pairs x nI x nJ (or
        nI x nJ) displacements; peak_uv: ... x P x 2 peak displacement
        increments; peak_count: ... per-tile peak totals (None: all P).  numpy
        inputs use the host entry point and return new arrays; torch CUDA
        tensors are updated in place on the context's stream.  Returns
        (u, v, invalid, stats)."""
        top = np.ascontiguousarray(top, np.int32)
        left = np.ascontiguousarray(left, np.int32)
        nI, nJ = top.shape
        cfg = _lib.valid_config(stages, max_iterations, until_no_more)
        pairs = int(np.prod(u.shape)) // (nI * nJ)
        P = 1 if peak_uv is None else int(peak_uv.shape[-2])
        if isinstance(u, np.ndarray):
            u = np.array(u, np.float32, copy=True, order="C")
            v = np.array(v, np.float32, copy=True, order="C")
            pk = None if peak_uv is None else np.ascontiguousarray(peak_uv, np.float32)
            pc = None if peak_count is None else np.ascontiguousarray(peak_count, np.int32)
            invalid = np.empty(u.shape, np.uint8)
            stats = np.empty((pairs, 4), np.int32)
            check(self.lib.qlp_validate_vectors(self.h, pairs, nI, nJ, _ptr(top), _ptr(left),
                                                _ptr(u), _ptr(v), P, _ptr(pk), _ptr(pc),
                                                ctypes.byref(cfg), _ptr(invalid), _ptr(stats)))
            return u, v, invalid.astype(bool), stats
        import torch

        invalid = torch.empty(u.shape, dtype=torch.uint8, device=u.device)
        stats = torch.empty((pairs, 4), dtype=torch.int32, device=u.device)
        check(self.lib.qlp_validate_vectors_dev(self.h, pairs, nI, nJ, _ptr(top), _ptr(left),
                                                _ptr(u), _ptr(v), P, _ptr(peak_uv),
                                                _ptr(peak_count), ctypes.byref(cfg),
                                                _ptr(invalid), _ptr(stats)))
        return u, v, invalid, stats


def inherit_velocity(ctx: Context, mode, parent, child, pu, pv, round_it=False):
    """IInterAreaVelocityInheritanceStrategy.reuseIterationStepTilesParameters on
    device tensors.  parent / child: (top, left, tile_size) with nI x nJ origin
    arrays; pu, pv: pairs x pnI x pnJ (or pnI x pnJ) CUDA tensors.  Returns the
    child displacements (u, v) as CUDA tensors."""
    import torch

    ptop, pleft, psz = parent
    ctop, cleft, csz = child
    ptop = np.ascontiguousarray(ptop, np.int32)
    pleft = np.ascontiguousarray(pleft, np.int32)
    ctop = np.ascontiguousarray(ctop, np.int32)
    cleft = np.ascontiguousarray(cleft, np.int32)
    pnI, pnJ = ptop.shape
    cnI, cnJ = ctop.shape
    pairs = int(pu.numel()) // (pnI * pnJ)
    u = torch.empty((pairs, cnI, cnJ), dtype=torch.float32, device=pu.device)
    v = torch.empty_like(u)
    check(ctx.lib.qlp_inherit_velocity_dev(ctx.h, int(mode), pairs, pnI, pnJ, int(psz), int(psz),
                                           _ptr(ptop), _ptr(pleft), cnI, cnJ, int(csz), int(csz),
                                           _ptr(ctop), _ptr(cleft), int(bool(round_it)),
                                           _ptr(pu), _ptr(pv), _ptr(u), _ptr(v)))
    return u, v


class JobResultEnum(Enum):
    """The keys of the reference's job protocol used on this path
    (jobs/JobResultEnum.java:14-22)."""
    JOB_RESULT_CLIPPED_TILES_A = "clipped_tiles_a"
    JOB_RESULT_CLIPPED_TILES_B = "clipped_tiles_b"
    JOB_RESULT_CROSS_MATRICES = "cross_matrices"
    JOB_RESULT_VALIDATOR_DATA = "validator_data"
    JOB_RESULT_VALIDATED_VECTORS = "validated_vectors"


class VectorValidatorJob:
    """Job<VectorValidatorJobParameters, IterationStepTiles> over libqlpiv
    (jobs/validator/VectorValidatorJob.java:43-410).

    Input (JOB_RESULT_VALIDATOR_DATA): a dict with the tile grid ``top``,
    ``left`` (nI x nJ), displacements ``u``, ``v`` (nI x nJ), the peak
    displacement increments ``peak_uv`` (nI x nJ x P x 2, peak 0 the main
    peak) and optionally ``peak_count`` (nI x nJ).  Output
    (JOB_RESULT_VALIDATED_VECTORS): a dict with the validated ``u``, ``v``,
    the ``invalid`` flags and the job ``stats``.  Configuration: the stage list
    of the combined validator ((validator, threshold, epsilon0, replacement,
    peak_index) tuples), maxCorrectionIterations and
    iterateUntilNoMoreCorrections.  analyze() raises the reference's
    JobAnalyzeException cases as RuntimeError (:52-74)."""

    def __init__(self, ctx: Context, stages, max_iterations=0, until_no_more=False):
        self.ctx = ctx
        self.stages = list(stages) if stages is not None else None
        self.max_iterations = max_iterations
        self.until_no_more = until_no_more
        self.inputs = {}
        self.results = {}

    def set_input_parameters(self, key: JobResultEnum, params):
        self.inputs[key] = params

    def analyze(self):
        if self.stages is None:
            raise RuntimeError("Cannot have a null vector validator or null vector replacement "
                               "strategy defined")
        for st in self.stages:
            if st is None or st[0] is None:
                raise RuntimeError("Validator strategy cannot be null")
        if JobResultEnum.JOB_RESULT_VALIDATOR_DATA not in self.inputs:
            raise RuntimeError("Validator input parameters are not set")

    def compute(self):
        p = self.inputs[JobResultEnum.JOB_RESULT_VALIDATOR_DATA]
        if not self.stages:
            self.results[JobResultEnum.JOB_RESULT_VALIDATED_VECTORS] = dict(p)
            return
        u, v, inv, st = self.ctx.validate_vectors(p["top"], p["left"], p["u"], p["v"],
                                                  self.stages, p.get("peak_uv"),
                                                  p.get("peak_count"), self.max_iterations,
                                                  self.until_no_more)
        out = dict(p)
        out.update(u=u, v=v, invalid=inv, stats=st)
        self.results[JobResultEnum.JOB_RESULT_VALIDATED_VECTORS] = out

    def get_job_result(self, key: JobResultEnum):
        return self.results.get(key)

    def dispose(self):
        self.inputs.clear()
        self.results.clear()


class CrossCorrelationJob:
    """Job<List<Tile>, XCorrelationResults> over libqlpiv (jobs/Job.java:14-67).

    Tiles are (pixels uint8 N x N, max value) pairs, i.e. Tile.getMatrix() data
    and Matrix.getMaxValue().  ``mode`` selects the reference job whose
    arithmetic is reproduced: QLP_NORM_PER_TILE for
    CrossCorrelationRealFFTParStdJob, QLP_NORM_A_MAX for
    CrossCorrelationSoftRealFFTJob.  Errors follow the reference: mismatched
    or non-square tile lists raise RuntimeError in analyze()
    (CrossCorrelationFFTTemplate.analyzeTilesHelper :594-642)."""

    def __init__(self, ctx: Context, mode=QLP_NORM_PER_TILE):
        self.ctx = ctx
        self.mode = mode
        self.inputs = {}
        self.results = {}
        self._analyzed = None

    def set_input_parameters(self, key: JobResultEnum, tiles):
        self.inputs[key] = tiles
        self._analyzed = None

    def analyze(self):
        ta = self.inputs.get(JobResultEnum.JOB_RESULT_CLIPPED_TILES_A)
        tb = self.inputs.get(JobResultEnum.JOB_RESULT_CLIPPED_TILES_B)
        if ta is None or tb is None:
            raise RuntimeError("Tiles A and tiles B must be set before analyze()")
        if len(ta) != len(tb):
            raise RuntimeError("Number of tiles in A differs from the number of tiles in B")
        if len(ta) == 0:
            self._analyzed = (0, 0, None, None, None, None)
            return
        N = ta[0][0].shape[0]
        for pix, _ in list(ta) + list(tb):
            if pix.shape != (N, N):
                raise RuntimeError("Tiles must be square and all of the same size")
        A = np.ascontiguousarray(np.stack([p for p, _ in ta]).astype(np.uint8))
        B = np.ascontiguousarray(np.stack([p for p, _ in tb]).astype(np.uint8))
        mA = np.array([m for _, m in ta], np.float32)
        mB = np.array([m for _, m in tb], np.float32)
        self._analyzed = (len(ta), N, A, B, mA, mB)

    def compute(self):
        if self._analyzed is None:
            self.analyze()
        n, N, A, B, mA, mB = self._analyzed
        planes = [] if n == 0 else list(self.ctx.xcorr(A, B, mA, mB, self.mode))
        self.results[JobResultEnum.JOB_RESULT_CROSS_MATRICES] = planes

    def get_job_result(self, key: JobResultEnum):
        return self.results.get(key)

    def dispose(self):
        self.inputs.clear()
        self.results.clear()


def composite_last_level_start_step(n_levels, interpolator_start_step, apply_first_at_last):
    """First level a plan interpolates for CompositeLastLevelInterpolator
    (interpolators/CompositeLastLevelInterpolator.java:37-54) with a
    cross-correlation first-levels interpolator and no last-level one.

    The reference computes isLastStep = currentStep != maxAdaptiveSteps - 1
    (:38), true on every level except the last, so the first-levels
    interpolator runs on the last level only, or on every level when
    applyMainInterpolatorAtLastStep is set; the job applies the composite from
    interpolatorStartStep on (jobs/LocalPIVOpenCLGpuJob.java:373).  The
    last-level interpolators it is meant for are the dense optical-flow ones
    (LucasKanade, LiuShenWithLucasKanade), which this library does not provide."""
    first = 0 if apply_first_at_last else n_levels - 1
    return max(int(interpolator_start_step), first)


def piv_config(**config):
    """qlp_piv_config from keyword arguments (the keys PivPlan documents)."""
    cfg = _lib.PivConfig()
    cfg.ia_start = config.get("ia_start", 64)
    cfg.ia_end = config.get("ia_end", 16)
    cfg.overlap_factor = config.get("overlap_factor", 0.5)
    m = config.get("margins", (0, 0, 0, 0))
    cfg.margin_top, cfg.margin_bottom, cfg.margin_left, cfg.margin_right = m
    cfg.norm_mode = config.get("norm_mode", QLP_NORM_PER_TILE)
    cfg.warp_mode = config.get("warp_mode", _lib.QLP_WARP_MICRO)
    cfg.subpix_kind = config.get("subpix_kind", QLP_SUBPIX_GUO)
    cfg.subpix_pixels = config.get("subpix_pixels", 3)
    cfg.subpix_iterations = config.get("subpix_iterations", 20)
    cfg.interpolator_start_step = config.get("interpolator_start_step", 0)
    cfg.keep_planes = int(config.get("keep_planes", False))
    cfg.peak_kind = config.get("peak_kind", QLP_PEAK_FROM_CENTER)
    cfg.n_peaks = config.get("n_peaks", 1)
    cfg.peak_kernel_width = config.get("peak_kernel_width", 3)
    cfg.valid = _lib.valid_config(config.get("validation") or [],
                                  config.get("valid_max_iterations", 0),
                                  config.get("valid_until_no_more", False))
    cfg.inherit_mode = config.get("inherit_mode", _lib.QLP_INHERIT_DIRECT)
    cfg.xcorr_form = config.get("xcorr_form", QLP_XCORR_EXACT)
    cfg.warp_order = config.get("warp_order", _lib.QLP_WARP_ORDER_SECOND)
    cfg.frame_type = config.get("frame_type", _lib.QLP_FRAME_U8)
    cfg.frame_bits = config.get("frame_bits", 16 if cfg.frame_type == _lib.QLP_FRAME_U16 else 0)
    return cfg


def plan_estimate(H, W, pairs=1, **config):
    """Device bytes a PivPlan of this configuration allocates (qlp_piv_plan_estimate)."""
    lib = _lib.load()
    cfg = piv_config(**config)
    b = ctypes.c_size_t()
    check(lib.qlp_piv_plan_estimate(int(H), int(W), int(pairs), ctypes.byref(cfg), ctypes.byref(b)))
    return b.value


def read_image(path):
    """An 8-bit grayscale TIFF, PNG or JPEG frame as an H x W uint8 array
    (qlp_image_read: ImageReaderJob for 8-bit images)."""
    lib = _lib.load()
    H, W = ctypes.c_int(), ctypes.c_int()
    p = str(path).encode()
    check(lib.qlp_image_read(p, None, 0, ctypes.byref(H), ctypes.byref(W)))
    out = np.empty((H.value, W.value), np.uint8)
    check(lib.qlp_image_read(p, _ptr(out), out.nbytes, ctypes.byref(H), ctypes.byref(W)))
    return out


def read_image16(path):
    """A 16-bit grayscale TIFF or PNG frame as an H x W uint16 array
    (qlp_image_read16: ImageReaderJob for 10/12/16-bit images)."""
    lib = _lib.load()
    H, W = ctypes.c_int(), ctypes.c_int()
    p = str(path).encode()
    check(lib.qlp_image_read16(p, None, 0, ctypes.byref(H), ctypes.byref(W)))
    out = np.empty((H.value, W.value), np.uint16)
    check(lib.qlp_image_read16(p, _ptr(out), out.size, ctypes.byref(H), ctypes.byref(W)))
    return out


class VelocityExporter:
    """The MATLAB Level 5 velocity file (qlp_export_*): openFile, per-frame
    exportDataToFile and closeFile of exporter/StructMultiFrameFloatVelocityExporter.

    image_size = (height, width); ia_start / ia_end = (I, J) IA sizes of the
    first and last level."""

    def __init__(self, filename, image_size, ia_start, ia_end, overlap_factor, total_maps=1,
                 swap_uv=False, mark_invalid_nan=False, multi_volume=False, next_filename=None,
                 date_text=None):
        self.lib = _lib.load()
        p = _lib.ExportParams()
        p.image_height, p.image_width = (int(x) for x in image_size)
        p.ia_start_i, p.ia_start_j = (int(x) for x in ia_start)
        p.ia_end_i, p.ia_end_j = (int(x) for x in ia_end)
        p.overlap_factor = float(overlap_factor)
        p.swap_uv = int(bool(swap_uv))
        p.mark_invalid_nan = int(bool(mark_invalid_nan))
        p.total_maps = int(total_maps)
        p.multi_volume = int(bool(multi_volume))
        p.next_filename = next_filename.encode() if next_filename else None
        h = ctypes.c_void_p()
        check(self.lib.qlp_export_open(str(filename).encode(), ctypes.byref(p),
                                       date_text.encode() if date_text else None, ctypes.byref(h)))
        self.h = h

    def export(self, frame, u, v, ia, adaptive_step, margins=(0, 0, 0, 0), invalid=None):
        """One level's map of frame 'frame': u, v (nI x nJ host arrays), ia =
        (height, width), margins = (top, left, bottom, right), invalid flags."""
        u = np.ascontiguousarray(u, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        nI, nJ = u.shape
        m = (ctypes.c_int * 4)(*[int(x) for x in margins])
        inv = None if invalid is None else np.ascontiguousarray(invalid, np.uint8)
        check(self.lib.qlp_export_frame(self.h, int(frame), int(ia[0]), int(ia[1]),
                                        int(adaptive_step), m, nI, nJ, _ptr(u), _ptr(v),
                                        _ptr(inv)))

    def export_results(self, frames, u, v, ia, n_maps, margins=(0, 0, 0, 0), start_offset=0):
        """One velocity map of whole results (exportDataToFile(PIVResults)):
        u, v (len(frames) x nI x nJ host arrays) go to the map indices
        'frames'; ia = (height, width); n_maps = PIVMap.getNumberOfMaps()."""
        u = np.ascontiguousarray(u, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        fr = np.ascontiguousarray(frames, np.int32)
        n, nI, nJ = u.shape
        m = (ctypes.c_int * 4)(*[int(x) for x in margins])
        check(self.lib.qlp_export_results(self.h, int(ia[0]), int(ia[1]), m, nI, nJ, int(n_maps), n,
                                          _ptr(fr), _ptr(u), _ptr(v), int(start_offset)))

    def close(self):
        if self.h:
            h, self.h = self.h, None
            check(self.lib.qlp_export_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Batch:
    """Frame pairs of mixed frame and IA sizes run by the step engine
    (qlp_batch): the per-frame job (clip or micro warp, correlation, FromCenter
    peak, sub-pixel fit, Direct inheritance between levels) of every pair, one
    graph launch per adaptive level, with the kernel shapes and the schedule
    picked by the library's autotuner on the first run.

    pairs: list of (H, W, ia) or (H, W, ia, levels) (ia: last-level IA size).
    config: the PivPlan keys (overlap_factor, margins, norm_mode, warp_mode,
    subpix_*, interpolator_start_step, xcorr_form)."""

    def __init__(self, ctx, pairs, **config):
        self.ctx = ctx
        self.lib = ctx.lib
        self.pairs = [tuple(int(x) for x in p) + ((1,) if len(p) == 3 else ()) for p in pairs]
        arr = (_lib.BatchPair * len(self.pairs))(*[_lib.BatchPair(*p) for p in self.pairs])
        cfg = piv_config(**config)
        h = ctypes.c_void_p()
        check(self.lib.qlp_batch_create(ctx.h, len(self.pairs), arr, ctypes.byref(cfg),
                                        ctypes.byref(h)))
        self.h = h
        self.grids = []
        total = ctypes.c_size_t()
        for k in range(len(self.pairs)):
            nI, nJ = ctypes.c_int(), ctypes.c_int()
            check(self.lib.qlp_batch_grid(self.h, k, ctypes.byref(nI), ctypes.byref(nJ),
                                          ctypes.byref(total)))
            self.grids.append((nI.value, nJ.value))
        self.total_ias = total.value

    def fix(self, ia=0, shape=_lib.QLP_SHAPE_AUTO, schedule=_lib.QLP_SCHEDULE_TUNED):
        """Fix a class's kernel shape (ia 0: all classes) and/or the schedule."""
        check(self.lib.qlp_batch_fix(self.h, int(ia), int(shape), int(schedule)))

    def run(self, framesA, framesB):
        """framesA, framesB: sequences of CUDA uint8 tensors (H x W each), kept
        alive by the caller until the run has completed."""
        n = len(self.pairs)
        pa = (ctypes.c_void_p * n)(*[_ptr(t) for t in framesA])
        pb = (ctypes.c_void_p * n)(*[_ptr(t) for t in framesB])
        check(self.lib.qlp_batch_run_dev(self.h, pa, pb))

    def result(self, k):
        """Host copies (after synchronising) of pair k's u, v (nI x nJ) and
        sub-pixel records (nI x nJ x 4)."""
        u, v, s = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        self.ctx.synchronize()
        check(self.lib.qlp_batch_result_dev(self.h, int(k), ctypes.byref(u), ctypes.byref(v),
                                            ctypes.byref(s)))
        nI, nJ = self.grids[k]
        n = nI * nJ
        return (_fetch(u, n, np.float32).reshape(nI, nJ), _fetch(v, n, np.float32).reshape(nI, nJ),
                _fetch(s, 4 * n, np.float32).reshape(nI, nJ, 4))

    def report(self):
        r = _lib.BatchReport()
        check(self.lib.qlp_batch_report_get(self.h, ctypes.byref(r)))
        cls = []
        for c in r.cls[:r.n_classes]:
            cls.append({"ia": c.ia, "n_ias": c.n_ias, "shape": c.shape,
                        "candidates": {int(c.cand_shape[k]): round(float(c.cand_ms[k]), 4)
                                       for k in range(c.n_candidates)}})
        return {"classes": cls, "schedule": r.schedule, "serial_ms": r.serial_ms,
                "concurrent_ms": r.concurrent_ms}

    def close(self):
        if self.h:
            self.lib.qlp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving form (qlp_server): a request loop over frame pairs of mixed
    sizes with HBM-budget admission and continuous batching (every step runs
    the next level of every admitted request in one graph launch; other
    configurations run on plan slots that overlap on the device).
    Replaces the reference's worker-thread pool over frame pairs
    (InputFiles.java:172-224, OpenClGpuManager.java:124-142).

    submit() takes the frames (numpy uint8 host arrays, preferably pinned, or
    CUDA tensors) and the PivPlan configuration keys; the frames must stay alive
    until the request has finished (wait())."""

    def __init__(self, device=0, hbm_budget=64 << 30, slots=4, batch_requests=0):
        self.lib = _lib.load()
        cfg = _lib.ServerConfig(int(hbm_budget), int(slots), int(batch_requests))
        h = ctypes.c_void_p()
        check(self.lib.qlp_server_create(int(device), ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self._keep = {}

    def submit(self, imgA, imgB, cfg=None, **config):
        """Queue one frame pair; cfg: a prebuilt piv_config(...) (then no keys),
        which saves rebuilding it for every request."""
        H, W = int(imgA.shape[0]), int(imgA.shape[1])
        on_dev = not isinstance(imgA, np.ndarray)
        if not on_dev:
            imgA = np.ascontiguousarray(imgA, np.uint8)
            imgB = np.ascontiguousarray(imgB, np.uint8)
        req = _lib.Request(H, W, _ptr(imgA), _ptr(imgB), int(on_dev),
                           cfg if cfg is not None else piv_config(**config))
        t = ctypes.c_uint64()
        check(self.lib.qlp_server_submit(self.h, ctypes.byref(req), ctypes.byref(t)))
        self._keep[t.value] = (imgA, imgB)
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        done = ctypes.c_int()
        check(self.lib.qlp_server_wait(self.h, ctypes.c_uint64(ticket), int(timeout_ms),
                                       ctypes.byref(done)))
        if done.value:
            self._keep.pop(ticket, None)
        return bool(done.value)

    def result(self, ticket, release=True):
        """(u, v, ia, times) of a finished request: the last level's field
        (nI x nJ numpy copies), its IA size and the request's timings in ms
        (submit, admit, done) and IA count."""
        u, v = ctypes.c_void_p(), ctypes.c_void_p()
        nI, nJ, ia = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        tm = _lib.RequestTimes()
        check(self.lib.qlp_server_result(self.h, ctypes.c_uint64(ticket), ctypes.byref(u),
                                         ctypes.byref(v), ctypes.byref(nI), ctypes.byref(nJ),
                                         ctypes.byref(ia), ctypes.byref(tm)))
        n = nI.value * nJ.value
        fu = np.ctypeslib.as_array((ctypes.c_float * n).from_address(u.value)).reshape(nI.value, nJ.value).copy()
        fv = np.ctypeslib.as_array((ctypes.c_float * n).from_address(v.value)).reshape(nI.value, nJ.value).copy()
        times = {"submit_ms": tm.submit_ms, "admit_ms": tm.admit_ms, "done_ms": tm.done_ms,
                 "ias": int(tm.ias)}
        if release:
            check(self.lib.qlp_server_release(self.h, ctypes.c_uint64(ticket)))
        return fu, fv, ia.value, times

    def stats(self):
        st = _lib.ServerStats()
        check(self.lib.qlp_server_get_stats(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self.h:
            self.lib.qlp_server_destroy(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PivPlan:
    """Multi-pass PIV over device-resident frame pairs (qlp_piv_plan): the
    adaptive per-level loop of jobs/LocalPIVOpenCLGpuJob.call :188-409 with
    device-resident displacement state.

    config keys: ia_start, ia_end, overlap_factor (1 - overlap), margins
    (top, bottom, left, right), norm_mode, warp_mode (QLP_WARP_NONE/MICRO),
    subpix_kind (QLP_SUBPIX_* or -1), subpix_pixels, subpix_iterations,
    interpolator_start_step, keep_planes, peak_kind (QLP_PEAK_*), n_peaks,
    peak_kernel_width, validation (list of stage tuples, see
    VectorValidatorJob), valid_max_iterations, valid_until_no_more,
    inherit_mode (QLP_INHERIT_*), xcorr_form (QLP_XCORR_*: the arithmetic of the
    16 and 32 px levels)."""

    def __init__(self, ctx: Context, H: int, W: int, pairs: int = 1, **config):
        cfg = piv_config(**config)
        self.P = cfg.n_peaks if cfg.peak_kind == QLP_PEAK_MULTI else 1
        self.ctx, self.H, self.W, self.pairs, self.cfg = ctx, H, W, pairs, cfg
        h = ctypes.c_void_p()
        check(ctx.lib.qlp_piv_plan_create(ctx.h, H, W, pairs, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        nl = ctypes.c_int()
        check(ctx.lib.qlp_piv_plan_levels(h, ctypes.byref(nl)))
        self.levels = []
        for lv in range(nl.value):
            ia, nI, nJ = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            check(ctx.lib.qlp_piv_plan_level(h, lv, ctypes.byref(ia), ctypes.byref(nI),
                                             ctypes.byref(nJ), None, None))
            top = np.empty(nI.value * nJ.value, np.int32)
            left = np.empty_like(top)
            check(ctx.lib.qlp_piv_plan_level(h, lv, None, None, None, _ptr(top), _ptr(left)))
            self.levels.append(dict(ia=ia.value, nI=nI.value, nJ=nJ.value,
                                    top=top.reshape(nI.value, nJ.value),
                                    left=left.reshape(nI.value, nJ.value)))

    def close(self):
        if self.h:
            self.ctx.lib.qlp_piv_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, imgsA, imgsB):
        """Enqueue all levels (device tensors pairs x H x W: uint8, float32 for a
        plan made with frame_type QLP_FRAME_F32, 16-bit samples (torch.uint16 or
        int16 holding them) for QLP_FRAME_U16)."""
        if self.cfg.frame_type == _lib.QLP_FRAME_F32:
            check(self.ctx.lib.qlp_piv_run_f32_dev(self.h, _ptr(imgsA), _ptr(imgsB)))
        elif self.cfg.frame_type == _lib.QLP_FRAME_U16:
            check(self.ctx.lib.qlp_piv_run_u16_dev(self.h, _ptr(imgsA), _ptr(imgsB)))
        else:
            check(self.ctx.lib.qlp_piv_run_dev(self.h, _ptr(imgsA), _ptr(imgsB)))

    def set_timing(self, enable=True):
        check(self.ctx.lib.qlp_piv_plan_set_timing(self.h, int(enable)))

    def timing(self):
        """[(pass_ms, fit_ms, valid_ms)] per level of the last run (waits for it)."""
        out = []
        for lv in range(len(self.levels)):
            a, b, c = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
            check(self.ctx.lib.qlp_piv_plan_timing(self.h, lv, ctypes.byref(a), ctypes.byref(b),
                                                   ctypes.byref(c)))
            out.append((a.value, b.value, c.value))
        return out

    def result(self, level=-1):
        """Host copies (after synchronising) of a level's u, v (pairs x nI x nJ)
        and peak records (pairs x nI x nJ x 4, or pairs x nI x nJ x P x 4 with
        several peaks)."""
        lv = level % len(self.levels)
        L = self.levels[lv]
        u, v, pk, pl = (ctypes.c_void_p() for _ in range(4))
        self.ctx.synchronize()
        check(self.ctx.lib.qlp_piv_result_dev(self.h, lv, ctypes.byref(u), ctypes.byref(v),
                                              ctypes.byref(pk), ctypes.byref(pl)))
        n = L["nI"] * L["nJ"] * self.pairs
        shp = (self.pairs, L["nI"], L["nJ"])
        pshape = shp + ((4,) if self.P == 1 else (self.P, 4))
        return (_fetch(u, n, np.float32).reshape(shp), _fetch(v, n, np.float32).reshape(shp),
                _fetch(pk, 4 * n * self.P, np.float32).reshape(pshape))

    def result_async(self, u_host, v_host, level=-1):
        """Enqueue copies of a level's u, v (pairs x nI x nJ float32) into the
        given host arrays on the context's stream, after the work enqueued so
        far; they are complete once the stream has passed them (an event
        recorded after this call).  Pinned host memory makes them asynchronous
        to the host."""
        lv = level % len(self.levels)
        L = self.levels[lv]
        n = L["nI"] * L["nJ"] * self.pairs
        for h in (u_host, v_host):
            if h.dtype != np.float32 or h.size != n or not h.flags["C_CONTIGUOUS"]:
                raise QlpError(_lib.QLP_ERR_ARG, "result_async: host arrays of pairs x nI x nJ "
                                                 "contiguous float32")
        u, v, pk, pl = (ctypes.c_void_p() for _ in range(4))
        check(self.ctx.lib.qlp_piv_result_dev(self.h, lv, ctypes.byref(u), ctypes.byref(v),
                                              ctypes.byref(pk), ctypes.byref(pl)))
        _hip_copy_to_host_async(u_host, u.value, self.ctx.stream_handle)
        _hip_copy_to_host_async(v_host, v.value, self.ctx.stream_handle)

    def planes(self, level=-1):
        """Host copy of a level's stored correlation planes (pairs x nI x nJ x D x
        D), or None when the level keeps no planes."""
        lv = level % len(self.levels)
        L = self.levels[lv]
        u, v, pk, pl = (ctypes.c_void_p() for _ in range(4))
        self.ctx.synchronize()
        check(self.ctx.lib.qlp_piv_result_dev(self.h, lv, ctypes.byref(u), ctypes.byref(v),
                                              ctypes.byref(pk), ctypes.byref(pl)))
        if not pl.value:
            return None
        D = 2 * L["ia"] - 1
        n = L["nI"] * L["nJ"] * self.pairs
        return _fetch(pl, n * D * D, np.float32).reshape(self.pairs, L["nI"], L["nJ"], D, D)

    def validation(self, level=-1):
        """Host copies of a level's validation outcome: invalid flags
        (pairs x nI x nJ bool) and stats (pairs x 4: sweeps, corrected, left
        invalid, flags); None when validation is off."""
        lv = level % len(self.levels)
        L = self.levels[lv]
        inv, st = ctypes.c_void_p(), ctypes.c_void_p()
        self.ctx.synchronize()
        check(self.ctx.lib.qlp_piv_validation_dev(self.h, lv, ctypes.byref(inv), ctypes.byref(st)))
        if not inv.value:
            return None
        n = L["nI"] * L["nJ"] * self.pairs
        return (_fetch(inv, n, np.uint8).reshape(self.pairs, L["nI"], L["nJ"]).astype(bool),
                _fetch(st, 4 * self.pairs, np.int32).reshape(self.pairs, 4))


def _fetch(ptr, count, dtype):
    out = np.empty(count, dtype)
    _hip_copy_to_host(out, ptr.value)
    return out


_hip_rt = None


def _hip():
    """The process's HIP runtime (the one torch loaded)."""
    global _hip_rt
    if _hip_rt is None:
        import torch  # noqa: F401  (loads libamdhip64.so.7 first)

        rt = ctypes.CDLL("libamdhip64.so.7")
        rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        rt.hipMemcpy.restype = ctypes.c_int
        rt.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.c_int, ctypes.c_void_p]
        rt.hipMemcpyAsync.restype = ctypes.c_int
        _hip_rt = rt
    return _hip_rt


def _hip_copy_to_host(out: np.ndarray, src: int):
    """Device-to-host copy (synchronous)."""
    rc = _hip().hipMemcpy(out.ctypes.data, ctypes.c_void_p(src), out.nbytes, 2)  # D2H
    if rc != 0:
        raise QlpError(_lib.QLP_ERR_DEVICE, f"hipMemcpy failed ({rc})")


def _hip_copy_to_host_async(out: np.ndarray, src: int, stream: int):
    """Device-to-host copy enqueued on `stream`."""
    rc = _hip().hipMemcpyAsync(out.ctypes.data, ctypes.c_void_p(src), out.nbytes, 2,
                               ctypes.c_void_p(stream))
    if rc != 0:
        raise QlpError(_lib.QLP_ERR_DEVICE, f"hipMemcpyAsync failed ({rc})")
