"""ctypes binding of libqlpiv.so (the C ABI declared in include/qlpiv.h).

The library is built in-tree by ``build()`` (hipcc for gfx950) into
``quicklabpiv-ng_amd/_build/libqlpiv.so``.  Loading fails loudly when it is
missing: there is no CPU fallback behind this binding.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(PKG_ROOT, "csrc")
LIB_PATH = os.path.join(PKG_ROOT, "_build", "libqlpiv.so")

QLP_OK = 0
QLP_ERR_ARG = 1
QLP_ERR_DEVICE = 2
QLP_ERR_COMPUTE = 3
QLP_ERR_BUDGET = 4
QLP_ERR_UNSUPPORTED = 5

QLP_NORM_PER_TILE = 0
QLP_NORM_A_MAX = 1
QLP_PEAK_FROM_CENTER = 0
QLP_PEAK_MULTI = 1
QLP_PEAK_SIMPLE = 2
QLP_SUBPIX_GAUSS1D = 0
QLP_SUBPIX_GUO = 1
QLP_SUBPIX_POLY = 2
QLP_XCORR_EXACT = 0
QLP_XCORR_MFMA_F16 = 1
QLP_XCORR_MFMA_BF16 = 2
QLP_XCORR_MFMA_F16_SINGLE = 3
QLP_XCORR_MFMA_BF16_SINGLE = 4

# every symbol include/qlpiv.h declares, with (restype, argtypes)
_vp, _i, _sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t
SIGNATURES = {
    "qlp_ctx_create": (_i, [_i, _sz, ctypes.POINTER(_vp)]),
    "qlp_ctx_destroy": (_i, [_vp]),
    "qlp_ctx_synchronize": (_i, [_vp]),
    "qlp_ctx_stream": (_vp, [_vp]),
    "qlp_last_error": (ctypes.c_char_p, []),
    "qlp_version": (ctypes.c_char_p, []),
    "qlp_xcorr_batch": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _i, _vp]),
    "qlp_xcorr_batch_dev": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _i, _vp]),
    "qlp_xcorr_batch_noalloc_dev": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _i, _vp]),
    "qlp_xcorr_f32": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _i, _vp]),
    "qlp_xcorr_f32_dev": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _i, _vp]),
    "qlp_clip_tiles_f32_dev": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "qlp_warp_micro_tiles_f32_dev": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp,
                                          _vp]),
    "qlp_xcorr_form_dev": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp]),
    "qlp_find_peaks": (_i, [_vp, _i, _i, _i, _vp, _i, _i, _vp, _vp]),
    "qlp_find_peaks_dev": (_i, [_vp, _i, _i, _i, _vp, _i, _i, _vp, _vp]),
    "qlp_subpixel": (_i, [_vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    "qlp_subpixel_dev": (_i, [_vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    "qlp_clip_tiles_dev": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    "qlp_warp_micro_tiles_dev": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "qlp_warp_mini_tiles_dev": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp,
                                     _vp]),
    "qlp_warp_image_dev": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "qlp_xcorr_images_dev": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, _i, _vp, _vp, _vp, _vp, _i,
                                  _vp, _vp, _vp, _vp, _i]),
    "qlp_xcorr_images_form_dev": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _i, _i, _vp, _vp, _vp, _vp,
                                       _i, _vp, _vp, _vp, _vp, _i]),
    "qlp_piv_plan_create": (_i, [_vp, _i, _i, _i, _vp, ctypes.POINTER(_vp)]),
    "qlp_piv_plan_destroy": (_i, [_vp]),
    "qlp_piv_plan_levels": (_i, [_vp, ctypes.POINTER(_i)]),
    "qlp_piv_plan_level": (_i, [_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_i),
                                ctypes.POINTER(_i), _vp, _vp]),
    "qlp_piv_run_dev": (_i, [_vp, _vp, _vp]),
    "qlp_piv_run_f32_dev": (_i, [_vp, _vp, _vp]),
    "qlp_piv_run_u16_dev": (_i, [_vp, _vp, _vp]),
    "qlp_piv_result_dev": (_i, [_vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "qlp_piv_validation_dev": (_i, [_vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "qlp_piv_plan_set_timing": (_i, [_vp, _i]),
    "qlp_piv_plan_timing": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_float),
                                 ctypes.POINTER(ctypes.c_float),
                                 ctypes.POINTER(ctypes.c_float)]),
    "qlp_validate_vectors_dev": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp,
                                      _vp, _vp]),
    "qlp_validate_vectors": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp,
                                  _vp, _vp]),
    "qlp_gaussian_filter_dev": (_i, [_vp, _i, _i, _i, _vp, _i, ctypes.c_float, _i, _vp, _vp]),
    "qlp_inherit_velocity_dev": (_i, [_vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _i, _i, _i, _i,
                                      _vp, _vp, _i, _vp, _vp, _vp, _vp]),
    "qlp_piv_plan_estimate": (_i, [_i, _i, _i, _vp, ctypes.POINTER(_sz)]),
    "qlp_server_create": (_i, [_i, _vp, ctypes.POINTER(_vp)]),
    "qlp_server_destroy": (_i, [_vp]),
    "qlp_server_submit": (_i, [_vp, _vp, ctypes.POINTER(ctypes.c_uint64)]),
    "qlp_server_wait": (_i, [_vp, ctypes.c_uint64, _i, ctypes.POINTER(_i)]),
    "qlp_server_result": (_i, [_vp, ctypes.c_uint64, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                               ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i), _vp]),
    "qlp_server_release": (_i, [_vp, ctypes.c_uint64]),
    "qlp_server_get_stats": (_i, [_vp, _vp]),
    "qlp_warp_clip_tiles_dev": (_i, [_vp, _i, _i, _vp, _vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp,
                                     _vp, _vp, _vp, _vp, _vp]),
    "qlp_warp_images_dev": (_i, [_vp, _i, _vp, _vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp,
                                 _vp, _vp]),
    "qlp_image_read": (_i, [ctypes.c_char_p, _vp, _sz, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "qlp_image_read16": (_i, [ctypes.c_char_p, _vp, _sz, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "qlp_export_open": (_i, [ctypes.c_char_p, _vp, ctypes.c_char_p, ctypes.POINTER(_vp)]),
    "qlp_export_frame": (_i, [_vp, _i, _i, _i, _i, _vp, _i, _i, _vp, _vp, _vp]),
    "qlp_export_results": (_i, [_vp, _i, _i, _vp, _i, _i, _i, _i, _vp, _vp, _vp, _i]),
    "qlp_export_close": (_i, [_vp]),
    "qlp_batch_create": (_i, [_vp, _i, _vp, _vp, ctypes.POINTER(_vp)]),
    "qlp_batch_fix": (_i, [_vp, _i, _i, _i]),
    "qlp_batch_run_dev": (_i, [_vp, _vp, _vp]),
    "qlp_batch_grid": (_i, [_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_sz)]),
    "qlp_batch_result_dev": (_i, [_vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                  ctypes.POINTER(_vp)]),
    "qlp_batch_report_get": (_i, [_vp, _vp]),
    "qlp_batch_destroy": (_i, [_vp]),
}

QLP_SHAPE_AUTO = 0
QLP_SHAPE_LANE = 1
QLP_SHAPE_LDS = 2
QLP_SHAPE_LDS_HALF = 3
QLP_SHAPE_LDS_WIDE = 4
QLP_SCHEDULE_TUNED = -1
QLP_SCHEDULE_SERIAL = 0
QLP_SCHEDULE_CONCURRENT = 1

QLP_WARP_NONE = 0
QLP_WARP_MICRO = 1
QLP_WARP_MINI = 2
QLP_WARP_FULL = 3
QLP_WARP_ORDER_SECOND = 0
QLP_WARP_ORDER_FIRST = 1
QLP_WARP_ORDER_BOTH = 2
QLP_SIDE_B_SECOND, QLP_SIDE_A_FIRST, QLP_SIDE_A_BOTH, QLP_SIDE_B_BOTH = 0, 1, 2, 3
QLP_FRAME_U8, QLP_FRAME_F32, QLP_FRAME_U16 = 0, 1, 2


QLP_INHERIT_DIRECT = 0
QLP_INHERIT_AREA = 1
QLP_INHERIT_DISTANCE = 2
QLP_INHERIT_BICUBIC = 3

QLP_VALIDATOR_MEDIAN = 0
QLP_VALIDATOR_DIFFERENCE = 1
QLP_REPLACE_NONE = 0
QLP_REPLACE_SECONDARY_PEAK = 1
QLP_REPLACE_BILINEAR = 2
QLP_VALID_MAX_STAGES = 8
QLP_VALID_MAX_SWEEPS = 256


class ValidStage(ctypes.Structure):
    """qlp_valid_stage (include/qlpiv.h)."""
    _fields_ = [("validator", ctypes.c_int), ("threshold", ctypes.c_float),
                ("epsilon0", ctypes.c_float), ("replacement", ctypes.c_int),
                ("peak_index", ctypes.c_int)]


class ValidConfig(ctypes.Structure):
    """qlp_valid_config (include/qlpiv.h)."""
    _fields_ = [("n_stages", ctypes.c_int), ("max_iterations", ctypes.c_int),
                ("until_no_more", ctypes.c_int),
                ("stage", ValidStage * QLP_VALID_MAX_STAGES)]


def valid_config(stages, max_iterations=1, until_no_more=False):
    """ValidConfig from (validator, threshold, epsilon0, replacement, peak_index) tuples."""
    if len(stages) > QLP_VALID_MAX_STAGES:
        raise ValueError("at most 8 validation stages")
    c = ValidConfig()
    c.n_stages = len(stages)
    c.max_iterations = int(max_iterations)
    c.until_no_more = int(bool(until_no_more))
    for k, (val, thr, eps, rep, pk) in enumerate(stages):
        c.stage[k] = ValidStage(int(val), float(thr), float(eps), int(rep), int(pk))
    return c


def multipeak_median_stages(n_peaks, threshold, epsilon0):
    """The stages of the reference UI's multi-peak normalised median model
    (PIVConfigurationConverterVisitor.visit(...MultiPeakNormalizedMedianModel)
    :607-641): median + secondary peak 1 .. P-1, then median + bilinear."""
    st = [(QLP_VALIDATOR_MEDIAN, threshold, epsilon0, QLP_REPLACE_SECONDARY_PEAK, k + 1)
          for k in range(n_peaks - 1)]
    st.append((QLP_VALIDATOR_MEDIAN, threshold, epsilon0, QLP_REPLACE_BILINEAR, 0))
    return st


class PivConfig(ctypes.Structure):
    """qlp_piv_config (include/qlpiv.h)."""
    _fields_ = [("ia_start", ctypes.c_int), ("ia_end", ctypes.c_int),
                ("overlap_factor", ctypes.c_float),
                ("margin_top", ctypes.c_int), ("margin_bottom", ctypes.c_int),
                ("margin_left", ctypes.c_int), ("margin_right", ctypes.c_int),
                ("norm_mode", ctypes.c_int), ("warp_mode", ctypes.c_int),
                ("subpix_kind", ctypes.c_int), ("subpix_pixels", ctypes.c_int),
                ("subpix_iterations", ctypes.c_int), ("interpolator_start_step", ctypes.c_int),
                ("keep_planes", ctypes.c_int), ("peak_kind", ctypes.c_int),
                ("n_peaks", ctypes.c_int), ("peak_kernel_width", ctypes.c_int),
                ("valid", ValidConfig), ("inherit_mode", ctypes.c_int),
                ("xcorr_form", ctypes.c_int), ("warp_order", ctypes.c_int),
                ("frame_type", ctypes.c_int), ("frame_bits", ctypes.c_int)]


class ServerConfig(ctypes.Structure):
    """qlp_server_config (include/qlpiv.h)."""
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("slots", ctypes.c_int),
                ("batch_requests", ctypes.c_int)]


class Request(ctypes.Structure):
    """qlp_request (include/qlpiv.h)."""
    _fields_ = [("H", ctypes.c_int), ("W", ctypes.c_int), ("imgA", ctypes.c_void_p),
                ("imgB", ctypes.c_void_p), ("frames_on_device", ctypes.c_int),
                ("cfg", PivConfig)]


class RequestTimes(ctypes.Structure):
    """qlp_request_times (include/qlpiv.h)."""
    _fields_ = [("submit_ms", ctypes.c_double), ("admit_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double), ("ias", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    """qlp_server_stats (include/qlpiv.h)."""
    _fields_ = [("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("pending", ctypes.c_uint64), ("evictions", ctypes.c_uint64),
                ("hbm_in_use", ctypes.c_size_t), ("hbm_peak", ctypes.c_size_t),
                ("ias_done", ctypes.c_double), ("running", ctypes.c_int)]


class ExportParams(ctypes.Structure):
    """qlp_export_params (include/qlpiv.h)."""
    _fields_ = [("image_width", ctypes.c_int), ("image_height", ctypes.c_int),
                ("ia_start_i", ctypes.c_int), ("ia_end_i", ctypes.c_int),
                ("ia_start_j", ctypes.c_int), ("ia_end_j", ctypes.c_int),
                ("overlap_factor", ctypes.c_float), ("swap_uv", ctypes.c_int),
                ("mark_invalid_nan", ctypes.c_int), ("total_maps", ctypes.c_int),
                ("multi_volume", ctypes.c_int), ("next_filename", ctypes.c_char_p)]


class BatchPair(ctypes.Structure):
    """qlp_batch_pair (include/qlpiv.h)."""
    _fields_ = [("H", ctypes.c_int), ("W", ctypes.c_int), ("ia", ctypes.c_int),
                ("levels", ctypes.c_int)]


class BatchClassReport(ctypes.Structure):
    """qlp_batch_class_report (include/qlpiv.h)."""
    _fields_ = [("ia", ctypes.c_int), ("n_ias", ctypes.c_int), ("shape", ctypes.c_int),
                ("n_candidates", ctypes.c_int), ("cand_shape", ctypes.c_int * 4),
                ("cand_ms", ctypes.c_float * 4)]


class BatchReport(ctypes.Structure):
    """qlp_batch_report (include/qlpiv.h)."""
    _fields_ = [("n_classes", ctypes.c_int), ("cls", BatchClassReport * 4),
                ("schedule", ctypes.c_int), ("serial_ms", ctypes.c_float),
                ("concurrent_ms", ctypes.c_float)]


_lib = None


class QlpError(RuntimeError):
    """A non-zero status from the library; .code holds the QLP_ERR_* value."""

    def __init__(self, code, msg):
        super().__init__(f"qlpiv error {code}: {msg}")
        self.code = code


def build(jobs=8):
    """Compile libqlpiv.so for gfx950 (cross-compiles without a GPU)."""
    subprocess.check_call(["make", "-s", f"-j{jobs}", "-C", CSRC])


def load():
    global _lib
    if _lib is None:
        # One HIP runtime per process: PyTorch ships its own libamdhip64.so.7.  When
        # torch is importable it is loaded first, so that libqlpiv's dependency on
        # the same soname resolves to torch's copy instead of a second runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libqlpiv.so not found at {LIB_PATH}: build it first (__graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc):
    if rc != QLP_OK:
        raise QlpError(rc, load().qlp_last_error().decode())
    return rc
