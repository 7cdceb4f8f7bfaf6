/*
 * qlpiv.h -- C ABI of the MI355X PIV correlation library (libqlpiv.so).
 *
 * Drop-in for QuickLabPIV-ng's correlation hot path.  Paths in the citations
 * are relative to the reference's src/main/java/pt/quickLabPIV/.  The Java
 * side binds these entry points through JNI or the FFM API (see
 * INTEGRATION.md); nothing here uses C++ or PyTorch types.
 *
 * Conventions
 *   - Every entry point returns an int status: QLP_OK (0) or a QLP_ERR_* code;
 *     qlp_last_error() returns the message of the calling thread's last error.
 *     QLP_ERR_ARG corresponds to the reference's RuntimeException on invalid
 *     tile lists (CrossCorrelationFFTTemplate.analyzeTilesHelper :594-642),
 *     QLP_ERR_DEVICE to JobAnalyzeException, QLP_ERR_COMPUTE to
 *     JobComputeException.  Per-IA fit fallbacks are not errors: they are
 *     reported in the status field of each peak record.
 *   - Entry points without a _dev suffix take host buffers, copy them to the
 *     device, run, copy results back and synchronise.  _dev entry points take
 *     device pointers and enqueue on the context's stream without
 *     synchronising (qlp_ctx_synchronize waits).
 *   - A context owns one HIP stream.  Calls on one context are serialised by
 *     the caller (one context per calling thread, as the reference keeps one
 *     xcorr job per worker thread: OpenClGpuManager.compute :124-140).
 *   - Tiles are N x N row-major uint8 (IA pixels); N is 16, 32, 64 or 128
 *     (128: the correlation keeps its 256 x 256 tile in a context-owned HBM
 *     workspace of 512 KB per resident work-group, at most 128 MB).
 *     Correlation planes are (2N-1) x (2N-1) row-major float32, element
 *     [i][j] = sum_nm F'[n][m] G'[n+i-(N-1)][m+j-(N-1)], zero displacement at
 *     [N-1][N-1] (MaxCrossResult.java:141-149).
 */
#ifndef QLPIV_H
#define QLPIV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QLP_OK 0
#define QLP_ERR_ARG 1         /* invalid shapes, counts or options */
#define QLP_ERR_DEVICE 2      /* HIP device/runtime failure or missing device */
#define QLP_ERR_COMPUTE 3     /* failure while computing */
#define QLP_ERR_BUDGET 4      /* request exceeds the context's HBM budget */
#define QLP_ERR_UNSUPPORTED 5 /* valid in the reference, not provided here */

/* Normalisation / arithmetic form of the correlation (the two cross-correlation
 * jobs the reference pipeline selects, OpenClGpuManager.java:124-140). */
#define QLP_NORM_PER_TILE 0 /* CrossCorrelationRealFFTParStdJob (GPU job): F' = F/max(F)*16+1,
                               G' = G/max(G)*16+1, product / (2N)^2, inverse by re/im swap */
#define QLP_NORM_A_MAX 1    /* CrossCorrelationSoftRealFFTJob (Java CPU job): both tiles by
                               max(A), inverse twiddles, /2N per 1D pass */

/* Peak finders (maximum/MaximumFinderFactoryEnum.java:17-41). */
#define QLP_PEAK_FROM_CENTER 0 /* FindMaximumFromCenter */
#define QLP_PEAK_MULTI 1       /* FindMaximumMultiPeaks(numberOfPeaks, kernelWidth) */
#define QLP_PEAK_SIMPLE 2      /* FindMaximumSimple (maximum/FindMaximumSimple.java:19-48) */

/* Sub-pixel interpolators (interpolators/CrossCorrelationInterpolatorFactoryEnum.java). */
#define QLP_SUBPIX_GAUSS1D 0 /* Gaussian1DInterpolator(pixels) */
#define QLP_SUBPIX_GUO 1     /* Gaussian1DHongweiGuoInterpolator(pixels, iterations) */
#define QLP_SUBPIX_POLY 2    /* Gaussian1DPolynomialInterpolator */

/* Status bits of a sub-pixel result record (i, j, value, status). */
#define QLP_FIT_I 1        /* row fit accepted */
#define QLP_FIT_J 2        /* column fit accepted */
#define QLP_FIT_SKIPPED 4  /* peak at (0,0) or too close to the border: integer peak kept */
#define QLP_FIT_FLOOR 8    /* plane floor below the fitter's limit (reference throws) */

typedef struct qlp_ctx qlp_ctx;

/* ---- context ------------------------------------------------------------ */

/* Create a context on HIP device 'hip_device' with an HBM budget in bytes for
 * the library's own allocations (0 = no limit).  Replaces the device discovery
 * and per-device geometry of device/DeviceManager.java:25-136 and
 * CrossCorrelationFFTTemplate.computeLocalGeometry :307-438. */
int qlp_ctx_create(int hip_device, size_t hbm_budget_bytes, qlp_ctx **out);
/* Release a context (Job.dispose, LocalPIVOpenCLGpuJob.java:438-443). */
int qlp_ctx_destroy(qlp_ctx *ctx);
/* Wait for the work enqueued on the context's stream. */
int qlp_ctx_synchronize(qlp_ctx *ctx);
/* The context's hipStream_t, for callers that order their own work with it. */
void *qlp_ctx_stream(qlp_ctx *ctx);
/* Message of the calling thread's last error ("" if none). */
const char *qlp_last_error(void);
/* Library version string. */
const char *qlp_version(void);

/* ---- cross-correlation ---------------------------------------------------- */

/* Correlate n IA pairs.  Replaces Job<List<Tile>, XCorrelationResults>:
 * setInputParameters(JOB_RESULT_CLIPPED_TILES_A / _B) + analyze() + compute() +
 * getJobResult(JOB_RESULT_CROSS_MATRICES) of CrossCorrelationRealFFTParStdJob
 * (:35-292; kernel CrossCorrelationRealFFTParallelKernel.run :214-713) for
 * mode QLP_NORM_PER_TILE, or of CrossCorrelationSoftRealFFTJob.compute :94-121
 * (FastRealFFTXCorr.computeXCorr :44-68) for QLP_NORM_A_MAX.
 *   tilesA, tilesB : n x N x N uint8 (Tile.getMatrix() bytes; warped tiles hold
 *                    integer values 0..255 and are passed as uint8 too)
 *   maxA, maxB     : n floats, Matrix.getMaxValue() of each tile (maxB unused
 *                    in QLP_NORM_A_MAX)
 *   planes_out     : n x (2N-1) x (2N-1) floats, same order as the inputs. */
int qlp_xcorr_batch(qlp_ctx *ctx, int n, int N, const uint8_t *tilesA, const uint8_t *tilesB,
                    const float *maxA, const float *maxB, int mode, float *planes_out);
int qlp_xcorr_batch_dev(qlp_ctx *ctx, int n, int N, const uint8_t *tilesA, const uint8_t *tilesB,
                        const float *maxA, const float *maxB, int mode, float *planes_out);

/* qlp_xcorr_batch_dev without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * layouts and values as qlp_xcorr_batch_dev.  The 128 px correlation keeps its
 * 512 KB tiles in a library-owned static device array of 64 slots (32 MiB per
 * GPU, held until the process ends) and runs on at most 64 work-groups; 16, 32
 * and 64 px tiles stay in LDS.  Calls that overlap on one GPU share that array
 * and race: serialise them (one stream, or one call at a time). */
int qlp_xcorr_batch_noalloc_dev(qlp_ctx *ctx, int n, int N, const uint8_t *tilesA,
                                const uint8_t *tilesB, const float *maxA, const float *maxB,
                                int mode, float *planes_out);

/* The same job on float tiles: the MatrixFloat tiles of frames filtered before
 * warping (images/ImageFloat.java clipImageMatrix :278-384, fed by
 * GaussianFilter2D.applyFilter(IImage, IImage) :116-126 through
 * JavaImageFilterJob.compute :50-53), normalised as v / max * 16 + 1 by
 * Matrix.copyMatrixToArrayAndNormalizeAndOffset like the 8-bit ones.
 *   tilesA, tilesB : n x N x N floats
 *   maxA, maxB     : Matrix.getMaxValue() of each tile (qlp_clip_tiles_f32_dev,
 *                    qlp_warp_micro_tiles_f32_dev)
 * Exact form only; IA sizes 16, 32, 64 and 128. */
int qlp_xcorr_f32(qlp_ctx *ctx, int n, int N, const float *tilesA, const float *tilesB,
                  const float *maxA, const float *maxB, int mode, float *planes_out);
int qlp_xcorr_f32_dev(qlp_ctx *ctx, int n, int N, const float *tilesA, const float *tilesB,
                      const float *maxA, const float *maxB, int mode, float *planes_out);

/* Arithmetic form of the correlation.  QLP_XCORR_EXACT repeats the reference's
 * float FFT operations (planes value-identical to the Java job).  The MFMA forms
 * compute the same correlation as dense DFT-basis products on the matrix cores
 * in fp16 or bf16 with fp32 accumulation, with split hi + lo operands (or single
 * operands): the throughput form, accurate to the stated error against an FP64
 * correlation but not value-identical.  IA sizes 16, 32 and 64. */
#define QLP_XCORR_EXACT 0
#define QLP_XCORR_MFMA_F16 1         /* fp16 hi + lo operands */
#define QLP_XCORR_MFMA_BF16 2        /* bf16 hi + lo operands */
#define QLP_XCORR_MFMA_F16_SINGLE 3  /* fp16 operands */
#define QLP_XCORR_MFMA_BF16_SINGLE 4 /* bf16 operands */

/* qlp_xcorr_batch_dev in the chosen form (same arguments and layouts).  No
 * counterpart in the reference: the matrix-core form of the correlation job. */
int qlp_xcorr_form_dev(qlp_ctx *ctx, int form, int n, int N, const uint8_t *tilesA,
                       const uint8_t *tilesB, const float *maxA, const float *maxB, int mode,
                       float *planes_out);

/* ---- integer peak search -------------------------------------------------- */

/* IMaximumFinder.findMaximum(Matrix) (maximum/IMaximumFinder.java:16-17) over n
 * planes of D x D.  kind QLP_PEAK_FROM_CENTER (FindMaximumFromCenter.findMaximum
 * :47-183; n_peaks must be 1), QLP_PEAK_MULTI (FindMaximumMultiPeaks.findMaximum
 * :42-162 with numberOfPeaks = n_peaks <= 16 and kernelWidth) or QLP_PEAK_SIMPLE
 * (FindMaximumSimple.findMaximum :19-48: the first strictly largest positive
 * value in row-major order, (D/2, D/2) with value 0 when none is positive;
 * n_peaks must be 1).
 *   peaks     : n x n_peaks x 3 floats (i, j, value), peak 0 the main peak;
 *               missing peaks are zero
 *   min_floor : n floats (MaxCrossResult.getMinFloor) */
int qlp_find_peaks(qlp_ctx *ctx, int kind, int n, int D, const float *planes, int n_peaks,
                   int kernel_width, float *peaks, float *min_floor);
int qlp_find_peaks_dev(qlp_ctx *ctx, int kind, int n, int D, const float *planes, int n_peaks,
                       int kernel_width, float *peaks, float *min_floor);

/* ---- sub-pixel refinement ------------------------------------------------- */

/* ICrossCorrelationInterpolator.interpolate(List<MaxCrossResult>)
 * (interpolators/ICrossCorrelationInterpolator.java:14-15, per-IA loop of
 * BasicInterpolatorAdapter.interpolate :26-39) for every peak of n planes.
 *   peaks_in  : n x n_peaks x 3 (i, j, value) from qlp_find_peaks
 *   peaks_out : n x n_peaks x 4 (i', j', value', status); the displacement is
 *               u = i' - (D/2), v = j' - (D/2) with integer D/2 (MaxCrossResult
 *               .setNthRelativeDisplacementFromPeak :304-326). */
int qlp_subpixel(qlp_ctx *ctx, int kind, int pixels, int iterations, int n, int D, int n_peaks,
                 const float *planes, const float *min_floor, const float *peaks_in,
                 float *peaks_out);
int qlp_subpixel_dev(qlp_ctx *ctx, int kind, int pixels, int iterations, int n, int D,
                     int n_peaks, const float *planes, const float *min_floor,
                     const float *peaks_in, float *peaks_out);

/* ---- tiles from images ---------------------------------------------------- */

/* Image.clipImageMatrix (images/Image.java:386-494) for n windows of N x N at
 * (top[k], left[k]) of an H x W uint8 image; pixels outside the image are 0;
 * max_out[k] = max over in-image pixels clamped to >= 16 (MatrixByte.java:16-39). */
int qlp_clip_tiles_dev(qlp_ctx *ctx, const uint8_t *img, int H, int W, int n, int N,
                       const int *top, const int *left, uint8_t *tiles, float *max_out);

/* Bilinear micro warp of second-image tiles
 * (images/BiLinearImageMicroAndMiniWarpingStrategy.warpAndClipImage :132-148,
 * SecondImage mode with createMicroDisplacements): window at
 * (top + floor(u), left + floor(v)), fractional remainder resampled;
 * max_out[k] = 255 (may be NULL). */
int qlp_warp_micro_tiles_dev(qlp_ctx *ctx, const uint8_t *img, int H, int W, int n, int N,
                             const int *top, const int *left, const float *u, const float *v,
                             uint8_t *tiles, float *max_out);

/* Tiles of a float frame (images/ImageFloat.java, e.g. a frame filtered by
 * GaussianFilter2D :116-126).  ImageFloat.clipImageMatrix :278-384: n windows
 * of N x N at (top[k], left[k]), or at (round(top[k] + u[k]), round(left[k] +
 * v[k])) when u and v are given (NoImageWarpingStrategy :80-81), zero outside;
 * max_out[k] = max over the tile, 0 -> 1 (MatrixFloat.java:45-60,
 * Matrix.setMaxValue :21-29; no >= 16 clamp). */
int qlp_clip_tiles_f32_dev(qlp_ctx *ctx, const float *img, int H, int W, int n, int N,
                           const int *top, const int *left, const float *u, const float *v,
                           float *tiles, float *max_out);

/* Sides of a Micro tile warp (BiLinearImageMicroAndMiniWarpingStrategy
 * .warpAndClipImage :63-167): the B tile in SecondImage order, the A tile in
 * FirstImage order, and the A or B tile in BothImages order. */
#define QLP_SIDE_B_SECOND 0
#define QLP_SIDE_A_FIRST 1
#define QLP_SIDE_A_BOTH 2
#define QLP_SIDE_B_BOTH 3

/* The Micro warp of one side of n tiles of a float frame, whose B tiles carry
 * the displacements (u, v) (doBiLinearWarp :292-334 reading ImageFloat
 * .readPixel :531-533).  Values are the rounded samples (they may exceed 255);
 * max_out[k] = 65535, the ImageFloat maximum (getMaximumValue :541-543,
 * doBiLinearWarp :298); max_out may be NULL. */
int qlp_warp_micro_tiles_f32_dev(qlp_ctx *ctx, const float *img, int H, int W, int n, int N,
                                 const int *top, const int *left, const float *u, const float *v,
                                 int side, float *tiles, float *max_out);

/* Bilinear mini warp of the second-image tiles of 'pairs' frames (img: pairs x
 * H x W) over a grid of nI x nJ tiles of N px (BiLinearImageMicroAndMiniWarping
 * Strategy.warpAndClipImage :132-148 with createMiniDisplacements :203-290):
 * window at (top + floor(u), left + floor(v)), each pixel resampled at the
 * spline field of the tile's 7 x 7 neighbourhood (neighbours' displacements
 * minus floor of the tile's own; BiCubicSplineInterpolatorWithBiLinearBackup).
 * top/left: device, nI*nJ origins; u, v: device, pairs x nI*nJ; tiles: pairs x
 * nI*nJ x N x N; max_out = 255 (may be NULL).  nI, nJ >= 2. */
int qlp_warp_mini_tiles_dev(qlp_ctx *ctx, const uint8_t *img, int H, int W, int pairs, int nI,
                            int nJ, int N, const int *top, const int *left, const float *u,
                            const float *v, uint8_t *tiles, float *max_out);
/* Full-image bilinear warp of 'pairs' second frames (BiLinearImageWarpingStrategy
 * .warpAndClipImage :38-141, SecondImage): every pixel resampled at the spline
 * field of the whole grid (createTileDisplacementInterpolator).  top/left: HOST
 * arrays of the grid's nI*nJ origins (the knots); u, v: device, pairs x nI*nJ;
 * out: device, pairs x H x W.  nI, nJ >= 2. */
int qlp_warp_image_dev(qlp_ctx *ctx, const uint8_t *img, int H, int W, int pairs, int nI, int nJ,
                       int N, const int *top, const int *left, const float *u, const float *v,
                       uint8_t *out);

/* ---- correlation pass straight from device-resident images ---------------- */

#define QLP_WARP_NONE 0  /* NoImageWarpingStrategy: B window at round(top+u), round(left+v) */
#define QLP_WARP_MICRO 1 /* BiLinearImageMicroAndMiniWarpingStrategy (micro) */
#define QLP_WARP_MINI 2  /* BiLinearImageMicroAndMiniWarpingStrategy (mini) */
#define QLP_WARP_FULL 3  /* BiLinearImageWarpingStrategy (whole frames) */

/* Warp orders (BiLinearImageWarpingModeEnum; WarpingModeFactoryEnum.create
 * :37-76 pairs each with Micro, Mini and Full): which frames are warped. */
#define QLP_WARP_ORDER_SECOND 0 /* B sampled at target + u; A clipped */
#define QLP_WARP_ORDER_FIRST 1  /* A sampled at target - u; B clipped */
#define QLP_WARP_ORDER_BOTH 2   /* A at target - u/2, B at target + u/2 (plus the
                                   reference's rounding offsets for Micro / Mini) */

/* IImageWarpingStrategy.warpAndClipImage of the Micro / Mini strategies at a
 * level > 0 (images/BiLinearImageMicroAndMiniWarpingStrategy.java:39-186) for
 * 'pairs' frame pairs on one nI x nJ grid of N px IAs: imgsA, imgsB pairs x H x W;
 * top, left: nI*nJ origins (device); u, v: pairs x nI*nJ B-tile displacements.
 * Writes tilesA/tilesB (pairs x nI*nJ x N x N) and their maxima: 255 for a warped
 * tile (H4), max(tile) >= 16 for a clipped one.  warp_mode QLP_WARP_MICRO or
 * QLP_WARP_MINI (Mini needs 2 tiles in each direction). */
int qlp_warp_clip_tiles_dev(qlp_ctx *ctx, int warp_mode, int order, const uint8_t *imgsA,
                            const uint8_t *imgsB, int H, int W, int pairs, int nI, int nJ, int N,
                            const int *top, const int *left, const float *u, const float *v,
                            uint8_t *tilesA, float *maxA, uint8_t *tilesB, float *maxB);
/* The frames BiLinearImageWarpingStrategy.warpAndClipImage warps at a level > 0
 * (images/BiLinearImageWarpingStrategy.java:38-183) for one warp order: outA
 * (FirstImage, BothImages) and outB (SecondImage, BothImages), pairs x H x W
 * each; unused ones may be NULL.  top, left: HOST arrays of nI*nJ origins; u, v:
 * device, pairs x nI*nJ. */
int qlp_warp_images_dev(qlp_ctx *ctx, int order, const uint8_t *imgsA, const uint8_t *imgsB, int H,
                        int W, int pairs, int nI, int nJ, int N, const int *top, const int *left,
                        const float *u, const float *v, uint8_t *outA, uint8_t *outB);

/* Clip/warp, correlate and FromCenter-search n IAs of one frame pair in a
 * single kernel (the fused form of qlp_clip_tiles_dev + qlp_warp_micro_tiles_dev
 * + qlp_xcorr_batch_dev + qlp_find_peaks_dev(FROM_CENTER)).  top/left: IA
 * origins; du/dv: B displacement per IA (NULL = none).  Tile maxima are computed
 * in the kernel (>= 16 clamp; 255 for warped tiles).  Outputs may be NULL:
 * planes (n x D x D), peaks (n x 3: i, j, value), min_floor (n), nbr (n x 2 x
 * (2*nbr_half+1): the plane column then row through the peak, for the fits). */
int qlp_xcorr_images_dev(qlp_ctx *ctx, int n, int N, int mode, const uint8_t *imgA,
                         const uint8_t *imgB, int H, int W, const int *top, const int *left,
                         const float *du, const float *dv, int warp_mode, float *planes,
                         float *peaks, float *min_floor, float *nbr, int nbr_half);
/* qlp_xcorr_images_dev in the chosen arithmetic form (QLP_XCORR_*).  No
 * counterpart in the reference: the matrix-core form of the fused pass. */
int qlp_xcorr_images_form_dev(qlp_ctx *ctx, int form, int n, int N, int mode, const uint8_t *imgA,
                              const uint8_t *imgB, int H, int W, const int *top, const int *left,
                              const float *du, const float *dv, int warp_mode, float *planes,
                              float *peaks, float *min_floor, float *nbr, int nbr_half);

/* ---- vector validation and replacement ------------------------------------ */

#define QLP_VALIDATOR_MEDIAN 0     /* iareas/validation/NormalizedMedianValidator.java */
#define QLP_VALIDATOR_DIFFERENCE 1 /* iareas/validation/DifferenceValidator.java */
#define QLP_REPLACE_NONE 0         /* no replacement (VectorReplacementFactoryEnum.None) */
#define QLP_REPLACE_SECONDARY_PEAK 1 /* iareas/replacement/SecondaryPeakReplacement.java */
#define QLP_REPLACE_BILINEAR 2     /* iareas/replacement/BilinearReplacement.java */
#define QLP_VALID_MAX_STAGES 8
/* Upper bound on validation sweeps when iterating until no more corrections
 * (the reference has none: a field whose corrections oscillate would not end). */
#define QLP_VALID_MAX_SWEEPS 256

/* One CombinedValidatorAndReplacementConfiguration
 * (iareas/validation/CombinedValidatorAndReplacementConfiguration.java:12-42). */
typedef struct qlp_valid_stage {
    int validator;    /* QLP_VALIDATOR_* */
    float threshold;  /* distance threshold, pixels */
    float epsilon0;   /* normalised median only */
    int replacement;  /* QLP_REPLACE_* */
    int peak_index;   /* QLP_REPLACE_SECONDARY_PEAK: peak used (1 = first secondary) */
} qlp_valid_stage;

/* VectorValidatorConfiguration (jobs/validator/VectorValidatorConfiguration.java)
 * plus the stage array of the CombinedValidator; n_stages = 0 turns validation
 * off.  The multi-peak normalised median model of the reference UI
 * (PIVConfigurationConverterVisitor.visit(...MultiPeakNormalizedMedianModel)
 * :607-641) is P stages: median + secondary peak 1 .. P-1, then median +
 * bilinear. */
typedef struct qlp_valid_config {
    int n_stages;
    int max_iterations;  /* maxCorrectionIterations */
    int until_no_more;   /* iterateUntilNoMoreCorrections */
    qlp_valid_stage stage[QLP_VALID_MAX_STAGES];
} qlp_valid_config;

/* VectorValidatorJob.compute (jobs/validator/VectorValidatorJob.java:243-405):
 * validate, replace and re-validate the displacement field of 'pairs' grids of
 * nI x nJ tiles in the reference's centre-out spiral visit order, then set the
 * vectors still invalid to (0, 0).
 *   top, left  : host arrays of nI*nJ tile origins (Tile.getTopPixel/getLeftPixel)
 *   u, v       : device, pairs x nI*nJ accumulated displacements (in/out)
 *   peak_uv    : device, pairs x nI*nJ x n_peaks x 2 displacement increments
 *                (MaxCrossResult.getNthDisplacementU/V)
 *   peak_count : device, pairs x nI*nJ (MaxCrossResult.getTotalPeaks) or NULL
 *                (every tile holds n_peaks peaks)
 *   invalid    : device, pairs x nI*nJ bytes, Tile.isInvalidDisplacement()
 *                after the job (may be NULL)
 *   stats      : device, pairs x 4 ints: sweeps, corrected vectors, vectors
 *                left invalid, flags (1: sweep limit reached, 2: a bilinear
 *                replacement met a grid the reference rejects with
 *                VectorReplacementException) (may be NULL) */
int qlp_validate_vectors_dev(qlp_ctx *ctx, int pairs, int nI, int nJ, const int *top,
                             const int *left, float *u, float *v, int n_peaks,
                             const float *peak_uv, const int *peak_count,
                             const qlp_valid_config *cfg, uint8_t *invalid, int *stats);
/* The same with every array in host memory (synchronous). */
int qlp_validate_vectors(qlp_ctx *ctx, int pairs, int nI, int nJ, const int *top,
                         const int *left, float *u, float *v, int n_peaks, const float *peak_uv,
                         const int *peak_count, const qlp_valid_config *cfg, uint8_t *invalid,
                         int *stats);

/* ---- image pre-filter ----------------------------------------------------- */

/* GaussianFilter2D.applyFilter (images/filters/GaussianFilter2D.java:82-114,
 * GaussianFilter2DInternal.applyFilter :31-171, DirectConvolution.convolve1D)
 * on n_img images of H x W (device; uint8, or float when in_is_float): a
 * separable Gaussian of kernel_px taps (<= 31) and the given sigma, rows then
 * columns; tmp and out are device float images of the same size (tmp holds
 * the row pass).  sigma < 1e-5 copies the input. */
int qlp_gaussian_filter_dev(qlp_ctx *ctx, int n_img, int H, int W, const void *in,
                            int in_is_float, float sigma, int kernel_px, float *tmp, float *out);

/* ---- velocity inheritance between adaptive levels ------------------------- */

#define QLP_INHERIT_DIRECT 0   /* AdaptiveInterVelocityInheritanceStrategyDirect */
#define QLP_INHERIT_AREA 1     /* AdaptiveInterVelocityInheritanceStrategyArea */
#define QLP_INHERIT_DISTANCE 2 /* AdaptiveInterVelocityInheritanceStrategyDistanceGeneric */
#define QLP_INHERIT_BICUBIC 3  /* AdaptiveInterVelocityInheritanceStrategyBiCubicSpline */

/* IInterAreaVelocityInheritanceStrategy.reuseIterationStepTilesParameters
 * (iareas/IInterAreaVelocityInheritanceStrategy.java) for 'pairs' frames: the
 * child level's tiles are reset and take the parent level's displacements.
 *   parent grid : pnI x pnJ tiles of pH x pW pixels, origins ptop/pleft (host,
 *                 row-major, pnI*pnJ each)
 *   child grid  : cnI x cnJ tiles of cH x cW pixels, origins ctop/cleft (host)
 *   round       : 1 for warping modes that need integer displacements
 *                 (WarpingModeFactoryEnum.isRequiresRounding)
 *   pu, pv      : device, pairs x pnI*pnJ parent displacements
 *   u, v        : device, pairs x cnI*cnJ child displacements (out) */
int qlp_inherit_velocity_dev(qlp_ctx *ctx, int mode, int pairs, int pnI, int pnJ, int pH, int pW,
                             const int *ptop, const int *pleft, int cnI, int cnJ, int cH, int cW,
                             const int *ctop, const int *cleft, int round, const float *pu,
                             const float *pv, float *u, float *v);

/* ---- multi-pass PIV over device-resident frame pairs ---------------------- */

/* The per-frame adaptive loop of jobs/LocalPIVOpenCLGpuJob.call :188-409 with
 * the super-position IA strategy (iareas/AdaptiveInterAreaStrategySuperPosition
 * .java:31-197), the chosen velocity inheritance between levels (Direct by
 * default, AdaptiveInterVelocityInheritanceStrategyDirect.java), warping,
 * peak search and sub-pixel fit, the displacement update
 * (IterationStepTiles.updateDisplacementsFromMaxCrossResults :317-371) and the
 * optional vector validation (:395-403). */
typedef struct qlp_piv_config {
    int ia_start, ia_end;     /* IA size of the first and of the last level (halving) */
    float overlap_factor;     /* 1 - overlap: 0.5 for 50 % overlap */
    int margin_top, margin_bottom, margin_left, margin_right;
    int norm_mode;            /* QLP_NORM_* */
    int warp_mode;            /* QLP_WARP_* */
    int subpix_kind;          /* QLP_SUBPIX_* or -1 for none */
    int subpix_pixels, subpix_iterations;
    int interpolator_start_step; /* first level that is interpolated */
    int keep_planes;          /* 0: no planes; 1: store the last level's planes; 2: store
                                 every level's planes (what the Java job returns) */
    int peak_kind;            /* QLP_PEAK_FROM_CENTER (fused into the pass), QLP_PEAK_MULTI
                                 or QLP_PEAK_SIMPLE (planes are then stored at every
                                 level) */
    int n_peaks;              /* MultiPeaks numberOfPeaks (1 for FromCenter) */
    int peak_kernel_width;    /* MultiPeaks kernelWidth */
    qlp_valid_config valid;   /* vector validation after every level (n_stages 0: off) */
    int inherit_mode;         /* QLP_INHERIT_* between levels (0: Direct) */
    int xcorr_form;           /* QLP_XCORR_* arithmetic of the 16, 32 and 64 px levels
                                 (128 px levels always use QLP_XCORR_EXACT) */
    int warp_order;           /* QLP_WARP_ORDER_* of the Micro, Mini and Full warps */
    int frame_type;           /* QLP_FRAME_U8 (qlp_piv_run_dev), QLP_FRAME_F32
                                 (qlp_piv_run_f32_dev) or QLP_FRAME_U16
                                 (qlp_piv_run_u16_dev) */
    int frame_bits;           /* QLP_FRAME_U16: bit depth 10, 12 or 16 */
} qlp_piv_config;

/* Pixel type of a plan's frames: 8-bit (images/Image.java), float
 * (images/ImageFloat.java, what GaussianFilter2D.applyFilter(IImage, IImage)
 * :116-126 writes when frames are filtered before warping), or 10/12/16-bit
 * (images/ImageInt16.java, pixel depths Image10Bit / Image12Bit / Image16Bit of
 * ImageFactoryEnum.create :20-31). */
#define QLP_FRAME_U8 0
#define QLP_FRAME_F32 1
#define QLP_FRAME_U16 2

typedef struct qlp_piv_plan qlp_piv_plan;

/* Build the level geometry for H x W frames and allocate the device state of
 * 'pairs' frame pairs processed together (inside the context's HBM budget). */
int qlp_piv_plan_create(qlp_ctx *ctx, int H, int W, int pairs, const qlp_piv_config *cfg,
                        qlp_piv_plan **out);
int qlp_piv_plan_destroy(qlp_piv_plan *plan);
/* Number of levels; IA size and grid of one level; the origins (host arrays of
 * nI * nJ ints, row-major) of its IAs. */
int qlp_piv_plan_levels(qlp_piv_plan *plan, int *n_levels);
int qlp_piv_plan_level(qlp_piv_plan *plan, int level, int *ia, int *nI, int *nJ, int *top,
                       int *left);
/* Run every level for the frames imgsA, imgsB (device, pairs x H x W uint8),
 * asynchronously on the context's stream. */
int qlp_piv_run_dev(qlp_piv_plan *plan, const uint8_t *imgsA, const uint8_t *imgsB);
/* The same loop for float frames (plan created with frame_type QLP_FRAME_F32;
 * imgsA, imgsB device, pairs x H x W floats): every level cuts float tiles
 * (ImageFloat.clipImageMatrix :278-384; the Micro warp of the order's sides,
 * maximum 65535; the Mini warp likewise; the Full warp's tiles cut from the
 * warped frames with the source frame's maximum, MatrixFloat.copyMatrixRegion
 * :437-455) and correlates them in the exact form.  Every warp mode and order. */
int qlp_piv_run_f32_dev(qlp_piv_plan *plan, const float *imgsA, const float *imgsB);
/* The loop for 10/12/16-bit frames (frame_type QLP_FRAME_U16; imgsA, imgsB
 * device, pairs x H x W uint16): the frames are converted to float in plan
 * buffers and run as float frames with the ImageInt16 rules: clipped tiles take
 * the MatrixInt16 maximum (>= 16, MatrixInt16.java:16-41, :371-405), Micro-warped
 * tiles the bit-depth maximum 2^bits - 1 (ImageInt16.getMaximumValue :557-559).
 * No warp, Micro or Mini; the reference's Full warp of these frames truncates
 * each sample to a byte (MatrixInt16.setElement(float) :95-101) and returns
 * QLP_ERR_UNSUPPORTED here. */
int qlp_piv_run_u16_dev(qlp_piv_plan *plan, const uint16_t *imgsA, const uint16_t *imgsB);
/* Device pointers of a level's results (valid after the run has completed):
 * u, v (pairs x n displacements, vertical then horizontal), peaks (pairs x n x
 * n_peaks x 4: i', j', value', status), planes (pairs x n x D x D when stored
 * for that level; else NULL). */
int qlp_piv_result_dev(qlp_piv_plan *plan, int level, const float **u, const float **v,
                       const float **peaks, const float **planes);
/* Device pointers of a level's validation outcome (NULL when validation is
 * off): invalid (pairs x n bytes) and stats (pairs x 4 ints), as written by
 * qlp_validate_vectors_dev. */
int qlp_piv_validation_dev(qlp_piv_plan *plan, int level, const uint8_t **invalid,
                           const int **stats);
/* Per-level device timing of later runs (HIP events on the context's stream):
 * the displacement-state kernel plus the correlation pass; the peak search (when
 * not fused), sub-pixel fit and displacement update; and the vector validation,
 * in milliseconds, of the last run.  Plays the role of the reference's per-job
 * wall-clock logging (ExecutionStatus). */
int qlp_piv_plan_set_timing(qlp_piv_plan *plan, int enable);
int qlp_piv_plan_timing(qlp_piv_plan *plan, int level, float *state_and_pass_ms, float *fit_ms,
                        float *valid_ms);
/* Device bytes a plan of this configuration allocates (host-only computation;
 * the budget check of qlp_piv_plan_create and of the server's admission). */
int qlp_piv_plan_estimate(int H, int W, int pairs, const qlp_piv_config *cfg, size_t *bytes);

/* ---- batches of frame pairs of mixed sizes, one graph launch per level -------
 *
 * No counterpart in the reference, which runs its per-frame job
 * (jobs/LocalPIVOpenCLGpuJob.call :188-409: clip or micro warp, correlation,
 * FromCenter peak, sub-pixel fit, Direct inheritance between the adaptive
 * levels) one frame pair at a time.  A batch holds many pairs, each with its
 * own frame size, last IA size (16, 32, 64 or 128) and number of adaptive
 * levels, and runs level k of every pair as ONE launch of a HIP graph of
 * persistent kernels, captured once and replayed for every level and every new
 * set of frames (a step table in device memory tells the kernels what each
 * launch holds).  Per IA the results equal those of a multi-pass plan of that
 * pair alone (ia_start = ia << (levels - 1), ia_end = ia) with the same
 * configuration.  Configurations: FromCenter peak, Direct inheritance, no warp
 * or the SecondImage micro warp, no validation, 8-bit frames, any fit and
 * arithmetic form (QLP_ERR_UNSUPPORTED otherwise).
 *
 * Autotuner: on the first run every IA-size class is timed with each kernel
 * shape that computes its correlation (QLP_SHAPE_*; the exact shapes give the
 * same values), the fastest is kept, and the graph is built twice, with the
 * classes in sequence on one stream and forked onto one stream each; the faster
 * schedule is kept.  Measured shapes are cached in the context by (IA size, IA
 * count, arithmetic form, normalisation), so later batches skip the timing. */
#define QLP_SHAPE_AUTO 0      /* the library's defaults, no measurement */
#define QLP_SHAPE_LANE 1      /* one lane per transform line (16 px IAs) */
#define QLP_SHAPE_LDS 2       /* one work-group per IA, default size */
#define QLP_SHAPE_LDS_HALF 3  /* one work-group per IA, half size */
#define QLP_SHAPE_LDS_WIDE 4  /* one work-group per IA, twice the size (16, 32, 64 px) */
#define QLP_SCHEDULE_TUNED (-1)
#define QLP_SCHEDULE_SERIAL 0
#define QLP_SCHEDULE_CONCURRENT 1

typedef struct qlp_batch qlp_batch;
typedef struct {
    int H, W;   /* frame size */
    int ia;     /* IA size of the last level: 16, 32, 64 or 128 */
    int levels; /* adaptive levels, the first at ia << (levels - 1) (0 or 1: one) */
} qlp_batch_pair;
typedef struct {
    int ia, n_ias;     /* IA size and IAs of the class */
    int shape;         /* QLP_SHAPE_* in use */
    int n_candidates;  /* shapes timed (0: fixed or from the context's cache) */
    int cand_shape[4];
    float cand_ms[4];  /* best of 3 timed runs of the class with each shape */
} qlp_batch_class_report;
typedef struct {
    int n_classes;
    qlp_batch_class_report cls[4];
    int schedule;                    /* QLP_SCHEDULE_SERIAL or _CONCURRENT in use */
    float serial_ms, concurrent_ms;  /* best of 3 graph launches (0 if not timed) */
} qlp_batch_report;

/* Lay out the IAs of every level of n_pairs pairs (the plan's level grids with
 * the config's overlap and margins) and allocate their device state.  Used
 * fields of cfg: overlap_factor, margins, norm_mode, warp_mode (NONE or MICRO),
 * subpix_kind / _pixels / _iterations, interpolator_start_step, xcorr_form
 * (16, 32 and 64 px classes); the others must keep their defaults. */
int qlp_batch_create(qlp_ctx *ctx, int n_pairs, const qlp_batch_pair *pairs,
                     const qlp_piv_config *cfg, qlp_batch **out);
/* Fix the kernel shape of the class of IA size 'ia' (ia 0: every class) and/or
 * the schedule (QLP_SCHEDULE_TUNED keeps it measured), before the first run. */
int qlp_batch_fix(qlp_batch *b, int ia, int shape, int schedule);
/* Run every level of every pair: imgsA[k], imgsB[k] are device pointers to
 * pair k's H x W uint8 frames (any address each run).  Asynchronous on the
 * context's stream; the first run tunes and builds the graph (synchronising). */
int qlp_batch_run_dev(qlp_batch *b, const uint8_t *const *imgsA, const uint8_t *const *imgsB);
/* Last-level grid of pair k and the IA count of the batch (every level). */
int qlp_batch_grid(qlp_batch *b, int pair, int *nI, int *nJ, size_t *total_ias);
/* Device pointers of pair k's last-level results (valid once the run has
 * completed): u, v (nI*nJ displacements, vertical then horizontal) and sub
 * (nI*nJ x 4: i', j', value', status of the sub-pixel fit). */
int qlp_batch_result_dev(qlp_batch *b, int pair, const float **u, const float **v,
                         const float **sub);
int qlp_batch_report_get(qlp_batch *b, qlp_batch_report *report);
int qlp_batch_destroy(qlp_batch *b);

/* ---- frame ingest --------------------------------------------------------------
 *
 * jobs/ImageReaderJob.compute :80-110 with images/Image.readImageToBuffer
 * :146-156 for 8-bit frames (the TYPE_BYTE_GRAY raster copied as it is), by
 * host code: 8-bit grayscale (black is zero) TIFF strips, uncompressed, PackBits,
 * LZW or Deflate (predictor 1 or 2), 8-bit grayscale non-interlaced PNG, and
 * baseline / extended sequential one-component 8-bit JPEG (with restart
 * intervals; the integer inverse DCT of IJG-derived decoders such as the one
 * javax.imageio wraps).  Colour, palette, alpha, other depths, tiles and
 * progressive or arithmetic-coded JPEG return QLP_ERR_UNSUPPORTED (the
 * reference converts them through Java2D). */
/* Decode 'path' into out (H x W bytes, row-major); out NULL only reports H, W. */
int qlp_image_read(const char *path, uint8_t *out, size_t out_bytes, int *H, int *W);
/* 10/12/16-bit frames (images/ImageInt16.readImageToBuffer :183-184, the
 * TYPE_USHORT_GRAY raster copied as it is): 16-bit grayscale TIFF strips (either
 * byte order; uncompressed, PackBits, LZW or Deflate; predictor 1 or 2) and
 * 16-bit grayscale non-interlaced PNG, into out (H x W uint16).  8-bit files,
 * which the reference widens through Java2D, return QLP_ERR_UNSUPPORTED. */
int qlp_image_read16(const char *path, uint16_t *out, size_t out_elems, int *H, int *W);

/* ---- result export -----------------------------------------------------------
 *
 * The MATLAB Level 5 velocity file of exporter/StructMultiFrameFloatVelocityExporter
 * .java (openFile :54-113, exportDataToFile(frameNumber, IterationStepTiles)
 * :131-180, closeFile :115-129), written by host code: "exportDate", the struct
 * "parameters" and per IA size the struct "velocities<H>x<W>" (iaHeight,
 * iaWidth, adaptiveStep, margins, u and v as single nI x nJ x total_maps),
 * uncompressed, in the reference's element layout. */
typedef struct qlp_exporter qlp_exporter;
typedef struct {
    int image_width, image_height;
    int ia_start_i, ia_end_i, ia_start_j, ia_end_j;
    float overlap_factor;
    int swap_uv;               /* PIVMapOptionalConfiguration.isSwapUVOrder */
    int mark_invalid_nan;      /* invalid tiles exported as NaN */
    int total_maps;            /* velocity maps (frames) in the file, >= 1 */
    int multi_volume;          /* the file is one volume of several */
    const char *next_filename; /* the next volume's file (multi_volume), NULL for the last */
} qlp_export_params;
/* Create (truncate) the file and record the parameters; date_text NULL takes the
 * current time as java.util.Date.toString() prints it. */
int qlp_export_open(const char *filename, const qlp_export_params *params, const char *date_text,
                    qlp_exporter **out);
/* One map (frame index < total_maps) of one level: u, v, invalid (may be NULL)
 * are HOST nI x nJ row-major arrays; margins = top, left, bottom, right. */
int qlp_export_frame(qlp_exporter *exporter, int frame, int ia_height, int ia_width,
                     int adaptive_step, const int margins[4], int nI, int nJ, const float *u,
                     const float *v, const uint8_t *invalid);
/* exportDataToFile(PIVResults) :185-238 for one velocity map of whole results
 * (PIVMap): a struct "velocities<H>x<W>" with iaHeight, iaWidth,
 * multiVolumeExport (and, for a multi-volume file, multiVolumeStartOffset =
 * start_offset, the first map's absolute frame offset), margins and u, v of
 * nI x nJ x n_maps singles; frame k of u, v (HOST, n_frames x nI x nJ
 * row-major) goes to map index frame_index[k] (Velocities.getFrameNumber).
 * Values are written as they are (no NaN marking, no swap). */
int qlp_export_results(qlp_exporter *exporter, int ia_height, int ia_width, const int margins[4],
                       int nI, int nJ, int n_maps, int n_frames, const int *frame_index,
                       const float *u, const float *v, int start_offset);
/* Write the file and release the exporter. */
int qlp_export_close(qlp_exporter *exporter);

/* ---- serving: a request loop over frame pairs of mixed sizes ----------------
 *
 * Replaces the reference's pool of per-thread PIV workers over a list of frame
 * pairs (InputFiles.java:172-224 with OpenClGpuManager.java:124-142: one job
 * and device queue per worker thread) by one server per GPU with continuous
 * batching.  Requests are admitted in arrival order as soon as the HBM they
 * need fits the server's budget.  Batched path (FromCenter, Direct inheritance,
 * no warp or the SecondImage micro warp, no validation, 8-bit frames, any fit
 * and arithmetic form): each step of the server runs the next adaptive level
 * of every admitted request, of every frame and IA size, as one launch of a
 * captured graph of persistent kernels (the batch engine of qlp_batch);
 * requests that arrive meanwhile join the following steps.  Other
 * configurations run on plan slots (a context with its own stream each), which
 * overlap on the device.  Frames, twiddles, the level grids and every level's
 * displacement state stay on the device for the request's whole multi-pass
 * run; the last level's field is returned in host memory.  Per request the
 * result equals a multi-pass plan of that request alone. */
#define QLP_SERVER_MAX_SLOTS 16

typedef struct qlp_server_config {
    size_t hbm_budget;  /* bytes of device memory the server may hold */
    int slots;          /* plan-path requests in flight at once (1..QLP_SERVER_MAX_SLOTS) */
    int batch_requests; /* batched-path requests admitted at once (0: 512) */
} qlp_server_config;

typedef struct qlp_request {
    int H, W;                   /* frame size */
    const uint8_t *imgA, *imgB; /* H x W uint8 frames; the caller keeps them until the
                                   request has finished */
    int frames_on_device;       /* 1: imgA/imgB are device pointers, 0: host memory */
    qlp_piv_config cfg;         /* the multi-pass configuration of this request */
} qlp_request;

typedef struct qlp_request_times {
    double submit_ms, admit_ms, done_ms; /* since the server was created */
    double ias;                          /* IAs correlated over all levels */
} qlp_request_times;

typedef struct qlp_server_stats {
    uint64_t completed, failed, pending, evictions;
    size_t hbm_in_use, hbm_peak;
    double ias_done;
    int running;
} qlp_server_stats;

typedef struct qlp_server qlp_server;

int qlp_server_create(int hip_device, const qlp_server_config *cfg, qlp_server **out);
/* Waits for the requests in flight and the queued ones, then frees everything. */
int qlp_server_destroy(qlp_server *server);
/* Queue a request; QLP_ERR_BUDGET if it alone exceeds the budget. */
int qlp_server_submit(qlp_server *server, const qlp_request *req, uint64_t *ticket);
/* Wait up to timeout_ms (-1: no limit) for a request; *done = 1 once finished.
 * Returns the request's own error if it failed. */
int qlp_server_wait(qlp_server *server, uint64_t ticket, int timeout_ms, int *done);
/* The last level's field of a finished request (host memory owned by the
 * server until qlp_server_release): u, v (nI x nJ), its IA size and timings. */
int qlp_server_result(qlp_server *server, uint64_t ticket, const float **u, const float **v,
                      int *nI, int *nJ, int *ia, qlp_request_times *times);
int qlp_server_release(qlp_server *server, uint64_t ticket);
int qlp_server_get_stats(qlp_server *server, qlp_server_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* QLPIV_H */
