// C ABI of libqlpiv: contexts, errors, host-buffer wrappers.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "qlpiv_internal.h"

namespace qlpiv {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

// The status of an exception that reached a C-ABI entry point (every entry
// point's body is a function-try-block ending here): no C++ exception may
// cross into a JNI or ctypes caller.
int fail_exception()
{
    try {
        throw;
    } catch (const std::bad_alloc &) {
        return fail(QLP_ERR_COMPUTE, "host memory allocation failed");
    } catch (const std::exception &e) {
        return fail(QLP_ERR_COMPUTE, std::string("internal error: ") + e.what());
    } catch (...) {
        return fail(QLP_ERR_COMPUTE, "internal error: unknown exception");
    }
}

int check_hip(hipError_t e, const char *what)
{
    if (e == hipSuccess) return QLP_OK;
    return fail(QLP_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// ws[n] = (float)sin(2*pi*n/M) in double (SimpleFFT.initEulerTable :323-330,
// CrossCorrelationFFTTemplate.initEulerTable :268-277), arranged as the
// (c, s) pairs of the three butterfly forms.
static void build_tables(int M, std::vector<float2> out[3])
{
    std::vector<float> ws(M);
    for (int k = 0; k < M; k++) ws[k] = (float)sin(2.0 * M_PI * (double)(float)k / (double)(float)M);
    for (int t = 0; t < 3; t++) out[t].resize(M / 2);
    for (int k = 0; k < M / 2; k++) {
        out[TW_FWD][k] = make_float2(ws[k + M / 4], -ws[k]);
        const int ic = (M / 4 - k >= 0) ? (M / 4 - k) : (M + (M / 4 - k));
        out[TW_INV_CPU][k] = make_float2(ws[ic], ws[k]);
        out[TW_INV_GPU][k] = make_float2(ws[k + M / 4], ws[k]);
    }
}

// Grow a context-owned device buffer to at least 'bytes' (contents are not
// kept), charging it to the context's HBM budget.
static int grow_buffer(Context *ctx, void *&buf, size_t &have, size_t bytes, const char *what)
{
    if (bytes <= have) return QLP_OK;
    if (ctx->hbm_budget && ctx->hbm_in_use - have + bytes > ctx->hbm_budget)
        return fail(QLP_ERR_BUDGET, "request exceeds the context's HBM budget");
    if (buf) {
        (void)hipFree(buf);
        ctx->hbm_in_use -= have;
        buf = nullptr;
        have = 0;
    }
    int rc = check_hip(hipMalloc(&buf, bytes), what);
    if (rc) return rc;
    have = bytes;
    ctx->hbm_in_use += bytes;
    return QLP_OK;
}

static int ensure_scratch(Context *ctx, size_t bytes)
{
    return grow_buffer(ctx, ctx->scratch, ctx->scratch_bytes, bytes, "hipMalloc scratch");
}

int ensure_iwork(Context *ctx, size_t bytes)
{
    return grow_buffer(ctx, ctx->iwork, ctx->iwork_bytes, bytes, "hipMalloc inheritance workspace");
}

int ensure_bigwork(Context *ctx, size_t bytes)
{
    return grow_buffer(ctx, ctx->bigwork, ctx->bigwork_bytes, bytes, "hipMalloc IA-128 workspace");
}

int ensure_mpwork(Context *ctx, size_t bytes)
{
    return grow_buffer(ctx, ctx->mpwork, ctx->mpwork_bytes, bytes, "hipMalloc MultiPeaks workspace");
}

bool mpwork_fits(const Context *ctx, size_t bytes)
{
    if (bytes <= ctx->mpwork_bytes) return true;
    return !ctx->hbm_budget || ctx->hbm_in_use - ctx->mpwork_bytes + bytes <= ctx->hbm_budget;
}

static int sync_ctx(Context *ctx) { return check_hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize"); }

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace qlpiv

using namespace qlpiv;

struct qlp_ctx {
    Context c;
};

Context *qlpiv::ctx_of(qlp_ctx *h) { return &h->c; }

extern "C" {

const char *qlp_last_error(void) { return g_last_error.c_str(); }

const char *qlp_version(void) { return "qlpiv-mi355x 0.1 (gfx950)"; }

int qlp_ctx_create(int hip_device, size_t hbm_budget_bytes, qlp_ctx **out)
try {
    if (!out) return fail(QLP_ERR_ARG, "qlp_ctx_create: null output pointer");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(QLP_ERR_DEVICE, "no HIP device available");
    if (hip_device < 0 || hip_device >= ndev)
        return fail(QLP_ERR_ARG, "qlp_ctx_create: device index out of range");
    int rc = check_hip(hipSetDevice(hip_device), "hipSetDevice");
    if (rc) return rc;
    qlp_ctx *h = new qlp_ctx();
    Context *ctx = &h->c;
    ctx->device = hip_device;
    ctx->hbm_budget = hbm_budget_bytes;
    const char *form = getenv("QLPIV_XCORR_FORM");
    ctx->xcorr_form = (form && strcmp(form, "lds") == 0) ? FORM_LDS : FORM_LANE;
    rc = check_hip(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
    if (rc) {
        delete h;
        return rc;
    }
    for (int li = 0; li < TW_SIZES; li++) {
        const int M = 32 << li;
        std::vector<float2> tabs[3];
        build_tables(M, tabs);
        for (int t = 0; t < 3; t++) {
            void *p = nullptr;
            rc = check_hip(hipMalloc(&p, sizeof(float2) * (M / 2)), "hipMalloc tables");
            if (!rc)
                rc = check_hip(hipMemcpy(p, tabs[t].data(), sizeof(float2) * (M / 2),
                                         hipMemcpyHostToDevice),
                               "upload tables");
            if (rc) {
                qlp_ctx_destroy(h);
                return rc;
            }
            ctx->tables.tw[li][t] = (float2 *)p;
            ctx->hbm_in_use += sizeof(float2) * (M / 2);
        }
    }
    // the allocation-free entry point's scratch address on this device (current)
    rc = init_noalloc_scratch(hip_device);
    if (rc) {
        qlp_ctx_destroy(h);
        return rc;
    }
    *out = h;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_ctx_destroy(qlp_ctx *h)
try {
    if (!h) return QLP_OK;
    Context *ctx = &h->c;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (int li = 0; li < TW_SIZES; li++)
        for (int t = 0; t < 3; t++)
            if (ctx->tables.tw[li][t]) (void)hipFree(ctx->tables.tw[li][t]);
    if (ctx->scratch) (void)hipFree(ctx->scratch);
    if (ctx->vwork) (void)hipFree(ctx->vwork);
    if (ctx->iwork) (void)hipFree(ctx->iwork);
    if (ctx->bigwork) (void)hipFree(ctx->bigwork);
    if (ctx->mpwork) (void)hipFree(ctx->mpwork);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 2; b++)
            if (ctx->mfma_tab[a][b]) (void)hipFree(ctx->mfma_tab[a][b]);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete h;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_ctx_synchronize(qlp_ctx *h)
try {
    if (!h) return fail(QLP_ERR_ARG, "null context");
    return sync_ctx(&h->c);
} catch (...) {
    return qlpiv::fail_exception();
}

void *qlp_ctx_stream(qlp_ctx *h) { return h ? (void *)h->c.stream : nullptr; }

static int xcorr_args(int n, int N, int mode)
{
    if (n < 0) return fail(QLP_ERR_ARG, "xcorr: negative IA count");
    if (N != 16 && N != 32 && N != 64 && N != 128)
        return fail(QLP_ERR_ARG, "xcorr: IA size must be 16, 32, 64 or 128 pixels");
    if (mode != QLP_NORM_PER_TILE && mode != QLP_NORM_A_MAX)
        return fail(QLP_ERR_ARG, "xcorr: unknown normalisation mode");
    return QLP_OK;
}

int qlp_xcorr_batch_dev(qlp_ctx *h, int n, int N, const uint8_t *tA, const uint8_t *tB,
                        const float *maxA, const float *maxB, int mode, float *planes)
try {
    if (!h) return fail(QLP_ERR_ARG, "null context");
    int rc = xcorr_args(n, N, mode);
    if (rc) return rc;
    if (n > 0 && (!tA || !tB || !maxA || !planes || (mode == QLP_NORM_PER_TILE && !maxB)))
        return fail(QLP_ERR_ARG, "xcorr: null buffer");
    return launch_xcorr(&h->c, n, N, tA, tB, maxA, maxB, mode, planes);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_xcorr_batch_noalloc_dev(qlp_ctx *h, int n, int N, const uint8_t *tA, const uint8_t *tB,
                                const float *maxA, const float *maxB, int mode, float *planes)
try {
    if (!h) return fail(QLP_ERR_ARG, "null context");
    int rc = xcorr_args(n, N, mode);
    if (rc) return rc;
    if (n > 0 && (!tA || !tB || !maxA || !planes || (mode == QLP_NORM_PER_TILE && !maxB)))
        return fail(QLP_ERR_ARG, "xcorr: null buffer");
    return launch_xcorr_noalloc(&h->c, n, N, tA, tB, maxA, maxB, mode, planes);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_xcorr_f32_dev(qlp_ctx *h, int n, int N, const float *tA, const float *tB, const float *maxA,
                      const float *maxB, int mode, float *planes)
try {
    if (!h) return fail(QLP_ERR_ARG, "null context");
    int rc = xcorr_args(n, N, mode);
    if (rc || n == 0) return rc;
    if (!tA || !tB || !maxA || !planes || (mode == QLP_NORM_PER_TILE && !maxB))
        return fail(QLP_ERR_ARG, "xcorr: null buffer");
    PassParams p;
    p.n = n;
    p.N = N;
    p.mode = mode;
    p.tAf = tA;
    p.tBf = tB;
    p.maxA = maxA;
    p.maxB = maxB ? maxB : maxA;
    p.planes = planes;
    return launch_pass(&h->c, p);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_xcorr_f32(qlp_ctx *h, int n, int N, const float *tA, const float *tB, const float *maxA,
                  const float *maxB, int mode, float *planes)
try {
    if (!h) return fail(QLP_ERR_ARG, "null context");
    int rc = xcorr_args(n, N, mode);
    if (rc || n == 0) return rc;
    if (!tA || !tB || !maxA || !planes || (mode == QLP_NORM_PER_TILE && !maxB))
        return fail(QLP_ERR_ARG, "xcorr: null buffer");
    Context *ctx = &h->c;
    const int D = 2 * N - 1;
    const size_t tbytes = sizeof(float) * (size_t)n * N * N;
    const size_t bt = align256(tbytes), bm = align256(sizeof(float) * n),
                 bp = align256(sizeof(float) * (size_t)n * D * D);
    rc = ensure_scratch(ctx, 2 * bt + 2 * bm + bp);
    if (rc) return rc;
    char *base = (char *)ctx->scratch;
    float *dA = (float *)base, *dB = (float *)(base + bt);
    float *dmA = (float *)(base + 2 * bt), *dmB = (float *)(base + 2 * bt + bm);
    float *dP = (float *)(base + 2 * bt + 2 * bm);
    hipStream_t s = ctx->stream;
    if ((rc = check_hip(hipMemcpyAsync(dA, tA, tbytes, hipMemcpyHostToDevice, s), "copy A")) ||
        (rc = check_hip(hipMemcpyAsync(dB, tB, tbytes, hipMemcpyHostToDevice, s), "copy B")) ||
        (rc = check_hip(hipMemcpyAsync(dmA, maxA, sizeof(float) * n, hipMemcpyHostToDevice, s), "copy maxA")))
        return rc;
    if (maxB && (rc = check_hip(hipMemcpyAsync(dmB, maxB, sizeof(float) * n, hipMemcpyHostToDevice, s), "copy maxB")))
        return rc;
    rc = qlp_xcorr_f32_dev(h, n, N, dA, dB, dmA, maxB ? dmB : dmA, mode, dP);
    if (rc) return rc;
    rc = check_hip(hipMemcpyAsync(planes, dP, sizeof(float) * (size_t)n * D * D, hipMemcpyDeviceToHost, s),
                   "copy planes");
    if (rc) return rc;
    return sync_ctx(ctx);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_xcorr_form_dev(qlp_ctx *h, int form, int n, int N, const uint8_t *tA, const uint8_t *tB,
                       const float *maxA, const float *maxB, int mode, float *planes)
try {
    if (form == QLP_XCORR_EXACT) return qlp_xcorr_batch_dev(h, n, N, tA, tB, maxA, maxB, mode, planes);
    if (!h) return fail(QLP_ERR_ARG, "null context");
    int rc = xcorr_args(n, N, mode);
    if (rc) return rc;
    if (form < QLP_XCORR_MFMA_F16 || form > QLP_XCORR_MFMA_BF16_SINGLE)
        return fail(QLP_ERR_ARG, "xcorr: unknown arithmetic form");
    if (N != 16 && N != 32 && N != 64)
        return fail(QLP_ERR_UNSUPPORTED, "xcorr: the MFMA forms take IA sizes 16, 32 and 64");
    if (n > 0 && (!tA || !tB || !maxA || !planes || (mode == QLP_NORM_PER_TILE && !maxB)))
        return fail(QLP_ERR_ARG, "xcorr: null buffer");
    PassParams p;
    p.n = n;
    p.N = N;
    p.mode = mode;
    p.tA = tA;
    p.tB = tB;
    p.maxA = maxA;
    p.maxB = mode == QLP_NORM_PER_TILE ? maxB : maxA;
    p.planes = planes;
    return launch_pass_mfma(&h->c, p, form, false, false);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_xcorr_batch(qlp_ctx *h, int n, int N, const uint8_t *tA, const uint8_t *tB,
                    const float *maxA, const float *maxB, int mode, float *planes)
try {
    if (!h) return fail(QLP_ERR_ARG, "null context");
    int rc = xcorr_args(n, N, mode);
    if (rc || n == 0) return rc;
    if (!tA || !tB || !maxA || !planes || (mode == QLP_NORM_PER_TILE && !maxB))
        return fail(QLP_ERR_ARG, "xcorr: null buffer");
    Context *ctx = &h->c;
    const int D = 2 * N - 1;
    const size_t bt = align256((size_t)n * N * N), bm = align256(sizeof(float) * n),
                 bp = align256(sizeof(float) * (size_t)n * D * D);
    rc = ensure_scratch(ctx, 2 * bt + 2 * bm + bp);
    if (rc) return rc;
    char *base = (char *)ctx->scratch;
    uint8_t *dA = (uint8_t *)base, *dB = (uint8_t *)(base + bt);
    float *dmA = (float *)(base + 2 * bt), *dmB = (float *)(base + 2 * bt + bm);
    float *dP = (float *)(base + 2 * bt + 2 * bm);
    hipStream_t s = ctx->stream;
    if ((rc = check_hip(hipMemcpyAsync(dA, tA, (size_t)n * N * N, hipMemcpyHostToDevice, s), "copy A")) ||
        (rc = check_hip(hipMemcpyAsync(dB, tB, (size_t)n * N * N, hipMemcpyHostToDevice, s), "copy B")) ||
        (rc = check_hip(hipMemcpyAsync(dmA, maxA, sizeof(float) * n, hipMemcpyHostToDevice, s), "copy maxA")))
        return rc;
    if (maxB && (rc = check_hip(hipMemcpyAsync(dmB, maxB, sizeof(float) * n, hipMemcpyHostToDevice, s), "copy maxB")))
        return rc;
    rc = launch_xcorr(ctx, n, N, dA, dB, dmA, maxB ? dmB : dmA, mode, dP);
    if (rc) return rc;
    rc = check_hip(hipMemcpyAsync(planes, dP, sizeof(float) * (size_t)n * D * D, hipMemcpyDeviceToHost, s),
                   "copy planes");
    if (rc) return rc;
    return sync_ctx(ctx);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_find_peaks_dev(qlp_ctx *h, int kind, int n, int D, const float *planes, int n_peaks,
                       int kernel_width, float *peaks, float *min_floor)
try {
    if (!h) return fail(QLP_ERR_ARG, "null context");
    if (n < 0) return fail(QLP_ERR_ARG, "find_peaks: negative count");
    if ((kind == QLP_PEAK_FROM_CENTER || kind == QLP_PEAK_SIMPLE) && n_peaks != 1)
        return fail(QLP_ERR_ARG, "find_peaks: FromCenter and Simple find exactly one peak");
    return launch_find_peaks(&h->c, kind, n, D, planes, n_peaks, kernel_width, peaks, min_floor);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_find_peaks(qlp_ctx *h, int kind, int n, int D, const float *planes, int n_peaks,
                   int kernel_width, float *peaks, float *min_floor)
try {
    if (!h) return fail(QLP_ERR_ARG, "null context");
    if (n < 0 || n_peaks < 1) return fail(QLP_ERR_ARG, "find_peaks: bad counts");
    if (n == 0) return QLP_OK;
    Context *ctx = &h->c;
    const size_t bp = align256(sizeof(float) * (size_t)n * D * D),
                 bk = align256(sizeof(float) * 3 * (size_t)n * n_peaks), bf = align256(sizeof(float) * n);
    int rc = ensure_scratch(ctx, bp + bk + bf);
    if (rc) return rc;
    char *base = (char *)ctx->scratch;
    float *dP = (float *)base, *dK = (float *)(base + bp), *dF = (float *)(base + bp + bk);
    hipStream_t s = ctx->stream;
    if ((rc = check_hip(hipMemcpyAsync(dP, planes, sizeof(float) * (size_t)n * D * D, hipMemcpyHostToDevice, s), "copy planes")))
        return rc;
    rc = qlp_find_peaks_dev(h, kind, n, D, dP, n_peaks, kernel_width, dK, dF);
    if (rc) return rc;
    if ((rc = check_hip(hipMemcpyAsync(peaks, dK, sizeof(float) * 3 * (size_t)n * n_peaks, hipMemcpyDeviceToHost, s), "copy peaks")) ||
        (rc = check_hip(hipMemcpyAsync(min_floor, dF, sizeof(float) * n, hipMemcpyDeviceToHost, s), "copy floor")))
        return rc;
    return sync_ctx(ctx);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_subpixel_dev(qlp_ctx *h, int kind, int pixels, int iterations, int n, int D, int n_peaks,
                     const float *planes, const float *min_floor, const float *peaks_in,
                     float *peaks_out)
try {
    if (!h) return fail(QLP_ERR_ARG, "null context");
    if (n < 0 || n_peaks < 1) return fail(QLP_ERR_ARG, "subpixel: bad counts");
    if (kind < 0) return fail(QLP_ERR_ARG, "subpixel: unknown interpolator kind");
    return launch_subpixel(&h->c, kind, pixels, iterations, n, D, n_peaks, planes, min_floor,
                           peaks_in, peaks_out);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_subpixel(qlp_ctx *h, int kind, int pixels, int iterations, int n, int D, int n_peaks,
                 const float *planes, const float *min_floor, const float *peaks_in,
                 float *peaks_out)
try {
    if (!h) return fail(QLP_ERR_ARG, "null context");
    if (n < 0 || n_peaks < 1) return fail(QLP_ERR_ARG, "subpixel: bad counts");
    if (kind < 0) return fail(QLP_ERR_ARG, "subpixel: unknown interpolator kind");
    if (n == 0) return QLP_OK;
    Context *ctx = &h->c;
    const size_t bp = align256(sizeof(float) * (size_t)n * D * D), bf = align256(sizeof(float) * n),
                 bi = align256(sizeof(float) * 3 * (size_t)n * n_peaks),
                 bo = align256(sizeof(float) * 4 * (size_t)n * n_peaks);
    int rc = ensure_scratch(ctx, bp + bf + bi + bo);
    if (rc) return rc;
    char *base = (char *)ctx->scratch;
    float *dP = (float *)base, *dF = (float *)(base + bp), *dI = (float *)(base + bp + bf),
          *dO = (float *)(base + bp + bf + bi);
    hipStream_t s = ctx->stream;
    if ((rc = check_hip(hipMemcpyAsync(dP, planes, sizeof(float) * (size_t)n * D * D, hipMemcpyHostToDevice, s), "copy planes")) ||
        (rc = check_hip(hipMemcpyAsync(dF, min_floor, sizeof(float) * n, hipMemcpyHostToDevice, s), "copy floor")) ||
        (rc = check_hip(hipMemcpyAsync(dI, peaks_in, sizeof(float) * 3 * (size_t)n * n_peaks, hipMemcpyHostToDevice, s), "copy peaks")))
        return rc;
    rc = launch_subpixel(ctx, kind, pixels, iterations, n, D, n_peaks, dP, dF, dI, dO);
    if (rc) return rc;
    rc = check_hip(hipMemcpyAsync(peaks_out, dO, sizeof(float) * 4 * (size_t)n * n_peaks, hipMemcpyDeviceToHost, s), "copy out");
    if (rc) return rc;
    return sync_ctx(ctx);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_clip_tiles_dev(qlp_ctx *h, const uint8_t *img, int H, int W, int n, int N, const int *top,
                       const int *left, uint8_t *tiles, float *max_out)
try {
    if (!h) return fail(QLP_ERR_ARG, "null context");
    if (n < 0 || N < 1 || H < 1 || W < 1) return fail(QLP_ERR_ARG, "clip: bad sizes");
    return launch_clip(&h->c, img, H, W, n, N, top, left, tiles, max_out);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_warp_micro_tiles_dev(qlp_ctx *h, const uint8_t *img, int H, int W, int n, int N,
                             const int *top, const int *left, const float *u, const float *v,
                             uint8_t *tiles, float *max_out)
try {
    if (!h) return fail(QLP_ERR_ARG, "null context");
    if (n < 0 || N < 1 || H < 1 || W < 1) return fail(QLP_ERR_ARG, "warp: bad sizes");
    return launch_warp_micro(&h->c, img, H, W, n, N, top, left, u, v, tiles, max_out);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_clip_tiles_f32_dev(qlp_ctx *h, const float *img, int H, int W, int n, int N, const int *top,
                           const int *left, const float *u, const float *v, float *tiles,
                           float *max_out)
try {
    if (!h) return fail(QLP_ERR_ARG, "null context");
    if (n < 0 || N < 1 || H < 1 || W < 1) return fail(QLP_ERR_ARG, "clip: bad sizes");
    if (n > 0 && (!img || !top || !left || !tiles || !max_out)) return fail(QLP_ERR_ARG, "clip: null buffer");
    return launch_clip_f32(&h->c, img, H, W, n, N, top, left, u, v, tiles, max_out);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_warp_micro_tiles_f32_dev(qlp_ctx *h, const float *img, int H, int W, int n, int N,
                                 const int *top, const int *left, const float *u, const float *v,
                                 int side, float *tiles, float *max_out)
try {
    if (!h) return fail(QLP_ERR_ARG, "null context");
    if (n < 0 || N < 1 || H < 1 || W < 1) return fail(QLP_ERR_ARG, "warp: bad sizes");
    if (side < QLP_SIDE_B_SECOND || side > QLP_SIDE_B_BOTH) return fail(QLP_ERR_ARG, "warp: unknown side");
    if (n > 0 && (!img || !top || !left || !u || !v || !tiles)) return fail(QLP_ERR_ARG, "warp: null buffer");
    return launch_warp_micro_f32(&h->c, img, H, W, n, N, top, left, u, v, tiles, max_out, side);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_xcorr_images_form_dev(qlp_ctx *h, int form, int n, int N, int mode, const uint8_t *imgA,
                              const uint8_t *imgB, int H, int W, const int *top, const int *left,
                              const float *du, const float *dv, int warp_mode, float *planes,
                              float *peaks, float *min_floor, float *nbr, int nbr_half)
try {
    if (!h) return fail(QLP_ERR_ARG, "null context");
    if (form < QLP_XCORR_EXACT || form > QLP_XCORR_MFMA_BF16_SINGLE)
        return fail(QLP_ERR_ARG, "xcorr: unknown correlation form");
    if (form != QLP_XCORR_EXACT && N != 16 && N != 32 && N != 64)
        return fail(QLP_ERR_UNSUPPORTED, "xcorr: the MFMA forms take IA sizes 16, 32 and 64");
    int rc = xcorr_args(n, N, mode);
    if (rc) return rc;
    if (n > 0 && (!imgA || !imgB || !top || !left)) return fail(QLP_ERR_ARG, "xcorr_images: null buffer");
    if (H < 1 || W < 1) return fail(QLP_ERR_ARG, "xcorr_images: bad image size");
    if ((du == nullptr) != (dv == nullptr)) return fail(QLP_ERR_ARG, "xcorr_images: du and dv go together");
    if (warp_mode != QLP_WARP_NONE && warp_mode != QLP_WARP_MICRO)
        return fail(QLP_ERR_ARG, "xcorr_images: unknown warp mode");
    PassParams p;
    p.n = n;
    p.N = N;
    p.mode = mode;
    p.imgA = imgA;
    p.imgB = imgB;
    p.H = H;
    p.W = W;
    p.top = top;
    p.left = left;
    p.du = du;
    p.dv = dv;
    p.bmode = warp_mode == QLP_WARP_MICRO ? B_WARP_MICRO : B_CLIP_SHIFT;
    p.planes = planes;
    p.peaks = peaks;
    p.min_floor = min_floor;
    p.nbr = nbr;
    p.nbr_half = nbr_half;
    p.arith = form;
    return launch_pass(&h->c, p);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_xcorr_images_dev(qlp_ctx *h, int n, int N, int mode, const uint8_t *imgA,
                         const uint8_t *imgB, int H, int W, const int *top, const int *left,
                         const float *du, const float *dv, int warp_mode, float *planes,
                         float *peaks, float *min_floor, float *nbr, int nbr_half)
try {
    return qlp_xcorr_images_form_dev(h, QLP_XCORR_EXACT, n, N, mode, imgA, imgB, H, W, top, left,
                                     du, dv, warp_mode, planes, peaks, min_floor, nbr, nbr_half);
} catch (...) {
    return qlpiv::fail_exception();
}

}  // extern "C"
