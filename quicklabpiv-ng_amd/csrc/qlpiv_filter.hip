// Gaussian image pre-filter (images/filters/GaussianFilter2D.java,
// GaussianFilter2DInternal.java, DirectConvolution.java).
//
// The reference filters row by row, then column by column over the row
// result, each line copied into a padded buffer (leading border mirrored,
// trailing border the last K/2 samples repeated in order, written after the
// data) and convolved by DirectConvolution, whose sum for one output runs over
// the kernel taps in order from 0.  One thread per output pixel repeats that
// sum: out[x] = sum_q buf(x + K - 1 - q) * k[q].
#include <math.h>

#include "qlpiv_internal.h"

namespace qlpiv {

constexpr int GF_MAX_K = 31;

struct GfKernel {
    int K;
    float k[GF_MAX_K];
};

// value of the padded line at buffer index b (line of n samples at stride s)
template <typename T>
__device__ __forceinline__ float gf_sample(const T *__restrict__ line, int n, int s, int K, int b)
{
    const int h = K / 2, bs = n + K - 1;
    int x;
    if (b >= bs - h) x = n - 1 - (bs - 1 - b);
    else if (b < h) x = h - 1 - b;
    else x = b - h;
    return (float)line[(size_t)x * s];
}

template <typename T>
__global__ void gf_rows_kernel(int n_img, int H, int W, const T *__restrict__ in, GfKernel gk,
                               float *__restrict__ out)
{
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t plane = (size_t)H * W;
    if (gid >= (size_t)n_img * plane) return;
    const size_t img = gid / plane;
    const int r = (int)((gid - img * plane) / W), c = (int)(gid - img * plane - (size_t)r * W);
    const T *line = in + img * plane + (size_t)r * W;
    float acc = 0.0f;
    for (int q = 0; q < gk.K; q++) acc += gf_sample(line, W, 1, gk.K, c + gk.K - 1 - q) * gk.k[q];
    out[gid] = acc;
}

__global__ void gf_cols_kernel(int n_img, int H, int W, const float *__restrict__ in, GfKernel gk,
                               float *__restrict__ out)
{
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t plane = (size_t)H * W;
    if (gid >= (size_t)n_img * plane) return;
    const size_t img = gid / plane;
    const int r = (int)((gid - img * plane) / W), c = (int)(gid - img * plane - (size_t)r * W);
    const float *line = in + img * plane + c;
    float acc = 0.0f;
    for (int q = 0; q < gk.K; q++) acc += gf_sample(line, H, W, gk.K, r + gk.K - 1 - q) * gk.k[q];
    out[gid] = acc;
}

template <typename T>
__global__ void gf_copy_kernel(size_t n, const T *__restrict__ in, float *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

// GaussianFilter2D.createKernel :36-50
static void gf_make_kernel(float sigma, int K, GfKernel &gk)
{
    double kd[GF_MAX_K];
    const int half = K / 2;
    double sum = 0.0;
    for (int i = 0; i < K; i++) {
        const double x = i - half;
        kd[i] = 1.0 / sqrt(2.0 * M_PI * sigma * sigma) * exp(-x * x / (2.0 * sigma * sigma));
        sum += kd[i];
    }
    gk.K = K;
    for (int i = 0; i < K; i++) gk.k[i] = (float)(kd[i] / sum);
}

}  // namespace qlpiv

using namespace qlpiv;

extern "C" {

int qlp_gaussian_filter_dev(qlp_ctx *h, int n_img, int H, int W, const void *in, int in_is_float,
                            float sigma, int kernel_px, float *tmp, float *out)
try {
    if (!h || !in || !out) return fail(QLP_ERR_ARG, "gaussian_filter: null argument");
    if (n_img < 0 || H < 1 || W < 1) return fail(QLP_ERR_ARG, "gaussian_filter: bad image size");
    if (kernel_px < 1 || kernel_px > GF_MAX_K)
        return fail(QLP_ERR_ARG, "gaussian_filter: kernel size must be 1..31 pixels");
    if (H < kernel_px || W < kernel_px)
        return fail(QLP_ERR_ARG, "gaussian_filter: image smaller than the kernel");
    if (n_img == 0) return QLP_OK;
    Context *ctx = ctx_of(h);
    const size_t total = (size_t)n_img * H * W;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    if (sigma < 1e-5f) {  // GaussianFilter2D.applyFilter :84-86: the input unchanged
        if (in_is_float)
            hipLaunchKernelGGL(gf_copy_kernel<float>, dim3(blocks), dim3(256), 0, ctx->stream,
                               total, (const float *)in, out);
        else
            hipLaunchKernelGGL(gf_copy_kernel<uint8_t>, dim3(blocks), dim3(256), 0, ctx->stream,
                               total, (const uint8_t *)in, out);
        return check_hip(hipGetLastError(), "gaussian filter launch");
    }
    if (!tmp) return fail(QLP_ERR_ARG, "gaussian_filter: null scratch image");
    GfKernel gk;
    gf_make_kernel(sigma, kernel_px, gk);
    if (in_is_float)
        hipLaunchKernelGGL(gf_rows_kernel<float>, dim3(blocks), dim3(256), 0, ctx->stream, n_img, H,
                           W, (const float *)in, gk, tmp);
    else
        hipLaunchKernelGGL(gf_rows_kernel<uint8_t>, dim3(blocks), dim3(256), 0, ctx->stream, n_img,
                           H, W, (const uint8_t *)in, gk, tmp);
    hipLaunchKernelGGL(gf_cols_kernel, dim3(blocks), dim3(256), 0, ctx->stream, n_img, H, W, tmp, gk,
                       out);
    return check_hip(hipGetLastError(), "gaussian filter launch");
} catch (...) {
    return qlpiv::fail_exception();
}

}  // extern "C"
