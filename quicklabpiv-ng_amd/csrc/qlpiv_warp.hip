// Interrogation-area tiles cut from device-resident 8-bit images.
//
// CLIP (images/Image.java clipImageMatrix :386-494 with the MatrixByte maximum,
//   MatrixByte.java:16-39, :278-317): the N x N window at (top, left), zero
//   outside the image; the tile maximum over the in-image pixels, clamped to
//   >= 16.
// WARP_MICRO (images/BiLinearImageMicroAndMiniWarpingStrategy.java, SecondImage
//   order: warpAndClipImage :132-148, createMicroDisplacements :188-201,
//   doBiLinearWarp :292-334, warpPixelOnImage :336-407): the window moves by
//   floor(u), floor(v) and every pixel is resampled bilinearly at the
//   fractional remainder, with round-half-up of the source position and of the
//   value (commons-math3 FastMath.round(float) = floor(x + 0.5f)).  Values stay
//   integers 0..255, so the tile is stored as u8; its maximum is the image
//   maximum 255 (doBiLinearWarp :298).
// Rows of the image are read by consecutive lanes (coalesced); one work-group
// per tile.
#include <algorithm>

#include "qlpiv_device.h"
#include "qlpiv_internal.h"

namespace qlpiv {

__global__ __launch_bounds__(256) void clip_kernel(const uint8_t *__restrict__ img, int H, int W,
                                                  int n, int N, const int *__restrict__ top,
                                                  const int *__restrict__ left,
                                                  uint8_t *__restrict__ tiles,
                                                  float *__restrict__ maxv)
{
    __shared__ int s_max[4];
    const int t = blockIdx.x;
    if (t >= n) return;
    const int y0 = top[t], x0 = left[t];
    int mx = 0;
    for (int idx = threadIdx.x; idx < N * N; idx += blockDim.x) {
        const int i = idx / N, j = idx - (idx / N) * N;
        const int y = y0 + i, x = x0 + j;
        uint8_t v = 0;
        if (y >= 0 && y < H && x >= 0 && x < W) {
            v = img[(size_t)y * W + x];
            mx = max(mx, (int)v);
        }
        tiles[(size_t)t * N * N + idx] = v;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) m = max(m, s_max[w]);
        maxv[t] = (float)(m < 16 ? 16 : m);
    }
}

// One work-group per tile: the Micro warp of one side (WarpSide) of tile t,
// whose B tile carries the displacement (du, dv)[t]; every pixel takes its
// fractional part (createMicroDisplacements :195-196).
__global__ __launch_bounds__(256) void warp_micro_kernel(const uint8_t *__restrict__ img, int H,
                                                        int W, int n, int N,
                                                        const int *__restrict__ top,
                                                        const int *__restrict__ left,
                                                        const float *__restrict__ du,
                                                        const float *__restrict__ dv, int side,
                                                        uint8_t *__restrict__ tiles,
                                                        float *__restrict__ maxv)
{
    const int t = blockIdx.x;
    if (t >= n) return;
    const float u = du[t], v = dv[t];
    const SideGeom g = side_geometry(side, top[t], left[t], u, v);
    const float ru = u - floorf(u), rv = v - floorf(v);
    for (int idx = threadIdx.x; idx < N * N; idx += blockDim.x) {
        const int i = idx / N, j = idx - (idx / N) * N;
        tiles[(size_t)t * N * N + idx] = (uint8_t)side_sample(img, H, W, g, ru, rv, i, j);
    }
    if (threadIdx.x == 0 && maxv) maxv[t] = 255.0f;
}

// Float frames (images/ImageFloat.java: the frames the Gaussian pre-filter
// writes, GaussianFilter2D.applyFilter(IImage, IImage) :116-126).
// CLIP_F32 (ImageFloat.clipImageMatrix :278-384): the N x N window, zero
// outside, at (top, left), or with a displacement at (round(top + u),
// round(left + v)) (NoImageWarpingStrategy :80-81); the tile maximum is the
// MatrixFloat one, over the tile with 0 -> 1 (MatrixFloat.java:45-60,
// Matrix.setMaxValue :21-29; no >= 16 clamp).
__global__ __launch_bounds__(256) void clip_f32_kernel(const float *__restrict__ img, int H, int W,
                                                      int n, int N, const int *__restrict__ top,
                                                      const int *__restrict__ left,
                                                      const float *__restrict__ du,
                                                      const float *__restrict__ dv,
                                                      int clamp16, float *__restrict__ tiles,
                                                      float *__restrict__ maxv)
{
    __shared__ float s_max[4];
    const int t = blockIdx.x;
    if (t >= n) return;
    int y0 = top[t], x0 = left[t];
    if (du) {
        y0 = fm_round((float)y0 + du[t]);
        x0 = fm_round((float)x0 + dv[t]);
    }
    float mx = 0.0f;
    for (int idx = threadIdx.x; idx < N * N; idx += blockDim.x) {
        const int i = idx / N, j = idx - (idx / N) * N;
        const int y = y0 + i, x = x0 + j;
        float v = 0.0f;
        if (y >= 0 && y < H && x >= 0 && x < W) {
            v = img[(size_t)y * W + x];
            if (v > mx) mx = v;
        }
        tiles[(size_t)t * N * N + idx] = v;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const float om = __shfl_xor(mx, o, 64);
        if (om > mx) mx = om;
    }
    if ((threadIdx.x & 63) == 0) s_max[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = 0.0f;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++)
            if (s_max[w] > m) m = s_max[w];
        maxv[t] = clamp16 ? (m < 16.0f ? 16.0f : m) : (m == 0.0f ? 1.0f : m);
    }
}

// 10/12/16-bit frames (ImageInt16) as float frames: every sample converted
// exactly.
__global__ __launch_bounds__(256) void u16_to_f32_kernel(const uint16_t *__restrict__ in,
                                                        float *__restrict__ out, size_t n)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x)
        out[k] = (float)in[k];
}

int launch_u16_to_f32(Context *ctx, const uint16_t *in, float *out, size_t n)
{
    if (n == 0) return QLP_OK;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(u16_to_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, in, out, n);
    return check_hip(hipGetLastError(), "frame conversion launch");
}

// WARP_MICRO_F32: the Micro warp of one side of a float frame's tile, sampled
// as for 8-bit frames (readPixel of ImageFloat :531-533); the rounded values
// can exceed 255, and the tile carries the frame's maximum wmax
// (doBiLinearWarp :298): 65535 for ImageFloat (getMaximumValue :541-543), the
// bit-depth maximum for ImageInt16 (:557-559).  With clamp16 the clip takes
// the MatrixInt16 maximum instead (>= 16, MatrixInt16.java:16-41, :371-405).
__global__ __launch_bounds__(256) void warp_micro_f32_kernel(const float *__restrict__ img, int H,
                                                            int W, int n, int N,
                                                            const int *__restrict__ top,
                                                            const int *__restrict__ left,
                                                            const float *__restrict__ du,
                                                            const float *__restrict__ dv, int side,
                                                            float wmax, float *__restrict__ tiles,
                                                            float *__restrict__ maxv)
{
    const int t = blockIdx.x;
    if (t >= n) return;
    const float u = du[t], v = dv[t];
    const SideGeom g = side_geometry(side, top[t], left[t], u, v);
    const float ru = u - floorf(u), rv = v - floorf(v);
    for (int idx = threadIdx.x; idx < N * N; idx += blockDim.x) {
        const int i = idx / N, j = idx - (idx / N) * N;
        tiles[(size_t)t * N * N + idx] = side_sample(img, H, W, g, ru, rv, i, j);
    }
    if (threadIdx.x == 0 && maxv) maxv[t] = wmax;
}

int launch_clip_f32(Context *ctx, const float *img, int H, int W, int n, int N, const int *top,
                    const int *left, const float *du, const float *dv, float *tiles, float *maxv,
                    int clamp16)
{
    if (n == 0) return QLP_OK;
    if ((du == nullptr) != (dv == nullptr)) return fail(QLP_ERR_ARG, "clip: u and v go together");
    hipLaunchKernelGGL(clip_f32_kernel, dim3(n), dim3(256), 0, ctx->stream, img, H, W, n, N, top, left,
                       du, dv, clamp16, tiles, maxv);
    return check_hip(hipGetLastError(), "float clip kernel launch");
}

int launch_warp_micro_f32(Context *ctx, const float *img, int H, int W, int n, int N, const int *top,
                          const int *left, const float *u, const float *v, float *tiles, float *maxv,
                          int side, float wmax)
{
    if (n == 0) return QLP_OK;
    hipLaunchKernelGGL(warp_micro_f32_kernel, dim3(n), dim3(256), 0, ctx->stream, img, H, W, n, N,
                       top, left, u, v, side, wmax, tiles, maxv);
    return check_hip(hipGetLastError(), "float warp kernel launch");
}

int launch_clip(Context *ctx, const uint8_t *img, int H, int W, int n, int N, const int *top,
                const int *left, uint8_t *tiles, float *maxv)
{
    if (n == 0) return QLP_OK;
    hipLaunchKernelGGL(clip_kernel, dim3(n), dim3(256), 0, ctx->stream, img, H, W, n, N, top, left,
                       tiles, maxv);
    return check_hip(hipGetLastError(), "clip kernel launch");
}

int launch_warp_micro(Context *ctx, const uint8_t *img, int H, int W, int n, int N, const int *top,
                      const int *left, const float *u, const float *v, uint8_t *tiles, float *maxv,
                      int side)
{
    if (n == 0) return QLP_OK;
    hipLaunchKernelGGL(warp_micro_kernel, dim3(n), dim3(256), 0, ctx->stream, img, H, W, n, N, top,
                       left, u, v, side, tiles, maxv);
    return check_hip(hipGetLastError(), "warp kernel launch");
}

}  // namespace qlpiv
