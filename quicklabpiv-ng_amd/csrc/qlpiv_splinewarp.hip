// Mini and Full-image bilinear warps with spline displacement fields
// (images/BiLinearImageMicroAndMiniWarpingStrategy.java:203-290 for Mini,
// images/BiLinearImageWarpingStrategy.java:38-183 for Full), in the three warp
// orders (SecondImage: B; FirstImage: A; BothImages: A and B with half the
// field, WarpSide in qlpiv_device.h).
//
// Both sample frame B at a per-pixel displacement taken from a displacement
// field over tile centres (iareas/BiCubicSplineInterpolatorWithBiLinearBackup
// .java): commons-math3's BicubicSplineInterpolator when both directions hold at
// least 3 knots (slopes from natural cubic splines along each knot row and
// column, the cross slope from the divided difference over the neighbouring
// knots, one bicubic Hermite patch per cell), else bilinear interpolation;
// queries are clamped to the knot range in float.  The field is evaluated in
// FP64, as the reference does.
//   Mini: per tile, the field of its 7 x 7 tile neighbourhood (cut to the grid),
//         whose values are the neighbours' displacements minus floor of the
//         tile's own, evaluated from the B reference window (shifted by
//         floor(u), floor(v)) in every order and applied at the side's window.
//         One work-group per tile: the neighbourhood's slopes in LDS, then
//         every pixel.  The warped tile carries the maximum 255.
//   Full: the field of the whole grid at every pixel of the frame: one thread
//         per knot row or column for the slopes (a global workspace), then one
//         thread per pixel for the field and the bilinear sample of each warped
//         frame; the pass cuts the tiles from the warped frames (maximum 255).
#include <math.h>

#include <vector>

#include "qlpiv_device.h"
#include "qlpiv_internal.h"

namespace qlpiv {

// Slopes at the knots of the natural cubic spline through (x[k], y[k*ys]),
// k < n, n >= 3 (SplineInterpolator.interpolate; the natural spline algorithm of
// Burden & Faires, Algorithm 3.4), written to s[k*ss].  ws: 4 (n + 1) doubles.
__device__ void natural_slopes(int n, const double *x, const double *y, int ys, double *s, int ss,
                               double *ws)
{
    const int m = n - 1;
    double *h = ws, *mu = ws + (n + 1), *z = ws + 2 * (n + 1), *c = ws + 3 * (n + 1);
    for (int i = 0; i < m; i++) h[i] = x[i + 1] - x[i];
    mu[0] = 0.0;
    z[0] = 0.0;
    for (int i = 1; i < m; i++) {
        const double g = 2.0 * (x[i + 1] - x[i - 1]) - h[i - 1] * mu[i - 1];
        mu[i] = h[i] / g;
        z[i] = (3.0 * (y[(i + 1) * ys] * h[i - 1] - y[i * ys] * (x[i + 1] - x[i - 1]) +
                       y[(i - 1) * ys] * h[i]) /
                    (h[i - 1] * h[i]) -
                h[i - 1] * z[i - 1]) /
               g;
    }
    z[m] = 0.0;
    c[m] = 0.0;
    double b = 0.0, d = 0.0;
    for (int j = m - 1; j >= 0; j--) {
        c[j] = z[j] - mu[j] * c[j + 1];
        b = (y[(j + 1) * ys] - y[j * ys]) / h[j] - h[j] * (c[j + 1] + 2.0 * c[j]) / 3.0;
        d = (c[j + 1] - c[j]) / (3.0 * h[j]);
        s[j * ss] = b;
        if (j == m - 1) s[m * ss] = (3.0 * d * h[m - 1] + 2.0 * c[m - 1]) * h[m - 1] + b;
    }
}

// commons BicubicSplineInterpolatingFunction.searchIndex over a clamped query:
// the first k >= 1 with c <= v[k], minus one (binary search)
__device__ __forceinline__ int cell_of(double c, const double *v, int n)
{
    int lo = 1, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (c <= v[mid])
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo - 1;
}

// The bicubic Hermite patch of cell (i, j) at (y, x): values f, slopes fy
// (along the rows' coordinate y), fx (along x) and the cross slope at the four
// corners; f* indexed [row * nx + col].
__device__ __forceinline__ double patch(const double *ys, const double *xs, int nx, int i, int j,
                                        double y, double x, const double (&f)[4],
                                        const double (&fy)[4], const double (&fx)[4],
                                        const double (&fyx)[4])
{
    const double dy = ys[i + 1] - ys[i], dx = xs[j + 1] - xs[j];
    const double s = (y - ys[i]) / dy, t = (x - xs[j]) / dx;
    const double hs[2] = {(1.0 + 2.0 * s) * (1.0 - s) * (1.0 - s), s * s * (3.0 - 2.0 * s)};
    const double gs[2] = {s * (1.0 - s) * (1.0 - s), s * s * (s - 1.0)};
    const double ht[2] = {(1.0 + 2.0 * t) * (1.0 - t) * (1.0 - t), t * t * (3.0 - 2.0 * t)};
    const double gt[2] = {t * (1.0 - t) * (1.0 - t), t * t * (t - 1.0)};
    double r = 0.0;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const int k = 2 * a + b;
            r += f[k] * hs[a] * ht[b] + fy[k] * dy * gs[a] * ht[b] + fx[k] * dx * hs[a] * gt[b] +
                 fyx[k] * dy * dx * gs[a] * gt[b];
        }
    (void)nx;
    return r;
}

// BiLinearInterpolatingFunction.value(y, x) :40-70
__device__ __forceinline__ double bilinear(const double *ys, int ny, const double *xs, int nx,
                                           const double *v, double y, double x)
{
    int I = 1, J = 1;
    while (I < ny - 1 && !(y <= ys[I])) I++;
    while (J < nx - 1 && !(x <= xs[J])) J++;
    const double wY = 1.0 - ((ys[I] - y) / (ys[I] - ys[I - 1]));
    const double wX = 1.0 - ((xs[J] - x) / (xs[J] - xs[J - 1]));
    return wY * (wX * v[I * nx + J] + (1.0 - wX) * v[I * nx + J - 1]) +
           (1.0 - wY) * (wX * v[(I - 1) * nx + J] + (1.0 - wX) * v[(I - 1) * nx + J - 1]);
}

// The field value of one component at (y, x) (already clamped), from the knot
// values f, their slopes fy (along y) and fx (along x); the cross slope is
// formed from f (BicubicSplineInterpolator: next/previous knot, clamped).
__device__ double field_at(const double *ys, int ny, const double *xs, int nx, const double *f,
                           const double *fy, const double *fx, bool spline, double y, double x)
{
    if (!spline) return bilinear(ys, ny, xs, nx, f, y, x);
    const int i = cell_of(y, ys, ny), j = cell_of(x, xs, nx);
    double cf[4], cfy[4], cfx[4], cfyx[4];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const int r = i + a, c = j + b, k = 2 * a + b;
            cf[k] = f[r * nx + c];
            cfy[k] = fy[r * nx + c];
            cfx[k] = fx[r * nx + c];
            const int nr = r + 1 < ny ? r + 1 : r, pr = r - 1 >= 0 ? r - 1 : 0;
            const int nc = c + 1 < nx ? c + 1 : c, pc = c - 1 >= 0 ? c - 1 : 0;
            cfyx[k] = (f[nr * nx + nc] - f[nr * nx + pc] - f[pr * nx + nc] + f[pr * nx + pc]) /
                      ((ys[nr] - ys[pr]) * (xs[nc] - xs[pc]));
        }
    return patch(ys, xs, nx, i, j, y, x, cf, cfy, cfx, cfyx);
}

// ---- Mini warp: one 64-thread work-group per tile -----------------------------

constexpr int MINI_R = 7;  // the tile neighbourhood (createMiniDisplacements :207)

// T: pixel type of the frames (uint8_t, or float for ImageFloat and converted
// ImageInt16 frames); O: tile type (uint8_t, or float); wmax: the maximum a
// warped tile carries (image.getMaximumValue(), doBiLinearWarp :298).
template <typename T, typename O>
__global__ __launch_bounds__(64) void warp_mini_kernel(int n, int nI, int nJ, int N, int H, int W,
                                                       const T *__restrict__ imgs,
                                                       const int *__restrict__ top,
                                                       const int *__restrict__ left,
                                                       const float *__restrict__ u,
                                                       const float *__restrict__ v, int side,
                                                       float wmax, O *__restrict__ tiles,
                                                       float *__restrict__ maxv)
{
    __shared__ double s_y[MINI_R], s_x[MINI_R];
    __shared__ double s_f[2][MINI_R * MINI_R], s_fy[2][MINI_R * MINI_R], s_fx[2][MINI_R * MINI_R];
    const int pair = blockIdx.x / n, t = blockIdx.x - pair * n;
    const int ti = t / nJ, tj = t - ti * nJ;
    const float *__restrict__ pu = u + (size_t)pair * n;
    const float *__restrict__ pv = v + (size_t)pair * n;
    const T *__restrict__ img = imgs + (size_t)pair * H * W;
    const float u0 = pu[t], v0 = pv[t];
    const double fu0 = floor((double)u0), fv0 = floor((double)v0);
    // the field is evaluated from the B reference window in every order
    const int tTop = (short)(top[t] + (int)floorf(u0));
    const int tLeft = (short)(left[t] + (int)floorf(v0));
    const SideGeom g = side_geometry(side, top[t], left[t], u0, v0);
    const int h = MINI_R / 2;
    const int i0 = ti - h < 0 ? 0 : ti - h, i1 = ti + h > nI - 1 ? nI - 1 : ti + h;
    const int j0 = tj - h < 0 ? 0 : tj - h, j1 = tj + h > nJ - 1 ? nJ - 1 : tj + h;
    const int ny = i1 - i0 + 1, nx = j1 - j0 + 1;
    const int k = threadIdx.x;
    // knots (tile centres as float, then double) and values
    if (k < ny) s_y[k] = (double)((float)top[(i0 + k) * nJ + j0] + (float)N / 2.0f - 0.5f);
    if (k < nx) s_x[k] = (double)((float)left[i0 * nJ + j0 + k] + (float)N / 2.0f - 0.5f);
    if (k < ny * nx) {
        const int gi = (i0 + k / nx) * nJ + j0 + k % nx;
        s_f[0][k] = (double)pu[gi] - fu0;
        s_f[1][k] = (double)pv[gi] - fv0;
    }
    __syncthreads();
    const bool spline = ny >= 3 && nx >= 3;
    if (spline && k < 2 * (nx + ny)) {
        double ws[4 * (MINI_R + 1)];
        const int c = k / (nx + ny), line = k - c * (nx + ny);
        if (line < nx)  // slopes along y of column 'line'
            natural_slopes(ny, s_y, &s_f[c][line], nx, &s_fy[c][line], nx, ws);
        else  // slopes along x of row 'line - nx'
            natural_slopes(nx, s_x, &s_f[c][(line - nx) * nx], 1, &s_fx[c][(line - nx) * nx], 1, ws);
    }
    __syncthreads();
    const float ymin = (float)s_y[0], ymax = (float)s_y[ny - 1];
    const float xmin = (float)s_x[0], xmax = (float)s_x[nx - 1];
    O *__restrict__ out = tiles + ((size_t)pair * n + t) * N * N;
    for (int p = k; p < N * N; p += 64) {
        const int i = p / N, j = p - (p / N) * N;
        float y = (float)(tTop + i), x = (float)(tLeft + j);
        x = x < xmin ? xmin : (x > xmax ? xmax : x);
        y = y < ymin ? ymin : (y > ymax ? ymax : y);
        const float du = (float)field_at(s_y, ny, s_x, nx, s_f[0], s_fy[0], s_fx[0], spline, y, x);
        const float dv = (float)field_at(s_y, ny, s_x, nx, s_f[1], s_fy[1], s_fx[1], spline, y, x);
        out[p] = (O)side_sample(img, H, W, g, du, dv, i, j);
    }
    if (k == 0 && maxv) maxv[(size_t)pair * n + t] = wmax;
}

int launch_warp_mini(Context *ctx, int pairs, const uint8_t *imgs, int H, int W, int nI, int nJ,
                     int N, const int *top, const int *left, const float *u, const float *v,
                     uint8_t *tiles, float *maxv, int side)
{
    const int n = nI * nJ;
    if (pairs <= 0 || n <= 0) return QLP_OK;
    // the bilinear backup needs two knots in each direction (with one the
    // reference's BiLinearInterpolatingFunction reads past its arrays)
    if (nI < 2 || nJ < 2)
        return fail(QLP_ERR_UNSUPPORTED, "mini warp: the grid needs 2 tiles in each direction");
    hipLaunchKernelGGL((warp_mini_kernel<uint8_t, uint8_t>), dim3(pairs * n), dim3(64), 0,
                       ctx->stream, n, nI, nJ, N, H, W, imgs, top, left, u, v, side, 255.0f, tiles,
                       maxv);
    return check_hip(hipGetLastError(), "mini warp launch");
}

int launch_warp_mini_f32(Context *ctx, int pairs, const float *imgs, int H, int W, int nI, int nJ,
                         int N, const int *top, const int *left, const float *u, const float *v,
                         float *tiles, float *maxv, int side, float wmax)
{
    const int n = nI * nJ;
    if (pairs <= 0 || n <= 0) return QLP_OK;
    if (nI < 2 || nJ < 2)
        return fail(QLP_ERR_UNSUPPORTED, "mini warp: the grid needs 2 tiles in each direction");
    hipLaunchKernelGGL((warp_mini_kernel<float, float>), dim3(pairs * n), dim3(64), 0, ctx->stream,
                       n, nI, nJ, N, H, W, imgs, top, left, u, v, side, wmax, tiles, maxv);
    return check_hip(hipGetLastError(), "mini warp launch");
}

// ---- Full-image warp -----------------------------------------------------------

// knots ys (nI) and xs (nJ) on the device; f, fy, fx: pairs x 2 x n doubles;
// ws: per line 4 (max(nI, nJ) + 1) doubles
__global__ void full_values_kernel(int pairs, int n, const float *__restrict__ u,
                                   const float *__restrict__ v, double *__restrict__ f)
{
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (size_t)pairs * n) return;
    const size_t pair = g / n, t = g - pair * n;
    f[(pair * 2 + 0) * n + t] = (double)u[g];
    f[(pair * 2 + 1) * n + t] = (double)v[g];
}

__global__ void full_slopes_kernel(int pairs, int nI, int nJ, const double *__restrict__ ys,
                                   const double *__restrict__ xs, const double *__restrict__ f,
                                   double *__restrict__ fy, double *__restrict__ fx,
                                   double *__restrict__ ws)
{
    const int lines = nI + nJ, n = nI * nJ;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= pairs * 2 * lines) return;
    const int pc = g / lines, line = g - pc * lines;  // pc = pair * 2 + component
    const int L = (nI > nJ ? nI : nJ) + 1;
    double *w = ws + (size_t)g * 4 * L;
    const double *ff = f + (size_t)pc * n;
    if (line < nJ)  // slopes along y of column 'line'
        natural_slopes(nI, ys, ff + line, nJ, fy + (size_t)pc * n + line, nJ, w);
    else  // slopes along x of row 'line - nJ'
        natural_slopes(nJ, xs, ff + (size_t)(line - nJ) * nJ, 1,
                       fx + (size_t)pc * n + (size_t)(line - nJ) * nJ, 1, w);
}

// One thread per pixel: the field there, then the sample of frame imgs for
// side 'side' (doBiLinearWarp :143-183: BothImages halves the field).
template <typename T, typename O>
__global__ void full_warp_kernel(int pairs, int H, int W, int nI, int nJ,
                                 const double *__restrict__ ys, const double *__restrict__ xs,
                                 const double *__restrict__ f, const double *__restrict__ fy,
                                 const double *__restrict__ fx, const T *__restrict__ imgs,
                                 int side, O *__restrict__ out)
{
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t plane = (size_t)H * W;
    if (g >= (size_t)pairs * plane) return;
    const size_t pair = g / plane, pix = g - pair * plane;
    const int i = (int)(pix / W), j = (int)(pix - (size_t)i * W);
    const int n = nI * nJ;
    const bool spline = nI >= 3 && nJ >= 3;
    const float ymin = (float)ys[0], ymax = (float)ys[nI - 1];
    const float xmin = (float)xs[0], xmax = (float)xs[nJ - 1];
    float y = (float)i, x = (float)j;
    x = x < xmin ? xmin : (x > xmax ? xmax : x);
    y = y < ymin ? ymin : (y > ymax ? ymax : y);
    const size_t b0 = pair * 2 * n, b1 = b0 + n;
    float du = (float)field_at(ys, nI, xs, nJ, f + b0, fy + b0, fx + b0, spline, y, x);
    float dv = (float)field_at(ys, nI, xs, nJ, f + b1, fy + b1, fx + b1, spline, y, x);
    if (side == SIDE_A_BOTH || side == SIDE_B_BOTH) {
        du /= 2.0f;
        dv /= 2.0f;
    }
    if (side == SIDE_A_FIRST || side == SIDE_A_BOTH) {  // target - field
        du = -du;
        dv = -dv;
    }
    out[g] = (O)warp_pixel(imgs + pair * plane, H, W, du, dv, (short)i, (short)j);
}

// Maximum over each pair's frame (strict >, from 0; 0 -> 1): the MatrixFloat
// maximum of clipImageMatrix(0, 0, H, W) (MatrixFloat.java:45-60), which the
// Full warp's tiles inherit (copyMatrixRegion :437-455).  One work-group per pair.
__global__ __launch_bounds__(1024) void frame_max_kernel(const float *__restrict__ imgs, size_t plane,
                                                         float *__restrict__ out)
{
    __shared__ float s_m[16];
    const float *img = imgs + (size_t)blockIdx.x * plane;
    float m = 0.0f;
    for (size_t k = threadIdx.x; k < plane; k += blockDim.x)
        if (img[k] > m) m = img[k];
    for (int o = 32; o >= 1; o >>= 1) {
        const float x = __shfl_xor(m, o, 64);
        if (x > m) m = x;
    }
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = 0.0f;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++)
            if (s_m[w] > r) r = s_m[w];
        out[blockIdx.x] = r == 0.0f ? 1.0f : r;
    }
}

// dst[pair * n + t] = per_pair[pair]
__global__ void spread_kernel(const float *__restrict__ per_pair, int n, int pairs,
                              float *__restrict__ dst)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n * pairs) dst[g] = per_pair[g / n];
}

int launch_frame_max_spread(Context *ctx, const float *imgs, int pairs, int H, int W, int n,
                            float *per_pair, float *dst)
{
    if (pairs <= 0) return QLP_OK;
    hipLaunchKernelGGL(frame_max_kernel, dim3(pairs), dim3(1024), 0, ctx->stream, imgs,
                       (size_t)H * W, per_pair);
    hipLaunchKernelGGL(spread_kernel, dim3((n * pairs + 255) / 256), dim3(256), 0, ctx->stream,
                       per_pair, n, pairs, dst);
    return check_hip(hipGetLastError(), "frame maximum launch");
}

size_t full_warp_work_bytes(int pairs, int nI, int nJ)
{
    const size_t n = (size_t)nI * nJ, L = (size_t)(nI > nJ ? nI : nJ) + 1;
    return ((size_t)nI + nJ) * 8 + (size_t)pairs * 2 * n * 8 * 3 +
           (size_t)pairs * 2 * (nI + nJ) * 4 * L * 8;
}

template <typename T>
static int launch_full_warp_t(Context *ctx, int pairs, const T *const *imgs, int H, int W, int nI,
                              int nJ, int tH, int tW, const int *top_h, const int *left_h,
                              const float *u, const float *v, void *work, T *const *outs,
                              const int *sides, int n_sides)
{
    if (pairs <= 0) return QLP_OK;
    if (nI < 2 || nJ < 2)
        return fail(QLP_ERR_UNSUPPORTED, "full-image warp: the grid needs 2 tiles in each direction");
    const size_t n = (size_t)nI * nJ, L = (size_t)(nI > nJ ? nI : nJ) + 1;
    double *ys = (double *)work, *xs = ys + nI;
    double *f = xs + nJ, *fy = f + (size_t)pairs * 2 * n, *fx = fy + (size_t)pairs * 2 * n;
    double *ws = fx + (size_t)pairs * 2 * n;
    (void)L;
    // knots: tile centres as float (createTileDisplacementInterpolator), then double
    std::vector<double> k(nI + nJ);
    for (int i = 0; i < nI; i++) k[i] = (double)((float)top_h[i * nJ] + (float)tH / 2.0f - 0.5f);
    for (int j = 0; j < nJ; j++) k[nI + j] = (double)((float)left_h[j] + (float)tW / 2.0f - 0.5f);
    if (int rc = check_hip(hipMemcpyAsync(ys, k.data(), k.size() * 8, hipMemcpyHostToDevice,
                                          ctx->stream), "full warp knots"))
        return rc;
    // the pageable source must stay valid until the copy has been staged
    if (int rc = check_hip(hipStreamSynchronize(ctx->stream), "full warp knots")) return rc;
    const size_t pn = (size_t)pairs * n;
    hipLaunchKernelGGL(full_values_kernel, dim3((unsigned)((pn + 255) / 256)), dim3(256), 0,
                       ctx->stream, pairs, (int)n, u, v, f);
    if (nI >= 3 && nJ >= 3) {
        const int lines = pairs * 2 * (nI + nJ);
        hipLaunchKernelGGL(full_slopes_kernel, dim3((lines + 63) / 64), dim3(64), 0, ctx->stream,
                           pairs, nI, nJ, ys, xs, f, fy, fx, ws);
    }
    const size_t total = (size_t)pairs * H * W;
    for (int k = 0; k < n_sides; k++)
        hipLaunchKernelGGL((full_warp_kernel<T, T>), dim3((unsigned)((total + 255) / 256)), dim3(256),
                           0, ctx->stream, pairs, H, W, nI, nJ, ys, xs, f, fy, fx, imgs[k], sides[k],
                           outs[k]);
    return check_hip(hipGetLastError(), "full warp launch");
}

int launch_full_warp(Context *ctx, int pairs, const uint8_t *const *imgs, int H, int W, int nI,
                     int nJ, int tH, int tW, const int *top_h, const int *left_h, const float *u,
                     const float *v, void *work, uint8_t *const *outs, const int *sides,
                     int n_sides)
{
    return launch_full_warp_t<uint8_t>(ctx, pairs, imgs, H, W, nI, nJ, tH, tW, top_h, left_h, u, v,
                                       work, outs, sides, n_sides);
}

int launch_full_warp_f32(Context *ctx, int pairs, const float *const *imgs, int H, int W, int nI,
                         int nJ, int tH, int tW, const int *top_h, const int *left_h, const float *u,
                         const float *v, void *work, float *const *outs, const int *sides,
                         int n_sides)
{
    return launch_full_warp_t<float>(ctx, pairs, imgs, H, W, nI, nJ, tH, tW, top_h, left_h, u, v,
                                     work, outs, sides, n_sides);
}

int warp_clip_tiles(Context *ctx, int mode, int order, const uint8_t *imgsA, const uint8_t *imgsB,
                    int H, int W, int pairs, int nI, int nJ, int N, const int *top, const int *left,
                    const float *u, const float *v, uint8_t *tA, float *mA, uint8_t *tB, float *mB)
{
    const int n = nI * nJ;
    const size_t tN = (size_t)N * N, plane = (size_t)H * W;
    auto clip = [&](const uint8_t *imgs, uint8_t *t, float *m) {
        int rc = QLP_OK;
        for (int k = 0; k < pairs && !rc; k++)
            rc = launch_clip(ctx, imgs + k * plane, H, W, n, N, top, left, t + k * n * tN, m + k * n);
        return rc;
    };
    auto warp = [&](const uint8_t *imgs, int side, uint8_t *t, float *m) {
        if (mode == QLP_WARP_MINI)
            return launch_warp_mini(ctx, pairs, imgs, H, W, nI, nJ, N, top, left, u, v, t, m, side);
        int rc = QLP_OK;
        for (int k = 0; k < pairs && !rc; k++)
            rc = launch_warp_micro(ctx, imgs + k * plane, H, W, n, N, top, left, u + k * n,
                                   v + k * n, t + k * n * tN, m + k * n, side);
        return rc;
    };
    int rc = order == QLP_WARP_ORDER_SECOND
                 ? clip(imgsA, tA, mA)
                 : warp(imgsA, order == QLP_WARP_ORDER_FIRST ? SIDE_A_FIRST : SIDE_A_BOTH, tA, mA);
    if (rc) return rc;
    return order == QLP_WARP_ORDER_FIRST
               ? clip(imgsB, tB, mB)
               : warp(imgsB, order == QLP_WARP_ORDER_SECOND ? SIDE_B_SECOND : SIDE_B_BOTH, tB, mB);
}

}  // namespace qlpiv

using namespace qlpiv;

extern "C" {

int qlp_warp_clip_tiles_dev(qlp_ctx *h, int warp_mode, int order, const uint8_t *imgsA,
                            const uint8_t *imgsB, int H, int W, int pairs, int nI, int nJ, int N,
                            const int *top, const int *left, const float *u, const float *v,
                            uint8_t *tilesA, float *maxA, uint8_t *tilesB, float *maxB)
try {
    if (!h || !imgsA || !imgsB || !top || !left || !u || !v || !tilesA || !maxA || !tilesB || !maxB)
        return fail(QLP_ERR_ARG, "warp_clip_tiles: null argument");
    if (warp_mode != QLP_WARP_MICRO && warp_mode != QLP_WARP_MINI)
        return fail(QLP_ERR_ARG, "warp_clip_tiles: warp mode must be Micro or Mini");
    if (order < QLP_WARP_ORDER_SECOND || order > QLP_WARP_ORDER_BOTH)
        return fail(QLP_ERR_ARG, "warp_clip_tiles: unknown warp order");
    if (H < 1 || W < 1 || pairs < 0 || nI < 1 || nJ < 1 || N < 1)
        return fail(QLP_ERR_ARG, "warp_clip_tiles: bad sizes");
    return warp_clip_tiles(ctx_of(h), warp_mode, order, imgsA, imgsB, H, W, pairs, nI, nJ, N, top,
                           left, u, v, tilesA, maxA, tilesB, maxB);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_warp_images_dev(qlp_ctx *h, int order, const uint8_t *imgsA, const uint8_t *imgsB, int H,
                        int W, int pairs, int nI, int nJ, int N, const int *top, const int *left,
                        const float *u, const float *v, uint8_t *outA, uint8_t *outB)
try {
    if (!h || !top || !left || !u || !v) return fail(QLP_ERR_ARG, "warp_images: null argument");
    if (order < QLP_WARP_ORDER_SECOND || order > QLP_WARP_ORDER_BOTH)
        return fail(QLP_ERR_ARG, "warp_images: unknown warp order");
    if (H < 1 || W < 1 || pairs < 0 || nI < 1 || nJ < 1 || N < 1)
        return fail(QLP_ERR_ARG, "warp_images: bad sizes");
    const uint8_t *src[2];
    uint8_t *dst[2];
    int sides[2], ns = 0;
    if (order != QLP_WARP_ORDER_SECOND) {
        if (!imgsA || !outA) return fail(QLP_ERR_ARG, "warp_images: null first frames");
        src[ns] = imgsA;
        dst[ns] = outA;
        sides[ns++] = order == QLP_WARP_ORDER_FIRST ? SIDE_A_FIRST : SIDE_A_BOTH;
    }
    if (order != QLP_WARP_ORDER_FIRST) {
        if (!imgsB || !outB) return fail(QLP_ERR_ARG, "warp_images: null second frames");
        src[ns] = imgsB;
        dst[ns] = outB;
        sides[ns++] = order == QLP_WARP_ORDER_SECOND ? SIDE_B_SECOND : SIDE_B_BOTH;
    }
    Context *ctx = ctx_of(h);
    if (int rc = ensure_iwork(ctx, full_warp_work_bytes(pairs, nI, nJ))) return rc;
    return launch_full_warp(ctx, pairs, src, H, W, nI, nJ, N, N, top, left, u, v, ctx->iwork, dst,
                            sides, ns);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_warp_mini_tiles_dev(qlp_ctx *h, const uint8_t *img, int H, int W, int pairs, int nI,
                            int nJ, int N, const int *top, const int *left, const float *u,
                            const float *v, uint8_t *tiles, float *maxv)
try {
    if (!h || !img || !top || !left || !u || !v || !tiles)
        return fail(QLP_ERR_ARG, "warp_mini: null argument");
    if (H < 1 || W < 1 || pairs < 0 || nI < 1 || nJ < 1 || N < 1)
        return fail(QLP_ERR_ARG, "warp_mini: bad sizes");
    return launch_warp_mini(ctx_of(h), pairs, img, H, W, nI, nJ, N, top, left, u, v, tiles, maxv);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_warp_image_dev(qlp_ctx *h, const uint8_t *img, int H, int W, int pairs, int nI, int nJ,
                       int N, const int *top, const int *left, const float *u, const float *v,
                       uint8_t *out)
try {
    if (!h || !img || !top || !left || !u || !v || !out)
        return fail(QLP_ERR_ARG, "warp_image: null argument");
    if (H < 1 || W < 1 || pairs < 0 || nI < 1 || nJ < 1 || N < 1)
        return fail(QLP_ERR_ARG, "warp_image: bad sizes");
    Context *ctx = ctx_of(h);
    if (int rc = ensure_iwork(ctx, full_warp_work_bytes(pairs, nI, nJ))) return rc;
    const int side = SIDE_B_SECOND;
    return launch_full_warp(ctx, pairs, &img, H, W, nI, nJ, N, N, top, left, u, v, ctx->iwork, &out,
                            &side, 1);
} catch (...) {
    return qlpiv::fail_exception();
}

}  // extern "C"
