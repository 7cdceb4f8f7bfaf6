// Device helpers shared by the kernels of libqlpiv.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qlpiv {

// commons-math3 FastMath.round(float) = (int)floor(x + 0.5f)
// IA count of a pass: from device memory for a persistent launch.
template <typename PP>
__device__ __forceinline__ int pass_count(const PP &p)
{
    return p.n_dev ? *p.n_dev : p.n;
}

__device__ __forceinline__ int fm_round(float x) { return (int)floorf(x + 0.5f); }

// BiLinearImageMicroAndMiniWarpingStrategy.warpPixelOnImage :336-407, SecondImage
// order (source = target + displacement): round-half-up source pixel, clamp to
// the image (which zeroes the fractional offset there), neighbour on the side
// of the offset, bilinear weights evaluated left to right in float, value
// rounded half up.
// I is the type of the offsets into the frame: size_t, or unsigned where the
// caller has bounded H * W below 2^32.
template <typename T, typename I = size_t>
__device__ __forceinline__ float warp_pixel(const T *__restrict__ img, int H, int W, float u,
                                            float v, int i, int j)
{
    float sI = (float)i + u, sJ = (float)j + v;
    int pI = (short)fm_round(sI), pJ = (short)fm_round(sJ);
    if (pI < 0) {
        sI = 0.0f;
        pI = 0;
    } else if (pI >= H) {
        sI = (float)(H - 1);
        pI = H - 1;
    }
    if (pJ < 0) {
        sJ = 0.0f;
        pJ = 0;
    } else if (pJ >= W) {
        sJ = (float)(W - 1);
        pJ = W - 1;
    }
    const float dU = sI - (float)pI, dV = sJ - (float)pJ;
    int oI = dU >= 0.0f ? 1 : -1, oJ = dV >= 0.0f ? 1 : -1;
    if (oI + pI < 0 || oI + pI >= H) oI = 0;
    if (oJ + pJ < 0 || oJ + pJ >= W) oJ = 0;
    const float wU = fabsf(dU), wV = fabsf(dV);
    float tl, tr, bl, br;
    if constexpr (sizeof(I) == sizeof(size_t)) {
        tl = img[(size_t)pI * W + pJ];
        tr = img[(size_t)pI * W + pJ + oJ];
        bl = img[(size_t)(pI + oI) * W + pJ];
        br = img[(size_t)(pI + oI) * W + pJ + oJ];
    } else {
        tl = img[(I)pI * (I)W + (I)pJ];
        tr = img[(I)pI * (I)W + (I)(pJ + oJ)];
        bl = img[(I)(pI + oI) * (I)W + (I)pJ];
        br = img[(I)(pI + oI) * (I)W + (I)(pJ + oJ)];
    }
    const float f = (1.0f - wU) * (1.0f - wV) * tl + (1.0f - wU) * wV * tr +
                    wU * (1.0f - wV) * bl + wU * wV * br;
    return (float)fm_round(f);
}

// Sides of a tile warp (BiLinearImageWarpingModeEnum, warpAndClipImage
// :63-167): the B tile in SecondImage order, the A tile in FirstImage order,
// the A and the B tile in BothImages order.
enum WarpSide { SIDE_B_SECOND = 0, SIDE_A_FIRST = 1, SIDE_A_BOTH = 2, SIDE_B_BOTH = 3 };

struct SideGeom {
    int top, left;     // origin of the warped window
    float offu, offv;  // BothImages offsets
    bool half, first;  // halve the field and add the offsets; sample at target - field
};

// The window and offsets of a side for a tile at (top, left) whose B tile
// carries the displacement (u, v): origins shifted by floor(u) (SecondImage),
// -floor(u) (FirstImage) or +-floor(u/2) (BothImages), the BothImages offsets
// evaluated in double and cast to float as the reference does (:95-96, :152-153).
__device__ __forceinline__ SideGeom side_geometry(int side, int top, int left, float u, float v)
{
    SideGeom g{0, 0, 0.0f, 0.0f, false, false};
    const float hu = u / 2.0f, hv = v / 2.0f;
    if (side == SIDE_A_FIRST) {
        g.top = (short)(top - (int)floorf(u));
        g.left = (short)(left - (int)floorf(v));
        g.first = true;
    } else if (side == SIDE_A_BOTH) {
        g.top = (short)(top - (int)floorf(hu));
        g.left = (short)(left - (int)floorf(hv));
        g.offu = (float)((double)hu - floor((double)hu)) - (float)((double)u - floor((double)u));
        g.offv = (float)((double)hv - floor((double)hv)) - (float)((double)v - floor((double)v));
        g.half = true;
        g.first = true;
    } else if (side == SIDE_B_BOTH) {
        g.top = (short)(top + (int)floorf(hu));
        g.left = (short)(left + (int)floorf(hv));
        g.offu = (float)((double)hu - floor((double)hu));
        g.offv = (float)((double)hv - floor((double)hv));
        g.half = true;
    } else {
        g.top = (short)(top + (int)floorf(u));
        g.left = (short)(left + (int)floorf(v));
    }
    return g;
}

// Pixel (i, j) of a side's window sampled with the field value (u, v) there
// (doBiLinearWarp :292-334).  FirstImage samples at target - field, which is
// target + (-field) exactly in IEEE arithmetic.
template <typename T>
__device__ __forceinline__ float side_sample(const T *__restrict__ img, int H, int W,
                                             const SideGeom &g, float u, float v, int i, int j)
{
    if (g.half) {
        u = u / 2.0f + g.offu;
        v = v / 2.0f + g.offv;
    }
    if (g.first) {
        u = -u;
        v = -v;
    }
    return warp_pixel(img, H, W, u, v, (short)(i + g.top), (short)(j + g.left));
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m)
{
    const int lo = __shfl_xor((int)(uint32_t)v, m, 64);
    const int hi = __shfl_xor((int)(uint32_t)(v >> 32), m, 64);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// First-visit order of cell (i, j) in FindMaximumFromCenter's scan
// (:53-174): the centre (h+1, h+1), its column outwards, its row outwards,
// then the four quadrants ring by ring; h = D / 2.
__device__ __forceinline__ int visit_order(int i, int j, int h)
{
    const int c = h + 1;
    if (i == c && j == c) return 0;
    if (j == c) return (i <= h) ? 1 + 2 * (h - i) : 2 + 2 * (i - h - 1);
    if (i == c) return (j <= h) ? 1 + 2 * h + 2 * (h - j) : 2 + 2 * h + 2 * (j - h - 1);
    const int ii = (i <= h) ? h - i : i - h - 1, qi = (i <= h) ? 0 : 1;
    const int jj = (j <= h) ? h - j : j - h - 1, qj = (j <= h) ? 0 : 1;
    return 1 + 4 * h + 4 * (ii * h + jj) + qi + 2 * qj;
}

// FromCenter ordering key of a visited cell (rows and columns 1..D-1):
// larger value first, then smaller squared distance to (h+1, h+1), then the
// earlier visit.  Cells the reference never selects (value < 0 or NaN) get 0;
// every selectable cell has a non-zero key.  -0 ties with +0 as in Java.
__device__ __forceinline__ uint64_t from_center_key(float v, int i, int j, int h)
{
    if (!(v >= 0.0f)) return 0;
    const uint32_t vb = (v == 0.0f) ? 0u : __float_as_uint(v);
    const int c = h + 1;
    const int d2 = (i - c) * (i - c) + (j - c) * (j - c);
    return ((uint64_t)vb << 32) | ((uint64_t)(0x7fff - d2) << 16) |
           (uint64_t)(0xffff - visit_order(i, j, h));
}

// The same argmax in two phases, for kernels that hold the plane: (1) the
// largest selectable value vmax over the visited cells (fmaxf from -1 ignores
// NaN; vmax < 0 means no selectable cell), (2) among the visited cells whose
// value equals vmax (+0 and -0 alike), the largest tie key below: nearest to
// (h+1, h+1), then the earliest visited.  The winner is the cell with the
// largest from_center_key, but the distance and visit order are computed only
// for the cells that reach the maximum.  The tie key is never 0.
__device__ __forceinline__ uint32_t from_center_tie(int i, int j, int h)
{
    const int c = h + 1;
    const int d2 = (i - c) * (i - c) + (j - c) * (j - c);
    return ((uint32_t)(0x7fff - d2) << 16) | (uint32_t)(0xffff - visit_order(i, j, h));
}

__device__ __forceinline__ float wave_max_float(float v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

}  // namespace qlpiv
