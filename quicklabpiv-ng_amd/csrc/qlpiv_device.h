// Device helpers shared by the kernels of libqlpiv.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qlpiv {

// commons-math3 FastMath.round(float) = (int)floor(x + 0.5f)
__device__ __forceinline__ int fm_round(float x) { return (int)floorf(x + 0.5f); }

// BiLinearImageMicroAndMiniWarpingStrategy.warpPixelOnImage :336-407, SecondImage
// order (source = target + displacement): round-half-up source pixel, clamp to
// the image (which zeroes the fractional offset there), neighbour on the side
// of the offset, bilinear weights evaluated left to right in float, value
// rounded half up.
__device__ __forceinline__ float warp_pixel(const uint8_t *__restrict__ img, int H, int W, float u,
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
    const float tl = img[(size_t)pI * W + pJ];
    const float tr = img[(size_t)pI * W + pJ + oJ];
    const float bl = img[(size_t)(pI + oI) * W + pJ];
    const float br = img[(size_t)(pI + oI) * W + pJ + oJ];
    const float f = (1.0f - wU) * (1.0f - wV) * tl + (1.0f - wU) * wV * tr +
                    wU * (1.0f - wV) * bl + wU * wV * br;
    return (float)fm_round(f);
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

}  // namespace qlpiv
