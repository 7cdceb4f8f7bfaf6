// FromCenter epilogue shared by the correlation kernels that keep a plane in
// registers or LDS (FindMaximumFromCenter.findMaximum :47-183 as an argmax of
// from_center_key).
#pragma once

#include "qlpiv_device.h"
#include "qlpiv_internal.h"

namespace qlpiv {

// FromCenter candidate of one plane cell (i, j) with linear index cell: row and
// column 0 are never visited; the running minimum covers the visited cells.
__device__ __forceinline__ void from_center_visit(float v, int i, int j, int h, int cell,
                                                  uint64_t &best, int &bestCell, float &mn)
{
    if (i == 0 || j == 0) return;
    if (v < mn) mn = v;
    const uint64_t key = from_center_key(v, i, j, h);
    if (key > best) {
        best = key;
        bestCell = cell;
    }
}

// Block reduction of the per-thread FromCenter candidates, then thread 0
// writes the peak record (i, j, value), the minimum and the neighbourhood
// (column through the peak, then the row) of IA ia.  Plane value (i, j) is
// read at src[i*rs + j*cs].  The keys are unique per cell, so the
// result does not depend on which thread visited which cell.
template <int NT, int D>
__device__ __forceinline__ void peak_epilogue(const PassParams &p, int ia, uint64_t best,
                                              int bestCell, float mn, const float *src, int rs,
                                              int cs, uint64_t *s_key, int *s_cell, float *s_min,
                                              int tid = -1)
{
    const int tx = tid < 0 ? (int)threadIdx.x : tid;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const uint64_t ob = shfl_xor_u64(best, o);
        const int oc = __shfl_xor(bestCell, o, 64);
        const float om = __shfl_xor(mn, o, 64);
        if (ob > best) {
            best = ob;
            bestCell = oc;
        }
        if (om < mn) mn = om;
    }
    if ((tx & 63) == 0) {
        s_key[tx >> 6] = best;
        s_cell[tx >> 6] = bestCell;
        s_min[tx >> 6] = mn;
    }
    __syncthreads();
    if (tx == 0) {
        for (int w = 1; w < NT / 64; w++) {
            if (s_key[w] > best) {
                best = s_key[w];
                bestCell = s_cell[w];
            }
            if (s_min[w] < mn) mn = s_min[w];
        }
        // key 0: no visited value >= 0, the reference keeps (0, 0, 0)
        int pi = 0, pj = 0;
        float pv = 0.0f;
        if (best != 0) {
            pi = bestCell / D;
            pj = bestCell % D;
            pv = src[pi * rs + pj * cs];
        }
        if (p.peaks) {
            p.peaks[(size_t)ia * 3 + 0] = (float)pi;
            p.peaks[(size_t)ia * 3 + 1] = (float)pj;
            p.peaks[(size_t)ia * 3 + 2] = pv;
        }
        if (p.min_floor) p.min_floor[ia] = mn;
        if (p.nbr) {
            const int L = 2 * p.nbr_half + 1;
            float *nb = p.nbr + (size_t)ia * 2 * L;
            for (int d = -p.nbr_half; d <= p.nbr_half; d++) {
                const int ri = pi + d, cj = pj + d;
                nb[d + p.nbr_half] = (ri >= 0 && ri < D) ? src[ri * rs + pj * cs] : 0.0f;
                nb[L + d + p.nbr_half] = (cj >= 0 && cj < D) ? src[pi * rs + cj * cs] : 0.0f;
            }
        }
    }
}

}  // namespace qlpiv
