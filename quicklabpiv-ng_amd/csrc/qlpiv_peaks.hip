// Integer peak search over correlation planes.
//
// FROM_CENTER (maximum/FindMaximumFromCenter.java:47-183): the reference scans
// from the cell (D/2+1, D/2+1) outwards -- its column, its row, then the four
// quadrants -- keeping a candidate that is replaced by a strictly larger value
// or by an equal value strictly closer to that cell.  Row 0 and column 0 are
// never visited and the running maximum starts at 0 at cell (0,0).  The result
// is therefore the visited cell that maximises the key
//     (value, -distance^2, -first-visit-order)
// over cells with value >= 0, or (0,0) with value 0 when no such cell exists.
// One wavefront per plane folds that 64-bit key with cross-lane shuffles.
//
// MULTI (maximum/FindMaximumMultiPeaks.java:42-162): block maxima of
// kernelWidth^2 blocks in the reference's four-quadrant visit order feed a
// bounded ascending list with neighbour suppression
// (util/SimpleFixedLengthFloatLinkedList.addInAscendOrder :98-237).  The block
// maxima are found by all lanes in parallel; lane 0 replays the insertions in
// visit order, including the reference's carry-over of the previous block's
// position when a block holds no value above 0.
#include <float.h>

#include "qlpiv_device.h"
#include "qlpiv_internal.h"

namespace qlpiv {

__global__ __launch_bounds__(256) void peak_from_center_kernel(int n, int D,
                                                               const float *__restrict__ planes,
                                                               float *__restrict__ peaks,
                                                               float *__restrict__ min_floor)
{
    const int lane = threadIdx.x & 63;
    const int ia = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (ia >= n) return;
    const float *m = planes + (size_t)ia * D * D;
    const int h = D / 2;
    uint64_t best = 0;
    int bestCell = 0;
    float mn = FLT_MAX;
    for (int idx = lane; idx < D * D; idx += 64) {
        const int i = idx / D, j = idx - (idx / D) * D;
        if (i == 0 || j == 0) continue;
        const float v = m[idx];
        if (v < mn) mn = v;
        const uint64_t key = from_center_key(v, i, j, h);
        if (key > best) {
            best = key;
            bestCell = idx;
        }
    }
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
    if (lane == 0) {
        // every valid cell has a non-zero key (0x7fff - d2 > 0), so key 0 means
        // no visited value >= 0: the reference then keeps its initial (0, 0, 0)
        float pi = 0.0f, pj = 0.0f, pv = 0.0f;
        if (best != 0) {
            pi = (float)(bestCell / D);
            pj = (float)(bestCell % D);
            pv = m[bestCell];
        }
        peaks[(size_t)ia * 3 + 0] = pi;
        peaks[(size_t)ia * 3 + 1] = pj;
        peaks[(size_t)ia * 3 + 2] = pv;
        min_floor[ia] = mn;
    }
}

// ---------------------------------------------------------------------------
// MultiPeaks
// ---------------------------------------------------------------------------

constexpr int MP_MAX_PEAKS = 16;
constexpr int MP_CHUNK = 64;

struct MpList {
    int size;
    float v[MP_MAX_PEAKS + 1];
    short i[MP_MAX_PEAKS + 1], j[MP_MAX_PEAKS + 1];
};

// addInAscendOrder :98-237 on a sorted array: the same accept/reject and
// delete decisions as the linked list, in the same order.
__device__ void mp_insert(MpList &L, int P, int kd, float value, short pi, short pj)
{
    if (L.size == 0) {
        L.v[0] = value;
        L.i[0] = pi;
        L.j[0] = pj;
        L.size = 1;
        return;
    }
    int pos = 0;
    uint32_t delmask = 0;
    while (pos < L.size && L.v[pos] < value) {
        const int di = pi - L.i[pos], dj = pj - L.j[pos];
        if ((float)(di * di + dj * dj) < (float)kd && value > L.v[pos]) delmask |= 1u << pos;
        pos++;
    }
    if (L.size == P && pos == 0) return;
    for (int e = pos; e < L.size; e++) {
        const int di = pi - L.i[e], dj = pj - L.j[e];
        if ((float)(di * di + dj * dj) < (float)kd && value <= L.v[e]) return;
    }
    // compact away the deleted entries before the insertion point
    int w = 0;
    for (int e = 0; e < L.size; e++) {
        if (e < pos && (delmask >> e) & 1u) continue;
        L.v[w] = L.v[e];
        L.i[w] = L.i[e];
        L.j[w] = L.j[e];
        w++;
    }
    const int ndel = L.size - w;
    L.size = w;
    const int ins = pos - ndel;
    if (ins == 0 && L.size == P) return;
    for (int e = L.size; e > ins; e--) {
        L.v[e] = L.v[e - 1];
        L.i[e] = L.i[e - 1];
        L.j[e] = L.j[e - 1];
    }
    L.v[ins] = value;
    L.i[ins] = pi;
    L.j[ins] = pj;
    L.size++;
    if (ins > 0 && L.size > P) {  // remove() drops the smallest (first) entry
        for (int e = 0; e + 1 < L.size; e++) {
            L.v[e] = L.v[e + 1];
            L.i[e] = L.i[e + 1];
            L.j[e] = L.j[e + 1];
        }
        L.size--;
    }
}

__global__ __launch_bounds__(64) void peak_multi_kernel(int n, int D, int P, int k,
                                                        const float *__restrict__ planes,
                                                        float *__restrict__ peaks,
                                                        float *__restrict__ min_floor,
                                                        int *__restrict__ count)
{
    __shared__ float s_v[MP_CHUNK];
    __shared__ short s_i[MP_CHUNK], s_j[MP_CHUNK];
    __shared__ int s_upd[MP_CHUNK];
    const int lane = threadIdx.x;
    const int ia = blockIdx.x;
    if (ia >= n) return;
    const float *m = planes + (size_t)ia * D * D;

    int iSlices = D / k;
    if (iSlices * k != D) iSlices += 1;
    const int iTop = iSlices - iSlices / 2;
    int jSlices = D / k;
    if (jSlices * k != D) jSlices += 1;
    const int jLeft = jSlices - jSlices / 2;
    const int nentries = iTop * jLeft * 4;

    // global minimum (every cell belongs to exactly one visited block)
    float mn = FLT_MAX;
    for (int idx = lane; idx < D * D; idx += 64) {
        const float v = m[idx];
        if (v < mn) mn = v;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const float om = __shfl_xor(mn, o, 64);
        if (om < mn) mn = om;
    }

    MpList L;
    L.size = 0;
    short staleI[4] = {0, 0, 0, 0}, staleJ[4] = {0, 0, 0, 0};
    for (int base = 0; base < nentries; base += MP_CHUNK) {
        const int e = base + lane;
        if (e < nentries) {
            const int blk = e >> 2, q = e & 3;
            const int si = blk / jLeft, sj = blk - (blk / jLeft) * jLeft;
            const int top = si * k + ((q & 2) ? iTop * k : 0);
            const int left = sj * k + ((q & 1) ? jLeft * k : 0);
            float bv = 0.0f;
            int bi = 0, bj = 0, upd = 0;
            for (int oi = 0; oi < k; oi++) {
                const int r = top + oi;
                if (r >= D) break;
                for (int oj = 0; oj < k; oj++) {
                    const int cc = left + oj;
                    if (cc >= D) break;
                    const float v = m[r * D + cc];
                    if (v > bv) {
                        bv = v;
                        bi = r;
                        bj = cc;
                        upd = 1;
                    }
                }
            }
            s_v[lane] = bv;
            s_i[lane] = (short)bi;
            s_j[lane] = (short)bj;
            s_upd[lane] = upd;
        }
        __syncthreads();
        if (lane == 0) {
            const int cnt = min(MP_CHUNK, nentries - base);
            for (int x = 0; x < cnt; x++) {
                const int q = (base + x) & 3;
                if (s_upd[x]) {
                    staleI[q] = s_i[x];
                    staleJ[q] = s_j[x];
                }
                mp_insert(L, P, (k - 1) * k, s_upd[x] ? s_v[x] : 0.0f, staleI[q], staleJ[q]);
            }
        }
        __syncthreads();
    }
    if (lane == 0) {
        float *out = peaks + (size_t)ia * P * 3;
        for (int p = 0; p < P; p++) out[3 * p] = out[3 * p + 1] = out[3 * p + 2] = 0.0f;
        // descending: peak 0 is the largest
        for (int e = 0; e < L.size; e++) {
            const int idx = L.size - 1 - e;
            out[3 * idx + 0] = (float)L.i[e];
            out[3 * idx + 1] = (float)L.j[e];
            out[3 * idx + 2] = L.v[e];
        }
        min_floor[ia] = mn;
        // MaxCrossResult.getTotalPeaks :109-115: the secondary arrays exist
        // only once a peak index >= 1 was stored
        if (count) count[ia] = L.size > 1 ? L.size : 1;
    }
}

int launch_find_peaks(Context *ctx, int kind, int n, int D, const float *planes, int n_peaks,
                      int kernel_width, float *peaks, float *min_floor, int *count)
{
    if (n == 0) return QLP_OK;
    if (D < 3 || D > 255) return fail(QLP_ERR_ARG, "find_peaks: plane size out of range 3..255");
    if (kind == QLP_PEAK_FROM_CENTER) {
        hipLaunchKernelGGL(peak_from_center_kernel, dim3((n + 3) / 4), dim3(256), 0, ctx->stream, n,
                           D, planes, peaks, min_floor);
    } else if (kind == QLP_PEAK_MULTI) {
        if (n_peaks < 1 || n_peaks > MP_MAX_PEAKS)
            return fail(QLP_ERR_ARG, "find_peaks: number of peaks must be 1..16");
        if (kernel_width < 1 || kernel_width > D)
            return fail(QLP_ERR_ARG, "find_peaks: kernel width out of range");
        int iSlices = (D + kernel_width - 1) / kernel_width;
        if ((iSlices - iSlices / 2) * kernel_width > D)
            return fail(QLP_ERR_ARG, "find_peaks: kernel width too large for the plane");
        hipLaunchKernelGGL(peak_multi_kernel, dim3(n), dim3(64), 0, ctx->stream, n, D, n_peaks,
                           kernel_width, planes, peaks, min_floor, count);
    } else {
        return fail(QLP_ERR_ARG, "find_peaks: unknown peak finder kind");
    }
    return check_hip(hipGetLastError(), "peak kernel launch");
}

}  // namespace qlpiv
