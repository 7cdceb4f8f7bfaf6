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
// maxima are found by all lanes in parallel; the insertions that can change the
// list are replayed in visit order, including the reference's carry-over of the
// previous block's position when a block holds no value above 0.
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

// Maximum (strict >, from 0, row-major first) and minimum of one block; rows
// and columns past the plane edge are not part of the block.  KW > 0 fixes the
// kernel width at compile time so that all KW^2 loads are in flight together.
template <int KW>
__device__ __forceinline__ void block_max(const float *__restrict__ m, int D, int k, int top,
                                          int left, float &bv, int &bi, int &bj, int &upd,
                                          float &mn)
{
    if (KW > 0) {
        float val[KW > 0 ? KW * KW : 1];
#pragma unroll
        for (int oi = 0; oi < KW; oi++)
#pragma unroll
            for (int oj = 0; oj < KW; oj++) {
                const int r = top + oi, cc = left + oj;
                val[oi * KW + oj] = (r < D && cc < D) ? m[r * D + cc] : 0.0f;
            }
#pragma unroll
        for (int oi = 0; oi < KW; oi++)
#pragma unroll
            for (int oj = 0; oj < KW; oj++) {
                const int r = top + oi, cc = left + oj;
                if (r < D && cc < D) {
                    const float v = val[oi * KW + oj];
                    if (v > bv) {
                        bv = v;
                        bi = r;
                        bj = cc;
                        upd = 1;
                    }
                    if (v < mn) mn = v;
                }
            }
    } else {
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
                if (v < mn) mn = v;
            }
        }
    }
}

// W lanes per plane, 64 / W planes per wavefront (W = 64 for few planes, fewer
// lanes per plane when there are many, so that the sequential insertions of
// several planes share one instruction stream).  Each plane's entries (blocks
// in visit order) are processed W at a time:
//  * every lane finds the maximum of one block (strict >, starting at 0) and
//    the minimum of its cells; a block without a value above 0 is inserted as
//    0 at the last position recorded for its quadrant, found with a ballot over
//    the earlier lanes of the same quadrant in the plane's segment or carried
//    from earlier chunks;
//  * the insertions are replayed in visit order, but only for entries that can
//    change the list: once the list holds P entries, a value not above its
//    smallest one is rejected without any other effect (addInAscendOrder
//    :127-129), so a ballot finds each plane's next entry that is either above
//    that minimum or meets a list with free room, and the segment's first lane
//    inserts it into the plane's list in LDS.  Entries skipped between two
//    insertions are exactly those the list would reject.
template <int KW, int W>
__global__ __launch_bounds__(64) void peak_multi_kernel(int n, int D, int P, int k,
                                                        const float *__restrict__ planes,
                                                        float *__restrict__ peaks,
                                                        float *__restrict__ min_floor,
                                                        int *__restrict__ count)
{
    constexpr int S = 64 / W;  // planes per wavefront
    __shared__ MpList s_list[S];
    const int lane = threadIdx.x;
    const int seg = lane / W, sl = lane % W;
    const int ia = blockIdx.x * S + seg;
    const bool act = ia < n;
    const uint64_t segmask = (W == 64) ? ~0ull : (((1ull << (W & 63)) - 1ull) << (seg * W));
    const float *m = planes + (size_t)(act ? ia : 0) * D * D;

    int iSlices = D / k;
    if (iSlices * k != D) iSlices += 1;
    const int iTop = iSlices - iSlices / 2;
    int jSlices = D / k;
    if (jSlices * k != D) jSlices += 1;
    const int jLeft = jSlices - jSlices / 2;
    const int nentries = iTop * jLeft * 4;
    const int q = sl & 3;  // chunks start at multiples of W (a multiple of 4)
    const uint64_t qlanes = 0x1111111111111111ull << (lane & 3);

    if (sl == 0) s_list[seg].size = 0;
    __syncthreads();
    float mn = FLT_MAX;
    int staleI[4] = {0, 0, 0, 0}, staleJ[4] = {0, 0, 0, 0};
    for (int base = 0; base < nentries; base += W) {
        const int e = base + sl;
        const bool valid = act && e < nentries;
        float bv = 0.0f;
        int bi = 0, bj = 0, upd = 0;
        if (valid) {
            const int blk = e >> 2;
            const int si = blk / jLeft, sj = blk - (blk / jLeft) * jLeft;
            const int top = si * k + ((q & 2) ? iTop * k : 0);
            const int left = sj * k + ((q & 1) ? jLeft * k : 0);
            block_max<KW>(m, D, k, top, left, bv, bi, bj, upd, mn);
        }
        // position carried by blocks without an update (FindMaximumMultiPeaks :66-139
        // keeps maxsI/maxsJ of the quadrant from its previous block)
        const uint64_t updm = __ballot(valid && upd);
        const uint64_t prior = updm & qlanes & segmask & ((1ull << lane) - 1ull);
        const int src = prior ? 63 - __clzll(prior) : lane;
        const int ci = __shfl(bi, src, 64), cj = __shfl(bj, src, 64);
        int carryI = staleI[0], carryJ = staleJ[0];
#pragma unroll
        for (int qq = 1; qq < 4; qq++)
            if (q == qq) {
                carryI = staleI[qq];
                carryJ = staleJ[qq];
            }
        const int pi = upd ? bi : (prior ? ci : carryI);
        const int pj = upd ? bj : (prior ? cj : carryJ);
        const float pv = upd ? bv : 0.0f;
#pragma unroll
        for (int qq = 0; qq < 4; qq++) {
            const uint64_t mq = updm & (0x1111111111111111ull << qq) & segmask;
            const int last = mq ? 63 - __clzll(mq) : lane;
            const int li = __shfl(bi, last, 64), lj = __shfl(bj, last, 64);
            if (mq) {
                staleI[qq] = li;
                staleJ[qq] = lj;
            }
        }
        // replay the insertions that can change the lists
        uint64_t pending = __ballot(valid);
        for (;;) {
            const int size = s_list[seg].size;
            const float lo = s_list[seg].v[0];
            const uint64_t cand = __ballot(size < P || pv > lo) & pending;
            if (!cand) break;
            const uint64_t mine = cand & segmask;
            const int c = mine ? __ffsll((unsigned long long)mine) - 1 : lane;
            const float cv = __shfl(pv, c, 64);
            const int cI = __shfl(pi, c, 64), cJ = __shfl(pj, c, 64);
            if (sl == 0 && mine) mp_insert(s_list[seg], P, (k - 1) * k, cv, (short)cI, (short)cJ);
            __syncthreads();
            pending &= ~__ballot(mine && lane <= c);
        }
    }
#pragma unroll
    for (int o = W / 2; o >= 1; o >>= 1) {
        const float om = __shfl_xor(mn, o, 64);
        if (om < mn) mn = om;
    }
    if (sl == 0 && act) {
        const MpList &L = s_list[seg];
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

template <int KW>
static void launch_multi(Context *ctx, int W, int n, int D, int P, int k, const float *planes,
                         float *peaks, float *min_floor, int *count)
{
    const unsigned grid = (unsigned)((n + 64 / W - 1) / (64 / W));
    switch (W) {
    case 8:
        hipLaunchKernelGGL((peak_multi_kernel<KW, 8>), dim3(grid), dim3(64), 0, ctx->stream, n, D,
                           P, k, planes, peaks, min_floor, count);
        break;
    case 16:
        hipLaunchKernelGGL((peak_multi_kernel<KW, 16>), dim3(grid), dim3(64), 0, ctx->stream, n,
                           D, P, k, planes, peaks, min_floor, count);
        break;
    case 32:
        hipLaunchKernelGGL((peak_multi_kernel<KW, 32>), dim3(grid), dim3(64), 0, ctx->stream, n,
                           D, P, k, planes, peaks, min_floor, count);
        break;
    default:
        hipLaunchKernelGGL((peak_multi_kernel<KW, 64>), dim3(grid), dim3(64), 0, ctx->stream, n,
                           D, P, k, planes, peaks, min_floor, count);
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
        // lanes per plane: the fewest that still give every SIMD about two
        // wavefronts.  Measured on MI355X (C2 planes, P = 3, k = 3; us for
        // W = 8/16/32/64): 65,025 x 31^2: 168/213/292/407; 16,129 x 63^2:
        // 188/148/156/185; 3,969 x 127^2: 388/235/158/115.
        // QLPIV_MULTIPEAK_LANES (8, 16, 32 or 64) overrides it for tuning.
        int W = 8;
        while (W < 64 && (size_t)n * W < (size_t)2048 * 64) W *= 2;
        if (const char *env = getenv("QLPIV_MULTIPEAK_LANES")) {
            const int w = atoi(env);
            if (w == 8 || w == 16 || w == 32 || w == 64) W = w;
        }
        if (kernel_width == 3)
            launch_multi<3>(ctx, W, n, D, n_peaks, kernel_width, planes, peaks, min_floor, count);
        else
            launch_multi<0>(ctx, W, n, D, n_peaks, kernel_width, planes, peaks, min_floor, count);
    } else {
        return fail(QLP_ERR_ARG, "find_peaks: unknown peak finder kind");
    }
    return check_hip(hipGetLastError(), "peak kernel launch");
}

}  // namespace qlpiv
