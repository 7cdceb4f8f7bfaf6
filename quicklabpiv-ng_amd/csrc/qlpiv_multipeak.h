// MultiPeaks (maximum/FindMaximumMultiPeaks.java:42-162) over one plane per
// segment of W lanes of a wavefront, shared by the stand-alone peak kernel and
// the correlation passes that still hold their planes in LDS.
//
// Block maxima of kernelWidth^2 blocks in the reference's four-quadrant visit
// order feed a bounded ascending list with neighbour suppression
// (util/SimpleFixedLengthFloatLinkedList.addInAscendOrder :98-237).  The block
// maxima are found by all lanes of the segment in parallel; the insertions that
// can change the list are replayed in visit order, including the reference's
// carry-over of the previous block's position when a block holds no value
// above 0.
#pragma once

#include <float.h>

#include "qlpiv_device.h"

namespace qlpiv {

constexpr int MP_MAX_PEAKS = 16;

struct MpList {
    int size;
    float v[MP_MAX_PEAKS + 1];
    short i[MP_MAX_PEAKS + 1], j[MP_MAX_PEAKS + 1];
};

// addInAscendOrder :98-237 on a sorted array: the same accept/reject and
// delete decisions as the linked list, in the same order.
__device__ __forceinline__ void mp_insert(MpList &L, int P, int kd, float value, short pi, short pj)
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

// addInAscendOrder :98-237 on a list in one lane's registers: PM >= P slots,
// ascending, positions packed as i << 8 | j.  The decisions are those of
// mp_insert; every update is a select gated by the accept decision, and the
// deletions and the shift around the new entry move values between
// neighbouring slots only, so that the slots stay in registers.
template <int PM>
__device__ __forceinline__ void mp_insert_reg(float (&lv)[PM], int (&lp)[PM], int &size, int P,
                                              int kd, float value, int pij)
{
    const int pi = pij >> 8, pj = pij & 255;
    int pos = 0, ndel = 0;
    bool reject = false, del[PM];
#pragma unroll
    for (int e = 0; e < PM; e++) {
        const bool have = e < size;
        const int di = pi - (lp[e] >> 8), dj = pj - (lp[e] & 255);
        const bool near = (float)(di * di + dj * dj) < (float)kd;
        const bool lt = have && lv[e] < value;  // the entries below the insertion point
        pos += lt ? 1 : 0;
        del[e] = lt && near;  // value > lv[e]: deleted if the insertion happens
        ndel += del[e] ? 1 : 0;
        reject = reject || (have && !lt && near && value <= lv[e]);
    }
    const int size1 = size - ndel, ins = pos - ndel;
    // a full list rejects a value not above its smallest; ins == 0 with a full
    // list after the deletions only happens without deletions
    const bool accept = !reject && !(size == P && pos == 0) && !(ins == 0 && size1 == P);
    // deletions, highest slot first (the last slot needs no shift)
#pragma unroll
    for (int e = PM - 2; e >= 0; e--) {
        const bool d = accept && del[e];
#pragma unroll
        for (int t = e; t < PM - 1; t++) {
            lv[t] = d ? lv[t + 1] : lv[t];
            lp[t] = d ? lp[t + 1] : lp[t];
        }
    }
    // insertion at ins; with a full list remove() drops the smallest entry, so
    // that the entries below the new one move down a slot instead
    const bool drop = ins > 0 && size1 + 1 > P;
    const int at = drop ? ins - 1 : ins;
    float nv[PM];
    int np[PM];
#pragma unroll
    for (int t = 0; t < PM; t++) {
        const float vdn = t + 1 < PM ? lv[t + 1] : 0.0f, vup = t > 0 ? lv[t - 1] : 0.0f;
        const int pdn = t + 1 < PM ? lp[t + 1] : 0, pup = t > 0 ? lp[t - 1] : 0;
        nv[t] = t < at ? (drop ? vdn : lv[t]) : (t == at ? value : (drop ? lv[t] : vup));
        np[t] = t < at ? (drop ? pdn : lp[t]) : (t == at ? pij : (drop ? lp[t] : pup));
    }
#pragma unroll
    for (int t = 0; t < PM; t++) {
        lv[t] = accept ? nv[t] : lv[t];
        lp[t] = accept ? np[t] : lp[t];
    }
    size = accept ? size1 + 1 - (drop ? 1 : 0) : size;
}

// MultiPeaks entries for kernel width 3: the blocks of a plane in the
// reference's visit order (slices of the top-left quadrant, then the four
// quadrants), each a block maximum (strict >, starting at 0: 0 for a block
// without a value above 0) and the position (i << 8 | j) of the first cell
// holding it; padded with -2 to a multiple of 16 (a chunk), each chunk's
// maximum beside them.  Geometry for D x D planes:
struct MpGeom {
    int iT;      // slices of the top (and left) half; the planes are square
    int BW;      // blocks per block row (and block rows), both halves
    int nent;    // entries per plane
    int stride;  // entries per plane in the workspace, padded to a chunk
    int nchunk;  // chunks of 16 entries
};

__host__ __device__ constexpr MpGeom mp_geom(int D)
{
    const int iS = (D + 2) / 3, iT = iS - iS / 2, BW = 2 * iT, nent = BW * BW;
    const int stride = (nent + 15) & ~15;
    return MpGeom{iT, BW, nent, stride, stride / 16};
}

// The entry of 3x3 block (br, bc) of a D x D plane from its nine cells c (row-
// major), read with rows and columns past the edge clamped to the last row or
// column: those cells repeat cells of the same block that come earlier in
// row-major order, so they change neither the maximum, nor the first cell
// holding it, nor the minimum.  *bmin is the block minimum (+inf for a block
// wholly past the edge, which has no update).
__device__ __forceinline__ uint2 mp_block_vals(int D, int br, int bc, const float *c, float &bmin)
{
    const int top = br * 3, left = bc * 3;
    const bool out = top >= D || left >= D;
    const float bmax = fmaxf(fmaxf(fmaxf(c[0], c[1]), fmaxf(c[2], c[3])),
                             fmaxf(fmaxf(c[4], c[5]), fmaxf(fmaxf(c[6], c[7]), c[8])));
    const float mn = fminf(fminf(fminf(c[0], c[1]), fminf(c[2], c[3])),
                           fminf(fminf(c[4], c[5]), fminf(fminf(c[6], c[7]), c[8])));
    bmin = out ? FLT_MAX : mn;
    int cell = 8;  // the first cell (row-major) holding the maximum
#pragma unroll
    for (int q = 7; q >= 0; q--) cell = c[q] == bmax ? q : cell;
    const int ci = cell >= 6 ? 2 : (cell >= 3 ? 1 : 0);
    const bool upd = !out && bmax > 0.0f;
    return make_uint2(__float_as_uint(upd ? bmax : 0.0f),
                      (uint32_t)(((top + ci) << 8) | (left + cell - 3 * ci)));
}

// Offsets (row * rs + column) of the nine clamped cells of block (br, bc).
__device__ __forceinline__ void mp_block_offsets(int D, int rs, int br, int bc, int (&o)[9])
{
    const int top = br * 3, left = bc * 3;
    const int r0 = min(top, D - 1) * rs, r1 = min(top + 1, D - 1) * rs,
              r2 = min(top + 2, D - 1) * rs;
    const int c0 = min(left, D - 1), c1 = min(left + 1, D - 1), c2 = min(left + 2, D - 1);
    o[0] = r0 + c0, o[1] = r0 + c1, o[2] = r0 + c2;
    o[3] = r1 + c0, o[4] = r1 + c1, o[5] = r1 + c2;
    o[6] = r2 + c0, o[7] = r2 + c1, o[8] = r2 + c2;
}

// Visit-order index of block (br, bc) (block rows and columns of both halves).
__device__ __forceinline__ int mp_entry_index(const MpGeom &G, int br, int bc)
{
    const int qd = (br >= G.iT ? 2 : 0) | (bc >= G.iT ? 1 : 0);
    const int si = br - ((qd & 2) ? G.iT : 0), sj = bc - ((qd & 1) ? G.iT : 0);
    return ((si * G.iT + sj) << 2) | qd;
}

// The entries and chunk maxima of a D x D plane held in LDS, cell (i, j) read by
// rd(i * RS + j), by NTH threads (a multiple of 16; t = 0 .. NTH-1): thread t
// takes entries t, t + NTH, ... in visit order, so that the 16 entries of a
// chunk sit in 16 consecutive lanes.  The entries go to ent in coalesced
// 8-byte stores and each chunk's maximum is folded across its lanes.  Returns
// the minimum of the thread's blocks.
template <int D, int RS, int NTH, typename F>
__device__ __forceinline__ float mp_entries_core(F rd, int t, uint2 *__restrict__ ent,
                                                 float *__restrict__ cmax)
{
    constexpr MpGeom G = mp_geom(D);
    static_assert(NTH % 16 == 0, "whole chunks of 16 lanes");
    float mn = FLT_MAX;
    for (int e = t; e < G.stride; e += NTH) {  // whole chunks of 16 lanes take part
        uint2 en = make_uint2(__float_as_uint(-2.0f), 0u);  // padding
        if (e < G.nent) {
            const int qd = e & 3, s = e >> 2, si = s / G.iT, sj = s - si * G.iT;
            const int br = si + ((qd & 2) ? G.iT : 0), bc = sj + ((qd & 1) ? G.iT : 0);
            float c[9];
            if constexpr (D % 3 == 0) {
                // every block lies wholly inside the plane or wholly past its
                // edge (no update, no minimum): no cell needs clamping
                const int base = 3 * br * RS + 3 * bc;
                const bool inside = 3 * br < D && 3 * bc < D;
#pragma unroll
                for (int k = 0; k < 9; k++) c[k] = inside ? rd(base + (k / 3) * RS + k % 3) : 0.0f;
            } else {
                int o[9];
                mp_block_offsets(D, RS, br, bc, o);
#pragma unroll
                for (int k = 0; k < 9; k++) c[k] = rd(o[k]);
            }
            float bmin;
            en = mp_block_vals(D, br, bc, c, bmin);
            if (bmin < mn) mn = bmin;
        }
        ent[e] = en;
        float cm = __uint_as_float(en.x);
#pragma unroll
        for (int o = 8; o >= 1; o >>= 1) cm = fmaxf(cm, __shfl_xor(cm, o, 64));
        if ((e & 15) == 0) cmax[e >> 4] = cm;
    }
    return mn;
}

// mp_entries_core by a whole work-group of NT threads (a multiple of 64) over
// a plane at pl[i * S + j]; the minimum is folded across the wavefronts through
// red (NT / 64 floats) and written to *min_floor.
template <int D, int S, int NT>
__device__ __forceinline__ void mp_entries_lds(const float *pl, float *red, uint2 *__restrict__ ent,
                                               float *__restrict__ cmax,
                                               float *__restrict__ min_floor, int tid)
{
    static_assert(NT % 64 == 0, "whole wavefronts");
    float mn = mp_entries_core<D, S, NT>([&](int o) { return pl[o]; }, tid, ent, cmax);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const float om = __shfl_xor(mn, o, 64);
        if (om < mn) mn = om;
    }
    if ((tid & 63) == 0) red[tid >> 6] = mn;
    __syncthreads();
    if (tid == 0) {
        float m = red[0];
        for (int w = 1; w < NT / 64; w++)
            if (red[w] < m) m = red[w];
        *min_floor = m;
    }
}

// The replay of one plane per lane (blockDim.x planes per work-group, the
// work-group's planes consecutive from blockIdx.x * blockDim.x): the chunk
// maxima of the work-group's planes are staged in s_cm (blockDim.x * nchunk
// floats), then each lane runs addInAscendOrder over its plane's entries in
// visit order, skipping the chunks and groups of four whose maximum is not above
// the list's smallest value while the list is full.  A block without an update
// is inserted as 0 at the position its quadrant recorded last
// (FindMaximumMultiPeaks :62-64 keep maxsI/maxsJ across blocks), looked up by a
// walk back over the quadrant's entries not seen yet.  Leaves the ascending list
// in lv / lp (positions i << 8 | j) with size entries.
template <int PM>
__device__ __forceinline__ void mp_replay(int n, int D, int P, const uint2 *__restrict__ ent,
                                          const float *__restrict__ cmax, float *s_cm,
                                          float (&lv)[PM], int (&lp)[PM], int &size)
{
    constexpr int KD = 6;  // (kernelWidth - 1) * kernelWidth
    const MpGeom G = mp_geom(D);
    const int g = threadIdx.x, GW = blockDim.x;
    const int ia = blockIdx.x * GW + g;
    const bool act = ia < n;
    const uint2 *__restrict__ pe = ent + (size_t)(act ? ia : 0) * G.stride;
    {
        const int ia0 = blockIdx.x * GW;
        const int cnt = (min(GW, n - ia0)) * G.nchunk;
        const float *src = cmax + (size_t)ia0 * G.nchunk;
        for (int i = g; i < cnt; i += GW) s_cm[i] = src[i];
        __syncthreads();
    }
    const float *cm = s_cm + (act ? g : 0) * G.nchunk;
#pragma unroll
    for (int t = 0; t < PM; t++) {
        lv[t] = 0.0f;
        lp[t] = 0;
    }
    size = 0;
    // an entry (>= 0) can change the list only if it is above lo: -1 while the
    // list has room, else the list's smallest value
    float lo = act ? -1.0f : FLT_MAX;
    int stale[4] = {0, 0, 0, 0};
    int seen[4] = {-4, -3, -2, -1};  // last entry of each quadrant taken into stale
    for (int c = 0; c < G.nchunk; c++) {
        if (!(cm[c] > lo)) continue;
        uint4 q[8];
        const uint4 *src = (const uint4 *)(pe + 16 * c);
#pragma unroll
        for (int k = 0; k < 8; k++) q[k] = src[k];
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const float v4[4] = {__uint_as_float(q[2 * b].x), __uint_as_float(q[2 * b].z),
                                 __uint_as_float(q[2 * b + 1].x), __uint_as_float(q[2 * b + 1].z)};
            const int p4[4] = {(int)q[2 * b].y, (int)q[2 * b].w, (int)q[2 * b + 1].y,
                               (int)q[2 * b + 1].w};
            if (!(fmaxf(fmaxf(v4[0], v4[1]), fmaxf(v4[2], v4[3])) > lo)) continue;
            const int e = 16 * c + 4 * b;
#pragma unroll
            for (int qd = 0; qd < 4; qd++) {
                const float v = v4[qd];
                if (v > lo) {
                    int pij = p4[qd];
                    if (!(v > 0.0f)) {  // no update: the quadrant's last recorded position
                        for (int f = e + qd - 4; f > seen[qd]; f -= 4)
                            if (__uint_as_float(pe[f].x) > 0.0f) {
                                stale[qd] = (int)pe[f].y;
                                break;
                            }
                        seen[qd] = e + qd;
                        pij = stale[qd];
                    }
                    mp_insert_reg<PM>(lv, lp, size, P, KD, v, pij);
                    lo = size < P ? -1.0f : lv[0];
                }
            }
        }
    }
}

// The peaks (P x 3: i, j, value, descending; slots past the list zero) and the
// peak count (MaxCrossResult.getTotalPeaks :109-115: the secondary arrays exist
// only once a peak index >= 1 was stored) of IA ia from its replayed list.
template <int PM>
__device__ __forceinline__ void mp_store_peaks(int ia, int P, const float (&lv)[PM],
                                               const int (&lp)[PM], int size,
                                               float *__restrict__ peaks, int *__restrict__ count)
{
    float *__restrict__ pk = peaks + (size_t)ia * P * 3;
#pragma unroll
    for (int s = 0; s < PM; s++) {
        if (s < size) {
            const int idx = size - 1 - s;
            pk[3 * idx + 0] = (float)(lp[s] >> 8);
            pk[3 * idx + 1] = (float)(lp[s] & 255);
            pk[3 * idx + 2] = lv[s];
        }
    }
    for (int s = size; s < P; s++) pk[3 * s] = pk[3 * s + 1] = pk[3 * s + 2] = 0.0f;
    if (count) count[ia] = size > 1 ? size : 1;
}

// Maximum (strict >, from 0, row-major first) and minimum of one block; rows
// and columns past the plane edge are not part of the block.  Cell (r, c) is
// m[r * rs + c * cs].  KW > 0 fixes the kernel width at compile time so that
// all KW^2 loads are in flight together.
template <int KW>
__device__ __forceinline__ void block_max(const float *__restrict__ m, int rs, int cs, int D, int k,
                                          int top, int left, float &bv, int &bi, int &bj, int &upd,
                                          float &mn)
{
    if (KW > 0) {
        float val[KW > 0 ? KW * KW : 1];
#pragma unroll
        for (int oi = 0; oi < KW; oi++)
#pragma unroll
            for (int oj = 0; oj < KW; oj++) {
                const int r = top + oi, cc = left + oj;
                val[oi * KW + oj] = (r < D && cc < D) ? m[r * rs + cc * cs] : 0.0f;
            }
        // the same comparisons in the same order, tracking one cell index; a
        // cell past the edge reads 0, which never exceeds bv (>= 0)
        int first = -1;
#pragma unroll
        for (int oi = 0; oi < KW; oi++)
#pragma unroll
            for (int oj = 0; oj < KW; oj++) {
                const float v = val[oi * KW + oj];
                if (v > bv) {
                    bv = v;
                    first = oi * KW + oj;
                }
                if (top + oi < D && left + oj < D && v < mn) mn = v;
            }
        if (first >= 0) {
            bi = top + first / KW;
            bj = left + first % KW;
            upd = 1;
        }
    } else {
        for (int oi = 0; oi < k; oi++) {
            const int r = top + oi;
            if (r >= D) break;
            for (int oj = 0; oj < k; oj++) {
                const int cc = left + oj;
                if (cc >= D) break;
                const float v = m[r * rs + cc * cs];
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

// Orders a wavefront's LDS writes before its later LDS reads by other lanes of
// the same wavefront (its LDS instructions execute in issue order; the fences
// keep the compiler from moving accesses across).
__device__ __forceinline__ void mp_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One plane per segment of W lanes (64 / W segments per wavefront; every lane
// of the wavefront calls this with its own segment's plane m and list L in
// LDS).  The plane's entries (blocks in visit order) are processed W at a time:
//  * every lane finds the maximum of one block (strict >, starting at 0) and
//    the minimum of its cells; a block without a value above 0 is inserted as
//    0 at the last position recorded for its quadrant, found with a ballot over
//    the earlier lanes of the same quadrant in the segment or carried from
//    earlier chunks;
//  * the insertions are replayed in visit order, but only for entries that can
//    change the list: once the list holds P entries, a value not above its
//    smallest one is rejected without any other effect (addInAscendOrder
//    :132-134), so a ballot finds each segment's next entry that is either
//    above that minimum or meets a list with free room, and the segment's first
//    lane inserts it.  Entries skipped between two insertions are exactly those
//    the list would reject.
// The segment's first lane writes peaks (P x 3: i, j, value, descending),
// *min_floor and *count when act.
template <int KW, int W>
__device__ __forceinline__ void multipeak_segment(const float *__restrict__ m, int rs, int cs,
                                                  int D, int P, int k, bool act, MpList &L,
                                                  float *__restrict__ peaks, float *min_floor,
                                                  int *count)
{
    const int lane = threadIdx.x & 63;
    const int seg = lane / W, sl = lane % W;
    const uint64_t segmask = (W == 64) ? ~0ull : (((1ull << (W & 63)) - 1ull) << (seg * W));

    int iSlices = D / k;
    if (iSlices * k != D) iSlices += 1;
    const int iTop = iSlices - iSlices / 2;
    int jSlices = D / k;
    if (jSlices * k != D) jSlices += 1;
    const int jLeft = jSlices - jSlices / 2;
    const int nentries = iTop * jLeft * 4;
    const int q = sl & 3;  // chunks start at multiples of W (a multiple of 4)
    const uint64_t qlanes = 0x1111111111111111ull << (lane & 3);

    if (sl == 0) L.size = 0;
    mp_wave_sync();
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
            block_max<KW>(m, rs, cs, D, k, top, left, bv, bi, bj, upd, mn);
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
            const int size = L.size;
            const float lo = L.v[0];
            const uint64_t cand = __ballot(size < P || pv > lo) & pending;
            if (!cand) break;
            const uint64_t mine = cand & segmask;
            const int c = mine ? __ffsll((unsigned long long)mine) - 1 : lane;
            const float cv = __shfl(pv, c, 64);
            const int cI = __shfl(pi, c, 64), cJ = __shfl(pj, c, 64);
            if (sl == 0 && mine) mp_insert(L, P, (k - 1) * k, cv, (short)cI, (short)cJ);
            mp_wave_sync();
            pending &= ~__ballot(mine && lane <= c);
        }
    }
#pragma unroll
    for (int o = W / 2; o >= 1; o >>= 1) {
        const float om = __shfl_xor(mn, o, 64);
        if (om < mn) mn = om;
    }
    if (sl == 0 && act) {
        for (int p = 0; p < P; p++) peaks[3 * p] = peaks[3 * p + 1] = peaks[3 * p + 2] = 0.0f;
        // descending: peak 0 is the largest
        for (int e = 0; e < L.size; e++) {
            const int idx = L.size - 1 - e;
            peaks[3 * idx + 0] = (float)L.i[e];
            peaks[3 * idx + 1] = (float)L.j[e];
            peaks[3 * idx + 2] = L.v[e];
        }
        *min_floor = mn;
        // MaxCrossResult.getTotalPeaks :109-115: the secondary arrays exist
        // only once a peak index >= 1 was stored
        if (count) *count = L.size > 1 ? L.size : 1;
    }
}

// The r-th (from 0) set bit of a mask.
__device__ __forceinline__ int mp_nth_bit(uint64_t m, int r)
{
    for (; r > 0; r--) m &= m - 1ull;
    return m ? __ffsll((unsigned long long)m) - 1 : 0;
}

// addInAscendOrder :98-237 with the list spread over the lanes of a segment:
// lane sl holds entry sl (ascending) while sl < size.  Every lane of the
// segment calls it with the same candidate; the accept/reject and delete
// decisions are those of mp_insert.
__device__ __forceinline__ void mp_insert_lanes(float &ev, int &ei, int &ej, int &size, int sl,
                                                uint64_t segmask, int P, int kd, float value,
                                                int pi, int pj)
{
    if (size == 0) {
        if (sl == 0) {
            ev = value;
            ei = pi;
            ej = pj;
        }
        size = 1;
        return;
    }
    const bool have = sl < size;
    const int di = pi - ei, dj = pj - ej;
    const bool near = have && (float)(di * di + dj * dj) < (float)kd;
    const bool lt = have && ev < value;  // entries below the insertion point
    const int pos = __popcll(__ballot(lt) & segmask);
    if (size == P && pos == 0) return;
    // an entry at or above the insertion point (value <= its value) nearby rejects it
    if (__ballot(near && have && !lt) & segmask) return;
    // nearby entries below it are deleted
    const uint64_t delm = __ballot(near && lt) & segmask;
    const int ndel = __popcll(delm);
    const int size1 = size - ndel;
    const int ins = pos - ndel;
    if (ins == 0 && size1 == P) return;
    const int drop = (ins > 0 && size1 + 1 > P) ? 1 : 0;  // remove() of the smallest entry
    const int nsize = size1 + 1 - drop;
    const uint64_t surv = __ballot(have) & segmask & ~delm;
    // target index sl: position u of the list after the insertion, before the drop
    const int u = sl + drop;
    const bool fresh = u == ins;
    const int src = mp_nth_bit(surv, (u < ins) ? u : (u > ins ? u - 1 : 0));
    const float sv = __shfl(ev, src, 64);
    const int si = __shfl(ei, src, 64), sj = __shfl(ej, src, 64);
    if (sl < nsize) {
        ev = fresh ? value : sv;
        ei = fresh ? pi : si;
        ej = fresh ? pj : sj;
    }
    size = nsize;
}

// multipeak_segment with the list in registers (needs P <= W): same visit
// order, insertion replay and outputs.
template <int KW, int W>
__device__ __forceinline__ void multipeak_segment_lanes(const float *__restrict__ m, int rs, int cs,
                                                        int D, int P, int k, bool act,
                                                        float *__restrict__ peaks,
                                                        float *min_floor, int *count)
{
    const int lane = threadIdx.x & 63;
    const int seg = lane / W, sl = lane % W;
    const uint64_t segmask = (W == 64) ? ~0ull : (((1ull << (W & 63)) - 1ull) << (seg * W));

    int iSlices = D / k;
    if (iSlices * k != D) iSlices += 1;
    const int iTop = iSlices - iSlices / 2;
    int jSlices = D / k;
    if (jSlices * k != D) jSlices += 1;
    const int jLeft = jSlices - jSlices / 2;
    const int nentries = iTop * jLeft * 4;
    const int q = sl & 3;
    const uint64_t qlanes = 0x1111111111111111ull << (lane & 3);
    const int kd = (k - 1) * k;

    float ev = 0.0f;
    int ei = 0, ej = 0, size = 0;
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
            block_max<KW>(m, rs, cs, D, k, top, left, bv, bi, bj, upd, mn);
        }
        const uint64_t updm = __ballot(valid && upd);
        const uint64_t prior = updm & qlanes & segmask & ((1ull << lane) - 1ull);
        const int psrc = prior ? 63 - __clzll(prior) : lane;
        const int ci = __shfl(bi, psrc, 64), cj = __shfl(bj, psrc, 64);
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
        uint64_t pending = __ballot(valid);
        for (;;) {
            const float lo = __shfl(ev, seg * W, 64);
            const uint64_t cand = __ballot(size < P || pv > lo) & pending;
            if (!cand) break;
            const uint64_t mine = cand & segmask;
            const int c = mine ? __ffsll((unsigned long long)mine) - 1 : lane;
            const float cv = __shfl(pv, c, 64);
            const int cI = __shfl(pi, c, 64), cJ = __shfl(pj, c, 64);
            if (mine) mp_insert_lanes(ev, ei, ej, size, sl, segmask, P, kd, cv, cI, cJ);
            pending &= ~__ballot(mine && lane <= c);
        }
    }
#pragma unroll
    for (int o = W / 2; o >= 1; o >>= 1) {
        const float om = __shfl_xor(mn, o, 64);
        if (om < mn) mn = om;
    }
    if (act && sl < P) {
        // descending: peak 0 is the largest; slots past the list are zero
        if (sl < size) {
            const int idx = size - 1 - sl;
            peaks[3 * idx + 0] = (float)ei;
            peaks[3 * idx + 1] = (float)ej;
            peaks[3 * idx + 2] = ev;
        } else {
            peaks[3 * sl] = peaks[3 * sl + 1] = peaks[3 * sl + 2] = 0.0f;
        }
    }
    if (act && sl == 0) {
        *min_floor = mn;
        if (count) *count = size > 1 ? size : 1;
    }
}

}  // namespace qlpiv
