// Vector validation and replacement of displacement fields
// (jobs/validator/VectorValidatorJob.java:105-405) on the GPU.
//
// The reference visits the tiles one at a time in a centre-out spiral
// (circularVisitor :105-196), and a visit may change the tile's displacement,
// which the later visits of its neighbours read.  Two facts make the job
// parallel without changing any result:
//  * The initializer visit (:268-281) only validates.  The normalised median
//    verdict of a tile depends on its neighbours' displacements, never on their
//    validity flags, so every tile is validated at once; tiles found valid are
//    locked (Tile.setLockedValidationState) and never change again.  With the
//    difference validator every tile stays invalid: all tiles start invalid
//    (Tile :56, resetDisplacements :74-81) and that validator only updates a
//    tile from a valid neighbour, so by induction over the visit order no
//    tile ever sees one.
//  * A visit of the validator loop (:283-353) reads and writes only its 3x3
//    neighbourhood.  Order matters only between two unlocked ("active") tiles
//    that are neighbours: the later one in spiral order must see the earlier
//    one's new state, the earlier one the later one's old state.  A sweep
//    therefore runs in rounds: an active tile is processed once every active
//    neighbour earlier in spiral order was processed in an earlier round of
//    the same sweep.  Two neighbours never share a round, so every visit reads
//    exactly what the sequential visit reads.
// The sweep kernel is one workgroup per frame pair; rounds are separated by a
// barrier, and the state other lanes wrote is read with device-coherent loads.
#include <float.h>
#include <math.h>

#include <vector>

#include "qlpiv_internal.h"

namespace qlpiv {

enum { F_INVALID = 1, F_LOCKED = 2 };

struct VStage {
    int validator;
    float threshold, eps;
    int replacement, peak_index;
};

struct VParams {
    int nI, nJ, n, P, n_stages, max_iter, until;
    VStage st[QLP_VALID_MAX_STAGES];
    const int *top, *left, *rank;
    float *u, *v, *bu, *bv;
    const float *pk;
    int pk_stride;
    float pk_off;
    const int *pcount;
    int *flags, *done, *act, *act_count;
    uint8_t *invalid;
    int *stats;
};

__device__ __forceinline__ float ldc(const float *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ldc(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Comparator of NormalizedMedianValidator.distanceComparator :19-35.
__device__ __forceinline__ int jcmp(float a, float b) { return a > b ? 1 : (a < b ? -1 : 0); }

// Collections.sort of at most 8 tiles: java.util.TimSort (OpenJDK 8) takes its
// short-array path, countRunAndMakeAscending followed by binarySort.  Only
// needed when a key is NaN (the comparator then calls NaN equal to
// everything and the order depends on the algorithm); otherwise every stable
// sort gives the same order and the rank form below is used.
__device__ void timsort_small(float *key, float *a, float *b, int n)
{
    if (n < 2) return;
    int runHi = 2;
    if (jcmp(key[1], key[0]) < 0) {
        while (runHi < n && jcmp(key[runHi], key[runHi - 1]) < 0) runHi++;
        for (int lo = 0, hi = runHi - 1; lo < hi; lo++, hi--) {
            float t = key[lo]; key[lo] = key[hi]; key[hi] = t;
            t = a[lo]; a[lo] = a[hi]; a[hi] = t;
            t = b[lo]; b[lo] = b[hi]; b[hi] = t;
        }
    } else {
        while (runHi < n && jcmp(key[runHi], key[runHi - 1]) >= 0) runHi++;
    }
    for (int start = runHi; start < n; start++) {
        const float pk = key[start], pa = a[start], pb = b[start];
        int left = 0, right = start;
        while (left < right) {
            const int mid = (left + right) >> 1;
            if (jcmp(pk, key[mid]) < 0) right = mid;
            else left = mid + 1;
        }
        for (int m = start; m > left; m--) {
            key[m] = key[m - 1];
            a[m] = a[m - 1];
            b[m] = b[m - 1];
        }
        key[left] = pk;
        a[left] = pa;
        b[left] = pb;
    }
}

// Median of the norm-sorted neighbours through the exact TimSort path.  Kept
// out of line and reloading the window, so that its arrays (private memory)
// stay off the register-resident common path.
__device__ __attribute__((noinline)) float2 median_timsort(const float *u, const float *v, int nI,
                                                           int nJ, size_t base, int i, int j)
{
    float nu[8], nv[8], key[8];
    int cnt = 0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            if (a == 1 && b == 1) continue;
            const int ii = i + a - 1, jj = j + b - 1;
            if (ii < 0 || ii >= nI || jj < 0 || jj >= nJ) continue;
            const size_t o = base + (size_t)ii * nJ + jj;
            nu[cnt] = ldc(u + o);
            nv[cnt] = ldc(v + o);
            key[cnt] = nu[cnt] * nu[cnt] + nv[cnt] * nv[cnt];
            cnt++;
        }
    timsort_small(key, nu, nv, cnt);
    float2 m = make_float2(0.0f, 0.0f);
    if (cnt > 0) {
        if (cnt % 2 == 0) {
            m.x = (nu[cnt / 2 - 1] + nu[cnt / 2]) / 2.0f;
            m.y = (nv[cnt / 2 - 1] + nv[cnt / 2]) / 2.0f;
        } else {
            m.x = nu[cnt / 2];
            m.y = nv[cnt / 2];
        }
    }
    return m;
}

// Arrays.sort(float[]) order: ascending with NaN after every number.
__device__ __forceinline__ bool flt_before(float x, float y)
{
    if (isnan(x)) return false;
    if (isnan(y)) return true;
    return x < y;
}

// The 3 x 3 window around tile (i, j) without the tile: slot s is window
// position (a, b) in reading order, (1, 1) skipped.  Its displacements (and
// validity flags) cannot change during a visit of (i, j) (no neighbour is
// visited in the same round), so a visit loads them once.  The loads are
// unconditional (a slot outside the grid reads the tile itself and is then
// ignored), so that they are in flight together.
struct Window {
    float nu[8], nv[8];
    int fl[8];  // flags (loaded by load_window_flags)
    bool present[8];
    int k[8];   // grid index, -1 outside the grid
};

__device__ __forceinline__ void window_slot(int s, int &a, int &b)
{
    a = s < 4 ? s / 3 : (s + 1) / 3;
    b = s < 4 ? s % 3 : (s + 1) % 3;
}

__device__ __forceinline__ Window load_window(const VParams &p, size_t base, int i, int j)
{
    Window w;
    size_t o[8];
#pragma unroll
    for (int s = 0; s < 8; s++) {
        int a, b;
        window_slot(s, a, b);
        const int ii = i + a - 1, jj = j + b - 1;
        w.present[s] = ii >= 0 && ii < p.nI && jj >= 0 && jj < p.nJ;
        w.k[s] = w.present[s] ? ii * p.nJ + jj : -1;
        o[s] = base + (size_t)(w.present[s] ? ii * p.nJ + jj : i * p.nJ + j);
    }
#pragma unroll
    for (int s = 0; s < 8; s++) {
        w.nu[s] = ldc(p.u + o[s]);
        w.nv[s] = ldc(p.v + o[s]);
        w.fl[s] = 0;
    }
#pragma unroll
    for (int s = 0; s < 8; s++) {
        if (!w.present[s]) w.nu[s] = w.nv[s] = 0.0f;
    }
    return w;
}

__device__ __forceinline__ void load_window_flags(const VParams &p, size_t base, int i, int j,
                                                  Window &w)
{
#pragma unroll
    for (int s = 0; s < 8; s++) w.fl[s] = ldc(p.flags + base + (w.k[s] >= 0 ? w.k[s] : i * p.nJ + j));
}

// NormalizedMedianValidator.validateVector :63-136: the new invalid state.
// The eight window slots stay in registers; sorted positions are found by
// ranks (ties by window order, which is the insertion order of the
// reference's list, so the stable sort's order).
__device__ bool median_invalid(const VParams &p, const Window &w, size_t base, int i, int j,
                               float tu, float tv, float thr, float eps)
{
    float nu[8], nv[8], key[8];
    bool present[8];
    int cnt = 0;
    bool anynan = false;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        present[s] = w.present[s];
        nu[s] = w.nu[s];
        nv[s] = w.nv[s];
        cnt += present[s] ? 1 : 0;
        key[s] = nu[s] * nu[s] + nv[s] * nv[s];
        anynan |= present[s] && isnan(key[s]);
    }
    float mu = 0.0f, mv = 0.0f;
    if (anynan) {
        const float2 m = median_timsort(p.u, p.v, p.nI, p.nJ, base, i, j);
        mu = m.x;
        mv = m.y;
    } else if (cnt > 0) {
        const int hi = cnt / 2, lo = cnt / 2 - 1;
        float au = 0.0f, av = 0.0f, bu = 0.0f, bv = 0.0f;
#pragma unroll
        for (int s = 0; s < 8; s++) {
            int rank = 0;
#pragma unroll
            for (int t = 0; t < 8; t++)
                if (t != s && present[t] && (key[t] < key[s] || (key[t] == key[s] && t < s))) rank++;
            if (present[s] && rank == lo) {
                au = nu[s];
                av = nv[s];
            }
            if (present[s] && rank == hi) {
                bu = nu[s];
                bv = nv[s];
            }
        }
        if (cnt % 2 == 0) {
            mu = (au + bu) / 2.0f;
            mv = (av + bv) / 2.0f;
        } else {
            mu = bu;
            mv = bv;
        }
    }
    // residuals in the reference's 8-slot array, unused slots Float.MAX_VALUE
    float r[8];
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const float dU = mu - nu[s], dV = mv - nv[s];
        r[s] = present[s] ? dU * dU + dV * dV : FLT_MAX;
    }
    // the reference's index arithmetic, kept as is (C and Java both truncate
    // -1 / 2 to 0 and give -1 % 2 == -1)
    const int ri = cnt;
    const int t1 = (ri - 1) / 2, t2 = (ri - 1) / 2 + 1;
    float r1 = 0.0f, r2 = 0.0f;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        int rank = 0;
#pragma unroll
        for (int t = 0; t < 8; t++)
            if (t != s && (flt_before(r[t], r[s]) ||
                           (t < s && !flt_before(r[s], r[t]) && !flt_before(r[t], r[s]))))
                rank++;
        if (rank == t1) r1 = r[s];
        if (rank == t2) r2 = r[s];
    }
    float rMed;
    if (ri - 1 == 0) rMed = 1.0f;
    else if ((ri - 1) % 2 == 0) rMed = (float)sqrt((double)((r1 + r2) / 2.0f));
    else rMed = (float)sqrt((double)r2);
    const float dU = mu - tu, dV = mv - tv;
    const float distance = (float)sqrt((double)(dU * dU + dV * dV)) / (rMed + eps);
    return !(distance < thr);
}

// DifferenceValidator.validateVector :34-63: the tile's own state is kept when
// none of the three neighbours it consults (upper-left, upper, left) is valid.
// (window slots 0, 1, 3: upper-left, upper, left; the window's flags loaded)
__device__ bool difference_invalid(const Window &w, float tu, float tv, float thr, bool cur)
{
    const float thr2 = thr * thr;
    bool inv = cur;
    const int slots[3] = {0, 1, 3};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const int s = slots[q];
        if (!w.present[s] || (w.fl[s] & F_INVALID)) continue;
        const float du = w.nu[s] - tu, dv = w.nv[s] - tv;
        const float distance = du * du + dv * dv;
        inv = !(distance < thr2);
    }
    return inv;
}

__device__ __forceinline__ bool stage_invalid(const VParams &p, const Window &w, int s, size_t base,
                                              int i, int j, float tu, float tv, bool cur)
{
    const VStage &st = p.st[s];
    if (st.validator == QLP_VALIDATOR_MEDIAN)
        return median_invalid(p, w, base, i, j, tu, tv, st.threshold, st.eps);
    return difference_invalid(w, tu, tv, st.threshold, cur);
}

// Neighbour of the 3x3 window (a, b in 0..2) or -1 outside the grid.
__device__ __forceinline__ int nb_index(const VParams &p, int i, int j, int a, int b)
{
    const int ii = i + a - 1, jj = j + b - 1;
    if (ii < 0 || ii >= p.nI || jj < 0 || jj >= p.nJ) return -1;
    return ii * p.nJ + jj;
}

struct NbTile {
    bool ok;  // exists and is valid
    float u, v, top, left;
};

constexpr int PK_PRE = 4;  // peaks of a tile loaded with its visit's other inputs

// Everything one sweep step reads for an active tile, loaded in one batch:
// its own state and peaks, the window (displacements and flags), the
// neighbours' spiral ranks, done rounds and origins.  When the tile turns out
// ready, no neighbour the loads saw can change before its visit ends: the
// earlier ones were visited in earlier rounds, the later ones wait for it.
struct VisitIn {
    Window w;
    int rk[8], dn[8];
    float ntop[8], nleft[8];
    int rank, done, flag, pcount;
    float tu, tv, bu, bv, ttop, tleft;
    float pk[PK_PRE][2];
};

// Window neighbour (a, b) from a visit's loaded inputs: validity, displacement
// and tile origin.
__device__ __forceinline__ NbTile nb_tile(const VisitIn &d, int a, int b)
{
    NbTile t{false, 0.0f, 0.0f, 0.0f, 0.0f};
    const int s = a * 3 + b - (a * 3 + b > 4 ? 1 : 0);
    if (d.w.k[s] < 0) return t;
    t.top = d.ntop[s];
    t.left = d.nleft[s];
    if (d.w.fl[s] & F_INVALID) return t;
    t.ok = true;
    t.u = d.w.nu[s];
    t.v = d.w.nv[s];
    return t;
}

// BilinearReplacement.interpV :31-80 (horizontal, over left pixels); interpU
// :82-135 is the same over top pixels.  err receives the cases the reference
// rejects with VectorReplacementException.
__device__ float interp_1d(float x1, float x2, float xi, float a1, float a2, int &err)
{
    float w1v = 0.0f, w2v = 0.0f;
    if ((x2 - xi) > 1e-6f && (xi - x1) > 1e-6f && (x2 - x1) > 1e-6f) {
        const float w1 = (x2 - xi) / (x2 - x1);
        if (w1 < 0.0f || w1 > 1.0f) err = 1;
        w1v = w1 * a1;
        const float w2 = (xi - x1) / (x2 - x1);
        if (w2 < 0.0f || w2 > 1.0f) err = 1;
        w2v = w2 * a2;
    } else {
        int count = 0;
        if (fabsf(xi - x2) < 1e-6f) {
            w2v = a2;
            count++;
        }
        if (fabsf(xi - x1) < 1e-6f) {
            w1v = a1;
            count++;
        }
        if (count == 0) err = 1;
        if (count > 1) {
            w1v /= (float)count;
            w2v /= (float)count;
        }
    }
    return w1v + w2v;
}

__device__ __forceinline__ float weight(float num, float den, int &err)
{
    const float w = num / den;
    if (w < 0.0f || w > 1.0f) err = 1;
    return w;
}

// BilinearReplacement.replaceVector :145-427 on an invalid tile.
__device__ void replace_bilinear(const VParams &p, const VisitIn &d, int i, int j, float ttop,
                                 float tleft, float &tu, float &tv, bool &inv, int &err)
{
    const NbTile left = nb_tile(d, 1, 0), right = nb_tile(d, 1, 2);
    const NbTile top = nb_tile(d, 0, 1), bottom = nb_tile(d, 2, 1);
    const NbTile tl = nb_tile(d, 0, 0), br = nb_tile(d, 2, 2);
    const NbTile bl = nb_tile(d, 2, 0), tr = nb_tile(d, 0, 2);
    const bool tlx = nb_index(p, i, j, 0, 0) >= 0, brx = nb_index(p, i, j, 2, 2) >= 0;
    const bool blx = nb_index(p, i, j, 2, 0) >= 0, trx = nb_index(p, i, j, 0, 2) >= 0;
    // tile alignment checks :174-189
    if (tlx && blx && tl.left - bl.left > 1.0e-6f) err = 1;
    if (trx && brx && tr.left - br.left > 1.0e-6f) err = 1;
    if (tlx && trx && tl.top - tr.top > 1.0e-6f) err = 1;
    if (blx && brx && bl.top - br.top > 1.0e-6f) err = 1;
    float uV[4] = {0.0f, 0.0f, 0.0f, 0.0f}, uU[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int cV[4] = {0, 0, 0, 0}, cU[4] = {0, 0, 0, 0};
    int count = 0;
    if (left.ok && right.ok) {
        uV[0] = interp_1d(left.left, right.left, tleft, left.v, right.v, err);
        uU[0] = interp_1d(left.top, right.top, ttop, left.u, right.u, err);
        cV[0] = 2;
        cU[0] = 1;
        count++;
    }
    if (bottom.ok && top.ok) {
        uV[1] = interp_1d(top.left, bottom.left, tleft, top.v, bottom.v, err);
        uU[1] = interp_1d(top.top, bottom.top, ttop, top.u, bottom.u, err);
        cV[1] = 1;
        cU[1] = 2;
        count++;
    }
    if (tl.ok && br.ok && bl.ok && tr.ok) {
        const float wTR = weight(tleft - tl.left, tr.left - tl.left, err);
        const float wTL = weight(tr.left - tleft, tr.left - tl.left, err);
        const float wBR = weight(tleft - bl.left, br.left - bl.left, err);
        const float wBL = weight(br.left - tleft, br.left - bl.left, err);
        const float wT = weight(bl.top - ttop, bl.top - tl.top, err);
        const float wB = weight(ttop - tl.top, bl.top - tl.top, err);
        if (fabs((double)(1.0f - wTR - wTL)) > 1.0e-6) err = 1;
        if (fabs((double)(1.0f - wBR - wBL)) > 1.0e-6) err = 1;
        if (fabs((double)(1.0f - wT - wB)) > 1.0e-6) err = 1;
        uV[2] = wT * (wTR * tr.v + wTL * tl.v) + wB * (wBR * br.v + wBL * bl.v);
        uU[2] = wT * (wTR * tr.u + wTL * tl.u) + wB * (wBR * br.u + wBL * bl.u);
        cV[2] = 3;
        cU[2] = 3;
        count++;
    } else {
        if (tl.ok && br.ok) {
            uV[2] = interp_1d(tl.left, br.left, tleft, tl.v, br.v, err);
            uU[2] = interp_1d(tl.top, br.top, ttop, tl.u, br.u, err);
            cV[2] = 1;
            cU[2] = 1;
            count++;
        }
        if (bl.ok && tr.ok) {
            uV[3] = interp_1d(bl.left, tr.left, tleft, bl.v, tr.v, err);
            uU[3] = interp_1d(tr.top, bl.top, ttop, tr.u, bl.u, err);
            cV[3] = 1;
            cU[3] = 1;
            count++;
        }
    }
    if (count == 0) {
        // :276-301: the first valid neighbour of each window row overwrites
        // the candidate (the break leaves only the inner loop)
        bool found = false;
        float ru = 0.0f, rv = 0.0f;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                if (a == 1 && b == 1) continue;
                const NbTile t = nb_tile(d, a, b);
                if (t.ok) {
                    ru = t.u;
                    rv = t.v;
                    found = true;
                    break;
                }
            }
        if (found) {
            tu = ru;
            tv = rv;
            inv = false;
        }
        return;
    }
    float accV = 0.0f, accU = 0.0f;
    int wU = 0, wV = 0;
    for (int k = 0; k < 4; k++) {
        accV += (float)cV[k] * uV[k];
        wV += cV[k];
        accU += (float)cU[k] * uU[k];
        wU += cU[k];
    }
    accV /= (float)wV;
    accU /= (float)wU;
    tu = accU;
    tv = accV;
    inv = false;
}

// SecondaryPeakReplacement.replaceVector :41-75: undo the main peak's
// increment, add the chosen peak's.
__device__ __forceinline__ void replace_secondary(const VParams &p, size_t gt, int k, int total,
                                                  const float (&pk)[PK_PRE][2], float &tu,
                                                  float &tv, bool &inv)
{
    if (total <= k || k >= p.P) return;
    const float *m = p.pk + gt * p.P * p.pk_stride;
    float ku = 0.0f, kv = 0.0f;
#pragma unroll
    for (int q = 1; q < PK_PRE; q++)
        if (q == k) {
            ku = pk[q][0];
            kv = pk[q][1];
        }
    if (k >= PK_PRE) {
        ku = m[(size_t)k * p.pk_stride];
        kv = m[(size_t)k * p.pk_stride + 1];
    }
    const float uOld = pk[0][0] - p.pk_off, vOld = pk[0][1] - p.pk_off;
    const float uNew = ku - p.pk_off;
    const float vNew = kv - p.pk_off;
    float u = tu, v = tv;
    u -= uOld;
    v -= vOld;
    u += uNew;
    v += vNew;
    tu = u;
    tv = v;
    inv = false;
}

// Backup, initializer visit (:243-282) and the active list.
__global__ __launch_bounds__(256) void valid_init_kernel(VParams p, int pairs)
{
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (size_t)pairs * p.n) return;
    const int pair = (int)(gid / p.n), t = (int)(gid - (size_t)pair * p.n);
    const size_t base = (size_t)pair * p.n;
    const float tu = p.u[gid], tv = p.v[gid];
    p.bu[gid] = tu;
    p.bv[gid] = tv;
    p.done[gid] = -1;
    bool inv = true;
    if (p.st[0].validator == QLP_VALIDATOR_MEDIAN) {
        const int i = t / p.nJ, j = t % p.nJ;
        inv = median_invalid(p, load_window(p, base, i, j), base, i, j, tu, tv, p.st[0].threshold,
                             p.st[0].eps);
    }
    p.flags[gid] = inv ? F_INVALID : F_LOCKED;
    // a locked tile ends valid; the sweep writes the flags of the active ones
    if (p.invalid) p.invalid[gid] = 0;
    if (inv) {
        const int slot = atomicAdd(p.act_count + pair, 1);
        p.act[base + slot] = t;
    }
}

__device__ __forceinline__ void load_visit(const VParams &p, size_t base, int t, VisitIn &d)
{
    const int i = t / p.nJ, j = t - (t / p.nJ) * p.nJ;
    const size_t gt = base + t;
    d.w = load_window(p, base, i, j);
    load_window_flags(p, base, i, j, d.w);
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int kk = d.w.k[s] < 0 ? t : d.w.k[s];
        d.rk[s] = p.rank[kk];
        d.dn[s] = ldc(p.done + base + kk);
        d.ntop[s] = (float)p.top[kk];
        d.nleft[s] = (float)p.left[kk];
    }
    d.rank = p.rank[t];
    d.done = p.done[gt];
    d.flag = p.flags[gt];
    d.tu = p.u[gt];
    d.tv = p.v[gt];
    d.bu = p.bu[gt];
    d.bv = p.bv[gt];
    d.ttop = (float)p.top[t];
    d.tleft = (float)p.left[t];
    d.pcount = p.pcount ? p.pcount[gt] : p.P;
    const float *m = p.pk + gt * p.P * p.pk_stride;
#pragma unroll
    for (int q = 0; q < PK_PRE; q++) {
        const int qq = q < p.P ? q : 0;
        d.pk[q][0] = m[(size_t)qq * p.pk_stride];
        d.pk[q][1] = m[(size_t)qq * p.pk_stride + 1];
    }
}

// A tile is ready once every neighbour visited before it in the spiral (and not
// locked) has been visited in this sweep, in an earlier round.
__device__ __forceinline__ bool ready(const VisitIn &d, int sweep_start, int round)
{
    bool ok = true;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        if (d.w.k[q] < 0 || d.rk[q] > d.rank || (d.w.fl[q] & F_LOCKED)) continue;
        if (d.dn[q] < sweep_start || d.dn[q] >= round) ok = false;
    }
    return ok;
}

// Visit of one active tile by the validator loop (:283-353).
__device__ void visit(const VParams &p, size_t base, int t, const VisitIn &d, int &corrected,
                      int &err)
{
    const int i = t / p.nJ, j = t - (t / p.nJ) * p.nJ;
    const size_t gt = base + t;
    float tu = d.tu, tv = d.tv;
    bool inv = (d.flag & F_INVALID) != 0;
    inv = stage_invalid(p, d.w, 0, base, i, j, tu, tv, inv);
    for (int s = 0; s < p.n_stages && inv; s++) {
        const int rep = p.st[s].replacement;
        if (rep != QLP_REPLACE_NONE) {
            tu = d.bu;  // Tile.restoreDisplacement
            tv = d.bv;
            if (rep == QLP_REPLACE_SECONDARY_PEAK)
                replace_secondary(p, gt, p.st[s].peak_index, d.pcount, d.pk, tu, tv, inv);
            else
                replace_bilinear(p, d, i, j, d.ttop, d.tleft, tu, tv, inv, err);
        }
        inv = stage_invalid(p, d.w, s, base, i, j, tu, tv, inv);
        if (!inv) corrected++;
    }
    p.u[gt] = tu;
    p.v[gt] = tv;
    p.flags[gt] = inv ? F_INVALID : 0;
}

// 256 lanes: the active tiles are a few percent of the grid, and a smaller
// group leaves the visit its registers (256 VGPRs, 80 bytes of scratch per
// lane on gfx950; a larger group would spill far more)
constexpr int SWEEP_THREADS = 256;

__global__ __launch_bounds__(SWEEP_THREADS) void valid_sweep_kernel(VParams p)
{
    __shared__ int s_prog, s_corr, s_err, s_inv;
    const int pair = blockIdx.x;
    const size_t base = (size_t)pair * p.n;
    const int cnt = p.act_count[pair];
    const int *act = p.act + base;
    int round = 0, sweeps = 0, total_corr = 0, capped = 0, err = 0;
    for (;;) {
        const int sweep_start = round;
        if (threadIdx.x == 0) s_corr = 0;
        __syncthreads();
        int processed = 0, corrected = 0;
        while (processed < cnt) {
            if (threadIdx.x == 0) s_prog = 0;
            __syncthreads();
            int mine = 0;
            for (int a = threadIdx.x; a < cnt; a += SWEEP_THREADS) {
                const int t = act[a];
                VisitIn d;
                load_visit(p, base, t, d);
                if (d.done >= sweep_start) continue;
                if (!ready(d, sweep_start, round)) continue;
                visit(p, base, t, d, corrected, err);
                p.done[base + t] = round;
                mine++;
            }
            if (mine) atomicAdd(&s_prog, mine);
            // the work-group barrier orders this round's stores before the next
            // round's loads (every tile of the pair is handled by this one
            // work-group; the device-coherent loads read past the L1)
            __syncthreads();
            processed += s_prog;
            round++;
            __syncthreads();
        }
        if (corrected) atomicAdd(&s_corr, corrected);
        __syncthreads();
        const int corr = s_corr;
        __syncthreads();
        sweeps++;
        total_corr += corr;
        // VectorValidatorJob.shouldDoAnotherAttempt :84-103
        if (corr == 0) break;
        if (!p.until && sweeps >= p.max_iter) break;
        if (sweeps >= QLP_VALID_MAX_SWEEPS) {
            capped = 1;
            break;
        }
    }
    // :372-383: vectors still invalid become (0, 0)
    if (threadIdx.x == 0) {
        s_inv = 0;
        s_err = 0;
    }
    __syncthreads();
    int inv = 0;
    for (int a = threadIdx.x; a < cnt; a += SWEEP_THREADS) {
        const size_t gt = base + act[a];
        const bool bad = (p.flags[gt] & F_INVALID) != 0;
        if (bad) {
            p.u[gt] = 0.0f;
            p.v[gt] = 0.0f;
            inv++;
        }
        if (p.invalid) p.invalid[gt] = bad ? 1 : 0;  // the other tiles: 0 from the initializer
    }
    if (inv) atomicAdd(&s_inv, inv);
    if (err) atomicOr(&s_err, 2);
    __syncthreads();
    // the active count returns to 0 for the next call (every thread read it at
    // the start, before the barriers above)
    if (threadIdx.x == 0) p.act_count[pair] = 0;
    if (p.stats && threadIdx.x == 0) {
        int *st = p.stats + (size_t)pair * 4;
        st[0] = sweeps;
        st[1] = total_corr;
        st[2] = s_inv;
        st[3] = capped | s_err;
    }
}

__global__ void zero_int_kernel(int *x, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0;
}

// Displacement update of the multi-pass loop from the main peak of a plane
// fit: u += i' - D/2, v += j' - D/2 (IterationStepTiles
// .updateDisplacementsFromMaxCrossResults :317-371 via Tile.accumulateDisplacement).
__global__ void accumulate_kernel(int n, int P, int D, const float *__restrict__ sub,
                                  float *__restrict__ u, float *__restrict__ v)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float c = (float)(D / 2);
    const float *s = sub + (size_t)t * P * 4;
    u[t] = u[t] + (s[0] - c);
    v[t] = v[t] + (s[1] - c);
}

// VectorValidatorJob.circularVisitor :105-196: position of every tile in the
// visit order.  The rings grow by one tile on every side per turn, so the walk
// reaches every tile; each tile is ranked once.
void spiral_rank(int nI, int nJ, int *rank)
{
    const int total = nI * nJ;
    for (int k = 0; k < total; k++) rank[k] = -1;
    int startI = nI / 2, startJ = nJ / 2;
    int sideI = 2, sideJ = 2;
    if (nI != startI * 2) sideI = 1;
    else startI--;
    if (nJ != startJ * 2) sideJ = 1;
    else startJ--;
    int visited = 0;
    auto visit_at = [&](int i, int j) {
        if (i >= 0 && i < nI && j >= 0 && j < nJ && rank[i * nJ + j] < 0) rank[i * nJ + j] = visited++;
    };
    while (visited < total) {
        const int mR = sideJ > 1 ? sideJ - 1 : 0, mD = sideI > 1 ? sideI - 1 : 0;
        const int mL = sideJ > 1 ? sideJ - 1 : 0, mU = sideI > 2 ? sideI - 2 : 0;
        int i = startI, j = startJ;
        visit_at(i, j);
        for (int o = 0; o < mR; o++) visit_at(i, ++j);
        for (int o = 0; o < mD; o++) visit_at(++i, j);
        if (mD > 0) {
            for (int o = 0; o < mL; o++) visit_at(i, --j);
            for (int o = 0; o < mU; o++) visit_at(--i, j);
        }
        sideI += 2;
        sideJ += 2;
        startJ--;
        startI--;
    }
}

int check_valid_config(const qlp_valid_config &c)
{
    if (c.n_stages < 0 || c.n_stages > QLP_VALID_MAX_STAGES)
        return fail(QLP_ERR_ARG, "validation: number of stages must be 0..8");
    for (int s = 0; s < c.n_stages; s++) {
        const qlp_valid_stage &st = c.stage[s];
        if (st.validator != QLP_VALIDATOR_MEDIAN && st.validator != QLP_VALIDATOR_DIFFERENCE)
            return fail(QLP_ERR_ARG, "Validator strategy cannot be null");
        if (st.replacement != QLP_REPLACE_NONE && st.replacement != QLP_REPLACE_SECONDARY_PEAK &&
            st.replacement != QLP_REPLACE_BILINEAR)
            return fail(QLP_ERR_ARG, "Unknown vector replacement method");
        if (st.replacement == QLP_REPLACE_SECONDARY_PEAK && st.peak_index < 1)
            return fail(QLP_ERR_ARG, "validation: secondary peak index must be >= 1");
    }
    return QLP_OK;
}

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t valid_work_bytes(int pairs, int n)
{
    const size_t pn = (size_t)pairs * n;
    return 2 * al256(pn * 4) + 3 * al256(pn * 4) + al256((size_t)pairs * 4);
}

void valid_work_carve(void *base, int pairs, int n, ValidLaunch &vl)
{
    const size_t pn = (size_t)pairs * n;
    char *b = (char *)base;
    vl.bu = (float *)b;
    b += al256(pn * 4);
    vl.bv = (float *)b;
    b += al256(pn * 4);
    vl.flags = (int *)b;
    b += al256(pn * 4);
    vl.done = (int *)b;
    b += al256(pn * 4);
    vl.act = (int *)b;
    b += al256(pn * 4);
    vl.act_count = (int *)b;
}

int launch_validate(Context *ctx, const qlp_valid_config &cfg, const ValidLaunch &vl)
{
    if (cfg.n_stages == 0 || vl.pairs == 0) return QLP_OK;
    VParams p;
    p.nI = vl.nI;
    p.nJ = vl.nJ;
    p.n = vl.nI * vl.nJ;
    p.P = vl.P;
    p.n_stages = cfg.n_stages;
    p.max_iter = cfg.max_iterations;
    p.until = cfg.until_no_more;
    for (int s = 0; s < QLP_VALID_MAX_STAGES; s++) {
        const qlp_valid_stage &st = cfg.stage[s < cfg.n_stages ? s : 0];
        p.st[s] = VStage{st.validator, st.threshold, st.epsilon0, st.replacement, st.peak_index};
    }
    p.top = vl.top;
    p.left = vl.left;
    p.rank = vl.rank;
    p.u = vl.u;
    p.v = vl.v;
    p.bu = vl.bu;
    p.bv = vl.bv;
    p.pk = vl.pk;
    p.pk_stride = vl.pk_stride;
    p.pk_off = vl.pk_off;
    p.pcount = vl.pcount;
    p.flags = vl.flags;
    p.done = vl.done;
    p.act = vl.act;
    p.act_count = vl.act_count;
    p.invalid = vl.invalid;
    p.stats = vl.stats;
    if (!vl.counts_zero)
        hipLaunchKernelGGL(zero_int_kernel, dim3((vl.pairs + 255) / 256), dim3(256), 0, ctx->stream,
                           vl.act_count, vl.pairs);
    const size_t total = (size_t)vl.pairs * p.n;
    hipLaunchKernelGGL(valid_init_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       ctx->stream, p, vl.pairs);
    hipLaunchKernelGGL(valid_sweep_kernel, dim3(vl.pairs), dim3(SWEEP_THREADS), 0, ctx->stream, p);
    return check_hip(hipGetLastError(), "validation kernel launch");
}

int launch_accumulate(Context *ctx, int n, int P, int D, const float *sub, float *u, float *v)
{
    if (n == 0) return QLP_OK;
    hipLaunchKernelGGL(accumulate_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, P, D,
                       sub, u, v);
    return check_hip(hipGetLastError(), "accumulate kernel launch");
}

}  // namespace qlpiv

using namespace qlpiv;

extern "C" {

int qlp_validate_vectors_dev(qlp_ctx *h, int pairs, int nI, int nJ, const int *top,
                             const int *left, float *u, float *v, int n_peaks,
                             const float *peak_uv, const int *peak_count,
                             const qlp_valid_config *cfg, uint8_t *invalid, int *stats)
try {
    if (!h || !cfg || !top || !left || !u || !v) return fail(QLP_ERR_ARG, "validate: null argument");
    if (pairs < 0 || nI < 1 || nJ < 1) return fail(QLP_ERR_ARG, "validate: bad grid");
    if (n_peaks < 1 || n_peaks > 16) return fail(QLP_ERR_ARG, "validate: number of peaks must be 1..16");
    int rc = check_valid_config(*cfg);
    if (rc) return rc;
    for (int s = 0; s < cfg->n_stages; s++)
        if (cfg->stage[s].replacement == QLP_REPLACE_SECONDARY_PEAK && !peak_uv)
            return fail(QLP_ERR_ARG, "validate: secondary peak replacement needs the peak displacements");
    if (pairs == 0) return QLP_OK;
    Context *ctx = ctx_of(h);
    const int n = nI * nJ;
    const size_t grid = al256((size_t)n * 4);
    const size_t need = 3 * grid + valid_work_bytes(pairs, n);
    if (need > ctx->vwork_bytes) {
        if (ctx->hbm_budget && ctx->hbm_in_use - ctx->vwork_bytes + need > ctx->hbm_budget)
            return fail(QLP_ERR_BUDGET, "request exceeds the context's HBM budget");
        if (ctx->vwork) {
            rc = check_hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
            if (rc) return rc;
            (void)hipFree(ctx->vwork);
            ctx->hbm_in_use -= ctx->vwork_bytes;
            ctx->vwork = nullptr;
            ctx->vwork_bytes = 0;
        }
        rc = check_hip(hipMalloc(&ctx->vwork, need), "hipMalloc validation workspace");
        if (rc) return rc;
        ctx->vwork_bytes = need;
        ctx->hbm_in_use += need;
    }
    std::vector<int> rank(n);
    spiral_rank(nI, nJ, rank.data());
    char *b = (char *)ctx->vwork;
    ValidLaunch vl;
    vl.pairs = pairs;
    vl.nI = nI;
    vl.nJ = nJ;
    vl.P = n_peaks;
    vl.top = (const int *)b;
    vl.left = (const int *)(b + grid);
    vl.rank = (const int *)(b + 2 * grid);
    // the grid arrays are pageable host memory: synchronous uploads, so the
    // workspace is not in use by an earlier call when it is overwritten
    rc = check_hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (!rc) rc = check_hip(hipMemcpy(b, top, (size_t)n * 4, hipMemcpyHostToDevice), "upload grid");
    if (!rc) rc = check_hip(hipMemcpy(b + grid, left, (size_t)n * 4, hipMemcpyHostToDevice), "upload grid");
    if (!rc) rc = check_hip(hipMemcpy(b + 2 * grid, rank.data(), (size_t)n * 4, hipMemcpyHostToDevice), "upload grid");
    if (rc) return rc;
    valid_work_carve(b + 3 * grid, pairs, n, vl);
    vl.u = u;
    vl.v = v;
    vl.pk = peak_uv;
    vl.pk_stride = 2;
    vl.pk_off = 0.0f;
    vl.pcount = peak_count;
    vl.invalid = invalid;
    vl.stats = stats;
    return launch_validate(ctx, *cfg, vl);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_validate_vectors(qlp_ctx *h, int pairs, int nI, int nJ, const int *top, const int *left,
                         float *u, float *v, int n_peaks, const float *peak_uv,
                         const int *peak_count, const qlp_valid_config *cfg, uint8_t *invalid,
                         int *stats)
try {
    if (!h || !cfg || !top || !left || !u || !v) return fail(QLP_ERR_ARG, "validate: null argument");
    if (pairs < 0 || nI < 1 || nJ < 1) return fail(QLP_ERR_ARG, "validate: bad grid");
    if (n_peaks < 1 || n_peaks > 16) return fail(QLP_ERR_ARG, "validate: number of peaks must be 1..16");
    if (pairs == 0) return QLP_OK;
    Context *ctx = ctx_of(h);
    const size_t pn = (size_t)pairs * nI * nJ;
    const size_t bu = al256(pn * 4), bk = al256(pn * n_peaks * 8), bc = al256(pn * 4),
                 bi = al256(pn), bs = al256((size_t)pairs * 16);
    void *d = nullptr;
    int rc = check_hip(hipMalloc(&d, 2 * bu + bk + bc + bi + bs), "hipMalloc validation buffers");
    if (rc) return rc;
    char *b = (char *)d;
    float *du = (float *)b, *dv = (float *)(b + bu), *dk = (float *)(b + 2 * bu);
    int *dc = (int *)(b + 2 * bu + bk);
    uint8_t *di = (uint8_t *)(b + 2 * bu + bk + bc);
    int *ds = (int *)(b + 2 * bu + bk + bc + bi);
    rc = check_hip(hipMemcpy(du, u, pn * 4, hipMemcpyHostToDevice), "copy u");
    if (!rc) rc = check_hip(hipMemcpy(dv, v, pn * 4, hipMemcpyHostToDevice), "copy v");
    if (!rc && peak_uv) rc = check_hip(hipMemcpy(dk, peak_uv, pn * n_peaks * 8, hipMemcpyHostToDevice), "copy peaks");
    if (!rc && peak_count) rc = check_hip(hipMemcpy(dc, peak_count, pn * 4, hipMemcpyHostToDevice), "copy counts");
    if (!rc)
        rc = qlp_validate_vectors_dev(h, pairs, nI, nJ, top, left, du, dv, n_peaks,
                                      peak_uv ? dk : nullptr, peak_count ? dc : nullptr, cfg, di, ds);
    if (!rc) rc = check_hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (!rc) rc = check_hip(hipMemcpy(u, du, pn * 4, hipMemcpyDeviceToHost), "copy u");
    if (!rc) rc = check_hip(hipMemcpy(v, dv, pn * 4, hipMemcpyDeviceToHost), "copy v");
    if (!rc && invalid) rc = check_hip(hipMemcpy(invalid, di, pn, hipMemcpyDeviceToHost), "copy invalid");
    if (!rc && stats) rc = check_hip(hipMemcpy(stats, ds, (size_t)pairs * 16, hipMemcpyDeviceToHost), "copy stats");
    (void)hipFree(d);
    return rc;
} catch (...) {
    return qlpiv::fail_exception();
}

}  // extern "C"
