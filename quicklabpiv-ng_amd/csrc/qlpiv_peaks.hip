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
// SIMPLE (maximum/FindMaximumSimple.java:19-48): see peak_simple_kernel.
//
// MULTI (maximum/FindMaximumMultiPeaks.java:42-162): the entries form (kernel
// width 3, at most 4 peaks) replays each plane's block maxima with one lane per
// plane, from entries the exact correlation pass writes or
// peak_multi_blocks_kernel finds; otherwise multipeak_segment
// (qlpiv_multipeak.h) with 8 to 64 lanes per plane.
#include <float.h>
#include <stdlib.h>
#include <string.h>

#include "qlpiv_device.h"
#include "qlpiv_internal.h"
#include "qlpiv_multipeak.h"

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

// SIMPLE (maximum/FindMaximumSimple.java:19-48): a row-major scan of every cell
// that keeps a strictly larger value, starting from 0; with no positive value
// the peak is (D/2, D/2) with value 0.  The first maximal cell in row-major
// order maximises the key (value bits, ~index) over cells with value > 0
// (positive floats order as their bit patterns; NaN never compares greater).
// minFloor is the minimum over every cell (NaN never compares smaller).
__global__ __launch_bounds__(256) void peak_simple_kernel(int n, int D,
                                                          const float *__restrict__ planes,
                                                          float *__restrict__ peaks,
                                                          float *__restrict__ min_floor)
{
    const int lane = threadIdx.x & 63;
    const int ia = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (ia >= n) return;
    const float *m = planes + (size_t)ia * D * D;
    uint64_t best = 0;
    float mn = FLT_MAX;
    for (int idx = lane; idx < D * D; idx += 64) {
        const float v = m[idx];
        if (v < mn) mn = v;
        if (v > 0.0f) {
            const uint64_t key = (uint64_t)__float_as_uint(v) << 32 | (uint32_t)~(uint32_t)idx;
            if (key > best) best = key;
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const uint64_t ob = shfl_xor_u64(best, o);
        const float om = __shfl_xor(mn, o, 64);
        if (ob > best) best = ob;
        if (om < mn) mn = om;
    }
    if (lane == 0) {
        float pi = (float)(D / 2), pj = (float)(D / 2), pv = 0.0f;
        if (best != 0) {
            const int cell = (int)~(uint32_t)best;
            pi = (float)(cell / D);
            pj = (float)(cell % D);
            pv = __uint_as_float((uint32_t)(best >> 32));
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

// W lanes per plane, 64 / W planes per wavefront (W = 64 for few planes, fewer
// lanes per plane when there are many, so that the sequential insertions of
// several planes share one instruction stream).
template <int KW, int W>
__global__ __launch_bounds__(64) void peak_multi_kernel(int n, int D, int P, int k, int lanes_list,
                                                        const float *__restrict__ planes,
                                                        float *__restrict__ peaks,
                                                        float *__restrict__ min_floor,
                                                        int *__restrict__ count)
{
    constexpr int S = 64 / W;  // planes per wavefront
    __shared__ MpList s_list[S];
    const int seg = (int)threadIdx.x / W;
    const int ia = blockIdx.x * S + seg;
    const bool act = ia < n;
    const size_t at = act ? (size_t)ia : 0;
    if (P <= W && lanes_list)  // the list in the segment's registers
        multipeak_segment_lanes<KW, W>(planes + at * D * D, D, 1, D, P, k, act, peaks + at * P * 3,
                                       min_floor + at, count ? count + at : nullptr);
    else
        multipeak_segment<KW, W>(planes + at * D * D, D, 1, D, P, k, act, s_list[seg],
                                 peaks + at * P * 3, min_floor + at, count ? count + at : nullptr);
}

// MultiPeaks for kernel width 3 and P <= 4 in two kernels over an entry
// workspace (multipeak_work_bytes): the entries of a plane are its blocks in
// the reference's visit order (slices of the top-left quadrant, then the four
// quadrants), each a block maximum (strict >, starting at 0: 0 for a block
// without a value above 0) and the position of the first cell holding it.
// Kernel 1 (planes in memory): one wavefront per plane.  The plane is copied
// into LDS with coalesced loads (consecutive lanes, consecutive cells), then its
// entries are formed from there as the exact passes form them (mp_entries_core:
// visit order, coalesced 8-byte entry stores, chunk maxima across 16 lanes),
// with the plane minimum at *min_floor.  The exact correlation passes write the
// same entries from the plane they still hold, and then only kernel 2 runs.
template <int DT>
__global__ __launch_bounds__(64) void peak_multi_blocks_kernel(int Dr,
                                                               const float *__restrict__ planes,
                                                               uint2 *__restrict__ ent,
                                                               float *__restrict__ cmax,
                                                               float *__restrict__ min_floor)
{
    extern __shared__ float s_pl[];
    static_assert(DT > 0, "compile-time plane size");
    constexpr int D = DT;
    constexpr MpGeom G = mp_geom(D);
    (void)Dr;
    const int lane = threadIdx.x, ia = blockIdx.x;
    const float *__restrict__ m = planes + (size_t)ia * D * D;
    constexpr int PER = (D * D + 63) / 64;
    float v[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int idx = lane + 64 * k;
        v[k] = m[idx < D * D ? idx : D * D - 1];
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int idx = lane + 64 * k;
        if (idx < D * D) s_pl[idx] = v[k];
    }
    mp_wave_sync();
    float mn = mp_entries_core<D, D, 64>([&](int o) { return s_pl[o]; }, lane,
                                         ent + (size_t)ia * G.stride, cmax + (size_t)ia * G.nchunk);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const float om = __shfl_xor(mn, o, 64);
        if (om < mn) mn = om;
    }
    if (lane == 0) min_floor[ia] = mn;
}

// Kernel 2: one lane per plane (blockDim.x planes per work-group) replays the
// plane's entries in visit order with its list in registers (mp_replay,
// mp_insert_reg).
// An entry changes the list only if the list has room or the entry is above the
// list's smallest value (addInAscendOrder :132-134), so a chunk whose maximum
// is not above that bound is skipped without reading its entries (the chunk
// maxima are staged in LDS first), and four entries are skipped together the
// same way.  A block without an update is inserted as 0 at the position its
// quadrant recorded last (FindMaximumMultiPeaks :62-64 keep maxsI/maxsJ across
// blocks): that position is looked up only when such an entry is replayed,
// walking back over the quadrant's entries not yet seen.
template <int PM>
__global__ __launch_bounds__(64) void peak_multi_replay_kernel(int n, int D, int P,
                                                               const uint2 *__restrict__ ent,
                                                               const float *__restrict__ cmax,
                                                               float *__restrict__ peaks,
                                                               int *__restrict__ count)
{
    extern __shared__ float s_cm[];
    const int ia = blockIdx.x * blockDim.x + threadIdx.x;
    float lv[PM];
    int lp[PM], size;
    mp_replay<PM>(n, D, P, ent, cmax, s_cm, lv, lp, size);
    if (ia < n) mp_store_peaks<PM>(ia, P, lv, lp, size, peaks, count);
}

bool multipeak_entries_form(int D, int n_peaks, int kernel_width)
{
    // QLPIV_MULTIPEAK_FORM = segment turns the form off everywhere (comparisons)
    if (const char *env = getenv("QLPIV_MULTIPEAK_FORM"))
        if (strcmp(env, "segment") == 0) return false;
    // kernel 1 holds one plane's entries in LDS
    return kernel_width == 3 && n_peaks >= 1 && n_peaks <= 4 && D >= 3 && D <= 255 &&
           (size_t)mp_geom(D).stride * sizeof(uint2) <= 65536;
}

size_t multipeak_work_bytes(int n, int D, int n_peaks, int kernel_width)
{
    if (n <= 0 || !multipeak_entries_form(D, n_peaks, kernel_width)) return 0;
    const MpGeom G = mp_geom(D);
    return (size_t)n * (G.stride * sizeof(uint2) + G.nchunk * sizeof(float));
}

void multipeak_work_carve(void *work, int n, int D, uint2 *&ent, float *&cmax)
{
    const MpGeom G = mp_geom(D);
    ent = (uint2 *)work;
    cmax = (float *)((char *)work + (size_t)n * G.stride * sizeof(uint2));
}

// Planes per work-group of the replay kernel (one lane each).  Measured on
// MI355X (C2 planes, P = 3; us for 4 / 8 / 16 / 32 / 64 planes): 65,025 x 31^2:
// 116 / 82 / 58 / 49 / 47; 16,129 x 63^2: 57 / 52 / 49 / 51 / 52; 3,969 x
// 127^2: 67 / 61 / 66 / 78 / 82.  QLPIV_MULTIPEAK_PLANES (1..64) overrides it.
static int replay_planes(int n)
{
    int G = n >= 32768 ? 64 : (n >= 8192 ? 16 : 8);
    if (const char *env = getenv("QLPIV_MULTIPEAK_PLANES")) {
        const int g = atoi(env);
        if (g >= 1 && g <= 64) G = g;
    }
    return G;
}

int launch_multipeak_replay(Context *ctx, int n, int D, int P, const uint2 *ent,
                            const float *cmax, float *peaks, int *count)
{
    if (n == 0) return QLP_OK;
    const MpGeom G = mp_geom(D);
    const int GW = replay_planes(n);
    const unsigned grid = (unsigned)((n + GW - 1) / GW);
    const size_t lds = (size_t)GW * G.nchunk * sizeof(float);
    switch (P) {
    case 1:
        hipLaunchKernelGGL(peak_multi_replay_kernel<1>, dim3(grid), dim3(GW), lds, ctx->stream, n,
                           D, P, ent, cmax, peaks, count);
        break;
    case 2:
        hipLaunchKernelGGL(peak_multi_replay_kernel<2>, dim3(grid), dim3(GW), lds, ctx->stream, n,
                           D, P, ent, cmax, peaks, count);
        break;
    case 3:
        hipLaunchKernelGGL(peak_multi_replay_kernel<3>, dim3(grid), dim3(GW), lds, ctx->stream, n,
                           D, P, ent, cmax, peaks, count);
        break;
    default:
        hipLaunchKernelGGL(peak_multi_replay_kernel<4>, dim3(grid), dim3(GW), lds, ctx->stream, n,
                           D, P, ent, cmax, peaks, count);
    }
    return check_hip(hipGetLastError(), "peak kernel launch");
}

// Both kernels over the context's workspace (planes in memory).
static int launch_entries_form(Context *ctx, int n, int D, int P, const float *planes,
                               float *peaks, float *min_floor, int *count)
{
    int rc = ensure_mpwork(ctx, multipeak_work_bytes(n, D, P, 3));
    if (rc) return rc;
    uint2 *ent;
    float *cm;
    multipeak_work_carve(ctx->mpwork, n, D, ent, cm);
    const size_t lds = (size_t)D * D * sizeof(float);
    switch (D) {
    case 31:
        hipLaunchKernelGGL(peak_multi_blocks_kernel<31>, dim3(n), dim3(64), lds, ctx->stream, D,
                           planes, ent, cm, min_floor);
        break;
    case 63:
        hipLaunchKernelGGL(peak_multi_blocks_kernel<63>, dim3(n), dim3(64), lds, ctx->stream, D,
                           planes, ent, cm, min_floor);
        break;
    case 127:
        hipLaunchKernelGGL(peak_multi_blocks_kernel<127>, dim3(n), dim3(64), lds, ctx->stream, D,
                           planes, ent, cm, min_floor);
        break;
    default:
        return fail(QLP_ERR_UNSUPPORTED, "find_peaks: entries form for 31, 63 and 127 only");
    }
    return launch_multipeak_replay(ctx, n, D, P, ent, cm, peaks, count);
}

template <int KW>
static void launch_multi(Context *ctx, int W, int n, int D, int P, int k, const float *planes,
                         float *peaks, float *min_floor, int *count)
{
    // the list in the segment's registers when it fits (P <= W) and the segment
    // is at most 32 lanes.  Measured on MI355X (C2 step, P = 3, k = 3; MultiPeaks
    // + Guo per level, ms, LDS list -> register list): 65,025 x 31^2 at W = 8:
    // 0.257 -> 0.219; 16,129 x 63^2 at W = 16: 0.191 -> 0.163; 3,969 x 127^2 at
    // W = 64: 0.140 -> 0.182.  QLPIV_MULTIPEAK_LIST (lds or lanes) overrides it.
    int lanes_list = W <= 32 ? 1 : 0;
    if (const char *env = getenv("QLPIV_MULTIPEAK_LIST")) {
        if (strcmp(env, "lds") == 0) lanes_list = 0;
        if (strcmp(env, "lanes") == 0) lanes_list = 1;
    }
    const unsigned grid = (unsigned)((n + 64 / W - 1) / (64 / W));
    switch (W) {
    case 8:
        hipLaunchKernelGGL((peak_multi_kernel<KW, 8>), dim3(grid), dim3(64), 0, ctx->stream, n, D,
                           P, k, lanes_list, planes, peaks, min_floor, count);
        break;
    case 16:
        hipLaunchKernelGGL((peak_multi_kernel<KW, 16>), dim3(grid), dim3(64), 0, ctx->stream, n,
                           D, P, k, lanes_list, planes, peaks, min_floor, count);
        break;
    case 32:
        hipLaunchKernelGGL((peak_multi_kernel<KW, 32>), dim3(grid), dim3(64), 0, ctx->stream, n,
                           D, P, k, lanes_list, planes, peaks, min_floor, count);
        break;
    default:
        hipLaunchKernelGGL((peak_multi_kernel<KW, 64>), dim3(grid), dim3(64), 0, ctx->stream, n,
                           D, P, k, lanes_list, planes, peaks, min_floor, count);
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
    } else if (kind == QLP_PEAK_SIMPLE) {
        hipLaunchKernelGGL(peak_simple_kernel, dim3((n + 3) / 4), dim3(256), 0, ctx->stream, n, D,
                           planes, peaks, min_floor);
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
        // planes in memory: the entries form (block maxima kernel, then the
        // replay) up to 63^2 planes, the segment form for 127^2, whose plane
        // does not leave room in LDS for more than two block wavefronts per
        // CU.  Measured on MI355X (C2 planes, us, segment -> blocks + replay):
        // 65,025 x 31^2 130 -> 51 + 46; 16,129 x 63^2 117 -> 62 + 49.
        // QLPIV_MULTIPEAK_FORM (segment or entries) overrides the choice.
        bool entries = D <= 63;
        if (const char *env = getenv("QLPIV_MULTIPEAK_FORM")) {
            if (strcmp(env, "segment") == 0) entries = false;
            if (strcmp(env, "entries") == 0) entries = true;
        }
        if (entries && (D == 31 || D == 63 || D == 127) &&
            multipeak_entries_form(D, n_peaks, kernel_width) &&
            mpwork_fits(ctx, multipeak_work_bytes(n, D, n_peaks, kernel_width)))
            return launch_entries_form(ctx, n, D, n_peaks, planes, peaks, min_floor, count);
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
