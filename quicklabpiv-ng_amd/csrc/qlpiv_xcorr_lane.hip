// Correlation pass, second form: one lane owns one line of the 2N x 2N transform.
//
// Same arithmetic as qlpiv_xcorr.hip (the reference's radix-2 decimation-in-time
// butterflies, twiddle entries and operation order, no contraction), laid out
// for a 64-wide wavefront:
//   - a lane holds a whole row or column (M = 2N complex values, or half of one
//     for M = 128 where two lanes share a line) in registers and runs every
//     butterfly stage there; the bit-reversal shuffle is a compile-time register
//     renaming and the twiddle index of every butterfly is a compile-time
//     constant, so twiddles are wave-uniform scalar loads;
//   - LDS only transposes between the row and the column passes (one float2
//     tile per IA, odd row stride: conflict-free rows and columns);
//   - butterflies with twiddle index 0 (c = 1, s = -0 or +0 in all three tables)
//     are plain add/sub, which gives the same values;
//   - the product is computed by the lane that runs the inverse row transform of
//     row l, straight from the forward spectrum (X[l][o] and X[-l][-o]);
//   - for M = 128 the last stage pairs the two halves of a line across lanes
//     l and l ^ 32 (one shuffle per value);
//   - the FromCenter key is folded while the final columns are in registers.
// M = 32 runs two IAs per 64-lane work-group, M = 64 one IA per wavefront,
// M = 128 one IA per 256-lane work-group.  Work-groups are assigned to IAs in
// contiguous runs per XCD, so overlapping IAs share an L2.
#include <type_traits>
#include "qlpiv_device.h"
#include "qlpiv_internal.h"
#include "qlpiv_multipeak.h"

namespace qlpiv {

template <int LOGM>
struct Lx {
    static constexpr int M = 1 << LOGM;
    static constexpr int N = M / 2;
    static constexpr int D = M - 1;
    static constexpr int S = M + 1;                   // LDS row stride (float2)
    static constexpr int LPL = (LOGM >= 6) ? 2 : 1;   // lanes per line
    static constexpr int E = M / LPL;                 // elements per lane
    static constexpr int LOGE = LOGM - (LPL == 2 ? 1 : 0);
    static constexpr int IAS = (LOGM == 5) ? 2 : 1;   // IAs per work-group
    static constexpr int TPI = M * LPL;               // threads per IA
    static constexpr int THREADS = TPI * IAS;
    static constexpr int PIX = N * N / TPI;           // tile pixels per thread
};

constexpr int brev_c(int x, int bits)
{
    int r = 0;
    for (int b = 0; b < bits; b++) r |= ((x >> b) & 1) << (bits - 1 - b);
    return r;
}

__device__ __forceinline__ void bfly_w(float &ar, float &ai, float &br, float &bi, float2 w)
{
    const float pr = w.x * br, pi = w.x * bi, qi = w.y * bi, qr = w.y * br;
    const float lr = (ar + pr) - qi;
    const float li = (ai + pi) + qr;
    const float ur = (ar - pr) + qi;
    const float ui = (ai - pi) - qr;
    ar = lr;
    ai = li;
    br = ur;
    bi = ui;
}

// twiddle index 0: c = 1 and s = -0 or +0, so c*b = b exactly and the s terms
// only add or subtract a zero
__device__ __forceinline__ void bfly_0(float &ar, float &ai, float &br, float &bi)
{
    const float lr = ar + br, li = ai + bi, ur = ar - br, ui = ai - bi;
    ar = lr;
    ai = li;
    br = ur;
    bi = ui;
}

// Stages S0.. of the transform of the lane's elements (register e holds
// position half*E + e of the bit-reversed order); natural-order output, with
// register e holding X[half*E + e].
template <int LOGM, int S0>
__device__ __forceinline__ void fft_reg(float (&re)[Lx<LOGM>::E], float (&im)[Lx<LOGM>::E], int half,
                                        const float2 *__restrict__ tw)
{
    using X = Lx<LOGM>;
    constexpr int M = X::M, E = X::E, LOGE = X::LOGE;
    #pragma unroll
    for (int s = S0; s < LOGE; s++) {
        const int L = 1 << s;
        #pragma unroll
        for (int e = 0; e < E; e++) {
            if (e & L) continue;
            const int k = (M >> (s + 1)) * (e & (L - 1));
            if (k == 0)
                bfly_0(re[e], im[e], re[e + L], im[e + L]);
            else
                bfly_w(re[e], im[e], re[e + L], im[e + L], tw[k]);
        }
    }
    if (X::LPL == 2) {
        // last stage: position e (half 0) with position E + e (half 1), k = e
        #pragma unroll
        for (int e = 0; e < E; e++) {
            const float or_ = __shfl_xor(re[e], 32, 64), oi_ = __shfl_xor(im[e], 32, 64);
            float ar = half ? or_ : re[e], ai = half ? oi_ : im[e];
            float br = half ? re[e] : or_, bi = half ? im[e] : oi_;
            if (e == 0)
                bfly_0(ar, ai, br, bi);
            else
                bfly_w(ar, ai, br, bi, tw[e]);
            re[e] = half ? br : ar;
            im[e] = half ? bi : ai;
        }
    }
}

// original (natural) index held by register e of a lane in half 'half'
template <int LOGM>
__device__ __forceinline__ int orig_index(int e, int half)
{
    using X = Lx<LOGM>;
    if (X::LPL == 1) return brev_c(e, LOGM);
    return 2 * brev_c(e, LOGM - 1) + half;
}

template <int LOGM, bool MODE_CPU, bool FROM_IMAGE, bool PEAK, bool PERSIST = false>
__global__ __launch_bounds__(Lx<LOGM>::THREADS) void xcorr_lane_kernel(
    PassParams p, int nblocks, const float2 *__restrict__ tw_fwd, const float2 *__restrict__ tw_inv)
{
    using X = Lx<LOGM>;
    constexpr int M = X::M, N = X::N, D = X::D, S = X::S, E = X::E, TPI = X::TPI;
    constexpr int NW = (TPI + 63) / 64;  // wavefronts per IA
    __shared__ float2 s_x[X::IAS][M * S];
    __shared__ int s_red[X::IAS][NW > 1 ? NW : 1];
    __shared__ uint64_t s_key[X::IAS][NW > 1 ? NW : 1];
    __shared__ int s_cell[X::IAS][NW > 1 ? NW : 1];
    __shared__ float s_min[X::IAS][NW > 1 ? NW : 1];
    __shared__ float s_vmax[X::IAS][NW > 1 ? NW : 1];

    // One block of IAs per work-group: contiguous runs of work-groups per XCD
    // (work-groups are dealt round-robin to the 8 XCDs, so block b runs on XCD
    // b % 8).  A persistent launch (IA count in device memory) strides over the
    // blocks instead.
    const int n = PERSIST ? *p.n_dev : p.n;
    const int nbl = PERSIST ? (n + X::IAS - 1) / X::IAS : nblocks;
    const int per = (nbl + 7) >> 3;
    for (int it = 0;; it++) {
        int blk;
        if (PERSIST) {
            blk = (int)blockIdx.x + it * (int)gridDim.x;
        } else {
            if (it) break;
            blk = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
        }
        if (blk >= nbl) break;

        // persistent: the thread index is hidden from the optimiser in each
        // iteration, so the LDS addresses built on it are recomputed there
        // instead of being hoisted out of the loop into registers
        int tx = (int)threadIdx.x;
        if (PERSIST) asm volatile("" : "+v"(tx));
        const int sub = (X::IAS == 1) ? 0 : tx / TPI;
        const int t = (X::IAS == 1) ? tx : tx % TPI;
        const int ia = blk * X::IAS + sub;
        const bool valid = ia < n;
        int line, half;
        if (X::LPL == 1) {
            line = t;
            half = 0;
        } else {
            line = (t >> 6) * 32 + (t & 31);
            half = (t >> 5) & 1;
        }
        float2 *__restrict__ sx = s_x[sub];

        // ---- tile formation and normalisation ------------------------------------
        float va[X::PIX], vb[X::PIX];
        float mA = 16.0f, mB = 16.0f;
        int amax = 0, bmax = 0;
        bool warp = false;
        if (FROM_IMAGE) {
            if (valid) {
                const int per_pair = p.per_pair > 0 ? p.per_pair : n;
                const int pair = ia / per_pair, tt = ia - pair * per_pair;
                const size_t img_off = (size_t)pair * p.H * p.W;
                const uint8_t *__restrict__ imA = p.imgA + img_off;
                const uint8_t *__restrict__ imB = p.imgB + img_off;
                const int y0 = p.top[tt], x0 = p.left[tt];
                int by0 = y0, bx0 = x0;
                float fu = 0.0f, fv = 0.0f;
                warp = (p.bmode == B_WARP_MICRO) && p.du != nullptr;
                if (p.du != nullptr) {
                    const float u = p.du[ia], v = p.dv[ia];
                    if (p.bmode == B_CLIP_SHIFT) {
                        by0 = fm_round((float)y0 + u);
                        bx0 = fm_round((float)x0 + v);
                    } else if (warp) {
                        const float flu = floorf(u), flv = floorf(v);
                        by0 = (short)(y0 + (int)flu);
                        bx0 = (short)(x0 + (int)flv);
                        fu = u - flu;  // createMicroDisplacements :195-196
                        fv = v - flv;
                    }
                }
                // unconditional loads from addresses clamped into the frame
                // (32-bit offsets: the launcher bounds H * W below 2^31), masked
                // afterwards, one loop per B form: all loads in flight together
                const unsigned Hu = (unsigned)p.H, Wu = (unsigned)p.W;
                auto load = [&](auto warp_c) {
                    #pragma unroll
                    for (int k = 0; k < X::PIX; k++) {
                        const int idx = t + k * TPI;
                        const int r = idx / N, c = idx % N;
                        const unsigned ya = (unsigned)(y0 + r), xa = (unsigned)(x0 + c);
                        const bool ina = ya < Hu && xa < Wu;
                        int a = imA[ina ? ya * Wu + xa : 0u];
                        a = ina ? a : 0;
                        int b;
                        if constexpr (decltype(warp_c)::value) {
                            b = (int)warp_pixel<uint8_t, unsigned>(imB, p.H, p.W, fu, fv, by0 + r,
                                                                   bx0 + c);
                        } else {
                            const unsigned yb = (unsigned)(by0 + r), xb = (unsigned)(bx0 + c);
                            const bool inb = yb < Hu && xb < Wu;
                            b = imB[inb ? yb * Wu + xb : 0u];
                            b = inb ? b : 0;
                            bmax = max(bmax, b);
                        }
                        amax = max(amax, a);
                        va[k] = (float)a;
                        vb[k] = (float)b;
                    }
                };
                if (warp)
                    load(std::true_type{});
                else
                    load(std::false_type{});
            } else {
                #pragma unroll
                for (int k = 0; k < X::PIX; k++) va[k] = vb[k] = 0.0f;
            }
            // per-IA maxima: lanes of one IA are a half-wave, a wave or four waves
            #pragma unroll
            for (int o = (TPI < 64 ? TPI : 64) / 2; o >= 1; o >>= 1) {
                amax = max(amax, __shfl_xor(amax, o, 64));
                bmax = max(bmax, __shfl_xor(bmax, o, 64));
            }
            if (NW > 1) {
                if ((t & 63) == 0) s_red[sub][t >> 6] = amax | (bmax << 16);
                __syncthreads();
                amax = bmax = 0;
                for (int w = 0; w < NW; w++) {
                    amax = max(amax, s_red[sub][w] & 0xffff);
                    bmax = max(bmax, s_red[sub][w] >> 16);
                }
            }
            mA = (float)(amax < 16 ? 16 : amax);
            mB = (warp || p.bmax255) ? 255.0f : (float)(bmax < 16 ? 16 : bmax);  // doBiLinearWarp :298
            if (MODE_CPU) mB = mA;
        } else {
            if (valid) {
                mA = p.maxA[ia];
                mB = MODE_CPU ? mA : p.maxB[ia];
                const uint8_t *a8 = p.tA + (size_t)ia * N * N, *b8 = p.tB + (size_t)ia * N * N;
                #pragma unroll
                for (int k = 0; k < X::PIX; k++) {
                    const int idx = t + k * TPI;
                    va[k] = (float)a8[idx];
                    vb[k] = (float)b8[idx];
                }
            } else {
                #pragma unroll
                for (int k = 0; k < X::PIX; k++) va[k] = vb[k] = 0.0f;
            }
        }
        #pragma unroll
        for (int k = 0; k < X::PIX; k++) {
            const int idx = t + k * TPI;
            const int mi = N * N - 1 - idx;  // F'[r][c] = A[N-1-r][N-1-c]
            sx[(mi / N) * S + (mi % N)].x = va[k] / mA * 16.0f + 1.0f;
            sx[(idx / N) * S + (idx % N)].y = vb[k] / mB * 16.0f + 1.0f;
        }
        __syncthreads();

        float re[E], im[E];
        // ---- forward rows: the N data rows; odd positions pair with padding -----
        if (line < N) {
            #pragma unroll
            for (int e = 0; e < E; e += 2) {
                const float2 v = sx[line * S + orig_index<LOGM>(e, half)];
                re[e] = v.x;
                im[e] = v.y;
            }
            __builtin_amdgcn_sched_barrier(0);
            #pragma unroll
            for (int e = 0; e < E; e += 2) {
                re[e + 1] = re[e];
                im[e + 1] = im[e];
            }
            fft_reg<LOGM, 1>(re, im, half, tw_fwd);
            #pragma unroll
            for (int e = 0; e < E; e++) sx[line * S + half * E + e] = make_float2(re[e], im[e]);
        }
        __syncthreads();
        // ---- forward columns (rows >= N are zero) --------------------------------
        {
            const int c = line;
            #pragma unroll
            for (int e = 0; e < E; e += 2) {
                const float2 v = sx[orig_index<LOGM>(e, half) * S + c];
                re[e] = v.x;
                im[e] = v.y;
            }
            __builtin_amdgcn_sched_barrier(0);
            #pragma unroll
            for (int e = 0; e < E; e += 2) {
                re[e + 1] = re[e];
                im[e + 1] = im[e];
            }
            fft_reg<LOGM, 1>(re, im, half, tw_fwd);
            #pragma unroll
            for (int e = 0; e < E; e++) sx[(half * E + e) * S + c] = make_float2(re[e], im[e]);
        }
        __syncthreads();
        // ---- product and inverse rows --------------------------------------------
        {
            const int l = line, nl = (M - line) & (M - 1);
            #pragma unroll
            for (int e = 0; e < E; e++) {
                const int o = orig_index<LOGM>(e, half), no = (M - o) & (M - 1);
                const float2 a = sx[l * S + o], b = sx[nl * S + no];
                const float fr = 0.5f * (a.x + b.x);
                const float fi = 0.5f * (a.y - b.y);
                const float gr = 0.5f * (a.y + b.y);
                const float gi = -0.5f * (a.x - b.x);
                float pre = fr * gr - fi * gi;
                float pim = fr * gi + fi * gr;
                if (!MODE_CPU) {
                    pre = pre / (float)(M * M);
                    pim = pim / (float)(M * M);
                }
                re[e] = pre;
                im[e] = pim;
            }
        }
        __syncthreads();
        {
            if (MODE_CPU) {
                #pragma unroll
                for (int e = 0; e < E; e++) {
                    re[e] = re[e] / (float)M;
                    im[e] = im[e] / (float)M;
                }
            }
            fft_reg<LOGM, 0>(re, im, half, tw_inv);
            #pragma unroll
            for (int e = 0; e < E; e++) sx[line * S + half * E + e] = make_float2(re[e], im[e]);
        }
        __syncthreads();
        // ---- inverse columns, plane store, FromCenter ------------------------------
        const int c = line;
        const bool col_ok = valid && c < D;
        if (c < D) {
            #pragma unroll
            for (int e = 0; e < E; e++) {
                const float2 v = sx[orig_index<LOGM>(e, half) * S + c];
                re[e] = v.x;
                im[e] = v.y;
            }
            if (MODE_CPU) {
                #pragma unroll
                for (int e = 0; e < E; e++) {
                    re[e] = re[e] / (float)M;
                    im[e] = im[e] / (float)M;
                }
            }
            fft_reg<LOGM, 0>(re, im, half, tw_inv);
            if (col_ok && p.planes) {
                float *out = p.planes + (size_t)ia * D * D;
                #pragma unroll
                for (int e = 0; e < E; e++) {
                    const int row = half * E + e;
                    if (row < D) out[row * D + c] = re[e];
                }
            }
        }
        if constexpr (!PEAK && !PERSIST && X::IAS * X::TPI == 64 && X::LPL == 1) {
            if (p.mp_ent) {
                // MultiPeaks entries: the IA's columns into its LDS tile (its last
                // reads are done: one wavefront holds the work-group), then the
                // blocks from there by the IA's lanes
                if (c < D) {
                    #pragma unroll
                    for (int e = 0; e < E; e++)
                        if (e < D) sx[e * S + c].x = re[e];
                }
                mp_wave_sync();
                if (valid) {
                    constexpr MpGeom MG = mp_geom(D);
                    float mn = mp_entries_core<D, S, TPI>(
                        [&](int o) { return sx[o].x; }, t, p.mp_ent + (size_t)ia * MG.stride,
                        p.mp_cmax + (size_t)ia * MG.nchunk);
                    #pragma unroll
                    for (int o = TPI / 2; o >= 1; o >>= 1) {
                        const float om = __shfl_xor(mn, o, 64);
                        if (om < mn) mn = om;
                    }
                    if (t == 0) p.mp_min[ia] = mn;
                }
            }
        }
        if (PEAK) {
            // FromCenter in two phases (from_center_tie): the largest value of
            // the IA's visited cells, then the tie keys of the cells equal to it
            constexpr int h = D / 2;
            float vmax = -1.0f, mn = 3.40282347e+38f;
            if (col_ok && c >= 1) {
                #pragma unroll
                for (int e = 0; e < E; e++) {
                    const int row = half * E + e;
                    if (row < 1 || row >= D) continue;
                    const float v = re[e];
                    vmax = fmaxf(vmax, v);
                    if (v < mn) mn = v;
                }
            }
            #pragma unroll
            for (int o = (TPI < 64 ? TPI : 64) / 2; o >= 1; o >>= 1)
                vmax = fmaxf(vmax, __shfl_xor(vmax, o, 64));
            if (NW > 1) {
                if ((t & 63) == 0) s_vmax[sub][t >> 6] = vmax;
                __syncthreads();
                #pragma unroll
                for (int w = 0; w < NW; w++) vmax = fmaxf(vmax, s_vmax[sub][w]);
            }
            uint64_t best = 0;
            int bestCell = 0;
            if (col_ok && c >= 1 && vmax >= 0.0f) {
                #pragma unroll
                for (int e = 0; e < E; e++) {
                    const int row = half * E + e;
                    if (row < 1 || row >= D || !(re[e] == vmax)) continue;
                    const uint32_t tie = from_center_tie(row, c, h);
                    if (tie > best) {
                        best = tie;
                        bestCell = row * D + c;
                    }
                }
            }
            if (p.nbr && c < D) {
                #pragma unroll
                for (int e = 0; e < E; e++) {
                    const int row = half * E + e;
                    if (row < D) sx[row * S + c].x = re[e];
                }
            }
            #pragma unroll
            for (int o = (TPI < 64 ? TPI : 64) / 2; o >= 1; o >>= 1) {
                const uint64_t ob = shfl_xor_u64(best, o);
                const int oc = __shfl_xor(bestCell, o, 64);
                const float om = __shfl_xor(mn, o, 64);
                if (ob > best) {
                    best = ob;
                    bestCell = oc;
                }
                if (om < mn) mn = om;
            }
            if (NW > 1) {
                if ((t & 63) == 0) {
                    s_key[sub][t >> 6] = best;
                    s_cell[sub][t >> 6] = bestCell;
                    s_min[sub][t >> 6] = mn;
                }
            }
            __syncthreads();
            if (t == 0 && valid) {
                for (int w = 1; w < NW; w++) {
                    if (s_key[sub][w] > best) {
                        best = s_key[sub][w];
                        bestCell = s_cell[sub][w];
                    }
                    if (s_min[sub][w] < mn) mn = s_min[sub][w];
                }
                int pi = 0, pj = 0;
                float pv = 0.0f;
                if (best != 0) {  // key 0: no visited value >= 0, the reference keeps (0, 0, 0)
                    pi = bestCell / D;
                    pj = bestCell % D;
                    pv = sx[pi * S + pj].x;
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
                        nb[d + p.nbr_half] = (ri >= 0 && ri < D) ? sx[ri * S + pj].x : 0.0f;
                        nb[L + d + p.nbr_half] = (cj >= 0 && cj < D) ? sx[pi * S + cj].x : 0.0f;
                    }
                }
            }
        }
        if (!PERSIST) break;
        __syncthreads();  // the next block of IAs reuses the LDS tiles
    }
}

template <int LOGM, bool MODE_CPU>
static void launch_lane_mc(Context *ctx, const PassParams &p, bool from_image, bool peak)
{
    using X = Lx<LOGM>;
    const int li = LOGM - 5;
    const float2 *twf = ctx->tables.tw[li][TW_FWD];
    const float2 *twi = ctx->tables.tw[li][MODE_CPU ? TW_INV_CPU : TW_INV_GPU];
    const int nblocks = (p.n + X::IAS - 1) / X::IAS;
    dim3 grid(((nblocks + 7) / 8) * 8), block(X::THREADS);
    if (p.n_dev) {  // persistent: the step engine's passes (tiles, fused peak)
        auto k = xcorr_lane_kernel<LOGM, MODE_CPU, false, true, true>;
        hipLaunchKernelGGL(k, dim3(pass_grid(p, X::IAS, resident_groups(k, X::THREADS))), block, 0,
                           ctx->stream, p, nblocks, twf, twi);
        return;
    }
    if (from_image) {
        if (peak)
            hipLaunchKernelGGL((xcorr_lane_kernel<LOGM, MODE_CPU, true, true>), grid, block, 0,
                               ctx->stream, p, nblocks, twf, twi);
        else
            hipLaunchKernelGGL((xcorr_lane_kernel<LOGM, MODE_CPU, true, false>), grid, block, 0,
                               ctx->stream, p, nblocks, twf, twi);
    } else {
        if (peak)
            hipLaunchKernelGGL((xcorr_lane_kernel<LOGM, MODE_CPU, false, true>), grid, block, 0,
                               ctx->stream, p, nblocks, twf, twi);
        else
            hipLaunchKernelGGL((xcorr_lane_kernel<LOGM, MODE_CPU, false, false>), grid, block, 0,
                               ctx->stream, p, nblocks, twf, twi);
    }
}

template <int LOGM>
static void launch_lane_m(Context *ctx, const PassParams &p, bool from_image, bool peak)
{
    if (p.mode == QLP_NORM_A_MAX)
        launch_lane_mc<LOGM, true>(ctx, p, from_image, peak);
    else
        launch_lane_mc<LOGM, false>(ctx, p, from_image, peak);
}

int launch_pass_lane(Context *ctx, const PassParams &p, bool from_image, bool peak)
{
    switch (p.N) {
    case 16: launch_lane_m<5>(ctx, p, from_image, peak); break;
    case 32: launch_lane_m<6>(ctx, p, from_image, peak); break;
    case 64: launch_lane_m<7>(ctx, p, from_image, peak); break;
    default: return fail(QLP_ERR_UNSUPPORTED, "xcorr: IA size N must be 16, 32 or 64");
    }
    return check_hip(hipGetLastError(), "xcorr lane pass launch");
}

}  // namespace qlpiv
