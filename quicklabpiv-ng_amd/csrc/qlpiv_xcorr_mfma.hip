// Correlation on the matrix cores: the throughput form of the correlation pass.
//
// The plane of one IA pair is the linear cross-correlation of the normalised
// tiles a, g (N x N):  R(p, q) = sum_x a(x) g(x + (p - (N-1), q - (N-1))),
// which the reference obtains by zero-padded FFTs of length M = 2N
// (CrossCorrelationRealFFTParallelKernel.java:214-713).  Here it is computed as
// dense products with a real trigonometric basis of M functions
//     h_r(x) = cos(2 pi r x / M)            r <  M/2
//     h_r(x) = cos(pi x)                    r == M/2
//     h_r(x) = sin(2 pi (r - M/2) x / M)   r >  M/2
// forward  E  = Hf a Hf^T   (Hf: M x N, the basis at x = 0 .. N-1),
// spectrum C  = the real-basis coefficients of conj(A^) G^ / M^2, formed
//               elementwise from the four (cos/sin x cos/sin) terms of every
//               frequency pair (k, l), (k, -l) of E_a and E_g,
// inverse  R  = Hout^T C Hout  (Hout: M x M, the basis at the lags p - (N-1)).
// That is 2 (M N^2 + M^2 N) + 2 M^3 multiply-adds per IA (0.92 M at N = 32),
// on MFMA in fp16 or bf16 with fp32 accumulation.  Split precision writes every
// operand as hi + lo (lo = x - hi, itself rounded) and accumulates
// hi*hi + hi*lo + lo*hi.  Power-of-two scalings keep fp16 operands in range;
// they are exact and undone on the fp32 result.
//
// One wavefront owns one IA (N = 16, 32) and the whole chain stays in its
// registers.  An MFMA contracts over the rows of an accumulator whether that
// accumulator enters as the A or the B operand (the K index of every product
// follows the row order of the accumulator layout, kphys, and the constant
// fragments are stored in that order); as the B operand its lane index stays
// the output column, as the A operand it becomes the output row.  The chain
//     P = x^T Hf^T  (pixels as A)     Q = Hf P = E^T  (P as B)
//     Y = C Hout    (spectrum as A)   R = Hout^T Y    (Y as B)
// therefore needs no transpose.  An A fragment and a B fragment of one matrix
// hold the same values per lane, so one table of Hf and Hout^T fragments
// serves both roles; it is staged in LDS once per work-group.  Operand
// fragments are kept as four packed 32-bit words (two 16-bit values each),
// converted in pairs.
#include "qlpiv_device.h"
#include "qlpiv_epilogue.h"
#include "qlpiv_internal.h"
#include "qlpiv_multipeak.h"

#include <math.h>
#include <string.h>

#include <type_traits>
#include <vector>

namespace qlpiv {

typedef _Float16 mh8 __attribute__((ext_vector_type(8)));
typedef __bf16 mb8 __attribute__((ext_vector_type(8)));
typedef float mf16 __attribute__((ext_vector_type(16)));
typedef float mf4 __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));  // one operand fragment
typedef unsigned int u2 __attribute__((ext_vector_type(2)));

constexpr int MF_WAVES = 4;  // wavefronts (IAs in flight) per work-group
#ifndef QLPIV_MF_MINW
#define QLPIV_MF_MINW 3  // work-groups per CU the register budget is sized for (16/32 px)
#endif

template <int N>
struct Mf {
    static constexpr int M = 2 * N, D = M - 1;
    static constexpr int MT = (M == 32) ? 16 : 32;  // MFMA tile: 16x16x32 or 32x32x16
    static constexpr int KB = (MT == 32) ? 16 : 32;  // K of one MFMA
    static constexpr int NS = MT * MT / 64;          // accumulator slots per lane
    static constexpr int RT = M / MT;                // tiles per matrix side (2)
    static constexpr int KN = (N + KB - 1) / KB;     // K blocks over the tile size
    static constexpr int KM = M / KB;                // K blocks over the transform size
    static constexpr int NF_HF = RT * KN;            // Hf fragments
    static constexpr int NF_HO = RT * KM;            // Hout^T fragments
    static constexpr int NF = NF_HF + NF_HO;
    // fp16 range: |a| <= 17, |E| <= 17 N^2 (kept in fp32), |C| <= 1156 N^2
    // before 2^-E1, |Y| <= M max|C| before 2^-E2: both stay below 2^15
    // (bounds of the normalised data).  Scaling C by the product of both
    // instead pushes its small coefficients into fp16 subnormals (measured:
    // 2e-6 instead of 7e-7 of the plane maximum).
    static constexpr int E1 = (N == 16) ? 4 : (N == 32 ? 6 : 8);
    // 2^-ES on each image's E gives their products 4 / M^2 / 2^E1
    static constexpr int ES = (N == 16) ? 6 : (N == 32 ? 8 : 10);
    static constexpr int E2 = (M == 32) ? 5 : (M == 64 ? 6 : 7);
};

template <int MT>
using Acc = typename std::conditional<MT == 32, mf16, mf4>::type;

// accumulator row of slot v for lane group g
template <int MT>
__host__ __device__ constexpr int mrow(int v, int g)
{
    return MT == 32 ? 8 * (v >> 2) + 4 * g + (v & 3) : 4 * g + v;
}

// K index held in element s of a B (and A) fragment of K block kb, lane group g
template <int MT>
__host__ __device__ constexpr int kphys(int kb, int g, int s)
{
    return MT == 32 ? 16 * kb + 8 * (s >> 2) + 4 * g + (s & 3)
                    : 32 * kb + 16 * (s >> 2) + 4 * g + (s & 3);
}

// An accumulator tile used as a B operand: slot v of row tile rt is element s
// of K block kb (the row it holds is kphys(kb, g, s)).
template <int MT>
__host__ __device__ constexpr int acc_kb(int rt, int v)
{
    return MT == 32 ? 2 * rt + (v >> 3) : rt >> 1;
}
template <int MT>
__host__ __device__ constexpr int acc_s(int rt, int v)
{
    return MT == 32 ? (v & 7) : 4 * (rt & 1) + v;
}

template <bool BF, int MT>
__device__ __forceinline__ Acc<MT> mma(u4 a, u4 b, Acc<MT> c)
{
    if constexpr (MT == 32) {
        if constexpr (BF)
            return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mb8, a),
                                                           __builtin_bit_cast(mb8, b), c, 0, 0, 0);
        else
            return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(mh8, a),
                                                          __builtin_bit_cast(mh8, b), c, 0, 0, 0);
    } else {
        if constexpr (BF)
            return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mb8, a),
                                                           __builtin_bit_cast(mb8, b), c, 0, 0, 0);
        else
            return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(mh8, a),
                                                          __builtin_bit_cast(mh8, b), c, 0, 0, 0);
    }
}

// x = hi + lo in the operand type (lo = x - hi, rounded), as 16-bit patterns
template <bool BF>
__device__ __forceinline__ void split1(float x, unsigned &h, unsigned &l)
{
    if constexpr (BF) {
        const __bf16 hh = (__bf16)x;
        h = __builtin_bit_cast(unsigned short, hh);
        l = __builtin_bit_cast(unsigned short, (__bf16)(x - (float)hh));
    } else {
        const _Float16 hh = (_Float16)x;
        h = __builtin_bit_cast(unsigned short, hh);
        l = __builtin_bit_cast(unsigned short, (_Float16)(x - (float)hh));
    }
}

typedef _Float16 mh2 __attribute__((ext_vector_type(2)));
typedef __bf16 mb2 __attribute__((ext_vector_type(2)));
typedef float mf2 __attribute__((ext_vector_type(2)));

// two values into packed hi and lo words (x0 in the low half): packed
// conversions, three instructions per value (two with MIX in fp16)
template <bool BF, bool MIX = true>
__device__ __forceinline__ void split2(float x0, float x1, unsigned &h, unsigned &l)
{
    const mf2 x = {x0, x1};
    if constexpr (BF) {
        const mb2 hi = __builtin_convertvector(x, mb2);
        const mb2 lo = __builtin_convertvector(x - __builtin_convertvector(hi, mf2), mb2);
        h = __builtin_bit_cast(unsigned, hi);
        l = __builtin_bit_cast(unsigned, lo);
    } else if constexpr (!MIX) {
        const mh2 hi = __builtin_convertvector(x, mh2);
        const mh2 lo = __builtin_convertvector(x - __builtin_convertvector(hi, mf2), mh2);
        h = __builtin_bit_cast(unsigned, hi);
        l = __builtin_bit_cast(unsigned, lo);
    } else {
        // x - hi as one mixed-precision fma per value (fp16 operand, exact
        // product by 1, one rounding: the same value as widening and subtracting)
        const mh2 hi = __builtin_convertvector(x, mh2);
        h = __builtin_bit_cast(unsigned, hi);
        float d0, d1;
        asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(d0) : "v"(h), "v"(x0));
        asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
            : "=v"(d1)
            : "v"(h), "v"(x1));
        const mf2 d = {d0, d1};
        l = __builtin_bit_cast(unsigned, __builtin_convertvector(d, mh2));
    }
}

// acc += A B (split: Ah Bh + Ah Bl + Al Bh)
template <bool BF, bool SPLIT, int MT>
__device__ __forceinline__ Acc<MT> mma3(u4 ah, u4 al, u4 bh, u4 bl, Acc<MT> c)
{
    if constexpr (SPLIT) {
        c = mma<BF, MT>(al, bh, c);
        c = mma<BF, MT>(ah, bl, c);
    }
    return mma<BF, MT>(ah, bh, c);
}

// f(integral_constant<I>) for I = B .. E-1
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// The FromCenter tie key (from_center_tie) and cell of the best of a lane's
// cells that equal the maximum: bit idx of mask marks cell idx, which cell()
// places at plane row a, column b.  The loop runs once per marked cell.
template <typename F>
__device__ __forceinline__ void best_tie(uint64_t mask, int h, int D, F cell, uint32_t &best,
                                         int &bestCell)
{
    while (mask) {
        const int idx = __builtin_ctzll(mask);
        mask &= mask - 1;
        int a, b;
        cell(idx, a, b);
        const uint32_t tie = from_center_tie(a, b, h);
        if (tie > best) {
            best = tie;
            bestCell = a * D + b;
        }
    }
}

struct MfmaParams {
    PassParams p;
    const u4 *frags;  // NF fragments x {hi, lo} x 64 lanes
};

template <int N, bool BF, bool SPLIT, bool FROM_IMAGE, bool PEAK>
__global__ __launch_bounds__(64 * MF_WAVES, QLPIV_MF_MINW) void xcorr_mfma_kernel(MfmaParams mp, int n_iter)
{
    using X = Mf<N>;
    static_assert(X::RT == 2, "one wavefront per IA holds 2 x 2 tiles");
    constexpr int D = X::D, MT = X::MT, NS = X::NS;
    const PassParams &p = mp.p;
    // the constant fragments, then one entry of row sums of Hf per lane
    // (then the Hout^T fragments again, scaled by 2^(E1 + E2): the inverse's
    // second product then returns the plane at its true scale)
    __shared__ u4 s_frag[(X::NF * 2 + 1 + X::NF_HO * 2) * 64];
    // MultiPeaks entries: 16 px, each wavefront's plane (3.8 KB); 32 px, a band
    // of six plane rows per wavefront (1.5 KB) and its chunk maxima
    __shared__ float s_plane[N == 16 ? MF_WAVES : 1][N == 16 ? D * D : 1];
    __shared__ float s_band[N == 32 ? MF_WAVES : 1][N == 32 ? 6 * D : 1];
    __shared__ uint32_t s_cmx[N == 32 ? MF_WAVES : 1][N == 32 ? 32 : 1];

    for (int k = threadIdx.x; k < (X::NF * 2 + 1 + X::NF_HO * 2) * 64; k += 64 * MF_WAVES)
        s_frag[k] = mp.frags[k];
    __syncthreads();

    // wave-uniform (scalar) wave index: the IA index and its global base
    // pointers are then scalar, and loads/stores take 32-bit lane offsets
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    // a persistent launch reads the IA count from device memory
    const int n = pass_count(p);
    if (p.n_dev) n_iter = ((n + MF_WAVES - 1) / MF_WAVES + (int)gridDim.x - 1) / (int)gridDim.x;

    for (int it = 0; it < n_iter; it++) {
        // The lane coordinates are made opaque once per iteration: everything
        // derived from them (LDS addresses of every slot, fragment offsets,
        // plane cells and their FromCenter order) is then recomputed from one
        // register plus constants instead of being hoisted out of the loop as
        // dozens of loop-invariant registers.
        // (The bound below keeps the lane's range known: the tile rows a lane
        // holds are then compile-time padding or not.)
        int lane = threadIdx.x & 63;
        asm volatile("" : "+v"(lane));
        __builtin_assume(lane >= 0 && lane < 64);
        const int j = lane % MT, g = lane / MT;
        // fragment f (hi / lo); f >= NF: the scaled Hout^T copy of f - NF, after the h row
        auto frag = [&](int f, int lo) -> u4 {
            return s_frag[(f * 2 + lo + (f >= X::NF ? 1 - X::NF_HF * 2 : 0)) * 64 + lane];
        };
        const int ia = (it * (int)gridDim.x + (int)blockIdx.x) * MF_WAVES + wave;
        const bool valid = ia < n;
        const int iac = valid ? ia : 0;  // the IA whose buffers an idle wave reads

        // ---- tiles in the A layout of x^T: element (kb, s) = row kphys, column j
        // Every load is unconditional, from an address clamped into its buffer,
        // and the value is masked afterwards: no branch separates the loads,
        // so all of them are in flight together.  An idle wave (no IA) works
        // on the buffers of IA 0 and stores nothing.
        float ta[X::KN][8], tg[X::KN][8];
        float mA = 16.0f, mB = 16.0f;
        if (FROM_IMAGE) {
            int amax = 0, bmax = 0;
            const int per_pair = p.per_pair > 0 ? p.per_pair : n;
            const int pair = iac / per_pair, tt = iac - pair * per_pair;
            const size_t img_off = (size_t)pair * p.H * p.W;
            const uint8_t *__restrict__ imA = p.imgA + img_off;
            const uint8_t *__restrict__ imB = p.imgB + img_off;
            const int y0 = p.top[tt], x0 = p.left[tt];
            int by0 = y0, bx0 = x0;
            float fu = 0.0f, fv = 0.0f;
            const bool warp = (p.bmode == B_WARP_MICRO) && p.du != nullptr;
            if (p.du != nullptr) {
                const float u = p.du[iac], v = p.dv[iac];
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
            // 32-bit offsets: the launcher bounds H * W below 2^31
            const unsigned H = (unsigned)p.H, Wd = (unsigned)p.W;
            // one loop per B form (the form is uniform; a branch inside the
            // loop would separate the loads)
            auto load = [&](auto warp_c) {
                constexpr bool WARP = decltype(warp_c)::value;
#pragma unroll
                for (int kb = 0; kb < X::KN; kb++)
#pragma unroll
                    for (int s = 0; s < 8; s++) {
                        const int r = kphys<MT>(kb, g, s);
                        const bool row = r < N;
                        const unsigned ya = (unsigned)(y0 + r), xa = (unsigned)(x0 + j);
                        const bool ina = row && ya < H && xa < Wd;
                        int a = imA[ina ? ya * Wd + xa : 0u];
                        a = ina ? a : 0;
                        int b;
                        if constexpr (WARP) {
                            b = (int)warp_pixel<uint8_t, unsigned>(imB, p.H, p.W, fu, fv, by0 + r,
                                                                   bx0 + j);
                            b = row ? b : 0;
                        } else {
                            const unsigned yb = (unsigned)(by0 + r), xb = (unsigned)(bx0 + j);
                            const bool inb = row && yb < H && xb < Wd;
                            b = imB[inb ? yb * Wd + xb : 0u];
                            b = inb ? b : 0;
                            bmax = max(bmax, b);
                        }
                        amax = max(amax, a);
                        ta[kb][s] = (float)a;
                        tg[kb][s] = (float)b;
                    }
            };
            if (warp)
                load(std::true_type{});
            else
                load(std::false_type{});
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                amax = max(amax, __shfl_xor(amax, o, 64));
                bmax = max(bmax, __shfl_xor(bmax, o, 64));
            }
            mA = (float)(amax < 16 ? 16 : amax);
            mB = (warp || p.bmax255) ? 255.0f : (float)(bmax < 16 ? 16 : bmax);
        } else {
            if (valid) {
                mA = p.maxA[ia];
                mB = p.maxB[ia];
            }
            const uint8_t *a8 = p.tA + (size_t)iac * N * N;
            const uint8_t *b8 = p.tB + (size_t)iac * N * N;
#pragma unroll
            for (int kb = 0; kb < X::KN; kb++)
#pragma unroll
                for (int s = 0; s < 8; s++) {
                    const int r = kphys<MT>(kb, g, s);
                    const bool in = r < N;
                    const unsigned off = in ? (unsigned)(r * N + j) : 0u;
                    const int a = a8[off], b = b8[off];
                    ta[kb][s] = in ? (float)a : 0.0f;
                    tg[kb][s] = in ? (float)b : 0.0f;
                }
        }
        if (p.mode == QLP_NORM_A_MAX) mB = mA;
        // the pixels as packed 16-bit operands (small integers: exact)
        u4 xt[2][X::KN];
#pragma unroll
        for (int kb = 0; kb < X::KN; kb++)
#pragma unroll
            for (int s = 0; s < 8; s += 2) {
                unsigned h, l;
                split2<BF>(ta[kb][s], ta[kb][s + 1], h, l);
                xt[0][kb][s >> 1] = h;
                split2<BF>(tg[kb][s], tg[kb][s + 1], h, l);
                xt[1][kb][s >> 1] = h;
            }

        // ---- forward: E = Hf x' Hf^T for x' = x * (16 / max) + 1, x = a, g ------
        // P = x'^T Hf^T with the pixels as the A operand (lane = tile column n,
        // K = tile row m).  The pixels are small integers, exact in fp16 and
        // bf16, so they need no lo part; the normalisation is affine and is
        // applied to P:  x'^T Hf^T = (16 / max) (x^T Hf^T + (max / 16) 1 h^T)
        // with h[rho] = sum_m Hf[rho][m], the second term added when P is
        // scaled.  Q = Hf P then takes P from the accumulator as its B
        // operand (K = n): Q[sigma][rho] = E[rho][sigma], lane = rho.
        // E leaves scaled by 2^-ES (folded into the scaling of P), so that the
        // spectrum's products carry the inverse's 4 / M^2 and the range
        // scaling 2^-E1 (4 / M^2 / 2^E1 = 2^-2ES).
        Acc<MT> W[2][2][2];  // [image][rt][ct]
        const float4 hrow = __builtin_bit_cast(float4, s_frag[X::NF * 2 * 64 + lane]);
#pragma unroll
        for (int im = 0; im < 2; im++) {
            const float mx = im == 0 ? mA : mB;
            // x / max * 16 + 1 (Matrix.copyMatrixToArrayAndNormalizeAndOffset
            // :175-181) as (x + max / 16) * (16 / max): one division per tile,
            // within an ulp or two of the reference's value, far below the
            // form's error
            const float sx = 16.0f / mx * (1.0f / (float)(1 << X::ES)), inv = mx * 0.0625f;
#pragma unroll
            for (int ct = 0; ct < 2; ct++) {
                // (x^T Hf^T + (max / 16) h) (16 / max) as one fma per value
                const float h0 = (ct == 0 ? hrow.x : hrow.y) * inv * sx;
                Acc<MT> acc = {};
#pragma unroll
                for (int kb = 0; kb < X::KN; kb++) {
                    if constexpr (SPLIT) acc = mma<BF, MT>(xt[im][kb], frag(ct * X::KN + kb, 1), acc);
                    acc = mma<BF, MT>(xt[im][kb], frag(ct * X::KN + kb, 0), acc);
                }
                u4 ph[X::KN], pl[X::KN];  // P's column tile ct as B fragments
#pragma unroll
                for (int kb = 0; kb < X::KN; kb++) ph[kb] = pl[kb] = (u4){0u, 0u, 0u, 0u};
#pragma unroll
                for (int v = 0; v < NS; v += 2) {
                    unsigned h, l;
                    split2<BF>(__builtin_fmaf(acc[v], sx, h0), __builtin_fmaf(acc[v + 1], sx, h0), h, l);
                    ph[acc_kb<MT>(0, v)][acc_s<MT>(0, v) >> 1] = h;
                    pl[acc_kb<MT>(0, v)][acc_s<MT>(0, v) >> 1] = l;
                }
#pragma unroll
                for (int rt = 0; rt < 2; rt++) {
                    Acc<MT> w = {};
#pragma unroll
                    for (int kb = 0; kb < X::KN; kb++)
                        w = mma3<BF, SPLIT, MT>(frag(rt * X::KN + kb, 0), frag(rt * X::KN + kb, 1),
                                                ph[kb], pl[kb], w);
                    W[im][rt][ct] = w;
                }
            }
        }

        // ---- spectrum: real-basis coefficients of conj(A^) G^ / M^2 ------------
        // W[.][rt][ct] at (lane, slot) = E[rho][sigma] with rho = MT ct + j
        // (row frequency k = j) and sigma = MT rt + row (column frequency l):
        // CC = [0][0], SC = [0][1], CS = [1][0], SS = [1][1].  Lane j = 0 holds
        // the two edge row frequencies k = 0 (ct 0) and M/2 (ct 1), which have
        // no sine partner, and slot 0 of lane group 0 the edge column
        // frequencies l = 0 (rt 0) and M/2 (rt 1).  There each output is the
        // general formula with A's terms substituted: the edge line's own pair
        // in the (CC, CS) or (CC, SC) roles and zeros for the missing partners.
        // The edge weights (1/2 per edge axis of 4 / M^2) are halvings of
        // Hout's rows 0 and M/2.
        u4 ch[2][X::KM], cl[2][X::KM];  // C as A fragments [ct][kb]
        const bool kk = (j == 0);
#pragma unroll
        for (int v0 = 0; v0 < NS; v0 += 2) {
            float o[2][2][2];  // [slot v0 + e][rt][ct]
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const int v = v0 + e;
                const bool ll = (v == 0) && (g == 0);  // mrow(v, g) == 0
                const bool edge = kk || ll;
                const float CC = W[0][0][0][v], SC = W[0][0][1][v], CS = W[0][1][0][v],
                            SS = W[0][1][1][v];
                const float cc = W[1][0][0][v], scg = W[1][0][1][v], cs = W[1][1][0][v],
                            ss = W[1][1][1][v];
                // A's (CC, CS, SC, SS) roles for the outputs 00, 10, 01, 11
                const float ssm = edge ? 0.0f : SS;
                const float a00[4] = {CC, ll ? 0.0f : CS, kk ? 0.0f : SC, ssm};
                const float a10[4] = {ll ? CS : CC, ll ? 0.0f : CS, kk ? 0.0f : (ll ? SS : SC), ssm};
                const float a01[4] = {kk ? SC : CC, ll ? 0.0f : (kk ? SS : CS), kk ? 0.0f : SC, ssm};
                const float a11[4] = {kk ? (ll ? SS : SC) : (ll ? CS : CC), ll ? 0.0f : (kk ? SS : CS),
                                      kk ? 0.0f : (ll ? SS : SC), ssm};
                o[e][0][0] = a00[0] * cc + a00[3] * ss + a00[2] * scg + a00[1] * cs;
                o[e][1][1] = a11[0] * ss + a11[3] * cc - a11[2] * cs - a11[1] * scg;
                o[e][0][1] = a01[0] * scg - a01[2] * cc + a01[1] * ss - a01[3] * cs;
                o[e][1][0] = a10[0] * cs - a10[1] * cc + a10[2] * ss - a10[3] * scg;
            }
#pragma unroll
            for (int rt = 0; rt < 2; rt++)
#pragma unroll
                for (int ct = 0; ct < 2; ct++) {
                    unsigned h, l;
                    split2<BF>(o[0][rt][ct], o[1][rt][ct], h, l);
                    ch[ct][acc_kb<MT>(rt, v0)][acc_s<MT>(rt, v0) >> 1] = h;
                    cl[ct][acc_kb<MT>(rt, v0)][acc_s<MT>(rt, v0) >> 1] = l;
                }
        }

        // ---- inverse: Y = C Hout, then R = Hout^T Y -------------------------------
        // C enters as the A operand (lane = rho, K = sigma), Y leaves with
        // lane = q and rho in the accumulator rows, which the second product
        // takes as its K index: no transpose.
        u4 yh[2][X::KM], yl[2][X::KM];  // Y as B fragments [qt][kb]
#pragma unroll
        for (int qt = 0; qt < 2; qt++)
#pragma unroll
            for (int kb = 0; kb < X::KM; kb++) yh[qt][kb] = yl[qt][kb] = (u4){0u, 0u, 0u, 0u};
#pragma unroll
        for (int ct = 0; ct < 2; ct++)
#pragma unroll
            for (int qt = 0; qt < 2; qt++) {
                Acc<MT> acc = {};
#pragma unroll
                for (int kb = 0; kb < X::KM; kb++) {
                    const int f = X::NF_HF + qt * X::KM + kb;
                    acc = mma3<BF, SPLIT, MT>(ch[ct][kb], cl[ct][kb], frag(f, 0), frag(f, 1), acc);
                }
#pragma unroll
                for (int v = 0; v < NS; v += 2) {
                    unsigned h, l;
                    split2<BF>(acc[v] * (1.0f / (float)(1 << X::E2)),
                               acc[v + 1] * (1.0f / (float)(1 << X::E2)), h, l);
                    yh[qt][acc_kb<MT>(ct, v)][acc_s<MT>(ct, v) >> 1] = h;
                    yl[qt][acc_kb<MT>(ct, v)][acc_s<MT>(ct, v) >> 1] = l;
                }
            }
        Acc<MT> R[2][2];
#pragma unroll
        for (int rt = 0; rt < 2; rt++)
#pragma unroll
            for (int qt = 0; qt < 2; qt++) {
                Acc<MT> acc = {};
#pragma unroll
                for (int kb = 0; kb < X::KM; kb++) {
                    const int f = X::NF + X::NF_HF + rt * X::KM + kb;  // the scaled copy
                    acc = mma3<BF, SPLIT, MT>(frag(f, 0), frag(f, 1), yh[qt][kb], yl[qt][kb], acc);
                }
                R[rt][qt] = acc;
            }

        // ---- epilogues: plane store, FromCenter -----------------------------------
        constexpr float unscale = 1.0f;  // carried by the scaled Hout^T copy
        if (valid && p.planes) {
            // row a = D (the last row of row tile 1) is only reached by the last
            // slot, column b = D only in column tile 1
            float *__restrict__ out = p.planes + (size_t)ia * D * D;
#pragma unroll
            for (int ct = 0; ct < 2; ct++) {
                const int b = MT * ct + j;
                if (ct == 0 || b < D) {
#pragma unroll
                    for (int rt = 0; rt < 2; rt++)
#pragma unroll
                        for (int v = 0; v < NS; v++) {
                            const int a = MT * rt + mrow<MT>(v, g);
                            if (rt == 0 || v < NS - 1 || a < D) out[(unsigned)(a * D + b)] = R[rt][ct][v] * unscale;
                        }
                }
            }
        }
        if constexpr (N == 16 && !PEAK) {
            if (p.mp_ent && valid) {
                // MultiPeaks entries: the wavefront's plane into its LDS slot,
                // then the blocks from there
                float *s_pl = s_plane[wave];
#pragma unroll
                for (int ct = 0; ct < 2; ct++) {
                    const int b = MT * ct + j;
#pragma unroll
                    for (int rt = 0; rt < 2; rt++)
#pragma unroll
                        for (int v = 0; v < NS; v++) {
                            const int a = MT * rt + mrow<MT>(v, g);
                            if (a < D && b < D) s_pl[a * D + b] = R[rt][ct][v] * unscale;
                        }
                }
                mp_wave_sync();
                constexpr MpGeom MG = mp_geom(D);
                float mn = mp_entries_core<D, D, 64>([&](int o) { return s_pl[o]; }, lane,
                                                     p.mp_ent + (size_t)ia * MG.stride,
                                                     p.mp_cmax + (size_t)ia * MG.nchunk);
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) {
                    const float om = __shfl_xor(mn, o, 64);
                    if (om < mn) mn = om;
                }
                if (lane == 0) p.mp_min[ia] = mn;
            }
        }
        if constexpr (N == 32 && !PEAK) {
            if (p.mp_ent && valid) {
                // MultiPeaks entries band by band: plane rows 6k .. 6k+5 (block
                // rows 2k, 2k+1 of both halves) into the wavefront's slot, then
                // those 2 x BW blocks by as many lanes; the chunk maxima by LDS
                // atomics (entries are >= 0, so their bit patterns order as the
                // values; the padding entries of the last chunk are not counted)
                constexpr MpGeom MG = mp_geom(D);
                static_assert(MG.nchunk <= 32 && 2 * MG.BW <= 64 && MG.stride - MG.nent <= 64,
                              "one lane per block of a band");
                float *slot = s_band[wave];
                uint32_t *cmx = s_cmx[wave];
                uint2 *__restrict__ ent = p.mp_ent + (size_t)ia * MG.stride;
                if (lane < MG.nchunk) cmx[lane] = 0u;
                if (MG.nent + lane < MG.stride)
                    ent[MG.nent + lane] = make_uint2(__float_as_uint(-2.0f), 0u);
                float mn = FLT_MAX;
#pragma unroll
                for (int k = 0; k < (D + 5) / 6; k++) {
                    // the lane holds rows a0 + 4 g of each slot: only the slots
                    // reaching this band are stored (a compile-time choice)
#pragma unroll
                    for (int rt = 0; rt < 2; rt++)
#pragma unroll
                        for (int v = 0; v < NS; v++) {
                            const int a0 = MT * rt + mrow<MT>(v, 0);
                            if (a0 + 4 < 6 * k || a0 >= 6 * k + 6) continue;
                            const int a = a0 + 4 * g;
#pragma unroll
                            for (int ct = 0; ct < 2; ct++) {
                                const int b = MT * ct + j;
                                if (a >= 6 * k && a < 6 * k + 6 && a < D && b < D)
                                    slot[(a - 6 * k) * D + b] = R[rt][ct][v] * unscale;
                            }
                        }
                    mp_wave_sync();
                    const int br = 2 * k + (lane >= MG.BW ? 1 : 0);
                    const int bc = lane - (lane >= MG.BW ? MG.BW : 0);
                    if (lane < 2 * MG.BW && br < MG.BW) {
                        // rows past the plane edge read the last row (in this band)
                        int o[9];
                        mp_block_offsets(D, D, br, bc, o);
                        float c[9];
#pragma unroll
                        for (int q = 0; q < 9; q++) c[q] = slot[o[q] - 6 * k * D];
                        float bmin;
                        const uint2 en = mp_block_vals(D, br, bc, c, bmin);
                        if (bmin < mn) mn = bmin;
                        const int e = mp_entry_index(MG, br, bc);
                        ent[e] = en;
                        atomicMax(&cmx[e >> 4], en.x);
                    }
                    mp_wave_sync();  // the slot's reads before the next band's stores
                }
                if (lane < MG.nchunk) p.mp_cmax[(size_t)ia * MG.nchunk + lane] = __uint_as_float(cmx[lane]);
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) {
                    const float om = __shfl_xor(mn, o, 64);
                    if (om < mn) mn = om;
                }
                if (lane == 0) p.mp_min[ia] = mn;
            }
        }
        if (PEAK) {
            // FromCenter in two phases (from_center_tie) on the accumulator
            // scale (unscale is a power of two: order and equality are kept).
            // Visited cells are rows and columns 1 .. D-1: column b = MT ct + j
            // leaves out j = 0 of tile 0 and j = MT-1 of tile 1 (b = D); row
            // a = MT rt + mrow(v, g) is 0 only in slot (0, 0) of group 0 and D
            // only in slot (1, NS-1) of the last group.
            constexpr int h = D / 2, G = 64 / MT;
            const bool colok[2] = {j != 0, j != MT - 1};
            const bool row0ok = g != 0, rowDok = g != G - 1;
            auto visited = [&](int rt, int ct, int v) -> bool {
                bool ok = colok[ct];
                if (rt == 0 && v == 0) ok = ok && row0ok;
                if (rt == 1 && v == NS - 1) ok = ok && rowDok;
                return ok;
            };
            float vmax = -1.0f, mn = 3.40282347e+38f;
#pragma unroll
            for (int rt = 0; rt < 2; rt++)
#pragma unroll
                for (int ct = 0; ct < 2; ct++)
#pragma unroll
                    for (int v = 0; v < NS; v++) {
                        const bool ok = visited(rt, ct, v);
                        const float x = R[rt][ct][v];
                        vmax = fmaxf(vmax, ok ? x : -1.0f);
                        mn = fminf(mn, ok ? x : 3.40282347e+38f);
                    }
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                vmax = fmaxf(vmax, __shfl_xor(vmax, o, 64));
                mn = fminf(mn, __shfl_xor(mn, o, 64));
            }
            // phase 2: each lane marks its cells equal to the maximum (almost
            // always none or one) and computes the tie keys of those only
            uint32_t best = 0;
            int bestCell = 0;
            if (vmax >= 0.0f) {
                uint64_t mask = 0;
#pragma unroll
                for (int rt = 0; rt < 2; rt++)
#pragma unroll
                    for (int ct = 0; ct < 2; ct++)
#pragma unroll
                        for (int v = 0; v < NS; v++) {
                            const bool m = visited(rt, ct, v) && R[rt][ct][v] == vmax;
                            mask |= m ? (uint64_t)1 << ((rt * 2 + ct) * NS + v) : 0;
                        }
                best_tie(mask, h, D,
                         [&](int idx, int &a, int &b) {
                             const int v = idx % NS, ct = (idx / NS) & 1, rt = idx / (2 * NS);
                             a = MT * rt + mrow<MT>(v, g);
                             b = MT * ct + j;
                         },
                         best, bestCell);
            }
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                const uint32_t ob = (uint32_t)__shfl_xor((int)best, o, 64);
                const int oc = __shfl_xor(bestCell, o, 64);
                if (ob > best) {
                    best = ob;
                    bestCell = oc;
                }
            }
            // the winner holds vmax; the neighbourhood: the row through the peak
            // sits in one slot (uniform) of one lane group, the column through
            // it in the G lanes of one column
            const int pc = __builtin_amdgcn_readfirstlane(best != 0 ? bestCell : 0);
            const int pi = pc / D, pj = pc % D;
            const float pv = best != 0 ? vmax * unscale : 0.0f;
            if (valid && p.nbr) {
                const int nh = p.nbr_half, L = 2 * nh + 1;
                float *nb = p.nbr + (size_t)ia * 2 * L;
                const int rts = pi / MT, ms = pi % MT;
                const int gs = (MT == 32) ? ((ms >> 2) & 1) : (ms >> 2);
                const int vs = (MT == 32) ? (((ms >> 3) << 2) | (ms & 3)) : (ms & 3);
                float rv[2] = {0.0f, 0.0f};
#pragma unroll
                for (int rt = 0; rt < 2; rt++)
#pragma unroll
                    for (int v = 0; v < NS; v++)
                        if (rt == rts && v == vs) {
                            rv[0] = R[rt][0][v];
                            rv[1] = R[rt][1][v];
                        }
#pragma unroll
                for (int ct = 0; ct < 2; ct++) {
                    const int b = MT * ct + j, dj = b - pj;
                    if (g == gs && b < D && dj >= -nh && dj <= nh) nb[L + dj + nh] = rv[ct] * unscale;
                }
                const int cts = pj / MT;
                if (j == pj % MT) {
#pragma unroll
                    for (int rt = 0; rt < 2; rt++)
#pragma unroll
                        for (int ct = 0; ct < 2; ct++) {
                            if (ct != cts) continue;
#pragma unroll
                            for (int v = 0; v < NS; v++) {
                                const int a = MT * rt + mrow<MT>(v, g), di = a - pi;
                                if (a < D && di >= -nh && di <= nh) nb[di + nh] = R[rt][ct][v] * unscale;
                            }
                        }
                }
            }
            if (valid && lane == 0) {
                if (p.peaks) {
                    p.peaks[(size_t)ia * 3 + 0] = (float)pi;
                    p.peaks[(size_t)ia * 3 + 1] = (float)pj;
                    p.peaks[(size_t)ia * 3 + 2] = pv;
                }
                if (p.min_floor) p.min_floor[ia] = mn * unscale;
            }
            // neighbourhood cells outside the plane are 0 (as in the exact pass)
            if (valid && p.nbr && lane < 2 * p.nbr_half + 1) {
                const int L = 2 * p.nbr_half + 1;
                float *nb = p.nbr + (size_t)ia * 2 * L;
                const int d = lane - p.nbr_half;
                if (pi + d < 0 || pi + d >= D) nb[lane] = 0.0f;
                if (pj + d < 0 || pj + d >= D) nb[L + lane] = 0.0f;
            }
        }
    }
}

// ---- N = 64: four wavefronts per IA ------------------------------------------
//
// At M = 128 every matrix of the chain is 4 x 4 tiles of 32 x 32, too many for
// one wavefront's registers.  The four wavefronts of a work-group share one IA:
// wave (sr, sc) owns the spectrum tiles sigma in {sr, sr + 2} (column
// frequency) x rho in {sc, sc + 2} (row frequency).  Basis tile t + 2 holds the
// sines of tile t's cosines (tile 2 starts with cos(pi x), the partner of
// k = 0), so each wave holds whole (cos/sin x cos/sin) quads and forms its part
// of the spectrum alone.  The chain is the one-wave kernel's (no transposes):
//   tiles : wave w forms two of the eight x^T A fragments of each image
//           (n tile sr, K blocks 2 sc, 2 sc + 1) -> LDS; every wave reads all
//   P = x^T Hf^T  for its rho tiles (both n tiles; the wave pairs with the same
//                 sc compute the same P)          Q = Hf P = E^T  (its quad)
//   C             its quad -> LDS as A fragments (rho tile, sigma K block)
//   Y = C Hout    rho tiles {sc, sc + 2} x q tiles {sr, sr + 2}, K = all sigma
//                 (its own and its partner's C)  -> LDS as B fragments
//   R = Hout^T Y  p tiles {sc, sc + 2} x q tiles {sr, sr + 2}, K = all rho
// The exchanges move whole 16-byte fragments (one LDS instruction per
// fragment and lane).  The constant fragments (Hf: 4 x 4, Hout^T: 4 x 8) are
// read from global memory (L2-resident).  Work-groups loop over the IAs.
namespace m64 {
constexpr int N = 64, D = 127, MT = 32, NS = 16, KN = 4, KM = 8;
constexpr int NT = 256;
constexpr int NF_HF = 16;       // Hf fragments first, then Hout^T (32)
constexpr int XCH = 64 * 64;    // u4 entries of the exchange buffer (64 KB)
}  // namespace m64

template <bool BF, bool SPLIT, bool FROM_IMAGE, bool PEAK>
__global__ __launch_bounds__(m64::NT, 2) void xcorr_mfma64_kernel(PassParams p,
                                                                  const u4 *__restrict__ frags)
{
    using namespace m64;
    using X = Mf<64>;
    // tile fragments, then C, then Y fragments, then (PEAK) the fit neighbourhood
    __shared__ u4 s_x[XCH];
    __shared__ uint32_t s_key[4];
    __shared__ int s_cell[4];
    __shared__ float s_min[4];
    __shared__ float s_vmax[4];
    __shared__ int s_max[2][4];

    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    const int sr = wave >> 1, sc = wave & 1;

    const int n = pass_count(p);
    for (int ia = blockIdx.x; ia < n; ia += gridDim.x) {
        int lane = threadIdx.x & 63;
        asm volatile("" : "+v"(lane));  // see xcorr_mfma_kernel
        __builtin_assume(lane >= 0 && lane < 64);
        const int j = lane % MT, g = lane / MT;
        auto frag = [&](int f, int lo) -> u4 { return frags[(f * 2 + lo) * 64 + lane]; };
        // Hout^T fragment f scaled by 2^(E1 + E2) (the table's copy after the h row)
        auto fscaled = [&](int f, int lo) -> u4 {
            return frags[(X::NF * 2 + 1 + (f - NF_HF) * 2 + lo) * 64 + lane];
        };
        auto xs = [&](int e) -> u4 & { return s_x[e * 64 + lane]; };

        // ---- pixel sources (as in the other forms) ------------------------------
        const uint8_t *a8 = nullptr, *b8 = nullptr;
        int y0 = 0, x0 = 0, by0 = 0, bx0 = 0;
        float fu = 0.0f, fv = 0.0f;
        bool warp = false;
        if (FROM_IMAGE) {
            const int per_pair = p.per_pair > 0 ? p.per_pair : n;
            const int pair = ia / per_pair, tt = ia - pair * per_pair;
            const size_t img_off = (size_t)pair * p.H * p.W;
            a8 = p.imgA + img_off;
            b8 = p.imgB + img_off;
            y0 = p.top[tt];
            x0 = p.left[tt];
            by0 = y0;
            bx0 = x0;
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
        } else {
            a8 = p.tA + (size_t)ia * N * N;
            b8 = p.tB + (size_t)ia * N * N;
        }

        // ---- tiles: this wave's x^T A fragments (n tile sr, K blocks 2 sc + kq):
        // lane (i, g) element s = x[kphys(kb, g, s)][32 sr + i].  Loads are
        // unconditional from clamped addresses (32-bit offsets: the launcher
        // bounds H * W below 2^31), one loop per B form.
        {
            const unsigned Hu = (unsigned)p.H, Wu = (unsigned)p.W;
            int mx[2] = {0, 0};
            float tv[2][2][8];  // [image][kq][s]
            // offsets first, then every load, then the masks: the loads issue
            // back to back
            auto load = [&](auto warp_c) {
                const uint8_t *__restrict__ src[2] = {a8, b8};
                unsigned off[2][2][8];
                bool in[2][2][8];
                int raw[2][2][8];
#pragma unroll
                for (int kq = 0; kq < 2; kq++)
#pragma unroll
                    for (int s = 0; s < 8; s++) {
                        const int r = kphys<MT>(2 * sc + kq, g, s), c = MT * sr + j;
#pragma unroll
                        for (int im = 0; im < 2; im++) {
                            if (!FROM_IMAGE) {
                                off[im][kq][s] = (unsigned)(r * N + c);
                                in[im][kq][s] = true;
                            } else {
                                const unsigned y = (unsigned)((im == 0 ? y0 : by0) + r);
                                const unsigned x = (unsigned)((im == 0 ? x0 : bx0) + c);
                                in[im][kq][s] = y < Hu && x < Wu;
                                off[im][kq][s] = in[im][kq][s] ? y * Wu + x : 0u;
                            }
                        }
                    }
#pragma unroll
                for (int kq = 0; kq < 2; kq++)
#pragma unroll
                    for (int s = 0; s < 8; s++)
#pragma unroll
                        for (int im = 0; im < 2; im++) {
                            if (FROM_IMAGE && decltype(warp_c)::value && im == 1) {
                                const int r = kphys<MT>(2 * sc + kq, g, s), c = MT * sr + j;
                                raw[im][kq][s] = (int)warp_pixel<uint8_t, unsigned>(
                                    b8, p.H, p.W, fu, fv, by0 + r, bx0 + c);
                                in[im][kq][s] = true;
                            } else {
                                raw[im][kq][s] = src[im][off[im][kq][s]];
                            }
                        }
#pragma unroll
                for (int kq = 0; kq < 2; kq++)
#pragma unroll
                    for (int s = 0; s < 8; s++)
#pragma unroll
                        for (int im = 0; im < 2; im++) {
                            const int v = in[im][kq][s] ? raw[im][kq][s] : 0;
                            mx[im] = max(mx[im], v);
                            tv[im][kq][s] = (float)v;
                        }
            };
            if (warp)
                load(std::true_type{});
            else
                load(std::false_type{});
#pragma unroll
            for (int im = 0; im < 2; im++) {
#pragma unroll
                for (int kq = 0; kq < 2; kq++) {
                    u4 f;
#pragma unroll
                    for (int s = 0; s < 8; s += 2) {
                        unsigned h, l;
                        split2<BF>(tv[im][kq][s], tv[im][kq][s + 1], h, l);  // exact: l == 0
                        f[s >> 1] = h;
                    }
                    xs(im * 8 + sr * 4 + 2 * sc + kq) = f;
                }
                if (FROM_IMAGE) {
#pragma unroll
                    for (int o = 32; o >= 1; o >>= 1) mx[im] = max(mx[im], __shfl_xor(mx[im], o, 64));
                    if (lane == 0) s_max[im][wave] = mx[im];
                }
            }
        }
        __syncthreads();  // the tile fragments and the partial maxima
        float mA, mB;
        if (FROM_IMAGE) {
            const int am = max(max(s_max[0][0], s_max[0][1]), max(s_max[0][2], s_max[0][3]));
            const int bm = max(max(s_max[1][0], s_max[1][1]), max(s_max[1][2], s_max[1][3]));
            mA = (float)(am < 16 ? 16 : am);
            mB = (warp || p.bmax255) ? 255.0f : (float)(bm < 16 ? 16 : bm);
        } else {
            mA = p.maxA[ia];
            mB = p.maxB[ia];
        }
        if (p.mode == QLP_NORM_A_MAX) mB = mA;

        // ---- forward: P = x'^T Hf^T (rho tiles sc, sc + 2), Q = Hf P (quad) ----
        // the normalisation folded as in the one-wave kernel; E leaves scaled
        // by 2^-ES
        Acc<MT> W[2][2][2];  // [image][ri: sigma tile sr + 2 ri][ci: rho tile sc + 2 ci]
        const float4 hrow = __builtin_bit_cast(float4, frags[X::NF * 2 * 64 + lane]);
        const float hv[4] = {hrow.x, hrow.y, hrow.z, hrow.w};
#pragma unroll
        for (int im = 0; im < 2; im++) {
            const float mx = im == 0 ? mA : mB;
            const float sx = 16.0f / mx * (1.0f / (float)(1 << X::ES)), inv = mx * 0.0625f;
#pragma unroll
            for (int ci = 0; ci < 2; ci++) {
                const int ct = sc + 2 * ci;
                const float h0 = hv[ct] * inv * sx;
                u4 ph[KN], pl[KN];  // P's rho tile ct as B fragments (K = n)
                {
                    // K block outermost (both n tiles per constant fragment);
                    // block kb + 1's fragment is loaded while block kb's run
                    Acc<MT> acc[2] = {};
                    u4 fh = frag(ct * KN, 0), fl = frag(ct * KN, 1);
#pragma unroll
                    for (int kb = 0; kb < KN; kb++) {
                        u4 nh = fh, nl = fl;
                        if (kb + 1 < KN) {
                            nh = frag(ct * KN + kb + 1, 0);
                            nl = frag(ct * KN + kb + 1, 1);
                        }
#pragma unroll
                        for (int nt = 0; nt < 2; nt++) {
                            const u4 xa = xs(im * 8 + nt * 4 + kb);  // tile fragment (n tile, K block)
                            if constexpr (SPLIT) acc[nt] = mma<BF, MT>(xa, fl, acc[nt]);
                            acc[nt] = mma<BF, MT>(xa, fh, acc[nt]);
                        }
                        fh = nh;
                        fl = nl;
                    }
#pragma unroll
                    for (int nt = 0; nt < 2; nt++)
#pragma unroll
                        for (int v = 0; v < NS; v += 2) {
                            unsigned h, l;
                            split2<BF, false>(__builtin_fmaf(acc[nt][v], sx, h0),
                                              __builtin_fmaf(acc[nt][v + 1], sx, h0), h, l);
                            ph[acc_kb<MT>(nt, v)][acc_s<MT>(nt, v) >> 1] = h;
                            pl[acc_kb<MT>(nt, v)][acc_s<MT>(nt, v) >> 1] = l;
                        }
                }
                {
                    Acc<MT> w[2] = {};
                    u4 fr[2][2];  // [ri][hi/lo] of block kb
#pragma unroll
                    for (int ri = 0; ri < 2; ri++)
#pragma unroll
                        for (int hl = 0; hl < 2; hl++) fr[ri][hl] = frag((sr + 2 * ri) * KN, hl);
#pragma unroll
                    for (int kb = 0; kb < KN; kb++) {
                        u4 nx[2][2];
#pragma unroll
                        for (int ri = 0; ri < 2; ri++)
#pragma unroll
                            for (int hl = 0; hl < 2; hl++)
                                nx[ri][hl] = kb + 1 < KN ? frag((sr + 2 * ri) * KN + kb + 1, hl)
                                                         : fr[ri][hl];
#pragma unroll
                        for (int ri = 0; ri < 2; ri++)
                            w[ri] = mma3<BF, SPLIT, MT>(fr[ri][0], fr[ri][1], ph[kb], pl[kb], w[ri]);
#pragma unroll
                        for (int ri = 0; ri < 2; ri++)
#pragma unroll
                            for (int hl = 0; hl < 2; hl++) fr[ri][hl] = nx[ri][hl];
                    }
#pragma unroll
                    for (int ri = 0; ri < 2; ri++) W[im][ri][ci] = w[ri];
                }
                // keep the scheduler from hoisting the next rho tile's constant
                // loads over this one (their registers spill)
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        __syncthreads();  // every wave has read the tile fragments

        // ---- spectrum of the quad (the one-wave kernel's formulas) -> LDS -------
        // C A fragment (rho tile ct, sigma K block kb8, hi/lo) at entry
        // (ct * KM + kb8) * 2 + hl; this wave's K blocks: 2 rt + (v >> 3)
        {
            const bool kk = (sc == 0) && (j == 0);
            u4 ch[2][4], cl[2][4];  // [ci][kq = 2 ri + (v >> 3)]
#pragma unroll
            for (int v0 = 0; v0 < NS; v0 += 2) {
                float o[2][2][2];  // [slot v0 + e][ri][ci]
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    const int v = v0 + e;
                    const bool ll = (sr == 0) && (v == 0) && (g == 0);  // mrow(v, g) == 0
                    const bool edge = kk || ll;
                    const float CC = W[0][0][0][v], SC = W[0][0][1][v], CS = W[0][1][0][v],
                                SS = W[0][1][1][v];
                    const float cc = W[1][0][0][v], scg = W[1][0][1][v], cs = W[1][1][0][v],
                                ss = W[1][1][1][v];
                    const float ssm = edge ? 0.0f : SS;
                    const float a00[4] = {CC, ll ? 0.0f : CS, kk ? 0.0f : SC, ssm};
                    const float a10[4] = {ll ? CS : CC, ll ? 0.0f : CS, kk ? 0.0f : (ll ? SS : SC), ssm};
                    const float a01[4] = {kk ? SC : CC, ll ? 0.0f : (kk ? SS : CS), kk ? 0.0f : SC, ssm};
                    const float a11[4] = {kk ? (ll ? SS : SC) : (ll ? CS : CC),
                                          ll ? 0.0f : (kk ? SS : CS), kk ? 0.0f : (ll ? SS : SC), ssm};
                    o[e][0][0] = a00[0] * cc + a00[3] * ss + a00[2] * scg + a00[1] * cs;
                    o[e][1][1] = a11[0] * ss + a11[3] * cc - a11[2] * cs - a11[1] * scg;
                    o[e][0][1] = a01[0] * scg - a01[2] * cc + a01[1] * ss - a01[3] * cs;
                    o[e][1][0] = a10[0] * cs - a10[1] * cc + a10[2] * ss - a10[3] * scg;
                }
#pragma unroll
                for (int ri = 0; ri < 2; ri++)
#pragma unroll
                    for (int ci = 0; ci < 2; ci++) {
                        unsigned h, l;
                        split2<BF, false>(o[0][ri][ci], o[1][ri][ci], h, l);
                        ch[ci][2 * ri + (v0 >> 3)][(v0 & 7) >> 1] = h;
                        cl[ci][2 * ri + (v0 >> 3)][(v0 & 7) >> 1] = l;
                    }
            }
#pragma unroll
            for (int ci = 0; ci < 2; ci++)
#pragma unroll
                for (int kq = 0; kq < 4; kq++) {
                    const int kb8 = 2 * (sr + 2 * (kq >> 1)) + (kq & 1);
                    const int e = ((sc + 2 * ci) * KM + kb8) * 2;
                    xs(e) = ch[ci][kq];
                    xs(e + 1) = cl[ci][kq];
                }
        }
        __syncthreads();

        // ---- Y = C Hout: rho tiles {sc, sc + 2} x q tiles {sr, sr + 2} ----------
        Acc<MT> Y[2][2];  // [ci][qi]
#pragma unroll
        for (int ci = 0; ci < 2; ci++)
#pragma unroll
            for (int qi = 0; qi < 2; qi++) Y[ci][qi] = Acc<MT>{};
        // K block outermost: the constant fragments of block kb + 1 are loaded
        // (from L2) while block kb's products run
        {
            u4 hq[2][2];  // [qi][hi/lo] of block kb
#pragma unroll
            for (int qi = 0; qi < 2; qi++)
#pragma unroll
                for (int hl = 0; hl < 2; hl++) hq[qi][hl] = frag(NF_HF + (sr + 2 * qi) * KM, hl);
#pragma unroll
            for (int kb = 0; kb < KM; kb++) {
                u4 hn[2][2];
                if (kb + 1 < KM) {
#pragma unroll
                    for (int qi = 0; qi < 2; qi++)
#pragma unroll
                        for (int hl = 0; hl < 2; hl++)
                            hn[qi][hl] = frag(NF_HF + (sr + 2 * qi) * KM + kb + 1, hl);
                }
#pragma unroll
                for (int ci = 0; ci < 2; ci++) {
                    const int ct = sc + 2 * ci;
                    const u4 ah = xs((ct * KM + kb) * 2), al = xs((ct * KM + kb) * 2 + 1);
#pragma unroll
                    for (int qi = 0; qi < 2; qi++)
                        Y[ci][qi] = mma3<BF, SPLIT, MT>(ah, al, hq[qi][0], hq[qi][1], Y[ci][qi]);
                }
                if (kb + 1 < KM) {
#pragma unroll
                    for (int qi = 0; qi < 2; qi++)
#pragma unroll
                        for (int hl = 0; hl < 2; hl++) hq[qi][hl] = hn[qi][hl];
                }
            }
        }
        __syncthreads();  // every wave has read C
        // Y B fragment (q tile qt, rho K block kb8, hi/lo) at entry (qt * KM + kb8) * 2 + hl
#pragma unroll
        for (int ci = 0; ci < 2; ci++)
#pragma unroll
            for (int qi = 0; qi < 2; qi++) {
                u4 yh[2], yl[2];
#pragma unroll
                for (int v = 0; v < NS; v += 2) {
                    unsigned h, l;
                    split2<BF, false>(Y[ci][qi][v] * (1.0f / (float)(1 << X::E2)),
                                      Y[ci][qi][v + 1] * (1.0f / (float)(1 << X::E2)), h, l);
                    yh[v >> 3][(v & 7) >> 1] = h;
                    yl[v >> 3][(v & 7) >> 1] = l;
                }
#pragma unroll
                for (int hb = 0; hb < 2; hb++) {
                    const int e = ((sr + 2 * qi) * KM + 2 * (sc + 2 * ci) + hb) * 2;
                    xs(e) = yh[hb];
                    xs(e + 1) = yl[hb];
                }
            }
        __syncthreads();

        // ---- R = Hout^T Y: p tiles {sc, sc + 2} x q tiles {sr, sr + 2} ---------
        Acc<MT> R[2][2];  // [pi][qi]
#pragma unroll
        for (int pi = 0; pi < 2; pi++)
#pragma unroll
            for (int qi = 0; qi < 2; qi++) R[pi][qi] = Acc<MT>{};
        {
            u4 hp[2][2];  // [pi][hi/lo] of block kb, the next block's loaded meanwhile
#pragma unroll
            for (int pi = 0; pi < 2; pi++)
#pragma unroll
                for (int hl = 0; hl < 2; hl++) hp[pi][hl] = fscaled(NF_HF + (sc + 2 * pi) * KM, hl);
#pragma unroll
            for (int kb = 0; kb < KM; kb++) {
                u4 hn[2][2];
                if (kb + 1 < KM) {
#pragma unroll
                    for (int pi = 0; pi < 2; pi++)
#pragma unroll
                        for (int hl = 0; hl < 2; hl++)
                            hn[pi][hl] = fscaled(NF_HF + (sc + 2 * pi) * KM + kb + 1, hl);
                }
#pragma unroll
                for (int qi = 0; qi < 2; qi++) {
                    const int qt = sr + 2 * qi;
                    const u4 bh = xs((qt * KM + kb) * 2), bl = xs((qt * KM + kb) * 2 + 1);
#pragma unroll
                    for (int pi = 0; pi < 2; pi++)
                        R[pi][qi] = mma3<BF, SPLIT, MT>(hp[pi][0], hp[pi][1], bh, bl, R[pi][qi]);
                }
                if (kb + 1 < KM) {
#pragma unroll
                    for (int pi = 0; pi < 2; pi++)
#pragma unroll
                        for (int hl = 0; hl < 2; hl++) hp[pi][hl] = hn[pi][hl];
                }
            }
        }
        __syncthreads();  // every wave has read Y: the buffer takes the plane

        // ---- epilogues: plane store, FromCenter -------------------------------------
        constexpr float unscale = 1.0f;  // carried by the scaled Hout^T copy
        if (p.planes) {
            float *__restrict__ out = p.planes + (size_t)ia * D * D;
#pragma unroll
            for (int pi = 0; pi < 2; pi++)
#pragma unroll
                for (int qi = 0; qi < 2; qi++)
#pragma unroll
                    for (int v = 0; v < NS; v++) {
                        const int a = MT * (sc + 2 * pi) + mrow<MT>(v, g), b = MT * (sr + 2 * qi) + j;
                        if (a < D && b < D) out[(unsigned)(a * D + b)] = R[pi][qi][v] * unscale;
                    }
        }
        if (!PEAK && p.mp_ent) {
            // MultiPeaks entries: the plane into the exchange buffer (free after
            // the barrier above; 127^2 floats fit its 16,384), then the blocks
            // from there by the work-group
            float *s_pl = (float *)&s_x[0];
            static_assert(D * D <= XCH * 4, "the plane fits the exchange buffer");
#pragma unroll
            for (int pi = 0; pi < 2; pi++)
#pragma unroll
                for (int qi = 0; qi < 2; qi++)
#pragma unroll
                    for (int v = 0; v < NS; v++) {
                        const int a = MT * (sc + 2 * pi) + mrow<MT>(v, g), b = MT * (sr + 2 * qi) + j;
                        if (a < D && b < D) s_pl[a * D + b] = R[pi][qi][v] * unscale;
                    }
            __syncthreads();
            constexpr MpGeom MG = mp_geom(D);
            mp_entries_lds<D, D, NT>(s_pl, s_min, p.mp_ent + (size_t)ia * MG.stride,
                                     p.mp_cmax + (size_t)ia * MG.nchunk, p.mp_min + ia,
                                     (int)threadIdx.x);
        }
        if (PEAK) {
            // FromCenter in two phases (from_center_tie) on the accumulator
            // scale (unscale is a power of two: order and equality are kept),
            // each wave over its four tiles, the waves joined through LDS
            auto visited = [&](int a, int b) { return a != 0 && b != 0 && a < D && b < D; };
            float vmax = -1.0f, mn = 3.40282347e+38f;
#pragma unroll
            for (int pi = 0; pi < 2; pi++)
#pragma unroll
                for (int qi = 0; qi < 2; qi++)
#pragma unroll
                    for (int v = 0; v < NS; v++) {
                        const int a = MT * (sc + 2 * pi) + mrow<MT>(v, g), b = MT * (sr + 2 * qi) + j;
                        const bool ok = visited(a, b);
                        const float x = R[pi][qi][v];
                        vmax = fmaxf(vmax, ok ? x : -1.0f);
                        mn = fminf(mn, ok ? x : 3.40282347e+38f);
                    }
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                vmax = fmaxf(vmax, __shfl_xor(vmax, o, 64));
                mn = fminf(mn, __shfl_xor(mn, o, 64));
            }
            if (lane == 0) {
                s_vmax[wave] = vmax;
                s_min[wave] = mn;
            }
            __syncthreads();
            vmax = fmaxf(fmaxf(s_vmax[0], s_vmax[1]), fmaxf(s_vmax[2], s_vmax[3]));
            mn = fminf(fminf(s_min[0], s_min[1]), fminf(s_min[2], s_min[3]));
            // phase 2: the tie keys of the cells equal to the maximum (a lane
            // holds none or one almost always)
            uint32_t best = 0;
            int bestCell = 0;
            if (vmax >= 0.0f) {
                uint64_t mask = 0;
#pragma unroll
                for (int pi = 0; pi < 2; pi++)
#pragma unroll
                    for (int qi = 0; qi < 2; qi++)
#pragma unroll
                        for (int v = 0; v < NS; v++) {
                            const int a = MT * (sc + 2 * pi) + mrow<MT>(v, g),
                                      b = MT * (sr + 2 * qi) + j;
                            const bool m = visited(a, b) && R[pi][qi][v] == vmax;
                            mask |= m ? (uint64_t)1 << ((pi * 2 + qi) * NS + v) : 0;
                        }
                best_tie(mask, D / 2, D,
                         [&](int idx, int &a, int &b) {
                             const int v = idx % NS, qi = (idx / NS) & 1, pi = idx / (2 * NS);
                             a = MT * (sc + 2 * pi) + mrow<MT>(v, g);
                             b = MT * (sr + 2 * qi) + j;
                         },
                         best, bestCell);
            }
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                const uint32_t ob = (uint32_t)__shfl_xor((int)best, o, 64);
                const int oc = __shfl_xor(bestCell, o, 64);
                if (ob > best) {
                    best = ob;
                    bestCell = oc;
                }
            }
            if (lane == 0) {
                s_key[wave] = best;
                s_cell[wave] = bestCell;
            }
            __syncthreads();
            best = 0;
            bestCell = 0;
#pragma unroll
            for (int w = 0; w < 4; w++)
                if ((uint32_t)s_key[w] > best) {  // keys are unique per cell
                    best = (uint32_t)s_key[w];
                    bestCell = s_cell[w];
                }
            // key 0: no visited value >= 0, the reference keeps (0, 0, 0)
            const int pc = best != 0 ? bestCell : 0;
            const int pI = pc / D, pJ = pc % D;
            // the neighbourhood (column through the peak, then the row) through
            // LDS: every cell inside the plane has one holder lane
            float *__restrict__ s_nb = (float *)&s_x[0];
            const int nh = p.nbr_half, L = 2 * nh + 1;
            if (p.nbr) {
                if (sc == ((pI >> 5) & 1)) {  // row pI: slot vs of lane group gs
                    const int ps = pI >> 6, ms = pI & 31;
                    const int gs = (ms >> 2) & 1, vs = ((ms >> 3) << 2) | (ms & 3);
                    float rv[2] = {0.0f, 0.0f};
#pragma unroll
                    for (int pi = 0; pi < 2; pi++)
#pragma unroll
                        for (int v = 0; v < NS; v++)
                            if (pi == ps && v == vs) {
                                rv[0] = R[pi][0][v];
                                rv[1] = R[pi][1][v];
                            }
#pragma unroll
                    for (int qi = 0; qi < 2; qi++) {
                        const int b = MT * (sr + 2 * qi) + j, d = b - pJ;
                        if (g == gs && b < D && d >= -nh && d <= nh) s_nb[L + d + nh] = rv[qi] * unscale;
                    }
                }
                if (sr == ((pJ >> 5) & 1) && j == (pJ & 31)) {  // column pJ
                    const int qs = pJ >> 6;
#pragma unroll
                    for (int pi = 0; pi < 2; pi++)
#pragma unroll
                        for (int v = 0; v < NS; v++) {
                            const int a = MT * (sc + 2 * pi) + mrow<MT>(v, g), d = a - pI;
                            if (a < D && d >= -nh && d <= nh)
                                s_nb[d + nh] = (qs == 0 ? R[pi][0][v] : R[pi][1][v]) * unscale;
                        }
                }
            }
            // the peak value is the maximum (the winning cell equals it)
            float pv = 0.0f;
            if (best != 0) pv = vmax * unscale;
            __syncthreads();
            if (wave == 0) {
                if (lane == 0) {
                    if (p.peaks) {
                        p.peaks[(size_t)ia * 3 + 0] = (float)pI;
                        p.peaks[(size_t)ia * 3 + 1] = (float)pJ;
                        p.peaks[(size_t)ia * 3 + 2] = pv;
                    }
                    if (p.min_floor) p.min_floor[ia] = mn * unscale;
                }
                // cells outside the plane are 0 (as in the exact pass)
                if (p.nbr && lane < L) {
                    float *nb = p.nbr + (size_t)ia * 2 * L;
                    const int d = lane - nh;
                    nb[lane] = (pI + d >= 0 && pI + d < D) ? s_nb[lane] : 0.0f;
                    nb[L + lane] = (pJ + d >= 0 && pJ + d < D) ? s_nb[L + lane] : 0.0f;
                }
            }
        }
        __syncthreads();  // the buffer is rewritten by the next IA
    }
}

// ---- host side ---------------------------------------------------------------

static double mf_basis(int M, int r, long x)
{
    const long xm = ((x % M) + M) % M;
    if (r < M / 2) return cos(2.0 * M_PI * (double)((r * xm) % M) / M);
    if (r == M / 2) return (xm & 1) ? -1.0 : 1.0;
    return sin(2.0 * M_PI * (double)(((r - M / 2) * xm) % M) / M);
}

static unsigned short to_bf16_bits(float f)
{
    uint32_t b;
    memcpy(&b, &f, 4);
    b += 0x7fffu + ((b >> 16) & 1u);  // round to nearest even (finite values)
    return (unsigned short)(b >> 16);
}

static float from_bf16_bits(unsigned short h)
{
    const uint32_t b = (uint32_t)h << 16;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

static void host_split(double x, bool bf, unsigned short &h, unsigned short &l)
{
    const float xf = (float)x;
    if (bf) {
        h = to_bf16_bits(xf);
        l = to_bf16_bits((float)(x - (double)from_bf16_bits(h)));
    } else {
        const _Float16 hh = (_Float16)xf;
        memcpy(&h, &hh, 2);
        const _Float16 ll = (_Float16)(float)(x - (double)(float)hh);
        memcpy(&l, &ll, 2);
    }
}

// Fragments of the constant A operands in kphys order: for fragment (rt, kb) of
// an R x K matrix, lane (i + MT g) element s holds A[MT rt + i][kphys(kb, g, s)].
// Table: the NF fragments (hi, lo), the row sums of Hf, then the Hout^T
// fragments again scaled by 2^(E1 + E2) (the final product's copy).
template <int N>
static void build_frags(bool bf, std::vector<unsigned short> &out)
{
    using X = Mf<N>;
    constexpr int M = X::M, MT = X::MT;
    const size_t scaled_at = (size_t)X::NF * 2 + 1;  // entry of the scaled copy
    out.assign((scaled_at + (size_t)X::NF_HO * 2) * 64 * 8, 0);
    // entry(ff, hl): table entry of fragment ff's hi (0) or lo (1) part
    auto put = [&](auto A, int rows, int cols, int kblocks, int fbase, auto entry) {
        for (int rt = 0; rt < X::RT; rt++)
            for (int kb = 0; kb < kblocks; kb++) {
                const int ff = fbase + rt * kblocks + kb;
                for (int lane = 0; lane < 64; lane++)
                    for (int s = 0; s < 8; s++) {
                        const int i = MT * rt + lane % MT, k = kphys<MT>(kb, lane / MT, s);
                        const double x = (i < rows && k < cols) ? A(i, k) : 0.0;
                        unsigned short h, l;
                        host_split(x, bf, h, l);
                        out[((size_t)entry(ff, 0) * 64 + lane) * 8 + s] = h;
                        out[((size_t)entry(ff, 1) * 64 + lane) * 8 + s] = l;
                    }
            }
    };
    auto plain = [](int ff, int hl) { return (size_t)ff * 2 + hl; };
    // Hf[rho][n] = h_rho(n)
    put([&](int r, int nn) { return mf_basis(M, r, nn); }, M, N, X::KN, 0, plain);
    // Hout^T[p][sigma] = h_sigma(p - (N - 1)); the edge frequencies sigma = 0
    // and M/2 carry their inverse weight 1/2 here
    auto hout = [&](double scale) {
        return [&, scale](int pp, int sg) {
            const double w = (sg == 0 || sg == M / 2) ? 0.5 : 1.0;
            return scale * w * mf_basis(M, sg, (long)pp - (N - 1));
        };
    };
    put(hout(1.0), M, M, X::KM, X::NF_HF, plain);
    put(hout((double)(1 << (X::E1 + X::E2))), M, M, X::KM, X::NF_HF,
        [&](int ff, int hl) { return scaled_at + (size_t)(ff - X::NF_HF) * 2 + hl; });
    // row sums h[rho] = sum_n Hf[rho][n], lane (i + MT g): h[MT c + i] for the
    // M / MT tiles c (2 or 4), as floats
    for (int lane = 0; lane < 64; lane++)
        for (int c = 0; c < M / MT; c++) {
            const int r = MT * c + lane % MT;
            double sum = 0.0;
            for (int nn = 0; nn < N; nn++) sum += mf_basis(M, r, nn);
            const float f = (float)sum;
            memcpy(&out[((size_t)X::NF * 2 * 64 + lane) * 8 + 2 * c], &f, 4);
        }
}

static int mfma_tables(Context *ctx, int N, bool bf, const u4 **out)
{
    const int ni = N == 16 ? 0 : (N == 32 ? 1 : 2), bi = bf ? 1 : 0;
    if (!ctx->mfma_tab[ni][bi]) {
        std::vector<unsigned short> h;
        if (N == 16)
            build_frags<16>(bf, h);
        else if (N == 32)
            build_frags<32>(bf, h);
        else
            build_frags<64>(bf, h);
        void *d = nullptr;
        const size_t bytes = h.size() * sizeof(unsigned short);
        if (int rc = check_hip(hipMalloc(&d, bytes), "hipMalloc(mfma tables)")) return rc;
        if (int rc = check_hip(hipMemcpy(d, h.data(), bytes, hipMemcpyHostToDevice),
                               "hipMemcpy(mfma tables)")) {
            (void)hipFree(d);
            return rc;
        }
        ctx->mfma_tab[ni][bi] = d;
    }
    *out = (const u4 *)ctx->mfma_tab[ni][bi];
    return QLP_OK;
}

template <int N, bool BF, bool SPLIT>
static void launch_mfma_t(Context *ctx, const MfmaParams &mp, bool from_image, bool peak)
{
    const int groups = (mp.p.n + MF_WAVES - 1) / MF_WAVES;
    const int grid = groups < 2048 ? groups : 2048;
    const int n_iter = (groups + grid - 1) / grid;
#define QLP_MF_LAUNCH(FI, PK)                                                                 \
    hipLaunchKernelGGL((xcorr_mfma_kernel<N, BF, SPLIT, FI, PK>), dim3(grid),                 \
                       dim3(64 * MF_WAVES), 0, ctx->stream, mp, n_iter)
    if (from_image) {
        if (peak)
            QLP_MF_LAUNCH(true, true);
        else
            QLP_MF_LAUNCH(true, false);
    } else {
        if (peak)
            QLP_MF_LAUNCH(false, true);
        else
            QLP_MF_LAUNCH(false, false);
    }
#undef QLP_MF_LAUNCH
}

template <bool BF, bool SPLIT>
static void launch_mfma64_t(Context *ctx, const PassParams &p, const u4 *frags, bool from_image,
                            bool peak)
{
    // two work-groups of 70 KB LDS per CU; each loops over IAs
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    const int grid = p.n < 2 * cus ? p.n : 2 * cus;
#define QLP_MF64_LAUNCH(FI, PK)                                                                \
    hipLaunchKernelGGL((xcorr_mfma64_kernel<BF, SPLIT, FI, PK>), dim3(grid), dim3(m64::NT), 0,   \
                       ctx->stream, p, frags)
    if (from_image) {
        if (peak)
            QLP_MF64_LAUNCH(true, true);
        else
            QLP_MF64_LAUNCH(true, false);
    } else {
        if (peak)
            QLP_MF64_LAUNCH(false, true);
        else
            QLP_MF64_LAUNCH(false, false);
    }
#undef QLP_MF64_LAUNCH
}

int launch_pass_mfma(Context *ctx, const PassParams &p, int form, bool from_image, bool peak)
{
    if (p.n <= 0) return QLP_OK;
    if (p.N != 16 && p.N != 32 && p.N != 64)
        return fail(QLP_ERR_UNSUPPORTED, "mfma correlation: IA size must be 16, 32 or 64");
    // frames are indexed with 32-bit offsets (the reference's frames are Java
    // arrays, below 2^31 elements too)
    if (from_image && (long long)p.H * p.W > 0x7fffffffLL)
        return fail(QLP_ERR_UNSUPPORTED, "mfma correlation: frames of 2^31 pixels or more");
    const bool bf = (form == QLP_XCORR_MFMA_BF16 || form == QLP_XCORR_MFMA_BF16_SINGLE);
    const bool split = (form == QLP_XCORR_MFMA_F16 || form == QLP_XCORR_MFMA_BF16);
    if (form < QLP_XCORR_MFMA_F16 || form > QLP_XCORR_MFMA_BF16_SINGLE)
        return fail(QLP_ERR_ARG, "mfma correlation: unknown form");
    MfmaParams mp;
    mp.p = p;
    if (int rc = mfma_tables(ctx, p.N, bf, &mp.frags)) return rc;
#define QLP_MF_N(NN)                                                                          \
    do {                                                                                      \
        if (bf) {                                                                             \
            if (split)                                                                        \
                launch_mfma_t<NN, true, true>(ctx, mp, from_image, peak);                     \
            else                                                                              \
                launch_mfma_t<NN, true, false>(ctx, mp, from_image, peak);                    \
        } else {                                                                              \
            if (split)                                                                        \
                launch_mfma_t<NN, false, true>(ctx, mp, from_image, peak);                    \
            else                                                                              \
                launch_mfma_t<NN, false, false>(ctx, mp, from_image, peak);                   \
        }                                                                                     \
    } while (0)
    if (p.N == 16) {
        QLP_MF_N(16);
    } else if (p.N == 32) {
        QLP_MF_N(32);
    } else if (bf) {
        if (split)
            launch_mfma64_t<true, true>(ctx, p, mp.frags, from_image, peak);
        else
            launch_mfma64_t<true, false>(ctx, p, mp.frags, from_image, peak);
    } else {
        if (split)
            launch_mfma64_t<false, true>(ctx, p, mp.frags, from_image, peak);
        else
            launch_mfma64_t<false, false>(ctx, p, mp.frags, from_image, peak);
    }
#undef QLP_MF_N
    return check_hip(hipGetLastError(), "mfma correlation launch");
}

}  // namespace qlpiv
