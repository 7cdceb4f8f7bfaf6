// Batched FFT cross-correlation of interrogation areas, exact Java arithmetic.
//
// One work-group correlates one IA pair entirely in LDS:
//   u8 tiles -> normalise (v / max * 16 + 1, F mirrored) -> 2N x 2N zero-padded
//   complex tile (F' real, G' imaginary) -> row FFTs -> column FFTs ->
//   two-for-one unpack and product F*G -> row IFFTs -> column IFFTs ->
//   (2N-1)^2 real plane.
//
// The transforms perform exactly the radix-2 decimation-in-time butterflies of
// the reference (SimpleFFT.computeFFTStep :122-221 / computeIFFTStep :255-320,
// CrossCorrelationRealFFTParallelKernel.computeFFTStep :96-195), with the same
// twiddle table entries and the same operation order inside each butterfly, so
// the planes equal the Java ones value for value.  Only the data movement is
// new: each thread keeps E = 8 (M <= 64) or 16 (M = 128) elements of a line in
// registers and runs log2(E) stages there; one LDS exchange separates the two
// register passes of a 1D transform, so a 64-point line costs two LDS round
// trips instead of six.  Butterflies whose second operand is zero padding
// (the first stage of the forward row and column passes) reduce to a copy,
// which gives the same values.
//
// Modes (template MODE_CPU):
//   true  : CrossCorrelationSoftRealFFTJob -> FastRealFFTXCorr (both tiles
//           normalised by A's max; IFFT with inverse twiddles, /M per 1D pass)
//   false : CrossCorrelationRealFFTParStdJob (each tile by its own max; product
//           divided by M^2 once; inverse through the re/im swap of the forward
//           transform)
#include "qlpiv_device.h"
#include "qlpiv_internal.h"

namespace qlpiv {

template <int LOGM>
struct XcShape {
    static constexpr int M = 1 << LOGM;
    static constexpr int N = M / 2;
    static constexpr int D = M - 1;
    static constexpr int LOGE = (LOGM >= 7) ? 4 : 3;
    static constexpr int E = 1 << LOGE;
    static constexpr int T = M / E;  // threads per line
    static constexpr int S = M + 1;  // LDS row stride in floats (odd: conflict-free columns)
    static constexpr int THREADS = (LOGM == 5) ? 128 : (LOGM == 6 ? 256 : 512);
    static constexpr int ROUNDS = (M * T + THREADS - 1) / THREADS;
};

__device__ __forceinline__ int bitrev_n(int x, int bits) { return (int)(__brev((unsigned)x) >> (32 - bits)); }

// The reference butterfly: lo = a + w*b, hi = a - w*b with w = c + i*s, both
// results computed from the inputs before either is stored.
__device__ __forceinline__ void bfly(float &ar, float &ai, float &br, float &bi, float2 w)
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

// One 1D transform over 'nlines' lines of M complex values held in LDS.
// Element (line l, index o) lives at l*LS + o*PS.  The input is in natural
// order (the bit-reversal shuffle is folded into the first gather); the output
// is in natural order, in place.
//   ZERO_HALF : indices >= M/2 are zero padding (forward passes)
//   PRESCALE  : divide every input by M first (the CPU IFFT's divideArray)
template <int LOGM, bool ZERO_HALF, bool PRESCALE>
__device__ __forceinline__ void fft_lines(float *__restrict__ re, float *__restrict__ im, int LS,
                                          int PS, int nlines, const float2 *__restrict__ tw)
{
    using X = XcShape<LOGM>;
    constexpr int M = X::M, E = X::E, LOGE = X::LOGE, T = X::T, B = LOGM - LOGE;
    const int ntasks = nlines * T;
    float rr[X::ROUNDS][E], ri[X::ROUNDS][E];

    // pass 1 gather: thread (l, t) takes positions p = t*E + q, which after the
    // perfect shuffle hold the original indices bitrev(p)
#pragma unroll
    for (int rd = 0; rd < X::ROUNDS; rd++) {
        const int task = threadIdx.x + rd * X::THREADS;
        if (task < ntasks) {
            const int l = task % nlines, t = task / nlines;
            const int base = l * LS;
#pragma unroll
            for (int q = 0; q < E; q += (ZERO_HALF ? 2 : 1)) {
                const int o = bitrev_n(t * E + q, LOGM);
                rr[rd][q] = re[base + o * PS];
                ri[rd][q] = im[base + o * PS];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int rd = 0; rd < X::ROUNDS; rd++) {
        const int task = threadIdx.x + rd * X::THREADS;
        if (task < ntasks) {
            const int l = task % nlines, t = task / nlines;
            const int base = l * LS;
            float *a = rr[rd], *b = ri[rd];
            if (PRESCALE) {
                const float fM = (float)M;
#pragma unroll
                for (int q = 0; q < E; q++) {
                    a[q] = a[q] / fM;
                    b[q] = b[q] / fM;
                }
            }
            int s0 = 0;
            if (ZERO_HALF) {
                // stage 0 pairs (2u, 2u+1); the odd partner is zero padding
#pragma unroll
                for (int q = 0; q < E; q += 2) {
                    a[q + 1] = a[q];
                    b[q + 1] = b[q];
                }
                s0 = 1;
            }
#pragma unroll
            for (int s = 0; s < LOGE; s++) {
                if (s < s0) continue;
                const int L = 1 << s;
#pragma unroll
                for (int q = 0; q < E; q++) {
                    if (q & L) continue;
                    const int k = (M >> (s + 1)) * (q & (L - 1));
                    if (k == 0) {
                        // c = 1, s = -0/+0 in all three tables: a plain add/sub, same values
                        const float lr = a[q] + a[q + L], li = b[q] + b[q + L];
                        const float ur = a[q] - a[q + L], ui = b[q] - b[q + L];
                        a[q] = lr;
                        b[q] = li;
                        a[q + L] = ur;
                        b[q + L] = ui;
                    } else {
                        bfly(a[q], b[q], a[q + L], b[q + L], tw[k]);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < E; q++) {
                const int p = t * E + q;
                re[base + p * PS] = a[q];
                im[base + p * PS] = b[q];
            }
        }
    }
    __syncthreads();
    // pass 2: positions p = t + (q << B), stages LOGE..LOGM-1, in place
    for (int task = threadIdx.x; task < ntasks; task += X::THREADS) {
        const int l = task % nlines, t = task / nlines;
        const int base = l * LS;
        float a[E], b[E];
#pragma unroll
        for (int q = 0; q < E; q++) {
            const int p = t + (q << B);
            a[q] = re[base + p * PS];
            b[q] = im[base + p * PS];
        }
#pragma unroll
        for (int s = LOGE; s < LOGM; s++) {
            const int Lq = 1 << (s - B);
#pragma unroll
            for (int q = 0; q < E; q++) {
                if (q & Lq) continue;
                const int k = (M >> (s + 1)) * (t + ((q & (Lq - 1)) << B));
                bfly(a[q], b[q], a[q + Lq], b[q + Lq], tw[k]);
            }
        }
#pragma unroll
        for (int q = 0; q < E; q++) {
            const int p = t + (q << B);
            re[base + p * PS] = a[q];
            im[base + p * PS] = b[q];
        }
    }
    __syncthreads();
}

// Two-for-one unpack and product (FastRealFFTXCorr.commonRealFFTXCorr :73-112;
// CrossCorrelationRealFFTParallelKernel.run :458-600).  Each conjugate pair
// (k, -k) is owned by one thread that reads both values before writing both.
// The formulas are symmetric (Fr, Gr) or antisymmetric (Fi, Gi) in the pair, so
// P[-k] = conj(P[k]) holds bit for bit whichever element owns the pair.
template <int LOGM, bool MODE_CPU>
__device__ __forceinline__ void unpack_product(float *__restrict__ re, float *__restrict__ im)
{
    using X = XcShape<LOGM>;
    constexpr int M = X::M, S = X::S;
    for (int idx = threadIdx.x; idx < M * M; idx += X::THREADS) {
        const int r = idx >> LOGM, c = idx & (M - 1);
        const int nr = (M - r) & (M - 1), nc = (M - c) & (M - 1);
        if (idx > nr * M + nc) continue;
        const float x = re[r * S + c], y = im[r * S + c];
        const float Nx = re[nr * S + nc], Ny = im[nr * S + nc];
        const float fr = 0.5f * (x + Nx);
        const float fi = 0.5f * (y - Ny);
        const float gr = 0.5f * (y + Ny);
        const float gi = -0.5f * (x - Nx);
        float pre = fr * gr - fi * gi;
        float pim = fr * gi + fi * gr;
        if (!MODE_CPU) {
            pre = pre / (float)(M * M);
            pim = pim / (float)(M * M);
        }
        re[r * S + c] = pre;
        im[r * S + c] = pim;
        re[nr * S + nc] = pre;
        im[nr * S + nc] = -pim;
    }
    __syncthreads();
}

// Block-wide maximum of an int (tile maxima).
template <int THREADS>
__device__ __forceinline__ int block_max_int(int v, int *s_red)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    int m = s_red[0];
#pragma unroll
    for (int w = 1; w < THREADS / 64; w++) m = max(m, s_red[w]);
    __syncthreads();
    return m;
}

// One correlation pass: tile formation, exact FFT correlation, and the
// optional epilogues (plane store, FromCenter peak with its neighbourhood).
//   FROM_IMAGE : tiles are cut from the images (A: clip at (top, left);
//                B: clip, shifted clip or micro warp by (du, dv)), the tile
//                maxima are computed here (>= 16 clamp; 255 for a warped B)
//   PEAK       : FromCenter on the plane while it is still in LDS
template <int LOGM, bool MODE_CPU, bool FROM_IMAGE, bool PEAK>
__global__ __launch_bounds__(XcShape<LOGM>::THREADS) void xcorr_pass_kernel(
    PassParams p, const float2 *__restrict__ tw_fwd, const float2 *__restrict__ tw_inv)
{
    using X = XcShape<LOGM>;
    constexpr int M = X::M, N = X::N, S = X::S, D = X::D, NT = X::THREADS;
    constexpr int PIX = N * N / NT;  // pixels per thread (2, 4 or 8)
    __shared__ float s_re[M * S];
    __shared__ float s_im[M * S];
    __shared__ float2 s_twf[M / 2], s_twi[M / 2];
    __shared__ int s_red[NT / 64];
    __shared__ uint64_t s_key[NT / 64];
    __shared__ int s_cell[NT / 64];
    __shared__ float s_min[NT / 64];

    const int ia = blockIdx.x;
    if (ia >= p.n) return;
    for (int k = threadIdx.x; k < M / 2; k += NT) {
        s_twf[k] = tw_fwd[k];
        s_twi[k] = tw_inv[k];
    }

    // ---- tile formation + normalisation: v / max * 16 + 1, F mirrored -------
    // (Matrix.copyMirroredMatrixToArrayNormalizeAndOffset :282-288,
    //  copyMatrixToArrayAndNormalizeAndOffset :175-181; only the N x N data
    //  block is written, the transforms never read the zero padding)
    float va[PIX], vb[PIX];
    float mA, mB;
    if (FROM_IMAGE) {
        const int per_pair = p.per_pair > 0 ? p.per_pair : p.n;
        const int pair = ia / per_pair, t = ia - pair * per_pair;
        const size_t img_off = (size_t)pair * p.H * p.W;
        const uint8_t *__restrict__ imA = p.imgA + img_off;
        const uint8_t *__restrict__ imB = p.imgB + img_off;
        const int y0 = p.top[t], x0 = p.left[t];
        int by0 = y0, bx0 = x0;
        float fu = 0.0f, fv = 0.0f;
        const bool warp = (p.bmode == B_WARP_MICRO) && p.du != nullptr;
        if (p.du != nullptr) {
            const float u = p.du[ia], v = p.dv[ia];  // per IA of the batch
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
        int amax = 0, bmax = 0;
#pragma unroll
        for (int k = 0; k < PIX; k++) {
            const int idx = threadIdx.x + k * NT;
            const int r = idx / N, c = idx - (idx / N) * N;
            int a = 0, b = 0;
            const int ya = y0 + r, xa = x0 + c;
            if (ya >= 0 && ya < p.H && xa >= 0 && xa < p.W) a = imA[(size_t)ya * p.W + xa];
            if (warp) {
                b = (int)warp_pixel(imB, p.H, p.W, fu, fv, by0 + r, bx0 + c);
            } else {
                const int yb = by0 + r, xb = bx0 + c;
                if (yb >= 0 && yb < p.H && xb >= 0 && xb < p.W) {
                    b = imB[(size_t)yb * p.W + xb];
                    bmax = max(bmax, b);
                }
            }
            amax = max(amax, a);
            va[k] = (float)a;
            vb[k] = (float)b;
        }
        amax = block_max_int<NT>(amax, s_red);
        mA = (float)(amax < 16 ? 16 : amax);
        if (warp || p.bmax255) {
            mB = 255.0f;  // warped tiles carry the image maximum (doBiLinearWarp :298)
        } else {
            bmax = block_max_int<NT>(bmax, s_red);
            mB = (float)(bmax < 16 ? 16 : bmax);
        }
        if (MODE_CPU) mB = mA;
    } else {
        mA = p.maxA[ia];
        mB = MODE_CPU ? mA : p.maxB[ia];
        const uint8_t *a8 = p.tA + (size_t)ia * N * N, *b8 = p.tB + (size_t)ia * N * N;
#pragma unroll
        for (int k = 0; k < PIX; k++) {
            const int idx = threadIdx.x + k * NT;
            va[k] = (float)a8[idx];
            vb[k] = (float)b8[idx];
        }
    }
#pragma unroll
    for (int k = 0; k < PIX; k++) {
        const int idx = threadIdx.x + k * NT;
        const int mi = N * N - 1 - idx;  // F'[r][c] = A[N-1-r][N-1-c]
        s_re[(mi / N) * S + (mi % N)] = va[k] / mA * 16.0f + 1.0f;
        s_im[(idx / N) * S + (idx % N)] = vb[k] / mB * 16.0f + 1.0f;
    }
    __syncthreads();

    // ---- transforms ----------------------------------------------------------
    // forward: rows (only the N data rows are non-zero), then all columns
    fft_lines<LOGM, true, false>(s_re, s_im, S, 1, N, s_twf);
    fft_lines<LOGM, true, false>(s_re, s_im, 1, S, M, s_twf);
    unpack_product<LOGM, MODE_CPU>(s_re, s_im);
    // inverse: all rows, then the D columns that reach the output
    fft_lines<LOGM, false, MODE_CPU>(s_re, s_im, S, 1, M, s_twi);
    fft_lines<LOGM, false, MODE_CPU>(s_re, s_im, 1, S, D, s_twi);

    // ---- epilogues -------------------------------------------------------------
    if (p.planes) {
        float *out = p.planes + (size_t)ia * D * D;
        for (int idx = threadIdx.x; idx < D * D; idx += NT) {
            const int i = idx / D, j = idx - (idx / D) * D;
            out[idx] = s_re[i * S + j];
        }
    }
    if (PEAK) {
        constexpr int h = D / 2;
        uint64_t best = 0;
        int bestCell = 0;
        float mn = 3.40282347e+38f;
        for (int idx = threadIdx.x; idx < D * D; idx += NT) {
            const int i = idx / D, j = idx - (idx / D) * D;
            if (i == 0 || j == 0) continue;
            const float v = s_re[i * S + j];
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
        if ((threadIdx.x & 63) == 0) {
            s_key[threadIdx.x >> 6] = best;
            s_cell[threadIdx.x >> 6] = bestCell;
            s_min[threadIdx.x >> 6] = mn;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
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
                pv = s_re[pi * S + pj];
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
                    nb[d + p.nbr_half] = (ri >= 0 && ri < D) ? s_re[ri * S + pj] : 0.0f;
                    nb[L + d + p.nbr_half] = (cj >= 0 && cj < D) ? s_re[pi * S + cj] : 0.0f;
                }
            }
        }
    }
}

template <int LOGM, bool MODE_CPU>
static void launch_pass_mc(Context *ctx, const PassParams &p, bool from_image, bool peak)
{
    using X = XcShape<LOGM>;
    const int li = LOGM - 5;
    const float2 *twf = ctx->tables.tw[li][TW_FWD];
    const float2 *twi = ctx->tables.tw[li][MODE_CPU ? TW_INV_CPU : TW_INV_GPU];
    dim3 grid(p.n), block(X::THREADS);
    if (from_image) {
        if (peak)
            hipLaunchKernelGGL((xcorr_pass_kernel<LOGM, MODE_CPU, true, true>), grid, block, 0,
                               ctx->stream, p, twf, twi);
        else
            hipLaunchKernelGGL((xcorr_pass_kernel<LOGM, MODE_CPU, true, false>), grid, block, 0,
                               ctx->stream, p, twf, twi);
    } else {
        if (peak)
            hipLaunchKernelGGL((xcorr_pass_kernel<LOGM, MODE_CPU, false, true>), grid, block, 0,
                               ctx->stream, p, twf, twi);
        else
            hipLaunchKernelGGL((xcorr_pass_kernel<LOGM, MODE_CPU, false, false>), grid, block, 0,
                               ctx->stream, p, twf, twi);
    }
}

template <int LOGM>
static void launch_pass_m(Context *ctx, const PassParams &p, bool from_image, bool peak)
{
    if (p.mode == QLP_NORM_A_MAX)
        launch_pass_mc<LOGM, true>(ctx, p, from_image, peak);
    else
        launch_pass_mc<LOGM, false>(ctx, p, from_image, peak);
}

int launch_pass(Context *ctx, const PassParams &p)
{
    if (p.n == 0) return QLP_OK;
    const bool from_image = p.imgA != nullptr;
    const bool peak = p.peaks != nullptr || p.min_floor != nullptr || p.nbr != nullptr;
    if (p.nbr && (p.nbr_half < 0 || p.nbr_half > 7))
        return fail(QLP_ERR_ARG, "pass: neighbourhood half-width must be 0..7");
    // measured on MI355X (C2 multipass): the lane-per-line form is faster for
    // N = 16 (0.293 vs 0.323 ms for 65,025 IAs), the LDS-exchange form for
    // N = 32 (0.381 vs 0.436 ms for 16,129 IAs) and N = 64, where more
    // wavefronts per IA hide latency better
    // the matrix-core forms cover N = 16 and 32; N = 64 keeps the exact form
    if (p.arith != QLP_XCORR_EXACT && (p.N == 16 || p.N == 32))
        return launch_pass_mfma(ctx, p, p.arith, from_image, peak);
    if (ctx->xcorr_form == FORM_LANE && p.N == 16) return launch_pass_lane(ctx, p, from_image, peak);
    switch (p.N) {
    case 16: launch_pass_m<5>(ctx, p, from_image, peak); break;
    case 32: launch_pass_m<6>(ctx, p, from_image, peak); break;
    case 64: launch_pass_m<7>(ctx, p, from_image, peak); break;
    default: return fail(QLP_ERR_UNSUPPORTED, "xcorr: IA size N must be 16, 32 or 64");
    }
    return check_hip(hipGetLastError(), "xcorr pass launch");
}

int launch_xcorr(Context *ctx, int n, int N, const uint8_t *tA, const uint8_t *tB,
                 const float *maxA, const float *maxB, int mode, float *planes)
{
    PassParams p;
    p.n = n;
    p.N = N;
    p.mode = mode;
    p.tA = tA;
    p.tB = tB;
    p.maxA = maxA;
    p.maxB = maxB;
    p.planes = planes;
    return launch_pass(ctx, p);
}

}  // namespace qlpiv
