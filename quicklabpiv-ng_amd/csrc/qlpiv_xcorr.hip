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
                    const float2 w = tw[(M >> (s + 1)) * (q & (L - 1))];
                    bfly(a[q], b[q], a[q + L], b[q + L], w);
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

template <int LOGM, bool MODE_CPU>
__global__ __launch_bounds__(XcShape<LOGM>::THREADS) void xcorr_exact_kernel(
    int n, const uint8_t *__restrict__ tA, const uint8_t *__restrict__ tB,
    const float *__restrict__ maxA, const float *__restrict__ maxB,
    const float2 *__restrict__ tw_fwd, const float2 *__restrict__ tw_inv,
    float *__restrict__ planes)
{
    using X = XcShape<LOGM>;
    constexpr int M = X::M, N = X::N, S = X::S, D = X::D;
    __shared__ float s_re[M * S];
    __shared__ float s_im[M * S];
    __shared__ float2 s_twf[M / 2], s_twi[M / 2];

    const int ia = blockIdx.x;
    if (ia >= n) return;
    for (int k = threadIdx.x; k < M / 2; k += X::THREADS) {
        s_twf[k] = tw_fwd[k];
        s_twi[k] = tw_inv[k];
    }
    // normalised, mirrored, zero-padded pack (only the N x N data block is
    // written; the transforms never read the padding)
    const float mA = maxA[ia];
    const float mB = MODE_CPU ? mA : maxB[ia];
    const uint32_t *a32 = reinterpret_cast<const uint32_t *>(tA + (size_t)ia * N * N);
    const uint32_t *b32 = reinterpret_cast<const uint32_t *>(tB + (size_t)ia * N * N);
    for (int w = threadIdx.x; w < N * N / 4; w += X::THREADS) {
        const uint32_t va = a32[w], vb = b32[w];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int idx = 4 * w + e;  // natural index of the pixel
            const float fa = (float)((va >> (8 * e)) & 0xffu);
            const float fb = (float)((vb >> (8 * e)) & 0xffu);
            // F' lands mirrored: F'[r][c] = A[N-1-r][N-1-c]
            const int mi = N * N - 1 - idx;
            const int fr = mi / N, fc = mi % N;
            s_re[fr * S + fc] = fa / mA * 16.0f + 1.0f;
            const int gr = idx / N, gc = idx % N;
            s_im[gr * S + gc] = fb / mB * 16.0f + 1.0f;
        }
    }
    __syncthreads();

    // forward: rows (only the N data rows are non-zero), then all columns
    fft_lines<LOGM, true, false>(s_re, s_im, S, 1, N, s_twf);
    fft_lines<LOGM, true, false>(s_re, s_im, 1, S, M, s_twf);
    unpack_product<LOGM, MODE_CPU>(s_re, s_im);
    // inverse: all rows, then the D columns that reach the output
    fft_lines<LOGM, false, MODE_CPU>(s_re, s_im, S, 1, M, s_twi);
    fft_lines<LOGM, false, MODE_CPU>(s_re, s_im, 1, S, D, s_twi);

    float *out = planes + (size_t)ia * D * D;
    for (int idx = threadIdx.x; idx < D * D; idx += X::THREADS) {
        const int i = idx / D, j = idx - (idx / D) * D;
        out[idx] = s_re[i * S + j];
    }
}

template <int LOGM>
static int launch_xcorr_m(Context *ctx, int n, const uint8_t *tA, const uint8_t *tB,
                          const float *maxA, const float *maxB, int mode, float *planes)
{
    using X = XcShape<LOGM>;
    const int li = LOGM - 5;
    dim3 grid(n), block(X::THREADS);
    if (mode == QLP_NORM_A_MAX) {
        hipLaunchKernelGGL((xcorr_exact_kernel<LOGM, true>), grid, block, 0, ctx->stream, n, tA,
                           tB, maxA, maxB, ctx->tables.tw[li][TW_FWD],
                           ctx->tables.tw[li][TW_INV_CPU], planes);
    } else {
        hipLaunchKernelGGL((xcorr_exact_kernel<LOGM, false>), grid, block, 0, ctx->stream, n, tA,
                           tB, maxA, maxB, ctx->tables.tw[li][TW_FWD],
                           ctx->tables.tw[li][TW_INV_GPU], planes);
    }
    return check_hip(hipGetLastError(), "xcorr kernel launch");
}

int launch_xcorr(Context *ctx, int n, int N, const uint8_t *tA, const uint8_t *tB,
                 const float *maxA, const float *maxB, int mode, float *planes, int peak_kind,
                 float *peaks_out)
{
    (void)peak_kind;
    (void)peaks_out;
    if (n == 0) return QLP_OK;
    switch (N) {
    case 16: return launch_xcorr_m<5>(ctx, n, tA, tB, maxA, maxB, mode, planes);
    case 32: return launch_xcorr_m<6>(ctx, n, tA, tB, maxA, maxB, mode, planes);
    case 64: return launch_xcorr_m<7>(ctx, n, tA, tB, maxA, maxB, mode, planes);
    default: return fail(QLP_ERR_UNSUPPORTED, "xcorr: IA size N must be 16, 32 or 64");
    }
}

}  // namespace qlpiv
