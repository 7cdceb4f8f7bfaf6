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
#include <type_traits>
#include <mutex>

#include "qlpiv_device.h"
#include "qlpiv_epilogue.h"
#include "qlpiv_internal.h"
#include "qlpiv_multipeak.h"

namespace qlpiv {

// Geometry of the in-LDS kernel for M = 2N = 2^LOGM; NT_ overrides the
// default work-group size.
template <int LOGM, int NT_ = 0>
struct XcShape {
    static constexpr int M = 1 << LOGM;
    static constexpr int N = M / 2;
    static constexpr int D = M - 1;
    static constexpr int LOGE = (LOGM >= 7) ? 4 : 3;
    static constexpr int E = 1 << LOGE;
    static constexpr int T = M / E;  // threads per line
    static constexpr int S = M + 1;  // LDS row stride in floats (odd: conflict-free columns)
    static constexpr int THREADS = NT_ ? NT_ : ((LOGM == 5) ? 128 : (LOGM == 6 ? 256 : 512));
};

__device__ __forceinline__ int bitrev_n(int x, int bits) { return (int)(__brev((unsigned)x) >> (32 - bits)); }

// Bit reversal of x over 'bits' bits, for indices known at compile time.
__host__ __device__ constexpr unsigned brev_const(int x, int bits)
{
    unsigned r = 0;
    for (int b = 0; b < bits; b++) r |= (unsigned)((x >> b) & 1) << (bits - 1 - b);
    return r;
}

// The thread index, hidden from the optimiser: address arithmetic built on it
// is then recomputed in each iteration of a loop over IAs (or line batches)
// instead of being hoisted out of the loop into registers for every unrolled
// load.  The helpers below take it as 'tid' (-1: threadIdx.x).
__device__ __forceinline__ int opaque_tid()
{
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

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
// is in natural order, in place.  NT threads share the work; at most MAXLINES
// lines per call.
//   ZERO_HALF : indices >= M/2 are zero padding (forward passes)
//   PRESCALE  : divide every input by M first (the CPU IFFT's divideArray)
template <int LOGM, bool ZERO_HALF, bool PRESCALE, int NT = XcShape<LOGM>::THREADS,
          int MAXLINES = XcShape<LOGM>::M>
__device__ __forceinline__ void fft_lines(float *__restrict__ re, float *__restrict__ im, int LS,
                                          int PS, int nlines, const float2 *__restrict__ tw,
                                          int tid = -1)
{
    const int tx = tid < 0 ? (int)threadIdx.x : tid;
    using X = XcShape<LOGM>;
    constexpr int M = X::M, E = X::E, LOGE = X::LOGE, T = X::T, B = LOGM - LOGE;
    constexpr int ROUNDS = (MAXLINES * T + NT - 1) / NT;
    const int ntasks = nlines * T;
    float rr[ROUNDS][E], ri[ROUNDS][E];

    // pass 1 gather: thread (l, t) takes positions p = t*E + q, which after the
    // perfect shuffle hold the original indices bitrev(p)
    #pragma unroll
    for (int rd = 0; rd < ROUNDS; rd++) {
        const int task = tx + rd * NT;
        if (task < ntasks) {
            const int l = task % nlines, t = task / nlines;
            // bitrev(t*E + q) over LOGM bits = bitrev(q) over LOGE bits, shifted
            // up by B, plus bitrev(t) over B bits (t < T = 2^B): one reversal
            // per line, and a compile-time offset per element
            const int base = l * LS + bitrev_n(t, B) * PS;
            #pragma unroll
            for (int q = 0; q < E; q += (ZERO_HALF ? 2 : 1)) {
                const int oq = (int)(brev_const(q, LOGE) << B);
                rr[rd][q] = re[base + oq * PS];
                ri[rd][q] = im[base + oq * PS];
            }
        }
    }
    __syncthreads();
    #pragma unroll
    for (int rd = 0; rd < ROUNDS; rd++) {
        const int task = tx + rd * NT;
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
    for (int task = tx; task < ntasks; task += NT) {
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
// Row stride S; NT threads.
template <int LOGM, bool MODE_CPU, int NT = XcShape<LOGM>::THREADS, int S = XcShape<LOGM>::S>
__device__ __forceinline__ void unpack_product(float *__restrict__ re, float *__restrict__ im,
                                               int tid = -1)
{
    constexpr int M = XcShape<LOGM>::M;
    for (int idx = tid < 0 ? (int)threadIdx.x : tid; idx < M * M; idx += NT) {
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
__device__ __forceinline__ int block_max_int(int v, int *s_red, int tid = -1)
{
    const int tx = tid < 0 ? (int)threadIdx.x : tid;
    #pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    if ((tx & 63) == 0) s_red[tx >> 6] = v;
    __syncthreads();
    int m = s_red[0];
    #pragma unroll
    for (int w = 1; w < THREADS / 64; w++) m = max(m, s_red[w]);
    __syncthreads();
    return m;
}

// Block-wide maximum of a float (NaN ignored), returned to every thread.
template <int THREADS>
__device__ __forceinline__ float block_max_float(float v, float *s_red, int tid)
{
    v = wave_max_float(v);
    if ((tid & 63) == 0) s_red[tid >> 6] = v;
    __syncthreads();
    float m = s_red[0];
#pragma unroll
    for (int w = 1; w < THREADS / 64; w++) m = fmaxf(m, s_red[w]);
    __syncthreads();
    return m;
}

// One correlation pass: tile formation, exact FFT correlation, and the
// optional epilogues (plane store, FromCenter peak with its neighbourhood).
//   FROM_IMAGE : tiles are cut from the images (A: clip at (top, left);
//                B: clip, shifted clip or micro warp by (du, dv)), the tile
//                maxima are computed here (>= 16 clamp; 255 for a warped B)
//   PEAK       : FromCenter on the plane while it is still in LDS
//   NT_        : work-group size (0: the default of XcShape)
//   PERSIST    : persistent launch: the IA count in device memory (p.n_dev),
//                every work-group loops over IAs; otherwise one IA per
//                work-group in straight-line code
template <int LOGM, bool MODE_CPU, bool FROM_IMAGE, bool PEAK, int NT_ = 0, bool PERSIST = false>
__global__ __launch_bounds__((XcShape<LOGM, NT_>::THREADS)) void xcorr_pass_kernel(
    PassParams p, const float2 *__restrict__ tw_fwd, const float2 *__restrict__ tw_inv)
{
    using X = XcShape<LOGM, NT_>;
    constexpr int M = X::M, N = X::N, S = X::S, D = X::D, NT = X::THREADS;
    constexpr int PIX = N * N / NT;  // pixels per thread (2 to 16)
    __shared__ float s_re[M * S];
    __shared__ float s_im[M * S];
    __shared__ float2 s_twf[M / 2], s_twi[M / 2];
    __shared__ int s_red[NT / 64];
    __shared__ uint64_t s_key[NT / 64];
    __shared__ int s_cell[NT / 64];
    __shared__ float s_min[NT / 64];

    const int n = PERSIST ? *p.n_dev : p.n;
    if ((int)blockIdx.x >= n) return;
    for (int k = threadIdx.x; k < M / 2; k += NT) {
        s_twf[k] = tw_fwd[k];
        s_twi[k] = tw_inv[k];
    }
    // persistent launch: this work-group's first tiles, loaded ahead
    float pa[PERSIST ? PIX : 1], pb[PERSIST ? PIX : 1], pmA = 0.0f, pmB = 0.0f;
    if constexpr (PERSIST && !FROM_IMAGE) {
        const int ia0 = blockIdx.x;
        pmA = p.maxA[ia0];
        pmB = MODE_CPU ? pmA : p.maxB[ia0];
        const size_t at = (size_t)ia0 * N * N;
        #pragma unroll
        for (int k = 0; k < PIX; k++) {
            pa[k] = (float)p.tA[at + threadIdx.x + k * NT];
            pb[k] = (float)p.tB[at + threadIdx.x + k * NT];
        }
    }
    // one IA per work-group, or (persistent launch) every gridDim.x-th IA
    for (int ia = blockIdx.x; ia < n; ia += gridDim.x) {
        const int tid = PERSIST ? opaque_tid() : (int)threadIdx.x;
        // ---- tile formation + normalisation: v / max * 16 + 1, F mirrored -------
        // (Matrix.copyMirroredMatrixToArrayNormalizeAndOffset :282-288,
        //  copyMatrixToArrayAndNormalizeAndOffset :175-181; only the N x N data
        //  block is written, the transforms never read the zero padding)
        float va[PIX], vb[PIX];
        float mA, mB;
        if (FROM_IMAGE) {
            const int per_pair = p.per_pair > 0 ? p.per_pair : n;
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
            // unconditional loads from addresses clamped into the frame (32-bit
            // offsets: the launcher bounds H * W below 2^31), masked afterwards,
            // one loop per B form: the loads of all pixels are in flight together
            const unsigned Hu = (unsigned)p.H, Wu = (unsigned)p.W;
            auto load = [&](auto warp_c) {
                #pragma unroll
                for (int k = 0; k < PIX; k++) {
                    const int idx = tid + k * NT;
                    const int r = idx / N, c = idx - (idx / N) * N;
                    const unsigned ya = (unsigned)(y0 + r), xa = (unsigned)(x0 + c);
                    const bool ina = ya < Hu && xa < Wu;
                    int a = imA[ina ? ya * Wu + xa : 0u];
                    a = ina ? a : 0;
                    int b;
                    if constexpr (decltype(warp_c)::value) {
                        b = (int)warp_pixel<uint8_t, unsigned>(imB, p.H, p.W, fu, fv, by0 + r, bx0 + c);
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
            amax = block_max_int<NT>(amax, s_red, tid);
            mA = (float)(amax < 16 ? 16 : amax);
            if (warp || p.bmax255) {
                mB = 255.0f;  // warped tiles carry the image maximum (doBiLinearWarp :298)
            } else {
                bmax = block_max_int<NT>(bmax, s_red, tid);
                mB = (float)(bmax < 16 ? 16 : bmax);
            }
            if (MODE_CPU) mB = mA;
        } else if (PERSIST) {
            // the tiles were loaded one IA ahead; load the next IA's now, so that
            // its loads are in flight during this IA's transforms
            mA = pmA;
            mB = MODE_CPU ? mA : pmB;
            #pragma unroll
            for (int k = 0; k < PIX; k++) {
                va[k] = pa[k];
                vb[k] = pb[k];
            }
            const int nx = ia + (int)gridDim.x;
            if (nx < n) {
                pmA = p.maxA[nx];
                pmB = MODE_CPU ? pmA : p.maxB[nx];
                const size_t at = (size_t)nx * N * N;
                #pragma unroll
                for (int k = 0; k < PIX; k++) {
                    const int idx = tid + k * NT;
                    pa[k] = (float)p.tA[at + idx];
                    pb[k] = (float)p.tB[at + idx];
                }
            }
        } else {
            mA = p.maxA[ia];
            mB = MODE_CPU ? mA : p.maxB[ia];
            // 8-bit tiles (MatrixByte) or float tiles (MatrixFloat)
            const size_t at = (size_t)ia * N * N;
            #pragma unroll
            for (int k = 0; k < PIX; k++) {
                const int idx = tid + k * NT;
                va[k] = p.tAf ? p.tAf[at + idx] : (float)p.tA[at + idx];
                vb[k] = p.tBf ? p.tBf[at + idx] : (float)p.tB[at + idx];
            }
        }
        #pragma unroll
        for (int k = 0; k < PIX; k++) {
            const int idx = tid + k * NT;
            const int mi = N * N - 1 - idx;  // F'[r][c] = A[N-1-r][N-1-c]
            s_re[(mi / N) * S + (mi % N)] = va[k] / mA * 16.0f + 1.0f;
            s_im[(idx / N) * S + (idx % N)] = vb[k] / mB * 16.0f + 1.0f;
        }
        __syncthreads();

        // ---- transforms ----------------------------------------------------------
        // forward: rows (only the N data rows are non-zero), then all columns
        fft_lines<LOGM, true, false, NT>(s_re, s_im, S, 1, N, s_twf, tid);
        fft_lines<LOGM, true, false, NT>(s_re, s_im, 1, S, M, s_twf, tid);
        unpack_product<LOGM, MODE_CPU, NT>(s_re, s_im, tid);
        // inverse: all rows, then the D columns that reach the output
        fft_lines<LOGM, false, MODE_CPU, NT>(s_re, s_im, S, 1, M, s_twi, tid);
        fft_lines<LOGM, false, MODE_CPU, NT>(s_re, s_im, 1, S, D, s_twi, tid);

        // ---- epilogues -------------------------------------------------------------
        if (p.planes) {
            float *out = p.planes + (size_t)ia * D * D;
            for (int idx = tid; idx < D * D; idx += NT) {
                const int i = idx / D, j = idx - (idx / D) * D;
                out[idx] = s_re[i * S + j];
            }
        }
        if (!PEAK && !PERSIST && p.mp_ent) {  // MultiPeaks entries from the plane in LDS
            constexpr MpGeom MG = mp_geom(D);
            mp_entries_lds<D, S, NT>(s_re, s_min, p.mp_ent + (size_t)ia * MG.stride,
                                     p.mp_cmax + (size_t)ia * MG.nchunk, p.mp_min + ia, tid);
        }
        if (PEAK) {
            // FromCenter in two phases (from_center_tie): the largest value of
            // the visited cells (rows and columns 1..D-1), then the tie keys of
            // the cells that equal it
            constexpr int h = D / 2;
            float vmax = -1.0f, mn = 3.40282347e+38f;
            for (int idx = tid; idx < D * D; idx += NT) {
                const int i = idx / D, j = idx - (idx / D) * D;
                if (i == 0 || j == 0) continue;
                const float v = s_re[i * S + j];
                vmax = fmaxf(vmax, v);
                if (v < mn) mn = v;
            }
            vmax = block_max_float<NT>(vmax, s_min, tid);
            uint64_t best = 0;
            int bestCell = 0;
            if (vmax >= 0.0f) {
                for (int idx = tid; idx < D * D; idx += NT) {
                    const int i = idx / D, j = idx - (idx / D) * D;
                    if (i == 0 || j == 0 || !(s_re[i * S + j] == vmax)) continue;
                    const uint32_t tie = from_center_tie(i, j, h);
                    if (tie > best) {
                        best = tie;
                        bestCell = idx;
                    }
                }
            }
            peak_epilogue<NT, D>(p, ia, best, bestCell, mn, s_re, S, 1, s_key, s_cell, s_min, tid);
        }
        if (!PERSIST) break;
        __syncthreads();  // the next IA reuses the LDS tile
    }
}

// ---------------------------------------------------------------------------
// IA 128 (M = 256).  The 256 x 256 complex tile (512 KB) does not fit in LDS,
// so each work-group keeps it in its own slot of an HBM workspace (complex
// float2 values, row-major) and runs every 1D pass on batches of BL lines
// staged through LDS.  Butterflies, twiddle entries, stage order, the unpack
// and the epilogues are those of the in-LDS kernel above, so the planes are the
// same values.  The passes, in place in the slot:
//   forward rows     : the N data rows -> slot rows 0..N-1
//   forward columns  : batches of columns closed under negation (index k and
//                      M-k together), so that the two-for-one unpack and the
//                      product run in LDS; only rows 0..M/2 of the product are
//                      stored (the others are their conjugate mirror)
//   inverse rows     : batches of rows closed under negation; a row above M/2
//                      is read as the conjugate of its mirror row, reversed
//   inverse columns  : the D columns that reach the plane
// Work-groups loop over the IAs (grid = workspace slots).
namespace big {
constexpr int LOGM = 8, M = 256, N = 128, D = 255, LS = M + 1;
constexpr size_t SLOT_BYTES = (size_t)M * M * sizeof(float2);
}  // namespace big

size_t big_work_estimate(int n) { return (size_t)(n < 256 ? n : 256) * big::SLOT_BYTES; }

// Index (row or column) held by line l of batch k of a paired pass with BL
// lines: batch 0 holds 0 .. BL/2-1, their negatives M-1 .. M-BL/2+1, and M/2;
// batch k >= 1 holds BL/2*k .. BL/2*k + BL/2-1 and their negatives.
template <int BL>
__device__ __forceinline__ int paired_index(int k, int l)
{
    constexpr int H = BL / 2, M = big::M;
    if (k == 0) return l < H ? l : (l < BL - 1 ? M + H - 1 - l : M / 2);
    return l < H ? H * k + l : M + H - H * k - l;
}

// Line of the same batch that holds the negative of line l's index.
template <int BL>
__device__ __forceinline__ int paired_line(int k, int l)
{
    constexpr int H = BL / 2;
    if (k == 0) return (l == 0 || l == BL - 1) ? l : (l < H ? l + H - 1 : l - (H - 1));
    return l < H ? l + H : l - H;
}

// NT threads, batches of BL lines (two shapes: 512 x 32 and 256 x 16).
template <bool MODE_CPU, bool FROM_IMAGE, bool PEAK, int NT, int BL>
__global__ __launch_bounds__(NT) void xcorr_big_kernel(PassParams p,
                                                       const float2 *__restrict__ tw_fwd,
                                                       const float2 *__restrict__ tw_inv,
                                                       float2 *__restrict__ work)
{
    using namespace big;
    constexpr int NB = M / BL;  // batches of a full pass
    // loads per thread of one batch; each batch issues all of its loads before
    // the first LDS write, so that a thread has them in flight together
    constexpr int IT_HALF = BL * N / NT, IT_FULL = BL * M / NT, IT_MAX = N * N / NT;
    static_assert(IT_HALF * NT == BL * N && IT_FULL * NT == BL * M && IT_MAX * NT == N * N, "shape");
    __shared__ float s_re[BL * LS];
    __shared__ float s_im[BL * LS];
    __shared__ float2 s_twf[M / 2], s_twi[M / 2];
    __shared__ int s_red[NT / 64];
    __shared__ uint64_t s_key[NT / 64];
    __shared__ int s_cell[NT / 64];
    __shared__ float s_min[NT / 64];

    const int tid = threadIdx.x;
    for (int k = tid; k < M / 2; k += NT) {
        s_twf[k] = tw_fwd[k];
        s_twi[k] = tw_inv[k];
    }
    float2 *__restrict__ ws = work + (size_t)blockIdx.x * (SLOT_BYTES / sizeof(float2));

    const int n = pass_count(p);
    for (int ia = blockIdx.x; ia < n; ia += gridDim.x) {
        // ---- pixel sources (as in the in-LDS kernel) and tile maxima ----------
        const uint8_t *a8 = nullptr, *b8 = nullptr;
        int y0 = 0, x0 = 0, by0 = 0, bx0 = 0;
        float fu = 0.0f, fv = 0.0f;
        bool warp = false;
        float mA, mB;
        if (FROM_IMAGE) {
            const int per_pair = p.per_pair > 0 ? p.per_pair : n;
            const int pair = ia / per_pair, t = ia - pair * per_pair;
            const size_t img_off = (size_t)pair * p.H * p.W;
            a8 = p.imgA + img_off;
            b8 = p.imgB + img_off;
            y0 = p.top[t];
            x0 = p.left[t];
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
                    fu = u - flu;
                    fv = v - flv;
                }
            }
        } else {
            a8 = p.tA + (size_t)ia * N * N;
            b8 = p.tB + (size_t)ia * N * N;
        }
        // 8-bit tiles (MatrixByte) or float tiles (MatrixFloat)
        const float *af = p.tAf ? p.tAf + (size_t)ia * N * N : nullptr;
        const float *bf = p.tBf ? p.tBf + (size_t)ia * N * N : nullptr;
        auto pixA = [&](int r, int c) -> float {
            if (!FROM_IMAGE) return af ? af[r * N + c] : (float)a8[r * N + c];
            const int y = y0 + r, x = x0 + c;
            return (y >= 0 && y < p.H && x >= 0 && x < p.W) ? (float)a8[(size_t)y * p.W + x] : 0.0f;
        };
        auto pixB = [&](int r, int c) -> float {
            if (!FROM_IMAGE) return bf ? bf[r * N + c] : (float)b8[r * N + c];
            if (warp) return warp_pixel(b8, p.H, p.W, fu, fv, by0 + r, bx0 + c);
            const int y = by0 + r, x = bx0 + c;
            return (y >= 0 && y < p.H && x >= 0 && x < p.W) ? (float)b8[(size_t)y * p.W + x] : 0.0f;
        };
        if (FROM_IMAGE) {
            const bool b255 = warp || p.bmax255;
            int amax = 0, bmax = 0;
#pragma unroll 8
            for (int q = 0; q < IT_MAX; q++) {
                const int idx = tid + q * NT;
                const int r = idx / N, c = idx - (idx / N) * N;
                amax = max(amax, (int)pixA(r, c));
                if (!b255) bmax = max(bmax, (int)pixB(r, c));
            }
            amax = block_max_int<NT>(amax, s_red);
            mA = (float)(amax < 16 ? 16 : amax);
            if (b255) {
                mB = 255.0f;
            } else {
                bmax = block_max_int<NT>(bmax, s_red);
                mB = (float)(bmax < 16 ? 16 : bmax);
            }
            if (MODE_CPU) mB = mA;
        } else {
            mA = p.maxA[ia];
            mB = MODE_CPU ? mA : p.maxB[ia];
        }

        // ---- forward rows: the N data rows, F'[r][c] = A[N-1-r][N-1-c] -------
        for (int rb = 0; rb < N; rb += BL) {
            const int tid = opaque_tid();
            float va[IT_HALF], vb[IT_HALF];
            #pragma unroll
            for (int q = 0; q < IT_HALF; q++) {
                const int idx = tid + q * NT;
                const int l = idx / N, c = idx - (idx / N) * N, r = rb + l;
                va[q] = pixA(N - 1 - r, N - 1 - c);
                vb[q] = pixB(r, c);
            }
            #pragma unroll
            for (int q = 0; q < IT_HALF; q++) {
                const int idx = tid + q * NT;
                const int l = idx / N, c = idx - (idx / N) * N;
                s_re[l * LS + c] = va[q] / mA * 16.0f + 1.0f;
                s_im[l * LS + c] = vb[q] / mB * 16.0f + 1.0f;
            }
            __syncthreads();
            fft_lines<LOGM, true, false, NT, BL>(s_re, s_im, LS, 1, BL, s_twf);
            for (int idx = tid; idx < BL * M; idx += NT) {
                const int l = idx / M, c = idx - (idx / M) * M;
                ws[(size_t)(rb + l) * M + c] = make_float2(s_re[l * LS + c], s_im[l * LS + c]);
            }
            __syncthreads();
        }
        // ---- forward columns (rows >= N are zero padding) + unpack/product ----
        for (int k = 0; k < NB; k++) {
            const int tid = opaque_tid();
            float2 x[IT_HALF];
            #pragma unroll
            for (int q = 0; q < IT_HALF; q++) {
                const int idx = tid + q * NT;
                const int r = idx / BL, l = idx - (idx / BL) * BL;
                x[q] = ws[(size_t)r * M + paired_index<BL>(k, l)];
            }
            #pragma unroll
            for (int q = 0; q < IT_HALF; q++) {
                const int idx = tid + q * NT;
                const int r = idx / BL, l = idx - (idx / BL) * BL;
                s_re[l * LS + r] = x[q].x;
                s_im[l * LS + r] = x[q].y;
            }
            __syncthreads();
            fft_lines<LOGM, true, false, NT, BL>(s_re, s_im, LS, 1, BL, s_twf);
            // unpack (FastRealFFTXCorr.commonRealFFTXCorr :73-112): the owner of
            // each conjugate pair (line l, row r) / (paired line, -r) writes both
            for (int idx = tid; idx < BL * M; idx += NT) {
                const int l = idx / M, r = idx - (idx / M) * M;
                const int pl = paired_line<BL>(k, l), nr = (M - r) & (M - 1);
                if (l > pl || (l == pl && r > nr)) continue;
                const float x = s_re[l * LS + r], y = s_im[l * LS + r];
                const float Nx = s_re[pl * LS + nr], Ny = s_im[pl * LS + nr];
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
                s_re[l * LS + r] = pre;
                s_im[l * LS + r] = pim;
                s_re[pl * LS + nr] = pre;
                s_im[pl * LS + nr] = -pim;
            }
            __syncthreads();
            for (int idx = tid; idx < (M / 2 + 1) * BL; idx += NT) {
                const int r = idx / BL, l = idx - (idx / BL) * BL;
                ws[(size_t)r * M + paired_index<BL>(k, l)] = make_float2(s_re[l * LS + r], s_im[l * LS + r]);
            }
            __syncthreads();
        }
        // ---- inverse rows: rows above M/2 from their conjugate mirror ---------
        for (int k = 0; k < NB; k++) {
            const int tid = opaque_tid();
            float2 x[IT_FULL];
            #pragma unroll
            for (int q = 0; q < IT_FULL; q++) {
                const int idx = tid + q * NT;
                const int l = idx / M, c = idx - (idx / M) * M;
                const int r = paired_index<BL>(k, l);
                const bool mirror = r > M / 2;
                x[q] = ws[mirror ? (size_t)(M - r) * M + ((M - c) & (M - 1)) : (size_t)r * M + c];
                if (mirror) x[q].y = -x[q].y;
            }
            #pragma unroll
            for (int q = 0; q < IT_FULL; q++) {
                const int idx = tid + q * NT;
                const int l = idx / M, c = idx - (idx / M) * M;
                s_re[l * LS + c] = x[q].x;
                s_im[l * LS + c] = x[q].y;
            }
            __syncthreads();
            fft_lines<LOGM, false, MODE_CPU, NT, BL>(s_re, s_im, LS, 1, BL, s_twi);
            for (int idx = tid; idx < BL * D; idx += NT) {
                const int l = idx / D, c = idx - (idx / D) * D;
                ws[(size_t)paired_index<BL>(k, l) * M + c] = make_float2(s_re[l * LS + c], s_im[l * LS + c]);
            }
            __syncthreads();
        }
        // ---- inverse columns (the D that reach the plane) and epilogues -------
        uint64_t best = 0;
        int bestCell = 0;
        float mn = 3.40282347e+38f;
        float *out = p.planes ? p.planes + (size_t)ia * D * D : nullptr;
        for (int cb = 0; cb < D; cb += BL) {
            const int nl = min(BL, D - cb);
            const int tid = opaque_tid();
            float2 x[IT_FULL];
            #pragma unroll
            for (int q = 0; q < IT_FULL; q++) {
                const int idx = tid + q * NT;
                const int r = idx / BL, l = idx - (idx / BL) * BL;
                if (l < nl) x[q] = ws[(size_t)r * M + cb + l];
            }
            #pragma unroll
            for (int q = 0; q < IT_FULL; q++) {
                const int idx = tid + q * NT;
                const int r = idx / BL, l = idx - (idx / BL) * BL;
                if (l < nl) {
                    s_re[l * LS + r] = x[q].x;
                    s_im[l * LS + r] = x[q].y;
                }
            }
            __syncthreads();
            fft_lines<LOGM, false, MODE_CPU, NT, BL>(s_re, s_im, LS, 1, nl, s_twi);
            for (int idx = tid; idx < D * BL; idx += NT) {
                const int i = idx / BL, l = idx - (idx / BL) * BL, j = cb + l;
                if (l >= nl) continue;
                const float v = s_re[l * LS + i];
                if (out) out[(size_t)i * D + j] = v;
                if (PEAK) {
                    // the neighbourhood is read back from the plane, or from the
                    // slot's column just consumed when no plane is stored
                    if (!out) ws[(size_t)i * M + j].x = v;
                    from_center_visit(v, i, j, D / 2, i * D + j, best, bestCell, mn);
                }
            }
            __syncthreads();
        }
        if (PEAK) {
            if (out)
                peak_epilogue<NT, D>(p, ia, best, bestCell, mn, out, D, 1, s_key, s_cell, s_min);
            else
                peak_epilogue<NT, D>(p, ia, best, bestCell, mn, (const float *)ws, 2 * M, 2, s_key,
                                     s_cell, s_min);
        }
        __syncthreads();  // the slot is rewritten by the next IA
    }
}

// need != null: only report the workspace bytes of this pass.
template <bool MODE_CPU, bool FROM_IMAGE, bool PEAK, int NT, int BL>
static int launch_big_k(Context *ctx, const PassParams &p, size_t *need)
{
    auto k = xcorr_big_kernel<MODE_CPU, FROM_IMAGE, PEAK, NT, BL>;
    static const int groups = resident_groups(k, NT);
    const int slots = p.n < groups ? p.n : groups;
    if (need) {
        *need = (size_t)slots * big::SLOT_BYTES;
        return QLP_OK;
    }
    int rc = ensure_bigwork(ctx, (size_t)slots * big::SLOT_BYTES);
    if (rc) return rc;
    const float2 *twf = ctx->tables.tw[big::LOGM - 5][TW_FWD];
    const float2 *twi = ctx->tables.tw[big::LOGM - 5][MODE_CPU ? TW_INV_CPU : TW_INV_GPU];
    hipLaunchKernelGGL(k, dim3(slots), dim3(NT), 0, ctx->stream, p, twf, twi, (float2 *)ctx->bigwork);
    return check_hip(hipGetLastError(), "IA-128 pass launch");
}

template <bool MODE_CPU>
static int launch_pass_big(Context *ctx, const PassParams &p, bool from_image, bool peak,
                           size_t *need = nullptr)
{
    if (p.shape == SHAPE_LDS_HALF && !from_image)
        return peak ? launch_big_k<MODE_CPU, false, true, 256, 16>(ctx, p, need)
                    : launch_big_k<MODE_CPU, false, false, 256, 16>(ctx, p, need);
    if (from_image)
        return peak ? launch_big_k<MODE_CPU, true, true, 512, 32>(ctx, p, need)
                    : launch_big_k<MODE_CPU, true, false, 512, 32>(ctx, p, need);
    return peak ? launch_big_k<MODE_CPU, false, true, 512, 32>(ctx, p, need)
                : launch_big_k<MODE_CPU, false, false, 512, 32>(ctx, p, need);
}

size_t big_work_need(int n, int mode, bool from_image, bool peak, int shape)
{
    PassParams p;
    p.n = n;
    p.N = 128;
    p.mode = mode;
    p.shape = shape;
    size_t need = 0;
    if (mode == QLP_NORM_A_MAX)
        (void)launch_pass_big<true>(nullptr, p, from_image, peak, &need);
    else
        (void)launch_pass_big<false>(nullptr, p, from_image, peak, &need);
    return need;
}

int prepare_big(Context *ctx, int n, int mode, bool from_image, bool peak, int shape)
{
    return ensure_bigwork(ctx, big_work_need(n, mode, from_image, peak, shape));
}

// One launch of xcorr_pass_kernel: a work-group per IA, or for a persistent
// launch the resident work-groups.
template <int LOGM, bool MODE_CPU, bool FROM_IMAGE, bool PEAK, int NT>
static void launch_pass_k(Context *ctx, const PassParams &p, const float2 *twf, const float2 *twi)
{
    if constexpr (!FROM_IMAGE && PEAK) {
        if (p.n_dev) {  // persistent: the step engine's passes (tiles, fused peak)
            auto k = xcorr_pass_kernel<LOGM, MODE_CPU, false, true, NT, true>;
            hipLaunchKernelGGL(k, dim3(pass_grid(p, 1, resident_groups(k, NT))), dim3(NT), 0,
                               ctx->stream, p, twf, twi);
            return;
        }
    }
    hipLaunchKernelGGL((xcorr_pass_kernel<LOGM, MODE_CPU, FROM_IMAGE, PEAK, NT>), dim3(p.n),
                       dim3(NT), 0, ctx->stream, p, twf, twi);
}

template <int LOGM, bool MODE_CPU, bool FROM_IMAGE, int NT>
static void launch_pass_fp(Context *ctx, const PassParams &p, bool peak, const float2 *twf,
                           const float2 *twi)
{
    if (peak)
        launch_pass_k<LOGM, MODE_CPU, FROM_IMAGE, true, NT>(ctx, p, twf, twi);
    else
        launch_pass_k<LOGM, MODE_CPU, FROM_IMAGE, false, NT>(ctx, p, twf, twi);
}

template <int LOGM, bool MODE_CPU>
static void launch_pass_mc(Context *ctx, const PassParams &p, bool from_image, bool peak)
{
    using X = XcShape<LOGM>;
    const int li = LOGM - 5;
    const float2 *twf = ctx->tables.tw[li][TW_FWD];
    const float2 *twi = ctx->tables.tw[li][MODE_CPU ? TW_INV_CPU : TW_INV_GPU];
    // from tiles, the default shape is the wide one too (the C3 batch tuner measured
    // it faster at 32 and 64 px: 2.17 vs 2.31 ms and 3.21 vs 3.95 ms per class)
    if (!from_image && (p.shape == SHAPE_LDS_WIDE || (p.shape == SHAPE_AUTO && LOGM >= 6))) {
        launch_pass_fp<LOGM, MODE_CPU, false, X::THREADS * 2>(ctx, p, peak, twf, twi);
        return;
    }
    if (p.shape == SHAPE_LDS_HALF && !from_image) {
        launch_pass_fp<LOGM, MODE_CPU, false, X::THREADS / 2>(ctx, p, peak, twf, twi);
        return;
    }
    if constexpr (LOGM >= 6) {
        if (from_image) {
            // twice the default work-group: every line of a 1D pass in one round
            // (measured on MI355X, C2 step: 32 px pass 0.298 -> 0.267 ms, 64 px
            // 0.475 -> 0.400 ms)
            launch_pass_fp<LOGM, MODE_CPU, true, X::THREADS * 2>(ctx, p, peak, twf, twi);
            return;
        }
        launch_pass_fp<LOGM, MODE_CPU, false, X::THREADS>(ctx, p, peak, twf, twi);
    } else {
        if (from_image)
            launch_pass_fp<LOGM, MODE_CPU, true, X::THREADS>(ctx, p, peak, twf, twi);
        else
            launch_pass_fp<LOGM, MODE_CPU, false, X::THREADS>(ctx, p, peak, twf, twi);
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

static bool lane_form(const Context *ctx, const PassParams &p)
{
    const bool ftiles = p.tAf != nullptr || p.tBf != nullptr;
    return !ftiles && p.N == 16 &&
           (p.shape == SHAPE_LANE || (p.shape == SHAPE_AUTO && ctx->xcorr_form == FORM_LANE));
}

bool pass_writes_entries(const Context *ctx, const PassParams &p)
{
    (void)ctx;  // both exact forms of 16 px (lane and LDS) write them
    if (p.n_dev || p.peaks || p.min_floor || p.nbr) return false;  // persistent or fused peak
    // the 64 px matrix-core pass puts its plane into its free exchange buffer,
    // the 16 px one into an LDS slot per wavefront (3.8 KB), the 32 px one six
    // rows at a time into a band slot per wavefront (1.5 KB)
    if (p.arith != QLP_XCORR_EXACT) return p.N == 16 || p.N == 32 || p.N == 64;
    return p.N == 16 || p.N == 32 || p.N == 64;
}

int launch_pass(Context *ctx, const PassParams &p)
{
    if (p.n == 0) return QLP_OK;
    const bool from_image = p.imgA != nullptr;
    const bool peak = p.peaks != nullptr || p.min_floor != nullptr || p.nbr != nullptr;
    if (p.n_dev && (from_image || !peak || p.tAf || p.tBf))
        return fail(QLP_ERR_ARG, "pass: persistent passes take 8-bit tiles and the fused peak");
    if (p.nbr && (p.nbr_half < 0 || p.nbr_half > 7))
        return fail(QLP_ERR_ARG, "pass: neighbourhood half-width must be 0..7");
    // the passes from frames index them with 32-bit offsets (the reference's
    // frames are Java arrays, below 2^31 elements too)
    if (from_image && (long long)p.H * p.W > 0x7fffffffLL)
        return fail(QLP_ERR_UNSUPPORTED, "pass: frames of 2^31 pixels or more");
    // measured on MI355X (C2 multipass): the lane-per-line form is faster for
    // N = 16 (0.293 vs 0.323 ms for 65,025 IAs), the LDS-exchange form for
    // N = 32 (0.381 vs 0.436 ms for 16,129 IAs) and N = 64, where more
    // wavefronts per IA hide latency better
    // the matrix-core forms cover N = 16, 32 and 64; N = 128 keeps the exact form
    // float tiles (MatrixFloat) run the exact form in the in-LDS or IA-128 kernels
    const bool ftiles = p.tAf != nullptr || p.tBf != nullptr;
    if (ftiles && (from_image || p.arith != QLP_XCORR_EXACT))
        return fail(QLP_ERR_UNSUPPORTED, "xcorr: float tiles take the exact form from tiles");
    if (p.mp_ent && !pass_writes_entries(ctx, p))
        return fail(QLP_ERR_ARG, "pass: MultiPeaks entries come from the exact passes only");
    if (p.arith != QLP_XCORR_EXACT && p.N <= 64)
        return launch_pass_mfma(ctx, p, p.arith, from_image, peak);
    if (lane_form(ctx, p)) return launch_pass_lane(ctx, p, from_image, peak);
    switch (p.N) {
    case 16: launch_pass_m<5>(ctx, p, from_image, peak); break;
    case 32: launch_pass_m<6>(ctx, p, from_image, peak); break;
    case 64: launch_pass_m<7>(ctx, p, from_image, peak); break;
    case 128:
        return p.mode == QLP_NORM_A_MAX ? launch_pass_big<true>(ctx, p, from_image, peak)
                                        : launch_pass_big<false>(ctx, p, from_image, peak);
    default: return fail(QLP_ERR_UNSUPPORTED, "xcorr: IA size N must be 16, 32, 64 or 128");
    }
    return check_hip(hipGetLastError(), "xcorr pass launch");
}

// ---------------------------------------------------------------------------
// Allocation-free correlation (qlp_xcorr_batch_noalloc_dev).  The IA-128 tile
// slots live in this file-scope device array instead of the context's
// hipMalloc'd workspace: NOALLOC_SLOTS slots of 512 KB (32 MiB per GPU, held
// until the process ends), so the 128 px pass runs on at most NOALLOC_SLOTS
// work-groups, each looping over its IAs.  16 / 32 / 64 px passes keep their
// tile in LDS and need no slot.  Every call on a GPU shares that GPU's array:
// calls that overlap on one GPU race.  The slots need no clearing: a pass
// writes every slot row it later reads.
constexpr int NOALLOC_SLOTS = 64;
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)NOALLOC_SLOTS * big::SLOT_BYTES;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];
// the array's address on each device (hipGetSymbolAddress with that device
// current); host code uses these, never the array's name
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

int init_noalloc_scratch(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        return fail(QLP_ERR_DEVICE, "no-alloc scratch: device index beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return QLP_OK;
    void *ptr = nullptr;
    int rc = check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (rc) return rc;
    g_scratch_ptr[device] = ptr;
    return QLP_OK;
}

template <bool MODE_CPU>
static int launch_big_noalloc(Context *ctx, const PassParams &p, float2 *work)
{
    auto k = xcorr_big_kernel<MODE_CPU, false, false, 512, 32>;
    static const int groups = resident_groups(k, 512);
    int slots = p.n < groups ? p.n : groups;
    if (slots > NOALLOC_SLOTS) slots = NOALLOC_SLOTS;
    const float2 *twf = ctx->tables.tw[big::LOGM - 5][TW_FWD];
    const float2 *twi = ctx->tables.tw[big::LOGM - 5][MODE_CPU ? TW_INV_CPU : TW_INV_GPU];
    hipLaunchKernelGGL(k, dim3(slots), dim3(512), 0, ctx->stream, p, twf, twi, work);
    return check_hip(hipGetLastError(), "IA-128 pass launch");
}

int launch_xcorr_noalloc(Context *ctx, int n, int N, const uint8_t *tA, const uint8_t *tB,
                         const float *maxA, const float *maxB, int mode, float *planes)
{
    if (N != 16 && N != 32 && N != 64 && N != 128)
        return fail(QLP_ERR_ARG, "xcorr_noalloc: IA size N must be 16, 32, 64 or 128 (a 128 px "
                                 "tile fills one 512 KB scratch slot)");
    PassParams p;
    p.n = n;
    p.N = N;
    p.mode = mode;
    p.tA = tA;
    p.tB = tB;
    p.maxA = maxA;
    p.maxB = maxB;
    p.planes = planes;
    if (n == 0) return QLP_OK;
    if (N < 128) return launch_pass(ctx, p);  // the tile stays in LDS: no scratch
    void *ptr = nullptr;
    if (ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES) {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        ptr = g_scratch_ptr[ctx->device];
    }
    if (!ptr) return fail(QLP_ERR_DEVICE, "xcorr_noalloc: scratch address not set up for this device");
    return mode == QLP_NORM_A_MAX ? launch_big_noalloc<true>(ctx, p, (float2 *)ptr)
                                  : launch_big_noalloc<false>(ctx, p, (float2 *)ptr);
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
