/*
 * qlpiv_oracle.c -- CPU restatement of QuickLabPIV-ng's PIV correlation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is built into
 * oracle/_build/libqlpiv_oracle.so and loaded by tests/, by the smoke check of
 * __graft_entry__.py and by the cpu_baseline leg of bench.py.  Nothing in the
 * product library links or calls it.
 *
 * Every function restates one piece of the reference's Java code with the same
 * operation order and the same float/double types, so that compiled with
 * -O2 -ffp-contract=off -fno-fast-math (no FMA contraction, SSE float
 * arithmetic, FLT_EVAL_METHOD 0) it produces the values the JVM produces.
 * Paths below are relative to the reference's src/main/java/pt/quickLabPIV/.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   or_sine_table            - closed form sin, float rounding-boundary margin test
 *   or_perfect_shuffle       - equals the bit-reversal permutation (all sizes)
 *   or_xcorr_javacpu/_javagpu- direct-sum correlation (KAT 4x4 of FastRealFFTXCorrTests,
 *                              brute force on random tiles), FP64 FFT oracle, and the
 *                              reference's golden GPU planes (1024 x 31x31, tol 1e-2)
 *   or_find_max_from_center  - brute-force argmax with the documented tie-break
 *   or_multipeaks / or_list_*- FindMaximumMultiPeakTests KAT and the 14
 *                              SimpleFixedLengthFloatLinkedListTests sequences
 *   or_gauss1d               - 3 px: closed form; 5/7 px: the least-squares optimum
 *                              equals scipy.optimize.least_squares on noisy samples
 *                              (<= 1e-6 px); GaussianTestHelper cases (567, tol 8e-3)
 *   or_guo                   - exact on noise-free Gaussians; HongWeiGuo cases; one
 *                              iteration = data-weighted least squares (numpy lstsq);
 *                              20 iterations = the re-weighting fixed point on noisy
 *                              5-15 px data (tests/test_oracle_fits.py)
 *   pinv_solve3 (SVD branch) - minimum-norm least squares (numpy pinv with
 *                              commons-math's threshold) on full-rank, small-valued
 *                              and exactly rank-1/rank-2 matrices
 *   or_poly                  - closed form
 *   or_warp_micro            - identity at zero displacement, pure shift at
 *                              integer displacement, separable bilinear otherwise
 *   or_clip_tile             - direct slicing with zero fill; max clamp >= 16
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>



/* ------------------------------------------------------------------------ */
/* FFT building blocks: jobs/xcorr/SimpleFFT.java                            */
/* ------------------------------------------------------------------------ */

/* SimpleFFT.initEulerTable :323-330 and CrossCorrelationFFTTemplate.initEulerTable
 * :268-277: ws[n] = (float)sin(2.0f*PI*(float)n/(float)N), evaluated in double. */
void or_sine_table(int n, float *ws)
{
    for (int k = 0; k < n; k++)
        ws[k] = (float)sin(2.0 * M_PI * (double)(float)k / (double)(float)n);
}

/* SimpleFFT.perfectShuffleFFTInput :32-69, restated loop for loop. */
void or_perfect_shuffle(float *a, int n)
{
    for (int osf = 1, ods = n; ods > 2; ods >>= 1, osf <<= 1) {
        for (int osi = 0; osi < osf; osi++) {
            int oo = osi * ods;
            for (int sf = 1, ds = ods; ds > 2; ds >>= 1, sf <<= 1) {
                for (int si = 0; si < sf; si++) {
                    int io = si * ds;
                    for (int i = 0; i < ds / 2; i += 2) {
                        int src = oo + io + i + 1;
                        int dst = oo + io + ds / 2 + i;
                        float t = a[src];
                        a[src] = a[dst];
                        a[dst] = t;
                    }
                }
            }
        }
    }
}

/* SimpleFFT.computeFFTStep :122-221 (the "Option A" butterfly with both
 * results held in temporaries). */
static void fft_step(float *xr, float *xi, int step, int level, int level2,
                     int tidx, int N, const float *ws)
{
    int k = step * (tidx % level);
    int i = (tidx / level) * level2 + (tidx % level);
    int j = i + level;
    float wr = ws[k + N / 4];
    float wi = -ws[k];
    float tlr = xr[i] + (wr * xr[j]) - (wi * xi[j]);
    float tli = xi[i] + (wr * xi[j]) + (wi * xr[j]);
    float tur = xr[i] - (wr * xr[j]) + (wi * xi[j]);
    float tui = xi[i] - (wr * xi[j]) - (wi * xr[j]);
    xr[i] = tlr;
    xi[i] = tli;
    xr[j] = tur;
    xi[j] = tui;
}

/* SimpleFFT.computeFFTSerial :97-110 */
void or_fft_serial(float *xr, float *xi, int N, const float *ws)
{
    int step = N;
    for (int level = 1; level < N; level <<= 1) {
        int level2 = level << 1;
        step >>= 1;
        for (int tidx = 0; tidx < N / 2; tidx++)
            fft_step(xr, xi, step, level, level2, tidx, N, ws);
    }
}

/* SimpleFFT.computeIFFTStep :255-320 */
static void ifft_step(float *xr, float *xi, int step, int level, int level2,
                      int tidx, int N, const float *ws)
{
    int k = step * (tidx % level);
    int i = (tidx / level) * level2 + (tidx % level);
    int j = i + level;
    float wr;
    if (N / 4 - k >= 0)
        wr = ws[N / 4 - k];
    else
        wr = ws[N + (N / 4 - k)];
    float wi = ws[k];
    float tur = xr[i] - (wr * xr[j]) + (wi * xi[j]);
    float tui = xi[i] - (wr * xi[j]) - (wi * xr[j]);
    xr[i] = xr[i] + (wr * xr[j]) - (wi * xi[j]);
    xi[i] = xi[i] + (wr * xi[j]) + (wi * xr[j]);
    xr[j] = tur;
    xi[j] = tui;
}

/* SimpleFFT.computeIFFTSerial :229-242 (divideArray :333-337 first) */
void or_ifft_serial(float *xr, float *xi, int N, const float *ws)
{
    float fN = (float)N;
    for (int q = 0; q < N; q++) xr[q] /= fN;
    for (int q = 0; q < N; q++) xi[q] /= fN;
    int step = N;
    for (int level = 1; level < N; level <<= 1) {
        int level2 = level << 1;
        step >>= 1;
        for (int tidx = 0; tidx < N / 2; tidx++)
            ifft_step(xr, xi, step, level, level2, tidx, N, ws);
    }
}

/* One 1D pass over every row of an M x M pair of arrays:
 * shuffle(xr row), shuffle(xi row), then the transform.
 * kind 0: forward (computeFFTSerial), 1: CPU inverse (computeIFFTSerial),
 * 2: GPU-kernel inverse = forward transform on the (xi, xr) swapped pair
 *    (CrossCorrelationRealFFTParallelKernel.java:606-623, 672-688). */
static void rows_pass(float *xr, float *xi, int M, const float *ws, int kind)
{
    for (int r = 0; r < M; r++) {
        float *a = xr + (size_t)r * M, *b = xi + (size_t)r * M;
        or_perfect_shuffle(a, M);
        or_perfect_shuffle(b, M);
        if (kind == 0)
            or_fft_serial(a, b, M, ws);
        else if (kind == 1)
            or_ifft_serial(a, b, M, ws);
        else
            or_fft_serial(b, a, M, ws);
    }
}

static void transpose_sq(float *x, int M)
{
    for (int i = 0; i < M; i++)
        for (int j = i + 1; j < M; j++) {
            float t = x[(size_t)i * M + j];
            x[(size_t)i * M + j] = x[(size_t)j * M + i];
            x[(size_t)j * M + i] = t;
        }
}

/* SimpleFFT.computeFFT2DSerial :379-432: rows, transpose, rows, transpose back. */
void or_fft2d(float *xr, float *xi, int M, int kind)
{
    float *ws = (float *)malloc(sizeof(float) * M);
    or_sine_table(M, ws); /* computeFFT2D :435-436 rebuilds the table per call */
    rows_pass(xr, xi, M, ws, kind);
    transpose_sq(xr, M);
    transpose_sq(xi, M);
    rows_pass(xr, xi, M, ws, kind);
    transpose_sq(xr, M);
    transpose_sq(xi, M);
    free(ws);
}

/* ------------------------------------------------------------------------ */
/* Cross-correlation                                                         */
/* ------------------------------------------------------------------------ */

/* Matrix.copyMirroredMatrixToArrayNormalizeAndOffset (Matrix.java:282-288),
 * Matrix.copyMatrixToArrayAndNormalizeAndOffset (:175-181) and the inline CPU
 * variant (FastRealFFTXCorr.java:48-60): v / max * 16.0f + 1.0f, F mirrored,
 * zero padded to 2N x 2N. */
static void pack_pair(const float *A, const float *B, float maxF, float maxG,
                      int N, float *tr, float *ti)
{
    int M = 2 * N;
    memset(tr, 0, sizeof(float) * M * M);
    memset(ti, 0, sizeof(float) * M * M);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            tr[i * M + j] = A[(N - 1 - i) * N + (N - 1 - j)] / maxF * 16.0f + 1.0f;
            ti[i * M + j] = B[i * N + j] / maxG * 16.0f + 1.0f;
        }
}

/* FastRealFFTXCorr.commonRealFFTXCorr :70-117 unpack + product loop, in place. */
static void unpack_product_cpu(float *tr, float *ti, int M)
{
#define TR(a, b) tr[(a) * M + (b)]
#define TI(a, b) ti[(a) * M + (b)]
    for (int i = 0; i <= M / 2; i++) {
        for (int j = 0; j <= M / 2; j++) {
            int k1 = i, Nk1 = (M - k1) % M, k2 = j, Nk2 = (M - k2) % M;
            float Gr = 0.5f * (TI(k1, k2) + TI(Nk1, Nk2));
            float Gi = -0.5f * (TR(k1, k2) - TR(Nk1, Nk2));
            float Fr = 0.5f * (TR(k1, k2) + TR(Nk1, Nk2));
            float Fi = 0.5f * (TI(k1, k2) - TI(Nk1, Nk2));
            TR(k1, k2) = Fr * Gr - Fi * Gi;
            TI(k1, k2) = Fr * Gi + Fi * Gr;
            TR(Nk1, Nk2) = TR(k1, k2);
            TI(Nk1, Nk2) = -TI(k1, k2);
            if (k1 != 0 && k2 != 0 && k1 < M / 2 && k2 < M / 2) {
                k1 = i + M / 2;
                Nk1 = (M - k1) % M;
                float Gr2 = 0.5f * (TI(k1, k2) + TI(Nk1, Nk2));
                float Gi2 = -0.5f * (TR(k1, k2) - TR(Nk1, Nk2));
                float Fr2 = 0.5f * (TR(k1, k2) + TR(Nk1, Nk2));
                float Fi2 = 0.5f * (TI(k1, k2) - TI(Nk1, Nk2));
                TR(k1, k2) = Fr2 * Gr2 - Fi2 * Gi2;
                TI(k1, k2) = Fr2 * Gi2 + Fi2 * Gr2;
                TR(Nk1, Nk2) = TR(k1, k2);
                TI(Nk1, Nk2) = -TI(k1, k2);
            }
        }
    }
#undef TR
#undef TI
}

/* CrossCorrelationRealFFTParallelKernel.run :458-527 unpack + product with the
 * single /(2N)^2 scaling, restated serially (the kernel's per-row work-items
 * touch disjoint element pairs, so the serial order gives the same values). */
static void unpack_product_gpu(float *xr, float *xi, int M)
{
    const float scale = (float)(M * M);
    for (int k1 = 0; k1 <= M / 2; k1++) {
        int Nk1 = (M - k1) % M;
        int lk1 = Nk1, Nlk1 = k1;
        for (int k2 = 0; k2 <= M / 2; k2++) {
            int Nk2 = (M - k2) % M;
            float x = xr[k1 * M + k2], Nx = xr[Nk1 * M + Nk2];
            float y = xi[k1 * M + k2], Ny = xi[Nk1 * M + Nk2];
            float fr = 0.5f * (x + Nx);
            float fi = 0.5f * (y - Ny);
            float gr = 0.5f * (y + Ny);
            float gi = -0.5f * (x - Nx);
            float re = (fr * gr - fi * gi) / scale;
            float im = (fr * gi + fi * gr) / scale;
            xr[k1 * M + k2] = re;
            xi[k1 * M + k2] = im;
            xr[Nk1 * M + Nk2] = re;
            xi[Nk1 * M + Nk2] = -im;
        }
        for (int k2 = 0; k2 < M; k2++) {
            int Nk2 = (M - k2) % M;
            if (k1 != 0 && k2 != 0 && k1 < M / 2 && k2 < M / 2) {
                float x = xr[lk1 * M + k2], Nx = xr[Nlk1 * M + Nk2];
                float y = xi[lk1 * M + k2], Ny = xi[Nlk1 * M + Nk2];
                float fr = 0.5f * (x + Nx);
                float fi = 0.5f * (y - Ny);
                float gr = 0.5f * (y + Ny);
                float gi = -0.5f * (x - Nx);
                float re = (fr * gr - fi * gi) / scale;
                float im = (fr * gi + fi * gr) / scale;
                xr[lk1 * M + k2] = re;
                xi[lk1 * M + k2] = im;
                xr[Nlk1 * M + Nk2] = re;
                xi[Nlk1 * M + Nk2] = -im;
            }
        }
    }
}

/* FastRealFFTXCorr.computeXCorr(Matrix, Matrix) :44-68 -- the Java CPU path
 * (CrossCorrelationSoftRealFFTJob.compute :94-121): both tiles normalised by
 * A's max, IFFT dividing by 2N in each 1D pass.  A, B: N x N row-major floats
 * (tile pixel values); out: (2N-1) x (2N-1). */
void or_xcorr_javacpu(const float *A, const float *B, float maxA, int N, float *out)
{
    int M = 2 * N;
    float *tr = (float *)malloc(sizeof(float) * M * M);
    float *ti = (float *)malloc(sizeof(float) * M * M);
    pack_pair(A, B, maxA, maxA, N, tr, ti);
    or_fft2d(tr, ti, M, 0);
    unpack_product_cpu(tr, ti, M);
    or_fft2d(tr, ti, M, 1);
    for (int i = 0; i < M - 1; i++)
        for (int j = 0; j < M - 1; j++)
            out[i * (M - 1) + j] = tr[i * M + j];
    free(tr);
    free(ti);
}

/* FastRealFFTXCorr.computeXCorr(float[][], float[][]) :15-42 -- the
 * un-normalised overload used by FastRealFFTXCorrTests (A mirrored, B as is). */
void or_xcorr_floats(const float *A, const float *B, int N, float *out)
{
    int M = 2 * N;
    float *tr = (float *)calloc((size_t)M * M, sizeof(float));
    float *ti = (float *)calloc((size_t)M * M, sizeof(float));
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            tr[i * M + j] = A[(N - 1 - i) * N + (N - 1 - j)];
            ti[i * M + j] = B[i * N + j];
        }
    or_fft2d(tr, ti, M, 0);
    unpack_product_cpu(tr, ti, M);
    or_fft2d(tr, ti, M, 1);
    for (int i = 0; i < M - 1; i++)
        for (int j = 0; j < M - 1; j++)
            out[i * (M - 1) + j] = tr[i * M + j];
    free(tr);
    free(ti);
}

/* CrossCorrelationRealFFTParStdJob.analyzeTemplate :91-136 packing (F by its own
 * max, G by its own max) + CrossCorrelationRealFFTParallelKernel.run :214-713
 * arithmetic as executed with Java float semantics (Aparapi JTP). */
void or_xcorr_javagpu(const float *A, const float *B, float maxA, float maxB, int N,
                      float *out)
{
    int M = 2 * N;
    float *xr = (float *)malloc(sizeof(float) * M * M);
    float *xi = (float *)malloc(sizeof(float) * M * M);
    pack_pair(A, B, maxA, maxB, N, xr, xi);
    or_fft2d(xr, xi, M, 0);
    unpack_product_gpu(xr, xi, M);
    or_fft2d(xr, xi, M, 2);
    for (int i = 0; i < M - 1; i++)
        for (int j = 0; j < M - 1; j++)
            out[i * (M - 1) + j] = xr[i * M + j];
    free(xr);
    free(xi);
}

/* Batched helpers over u8 tiles (for tests and the CPU baseline):
 * mode 0 = Java CPU path (A max for both), mode 1 = GPU job path (own maxima). */
void or_xcorr_batch_u8(int n, int N, const uint8_t *tA, const uint8_t *tB,
                       const float *maxA, const float *maxB, int mode, float *planes)
{
    int D = 2 * N - 1;
    float *A = (float *)malloc(sizeof(float) * N * N);
    float *B = (float *)malloc(sizeof(float) * N * N);
    for (int t = 0; t < n; t++) {
        for (int q = 0; q < N * N; q++) {
            A[q] = (float)tA[(size_t)t * N * N + q];
            B[q] = (float)tB[(size_t)t * N * N + q];
        }
        if (mode == 0)
            or_xcorr_javacpu(A, B, maxA[t], N, planes + (size_t)t * D * D);
        else
            or_xcorr_javagpu(A, B, maxA[t], maxB[t], N, planes + (size_t)t * D * D);
    }
    free(A);
    free(B);
}

/* ------------------------------------------------------------------------ */
/* Peak search                                                               */
/* ------------------------------------------------------------------------ */

typedef struct {
    float i, j, value;
} or_cand;

/* MaxCrossResult.computeDistance :48-51 */
static float cand_dist(const or_cand *c, short ci, short cj)
{
    float di = c->i - ci, dj = c->j - cj;
    return (float)sqrt((double)(di * di + dj * dj));
}

/* FindMaximumFromCenter.findMaximum :47-183 with updateMaxResults :33-44.
 * out[0..3] = peak i, peak j, peak value, minFloor. */
/* FindMaximumSimple.findMaximum :19-48: row-major scan, strictly larger than the
 * running maximum (initially 0); no positive value -> (height/2, width/2) with
 * value 0; minFloor = the minimum (initially Float.MAX_VALUE). */
void or_find_max_simple(const float *m, int dimI, int dimJ, float *out)
{
    float mx = 0.0f, mn = 3.40282347e+38f;
    int pi = 0, pj = 0;
    for (int i = 0; i < dimI; i++)
        for (int j = 0; j < dimJ; j++) {
            float value = m[i * dimJ + j];
            if (value > mx) {
                mx = value;
                pi = i;
                pj = j;
            }
            if (value < mn) mn = value;
        }
    if (mx == 0.0f) {
        pi = dimI / 2;
        pj = dimJ / 2;
    }
    out[0] = (float)pi;
    out[1] = (float)pj;
    out[2] = mx;
    out[3] = mn;
}

void or_find_max_from_center(const float *m, int dimI, int dimJ, float *out)
{
#define EL(a, b) m[(a) * dimJ + (b)]
    float maximum = 0.0f;
    or_cand best = {0.0f, 0.0f, 0.0f};
    short ci = (short)(dimI / 2 + 1), cj = (short)(dimJ / 2 + 1);
#define UPDATE(II, JJ, VV)                                                    \
    do {                                                                      \
        or_cand nw = {(float)(II), (float)(JJ), (VV)};                        \
        if (nw.value > maximum) {                                             \
            best = nw;                                                        \
            maximum = nw.value;                                               \
        } else if (nw.value == maximum) {                                     \
            if (cand_dist(&nw, ci, cj) < cand_dist(&best, ci, cj)) best = nw; \
        }                                                                     \
    } while (0)
    UPDATE(dimI / 2 + 1, dimJ / 2 + 1, EL(dimI / 2 + 1, dimJ / 2 + 1));
    float minValue = 3.40282347e+38f;
    for (int i = 0; i < dimI / 2; i++) {
        float v1 = EL(dimI / 2 - i, dimJ / 2 + 1), v2 = EL(dimI / 2 + 1 + i, dimJ / 2 + 1);
        if (v1 < minValue) minValue = v1;
        if (v2 < minValue) minValue = v2;
        if (v1 >= maximum) UPDATE(dimI / 2 - i, dimJ / 2 + 1, v1);
        if (v2 >= maximum) UPDATE(dimI / 2 + 1 + i, dimJ / 2 + 1, v2);
    }
    for (int j = 0; j < dimJ / 2; j++) {
        float v1 = EL(dimI / 2 + 1, dimJ / 2 - j), v2 = EL(dimI / 2 + 1, dimJ / 2 + 1 + j);
        if (v1 < minValue) minValue = v1;
        if (v2 < minValue) minValue = v2;
        if (v1 >= maximum) UPDATE(dimI / 2 + 1, dimJ / 2 - j, v1);
        if (v2 >= maximum) UPDATE(dimI / 2 + 1, dimJ / 2 + 1 + j, v2);
    }
    for (int i = 0; i < dimI / 2; i++) {
        for (int j = 0; j < dimJ / 2; j++) {
            float v1 = EL(dimI / 2 - i, dimJ / 2 - j);
            float v2 = EL(dimI / 2 + 1 + i, dimJ / 2 - j);
            float v3 = EL(dimI / 2 - i, dimJ / 2 + 1 + j);
            float v4 = EL(dimI / 2 + 1 + i, dimJ / 2 + 1 + j);
            if (v1 < minValue) minValue = v1;
            if (v2 < minValue) minValue = v2;
            if (v3 < minValue) minValue = v3;
            if (v4 < minValue) minValue = v4;
            if (v1 >= maximum) UPDATE(dimI / 2 - i, dimJ / 2 - j, v1);
            if (v2 >= maximum) UPDATE(dimI / 2 + 1 + i, dimJ / 2 - j, v2);
            if (v3 >= maximum) UPDATE(dimI / 2 - i, dimJ / 2 + 1 + j, v3);
            if (v4 >= maximum) UPDATE(dimI / 2 + 1 + i, dimJ / 2 + 1 + j, v4);
        }
    }
#undef UPDATE
#undef EL
    out[0] = best.i;
    out[1] = best.j;
    out[2] = best.value;
    out[3] = minValue;
}

/* util/SimpleFixedLengthFloatLinkedList.java, restated with a node pool and
 * next indices that follow the Java object links step for step. */
#define OR_LIST_MAXP 64
typedef struct {
    short i, j;
    float value;
    int next;
} or_node;
typedef struct {
    int size, maxSize;
    float kernelDistance;
    int first, last, removed;
    or_node pool[OR_LIST_MAXP + 2];
    int npool;
    int freelist[OR_LIST_MAXP + 2];
    int nfree;
} or_list;

void or_list_init(or_list *L, int maxSize, int kernelWidth)
{
    memset(L, 0, sizeof(*L));
    L->maxSize = maxSize;
    L->kernelDistance = (float)((kernelWidth - 1) * kernelWidth); /* ctor :69-76 */
    L->first = L->last = L->removed = -1;
}

static int list_alloc(or_list *L)
{
    int f = L->removed; /* getFreeElement :78-86 */
    L->removed = -1;
    if (f < 0) {
        if (L->nfree > 0)
            f = L->freelist[--L->nfree];
        else
            f = L->npool++;
    }
    return f;
}

static void list_release(or_list *L, int idx)
{
    /* a node dropped from 'removed' by a later removal becomes garbage in Java */
    if (idx >= 0) L->freelist[L->nfree++] = idx;
}

/* computeDistance :49-51 */
static float list_dist(short i, short j, const or_node *e)
{
    return (float)(pow((double)(i - e->i), 2) + pow((double)(j - e->j), 2));
}

/* remove :239-258 (drops the smallest, i.e. the first entry) */
float or_list_remove(or_list *L)
{
    float v = L->pool[L->first].value;
    if (L->removed >= 0) list_release(L, L->removed);
    L->removed = L->first;
    if (L->first == L->last)
        L->first = L->last = -1;
    else
        L->first = L->pool[L->first].next;
    L->pool[L->removed].next = -1;
    L->size--;
    return v;
}

/* addInAscendOrder :98-237 */
int or_list_add_ascend(or_list *L, float value, short i, short j)
{
    if (L->first < 0 || L->last < 0) {
        int e = list_alloc(L);
        L->pool[e].value = value;
        L->pool[e].i = i;
        L->pool[e].j = j;
        L->pool[e].next = -1;
        L->first = L->last = e;
        L->size = 1;
        return 1;
    }
    int previous = -1, iter = L->first;
    int delTo[OR_LIST_MAXP + 2], delPrev[OR_LIST_MAXP + 2];
    int deleteSize = 0;
    while (iter >= 0 && L->pool[iter].value < value) {
        const or_node *e = &L->pool[iter];
        if (list_dist(i, j, e) < L->kernelDistance && value > e->value) { /* isToRemoveAdjacent :61-67 */
            delPrev[deleteSize] = previous;
            delTo[deleteSize] = iter;
            deleteSize++;
        }
        previous = iter;
        iter = L->pool[iter].next;
    }
    if (L->size == L->maxSize && iter == L->first) return 0;
    while (iter >= 0) {
        const or_node *e = &L->pool[iter];
        if (list_dist(i, j, e) < L->kernelDistance && value <= e->value) return 0; /* :53-59 */
        iter = L->pool[iter].next;
    }
    for (int d = 0; d < deleteSize; d++) {
        int deleted = delTo[d], deletedPrevious = delPrev[d];
        if (deleted == L->first) {
            if (L->first == L->last) {
                L->first = L->last = -1;
                previous = -1;
            } else {
                L->first = L->pool[L->first].next;
                if (previous == deleted) previous = deletedPrevious;
            }
            if (L->removed >= 0 && L->removed != deleted) list_release(L, L->removed);
            L->removed = deleted;
        } else {
            L->pool[deletedPrevious].next = L->pool[deleted].next;
            if (L->last == deleted) L->last = deletedPrevious;
            if (previous == deleted) previous = deletedPrevious;
            list_release(L, deleted);
        }
        if (d + 1 < deleteSize && deleted == delPrev[d + 1]) delPrev[d + 1] = deletedPrevious;
        L->pool[deleted].i = -1;
        L->pool[deleted].j = -1;
        L->pool[deleted].value = 0.0f;
        L->pool[deleted].next = -1;
        L->size--;
    }
    if (previous >= 0) {
        int ne = list_alloc(L);
        L->pool[ne].value = value;
        L->pool[ne].i = i;
        L->pool[ne].j = j;
        L->pool[ne].next = L->pool[previous].next;
        L->pool[previous].next = ne;
        if (previous == L->last) L->last = ne;
        L->size++;
        if (L->size > L->maxSize) or_list_remove(L);
    } else {
        if (L->size == L->maxSize) return 0;
        int e = list_alloc(L);
        L->pool[e].value = value;
        L->pool[e].i = i;
        L->pool[e].j = j;
        L->pool[e].next = -1;
        if (L->first < 0 || L->last < 0) {
            L->first = L->last = e;
        } else {
            L->pool[e].next = L->first;
            L->first = e;
        }
        L->size++;
    }
    return 1;
}

/* Dump the list in iteration order (ascending): out = [size, (value,i,j)*] */
int or_list_dump(const or_list *L, float *out)
{
    int n = 0;
    for (int it = L->first; it >= 0; it = L->pool[it].next) {
        out[3 * n + 0] = L->pool[it].value;
        out[3 * n + 1] = L->pool[it].i;
        out[3 * n + 2] = L->pool[it].j;
        n++;
    }
    return n;
}

/* Run a scripted sequence of operations (for the list KATs):
 * ops[k] = (opcode, value, i, j) with opcode 0 = addInAscendOrder, 1 = remove,
 * 2 = clear.  Returns final size and the dumped entries; accepted[k] holds the
 * boolean result of each add. */
int or_list_script(int maxSize, int kernelWidth, int nops, const float *ops, int *accepted,
                   float *out)
{
    or_list L;
    or_list_init(&L, maxSize, kernelWidth);
    for (int k = 0; k < nops; k++) {
        int op = (int)ops[4 * k];
        if (op == 0)
            accepted[k] = or_list_add_ascend(&L, ops[4 * k + 1], (short)ops[4 * k + 2],
                                             (short)ops[4 * k + 3]);
        else if (op == 1)
            or_list_remove(&L), accepted[k] = 1;
        else {
            while (L.size > 0) or_list_remove(&L); /* clear :367-371 */
            accepted[k] = 1;
        }
    }
    return or_list_dump(&L, out);
}

/* FindMaximumMultiPeaks.findMaximum :42-162.  out: P x (i, j, value) in
 * descending order (peak 0 is the main peak), entries beyond the found count
 * are zero; returns peaksFound; *minFloor receives the global minimum. */
int or_find_max_multipeaks(const float *m, int dimI, int dimJ, int numberOfPeaks,
                           int kernelWidth, float *out, float *minFloor)
{
#define EL(a, b) m[(a) * dimJ + (b)]
    short k = (short)kernelWidth;
    short iSlices = (short)(dimI / k);
    if (iSlices * k != dimI) iSlices += 1;
    short iSlicesBottom = (short)(iSlices / 2), iSlicesTop = (short)(iSlices - iSlicesBottom);
    short jSlices = (short)(dimJ / k);
    if (jSlices * k != dimJ) jSlices += 1;
    short jSlicesRight = (short)(jSlices / 2), jSlicesLeft = (short)(jSlices - jSlicesRight);
    (void)iSlicesBottom;
    (void)jSlicesRight;
    or_list L;
    or_list_init(&L, numberOfPeaks, kernelWidth);
    float maxs[4] = {0, 0, 0, 0};
    short maxsI[4] = {0, 0, 0, 0}, maxsJ[4] = {0, 0, 0, 0};
    float mn = 3.40282347e+38f;
    for (short si = 0; si < iSlicesTop; si++) {
        for (short sj = 0; sj < jSlicesLeft; sj++) {
            short top = (short)(si * k), left = (short)(sj * k);
            short bottom = (short)(si * k + iSlicesTop * k), right = (short)(sj * k + jSlicesLeft * k);
            for (int q = 0; q < 4; q++) maxs[q] = 0.0f;
            for (short oi = 0; oi < k; oi++) {
                for (short oj = 0; oj < k; oj++) {
                    short pI1 = (short)(top + oi), pJ1 = (short)(left + oj);
                    float v1 = EL(pI1, pJ1);
                    if (v1 > maxs[0]) { maxs[0] = v1; maxsI[0] = pI1; maxsJ[0] = pJ1; }
                    if (v1 < mn) mn = v1;
                    short pI2 = (short)(top + oi), pJ2 = (short)(right + oj);
                    if (pJ2 < dimJ) {
                        float v2 = EL(pI2, pJ2);
                        if (v2 > maxs[1]) { maxs[1] = v2; maxsI[1] = pI2; maxsJ[1] = pJ2; }
                        if (v2 < mn) mn = v2;
                    }
                    short pI3 = (short)(bottom + oi), pJ3 = (short)(left + oj);
                    if (pI3 < dimI) {
                        float v3 = EL(pI3, pJ3);
                        if (v3 > maxs[2]) { maxs[2] = v3; maxsI[2] = pI3; maxsJ[2] = pJ3; }
                        if (v3 < mn) mn = v3;
                    }
                    short pI4 = (short)(bottom + oi), pJ4 = (short)(right + oj);
                    if (pI4 < dimI && pJ4 < dimJ) {
                        float v4 = EL(pI4, pJ4);
                        if (v4 > maxs[3]) { maxs[3] = v4; maxsI[3] = pI4; maxsJ[3] = pJ4; }
                        if (v4 < mn) mn = v4;
                    }
                }
            }
            for (int q = 0; q < 4; q++) or_list_add_ascend(&L, maxs[q], maxsI[q], maxsJ[q]);
        }
    }
#undef EL
    float tmp[3 * (OR_LIST_MAXP + 2)];
    int found = or_list_dump(&L, tmp);
    for (int p = 0; p < numberOfPeaks; p++) out[3 * p] = out[3 * p + 1] = out[3 * p + 2] = 0.0f;
    int index = found - 1;
    for (int e = 0; e < found; e++, index--) {
        out[3 * index + 0] = tmp[3 * e + 1];
        out[3 * index + 1] = tmp[3 * e + 2];
        out[3 * index + 2] = tmp[3 * e + 0];
    }
    *minFloor = mn;
    return found;
}

/* ------------------------------------------------------------------------ */
/* Sub-pixel fits.  Peak records are (i, j, value); outputs (i', j', value',    */
/* status) with status bit0 = I fit accepted, bit1 = J fit accepted,          */
/* bit2 = skipped (border or (0,0)).                                          */
/* ------------------------------------------------------------------------ */

/* Closed-form 3-point Gaussian: the exact interpolant that the Levenberg-
 * Marquardt GaussianCurveFitter of Gaussian1DInterpolator (:78-94) converges to
 * for 3 samples. */
static double gauss3_mean(double x0, double ym, double y0, double yp)
{
    double lm = log(ym), l0 = log(y0), lp = log(yp);
    return x0 + (lm - lp) / (2.0 * lm - 4.0 * l0 + 2.0 * lp);
}

/* Least-squares Gaussian y = a exp(-(x-b)^2/(2c^2)) through n samples by
 * damped Gauss-Newton (Levenberg-Marquardt) in double, started from the
 * 3-point interpolant: the optimum that GaussianCurveFitter converges to. */
static int gauss_lsq(int n, const double *x, const double *y, double *par)
{
    int c = n / 2;
    double ymax = y[c];
    double b = gauss3_mean(x[c], y[c - 1], y[c], y[c + 1]);
    double l = log(y[c - 1]) - 2 * log(y[c]) + log(y[c + 1]);
    double s = (l < 0) ? sqrt(-1.0 / l) : 1.0;
    if (!isfinite(b) || fabs(b - x[c]) > 1.0) b = x[c];
    if (!isfinite(s) || s <= 0) s = 1.0;
    double a = ymax / exp(-(x[c] - b) * (x[c] - b) / (2 * s * s));
    double p[3] = {a, b, s};
    double lambda = 1e-3;
    double cost = 0;
    for (int k = 0; k < n; k++) {
        double r = y[k] - p[0] * exp(-(x[k] - p[1]) * (x[k] - p[1]) / (2 * p[2] * p[2]));
        cost += r * r;
    }
    for (int it = 0; it < 200; it++) {
        double JtJ[3][3] = {{0}}, Jtr[3] = {0};
        for (int k = 0; k < n; k++) {
            double d = x[k] - p[1];
            double e = exp(-d * d / (2 * p[2] * p[2]));
            double f = p[0] * e;
            double J[3] = {e, f * d / (p[2] * p[2]), f * d * d / (p[2] * p[2] * p[2])};
            double r = y[k] - f;
            for (int u = 0; u < 3; u++) {
                Jtr[u] += J[u] * r;
                for (int v = 0; v < 3; v++) JtJ[u][v] += J[u] * J[v];
            }
        }
        int improved = 0;
        for (int tries = 0; tries < 30 && !improved; tries++) {
            double Am[3][4];
            for (int u = 0; u < 3; u++) {
                for (int v = 0; v < 3; v++) Am[u][v] = JtJ[u][v] + (u == v ? lambda * JtJ[u][u] : 0.0);
                Am[u][3] = Jtr[u];
            }
            for (int col = 0; col < 3; col++) { /* Gaussian elimination, partial pivot */
                int piv = col;
                for (int r2 = col + 1; r2 < 3; r2++)
                    if (fabs(Am[r2][col]) > fabs(Am[piv][col])) piv = r2;
                for (int q = 0; q < 4; q++) {
                    double t = Am[col][q]; Am[col][q] = Am[piv][q]; Am[piv][q] = t;
                }
                for (int r2 = col + 1; r2 < 3; r2++) {
                    double fct = Am[r2][col] / Am[col][col];
                    for (int q = col; q < 4; q++) Am[r2][q] -= fct * Am[col][q];
                }
            }
            double dp[3];
            for (int r2 = 2; r2 >= 0; r2--) {
                double sum = Am[r2][3];
                for (int q = r2 + 1; q < 3; q++) sum -= Am[r2][q] * dp[q];
                dp[r2] = sum / Am[r2][r2];
            }
            double np[3] = {p[0] + dp[0], p[1] + dp[1], p[2] + dp[2]};
            double nc = 0;
            for (int k = 0; k < n; k++) {
                double r = y[k] - np[0] * exp(-(x[k] - np[1]) * (x[k] - np[1]) / (2 * np[2] * np[2]));
                nc += r * r;
            }
            if (isfinite(nc) && nc <= cost) {
                double rel = (cost - nc) / (cost > 0 ? cost : 1);
                p[0] = np[0]; p[1] = np[1]; p[2] = np[2];
                cost = nc;
                lambda *= 0.3;
                improved = 1;
                if (rel < 1e-15 && fabs(dp[1]) < 1e-13) it = 1000;
            } else {
                lambda *= 10;
            }
        }
        if (!improved) break;
    }
    par[0] = p[0];
    par[1] = p[1];
    par[2] = fabs(p[2]);
    return 0;
}

/* Gaussian1DInterpolator.interpolate :36-114 for one peak. */
void or_gauss1d(const float *m, int dimI, int dimJ, int pixels, float minFloor,
                const float *peak, float *res)
{
    int maxI = (int)peak[0], maxJ = (int)peak[1];
    res[0] = peak[0];
    res[1] = peak[1];
    res[2] = peak[2];
    res[3] = 4.0f;
    if (maxI == 0.0f && maxJ == 0.0f) return;
    if (maxI - pixels / 2 < 0 || maxI + pixels / 2 >= dimI) return;
    if (maxJ - pixels / 2 < 0 || maxJ + pixels / 2 >= dimJ) return;
    if (minFloor < -3.15f) { res[3] = 8.0f; return; } /* InterpolateException */
    float add = 0.0f;
    if (minFloor < 0.0f) add = -minFloor + 1e-8f;
    double xs[16], ys[16], par[3];
    int n = 0;
    for (int i = maxI - pixels / 2; i <= maxI + pixels / 2; i++, n++) {
        xs[n] = i;
        ys[n] = (double)(m[i * dimJ + maxJ] + add);
    }
    double meanI = (pixels == 3) ? gauss3_mean(xs[0] + 1, ys[0], ys[1], ys[2])
                                 : (gauss_lsq(n, xs, ys, par), par[1]);
    n = 0;
    for (int j = maxJ - pixels / 2; j <= maxJ + pixels / 2; j++, n++) {
        xs[n] = j;
        ys[n] = (double)(m[maxI * dimJ + j] + add);
    }
    double meanJ = (pixels == 3) ? gauss3_mean(xs[0] + 1, ys[0], ys[1], ys[2])
                                 : (gauss_lsq(n, xs, ys, par), par[1]);
    int st = 0;
    if (fabs(meanI - peak[0]) < 2.0f) { res[0] = (float)meanI; st |= 1; }
    if (fabs(meanJ - peak[1]) < 2.0f) { res[1] = (float)meanJ; st |= 2; }
    res[3] = (float)st;
}

/* commons-math3 3.6.1 LUDecomposition (partial pivoting, singularity
 * threshold 1e-11) and its solver, used by Gaussian1DHongweiGuoInterpolator
 * .estimateGaussianParameters :113-116.  Returns 0 when singular. */
static int lu_solve3(const double A[3][3], const double b[3], double x[3])
{
    double lu[3][3];
    int piv[3] = {0, 1, 2};
    memcpy(lu, A, sizeof(lu));
    for (int col = 0; col < 3; col++) {
        for (int row = 0; row < col; row++) {
            double sum = lu[row][col];
            for (int i = 0; i < row; i++) sum -= lu[row][i] * lu[i][col];
            lu[row][col] = sum;
        }
        int max = col;
        double largest = -INFINITY;
        for (int row = col; row < 3; row++) {
            double sum = lu[row][col];
            for (int i = 0; i < col; i++) sum -= lu[row][i] * lu[i][col];
            lu[row][col] = sum;
            if (fabs(sum) > largest) { largest = fabs(sum); max = row; }
        }
        if (fabs(lu[max][col]) < 1e-11) return 0;
        if (max != col) {
            for (int i = 0; i < 3; i++) {
                double t = lu[max][i]; lu[max][i] = lu[col][i]; lu[col][i] = t;
            }
            int t = piv[max]; piv[max] = piv[col]; piv[col] = t;
        }
        double d = lu[col][col];
        for (int row = col + 1; row < 3; row++) lu[row][col] /= d;
    }
    double bp[3];
    for (int r = 0; r < 3; r++) bp[r] = b[piv[r]];
    for (int col = 0; col < 3; col++) {
        double bc = bp[col];
        for (int i = col + 1; i < 3; i++) bp[i] -= bc * lu[i][col];
    }
    for (int col = 2; col >= 0; col--) {
        bp[col] /= lu[col][col];
        double bc = bp[col];
        for (int i = 0; i < col; i++) bp[i] -= bc * lu[i][col];
    }
    for (int r = 0; r < 3; r++) x[r] = bp[r];
    return 1;
}

/* Pseudo-inverse solve of the symmetric normal matrix through a cyclic Jacobi
 * eigen-decomposition: the SingularValueDecomposition fallback (:117-122).
 * commons-math3 3.6.1 SingularValueDecomposition keeps the singular values
 * above tol = max(m * s_max * 2^-52, sqrt(Double.MIN_NORMAL)) with m = 3 rows
 * and its solver applies the pseudo-inverse; for a symmetric matrix the
 * singular values are the absolute eigenvalues.  Pinned by the minimum-norm
 * least-squares solution (numpy pinv / lstsq) on well-conditioned,
 * small-valued and exactly rank-1 / rank-2 matrices; commons-math's own SVD
 * iteration is not restated, so last-ulp agreement is not claimed. */
static void pinv_solve3(const double A[3][3], const double b[3], double x[3])
{
    double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    memcpy(a, A, sizeof(a));
    for (int sweep = 0; sweep < 50; sweep++) {
        double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        if (off < 1e-300) break;
        for (int p = 0; p < 2; p++)
            for (int q = p + 1; q < 3; q++) {
                if (a[p][q] == 0.0) continue;
                double th = (a[q][q] - a[p][p]) / (2 * a[p][q]);
                double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1));
                double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < 3; k++) {
                    double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; k++) {
                    double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; k++) {
                    double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
    }
    double smax = 0;
    for (int k = 0; k < 3; k++) smax = fmax(smax, fabs(a[k][k]));
    double tol = fmax(3 * smax * 2.220446049250313e-16, 1.4916681462400413e-154);
    for (int r = 0; r < 3; r++) x[r] = 0;
    for (int k = 0; k < 3; k++) {
        double s = a[k][k];
        if (fabs(s) <= tol) continue;
        double proj = 0;
        for (int r = 0; r < 3; r++) proj += v[r][k] * b[r];
        for (int r = 0; r < 3; r++) x[r] += v[r][k] * proj / s;
    }
}

/* Gaussian1DHongweiGuoInterpolator.initMatrices :47-70 / updateMatrices :72-102
 * / estimateGaussianParameters :104-130.  xs, ys are float arrays (Java float[]);
 * the y^2 of the data term is a float product, everything else double. */
static void guo_coefs(int iterations, int n, const float *xs, const float *ys, double coefs[3])
{
    double A[3][3] = {{0}}, b[3] = {0};
    coefs[0] = coefs[1] = coefs[2] = 0.0;
    for (int i = 0; i < n; i++) {
        double ySquare = (double)(ys[i] * ys[i]);
        double xSquare = (double)(xs[i] * xs[i]);
        double xCubic = xSquare * xs[i];
        double xFourth = xCubic * xs[i];
        double logy = log((double)ys[i]);
        A[0][0] += ySquare;
        A[0][1] += xs[i] * ySquare;
        A[0][2] += xSquare * ySquare;
        b[0] += ySquare * logy;
        A[1][0] += xs[i] * ySquare;
        A[1][1] += xSquare * ySquare;
        A[1][2] += xCubic * ySquare;
        b[1] += xs[i] * ySquare * logy;
        A[2][0] += xSquare * ySquare;
        A[2][1] += xCubic * ySquare;
        A[2][2] += xFourth * ySquare;
        b[2] += xSquare * ySquare * logy;
    }
    for (int it = 0; it < iterations; it++) {
        if (!lu_solve3(A, b, coefs)) pinv_solve3(A, b, coefs);
        memset(A, 0, sizeof(A));
        memset(b, 0, sizeof(b));
        for (int i = 0; i < n; i++) {
            double yk = exp(coefs[0] + coefs[1] * xs[i] + coefs[2] * xs[i] * xs[i]);
            double ySquare = yk * yk;
            double xSquare = (double)(xs[i] * xs[i]);
            double xCubic = xSquare * xs[i];
            double xFourth = xCubic * xs[i];
            double logy = log((double)ys[i]);
            A[0][0] += ySquare;
            A[0][1] += xs[i] * ySquare;
            A[0][2] += xSquare * ySquare;
            b[0] += ySquare * logy;
            A[1][0] += xs[i] * ySquare;
            A[1][1] += xSquare * ySquare;
            A[1][2] += xCubic * ySquare;
            b[1] += xs[i] * ySquare * logy;
            A[2][0] += xSquare * ySquare;
            A[2][1] += xCubic * ySquare;
            A[2][2] += xFourth * ySquare;
            b[2] += xSquare * ySquare * logy;
        }
    }
}

static void guo_estimate(int iterations, int n, const float *xs, const float *ys, float *mean,
                         float *amp)
{
    double coefs[3];
    guo_coefs(iterations, n, xs, ys, coefs);
    *mean = (float)(-coefs[1] / (2 * coefs[2]));
    *amp = (float)(exp(coefs[0] - coefs[1] * coefs[1] / (4 * coefs[2])));
}

/* Test hooks: the Guo coefficients (a, b, c of ln y = a + b x + c x^2) after
 * 'iterations' solves, and the pseudo-inverse solve on its own (A row-major). */
void or_guo_coefs(int iterations, int n, const float *xs, const float *ys, double *coefs)
{
    guo_coefs(iterations, n, xs, ys, coefs);
}

void or_pinv_solve3(const double *A9, const double *b, double *x)
{
    double A[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) A[r][c] = A9[r * 3 + c];
    pinv_solve3(A, b, x);
}

/* 1 when commons-math's LUDecomposition finds the matrix singular (pivot
 * below 1e-11), i.e. when Guo takes the SVD fallback. */
int or_lu_singular3(const double *A9)
{
    double A[3][3], b[3] = {0, 0, 0}, x[3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) A[r][c] = A9[r * 3 + c];
    return !lu_solve3(A, b, x);
}

/* The least-squares Gaussian of the 5/7-point Gaussian1D fit: (a, b, c). */
void or_gauss_lsq(int n, const double *x, const double *y, double *par)
{
    gauss_lsq(n, x, y, par);
}

/* Gaussian1DHongweiGuoInterpolator.interpolate :133-248 for one peak. */
void or_guo(const float *m, int dimI, int dimJ, int pixels, int iterations, float minFloor,
            const float *peak, float *res)
{
    int maxI = (int)peak[0], maxJ = (int)peak[1];
    res[0] = peak[0];
    res[1] = peak[1];
    res[2] = peak[2];
    res[3] = 4.0f;
    if (maxI == 0.0f && maxJ == 0.0f) return;
    if (maxI - pixels / 2 < 0 || maxI + pixels / 2 >= dimI) return;
    if (maxJ - pixels / 2 < 0 || maxJ + pixels / 2 >= dimJ) return;
    if (minFloor < -5.15f) { res[3] = 8.0f; return; }
    float add = 0.0f;
    if (minFloor < 0.0f) add = -minFloor + 1e-8f;
    float xs[16], ys[16];
    int n = 0;
    for (int mi = maxI - pixels / 2; mi <= maxI + pixels / 2; mi++, n++) {
        xs[n] = (float)mi;
        ys[n] = m[mi * dimJ + maxJ] + add;
    }
    float yMean, yAmp, xMean, xAmp;
    guo_estimate(iterations, n, xs, ys, &yMean, &yAmp);
    n = 0;
    for (int mj = maxJ - pixels / 2; mj <= maxJ + pixels / 2; mj++, n++) {
        xs[n] = (float)mj;
        ys[n] = m[maxI * dimJ + mj] + add;
    }
    guo_estimate(iterations, n, xs, ys, &xMean, &xAmp);
    float finalY = yMean, finalX = xMean;
    int yValid = 1, xValid = 1;
    if (fabsf(yMean - (float)maxI) > pixels / 2.0f) { finalY = (float)maxI; yValid = 0; }
    if (fabsf(xMean - (float)maxJ) > pixels / 2.0f) { finalX = (float)maxJ; xValid = 0; }
    /* NaN means fail the '>' test and are kept, exactly as in Java */
    float mx = peak[2];
    if (yValid) {
        mx = yAmp;
        if (xValid && xAmp > mx) mx = xAmp;
    } else if (xValid) {
        mx = xAmp;
    }
    res[0] = finalY;
    res[1] = finalX;
    res[2] = mx;
    res[3] = (float)(yValid | (xValid << 1));
}

/* Gaussian1DPolynomialInterpolator.interpolate :25-106 for one peak. */
void or_poly(const float *m, int dimI, int dimJ, const float *peak, float *res)
{
    int maxI = (int)peak[0], maxJ = (int)peak[1];
    res[0] = peak[0];
    res[1] = peak[1];
    res[2] = peak[2];
    res[3] = 4.0f;
    if (maxI == 0.0f && maxJ == 0.0f) return;
    if (maxI - 1 < 0 || maxI + 1 >= dimI) return;
    if (maxJ - 1 < 0 || maxJ + 1 >= dimJ) return;
    int doGaussian = 1;
    float vy[3], vx[3];
    for (int i = -1; i < 2; i++)
        for (int j = -1; j < 2; j++) {
            float value = m[(maxI + i) * dimJ + (j + maxJ)];
            if (j == 0) {
                if (value < 0.0f) doGaussian = 0;
                vy[i + 1] = value + 1e-8f;
            }
            if (i == 0) {
                if (value < 0.0f) doGaussian = 0;
                vx[j + 1] = value + 1e-8f;
            }
        }
    float peakI = peak[0], peakJ = peak[1];
    int st = 0;
    if (doGaussian) {
        float ti = maxI + (float)(log(vy[0]) - log(vy[2])) /
                              (float)(2.0f * log(vy[0]) - 4.0f * log(vy[1]) + 2.0f * log(vy[2]));
        if (isfinite(ti) && fabsf(ti - maxI) <= 1.0f) { peakI = ti; st |= 1; }
        float tj = maxJ + (float)(log(vx[0]) - log(vx[2])) /
                              (float)(2.0f * log(vx[0]) - 4.0f * log(vx[1]) + 2.0f * log(vx[2]));
        if (isfinite(tj) && fabsf(tj - maxJ) <= 1.0f) { peakJ = tj; st |= 2; }
    } else {
        float ti = maxI + (float)(vy[0] - vy[2]) / (float)(2.0f * vy[0] - 4.0f * vy[1] + 2.0f * vy[2]);
        if (isfinite(ti) && fabsf(ti - maxI) <= 1.0f) { peakI = ti; st |= 1; }
        float tj = maxJ + (float)(vx[0] - vx[2]) / (float)(2.0f * vx[0] - 4.0f * vx[1] + 2.0f * vx[2]);
        if (isfinite(tj) && fabsf(tj - maxJ) <= 1.0f) { peakJ = tj; st |= 2; }
    }
    res[0] = peakI;
    res[1] = peakJ;
    res[3] = (float)st;
}

/* ------------------------------------------------------------------------ */
/* Tiles: clipping and bilinear micro warping                               */
/* ------------------------------------------------------------------------ */

/* Image.clipImageMatrix :386-494 (new matrix, partial clipping allowed) with the
 * MatrixByte maximum (clamped >= 16; MatrixByte.java:16-39, :278-317).  Returns
 * the tile max. */
float or_clip_tile(const uint8_t *img, int H, int W, int top, int left, int N, uint8_t *tile)
{
    int maxv = 0;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            int y = top + i, x = left + j;
            uint8_t v = 0;
            if (y >= 0 && y < H && x >= 0 && x < W) {
                v = img[(size_t)y * W + x];
                if (v > maxv) maxv = v;
            }
            tile[i * N + j] = v;
        }
    if (maxv < 16) maxv = 16;
    return (float)maxv;
}

/* FastMath.round(float) = (int)floor(x + 0.5f) (commons-math3 3.6.1). */
static int fm_round(float x) { return (int)floorf(x + 0.5f); }

/* BiLinearImageMicroAndMiniWarpingStrategy.warpPixelOnImage :336-407: the
 * bilinear sample at source = target + displacement (SecondImage order) or
 * target - displacement (FirstImage order, first != 0), round-half-up. */
/* A frame of 8-bit pixels (Image) or of floats (ImageFloat): readPixel. */
typedef struct {
    const uint8_t *u8;
    const float *f32;
    const uint16_t *u16;
} or_frame;

static float frame_px(const or_frame *fr, size_t k)
{
    return fr->f32 ? fr->f32[k] : (fr->u16 ? (float)fr->u16[k] : (float)fr->u8[k]);
}

static float warp_pixel_frame(const or_frame *img, int H, int W, float u, float v, short i, short j,
                              int first);

static float warp_pixel_order(const uint8_t *img, int H, int W, float u, float v, short i, short j,
                              int first)
{
    const or_frame fr = {img, NULL, NULL};
    return warp_pixel_frame(&fr, H, W, u, v, i, j, first);
}

static float warp_pixel_frame(const or_frame *img, int H, int W, float u, float v, short i, short j,
                              int first)
{
    float sI = i + u, sJ = j + v;
    if (first) {
        sI = i - u;
        sJ = j - v;
    }
    short pI = (short)fm_round(sI), pJ = (short)fm_round(sJ);
    if (pI < 0) { sI = 0; pI = 0; }
    else if (pI >= H) { sI = H - 1; pI = (short)(H - 1); }
    if (pJ < 0) { sJ = 0; pJ = 0; }
    else if (pJ >= W) { sJ = W - 1; pJ = (short)(W - 1); }
    float dU = sI - pI, dV = sJ - pJ;
    short oI = dU >= 0 ? 1 : -1, oJ = dV >= 0 ? 1 : -1;
    if (oI + pI < 0 || oI + pI >= H) oI = 0;
    if (oJ + pJ < 0 || oJ + pJ >= W) oJ = 0;
    float wU = fabsf(dU), wV = fabsf(dV);
    float tl = frame_px(img, (size_t)pI * W + pJ);
    float tr = frame_px(img, (size_t)pI * W + pJ + oJ);
    float bl = frame_px(img, (size_t)(pI + oI) * W + pJ);
    float br = frame_px(img, (size_t)(pI + oI) * W + pJ + oJ);
    float f = (1.0f - wU) * (1.0f - wV) * tl + (1.0f - wU) * wV * tr + wU * (1.0f - wV) * bl +
              wU * wV * br;
    return (float)fm_round(f);
}

static float warp_pixel(const uint8_t *img, int H, int W, float u, float v, short i, short j)
{
    return warp_pixel_order(img, H, W, u, v, i, j, 0);
}

/* Which frame a tile warp samples, and how (BiLinearImageWarpingModeEnum,
 * warpAndClipImage :63-167): the B tile in SecondImage order, the A tile in
 * FirstImage order, and the A and B tiles in BothImages order. */
enum { OR_SIDE_B_SECOND = 0, OR_SIDE_A_FIRST = 1, OR_SIDE_A_BOTH = 2, OR_SIDE_B_BOTH = 3 };

typedef struct {
    short top, left;   /* origin of the warped window */
    float offU, offV;  /* BothImages offsets */
    int half, first;   /* halve the field and add the offsets; sample at target - field */
} or_side;

/* warpAndClipImage :65-86 (FirstImage A), :87-123 (BothImages A), :132-148
 * (SecondImage B), :149-167 (BothImages B), for a tile at (top, left) whose B
 * tile carries the displacement (dispU, dispV). */
static or_side side_geometry(int side, int top, int left, float dispU, float dispV)
{
    or_side g = {0, 0, 0.0f, 0.0f, 0, 0};
    switch (side) {
    case OR_SIDE_A_FIRST:
        g.top = (short)(top - floor(dispU));
        g.left = (short)(left - floor(dispV));
        g.first = 1;
        break;
    case OR_SIDE_A_BOTH:
        g.top = (short)(top - floor(dispU / 2.0f));
        g.left = (short)(left - floor(dispV / 2.0f));
        g.offU = (float)(dispU / 2.0f - floor(dispU / 2.0f)) - (float)(dispU - floor(dispU));
        g.offV = (float)(dispV / 2.0f - floor(dispV / 2.0f)) - (float)(dispV - floor(dispV));
        g.half = 1;
        g.first = 1;
        break;
    case OR_SIDE_B_BOTH:
        g.top = (short)(top + floor(dispU / 2.0f));
        g.left = (short)(left + floor(dispV / 2.0f));
        g.offU = (float)(dispU / 2.0f - floor(dispU / 2.0f));
        g.offV = (float)(dispV / 2.0f - floor(dispV / 2.0f));
        g.half = 1;
        break;
    default: /* OR_SIDE_B_SECOND */
        g.top = (short)(top + floor(dispU));
        g.left = (short)(left + floor(dispV));
        break;
    }
    return g;
}

/* doBiLinearWarp :292-334 for one pixel of a side: the field value (u, v) at
 * the pixel, halved and offset in BothImages order, sampled in the side's order. */
static uint8_t side_pixel(const uint8_t *img, int H, int W, const or_side *g, float u, float v,
                          int i, int j)
{
    if (g->half) {
        u = u / 2.0f + g->offU;
        v = v / 2.0f + g->offV;
    }
    return (uint8_t)warp_pixel_order(img, H, W, u, v, (short)(i + g->top), (short)(j + g->left),
                                     g->first);
}

/* Micro warp of one tile (warpAndClipImage :39-186, createMicroDisplacements
 * :188-201, doBiLinearWarp :292-334) for a side (OR_SIDE_*): every pixel takes
 * the fractional part of the B tile's displacement.  Output tile values are
 * integers 0..255; the reference gives a warped tile the maximum 255
 * (image.getMaximumValue(), :298). */
void or_warp_micro_side(const uint8_t *img, int H, int W, int top, int left, int N, float dispU,
                        float dispV, int side, uint8_t *tile)
{
    const or_side g = side_geometry(side, top, left, dispU, dispV);
    const float fu = dispU - (float)floor(dispU);
    const float fv = dispV - (float)floor(dispV);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) tile[i * N + j] = side_pixel(img, H, W, &g, fu, fv, i, j);
}

/* ImageFloat.clipImageMatrix :278-384 (new matrix, partial clipping allowed)
 * with the MatrixFloat maximum over the tile (MatrixFloat.java:45-60,
 * :375-400; a maximum of 0 becomes 1, Matrix.setMaxValue :21-29).  Returns the
 * tile max. */
float or_clip_tile_f32(const float *img, int H, int W, int top, int left, int N, float *tile)
{
    float maxv = 0.0f;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            int y = top + i, x = left + j;
            float v = 0.0f;
            if (y >= 0 && y < H && x >= 0 && x < W) {
                v = img[(size_t)y * W + x];
                if (v > maxv) maxv = v;
            }
            tile[i * N + j] = v;
        }
    return maxv == 0.0f ? 1.0f : maxv;
}

/* Micro warp of one tile of a float frame (the same strategy as
 * or_warp_micro_side, reading ImageFloat.readPixel :531-533).  Values are the
 * rounded samples; the reference gives the tile the image maximum 65535
 * (ImageFloat.getMaximumValue :541-543, doBiLinearWarp :298), returned here. */
float or_warp_micro_side_f32(const float *img, int H, int W, int top, int left, int N, float dispU,
                             float dispV, int side, float *tile)
{
    const or_frame fr = {NULL, img, NULL};
    const or_side g = side_geometry(side, top, left, dispU, dispV);
    const float fu = dispU - (float)floor(dispU);
    const float fv = dispV - (float)floor(dispV);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            float u = fu, v = fv;
            if (g.half) {
                u = u / 2.0f + g.offU;
                v = v / 2.0f + g.offV;
            }
            tile[i * N + j] = warp_pixel_frame(&fr, H, W, u, v, (short)(i + g.top), (short)(j + g.left),
                                               g.first);
        }
    return 65535.0f;
}

/* ImageInt16.clipImageMatrix (images/ImageInt16.java, MatrixInt16 tiles):
 * the MatrixByte rules with 16-bit samples, the maximum over the in-frame
 * pixels clamped to >= 16 (MatrixInt16.java:16-41, copySubMatrixFromArray
 * :371-405).  Tile values are returned as floats. */
float or_clip_tile_u16(const uint16_t *img, int H, int W, int top, int left, int N, float *tile)
{
    int maxv = 0;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            int y = top + i, x = left + j;
            int v = 0;
            if (y >= 0 && y < H && x >= 0 && x < W) {
                v = img[(size_t)y * W + x];
                if (v > maxv) maxv = v;
            }
            tile[i * N + j] = (float)v;
        }
    if (maxv < 16) maxv = 16;
    return (float)maxv;
}

/* Micro warp of one tile of a 10/12/16-bit frame: the rounded samples; the
 * tile maximum is the frame's bit-depth maximum (ImageInt16.getMaximumValue
 * :557-559, ImageBitDepthEnum: 1023, 4095, 65535), returned here. */
float or_warp_micro_side_u16(const uint16_t *img, int H, int W, int bits, int top, int left, int N,
                             float dispU, float dispV, int side, float *tile)
{
    const or_frame fr = {NULL, NULL, img};
    const or_side g = side_geometry(side, top, left, dispU, dispV);
    const float fu = dispU - (float)floor(dispU);
    const float fv = dispV - (float)floor(dispV);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            float u = fu, v = fv;
            if (g.half) {
                u = u / 2.0f + g.offU;
                v = v / 2.0f + g.offV;
            }
            tile[i * N + j] = warp_pixel_frame(&fr, H, W, u, v, (short)(i + g.top), (short)(j + g.left),
                                               g.first);
        }
    return (float)((1 << bits) - 1);
}

/* SecondImage order: the B tile shifted by floor(u), floor(v), sampled at the
 * fractional part. */
void or_warp_micro(const uint8_t *img, int H, int W, int top, int left, int N, float dispU,
                   float dispV, uint8_t *tile)
{
    or_warp_micro_side(img, H, W, top, left, N, dispU, dispV, OR_SIDE_B_SECOND, tile);
}

/* ------------------------------------------------------------------------ */
/* The reference CPU path per IA, for the timed CPU baseline:               */
/* CrossCorrelationSoftRealFFTJob.compute :94-121 (FastRealFFTXCorr),        */
/* FindMaximumFromCenter.findMaximum :47-183, then the sub-pixel fit         */
/* (fit 0 = Gaussian1D 3 px, 1 = Guo 3 px / 20 iterations) on one thread.    */
/* out: n x 4 (i', j', value', status).                                      */
/* ------------------------------------------------------------------------ */
void or_cpu_path_batch(int n, int N, const uint8_t *tA, const uint8_t *tB, const float *maxA,
                       int fit, float *out)
{
    int D = 2 * N - 1;
    float *A = (float *)malloc(sizeof(float) * N * N);
    float *B = (float *)malloc(sizeof(float) * N * N);
    float *plane = (float *)malloc(sizeof(float) * D * D);
    for (int t = 0; t < n; t++) {
        for (int q = 0; q < N * N; q++) {
            A[q] = (float)tA[(size_t)t * N * N + q];
            B[q] = (float)tB[(size_t)t * N * N + q];
        }
        or_xcorr_javacpu(A, B, maxA[t], N, plane);
        float pk[4];
        or_find_max_from_center(plane, D, D, pk);
        if (fit == 0)
            or_gauss1d(plane, D, D, 3, pk[3], pk, out + 4 * (size_t)t);
        else
            or_guo(plane, D, D, 3, 20, pk[3], pk, out + 4 * (size_t)t);
    }
    free(A);
    free(B);
    free(plane);
}

/* One adaptive level of the reference CPU pipeline per IA, for the timed CPU
 * baseline of the multi-pass workload: clip tile A, micro-warp tile B by the
 * inherited displacement (SecondImage), correlate with the Java CPU job
 * (A-max normalisation), FromCenter, Guo 3 px / 20 iterations.
 * out: n x 4 (i', j', value', status). */
void or_cpu_level_batch_multi(const uint8_t *imgA, const uint8_t *imgB, int H, int W, int n,
                              int N, const int *top, const int *left, const float *u,
                              const float *v, int P, int k, float *out, int *count);

void or_cpu_level_batch(const uint8_t *imgA, const uint8_t *imgB, int H, int W, int n, int N,
                        const int *top, const int *left, const float *u, const float *v,
                        float *out)
{
    or_cpu_level_batch_multi(imgA, imgB, H, W, n, N, top, left, u, v, 1, 3, out, NULL);
}

/* The same with FindMaximumMultiPeaks (numberOfPeaks = P > 1, kernelWidth k)
 * and the Guo fit of every peak found (Gaussian1DHongweiGuoInterpolator
 * .interpolate :134-170); P = 1 keeps FindMaximumFromCenter.
 * out: n x P x 4; count: n peaks found (MaxCrossResult.getTotalPeaks), may be NULL. */
void or_cpu_level_batch_multi(const uint8_t *imgA, const uint8_t *imgB, int H, int W, int n,
                              int N, const int *top, const int *left, const float *u,
                              const float *v, int P, int k, float *out, int *count)
{
    int D = 2 * N - 1;
    uint8_t *ta = (uint8_t *)malloc((size_t)N * N), *tb = (uint8_t *)malloc((size_t)N * N);
    float *A = (float *)malloc(sizeof(float) * N * N);
    float *B = (float *)malloc(sizeof(float) * N * N);
    float *plane = (float *)malloc(sizeof(float) * D * D);
    for (int t = 0; t < n; t++) {
        float mA = or_clip_tile(imgA, H, W, top[t], left[t], N, ta);
        or_warp_micro(imgB, H, W, top[t], left[t], N, u[t], v[t], tb);
        for (int q = 0; q < N * N; q++) {
            A[q] = (float)ta[q];
            B[q] = (float)tb[q];
        }
        or_xcorr_javacpu(A, B, mA, N, plane);
        if (P <= 1) {
            float pk[4];
            or_find_max_from_center(plane, D, D, pk);
            or_guo(plane, D, D, 3, 20, pk[3], pk, out + 4 * (size_t)t);
            if (count) count[t] = 1;
        } else {
            float pk[3 * OR_LIST_MAXP], mf;
            int found = or_find_max_multipeaks(plane, D, D, P, k, pk, &mf);
            for (int q = 0; q < P; q++) {
                float *o = out + 4 * ((size_t)t * P + q);
                if (q < found) {
                    or_guo(plane, D, D, 3, 20, mf, pk + 3 * q, o);
                } else {
                    o[0] = o[1] = o[2] = o[3] = 0.0f;
                }
            }
            if (count) count[t] = found > 1 ? found : 1;
        }
    }
    free(ta);
    free(tb);
    free(A);
    free(B);
    free(plane);
}

/* One level of the per-frame loop (LocalPIVOpenCLGpuJob.call :188-409) for n
 * IAs of one frame pair, in any configuration the device plan runs:
 *   warp       0: B clipped at (top, left) (the first level); 1: B clipped at
 *              (round(top + u), round(left + v)) (NoImageWarpingStrategy
 *              :79-81); 2: B micro-warped by (u, v), maximum 255
 *   mode       0: CrossCorrelationRealFFTParStdJob arithmetic (each tile by its
 *              own maximum); 1: CrossCorrelationSoftRealFFTJob (both by A's)
 *   peak_kind  0: FindMaximumFromCenter; 1: FindMaximumMultiPeaks(P, k);
 *              2: FindMaximumSimple
 *   fit_kind   -1: none (the integer peak, status 4); 0: Gaussian1D(pixels);
 *              1: Guo(pixels, iterations); 2: Polynomial -- on every peak found
 * planes (n x D x D) may be NULL; out: n x P x 4 (i', j', value', status; peaks
 * not found are zero); count: n peaks found (at least 1), may be NULL. */
void or_level_batch(const uint8_t *imgA, const uint8_t *imgB, int H, int W, int n, int N,
                    const int *top, const int *left, const float *u, const float *v, int warp,
                    int mode, int peak_kind, int P, int k, int fit_kind, int pixels,
                    int iterations, float *planes, float *out, int *count)
{
    const int D = 2 * N - 1;
    uint8_t *ta = (uint8_t *)malloc((size_t)N * N), *tb = (uint8_t *)malloc((size_t)N * N);
    float *A = (float *)malloc(sizeof(float) * N * N);
    float *B = (float *)malloc(sizeof(float) * N * N);
    float *own = (float *)malloc(sizeof(float) * D * D);
    float pk[3 * OR_LIST_MAXP + 1];
    for (int t = 0; t < n; t++) {
        float *plane = planes ? planes + (size_t)t * D * D : own;
        const float mA = or_clip_tile(imgA, H, W, top[t], left[t], N, ta);
        float mB;
        if (warp == 2) {
            or_warp_micro(imgB, H, W, top[t], left[t], N, u[t], v[t], tb);
            mB = 255.0f;
        } else if (warp == 1) {
            const int tt = (int)floorf((float)top[t] + u[t] + 0.5f);
            const int ll = (int)floorf((float)left[t] + v[t] + 0.5f);
            mB = or_clip_tile(imgB, H, W, tt, ll, N, tb);
        } else {
            mB = or_clip_tile(imgB, H, W, top[t], left[t], N, tb);
        }
        for (int q = 0; q < N * N; q++) {
            A[q] = (float)ta[q];
            B[q] = (float)tb[q];
        }
        if (mode == 0)
            or_xcorr_javagpu(A, B, mA, mB, N, plane);
        else
            or_xcorr_javacpu(A, B, mA, N, plane);
        int found = 1;
        float mf;
        if (peak_kind == 1) {
            found = or_find_max_multipeaks(plane, D, D, P, k, pk, &mf);
        } else {
            float r[4];
            if (peak_kind == 2)
                or_find_max_simple(plane, D, D, r);
            else
                or_find_max_from_center(plane, D, D, r);
            pk[0] = r[0];
            pk[1] = r[1];
            pk[2] = r[2];
            mf = r[3];
        }
        const int np = peak_kind == 1 ? P : 1;
        for (int q = 0; q < np; q++) {
            float *o = out + 4 * ((size_t)t * np + q);
            if (q >= found) {
                o[0] = o[1] = o[2] = o[3] = 0.0f;
                continue;
            }
            const float *pq = pk + 3 * q;
            if (fit_kind == 0)
                or_gauss1d(plane, D, D, pixels, mf, pq, o);
            else if (fit_kind == 1)
                or_guo(plane, D, D, pixels, iterations, mf, pq, o);
            else if (fit_kind == 2)
                or_poly(plane, D, D, pq, o);
            else {
                o[0] = pq[0];
                o[1] = pq[1];
                o[2] = pq[2];
                o[3] = 4.0f;
            }
        }
        if (count) count[t] = found > 1 ? found : 1;
    }
    free(ta);
    free(tb);
    free(A);
    free(B);
    free(own);
}

/* ------------------------------------------------------------------------ */
/* Vector validation: jobs/validator/VectorValidatorJob.java,               */
/* iareas/validation/{NormalizedMedian,Difference}Validator.java,           */
/* iareas/replacement/{SecondaryPeak,Bilinear}Replacement.java, Tile.java   */
/* ------------------------------------------------------------------------ */

typedef struct {
    int validator;   /* 0 normalised median, 1 difference */
    float threshold;
    float epsilon0;
    int replacement; /* 0 none, 1 secondary peak, 2 bilinear */
    int peak_index;
} or_vstage;

typedef struct {
    int nI, nJ;
    const int *top, *left;
    float *u, *v, *bu, *bv;
    unsigned char *inv, *locked;
    int P;
    const float *puv;  /* nI*nJ x P x 2 */
    const int *pcount; /* or NULL */
    int err;
} or_vgrid;

/* VectorValidatorJob.circularVisitor :105-196 + visitTileHelper :198-213.
 * order[k] = row-major index of the k-th visited tile.  Returns the number of
 * tiles visited, or -1 when a tile would be visited twice (the reference
 * throws VectorValidatorException). */
int or_spiral_order(int nI, int nJ, int *order)
{
    int total = nI * nJ;
    int startI = nI / 2, startJ = nJ / 2;
    int sideI = 2, sideJ = 2;
    if (nI != startI * 2) sideI = 1; else startI--;
    if (nJ != startJ * 2) sideJ = 1; else startJ--;
    unsigned char *seen = (unsigned char *)calloc((size_t)total, 1);
    int visited = 0, bad = 0;
    int moves[4];
#define OR_VISIT(ii, jj)                                                     \
    do {                                                                     \
        int vi_ = (ii), vj_ = (jj);                                          \
        if (vi_ >= 0 && vi_ < nI && vj_ >= 0 && vj_ < nJ) {                  \
            if (seen[vi_ * nJ + vj_]) bad = 1;                               \
            seen[vi_ * nJ + vj_] = 1;                                        \
            if (visited < total) order[visited] = vi_ * nJ + vj_;            \
            visited++;                                                       \
        }                                                                    \
    } while (0)
    while (visited < total && !bad) {
        moves[0] = sideJ > 1 ? sideJ - 1 : 0;
        moves[1] = sideI > 1 ? sideI - 1 : 0;
        moves[2] = sideJ > 1 ? sideJ - 1 : 0;
        moves[3] = sideI > 2 ? sideI - 2 : 0;
        int j = startJ, i = startI;
        OR_VISIT(i, j);
        for (int o = 0; o < moves[0]; o++) { ++j; OR_VISIT(i, j); }
        for (int o = 0; o < moves[1]; o++) { ++i; OR_VISIT(i, j); }
        if (moves[1] > 0) {
            for (int o = 0; o < moves[2]; o++) { --j; OR_VISIT(i, j); }
            for (int o = 0; o < moves[3]; o++) { --i; OR_VISIT(i, j); }
        }
        sideI += 2;
        sideJ += 2;
        startJ--;
        startI--;
    }
#undef OR_VISIT
    free(seen);
    return bad ? -1 : visited;
}

/* Tile.setInvalidDisplacement :204-209 */
static void vt_set_invalid(or_vgrid *g, int t, int state)
{
    if (g->locked[t]) return;
    g->inv[t] = (unsigned char)state;
}

/* VectorValidatorJob.getNeighbors :215-238: adj[a][b] = tile index or -1 */
static void vt_neighbors(const or_vgrid *g, int i, int j, int adj[3][3])
{
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            int ii = i + (a - 1), jj = j + (b - 1);
            adj[a][b] = (ii < 0 || ii >= g->nI || jj < 0 || jj >= g->nJ) ? -1 : ii * g->nJ + jj;
        }
}

/* distanceComparator of NormalizedMedianValidator :19-35 */
static int vt_cmp(const or_vgrid *g, int t1, int t2)
{
    float u1 = g->u[t1], v1 = g->v[t1], u2 = g->u[t2], v2 = g->v[t2];
    float n1 = u1 * u1 + v1 * v1;
    float n2 = u2 * u2 + v2 * v2;
    if (n1 > n2) return 1;
    if (n1 < n2) return -1;
    return 0;
}

/* java.util.TimSort (OpenJDK 8) for fewer than 32 elements:
 * countRunAndMakeAscending, then binarySort (Collections.sort -> ArrayList.sort
 * -> Arrays.sort(T[], lo, hi, Comparator)). */
static void vt_timsort_small(const or_vgrid *g, int *a, int n)
{
    if (n < 2) return;
    int runHi = 1;
    if (vt_cmp(g, a[runHi++], a[0]) < 0) {
        while (runHi < n && vt_cmp(g, a[runHi], a[runHi - 1]) < 0) runHi++;
        for (int lo = 0, hi = runHi - 1; lo < hi; lo++, hi--) {
            int t = a[lo]; a[lo] = a[hi]; a[hi] = t;
        }
    } else {
        while (runHi < n && vt_cmp(g, a[runHi], a[runHi - 1]) >= 0) runHi++;
    }
    for (int start = runHi; start < n; start++) {
        int pivot = a[start];
        int left = 0, right = start;
        while (left < right) {
            int mid = (int)((unsigned)(left + right) >> 1);
            if (vt_cmp(g, pivot, a[mid]) < 0) right = mid; else left = mid + 1;
        }
        memmove(a + left + 1, a + left, sizeof(int) * (size_t)(start - left));
        a[left] = pivot;
    }
}

/* Arrays.sort(float[]): NaN after every number */
static int vt_fcmp(const void *x, const void *y)
{
    float a = *(const float *)x, b = *(const float *)y;
    int na = isnan(a), nb = isnan(b);
    if (na || nb) return na - nb;
    return (a > b) - (a < b);
}

/* NormalizedMedianValidator.validateVector :63-136 */
static void vt_validate_median(or_vgrid *g, int t, int adj[3][3], float threshold, float epsilon0)
{
    int sorted[8], cnt = 0;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            if (a == 1 && b == 1) continue;
            if (adj[a][b] < 0) continue;
            sorted[cnt++] = adj[a][b];
        }
    vt_timsort_small(g, sorted, cnt);
    float medianU = 0.0f, medianV = 0.0f;
    if (cnt > 0) {
        if (cnt % 2 == 0) {
            medianU = (g->u[sorted[cnt / 2 - 1]] + g->u[sorted[cnt / 2]]) / 2.0f;
            medianV = (g->v[sorted[cnt / 2 - 1]] + g->v[sorted[cnt / 2]]) / 2.0f;
        } else {
            medianU = g->u[sorted[cnt / 2]];
            medianV = g->v[sorted[cnt / 2]];
        }
    }
    float rIs[8];
    for (int k = 0; k < 8; k++) rIs[k] = 3.40282347e+38f;
    int rIndex = 0;
    for (int k = 0; k < cnt; k++) {
        float dU = medianU - g->u[sorted[k]];
        float dV = medianV - g->v[sorted[k]];
        float rI = dU * dU + dV * dV;
        rIs[rIndex++] = rI;
    }
    qsort(rIs, 8, sizeof(float), vt_fcmp);
    float rMed;
    if (rIndex - 1 == 0) {
        rMed = 1.0f;
    } else if ((rIndex - 1) % 2 == 0) {
        rMed = (float)sqrt((double)((rIs[(rIndex - 1) / 2] + rIs[(rIndex - 1) / 2 + 1]) / 2.0f));
    } else {
        rMed = (float)sqrt((double)rIs[(rIndex - 1) / 2 + 1]);
    }
    float dU = medianU - g->u[t];
    float dV = medianV - g->v[t];
    float distance = (float)sqrt((double)(dU * dU + dV * dV)) / (rMed + epsilon0);
    vt_set_invalid(g, t, distance < threshold ? 0 : 1);
}

/* DifferenceValidator.validateVector :34-63 (its loops stop at index 1, so it
 * consults the upper-left, upper and left neighbours only) */
static void vt_validate_difference(or_vgrid *g, int t, int adj[3][3], float distanceThreshold)
{
    float threshold = distanceThreshold * distanceThreshold;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            if (a == 1 && b == 1) continue;
            int n = adj[a][b];
            if (n < 0 || g->inv[n]) continue;
            float du = g->u[n] - g->u[t];
            float dv = g->v[n] - g->v[t];
            float distance = du * du + dv * dv;
            vt_set_invalid(g, t, distance < threshold ? 0 : 1);
        }
}

static void vt_validate(or_vgrid *g, const or_vstage *s, int t, int adj[3][3])
{
    if (s->validator == 0) vt_validate_median(g, t, adj, s->threshold, s->epsilon0);
    else vt_validate_difference(g, t, adj, s->threshold);
}

/* SecondaryPeakReplacement.replaceVector :41-75 */
static void vt_replace_secondary(or_vgrid *g, int t, int peakIndex)
{
    if (!g->inv[t]) return;
    int total = g->pcount ? g->pcount[t] : g->P;
    if (total <= peakIndex) return;
    const float *pk = g->puv + (size_t)t * g->P * 2;
    float u = g->u[t], v = g->v[t];
    float uOld = pk[0], vOld = pk[1];
    u -= uOld;
    v -= vOld;
    float uNew = pk[2 * peakIndex], vNew = pk[2 * peakIndex + 1];
    u += uNew;
    v += vNew;
    g->u[t] = u; /* Tile.replaceDisplacement, out-of-bound clipping allowed */
    g->v[t] = v;
    vt_set_invalid(g, t, 0);
}

/* BilinearReplacement.interpV :31-80 (horizontal, left pixels) */
static float vt_interpV(or_vgrid *g, int t1, int t2, int ti)
{
    float v1 = g->v[t1], v2 = g->v[t2];
    float leftPixel1 = (float)g->left[t1], leftPixel2 = (float)g->left[t2];
    float leftPixelInterp = (float)g->left[ti];
    float weightedV1 = 0.0f, weightedV2 = 0.0f;
    if ((leftPixel2 - leftPixelInterp) > 1e-6f && (leftPixelInterp - leftPixel1) > 1e-6f &&
        (leftPixel2 - leftPixel1) > 1e-6f) {
        float weightV1 = (leftPixel2 - leftPixelInterp) / (leftPixel2 - leftPixel1);
        if (weightV1 < 0.0f || weightV1 > 1.0f) g->err = 1;
        weightedV1 = weightV1 * v1;
        float weightV2 = (leftPixelInterp - leftPixel1) / (leftPixel2 - leftPixel1);
        if (weightV2 < 0.0f || weightV2 > 1.0f) g->err = 1;
        weightedV2 = weightV2 * v2;
    } else {
        int count = 0;
        if (fabsf(leftPixelInterp - leftPixel2) < 1e-6f) { weightedV2 = v2; count++; }
        if (fabsf(leftPixelInterp - leftPixel1) < 1e-6f) { weightedV1 = v1; count++; }
        if (count == 0) g->err = 1;
        if (count > 1) {
            weightedV1 /= (float)count;
            weightedV2 /= (float)count;
        }
    }
    return weightedV1 + weightedV2;
}

/* BilinearReplacement.interpU :82-135 (vertical, top pixels) */
static float vt_interpU(or_vgrid *g, int t1, int t2, int ti)
{
    float u1 = g->u[t1], u2 = g->u[t2];
    float topPixel1 = (float)g->top[t1], topPixel2 = (float)g->top[t2];
    float topPixelInterp = (float)g->top[ti];
    float weightedU1 = 0.0f, weightedU2 = 0.0f;
    if ((topPixel2 - topPixelInterp) > 1e-6f && (topPixelInterp - topPixel1) > 1e-6f &&
        (topPixel2 - topPixel1) > 1e-6f) {
        float weightU1 = (topPixel2 - topPixelInterp) / (topPixel2 - topPixel1);
        if (weightU1 < 0.0f || weightU1 > 1.0f) g->err = 1;
        weightedU1 = weightU1 * u1;
        float weightU2 = (topPixelInterp - topPixel1) / (topPixel2 - topPixel1);
        if (weightU2 < 0.0f || weightU2 > 1.0f) g->err = 1;
        weightedU2 = weightU2 * u2;
    } else {
        int count = 0;
        if (fabsf(topPixelInterp - topPixel2) < 1e-6f) { weightedU2 = u2; count++; }
        if (fabsf(topPixelInterp - topPixel1) < 1e-6f) { weightedU1 = u1; count++; }
        if (count == 0) g->err = 1;
        if (count > 1) {
            weightedU1 /= (float)count;
            weightedU2 /= (float)count;
        }
    }
    return weightedU1 + weightedU2;
}

#define VT_OK(x) ((x) >= 0 && !g->inv[x])

/* BilinearReplacement.replaceVector :145-427 */
static void vt_replace_bilinear(or_vgrid *g, int t, int adj[3][3])
{
    if (!g->inv[t]) return;
    int left = adj[1][0], right = adj[1][2], top = adj[0][1], bottom = adj[2][1];
    int topLeft = adj[0][0], bottomRight = adj[2][2], bottomLeft = adj[2][0], topRight = adj[0][2];
    float uV[4] = {0, 0, 0, 0}, uU[4] = {0, 0, 0, 0};
    int classificationU[4] = {0, 0, 0, 0}, classificationV[4] = {0, 0, 0, 0};
    int count = 0;
    if (topLeft >= 0 && bottomLeft >= 0 && (float)g->left[topLeft] - (float)g->left[bottomLeft] > 1.0e-6f) g->err = 1;
    if (topRight >= 0 && bottomRight >= 0 && (float)g->left[topRight] - (float)g->left[bottomRight] > 1.0e-6f) g->err = 1;
    if (topLeft >= 0 && topRight >= 0 && (float)g->top[topLeft] - (float)g->top[topRight] > 1.0e-6f) g->err = 1;
    if (bottomLeft >= 0 && bottomRight >= 0 && (float)g->top[bottomLeft] - (float)g->top[bottomRight] > 1.0e-6f) g->err = 1;
    if (VT_OK(left) && VT_OK(right)) {
        uV[0] = vt_interpV(g, left, right, t);
        uU[0] = vt_interpU(g, left, right, t);
        classificationV[0] = 2;
        classificationU[0] = 1;
        count++;
    }
    if (VT_OK(bottom) && VT_OK(top)) {
        uV[1] = vt_interpV(g, top, bottom, t);
        uU[1] = vt_interpU(g, top, bottom, t);
        classificationV[1] = 1;
        classificationU[1] = 2;
        count++;
    }
    if (VT_OK(topLeft) && VT_OK(bottomRight) && VT_OK(bottomLeft) && VT_OK(topRight)) {
        float wTR = (float)(g->left[t] - g->left[topLeft]) / (float)(g->left[topRight] - g->left[topLeft]);
        if (wTR < 0.0f || wTR > 1.0f) g->err = 1;
        float wTL = (float)(g->left[topRight] - g->left[t]) / (float)(g->left[topRight] - g->left[topLeft]);
        if (wTL < 0.0f || wTL > 1.0f) g->err = 1;
        float wBR = (float)(g->left[t] - g->left[bottomLeft]) / (float)(g->left[bottomRight] - g->left[bottomLeft]);
        if (wBR < 0.0f || wBR > 1.0f) g->err = 1;
        float wBL = (float)(g->left[bottomRight] - g->left[t]) / (float)(g->left[bottomRight] - g->left[bottomLeft]);
        if (wBL < 0.0f || wBL > 1.0f) g->err = 1;
        float wT = (float)(g->top[bottomLeft] - g->top[t]) / (float)(g->top[bottomLeft] - g->top[topLeft]);
        if (wT < 0.0f || wT > 1.0f) g->err = 1;
        float wB = (float)(g->top[t] - g->top[topLeft]) / (float)(g->top[bottomLeft] - g->top[topLeft]);
        if (wB < 0.0f || wB > 1.0f) g->err = 1;
        if (fabs((double)(1.0f - wTR - wTL)) > 1.0e-6) g->err = 1;
        if (fabs((double)(1.0f - wBR - wBL)) > 1.0e-6) g->err = 1;
        if (fabs((double)(1.0f - wT - wB)) > 1.0e-6) g->err = 1;
        uV[2] = wT * (wTR * g->v[topRight] + wTL * g->v[topLeft]) +
                wB * (wBR * g->v[bottomRight] + wBL * g->v[bottomLeft]);
        uU[2] = wT * (wTR * g->u[topRight] + wTL * g->u[topLeft]) +
                wB * (wBR * g->u[bottomRight] + wBL * g->u[bottomLeft]);
        classificationV[2] = 3;
        classificationU[2] = 3;
        count++;
    } else {
        if (VT_OK(topLeft) && VT_OK(bottomRight)) {
            uV[2] = vt_interpV(g, topLeft, bottomRight, t);
            uU[2] = vt_interpU(g, topLeft, bottomRight, t);
            classificationV[2] = 1;
            classificationU[2] = 1;
            count++;
        }
        if (VT_OK(bottomLeft) && VT_OK(topRight)) {
            uV[3] = vt_interpV(g, bottomLeft, topRight, t);
            uU[3] = vt_interpU(g, topRight, bottomLeft, t);
            classificationV[3] = 1;
            classificationU[3] = 1;
            count++;
        }
    }
    if (count == 0) {
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) {
                if (i == 1 && j == 1) continue;
                if (VT_OK(adj[i][j])) {
                    uV[0] = g->v[adj[i][j]];
                    uU[0] = g->u[adj[i][j]];
                    count = 1;
                    break;
                }
            }
        }
        if (count == 0) {
            vt_set_invalid(g, t, 1);
        } else {
            g->u[t] = uU[0];
            g->v[t] = uV[0];
            vt_set_invalid(g, t, 0);
        }
    } else {
        float accumV = 0.0f, accumU = 0.0f;
        int weightU = 0, weightV = 0;
        for (int i = 0; i < 4; i++) {
            accumV += (float)classificationV[i] * uV[i];
            weightV += classificationV[i];
            accumU += (float)classificationU[i] * uU[i];
            weightU += classificationU[i];
        }
        accumV /= (float)weightV;
        accumU /= (float)weightU;
        g->u[t] = accumU;
        g->v[t] = accumV;
        vt_set_invalid(g, t, 0);
    }
}
#undef VT_OK

/* VectorValidatorJob.compute :243-405 over one nI x nJ grid.
 *   u, v    : tile displacements (in/out)
 *   puv     : nI*nJ x P x 2 peak displacement increments (peak 0 the main peak)
 *   pcount  : MaxCrossResult.getTotalPeaks per tile, or NULL (= P)
 *   invalid : Tile.isInvalidDisplacement after the job
 *   stats   : sweeps, corrected vectors, vectors left invalid,
 *             flags (1 sweep limit, 2 replacement exception, 4 spiral error)
 * The sweep limit (max_sweeps) bounds the loop the reference leaves open. */
void or_validate(int nI, int nJ, const int *top, const int *left, float *u, float *v, int P,
                 const float *puv, const int *pcount, int n_stages, const or_vstage *stages,
                 int max_iterations, int until_no_more, int max_sweeps, unsigned char *invalid,
                 int *stats)
{
    int n = nI * nJ;
    or_vgrid G;
    or_vgrid *g = &G;
    g->nI = nI;
    g->nJ = nJ;
    g->top = top;
    g->left = left;
    g->u = u;
    g->v = v;
    g->bu = (float *)malloc(sizeof(float) * n);
    g->bv = (float *)malloc(sizeof(float) * n);
    g->inv = (unsigned char *)malloc((size_t)n);
    g->locked = (unsigned char *)calloc((size_t)n, 1);
    g->P = P;
    g->puv = puv;
    g->pcount = pcount;
    g->err = 0;
    int *order = (int *)malloc(sizeof(int) * n);
    int flags = 0;
    if (or_spiral_order(nI, nJ, order) != n) {
        if (stats) {
            stats[0] = stats[1] = stats[2] = 0;
            stats[3] = 4;
        }
        free(order);
        free(g->bu);
        free(g->bv);
        free(g->inv);
        free(g->locked);
        return;
    }
    for (int t = 0; t < n; t++) {
        g->inv[t] = 1; /* Tile() / resetDisplacements */
        g->bu[t] = u[t]; /* backupDisplacement */
        g->bv[t] = v[t];
    }
    int adj[3][3];
    /* initializer visitor :268-281 */
    for (int k = 0; k < n; k++) {
        int t = order[k];
        vt_neighbors(g, t / nJ, t % nJ, adj);
        vt_validate(g, &stages[0], t, adj);
        if (!g->inv[t]) g->locked[t] = 1;
    }
    int sweeps = 0, total = 0;
    for (;;) {
        int corrected = 0;
        for (int k = 0; k < n; k++) {
            int t = order[k];
            vt_neighbors(g, t / nJ, t % nJ, adj);
            vt_validate(g, &stages[0], t, adj);
            for (int index = 0; index < n_stages && g->inv[t]; index++) {
                if (stages[index].replacement != 0) {
                    g->u[t] = g->bu[t]; /* restoreDisplacement */
                    g->v[t] = g->bv[t];
                    if (stages[index].replacement == 1) vt_replace_secondary(g, t, stages[index].peak_index);
                    else vt_replace_bilinear(g, t, adj);
                }
                vt_validate(g, &stages[index], t, adj);
                if (!g->inv[t]) {
                    corrected++;
                    break;
                }
            }
        }
        sweeps++;
        total += corrected;
        /* shouldDoAnotherAttempt :84-103 */
        if (corrected == 0) break;
        if (!until_no_more && sweeps >= max_iterations) break;
        if (sweeps >= max_sweeps) {
            flags |= 1;
            break;
        }
    }
    int invalids = 0;
    for (int t = 0; t < n; t++) {
        if (g->inv[t]) {
            invalids++;
            u[t] = 0.0f;
            v[t] = 0.0f;
        }
        if (invalid) invalid[t] = g->inv[t];
    }
    if (g->err) flags |= 2;
    if (stats) {
        stats[0] = sweeps;
        stats[1] = total;
        stats[2] = invalids;
        stats[3] = flags;
    }
    free(order);
    free(g->bu);
    free(g->bv);
    free(g->inv);
    free(g->locked);
}

/* ------------------------------------------------------------------------ */
/* Velocity inheritance between adaptive levels                             */
/* ------------------------------------------------------------------------ */

/* Tile.accumulateDisplacement :103-136 on a reset tile (clipping allowed),
 * with the rounding of warping modes that need integer displacements. */
static void inherit_store(float du, float dv, int round_it, float *u, float *v)
{
    if (round_it) {
        du = (float)fm_round(du);
        dv = (float)fm_round(dv);
    }
    *u = 0.0f + du;
    *v = 0.0f + dv;
}

typedef struct {
    int parent;
    float overlapFactor, overlapArea;
    int moveUp, moveDown, moveLeft, moveRight;
} or_overlap;

/* AdaptiveInterVelocityInheritanceStrategyArea.checkOverlappingFactor :224-299 */
static void or_check_overlap(int pTop, int pLeft, int pH, int pW, int cTop, int cLeft, int cH,
                             int cW, or_overlap *r)
{
    int moveDown = 0, moveUp = 0, moveRight = 0, moveLeft = 0;
    int vert = 0, horiz = 0;
    if (pTop <= cTop) {
        if (pTop + pH > cTop) {
            int end = cTop + cH;
            if (end - 1 >= pTop + pH) end = pTop + pH;
            vert = end - cTop;
        } else {
            moveDown = 1;
        }
    } else {
        if (cTop + cH > pTop) vert = cTop + cH - pTop;
        else moveUp = 1;
    }
    if (pLeft <= cLeft) {
        if (pLeft + pW > cLeft) {
            int end = cLeft + cW;
            if (end - 1 >= pLeft + pW) end = pLeft + pW;
            horiz = end - cLeft;
        } else {
            moveRight = 1;
        }
    } else {
        if (cLeft + cW > pLeft) horiz = cLeft + cW - pLeft;
        else moveLeft = 1;
    }
    r->overlapArea = (float)(vert * horiz);
    r->overlapFactor = r->overlapArea / (float)(cW * cH);
    r->moveUp = moveUp;
    r->moveDown = moveDown;
    r->moveLeft = moveLeft;
    r->moveRight = moveRight;
}

/* AdaptiveInterVelocityInheritanceStrategyArea.reuseIterationStepTilesParameters
 * :47-214 for one frame.  Returns 0, or -1 where the reference throws
 * IllegalStateException. */
int or_inherit_area(int pnI, int pnJ, int pH, int pW, const int *ptop, const int *pleft,
                    const float *pu, const float *pv, int cnI, int cnJ, int cH, int cW,
                    const int *ctop, const int *cleft, int round_it, float *u, float *v)
{
    int parentIndexI = 0, parentIndexJ = 0;
    int cap = 64;
    or_overlap *areas = (or_overlap *)malloc(sizeof(or_overlap) * cap);
#define PT(I, J) ((I) * pnJ + (J))
    for (int indexI = 0; indexI < cnI; indexI++) {
        parentIndexJ = 0;
        int ref = PT(parentIndexI, parentIndexJ);
        for (int indexJ = 0; indexJ < cnJ; indexJ++) {
            int cur = indexI * cnJ + indexJ;
            or_overlap ra;
            or_check_overlap(ptop[ref], pleft[ref], pH, pW, ctop[cur], cleft[cur], cH, cW, &ra);
            if (ra.moveUp || ra.moveLeft) goto fail;
            if ((ra.moveRight && ra.moveLeft) || (ra.moveUp && ra.moveDown)) goto fail;
            if (ra.moveRight) {
                if (ra.moveDown) goto fail;
                if (parentIndexJ + 1 == pnJ) goto fail;
                parentIndexJ++;
                ref = PT(parentIndexI, parentIndexJ);
            }
            if (ra.moveDown) {
                if (parentIndexI + 1 == pnI) goto fail;
                parentIndexI++;
                ref = PT(parentIndexI, parentIndexJ);
            }
            float accumulatedArea = 0.0f;
            int offsetI = 0, offsetJ = 0, maxTilesInJ = 0, endCycles = 0, resultIndex = 0;
            do {
                int pt = PT(parentIndexI + offsetI, parentIndexJ + offsetJ);
                if (resultIndex >= cap) {
                    cap *= 2;
                    areas = (or_overlap *)realloc(areas, sizeof(or_overlap) * cap);
                }
                or_overlap *ca = &areas[resultIndex];
                or_check_overlap(ptop[pt], pleft[pt], pH, pW, ctop[cur], cleft[cur], cH, cW, ca);
                ca->parent = pt;
                if (ca->overlapFactor == 0) {
                    if ((ca->moveRight && ca->moveLeft) || (ca->moveUp && ca->moveDown)) goto fail;
                    if (ca->moveRight) goto fail;
                    if (ca->moveDown) goto fail;
                    if (ca->moveLeft) {
                        maxTilesInJ = offsetJ;
                        offsetJ = 0;
                        offsetI++;
                        if (parentIndexI + offsetI == pnI) endCycles = 1;
                    } else if (ca->moveUp) {
                        if (offsetJ != 0) goto fail;
                        endCycles = 1;
                    }
                } else {
                    accumulatedArea += ca->overlapArea;
                    resultIndex++;
                    offsetJ++;
                    int incrementI = 0;
                    if (parentIndexJ + offsetJ == pnJ) incrementI = 1;
                    if (maxTilesInJ > 0 && offsetJ == maxTilesInJ) incrementI = 1;
                    if (incrementI) {
                        offsetI++;
                        offsetJ = 0;
                        if (parentIndexI + offsetI == pnI) endCycles = 1;
                    }
                }
            } while (!endCycles);
            float du = 0.0f, dv = 0.0f;
            for (int k = 0; k < resultIndex; k++) {
                float weight = areas[k].overlapArea / accumulatedArea;
                du += weight * pu[areas[k].parent];
                dv += weight * pv[areas[k].parent];
            }
            inherit_store(du, dv, round_it, &u[cur], &v[cur]);
        }
    }
#undef PT
    free(areas);
    return 0;
fail:
    free(areas);
    return -1;
}

typedef struct {
    int parent; /* -1: null */
    float distance;
    int isUp, isDown, isLeft, isRight;
} or_dist;

/* AdaptiveInterVelocityInheritanceStrategyDistanceGeneric.checkDistance :220-265 */
static void or_check_distance(int pTop, int pLeft, int pH, int pW, int cTop, int cLeft, int cH,
                              int cW, or_dist *r)
{
    float centerY = (float)pTop + (float)pH / 2.0f;
    float centerX = (float)pLeft + (float)pW / 2.0f;
    float currentCenterY = (float)cTop + (float)cH / 2.0f;
    float currentCenterX = (float)cLeft + (float)cW / 2.0f;
    r->isUp = centerY <= currentCenterY;
    r->isDown = !r->isUp;
    r->isLeft = centerX <= currentCenterX;
    r->isRight = !r->isLeft;
    double dx = (double)(currentCenterX - centerX), dy = (double)(currentCenterY - centerY);
    /* FastMath.pow(d, 2) of a float operand: the exact square (representable in
     * double, and FastMath.pow errs by less than 1 ulp) */
    r->distance = (float)sqrt(dx * dx + dy * dy);
}

/* AdaptiveInterVelocityInheritanceStrategyDistanceGeneric.reuseIterationStepTilesParameters
 * :30-208 for one frame.  Returns 0, or -1 where the reference throws. */
int or_inherit_distance(int pnI, int pnJ, int pH, int pW, const int *ptop, const int *pleft,
                        const float *pu, const float *pv, int cnI, int cnJ, int cH, int cW,
                        const int *ctop, const int *cleft, int round_it, float *u, float *v)
{
    int parentIndexI = 0, parentIndexJ = 0;
    or_dist ca[4];
#define PT(I, J) ((I) * pnJ + (J))
    for (int indexI = 0; indexI < cnI; indexI++) {
        parentIndexJ = 0;
        for (int indexJ = 0; indexJ < cnJ; indexJ++) {
            int cur = indexI * cnJ + indexJ;
            int more = 1;
            while (more) {
                int adjusted = 0;
                int ref = PT(parentIndexI, parentIndexJ);
                or_check_distance(ptop[ref], pleft[ref], pH, pW, ctop[cur], cleft[cur], cH, cW, &ca[0]);
                ca[0].parent = ref;
                if (parentIndexI + 1 < pnI && parentIndexJ + 1 < pnJ) {
                    int dr = PT(parentIndexI + 1, parentIndexJ + 1);
                    or_check_distance(ptop[dr], pleft[dr], pH, pW, ctop[cur], cleft[cur], cH, cW, &ca[3]);
                    ca[3].parent = dr;
                    if (!ca[3].isRight) {
                        parentIndexJ++;
                        adjusted = 1;
                    }
                    if (!ca[3].isDown) {
                        parentIndexI++;
                        adjusted = 1;
                    }
                    if (ca[3].isDown && ca[3].isRight) more = 0;
                    else continue;
                } else {
                    ca[3].parent = -1;
                }
                if (parentIndexJ + 1 < pnJ) {
                    int rt = PT(parentIndexI, parentIndexJ + 1);
                    or_check_distance(ptop[rt], pleft[rt], pH, pW, ctop[cur], cleft[cur], cH, cW, &ca[1]);
                    ca[1].parent = rt;
                    if (!ca[1].isRight) {
                        if (!more) return -1;
                        parentIndexJ++;
                        adjusted = 1;
                        continue;
                    } else {
                        more = 0;
                    }
                } else {
                    ca[1].parent = -1;
                }
                if (parentIndexI + 1 < pnI) {
                    int dn = PT(parentIndexI + 1, parentIndexJ);
                    or_check_distance(ptop[dn], pleft[dn], pH, pW, ctop[cur], cleft[cur], cH, cW, &ca[2]);
                    ca[2].parent = dn;
                    if (!ca[2].isDown) {
                        if (!more) return -1;
                        adjusted = 1;
                        parentIndexI++;
                        continue;
                    } else {
                        more = 0;
                    }
                } else {
                    ca[2].parent = -1;
                }
                if (more && !adjusted) more = 0;
            }
            if (ca[0].isRight) {
                ca[1].parent = -1;
                ca[3].parent = -1;
            }
            if (ca[0].isDown) {
                ca[2].parent = -1;
                ca[3].parent = -1;
            }
            float total = 0.0f;
            int zero = -1;
            for (int k = 0; k < 4; k++) {
                if (ca[k].parent >= 0) {
                    if (ca[k].distance == 0.0f) {
                        zero = k;
                        break;
                    }
                    total += 1.0f / ca[k].distance;
                }
            }
            float du = 0.0f, dv = 0.0f;
            if (zero < 0) {
                for (int k = 0; k < 4; k++) {
                    if (ca[k].parent >= 0) {
                        float weight = (1.0f / ca[k].distance) / total;
                        du += weight * pu[ca[k].parent];
                        dv += weight * pv[ca[k].parent];
                    }
                }
            } else {
                du = pu[ca[zero].parent];
                dv = pv[ca[zero].parent];
            }
            inherit_store(du, dv, round_it, &u[cur], &v[cur]);
        }
    }
#undef PT
    return 0;
}

/* ---- commons-math3 3.6.1 AkimaSplineInterpolator + PolynomialSplineFunction,
 *      restated from the library's published algorithm ---------------------- */

/* Precision.equals(x, 0.0) (one ulp) for x >= 0 */
static int cm_equals_zero(double x)
{
    uint64_t b;
    memcpy(&b, &x, 8);
    return b <= 1;
}

static double cm_diff3(const double *xv, const double *yv, int d, int i0, int i1, int i2)
{
    double x0 = yv[i0], x1 = yv[i1], x2 = yv[i2];
    double t = xv[d] - xv[i0];
    double t1 = xv[i1] - xv[i0];
    double t2 = xv[i2] - xv[i0];
    double a = (x2 - x0 - (t2 / t1 * (x1 - x0))) / (t2 * t2 - t1 * t2);
    double b = (x1 - x0 - a * t1 * t1) / t1;
    return (2 * a * t) + b;
}

/* Java's Arrays.binarySearch over sorted doubles (no NaN, no signed zeros in use) */
static int cm_bsearch(const double *a, int n, double key)
{
    int lo = 0, hi = n - 1;
    while (lo <= hi) {
        int mid = (int)((unsigned)(lo + hi) >> 1);
        double mv = a[mid];
        if (mv < key) lo = mid + 1;
        else if (mv > key) hi = mid - 1;
        else return mid;
    }
    return -(lo + 1);
}

/* PolynomialFunction.value with the constructor's trailing-zero trim */
static double cm_poly(const double *c, int n, double x)
{
    while (n > 1 && c[n - 1] == 0) n--;
    double r = c[n - 1];
    for (int j = n - 2; j >= 0; j--) r = x * r + c[j];
    return r;
}

/* AkimaSplineInterpolator.interpolate(x, y) over 5 points, evaluated at v
 * (PolynomialSplineFunction.value). */
static double cm_akima5(const double *xv, const double *yv, double v)
{
    const int n = 5;
    double diff[4], w[4], fd[5];
    for (int i = 0; i < 4; i++) diff[i] = (yv[i + 1] - yv[i]) / (xv[i + 1] - xv[i]);
    w[0] = 0.0;
    for (int i = 1; i < 4; i++) w[i] = fabs(diff[i] - diff[i - 1]);
    for (int i = 2; i < n - 2; i++) {
        double wP = w[i + 1], wM = w[i - 1];
        if (cm_equals_zero(wP) && cm_equals_zero(wM)) {
            double xvv = xv[i], xvP = xv[i + 1], xvM = xv[i - 1];
            fd[i] = (((xvP - xvv) * diff[i - 1]) + ((xvv - xvM) * diff[i])) / (xvP - xvM);
        } else {
            fd[i] = ((wP * diff[i - 1]) + (wM * diff[i])) / (wP + wM);
        }
    }
    fd[0] = cm_diff3(xv, yv, 0, 0, 1, 2);
    fd[1] = cm_diff3(xv, yv, 1, 0, 1, 2);
    fd[n - 2] = cm_diff3(xv, yv, n - 2, n - 3, n - 2, n - 1);
    fd[n - 1] = cm_diff3(xv, yv, n - 1, n - 3, n - 2, n - 1);
    /* PolynomialSplineFunction.value: the piece holding v */
    int i = cm_bsearch(xv, n, v);
    if (i < 0) i = -i - 2;
    if (i >= n - 1) i--;
    double wd = xv[i + 1] - xv[i];
    double w2 = wd * wd;
    double yvv = yv[i], yvP = yv[i + 1], f0 = fd[i], fP = fd[i + 1];
    double c[4];
    c[0] = yvv;
    c[1] = fd[i];
    c[2] = (3 * (yvP - yvv) / wd - 2 * f0 - fP) / wd;
    c[3] = (2 * (yvv - yvP) / wd + f0 + fP) / w2;
    return cm_poly(c, 4, v - xv[i]);
}

/* PiecewiseBicubicSplineInterpolatingFunction.searchIndex (offset 2, count 5) */
static int cm_search5(double c, const double *val, int len)
{
    int r = cm_bsearch(val, len, c);
    if (r < 0) r = -r - 2 - 1;
    else r -= 2;
    if (r < 0) r = 0;
    if (r + 5 >= len) r = len - 5;
    return r;
}

/* PiecewiseBicubicSplineInterpolatingFunction.value(x, y): f[i][j] at (xval[i], yval[j]) */
double or_piecewise_bicubic(const double *xval, int xLen, const double *yval, int yLen,
                            const double *f, double x, double y)
{
    int i = cm_search5(x, xval, xLen);
    int j = cm_search5(y, yval, yLen);
    double xa[5], ya[5], za[5], interp[5];
    for (int k = 0; k < 5; k++) {
        xa[k] = xval[i + k];
        ya[k] = yval[j + k];
    }
    for (int zi = 0; zi < 5; zi++) {
        for (int k = 0; k < 5; k++) za[k] = f[(size_t)(i + k) * yLen + (j + zi)];
        interp[zi] = cm_akima5(xa, za, x);
    }
    return cm_akima5(ya, interp, y);
}

/* AdaptiveInterVelocityInheritanceStrategyBiCubicSpline.reuseIterationStepTilesParameters
 * :29-118 for one frame: parent displacements on the grid of parent tile
 * centres, PiecewiseBicubicSplineInterpolator (5 x 5 Akima windows), child
 * centres clamped to the parent range.  With fewer than 5 parent tiles in a
 * direction the reference falls back to a bilinear function whose values it
 * then never stores: the children keep zero displacement. */
void or_inherit_bicubic(int pnI, int pnJ, int pH, int pW, const int *ptop, const int *pleft,
                        const float *pu, const float *pv, int cnI, int cnJ, int cH, int cW,
                        const int *ctop, const int *cleft, int round_it, float *u, float *v)
{
    int n = cnI * cnJ;
    if (pnI < 5 || pnJ < 5) {
        for (int t = 0; t < n; t++) u[t] = v[t] = 0.0f;
        return;
    }
    double *us = (double *)malloc(sizeof(double) * pnI * pnJ);
    double *vs = (double *)malloc(sizeof(double) * pnI * pnJ);
    double *xs = (double *)malloc(sizeof(double) * pnJ);
    double *ys = (double *)malloc(sizeof(double) * pnI);
    for (int i = 0; i < pnI; i++)
        for (int j = 0; j < pnJ; j++) {
            us[i * pnJ + j] = pu[i * pnJ + j];
            vs[i * pnJ + j] = pv[i * pnJ + j];
            if (i == 0) xs[j] = (double)((float)pleft[j] + (float)pW / 2.0f - 0.5f);
            if (j == 0) ys[i] = (double)((float)ptop[i * pnJ] + (float)pH / 2.0f - 0.5f);
        }
    float Xmin = (float)xs[0], Xmax = (float)xs[pnJ - 1];
    float Ymin = (float)ys[0], Ymax = (float)ys[pnI - 1];
    for (int t = 0; t < n; t++) {
        float x = (float)cleft[t] + (float)cW / 2.0f - 0.5f;
        float y = (float)ctop[t] + (float)cH / 2.0f - 0.5f;
        if (x < Xmin) x = Xmin;
        else if (x > Xmax) x = Xmax;
        if (y < Ymin) y = Ymin;
        else if (y > Ymax) y = Ymax;
        float iu = (float)or_piecewise_bicubic(ys, pnI, xs, pnJ, us, (double)y, (double)x);
        float iv = (float)or_piecewise_bicubic(ys, pnI, xs, pnJ, vs, (double)y, (double)x);
        inherit_store(iu, iv, round_it, &u[t], &v[t]);
    }
    free(us);
    free(vs);
    free(xs);
    free(ys);
}

/* ------------------------------------------------------------------------ */
/* Gaussian pre-filter: images/filters/GaussianFilter2D.java,               */
/* GaussianFilter2DInternal.java, DirectConvolution.java                    */
/* ------------------------------------------------------------------------ */

/* GaussianFilter2D.createKernel :36-50 (FastMath.sqrt/exp/PI in double) */
void or_gauss_kernel(float sigma, int kernelPx, float *k)
{
    double kd[64];
    int half = kernelPx / 2;
    double sum = 0.0;
    for (int i = 0; i < kernelPx; i++) {
        double x = i - half;
        kd[i] = 1.0 / sqrt(2.0 * M_PI * sigma * sigma) * exp(-x * x / (2.0 * sigma * sigma));
        sum += kd[i];
    }
    for (int i = 0; i < kernelPx; i++) k[i] = (float)(kd[i] / sum);
}

/* DirectConvolution.convolve1D :15-49 */
void or_convolve1d(const float *a, int sizeA, const float *b, int sizeB, float *result)
{
    int minSize = sizeA, otherSize = sizeB;
    const float *minVec = a, *otherVec = b;
    if (sizeB < minSize) {
        minSize = sizeB;
        otherSize = sizeA;
        minVec = b;
        otherVec = a;
    }
    for (int i = 0; i < otherSize; i++) result[i] = 0.0f;
    int halfSize = minSize / 2;
    for (int i = 0; i < sizeA + sizeB - 1; i++) {
        for (int j = 0; j < minSize; j++) {
            if (i - j >= 0 && i - j < otherSize) {
                if (halfSize * 2 == minSize) {
                    if (i >= halfSize - 1 && i < otherSize + halfSize - 1)
                        result[i - (halfSize - 1)] += otherVec[i - j] * minVec[j];
                } else {
                    if (i >= halfSize && i < otherSize + halfSize)
                        result[i - halfSize] += otherVec[i - j] * minVec[j];
                }
            }
        }
    }
}

/* GaussianFilter2DInternal.applyFilter(float[][], float[][]) :120-171: rows,
 * then columns of the row result; borders: mirrored on the leading side, the
 * last half-kernel samples repeated in order on the trailing side. */
void or_gaussian_filter2d(const float *in, int H, int W, const float *kernel, int K, float *out)
{
    int half = K / 2;
    int nmax = (H > W ? H : W) + K - 1;
    float *buf = (float *)malloc(sizeof(float) * nmax);
    float *res = (float *)malloc(sizeof(float) * nmax);
    int bs = W + K - 1;
    for (int i = 0; i < H; i++) {
        for (int j = 0; j < W; j++) buf[j + half] = in[(size_t)i * W + j];
        for (int j = 0; j < half; j++) {
            buf[half - 1 - j] = in[(size_t)i * W + j];
            buf[bs - 1 - j] = in[(size_t)i * W + (W - 1 - j)];
        }
        or_convolve1d(buf, bs, kernel, K, res);
        for (int j = 0; j < W; j++) out[(size_t)i * W + j] = res[j + half];
    }
    bs = H + K - 1;
    for (int j = 0; j < W; j++) {
        for (int i = 0; i < H; i++) buf[i + half] = out[(size_t)i * W + j];
        for (int i = 0; i < half; i++) {
            buf[half - 1 - i] = out[(size_t)i * W + j];
            buf[bs - 1 - i] = out[(size_t)(H - 1 - i) * W + j];
        }
        or_convolve1d(buf, bs, kernel, K, res);
        for (int i = 0; i < H; i++) out[(size_t)i * W + j] = res[i + half];
    }
    free(buf);
    free(res);
}

/* ------------------------------------------------------------------------ */
/* Spline displacement fields of the Mini and Full-image warps:              */
/* iareas/BiCubicSplineInterpolatorWithBiLinearBackup.java over commons-math3 */
/* 3.6.1 BicubicSplineInterpolator (absent here; restated from its published  */
/* algorithm: natural cubic splines give the knot slopes, the cross slope is  */
/* the divided difference over the neighbouring knots, one bicubic Hermite    */
/* patch per cell), and iareas/BiLinearInterpolatingFunction.java.  FP64.     */
/* ------------------------------------------------------------------------ */

#define OR_MAX_KNOTS 1024

/* Natural cubic spline through (x[k], y[k]), k < n, n >= 3: its slope at every
 * knot.  SplineInterpolator.interpolate (Burden & Faires, Numerical Analysis,
 * Algorithm 3.4: natural end conditions, forward elimination of the tridiagonal
 * system, back substitution), then PolynomialSplineFunction.derivative() at the
 * knots: each knot's own piece at 0, the last knot at the end of the last piece. */
void or_natural_slopes(int n, const double *x, const double *y, double *slope)
{
    const int m = n - 1; /* intervals (n <= OR_MAX_KNOTS) */
    double h[OR_MAX_KNOTS], mu[OR_MAX_KNOTS], z[OR_MAX_KNOTS + 1], b[OR_MAX_KNOTS],
        c[OR_MAX_KNOTS + 1], d[OR_MAX_KNOTS];
    for (int i = 0; i < m; i++) h[i] = x[i + 1] - x[i];
    mu[0] = 0.0;
    z[0] = 0.0;
    for (int i = 1; i < m; i++) {
        const double g = 2.0 * (x[i + 1] - x[i - 1]) - h[i - 1] * mu[i - 1];
        mu[i] = h[i] / g;
        z[i] = (3.0 * (y[i + 1] * h[i - 1] - y[i] * (x[i + 1] - x[i - 1]) + y[i - 1] * h[i]) /
                    (h[i - 1] * h[i]) -
                h[i - 1] * z[i - 1]) /
               g;
    }
    z[m] = 0.0;
    c[m] = 0.0;
    for (int j = m - 1; j >= 0; j--) {
        c[j] = z[j] - mu[j] * c[j + 1];
        b[j] = (y[j + 1] - y[j]) / h[j] - h[j] * (c[j + 1] + 2.0 * c[j]) / 3.0;
        d[j] = (c[j + 1] - c[j]) / (3.0 * h[j]);
    }
    for (int j = 0; j < m; j++) slope[j] = b[j];
    /* derivative polynomial b + 2c t + 3d t^2 of the last piece at t = h (Horner) */
    slope[m] = (3.0 * d[m - 1] * h[m - 1] + 2.0 * c[m - 1]) * h[m - 1] + b[m - 1];
}

/* BicubicSplineInterpolator.interpolate(xs, ys, f): f[i * ny + j] at (xs[i], ys[j]);
 * slopes along x (splines over xs of every column j), along y (splines over ys
 * of every row i), and the cross slope
 *   (f[nI][nJ] - f[nI][pJ] - f[pI][nJ] + f[pI][pJ]) / ((xs[nI]-xs[pI]) (ys[nJ]-ys[pJ]))
 * with nI = i + 1 (i at the last knot), pI = i - 1 (i at the first). */
void or_spline_surface(int nx, const double *xs, int ny, const double *ys, const double *f,
                       double *fx, double *fy, double *fxy)
{
    double col[OR_MAX_KNOTS], s[OR_MAX_KNOTS];
    for (int j = 0; j < ny; j++) {
        for (int i = 0; i < nx; i++) col[i] = f[i * ny + j];
        or_natural_slopes(nx, xs, col, s);
        for (int i = 0; i < nx; i++) fx[i * ny + j] = s[i];
    }
    for (int i = 0; i < nx; i++) {
        or_natural_slopes(ny, ys, f + i * ny, s);
        for (int j = 0; j < ny; j++) fy[i * ny + j] = s[j];
    }
    for (int i = 0; i < nx; i++) {
        const int nI = i + 1 < nx ? i + 1 : i, pI = i - 1 >= 0 ? i - 1 : 0;
        for (int j = 0; j < ny; j++) {
            const int nJ = j + 1 < ny ? j + 1 : j, pJ = j - 1 >= 0 ? j - 1 : 0;
            fxy[i * ny + j] = (f[nI * ny + nJ] - f[nI * ny + pJ] - f[pI * ny + nJ] + f[pI * ny + pJ]) /
                              ((xs[nI] - xs[pI]) * (ys[nJ] - ys[pJ]));
        }
    }
}

/* commons BicubicSplineInterpolatingFunction.searchIndex: the cell of c */
static int spline_cell(double c, const double *v, int n)
{
    if (c < v[0]) return -1;
    for (int i = 1; i < n; i++)
        if (c <= v[i]) return i - 1;
    return -1;
}

/* BicubicSplineInterpolatingFunction.value(x, y): the cell's bicubic patch, in
 * Hermite form on the normalised cell coordinates s, t (the patch matching the
 * corner values, slopes times the cell size and cross slopes times both sizes). */
double or_spline_surface_value(int nx, const double *xs, int ny, const double *ys, const double *f,
                               const double *fx, const double *fy, const double *fxy, double x,
                               double y)
{
    const int i = spline_cell(x, xs, nx), j = spline_cell(y, ys, ny);
    if (i < 0 || j < 0) return NAN;
    const double dx = xs[i + 1] - xs[i], dy = ys[j + 1] - ys[j];
    const double s = (x - xs[i]) / dx, t = (y - ys[j]) / dy;
    const double hs[2] = {(1.0 + 2.0 * s) * (1.0 - s) * (1.0 - s), s * s * (3.0 - 2.0 * s)};
    const double gs[2] = {s * (1.0 - s) * (1.0 - s), s * s * (s - 1.0)};
    const double ht[2] = {(1.0 + 2.0 * t) * (1.0 - t) * (1.0 - t), t * t * (3.0 - 2.0 * t)};
    const double gt[2] = {t * (1.0 - t) * (1.0 - t), t * t * (t - 1.0)};
    double r = 0.0;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            const int k = (i + a) * ny + (j + b);
            r += f[k] * hs[a] * ht[b] + fx[k] * dx * gs[a] * ht[b] + fy[k] * dy * hs[a] * gt[b] +
                 fxy[k] * dx * dy * gs[a] * gt[b];
        }
    return r;
}

/* BiLinearInterpolatingFunction.value(y, x) :40-70 (ys rows, xs columns) */
double or_bilinear_value(int ny, const double *ys, int nx, const double *xs, const double *v,
                         double y, double x)
{
    int I, J;
    for (I = 1; I < ny; I++)
        if (y <= ys[I]) break;
    for (J = 1; J < nx; J++)
        if (x <= xs[J]) break;
    if (I >= ny) I = ny - 1;
    if (J >= nx) J = nx - 1;
    const double wY = 1.0 - ((ys[I] - y) / (ys[I] - ys[I - 1]));
    const double wX = 1.0 - ((xs[J] - x) / (xs[J] - xs[J - 1]));
    return wY * (wX * v[I * nx + J] + (1.0 - wX) * v[I * nx + J - 1]) +
           (1.0 - wY) * (wX * v[(I - 1) * nx + J] + (1.0 - wX) * v[(I - 1) * nx + J - 1]);
}

/* A displacement field over knots ys (rows) x xs (columns): the bicubic spline
 * when both directions hold at least 3 knots (the natural spline's minimum),
 * otherwise the bilinear backup; queries are clamped to the knot range in
 * float (interpolateDisplacements :55-104). */
typedef struct {
    int ny, nx, spline;
    double *ys, *xs;
    double *f[2], *fx[2], *fy[2], *fxy[2];
    float ymin, ymax, xmin, xmax;
} or_field;

static void field_alloc(or_field *F, int ny, int nx)
{
    F->ny = ny;
    F->nx = nx;
    F->ys = (double *)malloc(sizeof(double) * ny);
    F->xs = (double *)malloc(sizeof(double) * nx);
    for (int c = 0; c < 2; c++) {
        F->f[c] = (double *)malloc(sizeof(double) * ny * nx);
        F->fx[c] = (double *)malloc(sizeof(double) * ny * nx);
        F->fy[c] = (double *)malloc(sizeof(double) * ny * nx);
        F->fxy[c] = (double *)malloc(sizeof(double) * ny * nx);
    }
}

static void field_free(or_field *F)
{
    free(F->ys);
    free(F->xs);
    for (int c = 0; c < 2; c++) {
        free(F->f[c]);
        free(F->fx[c]);
        free(F->fy[c]);
        free(F->fxy[c]);
    }
}

static void field_build(or_field *F)
{
    F->spline = F->ny >= 3 && F->nx >= 3;
    if (F->spline)
        for (int c = 0; c < 2; c++)
            or_spline_surface(F->ny, F->ys, F->nx, F->xs, F->f[c], F->fx[c], F->fy[c], F->fxy[c]);
    F->ymin = (float)F->ys[0];
    F->ymax = (float)F->ys[F->ny - 1];
    F->xmin = (float)F->xs[0];
    F->xmax = (float)F->xs[F->nx - 1];
}

static float field_value(const or_field *F, int c, float y, float x)
{
    if (x < F->xmin) x = F->xmin;
    else if (x > F->xmax) x = F->xmax;
    if (y < F->ymin) y = F->ymin;
    else if (y > F->ymax) y = F->ymax;
    if (F->spline)
        return (float)or_spline_surface_value(F->ny, F->ys, F->nx, F->xs, F->f[c], F->fx[c],
                                              F->fy[c], F->fxy[c], (double)y, (double)x);
    return (float)or_bilinear_value(F->ny, F->ys, F->nx, F->xs, F->f[c], (double)y, (double)x);
}

/* Mini warp of tile (ti, tj) for a side (warpAndClipImage :39-186,
 * createMiniDisplacements :203-290, doBiLinearWarp :292-334): the field of the
 * tile's 7 x 7 neighbourhood (cut to the grid), whose knots are the tile
 * centres and whose values are the neighbours' displacements minus floor of
 * this tile's, is evaluated from the B reference window (top + floor(u),
 * left + floor(v)) in every order, and applied at the side's window.  Grid of
 * nI x nJ tiles of N px, row-major tops/lefts. */
/* The Mini warp of tile (ti, tj) for a side, from a frame of any pixel type,
 * into float samples. */
static void warp_mini_frame(const or_frame *fr, int H, int W, int nI, int nJ, const int *top,
                            const int *left, int N, const float *u, const float *v, int ti, int tj,
                            int side, float *tile)
{
    or_field F;
    const int t = ti * nJ + tj;
    const double fu0 = floor((double)u[t]), fv0 = floor((double)v[t]);
    const short refTop = (short)(top[t] + floor(u[t]));
    const short refLeft = (short)(left[t] + floor(v[t]));
    const or_side g = side_geometry(side, top[t], left[t], u[t], v[t]);
    const int i0 = ti - 3 < 0 ? 0 : ti - 3, i1 = ti + 3 > nI - 1 ? nI - 1 : ti + 3;
    const int j0 = tj - 3 < 0 ? 0 : tj - 3, j1 = tj + 3 > nJ - 1 ? nJ - 1 : tj + 3;
    field_alloc(&F, i1 - i0 + 1, j1 - j0 + 1);
    for (int i = i0; i <= i1; i++) F.ys[i - i0] = (double)(top[i * nJ + j0] + N / 2.0f - 0.5f);
    for (int j = j0; j <= j1; j++) F.xs[j - j0] = (double)(left[i0 * nJ + j] + N / 2.0f - 0.5f);
    for (int i = i0; i <= i1; i++)
        for (int j = j0; j <= j1; j++) {
            F.f[0][(i - i0) * F.nx + (j - j0)] = (double)u[i * nJ + j] - fu0;
            F.f[1][(i - i0) * F.nx + (j - j0)] = (double)v[i * nJ + j] - fv0;
        }
    field_build(&F);
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            const float y = (float)(refTop + i), x = (float)(refLeft + j);
            float du = field_value(&F, 0, y, x), dv = field_value(&F, 1, y, x);
            if (g.half) {
                du = du / 2.0f + g.offU;
                dv = dv / 2.0f + g.offV;
            }
            tile[i * N + j] = warp_pixel_frame(fr, H, W, du, dv, (short)(i + g.top), (short)(j + g.left),
                                               g.first);
        }
    field_free(&F);
}

void or_warp_mini_side(const uint8_t *img, int H, int W, int nI, int nJ, const int *top,
                       const int *left, int N, const float *u, const float *v, int ti, int tj,
                       int side, uint8_t *tile)
{
    const or_frame fr = {img, NULL, NULL};
    float *t = (float *)malloc(sizeof(float) * N * N);
    warp_mini_frame(&fr, H, W, nI, nJ, top, left, N, u, v, ti, tj, side, t);
    for (int k = 0; k < N * N; k++) tile[k] = (uint8_t)t[k];
    free(t);
}

/* The Mini warp of a float frame (ImageFloat) or a 10/12/16-bit frame
 * (ImageInt16): the rounded samples as floats (doBiLinearWarp :292-334 writes
 * them into a MatrixFloat whose maximum is image.getMaximumValue(), :298). */
void or_warp_mini_side_f32(const float *img, int H, int W, int nI, int nJ, const int *top,
                           const int *left, int N, const float *u, const float *v, int ti, int tj,
                           int side, float *tile)
{
    const or_frame fr = {NULL, img, NULL};
    warp_mini_frame(&fr, H, W, nI, nJ, top, left, N, u, v, ti, tj, side, tile);
}

void or_warp_mini_side_u16(const uint16_t *img, int H, int W, int nI, int nJ, const int *top,
                           const int *left, int N, const float *u, const float *v, int ti, int tj,
                           int side, float *tile)
{
    const or_frame fr = {NULL, NULL, img};
    warp_mini_frame(&fr, H, W, nI, nJ, top, left, N, u, v, ti, tj, side, tile);
}

void or_warp_mini(const uint8_t *img, int H, int W, int nI, int nJ, const int *top,
                  const int *left, int N, const float *u, const float *v, int ti, int tj,
                  uint8_t *tile)
{
    or_warp_mini_side(img, H, W, nI, nJ, top, left, N, u, v, ti, tj, OR_SIDE_B_SECOND, tile);
}

/* Full-image warp field (BiLinearImageWarpingStrategy.warpAndClipImage :46-57,
 * createTileDisplacementInterpolator): the field over the whole grid's tile
 * centres, evaluated at every pixel (i rows, j columns) of an H x W image. */
void or_full_field(int nI, int nJ, const int *top, const int *left, int tH, int tW,
                   const float *u, const float *v, int H, int W, float *fu, float *fv)
{
    or_field F;
    field_alloc(&F, nI, nJ);
    for (int i = 0; i < nI; i++) F.ys[i] = (double)(top[i * nJ] + tH / 2.0f - 0.5f);
    for (int j = 0; j < nJ; j++) F.xs[j] = (double)(left[j] + tW / 2.0f - 0.5f);
    for (int k = 0; k < nI * nJ; k++) {
        F.f[0][k] = (double)u[k];
        F.f[1][k] = (double)v[k];
    }
    field_build(&F);
    for (int i = 0; i < H; i++)
        for (int j = 0; j < W; j++) {
            fu[(size_t)i * W + j] = field_value(&F, 0, (float)i, (float)j);
            fv[(size_t)i * W + j] = field_value(&F, 1, (float)i, (float)j);
        }
    field_free(&F);
}

/* BiLinearImageWarpingStrategy.doBiLinearWarp :143-183: every pixel of the
 * image sampled at its own field displacement; side OR_SIDE_B_SECOND (target
 * + field), OR_SIDE_A_FIRST (target - field), OR_SIDE_A_BOTH / _B_BOTH (half
 * the field, minus / plus). */
void or_warp_image_side(const uint8_t *img, int H, int W, const float *fu, const float *fv,
                        int side, uint8_t *out)
{
    const int half = side == OR_SIDE_A_BOTH || side == OR_SIDE_B_BOTH;
    const int first = side == OR_SIDE_A_FIRST || side == OR_SIDE_A_BOTH;
    for (int i = 0; i < H; i++)
        for (int j = 0; j < W; j++) {
            float u = fu[(size_t)i * W + j], v = fv[(size_t)i * W + j];
            if (half) {
                u /= 2.0f;
                v /= 2.0f;
            }
            out[(size_t)i * W + j] = (uint8_t)warp_pixel_order(img, H, W, u, v, (short)i, (short)j, first);
        }
}

/* The same for a float frame (ImageFloat): the rounded samples as floats. */
void or_warp_image_side_f32(const float *img, int H, int W, const float *fu, const float *fv,
                            int side, float *out)
{
    const or_frame fr = {NULL, img, NULL};
    const int half = side == OR_SIDE_A_BOTH || side == OR_SIDE_B_BOTH;
    const int first = side == OR_SIDE_A_FIRST || side == OR_SIDE_A_BOTH;
    for (int i = 0; i < H; i++)
        for (int j = 0; j < W; j++) {
            float u = fu[(size_t)i * W + j], v = fv[(size_t)i * W + j];
            if (half) {
                u /= 2.0f;
                v /= 2.0f;
            }
            out[(size_t)i * W + j] = warp_pixel_frame(&fr, H, W, u, v, (short)i, (short)j, first);
        }
}

void or_warp_image(const uint8_t *img, int H, int W, const float *fu, const float *fv,
                   uint8_t *out)
{
    or_warp_image_side(img, H, W, fu, fv, OR_SIDE_B_SECOND, out);
}
