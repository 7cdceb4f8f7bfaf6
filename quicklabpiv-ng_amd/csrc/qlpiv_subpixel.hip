// Sub-pixel peak refinement, one thread per (IA, peak).
//
// Input peak records are (i, j, value) triples as produced by the peak search;
// outputs are (i', j', value', status) with status bit0 = row fit accepted,
// bit1 = column fit accepted, bit2 = skipped (peak at (0,0) or too close to
// the border: integer peak kept), bit3 = floor below the fitter's limit (the
// reference throws InterpolateException there).
//
// GAUSS1D (interpolators/Gaussian1DInterpolator.java:36-114): a Gaussian
//   through 'pixels' samples along each axis.  For 3 samples the
//   Levenberg-Marquardt fitter of the reference converges to the exact
//   interpolant; it is evaluated here in float from sample ratios,
//   ln(y-/y+) / (2 (ln(y-/y0) + ln(y+/y0))), which keeps full relative
//   precision.  5 and 7 samples use a damped Gauss-Newton least-squares fit.
// GUO (Gaussian1DHongweiGuoInterpolator.java:104-248): iteratively reweighted
//   least squares of ln y = a + b x + c x^2 with the 3x3 normal equations
//   solved in FP64 by LU with partial pivoting (the commons-math3 3.6.1
//   LUDecomposition algorithm); the peak value becomes the fitted amplitude.
// POLY (Gaussian1DPolynomialInterpolator.java:25-106): closed-form 3-point
//   log-parabola with the parabolic fallback for negative samples.
#include <float.h>

#include "qlpiv_internal.h"

namespace qlpiv {

enum { ST_I = 1, ST_J = 2, ST_SKIP = 4, ST_FLOOR = 8 };

__device__ __forceinline__ float gauss3_offset(float ym, float y0, float yp)
{
    const float lmp = logf(ym / yp);
    const float den = 2.0f * (logf(ym / y0) + logf(yp / y0));
    return lmp / den;
}

// Gauss-Newton / Levenberg-Marquardt fit of y = a exp(-(x-b)^2 / (2 c^2)) to n
// samples (n = 5 or 7), x relative to the centre sample; returns b.
__device__ double gauss_lsq_mean(int n, const double *y)
{
    const int c = n / 2;
    double b = 0.0;
    {
        const double l = log(y[c - 1]) - 2.0 * log(y[c]) + log(y[c + 1]);
        const double bb = (log(y[c - 1]) - log(y[c + 1])) / (2.0 * l);
        if (isfinite(bb) && fabs(bb) <= 1.0) b = bb;
    }
    double s;
    {
        const double l = log(y[c - 1]) - 2.0 * log(y[c]) + log(y[c + 1]);
        s = (l < 0.0) ? sqrt(-1.0 / l) : 1.0;
        if (!isfinite(s) || s <= 0.0) s = 1.0;
    }
    double a = y[c] / exp(-(b * b) / (2.0 * s * s));
    double p0 = a, p1 = b, p2 = s;
    double lambda = 1e-3;
    double cost = 0.0;
    for (int k = 0; k < n; k++) {
        const double x = (double)(k - c);
        const double r = y[k] - p0 * exp(-(x - p1) * (x - p1) / (2.0 * p2 * p2));
        cost += r * r;
    }
    for (int it = 0; it < 100; it++) {
        double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, g[3] = {0, 0, 0};
        for (int k = 0; k < n; k++) {
            const double x = (double)(k - c);
            const double d = x - p1;
            const double e = exp(-d * d / (2.0 * p2 * p2));
            const double f = p0 * e;
            const double J[3] = {e, f * d / (p2 * p2), f * d * d / (p2 * p2 * p2)};
            const double r = y[k] - f;
            for (int u = 0; u < 3; u++) {
                g[u] += J[u] * r;
                for (int v = 0; v < 3; v++) A[u][v] += J[u] * J[v];
            }
        }
        bool improved = false;
        for (int tries = 0; tries < 30 && !improved; tries++) {
            double Am[3][4];
            for (int u = 0; u < 3; u++) {
                for (int v = 0; v < 3; v++) Am[u][v] = A[u][v] + (u == v ? lambda * A[u][u] : 0.0);
                Am[u][3] = g[u];
            }
            for (int col = 0; col < 3; col++) {
                int piv = col;
                for (int r2 = col + 1; r2 < 3; r2++)
                    if (fabs(Am[r2][col]) > fabs(Am[piv][col])) piv = r2;
                for (int q = 0; q < 4; q++) {
                    const double t = Am[col][q];
                    Am[col][q] = Am[piv][q];
                    Am[piv][q] = t;
                }
                for (int r2 = col + 1; r2 < 3; r2++) {
                    const double fct = Am[r2][col] / Am[col][col];
                    for (int q = col; q < 4; q++) Am[r2][q] -= fct * Am[col][q];
                }
            }
            double dp[3];
            for (int r2 = 2; r2 >= 0; r2--) {
                double sum = Am[r2][3];
                for (int q = r2 + 1; q < 3; q++) sum -= Am[r2][q] * dp[q];
                dp[r2] = sum / Am[r2][r2];
            }
            const double n0 = p0 + dp[0], n1 = p1 + dp[1], n2 = p2 + dp[2];
            double nc = 0.0;
            for (int k = 0; k < n; k++) {
                const double x = (double)(k - c);
                const double r = y[k] - n0 * exp(-(x - n1) * (x - n1) / (2.0 * n2 * n2));
                nc += r * r;
            }
            if (isfinite(nc) && nc <= cost) {
                const double rel = (cost - nc) / (cost > 0.0 ? cost : 1.0);
                p0 = n0;
                p1 = n1;
                p2 = n2;
                cost = nc;
                lambda *= 0.3;
                improved = true;
                if (rel < 1e-15 && fabs(dp[1]) < 1e-13) it = 1000;
            } else {
                lambda *= 10.0;
            }
        }
        if (!improved) break;
    }
    return p1;
}

// commons-math3 LUDecomposition + solver, 3x3 (singularity threshold 1e-11).
__device__ bool lu_solve3(const double A[3][3], const double b[3], double x[3])
{
    double lu[3][3];
    int piv[3] = {0, 1, 2};
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) lu[r][c] = A[r][c];
    for (int col = 0; col < 3; col++) {
        for (int row = 0; row < col; row++) {
            double sum = lu[row][col];
            for (int i = 0; i < row; i++) sum -= lu[row][i] * lu[i][col];
            lu[row][col] = sum;
        }
        int mx = col;
        double largest = -DBL_MAX;
        for (int row = col; row < 3; row++) {
            double sum = lu[row][col];
            for (int i = 0; i < col; i++) sum -= lu[row][i] * lu[i][col];
            lu[row][col] = sum;
            if (fabs(sum) > largest) {
                largest = fabs(sum);
                mx = row;
            }
        }
        if (fabs(lu[mx][col]) < 1e-11) return false;
        if (mx != col) {
            for (int i = 0; i < 3; i++) {
                const double t = lu[mx][i];
                lu[mx][i] = lu[col][i];
                lu[col][i] = t;
            }
            const int t = piv[mx];
            piv[mx] = piv[col];
            piv[col] = t;
        }
        const double d = lu[col][col];
        for (int row = col + 1; row < 3; row++) lu[row][col] /= d;
    }
    double bp[3];
    for (int r = 0; r < 3; r++) bp[r] = b[piv[r]];
    for (int col = 0; col < 3; col++) {
        const double bc = bp[col];
        for (int i = col + 1; i < 3; i++) bp[i] -= bc * lu[i][col];
    }
    for (int col = 2; col >= 0; col--) {
        bp[col] /= lu[col][col];
        const double bc = bp[col];
        for (int i = 0; i < col; i++) bp[i] -= bc * lu[i][col];
    }
    for (int r = 0; r < 3; r++) x[r] = bp[r];
    return true;
}

// Minimum-norm least-squares solve of the symmetric normal matrix by a cyclic
// Jacobi eigen-decomposition (the reference's SVD fallback for a singular LU).
__device__ void pinv_solve3(const double Ain[3][3], const double b[3], double x[3])
{
    double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) a[r][c] = Ain[r][c];
    for (int sweep = 0; sweep < 50; sweep++) {
        const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
        if (off < 1e-300) break;
        for (int p = 0; p < 2; p++)
            for (int q = p + 1; q < 3; q++) {
                if (a[p][q] == 0.0) continue;
                const double th = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; k++) {
                    const double akp = a[k][p], akq = a[k][q];
                    a[k][p] = c * akp - s * akq;
                    a[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; k++) {
                    const double apk = a[p][k], aqk = a[q][k];
                    a[p][k] = c * apk - s * aqk;
                    a[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; k++) {
                    const double vkp = v[k][p], vkq = v[k][q];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
            }
    }
    double smax = 0.0;
    for (int k = 0; k < 3; k++) smax = fmax(smax, fabs(a[k][k]));
    const double tol = 3.0 * smax * 2.220446049250313e-16;
    for (int r = 0; r < 3; r++) x[r] = 0.0;
    for (int k = 0; k < 3; k++) {
        const double s = a[k][k];
        if (fabs(s) <= tol) continue;
        double proj = 0.0;
        for (int r = 0; r < 3; r++) proj += v[r][k] * b[r];
        for (int r = 0; r < 3; r++) x[r] += v[r][k] * proj / s;
    }
}

__device__ void guo_accumulate(double A[3][3], double b[3], float xs, double ySquare, float ys)
{
    const double xSquare = (double)(xs * xs);
    const double xCubic = xSquare * xs;
    const double xFourth = xCubic * xs;
    const double logy = log((double)ys);
    A[0][0] += ySquare;
    A[0][1] += xs * ySquare;
    A[0][2] += xSquare * ySquare;
    b[0] += ySquare * logy;
    A[1][0] += xs * ySquare;
    A[1][1] += xSquare * ySquare;
    A[1][2] += xCubic * ySquare;
    b[1] += xs * ySquare * logy;
    A[2][0] += xSquare * ySquare;
    A[2][1] += xCubic * ySquare;
    A[2][2] += xFourth * ySquare;
    b[2] += xSquare * ySquare * logy;
}

// estimateGaussianParameters :104-130
__device__ void guo_estimate(int iterations, int n, const float *xs, const float *ys, float &mean,
                             float &amp)
{
    double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, b[3] = {0, 0, 0}, coefs[3] = {0, 0, 0};
    for (int i = 0; i < n; i++) guo_accumulate(A, b, xs[i], (double)(ys[i] * ys[i]), ys[i]);
    for (int it = 0; it < iterations; it++) {
        if (!lu_solve3(A, b, coefs)) pinv_solve3(A, b, coefs);
        for (int r = 0; r < 3; r++) {
            b[r] = 0.0;
            for (int c = 0; c < 3; c++) A[r][c] = 0.0;
        }
        for (int i = 0; i < n; i++) {
            const double yk = exp(coefs[0] + coefs[1] * xs[i] + coefs[2] * xs[i] * xs[i]);
            guo_accumulate(A, b, xs[i], yk * yk, ys[i]);
        }
    }
    mean = (float)(-coefs[1] / (2.0 * coefs[2]));
    amp = (float)(exp(coefs[0] - coefs[1] * coefs[1] / (4.0 * coefs[2])));
}

constexpr int NBR_MAX_HALF = 7;

// Fit one peak.  col[NBR_MAX_HALF + d] = plane[maxI + d][maxJ] and
// row[NBR_MAX_HALF + d] = plane[maxI][maxJ + d] for |d| <= half; out = (i', j',
// value', status).  The border and (0,0) checks use the plane size D.
__device__ void fit_peak(int kind, int pixels, int iterations, int D, float pI, float pJ, float pV,
                         float mf, const float *col, const float *row, float *out)
{
    out[0] = pI;
    out[1] = pJ;
    out[2] = pV;
    out[3] = (float)ST_SKIP;
    const int maxI = (int)pI, maxJ = (int)pJ;
    const int half = (kind == QLP_SUBPIX_POLY) ? 1 : pixels / 2;
    if (maxI == 0 && maxJ == 0) return;
    if (maxI - half < 0 || maxI + half >= D) return;
    if (maxJ - half < 0 || maxJ + half >= D) return;
    const float *cy = col + NBR_MAX_HALF, *rx = row + NBR_MAX_HALF;

    if (kind == QLP_SUBPIX_POLY) {
        float vy[3], vx[3];
        bool gauss = true;
        for (int d = -1; d <= 1; d++) {
            const float a = cy[d], c = rx[d];
            if (a < 0.0f || c < 0.0f) gauss = false;
            vy[d + 1] = a + 1e-8f;
            vx[d + 1] = c + 1e-8f;
        }
        int st = 0;
        float ri = pI, rj = pJ;
        float ti, tj;
        if (gauss) {
            ti = (float)maxI + gauss3_offset(vy[0], vy[1], vy[2]);
            tj = (float)maxJ + gauss3_offset(vx[0], vx[1], vx[2]);
        } else {
            ti = (float)maxI + (vy[0] - vy[2]) / (2.0f * vy[0] - 4.0f * vy[1] + 2.0f * vy[2]);
            tj = (float)maxJ + (vx[0] - vx[2]) / (2.0f * vx[0] - 4.0f * vx[1] + 2.0f * vx[2]);
        }
        if (isfinite(ti) && fabsf(ti - (float)maxI) <= 1.0f) {
            ri = ti;
            st |= ST_I;
        }
        if (isfinite(tj) && fabsf(tj - (float)maxJ) <= 1.0f) {
            rj = tj;
            st |= ST_J;
        }
        out[0] = ri;
        out[1] = rj;
        out[3] = (float)st;
        return;
    }

    if (kind == QLP_SUBPIX_GAUSS1D) {
        if (mf < -3.15f) {
            out[3] = (float)ST_FLOOR;
            return;
        }
        const float add = (mf < 0.0f) ? (-mf + 1e-8f) : 0.0f;
        double mi, mj;
        if (pixels == 3) {
            const float ym = cy[-1] + add, y0 = cy[0] + add, yp = cy[1] + add;
            const float xm = rx[-1] + add, x0 = rx[0] + add, xp = rx[1] + add;
            mi = (double)((float)maxI + gauss3_offset(ym, y0, yp));
            mj = (double)((float)maxJ + gauss3_offset(xm, x0, xp));
        } else {
            double yv[2 * NBR_MAX_HALF + 1], xv[2 * NBR_MAX_HALF + 1];
            for (int d = -half; d <= half; d++) {
                yv[d + half] = (double)(cy[d] + add);
                xv[d + half] = (double)(rx[d] + add);
            }
            mi = (double)maxI + gauss_lsq_mean(pixels, yv);
            mj = (double)maxJ + gauss_lsq_mean(pixels, xv);
        }
        int st = 0;
        if (fabs(mi - (double)pI) < 2.0) {
            out[0] = (float)mi;
            st |= ST_I;
        }
        if (fabs(mj - (double)pJ) < 2.0) {
            out[1] = (float)mj;
            st |= ST_J;
        }
        out[3] = (float)st;
        return;
    }

    // GUO
    if (mf < -5.15f) {
        out[3] = (float)ST_FLOOR;
        return;
    }
    const float add = (mf < 0.0f) ? (-mf + 1e-8f) : 0.0f;
    float xs[2 * NBR_MAX_HALF + 1], ys[2 * NBR_MAX_HALF + 1];
    for (int d = -half; d <= half; d++) {
        xs[d + half] = (float)(maxI + d);
        ys[d + half] = cy[d] + add;
    }
    float yMean, yAmp, xMean, xAmp;
    guo_estimate(iterations, pixels, xs, ys, yMean, yAmp);
    for (int d = -half; d <= half; d++) {
        xs[d + half] = (float)(maxJ + d);
        ys[d + half] = rx[d] + add;
    }
    guo_estimate(iterations, pixels, xs, ys, xMean, xAmp);
    float fy = yMean, fx = xMean;
    bool yValid = true, xValid = true;
    if (fabsf(yMean - (float)maxI) > pixels / 2.0f) {
        fy = (float)maxI;
        yValid = false;
    }
    if (fabsf(xMean - (float)maxJ) > pixels / 2.0f) {
        fx = (float)maxJ;
        xValid = false;
    }
    float mx = pV;
    if (yValid) {
        mx = yAmp;
        if (xValid && xAmp > mx) mx = xAmp;
    } else if (xValid) {
        mx = xAmp;
    }
    out[0] = fy;
    out[1] = fx;
    out[2] = mx;
    out[3] = (float)((yValid ? ST_I : 0) | (xValid ? ST_J : 0));
}

// Fits for every (IA, peak) of planes stored in HBM.
__global__ __launch_bounds__(256) void subpixel_kernel(int kind, int pixels, int iterations, int n,
                                                      int D, int P, const float *__restrict__ planes,
                                                      const float *__restrict__ min_floor,
                                                      const float *__restrict__ peaks_in,
                                                      float *__restrict__ peaks_out)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n * P) return;
    const int ia = gid / P;
    const float *m = planes + (size_t)ia * D * D;
    const float pI = peaks_in[(size_t)gid * 3 + 0];
    const float pJ = peaks_in[(size_t)gid * 3 + 1];
    const float pV = peaks_in[(size_t)gid * 3 + 2];
    if (kind < 0) {  // level before the interpolator start step: the integer peak
        float *o = peaks_out + (size_t)gid * 4;
        o[0] = pI;
        o[1] = pJ;
        o[2] = pV;
        o[3] = (float)ST_SKIP;
        return;
    }
    const int maxI = (int)pI, maxJ = (int)pJ;
    const int half = (kind == QLP_SUBPIX_POLY) ? 1 : pixels / 2;
    float col[2 * NBR_MAX_HALF + 1], row[2 * NBR_MAX_HALF + 1];
    for (int d = -half; d <= half; d++) {
        const int ri = maxI + d, cj = maxJ + d;
        col[NBR_MAX_HALF + d] = (ri >= 0 && ri < D && maxJ >= 0 && maxJ < D) ? m[ri * D + maxJ] : 0.0f;
        row[NBR_MAX_HALF + d] = (cj >= 0 && cj < D && maxI >= 0 && maxI < D) ? m[maxI * D + cj] : 0.0f;
    }
    fit_peak(kind, pixels, iterations, D, pI, pJ, pV, min_floor[ia], col, row,
             peaks_out + (size_t)gid * 4);
}

// Fits from the neighbourhoods exported by the fused correlation pass, and the
// displacement update of the multi-pass loop: u += i' - D/2, v += j' - D/2
// (MaxCrossResult.getNthDisplacementU :141-149, Tile.accumulateDisplacement
// :103-136 with out-of-bound clipping allowed).
__global__ __launch_bounds__(256) void subpixel_nbr_kernel(int kind, int pixels, int iterations,
                                                          int n, int D, int nbr_half,
                                                          const float *__restrict__ nbr,
                                                          const float *__restrict__ min_floor,
                                                          const float *__restrict__ peaks_in,
                                                          float *__restrict__ peaks_out,
                                                          float *__restrict__ du,
                                                          float *__restrict__ dv)
{
    const int ia = blockIdx.x * blockDim.x + threadIdx.x;
    if (ia >= n) return;
    const int L = 2 * nbr_half + 1;
    const float *nb = nbr + (size_t)ia * 2 * L;
    float col[2 * NBR_MAX_HALF + 1], row[2 * NBR_MAX_HALF + 1];
    for (int d = -nbr_half; d <= nbr_half; d++) {
        col[NBR_MAX_HALF + d] = nb[d + nbr_half];
        row[NBR_MAX_HALF + d] = nb[L + d + nbr_half];
    }
    float res[4];
    const float pI = peaks_in[(size_t)ia * 3 + 0], pJ = peaks_in[(size_t)ia * 3 + 1];
    if (kind < 0) {  // no interpolator: the integer peak
        res[0] = pI;
        res[1] = pJ;
        res[2] = peaks_in[(size_t)ia * 3 + 2];
        res[3] = (float)ST_SKIP;
    } else {
        fit_peak(kind, pixels, iterations, D, pI, pJ, peaks_in[(size_t)ia * 3 + 2], min_floor[ia],
                 col, row, res);
    }
    if (peaks_out) {
        float *o = peaks_out + (size_t)ia * 4;
        o[0] = res[0];
        o[1] = res[1];
        o[2] = res[2];
        o[3] = res[3];
    }
    if (du) {
        const int c = D / 2;
        du[ia] = du[ia] + (res[0] - (float)c);
        dv[ia] = dv[ia] + (res[1] - (float)c);
    }
}

static int check_subpixel_args(int kind, int pixels, int iterations)
{
    if (kind != QLP_SUBPIX_GAUSS1D && kind != QLP_SUBPIX_GUO && kind != QLP_SUBPIX_POLY)
        return fail(QLP_ERR_ARG, "subpixel: unknown interpolator kind");
    if (kind != QLP_SUBPIX_POLY && (pixels < 3 || pixels > 2 * NBR_MAX_HALF + 1 || (pixels & 1) == 0))
        return fail(QLP_ERR_ARG, "subpixel: pixels must be odd, 3..15");
    if (kind == QLP_SUBPIX_GUO && iterations < 1)
        return fail(QLP_ERR_ARG, "subpixel: Guo iterations must be >= 1");
    return QLP_OK;
}

int launch_subpixel(Context *ctx, int kind, int pixels, int iterations, int n, int D, int n_peaks,
                    const float *planes, const float *min_floor, const float *peaks_in,
                    float *peaks_out)
{
    if (n == 0) return QLP_OK;
    if (kind >= 0) {  // kind < 0: integer peaks copied (multi-pass levels before the start step)
        int rc = check_subpixel_args(kind, pixels, iterations);
        if (rc) return rc;
    }
    const int total = n * n_peaks;
    hipLaunchKernelGGL(subpixel_kernel, dim3((total + 255) / 256), dim3(256), 0, ctx->stream, kind,
                       pixels, iterations, n, D, n_peaks, planes, min_floor, peaks_in, peaks_out);
    return check_hip(hipGetLastError(), "subpixel kernel launch");
}

int launch_subpixel_nbr(Context *ctx, int kind, int pixels, int iterations, int n, int D,
                        int nbr_half, const float *nbr, const float *min_floor,
                        const float *peaks_in, float *peaks_out, float *du, float *dv)
{
    if (n == 0) return QLP_OK;
    if (kind >= 0) {
        int rc = check_subpixel_args(kind, pixels, iterations);
        if (rc) return rc;
        const int half = (kind == QLP_SUBPIX_POLY) ? 1 : pixels / 2;
        if (half > nbr_half) return fail(QLP_ERR_ARG, "subpixel: neighbourhood too small");
    }
    hipLaunchKernelGGL(subpixel_nbr_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, kind,
                       pixels, iterations, n, D, nbr_half, nbr, min_floor, peaks_in, peaks_out, du,
                       dv);
    return check_hip(hipGetLastError(), "subpixel kernel launch");
}

}  // namespace qlpiv
