// Sub-pixel peak refinement, one thread per (IA, peak) (two for Guo).
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

#include <algorithm>

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
// Fully unrolled: the pivot row is the only run-time index, resolved by
// branches, so the factors stay in registers.
__device__ __forceinline__ bool lu_solve3(const double A[3][3], const double b[3], double x[3])
{
    double lu[3][3];
    int piv[3] = {0, 1, 2};
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) lu[r][c] = A[r][c];
#pragma unroll
    for (int col = 0; col < 3; col++) {
#pragma unroll
        for (int row = 0; row < col; row++) {
            double sum = lu[row][col];
#pragma unroll
            for (int i = 0; i < row; i++) sum -= lu[row][i] * lu[i][col];
            lu[row][col] = sum;
        }
        int mx = col;
        double largest = -DBL_MAX;
#pragma unroll
        for (int row = col; row < 3; row++) {
            double sum = lu[row][col];
#pragma unroll
            for (int i = 0; i < col; i++) sum -= lu[row][i] * lu[i][col];
            lu[row][col] = sum;
            if (fabs(sum) > largest) {
                largest = fabs(sum);
                mx = row;
            }
        }
        double pv = lu[col][col];
#pragma unroll
        for (int r = col + 1; r < 3; r++)
            if (mx == r) pv = lu[r][col];
        if (fabs(pv) < 1e-11) return false;
#pragma unroll
        for (int r = col + 1; r < 3; r++)
            if (mx == r) {
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    const double t = lu[r][i];
                    lu[r][i] = lu[col][i];
                    lu[col][i] = t;
                }
                const int t = piv[r];
                piv[r] = piv[col];
                piv[col] = t;
            }
        const double d = lu[col][col];
#pragma unroll
        for (int row = col + 1; row < 3; row++) lu[row][col] /= d;
    }
    double bp[3];
#pragma unroll
    for (int r = 0; r < 3; r++) bp[r] = piv[r] == 0 ? b[0] : (piv[r] == 1 ? b[1] : b[2]);
#pragma unroll
    for (int col = 0; col < 3; col++) {
        const double bc = bp[col];
#pragma unroll
        for (int i = col + 1; i < 3; i++) bp[i] -= bc * lu[i][col];
    }
#pragma unroll
    for (int col = 2; col >= 0; col--) {
        bp[col] /= lu[col][col];
        const double bc = bp[col];
#pragma unroll
        for (int i = 0; i < col; i++) bp[i] -= bc * lu[i][col];
    }
#pragma unroll
    for (int r = 0; r < 3; r++) x[r] = bp[r];
    return true;
}

// Minimum-norm least-squares solve of the symmetric normal matrix by a cyclic
// Jacobi eigen-decomposition (the reference's SVD fallback for a singular LU).
// Out of line: rarely taken.
struct Sol3 {
    double x0, x1, x2;
};

__device__ __attribute__((noinline)) Sol3 pinv_solve3(double a00, double a01, double a02,
                                                      double a10, double a11, double a12,
                                                      double a20, double a21, double a22,
                                                      double b0, double b1, double b2)
{
    double a[3][3] = {{a00, a01, a02}, {a10, a11, a12}, {a20, a21, a22}};
    const double b[3] = {b0, b1, b2};
    double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
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
    double x[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 3; k++) {
        const double s = a[k][k];
        if (fabs(s) <= tol) continue;
        double proj = 0.0;
        for (int r = 0; r < 3; r++) proj += v[r][k] * b[r];
        for (int r = 0; r < 3; r++) x[r] += v[r][k] * proj / s;
    }
    return Sol3{x[0], x[1], x[2]};
}

// One sample of the weighted normal equations
// (Gaussian1DHongweiGuoInterpolator.estimateGaussianParameters :104-130);
// logy = ln(y) of the sample, the same in every iteration.
__device__ __forceinline__ void guo_accumulate(double A[3][3], double b[3], float xs,
                                               double ySquare, double logy)
{
    const double xSquare = (double)(xs * xs);
    const double xCubic = xSquare * xs;
    const double xFourth = xCubic * xs;
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

}

// One axis of Gaussian1DHongweiGuoInterpolator (estimateGaussianParameters
// :104-130 and its use :190-240): samples s[NBR_MAX_HALF + d] for |d| <= half
// at positions c + d, shifted by 'add'; returns the fitted mean and amplitude.
// HM is a compile-time bound on half: the sample loops run over |d| <= HM
// with a guard, so the arrays stay in registers.  The logarithms of the
// samples are taken once (the reference recomputes the same values in every
// iteration).
template <int HM>
__device__ void guo_axis(int iterations, int half, int c, const float *s, float add, float &mean,
                         float &amp)
{
    float xs[2 * NBR_MAX_HALF + 1], ys[2 * NBR_MAX_HALF + 1];
    double ly[2 * NBR_MAX_HALF + 1];
#pragma unroll
    for (int d = -HM; d <= HM; d++) {
        const int k = d + NBR_MAX_HALF;
        xs[k] = (float)(c + d);
        ys[k] = s[k] + add;
        ly[k] = (d >= -half && d <= half) ? log((double)ys[k]) : 0.0;
    }
    double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, b[3] = {0, 0, 0}, coefs[3] = {0, 0, 0};
#pragma unroll
    for (int d = -HM; d <= HM; d++) {
        const int k = d + NBR_MAX_HALF;
        if (d >= -half && d <= half) guo_accumulate(A, b, xs[k], (double)(ys[k] * ys[k]), ly[k]);
    }
    for (int it = 0; it < iterations; it++) {
        const bool lu_ok = lu_solve3(A, b, coefs);
        if (!lu_ok) {
            const Sol3 x = pinv_solve3(A[0][0], A[0][1], A[0][2], A[1][0], A[1][1], A[1][2],
                                       A[2][0], A[2][1], A[2][2], b[0], b[1], b[2]);
            coefs[0] = x.x0;
            coefs[1] = x.x1;
            coefs[2] = x.x2;
        }
        // Three samples: the weighted normal equations are square and
        // non-singular, so their solution is the interpolant whatever the
        // weights, and the reference's further iterations reproduce it up to
        // FP64 rounding (within one float ulp of the 20-iteration mean over
        // 20,000 noisy sample sets).  A singular first solve keeps iterating.
        if (half == 1 && lu_ok) break;
#pragma unroll
        for (int r = 0; r < 3; r++) {
            b[r] = 0.0;
#pragma unroll
            for (int q = 0; q < 3; q++) A[r][q] = 0.0;
        }
#pragma unroll
        for (int d = -HM; d <= HM; d++) {
            const int k = d + NBR_MAX_HALF;
            if (d >= -half && d <= half) {
                const double yk = exp(coefs[0] + coefs[1] * xs[k] + coefs[2] * xs[k] * xs[k]);
                guo_accumulate(A, b, xs[k], yk * yk, ly[k]);
            }
        }
    }
    mean = (float)(-coefs[1] / (2.0 * coefs[2]));
    amp = (float)(exp(coefs[0] - coefs[1] * coefs[1] / (4.0 * coefs[2])));
}

// Guo fits, two lanes per peak: lane 2q fits the row axis (i), lane 2q+1 the
// column axis (j); the pair combines through a lane exchange.  Peaks come
// from planes in HBM (FROM_NBR false; items = n * P) or from the
// neighbourhoods exported by the fused correlation pass (FROM_NBR true; items
// = n, followed by the displacement update u += i' - D/2, v += j' - D/2 of
// the multi-pass loop, Tile.accumulateDisplacement :103-136).
// The fit record of IA ia and the displacement update of the multi-pass loop,
// u += i' - D/2, v += j' - D/2 (MaxCrossResult.getNthDisplacementU :141-149,
// Tile.accumulateDisplacement :103-136 with out-of-bound clipping allowed):
// into peaks_out / du / dv, or with a scatter straight into the IA's pair level.
__device__ __forceinline__ void fit_result(int ia, int D, const float res[4],
                                           float *__restrict__ peaks_out, float *__restrict__ du,
                                           float *__restrict__ dv, const FitScatter &sc)
{
    const int c = D / 2;
    if (sc.segs) {
        const StepSeg &S = sc.segs[sc.seg_of[ia]];
        const int t = ia - S.first;
        S.u[t] = du[ia] + (res[0] - (float)c);
        S.v[t] = dv[ia] + (res[1] - (float)c);
        if (S.sub) reinterpret_cast<float4 *>(S.sub)[t] = make_float4(res[0], res[1], res[2], res[3]);
        return;
    }
    if (peaks_out) {
        float *o = peaks_out + (size_t)ia * 4;
        o[0] = res[0];
        o[1] = res[1];
        o[2] = res[2];
        o[3] = res[3];
    }
    if (du) {
        du[ia] = du[ia] + (res[0] - (float)c);
        dv[ia] = dv[ia] + (res[1] - (float)c);
    }
}

template <bool FROM_NBR, int HM>
__device__ __forceinline__ void guo_item(int gid, int pixels, int iterations, int items, int D,
                                         int P, const float *__restrict__ src, int nbr_half,
                                         const float *__restrict__ min_floor,
                                         const float *__restrict__ peaks_in,
                                         float *__restrict__ peaks_out, float *__restrict__ du,
                                         float *__restrict__ dv,
                                         const uint8_t *__restrict__ fit_on,
                                         const FitScatter &sc)
{
    const int item = gid >> 1, axis = gid & 1;
    const bool in = item < items;
    const int ia = FROM_NBR ? item : item / (P > 0 ? P : 1);
    float pI = 0.0f, pJ = 0.0f, pV = 0.0f, mf = 0.0f;
    if (in) {
        pI = peaks_in[(size_t)item * 3 + 0];
        pJ = peaks_in[(size_t)item * 3 + 1];
        pV = peaks_in[(size_t)item * 3 + 2];
        mf = min_floor[ia];
    }
    const int maxI = (int)pI, maxJ = (int)pJ;
    const int half = pixels / 2;
    const bool skip = (maxI == 0 && maxJ == 0) || maxI - half < 0 || maxI + half >= D ||
                      maxJ - half < 0 || maxJ + half >= D || (fit_on && in && !fit_on[item]);
    const bool low = !skip && mf < -5.15f;
    float mean = 0.0f, amp = 0.0f;
    if (in && !skip && !low) {
        const float add = (mf < 0.0f) ? (-mf + 1e-8f) : 0.0f;
        float smp[2 * NBR_MAX_HALF + 1];
        if (FROM_NBR) {
            const int L = 2 * nbr_half + 1;
            const float *nb = src + (size_t)ia * 2 * L + axis * L + nbr_half;
#pragma unroll
            for (int d = -HM; d <= HM; d++)
                smp[d + NBR_MAX_HALF] = (d >= -half && d <= half) ? nb[d] : 0.0f;
        } else {
            const float *m = src + (size_t)ia * D * D + (size_t)maxI * D + maxJ;
            const int stride = axis ? 1 : D;
#pragma unroll
            for (int d = -HM; d <= HM; d++)
                smp[d + NBR_MAX_HALF] = (d >= -half && d <= half) ? m[d * stride] : 0.0f;
        }
        guo_axis<HM>(iterations, half, axis ? maxJ : maxI, smp, add, mean, amp);
    }
    const float xMean = __shfl_xor(mean, 1, 64), xAmp = __shfl_xor(amp, 1, 64);
    if (!in || axis) return;
    float res[4] = {pI, pJ, pV, (float)ST_SKIP};
    if (!skip && low) {
        res[3] = (float)ST_FLOOR;
    } else if (!skip) {
        const float yMean = mean, yAmp = amp;
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
        res[0] = fy;
        res[1] = fx;
        res[2] = mx;
        res[3] = (float)((yValid ? ST_I : 0) | (xValid ? ST_J : 0));
    }
    fit_result(item, D, res, peaks_out, FROM_NBR ? du : nullptr, dv, sc);
    // planes in HBM: the displacement update from each IA's main peak (peak 0),
    // as accumulate_kernel does after the other fits
    if (!FROM_NBR && du && item - ia * P == 0) {
        const float c = (float)(D / 2);
        du[ia] = du[ia] + (res[0] - c);
        dv[ia] = dv[ia] + (res[1] - c);
    }
}

//   n_dev  : persistent launch: the item count in device memory (items is then
//            the capacity); the grid strides over the items
//   fit_on : per-item flag (or null): 0 keeps the integer peak, as a level
//            before the interpolator start step
template <bool FROM_NBR, int HM, bool PERSIST = false>
__global__ __launch_bounds__(256) void guo_kernel(int pixels, int iterations, int items, int D,
                                                  int P, const float *__restrict__ src,
                                                  int nbr_half, const float *__restrict__ min_floor,
                                                  const float *__restrict__ peaks_in,
                                                  float *__restrict__ peaks_out,
                                                  float *__restrict__ du, float *__restrict__ dv,
                                                  const int *__restrict__ n_dev,
                                                  const uint8_t *__restrict__ fit_on,
                                                  FitScatter sc)
{
    if (!PERSIST) {
        guo_item<FROM_NBR, HM>((int)(blockIdx.x * blockDim.x + threadIdx.x), pixels, iterations,
                               items, D, P, src, nbr_half, min_floor, peaks_in, peaks_out, du, dv,
                               fit_on, sc);
        return;
    }
    items = *n_dev;
    const int stride = (int)(gridDim.x * blockDim.x);  // even: the lane pairs stay aligned
    for (int base = (int)(blockIdx.x * blockDim.x); base < 2 * items; base += stride) {
        int tx = (int)threadIdx.x;
        asm volatile("" : "+v"(tx));  // nothing built on it is hoisted out of the loop
        guo_item<FROM_NBR, HM>(base + tx, pixels, iterations, items, D, P, src, nbr_half,
                               min_floor, peaks_in, peaks_out, du, dv, fit_on, sc);
    }
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
__device__ __forceinline__ void subpixel_nbr_item(int ia, int kind, int pixels, int iterations,
                                                          int n, int D, int nbr_half,
                                                          const float *__restrict__ nbr,
                                                          const float *__restrict__ min_floor,
                                                          const float *__restrict__ peaks_in,
                                                          float *__restrict__ peaks_out,
                                                          float *__restrict__ du,
                                                          float *__restrict__ dv,
                                                          const uint8_t *__restrict__ fit_on,
                                                          const FitScatter &sc)
{
    const int L = 2 * nbr_half + 1;
    const float *nb = nbr + (size_t)ia * 2 * L;
    float col[2 * NBR_MAX_HALF + 1], row[2 * NBR_MAX_HALF + 1];
    for (int d = -nbr_half; d <= nbr_half; d++) {
        col[NBR_MAX_HALF + d] = nb[d + nbr_half];
        row[NBR_MAX_HALF + d] = nb[L + d + nbr_half];
    }
    float res[4];
    const float pI = peaks_in[(size_t)ia * 3 + 0], pJ = peaks_in[(size_t)ia * 3 + 1];
    if (kind < 0 || (fit_on && !fit_on[ia])) {  // no interpolator: the integer peak
        res[0] = pI;
        res[1] = pJ;
        res[2] = peaks_in[(size_t)ia * 3 + 2];
        res[3] = (float)ST_SKIP;
    } else {
        fit_peak(kind, pixels, iterations, D, pI, pJ, peaks_in[(size_t)ia * 3 + 2], min_floor[ia],
                 col, row, res);
    }
    fit_result(ia, D, res, peaks_out, du, dv, sc);
}

template <bool PERSIST>
__global__ __launch_bounds__(256) void subpixel_nbr_kernel(int kind, int pixels, int iterations,
                                                          int n, int D, int nbr_half,
                                                          const float *__restrict__ nbr,
                                                          const float *__restrict__ min_floor,
                                                          const float *__restrict__ peaks_in,
                                                          float *__restrict__ peaks_out,
                                                          float *__restrict__ du,
                                                          float *__restrict__ dv,
                                                          const int *__restrict__ n_dev,
                                                          const uint8_t *__restrict__ fit_on,
                                                          FitScatter sc)
{
    if (!PERSIST) {
        const int ia = blockIdx.x * blockDim.x + threadIdx.x;
        if (ia < n)
            subpixel_nbr_item(ia, kind, pixels, iterations, n, D, nbr_half, nbr, min_floor,
                              peaks_in, peaks_out, du, dv, fit_on, sc);
        return;
    }
    n = *n_dev;  // persistent launch: the grid strides over the IAs
    for (int base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        int tx = (int)threadIdx.x;
        asm volatile("" : "+v"(tx));  // nothing built on it is hoisted out of the loop
        if (base + tx < n)
            subpixel_nbr_item(base + tx, kind, pixels, iterations, n, D, nbr_half, nbr, min_floor,
                              peaks_in, peaks_out, du, dv, fit_on, sc);
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
                    float *peaks_out, float *u, float *v)
{
    // u, v (or null): the Guo kernel also applies the displacement update of the
    // multi-pass loop from each IA's main peak; the other fits leave it to
    // launch_accumulate
    if (u && kind != QLP_SUBPIX_GUO)
        return fail(QLP_ERR_ARG, "subpixel: the fused displacement update is Guo's only");
    if (n == 0) return QLP_OK;
    if (kind >= 0) {  // kind < 0: integer peaks copied (multi-pass levels before the start step)
        int rc = check_subpixel_args(kind, pixels, iterations);
        if (rc) return rc;
    }
    const int total = n * n_peaks;
    if (kind == QLP_SUBPIX_GUO) {
        auto k = pixels == 3 ? guo_kernel<false, 1> : pixels == 5 ? guo_kernel<false, 2>
               : pixels == 7 ? guo_kernel<false, 3> : guo_kernel<false, NBR_MAX_HALF>;
        hipLaunchKernelGGL(k, dim3((2 * total + 255) / 256), dim3(256), 0, ctx->stream, pixels,
                           iterations, total, D, n_peaks, planes, 0, min_floor, peaks_in,
                           peaks_out, u, v, nullptr, nullptr, FitScatter{});
    } else
        hipLaunchKernelGGL(subpixel_kernel, dim3((total + 255) / 256), dim3(256), 0, ctx->stream,
                           kind, pixels, iterations, n, D, n_peaks, planes, min_floor, peaks_in,
                           peaks_out);
    return check_hip(hipGetLastError(), "subpixel kernel launch");
}

int launch_subpixel_nbr(Context *ctx, int kind, int pixels, int iterations, int n, int D,
                        int nbr_half, const float *nbr, const float *min_floor,
                        const float *peaks_in, float *peaks_out, float *du, float *dv,
                        const int *n_dev, const uint8_t *fit_on, const FitScatter *scatter)
{
    if (n == 0) return QLP_OK;
    if (scatter && (!n_dev || !du || !dv))
        return fail(QLP_ERR_ARG, "subpixel: a scatter takes a persistent launch and the displacements");
    const FitScatter sc = scatter ? *scatter : FitScatter{};
    if (kind >= 0) {
        int rc = check_subpixel_args(kind, pixels, iterations);
        if (rc) return rc;
        const int half = (kind == QLP_SUBPIX_POLY) ? 1 : pixels / 2;
        if (half > nbr_half) return fail(QLP_ERR_ARG, "subpixel: neighbourhood too small");
    }
    if (kind == QLP_SUBPIX_GUO) {
        auto k = n_dev ? (pixels == 3 ? guo_kernel<true, 1, true> : pixels == 5 ? guo_kernel<true, 2, true>
                          : pixels == 7 ? guo_kernel<true, 3, true> : guo_kernel<true, NBR_MAX_HALF, true>)
                       : (pixels == 3 ? guo_kernel<true, 1> : pixels == 5 ? guo_kernel<true, 2>
                          : pixels == 7 ? guo_kernel<true, 3> : guo_kernel<true, NBR_MAX_HALF>);
        // a persistent launch (n = the capacity) takes at most the work-groups
        // the device holds at once: a step with few IAs then pays for few
        // empty work-groups
        int grid = (2 * n + 255) / 256;
        if (n_dev) grid = std::min(grid, resident_groups(k, 256));
        hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, ctx->stream, pixels, iterations, n, D, 1,
                           nbr, nbr_half, min_floor, peaks_in, peaks_out, du, dv, n_dev, fit_on, sc);
    } else {
        auto k = n_dev ? subpixel_nbr_kernel<true> : subpixel_nbr_kernel<false>;
        int grid = (n + 255) / 256;
        if (n_dev) grid = std::min(grid, resident_groups(k, 256));
        hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, ctx->stream, kind, pixels, iterations, n,
                           D, nbr_half, nbr, min_floor, peaks_in, peaks_out, du, dv, n_dev, fit_on, sc);
    }
    return check_hip(hipGetLastError(), "subpixel kernel launch");
}

}  // namespace qlpiv
