// Velocity inheritance between adaptive levels
// (iareas/IInterAreaVelocityInheritanceStrategy.reuseIterationStepTilesParameters).
//
// Area (AdaptiveInterVelocityInheritanceStrategyArea.java:47-299) and
// Distance (AdaptiveInterVelocityInheritanceStrategyDistanceGeneric.java:30-265)
// weigh parent displacements with factors that depend only on the tile
// geometry, so the host builds a stencil per level once -- contributing
// parents in the reference's order and their float weights -- and a kernel
// forms the weighted sums of each frame.  The reference finds the contributing
// parents with a pointer that walks the parent grid; on the monotone grids of
// the super-position strategy that walk visits exactly the parents that
// overlap the child (Area, row-major) or the parent whose centre is the last
// one up-left of the child's plus its right, lower and lower-right neighbours
// (Distance), which the stencil builders compute directly.
//
// BiCubicSpline (AdaptiveInterVelocityInheritanceStrategyBiCubicSpline.java
// :29-118) evaluates commons-math3 3.6.1's PiecewiseBicubicSplineInterpolator
// (5 x 5 Akima windows) on the parent displacement grid, in FP64 as the
// reference does, one thread per child.
#include <math.h>
#include <string.h>

#include <vector>

#include "qlpiv_internal.h"

namespace qlpiv {

// checkOverlappingFactor :224-299, the overlap lengths
static void overlap_1d(int p0, int pl, int c0, int cl, int &len)
{
    len = 0;
    if (p0 <= c0) {
        if (p0 + pl > c0) {
            int end = c0 + cl;
            if (end - 1 >= p0 + pl) end = p0 + pl;
            len = end - c0;
        }
    } else if (c0 + cl > p0) {
        len = c0 + cl - p0;
    }
}

int build_inherit_stencil(int mode, const InheritGrid &pg, const InheritGrid &cg,
                          std::vector<int> &idx, std::vector<float> &w, int &K)
{
    const int cn = cg.nI * cg.nJ;
    if (mode == QLP_INHERIT_AREA) {
        // rows / columns of parents overlapping each child row / column
        std::vector<int> r0(cg.nI), r1(cg.nI), c0(cg.nJ), c1(cg.nJ);
        std::vector<int> vlen((size_t)cg.nI * pg.nI), hlen((size_t)cg.nJ * pg.nJ);
        K = 0;
        int kr = 0, kc = 0;
        for (int i = 0; i < cg.nI; i++) {
            r0[i] = -1;
            r1[i] = -2;
            for (int r = 0; r < pg.nI; r++) {
                int len;
                overlap_1d(pg.top[r * pg.nJ], pg.H, cg.top[i * cg.nJ], cg.H, len);
                vlen[(size_t)i * pg.nI + r] = len;
                if (len > 0) {
                    if (r0[i] < 0) r0[i] = r;
                    r1[i] = r;
                }
            }
            if (r0[i] < 0) return fail(QLP_ERR_ARG, "inherit: a child row overlaps no parent tile");
            kr = std::max(kr, r1[i] - r0[i] + 1);
        }
        for (int j = 0; j < cg.nJ; j++) {
            c0[j] = -1;
            c1[j] = -2;
            for (int c = 0; c < pg.nJ; c++) {
                int len;
                overlap_1d(pg.left[c], pg.W, cg.left[j], cg.W, len);
                hlen[(size_t)j * pg.nJ + c] = len;
                if (len > 0) {
                    if (c0[j] < 0) c0[j] = c;
                    c1[j] = c;
                }
            }
            if (c0[j] < 0) return fail(QLP_ERR_ARG, "inherit: a child column overlaps no parent tile");
            kc = std::max(kc, c1[j] - c0[j] + 1);
        }
        K = kr * kc;
        idx.assign((size_t)cn * K, -1);
        w.assign((size_t)cn * K, 0.0f);
        std::vector<float> area(K);
        for (int i = 0; i < cg.nI; i++)
            for (int j = 0; j < cg.nJ; j++) {
                const int t = i * cg.nJ + j;
                float acc = 0.0f;
                int k = 0;
                for (int r = r0[i]; r <= r1[i]; r++)
                    for (int c = c0[j]; c <= c1[j]; c++) {
                        area[k] = (float)(vlen[(size_t)i * pg.nI + r] * hlen[(size_t)j * pg.nJ + c]);
                        acc += area[k];
                        idx[(size_t)t * K + k] = r * pg.nJ + c;
                        k++;
                    }
                for (int q = 0; q < k; q++) w[(size_t)t * K + q] = area[q] / acc;
            }
        return QLP_OK;
    }
    if (mode == QLP_INHERIT_DISTANCE) {
        K = 4;
        idx.assign((size_t)cn * K, -1);
        w.assign((size_t)cn * K, 0.0f);
        auto cy_p = [&](int r) { return (float)pg.top[r * pg.nJ] + (float)pg.H / 2.0f; };
        auto cx_p = [&](int c) { return (float)pg.left[c] + (float)pg.W / 2.0f; };
        for (int i = 0; i < cg.nI; i++)
            for (int j = 0; j < cg.nJ; j++) {
                const int t = i * cg.nJ + j;
                const float cy = (float)cg.top[t] + (float)cg.H / 2.0f;
                const float cx = (float)cg.left[t] + (float)cg.W / 2.0f;
                int I = 0, J = 0;
                while (I + 1 < pg.nI && cy_p(I + 1) <= cy) I++;
                while (J + 1 < pg.nJ && cx_p(J + 1) <= cx) J++;
                const bool refRight = !(cx_p(J) <= cx), refDown = !(cy_p(I) <= cy);
                int par[4] = {I * pg.nJ + J, -1, -1, -1};
                if (J + 1 < pg.nJ && !refRight) par[1] = I * pg.nJ + J + 1;
                if (I + 1 < pg.nI && !refDown) par[2] = (I + 1) * pg.nJ + J;
                if (I + 1 < pg.nI && J + 1 < pg.nJ && !refRight && !refDown)
                    par[3] = (I + 1) * pg.nJ + J + 1;
                float dist[4];
                for (int k = 0; k < 4; k++) {
                    if (par[k] < 0) continue;
                    const int r = par[k] / pg.nJ, c = par[k] % pg.nJ;
                    const double dx = (double)(cx - cx_p(c)), dy = (double)(cy - cy_p(r));
                    dist[k] = (float)sqrt(dx * dx + dy * dy);
                }
                float total = 0.0f;
                int zero = -1;
                for (int k = 0; k < 4; k++) {
                    if (par[k] < 0) continue;
                    if (dist[k] == 0.0f) {
                        zero = k;
                        break;
                    }
                    total += 1.0f / dist[k];
                }
                int q = 0;
                if (zero >= 0) {
                    idx[(size_t)t * K] = par[zero];
                    w[(size_t)t * K] = 1.0f;
                } else {
                    for (int k = 0; k < 4; k++) {
                        if (par[k] < 0) continue;
                        idx[(size_t)t * K + q] = par[k];
                        w[(size_t)t * K + q] = (1.0f / dist[k]) / total;
                        q++;
                    }
                }
            }
        return QLP_OK;
    }
    return fail(QLP_ERR_ARG, "inherit: not a stencil strategy");
}

// Weighted sums of the stencil, then Tile.accumulateDisplacement :103-136 on
// the reset tile (out-of-bound clipping allowed), rounded for warping modes
// that need integer displacements (FastMath.round).
__global__ void inherit_stencil_kernel(int pairs, int cn, int pn, int K, const int *__restrict__ idx,
                                       const float *__restrict__ w, int round_it,
                                       const float *__restrict__ pu, const float *__restrict__ pv,
                                       float *__restrict__ u, float *__restrict__ v)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= pairs * cn) return;
    const int pair = gid / cn, t = gid - pair * cn;
    const float *qu = pu + (size_t)pair * pn, *qv = pv + (size_t)pair * pn;
    float du = 0.0f, dv = 0.0f;
    for (int k = 0; k < K; k++) {
        const int p = idx[(size_t)t * K + k];
        if (p < 0) break;
        const float wk = w[(size_t)t * K + k];
        du += wk * qu[p];
        dv += wk * qv[p];
    }
    if (round_it) {
        du = (float)(int)floorf(du + 0.5f);
        dv = (float)(int)floorf(dv + 0.5f);
    }
    u[gid] = 0.0f + du;
    v[gid] = 0.0f + dv;
}

// ---- commons-math3 3.6.1 PiecewiseBicubicSplineInterpolatingFunction -------

__device__ __forceinline__ int cm_bsearch(const double *a, int n, double key)
{
    int lo = 0, hi = n - 1;
    while (lo <= hi) {
        const int mid = (int)((unsigned)(lo + hi) >> 1);
        const double mv = a[mid];
        if (mv < key) lo = mid + 1;
        else if (mv > key) hi = mid - 1;
        else return mid;
    }
    return -(lo + 1);
}

// Precision.equals(x, 0.0) with one ulp, for x >= 0
__device__ __forceinline__ bool cm_zero(double x) { return (unsigned long long)__double_as_longlong(x) <= 1ull; }

__device__ __forceinline__ double cm_diff3(const double *xv, const double *yv, int d, int i0, int i1,
                                           int i2)
{
    const double x0 = yv[i0], x1 = yv[i1], x2 = yv[i2];
    const double t = xv[d] - xv[i0];
    const double t1 = xv[i1] - xv[i0];
    const double t2 = xv[i2] - xv[i0];
    const double a = (x2 - x0 - (t2 / t1 * (x1 - x0))) / (t2 * t2 - t1 * t2);
    const double b = (x1 - x0 - a * t1 * t1) / t1;
    return (2 * a * t) + b;
}

// AkimaSplineInterpolator.interpolate over 5 points and
// PolynomialSplineFunction.value at v (the constructor of PolynomialFunction
// trims trailing zero coefficients before Horner's rule).
__device__ double cm_akima5(const double *xv, const double *yv, double v)
{
    double diff[4], wt[4], fd[5];
#pragma unroll
    for (int i = 0; i < 4; i++) diff[i] = (yv[i + 1] - yv[i]) / (xv[i + 1] - xv[i]);
    wt[0] = 0.0;
#pragma unroll
    for (int i = 1; i < 4; i++) wt[i] = fabs(diff[i] - diff[i - 1]);
    {
        const double wP = wt[3], wM = wt[1];
        if (cm_zero(wP) && cm_zero(wM)) {
            const double xvv = xv[2], xvP = xv[3], xvM = xv[1];
            fd[2] = (((xvP - xvv) * diff[1]) + ((xvv - xvM) * diff[2])) / (xvP - xvM);
        } else {
            fd[2] = ((wP * diff[1]) + (wM * diff[2])) / (wP + wM);
        }
    }
    fd[0] = cm_diff3(xv, yv, 0, 0, 1, 2);
    fd[1] = cm_diff3(xv, yv, 1, 0, 1, 2);
    fd[3] = cm_diff3(xv, yv, 3, 2, 3, 4);
    fd[4] = cm_diff3(xv, yv, 4, 2, 3, 4);
    int i = cm_bsearch(xv, 5, v);
    if (i < 0) i = -i - 2;
    if (i >= 4) i--;
    double x0 = xv[0], x1 = xv[1], y0 = yv[0], y1 = yv[1], f0 = fd[0], f1 = fd[1];
#pragma unroll
    for (int k = 1; k < 4; k++)
        if (i == k) {
            x0 = xv[k];
            x1 = xv[k + 1];
            y0 = yv[k];
            y1 = yv[k + 1];
            f0 = fd[k];
            f1 = fd[k + 1];
        }
    const double wd = x1 - x0;
    const double w2 = wd * wd;
    double c[4];
    c[0] = y0;
    c[1] = f0;
    c[2] = (3 * (y1 - y0) / wd - 2 * f0 - f1) / wd;
    c[3] = (2 * (y0 - y1) / wd + f0 + f1) / w2;
    const double s = v - x0;
    if (c[3] != 0) return ((c[3] * s + c[2]) * s + c[1]) * s + c[0];
    if (c[2] != 0) return (c[2] * s + c[1]) * s + c[0];
    if (c[1] != 0) return c[1] * s + c[0];
    return c[0];
}

// searchIndex(c, val, offset 2, count 5)
__device__ __forceinline__ int cm_search5(double c, const double *val, int len)
{
    int r = cm_bsearch(val, len, c);
    if (r < 0) r = -r - 3;
    else r -= 2;
    if (r < 0) r = 0;
    if (r + 5 >= len) r = len - 5;
    return r;
}

__global__ void inherit_bicubic_kernel(int pairs, int pnI, int pnJ, int cn,
                                       const double *__restrict__ ys, const double *__restrict__ xs,
                                       const float *__restrict__ cy, const float *__restrict__ cx,
                                       int round_it, const float *__restrict__ pu,
                                       const float *__restrict__ pv, float *__restrict__ u,
                                       float *__restrict__ v)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= pairs * cn) return;
    const int pair = gid / cn, t = gid - pair * cn;
    const size_t pn = (size_t)pnI * pnJ;
    // PiecewiseBicubicSplineInterpolatingFunction.value(y, x): first variable
    // along the rows (xval = ys), second along the columns (yval = xs)
    const double y = (double)cy[t], x = (double)cx[t];
    const int i = cm_search5(y, ys, pnI), j = cm_search5(x, xs, pnJ);
    double ya[5], xa[5];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        ya[k] = ys[i + k];
        xa[k] = xs[j + k];
    }
    float res[2];
#pragma unroll
    for (int comp = 0; comp < 2; comp++) {
        const float *f = (comp == 0 ? pu : pv) + (size_t)pair * pn;
        double interp[5];
#pragma unroll
        for (int zi = 0; zi < 5; zi++) {
            double za[5];
#pragma unroll
            for (int k = 0; k < 5; k++) za[k] = (double)f[(size_t)(i + k) * pnJ + (j + zi)];
            interp[zi] = cm_akima5(ya, za, y);
        }
        res[comp] = (float)cm_akima5(xa, interp, x);
    }
    float du = res[0], dv = res[1];
    if (round_it) {
        du = (float)(int)floorf(du + 0.5f);
        dv = (float)(int)floorf(dv + 0.5f);
    }
    u[gid] = 0.0f + du;
    v[gid] = 0.0f + dv;
}

int build_inherit_bicubic(const InheritGrid &pg, const InheritGrid &cg, std::vector<double> &ys,
                          std::vector<double> &xs, std::vector<float> &cy, std::vector<float> &cx)
{
    ys.resize(pg.nI);
    xs.resize(pg.nJ);
    for (int r = 0; r < pg.nI; r++) ys[r] = (double)((float)pg.top[r * pg.nJ] + (float)pg.H / 2.0f - 0.5f);
    for (int c = 0; c < pg.nJ; c++) xs[c] = (double)((float)pg.left[c] + (float)pg.W / 2.0f - 0.5f);
    const float Xmin = (float)xs[0], Xmax = (float)xs[pg.nJ - 1];
    const float Ymin = (float)ys[0], Ymax = (float)ys[pg.nI - 1];
    const int cn = cg.nI * cg.nJ;
    cy.resize(cn);
    cx.resize(cn);
    for (int t = 0; t < cn; t++) {
        float x = (float)cg.left[t] + (float)cg.W / 2.0f - 0.5f;
        float y = (float)cg.top[t] + (float)cg.H / 2.0f - 0.5f;
        if (x < Xmin) x = Xmin;
        else if (x > Xmax) x = Xmax;
        if (y < Ymin) y = Ymin;
        else if (y > Ymax) y = Ymax;
        cx[t] = x;
        cy[t] = y;
    }
    return QLP_OK;
}

__global__ void fill_zero_kernel(float *__restrict__ x, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0f;
}

int launch_inherit_stencil(Context *ctx, int pairs, int cn, int pn, int K, const int *idx,
                           const float *w, int round_it, const float *pu, const float *pv, float *u,
                           float *v)
{
    if (pairs * cn == 0) return QLP_OK;
    hipLaunchKernelGGL(inherit_stencil_kernel, dim3((pairs * cn + 255) / 256), dim3(256), 0,
                       ctx->stream, pairs, cn, pn, K, idx, w, round_it, pu, pv, u, v);
    return check_hip(hipGetLastError(), "inherit kernel launch");
}

int launch_inherit_bicubic(Context *ctx, int pairs, int pnI, int pnJ, int cn, const double *ys,
                           const double *xs, const float *cy, const float *cx, int round_it,
                           const float *pu, const float *pv, float *u, float *v)
{
    if (pairs * cn == 0) return QLP_OK;
    if (pnI < 5 || pnJ < 5) {
        // InsufficientDataException: the reference's bilinear backup values are
        // never stored, the children keep their reset (zero) displacement
        hipLaunchKernelGGL(fill_zero_kernel, dim3((pairs * cn + 255) / 256), dim3(256), 0,
                           ctx->stream, u, pairs * cn);
        hipLaunchKernelGGL(fill_zero_kernel, dim3((pairs * cn + 255) / 256), dim3(256), 0,
                           ctx->stream, v, pairs * cn);
    } else {
        hipLaunchKernelGGL(inherit_bicubic_kernel, dim3((pairs * cn + 255) / 256), dim3(256), 0,
                           ctx->stream, pairs, pnI, pnJ, cn, ys, xs, cy, cx, round_it, pu, pv, u, v);
    }
    return check_hip(hipGetLastError(), "inherit kernel launch");
}

}  // namespace qlpiv

using namespace qlpiv;

extern "C" {

int qlp_inherit_velocity_dev(qlp_ctx *h, int mode, int pairs, int pnI, int pnJ, int pH, int pW,
                             const int *ptop, const int *pleft, int cnI, int cnJ, int cH, int cW,
                             const int *ctop, const int *cleft, int round_it, const float *pu,
                             const float *pv, float *u, float *v)
try {
    if (!h || !ptop || !pleft || !ctop || !cleft || !pu || !pv || !u || !v)
        return fail(QLP_ERR_ARG, "inherit: null argument");
    if (pairs < 0 || pnI < 1 || pnJ < 1 || cnI < 1 || cnJ < 1 || pH < 1 || pW < 1 || cH < 1 || cW < 1)
        return fail(QLP_ERR_ARG, "inherit: bad grid");
    if (pairs == 0) return QLP_OK;
    Context *ctx = ctx_of(h);
    InheritGrid pg{pnI, pnJ, pH, pW, ptop, pleft}, cg{cnI, cnJ, cH, cW, ctop, cleft};
    const int cn = cnI * cnJ, pn = pnI * pnJ;
    if (mode == QLP_INHERIT_DIRECT)
        return launch_inherit_direct(ctx, pairs, cnI, cnJ, pnI, pnJ, round_it, pu, pv, u, v);
    int rc = check_hip(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (rc) return rc;
    if (mode == QLP_INHERIT_AREA || mode == QLP_INHERIT_DISTANCE) {
        std::vector<int> idx;
        std::vector<float> w;
        int K = 0;
        rc = build_inherit_stencil(mode, pg, cg, idx, w, K);
        if (rc) return rc;
        const size_t bi = ((size_t)cn * K * 4 + 255) & ~(size_t)255;
        rc = ensure_iwork(ctx, 2 * bi);
        if (rc) return rc;
        char *b = (char *)ctx->iwork;
        rc = check_hip(hipMemcpy(b, idx.data(), (size_t)cn * K * 4, hipMemcpyHostToDevice), "upload stencil");
        if (!rc) rc = check_hip(hipMemcpy(b + bi, w.data(), (size_t)cn * K * 4, hipMemcpyHostToDevice), "upload stencil");
        if (rc) return rc;
        return launch_inherit_stencil(ctx, pairs, cn, pn, K, (const int *)b, (const float *)(b + bi),
                                      round_it, pu, pv, u, v);
    }
    if (mode == QLP_INHERIT_BICUBIC) {
        std::vector<double> ys, xs;
        std::vector<float> cy, cx;
        rc = build_inherit_bicubic(pg, cg, ys, xs, cy, cx);
        if (rc) return rc;
        const size_t b1 = ((size_t)pnI * 8 + 255) & ~(size_t)255, b2 = ((size_t)pnJ * 8 + 255) & ~(size_t)255,
                     b3 = ((size_t)cn * 4 + 255) & ~(size_t)255;
        rc = ensure_iwork(ctx, b1 + b2 + 2 * b3);
        if (rc) return rc;
        char *b = (char *)ctx->iwork;
        rc = check_hip(hipMemcpy(b, ys.data(), (size_t)pnI * 8, hipMemcpyHostToDevice), "upload grid");
        if (!rc) rc = check_hip(hipMemcpy(b + b1, xs.data(), (size_t)pnJ * 8, hipMemcpyHostToDevice), "upload grid");
        if (!rc) rc = check_hip(hipMemcpy(b + b1 + b2, cy.data(), (size_t)cn * 4, hipMemcpyHostToDevice), "upload grid");
        if (!rc) rc = check_hip(hipMemcpy(b + b1 + b2 + b3, cx.data(), (size_t)cn * 4, hipMemcpyHostToDevice), "upload grid");
        if (rc) return rc;
        return launch_inherit_bicubic(ctx, pairs, pnI, pnJ, cn, (const double *)b,
                                      (const double *)(b + b1), (const float *)(b + b1 + b2),
                                      (const float *)(b + b1 + b2 + b3), round_it, pu, pv, u, v);
    }
    return fail(QLP_ERR_ARG, "inherit: unknown strategy");
} catch (...) {
    return qlpiv::fail_exception();
}

}  // extern "C"
