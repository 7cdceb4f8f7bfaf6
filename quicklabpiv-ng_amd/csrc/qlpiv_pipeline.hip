// Multi-pass PIV over device-resident frame pairs.
//
// Host side: the IA geometry of every adaptive level
// (iareas/AdaptiveInterAreaStrategySuperPosition.java:31-197, evaluated in the
// reference's float arithmetic) and the per-level loop of
// jobs/LocalPIVOpenCLGpuJob.call :188-409.  Device side, per level:
//   1. displacement state: zero at level 0, else inheritance from the parent
//      level: Direct (AdaptiveInterVelocityInheritanceStrategyDirect: child
//      (2I+a, 2J+b) takes parent (I, J); children without a parent keep 0),
//      Area, Distance or BiCubicSpline (qlpiv_inherit.hip); NoImageWarping
//      rounds the inherited value;
//   2. one fused pass: clip A, clip/shift/micro-warp B, correlate, FromCenter,
//      export the peak neighbourhood (with MultiPeaks or the Simple finder: the
//      pass stores the planes and the peak search runs over them);
//   3. the sub-pixel fit of every peak and the displacement update
//      u += i' - D/2 from the main peak;
//   4. optionally the vector validation with secondary-peak and bilinear
//      replacement (jobs/validator/VectorValidatorJob.java).
// Everything stays in HBM between levels; the host only enqueues launches.
#include <math.h>

#include <algorithm>
#include <vector>

#include "qlpiv_device.h"
#include "qlpiv_internal.h"

namespace qlpiv {

struct Level {
    int N = 0, nI = 0, nJ = 0, n = 0;
    std::vector<int> top_h, left_h;
    int *top = nullptr, *left = nullptr;  // n
    float *u = nullptr, *v = nullptr;     // pairs * n
    float *peaks = nullptr;               // pairs * n * P * 3 (integer search)
    float *mf = nullptr;                  // pairs * n
    float *nbr = nullptr;                 // pairs * n * 2 * NBR_L
    float *sub = nullptr;                 // pairs * n * P * 4
    float *planes = nullptr;              // pairs * n * D * D (optional)
    int *count = nullptr;                 // pairs * n peaks found (MultiPeaks)
    int *rank = nullptr;                  // n, spiral visit order of the validation
    void *vwork = nullptr;                // validation workspace
    uint8_t *invalid = nullptr;           // pairs * n
    int *stats = nullptr;                 // pairs * 4
    // inheritance from the parent level (Area / Distance stencil, BiCubicSpline grid)
    int K = 0;
    int *sidx = nullptr;                  // n * K
    float *sw = nullptr;                  // n * K
    double *bys = nullptr, *bxs = nullptr;  // parent centres
    float *bcy = nullptr, *bcx = nullptr;   // clamped child centres (n)
};

constexpr int NBR_HALF = 7;
constexpr int NBR_L = 2 * NBR_HALF + 1;

static inline float f32(double x) { return (float)x; }

static int round_half_up(float x) { return (int)floorf(x + 0.5f); }

// One level of AdaptiveInterAreaStrategySuperPosition (level0: positions
// margin + round(k * inc); later levels accumulate the float position).
static void level_grid(int H, int W, int tile, float of, const int m[4], bool level0, Level &L)
{
    const int mt = m[0], mb = m[1], ml = m[2], mr = m[3];
    const float hinc = of * (float)tile, winc = of * (float)tile;
    const int nJ = (int)(0.01f + ((float)(W - (ml + mr)) - ((float)tile - winc)) / winc);
    const int nI = (int)(0.01f + ((float)(H - (mt + mb)) - ((float)tile - hinc)) / hinc);
    const int remI = H - (int)ceilf((float)(nI - 1) * hinc + (float)tile) - (mt + mb);
    const int remJ = W - (int)ceilf((float)(nJ - 1) * winc + (float)tile) - (ml + mr);
    const int top0 = mt + (int)floor(remI / 2.0);  // FastMath.floorDiv
    const int left0 = ml + (int)floor(remJ / 2.0);
    L.N = tile;
    L.nI = nI;
    L.nJ = nJ;
    L.n = nI * nJ;
    L.top_h.resize(L.n);
    L.left_h.resize(L.n);
    if (level0) {
        for (int i = 0; i < nI; i++)
            for (int j = 0; j < nJ; j++) {
                L.top_h[i * nJ + j] = top0 + round_half_up((float)i * hinc);
                L.left_h[i * nJ + j] = left0 + round_half_up((float)j * winc);
            }
    } else {
        float tp = (float)top0;
        for (int i = 0; i < nI; i++, tp += hinc) {
            float lp = (float)left0;
            for (int j = 0; j < nJ; j++, lp += winc) {
                L.top_h[i * nJ + j] = round_half_up(tp);
                L.left_h[i * nJ + j] = round_half_up(lp);
            }
        }
    }
}

void level_origins(int H, int W, int tile, float overlap_factor, const int margins[4], bool first,
                   int &nI, int &nJ, std::vector<int> &top, std::vector<int> &left)
{
    Level L;
    level_grid(H, W, tile, overlap_factor, margins, first, L);
    nI = L.nI;
    nJ = L.nJ;
    top = std::move(L.top_h);
    left = std::move(L.left_h);
}

void grid_origins(int H, int W, int tile, float overlap_factor, const int margins[4], int &nI,
                  int &nJ, std::vector<int> &top, std::vector<int> &left)
{
    Level L;
    level_grid(H, W, tile, overlap_factor, margins, true, L);
    nI = L.nI;
    nJ = L.nJ;
    top = std::move(L.top_h);
    left = std::move(L.left_h);
}

__global__ void fill_kernel(float *__restrict__ x, float value, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = value;
}

__global__ void inherit_direct_kernel(int pairs, int nI, int nJ, int pI, int pJ, int round_it,
                                      const float *__restrict__ pu, const float *__restrict__ pv,
                                      float *__restrict__ u, float *__restrict__ v)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = nI * nJ;
    if (gid >= pairs * n) return;
    const int pair = gid / n, t = gid - pair * n;
    const int ci = t / nJ, cj = t - (t / nJ) * nJ;
    const int qi = ci >> 1, qj = cj >> 1;
    float iu = 0.0f, iv = 0.0f;
    if (qi < pI && qj < pJ) {
        const float a = pu[(size_t)pair * pI * pJ + qi * pJ + qj];
        const float b = pv[(size_t)pair * pI * pJ + qi * pJ + qj];
        // accumulateDisplacement onto a reset tile: 0 + increment
        iu = round_it ? (float)(int)floorf(a + 0.5f) : a;
        iv = round_it ? (float)(int)floorf(b + 0.5f) : b;
        iu = 0.0f + iu;
        iv = 0.0f + iv;
    }
    u[gid] = iu;
    v[gid] = iv;
}

}  // namespace qlpiv

using namespace qlpiv;

struct qlp_piv_plan {
    Context *ctx = nullptr;
    int H = 0, W = 0, pairs = 0;
    int P = 1;  // peaks per IA
    qlp_piv_config cfg{};
    std::vector<Level> levels;
    size_t bytes = 0;
    // warp buffers of the levels after the first: Mini warp A and B tiles and
    // their maxima; Full-image warp: the warped second frames and the field
    // workspace (sized for the largest level)
    uint8_t *tA = nullptr, *tB = nullptr, *wimg = nullptr;
    float *mA = nullptr, *mB = nullptr;
    void *fwork = nullptr;
    size_t tiles_bytes = 0, fwork_bytes = 0;
    // float and 10/12/16-bit frames: the A and B float tiles of every level
    // (sized for the largest); 10/12/16-bit frames converted to float
    float *fA = nullptr, *fB = nullptr, *fframes = nullptr;
    float *fmaxp = nullptr;  // Full warp of float frames: each pair's frame maximum
    size_t ftiles_bytes = 0;
    // IAs of the largest IA-128 pass: sizes the context's tile-slot workspace
    int big_ias = 0;
    size_t mp_work = 0;  // MultiPeaks entry workspace of the largest level (context-owned)
    // optional per-level timing: events before the state kernel, after the
    // correlation pass, after the fit and after the validation of every level
    bool timing = false;
    std::vector<hipEvent_t> ev;
};

// Levels after the first correlate from tile buffers (warp kernels in front of
// the pass) unless the warp is SecondImage Micro or Full, which the pass reads
// from the frames itself.
static bool tiles_route(const qlp_piv_config &c)
{
    return c.warp_mode == QLP_WARP_MINI ||
           ((c.warp_mode == QLP_WARP_MICRO || c.warp_mode == QLP_WARP_FULL) &&
            c.warp_order != QLP_WARP_ORDER_SECOND);
}

// Warped whole frames a Full-warp plan keeps: B (SecondImage), A (FirstImage), both.
static int full_frames(const qlp_piv_config &c)
{
    return c.warp_mode != QLP_WARP_FULL ? 0 : (c.warp_order == QLP_WARP_ORDER_BOTH ? 2 : 1);
}

static void plan_free(qlp_piv_plan *p)
{
    for (auto &L : p->levels) {
        void *bufs[] = {L.top, L.left, L.u, L.v, L.peaks, L.mf, L.nbr, L.sub, L.planes,
                        L.count, L.rank, L.vwork, L.invalid, L.stats, L.sidx, L.sw,
                        L.bys, L.bxs, L.bcy, L.bcx};
        for (void *b : bufs)
            if (b) (void)hipFree(b);
    }
    for (void *b : {(void *)p->tA, (void *)p->tB, (void *)p->wimg, (void *)p->mA, (void *)p->mB,
                    p->fwork, (void *)p->fA, (void *)p->fB, (void *)p->fframes, (void *)p->fmaxp})
        if (b) (void)hipFree(b);
    p->tA = p->tB = p->wimg = nullptr;
    p->mA = p->mB = p->fA = p->fB = p->fframes = p->fmaxp = nullptr;
    p->fwork = nullptr;
    for (hipEvent_t e : p->ev) (void)hipEventDestroy(e);
    p->ev.clear();
    if (p->ctx) p->ctx->hbm_in_use -= p->bytes;
    p->levels.clear();
}

// Host-side inheritance set-up of one level from its parent (stencil or spline grid).
static int prepare_inherit(int mode, const Level &P, Level &L, size_t &bytes)
{
    InheritGrid pg{P.nI, P.nJ, P.N, P.N, P.top_h.data(), P.left_h.data()};
    InheritGrid cg{L.nI, L.nJ, L.N, L.N, L.top_h.data(), L.left_h.data()};
    const size_t n = L.n;
    int rc;
    if (mode == QLP_INHERIT_BICUBIC) {
        std::vector<double> ys, xs;
        std::vector<float> cy, cx;
        rc = build_inherit_bicubic(pg, cg, ys, xs, cy, cx);
        if (!rc) rc = check_hip(hipMalloc(&L.bys, ys.size() * 8), "hipMalloc plan");
        if (!rc) rc = check_hip(hipMalloc(&L.bxs, xs.size() * 8), "hipMalloc plan");
        if (!rc) rc = check_hip(hipMalloc(&L.bcy, n * 4), "hipMalloc plan");
        if (!rc) rc = check_hip(hipMalloc(&L.bcx, n * 4), "hipMalloc plan");
        if (!rc) rc = check_hip(hipMemcpy(L.bys, ys.data(), ys.size() * 8, hipMemcpyHostToDevice), "upload grid");
        if (!rc) rc = check_hip(hipMemcpy(L.bxs, xs.data(), xs.size() * 8, hipMemcpyHostToDevice), "upload grid");
        if (!rc) rc = check_hip(hipMemcpy(L.bcy, cy.data(), n * 4, hipMemcpyHostToDevice), "upload grid");
        if (!rc) rc = check_hip(hipMemcpy(L.bcx, cx.data(), n * 4, hipMemcpyHostToDevice), "upload grid");
        bytes += (ys.size() + xs.size()) * 8 + n * 8;
        return rc;
    }
    std::vector<int> idx;
    std::vector<float> w;
    rc = build_inherit_stencil(mode, pg, cg, idx, w, L.K);
    if (!rc) rc = check_hip(hipMalloc(&L.sidx, idx.size() * 4), "hipMalloc plan");
    if (!rc) rc = check_hip(hipMalloc(&L.sw, w.size() * 4), "hipMalloc plan");
    if (!rc) rc = check_hip(hipMemcpy(L.sidx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice), "upload stencil");
    if (!rc) rc = check_hip(hipMemcpy(L.sw, w.data(), w.size() * 4, hipMemcpyHostToDevice), "upload stencil");
    bytes += idx.size() * 8;
    return rc;
}

int qlpiv::launch_inherit_direct(Context *ctx, int pairs, int nI, int nJ, int pI, int pJ,
                                 int round_it, const float *pu, const float *pv, float *u, float *v)
{
    const int pn = pairs * nI * nJ;
    if (pn == 0) return QLP_OK;
    hipLaunchKernelGGL(inherit_direct_kernel, dim3((pn + 255) / 256), dim3(256), 0, ctx->stream,
                       pairs, nI, nJ, pI, pJ, round_it, pu, pv, u, v);
    return check_hip(hipGetLastError(), "inherit kernel launch");
}

extern "C" {

// Checks a configuration and lays out the level geometry of a plan (host only):
// fills p->levels, p->P and the device bytes the plan's buffers need.
static int plan_layout(int H, int W, int pairs, const qlp_piv_config *cfg, qlp_piv_plan *p,
                       size_t &need)
{
    if (H < 1 || W < 1 || pairs < 1) return fail(QLP_ERR_ARG, "piv_plan_create: bad frame size");
    if (cfg->overlap_factor < 0.1f || cfg->overlap_factor > 0.9f)
        return fail(QLP_ERR_ARG, "Overlap factor must greater than 1/10 and less than 9/10");
    if (cfg->ia_end < 1 || cfg->ia_start % cfg->ia_end)
        return fail(QLP_ERR_ARG, "Interrogation Area start pixels is not a multiple of Interrogation Area end pixels");
    int steps = cfg->ia_start / cfg->ia_end, nlev = 1;
    while ((1 << (nlev - 1)) < steps) nlev++;
    if ((1 << (nlev - 1)) != steps)
        return fail(QLP_ERR_ARG, "Total number of steps could not be properly computed since step sizes ratio is not a power of two");
    if (cfg->warp_mode < QLP_WARP_NONE || cfg->warp_mode > QLP_WARP_FULL)
        return fail(QLP_ERR_ARG, "piv_plan_create: unknown warp mode");
    if (cfg->warp_order < QLP_WARP_ORDER_SECOND || cfg->warp_order > QLP_WARP_ORDER_BOTH)
        return fail(QLP_ERR_ARG, "piv_plan_create: unknown warp order");
    if (cfg->norm_mode != QLP_NORM_PER_TILE && cfg->norm_mode != QLP_NORM_A_MAX)
        return fail(QLP_ERR_ARG, "piv_plan_create: unknown normalisation mode");
    for (int tile = cfg->ia_start; tile >= cfg->ia_end; tile /= 2)
        if (tile != 16 && tile != 32 && tile != 64 && tile != 128)
            return fail(QLP_ERR_UNSUPPORTED, "piv_plan_create: IA sizes must be 16, 32, 64 or 128");
    if (cfg->subpix_kind >= 0) {
        const int half = cfg->subpix_kind == QLP_SUBPIX_POLY ? 1 : cfg->subpix_pixels / 2;
        if (half > NBR_HALF || (cfg->subpix_kind != QLP_SUBPIX_POLY && cfg->subpix_pixels % 2 == 0))
            return fail(QLP_ERR_ARG, "piv_plan_create: sub-pixel window must be odd and <= 15");
    }
    const bool multi = cfg->peak_kind == QLP_PEAK_MULTI;
    // the Simple and MultiPeaks finders run on stored planes
    const bool separate = cfg->peak_kind != QLP_PEAK_FROM_CENTER;
    if (cfg->peak_kind != QLP_PEAK_FROM_CENTER && !multi && cfg->peak_kind != QLP_PEAK_SIMPLE)
        return fail(QLP_ERR_ARG, "piv_plan_create: unknown peak finder kind");
    if (!multi && cfg->n_peaks > 1)
        return fail(QLP_ERR_ARG, "find_peaks: FromCenter and Simple find exactly one peak");
    const int P = multi ? cfg->n_peaks : 1;
    if (multi) {
        if (P < 1 || P > 16) return fail(QLP_ERR_ARG, "find_peaks: number of peaks must be 1..16");
        const int k = cfg->peak_kernel_width;
        for (int tile = cfg->ia_start; tile >= cfg->ia_end; tile /= 2) {
            const int D = 2 * tile - 1;
            const int iSlices = k >= 1 ? (D + k - 1) / k : 0;
            if (k < 1 || k > D || (iSlices - iSlices / 2) * k > D)
                return fail(QLP_ERR_ARG, "find_peaks: kernel width out of range");
        }
    }
    int rc = check_valid_config(cfg->valid);
    if (rc) return rc;
    if (cfg->inherit_mode < QLP_INHERIT_DIRECT || cfg->inherit_mode > QLP_INHERIT_BICUBIC)
        return fail(QLP_ERR_ARG, "piv_plan_create: unknown velocity inheritance strategy");
    if (cfg->xcorr_form < QLP_XCORR_EXACT || cfg->xcorr_form > QLP_XCORR_MFMA_BF16_SINGLE)
        return fail(QLP_ERR_ARG, "piv_plan_create: unknown correlation form");
    if (cfg->frame_type != QLP_FRAME_U8 && cfg->frame_type != QLP_FRAME_F32 &&
        cfg->frame_type != QLP_FRAME_U16)
        return fail(QLP_ERR_ARG, "piv_plan_create: unknown frame type");
    if (cfg->frame_type == QLP_FRAME_U16 && cfg->frame_bits != 10 && cfg->frame_bits != 12 &&
        cfg->frame_bits != 16)
        return fail(QLP_ERR_ARG, "piv_plan_create: 10/12/16-bit frames need a bit depth of 10, 12 or 16");
    const bool f32 = cfg->frame_type != QLP_FRAME_U8;
    // the Full warp of an ImageInt16 stores each sample as (byte) into its short
    // matrix (MatrixInt16.setElement(float) :95-101): not provided
    if (cfg->frame_type == QLP_FRAME_U16 && cfg->warp_mode == QLP_WARP_FULL)
        return fail(QLP_ERR_UNSUPPORTED,
                    "piv_plan_create: 10/12/16-bit frames take no warp, Micro or Mini");
    if (f32 && cfg->xcorr_form != QLP_XCORR_EXACT)
        return fail(QLP_ERR_UNSUPPORTED, "piv_plan_create: float and 10/12/16-bit frames take the exact correlation form");
    const bool valid = cfg->valid.n_stages > 0;
    p->H = H;
    p->W = W;
    p->pairs = pairs;
    p->cfg = *cfg;
    p->P = P;
    const int m[4] = {cfg->margin_top, cfg->margin_bottom, cfg->margin_left, cfg->margin_right};
    need = 0;
    for (int l = 0, tile = cfg->ia_start; l < nlev; l++, tile /= 2) {
        Level L;
        level_grid(H, W, tile, cfg->overlap_factor, m, l == 0, L);
        if (L.nI < 1 || L.nJ < 1) return fail(QLP_ERR_ARG, "piv_plan_create: no IA fits the frame");
        const size_t n = L.n, pn = (size_t)pairs * n, D = 2 * tile - 1;
        need += n * 8 + pn * (8 + 12 * P + 4 + 4 * 2 * NBR_L + 16 * P);
        if (separate || cfg->keep_planes >= 2 || (cfg->keep_planes == 1 && l == nlev - 1))
            need += pn * D * D * 4;
        if (multi) need += pn * 4;
        if (valid) need += n * 4 + valid_work_bytes(pairs, (int)n) + pn + (size_t)pairs * 16;
        if (l > 0 && (cfg->warp_mode == QLP_WARP_MINI || cfg->warp_mode == QLP_WARP_FULL) &&
            (L.nI < 2 || L.nJ < 2))
            return fail(QLP_ERR_UNSUPPORTED,
                        "piv_plan_create: spline warps need 2 tiles in each direction");
        if (tile == 128) p->big_ias = std::max(p->big_ias, (int)pn);
        // MultiPeaks entries: the exact 16 / 32 / 64 px passes and the 64 px
        // matrix-core pass write them, the finder forms them from planes of 63^2
        // or less
        if (multi && tile <= 64)
            p->mp_work = std::max(p->mp_work, multipeak_work_bytes((int)pn, (int)D, P,
                                                                   cfg->peak_kernel_width));
        if (l > 0 && tiles_route(*cfg) && !f32)
            p->tiles_bytes = std::max(p->tiles_bytes, pn * (size_t)tile * tile);
        if (f32) p->ftiles_bytes = std::max(p->ftiles_bytes, pn * (size_t)tile * tile * sizeof(float));
        if (l > 0 && cfg->warp_mode == QLP_WARP_FULL)
            p->fwork_bytes = std::max(p->fwork_bytes, full_warp_work_bytes(pairs, L.nI, L.nJ));
        p->levels.push_back(std::move(L));
    }
    if ((tiles_route(*cfg) && p->tiles_bytes) || f32) {
        size_t mx = 0;
        for (const Level &L : p->levels) mx = std::max(mx, (size_t)pairs * L.n);
        need += 2 * p->tiles_bytes + 2 * p->ftiles_bytes + 2 * mx * 4;
    }
    if (cfg->warp_mode == QLP_WARP_FULL && p->fwork_bytes)
        need += (size_t)full_frames(*cfg) * pairs * H * W * (f32 ? sizeof(float) : 1) +
                p->fwork_bytes + (f32 ? sizeof(float) * pairs : 0);
    if (cfg->frame_type == QLP_FRAME_U16) need += 2 * sizeof(float) * (size_t)pairs * H * W;
    return QLP_OK;
}

int qlp_piv_plan_estimate(int H, int W, int pairs, const qlp_piv_config *cfg, size_t *bytes)
try {
    if (!cfg || !bytes) return fail(QLP_ERR_ARG, "piv_plan_estimate: null argument");
    qlp_piv_plan p;
    size_t need = 0;
    if (int rc = plan_layout(H, W, pairs, cfg, &p, need)) return rc;
    // the inheritance stencils (at most 9 parents per tile) or spline tables
    if (cfg->inherit_mode != QLP_INHERIT_DIRECT)
        for (const Level &L : p.levels) need += L.n * 9 * 8 + 64 * 1024;
    // the context's IA-128 workspace (shared by the context's plans)
    if (p.big_ias) need += big_work_estimate(p.big_ias);
    // the context's MultiPeaks workspace (shared by the context's plans)
    need += p.mp_work;
    *bytes = need;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_piv_plan_create(qlp_ctx *h, int H, int W, int pairs, const qlp_piv_config *cfg,
                        qlp_piv_plan **out)
try {
    if (!h || !cfg || !out) return fail(QLP_ERR_ARG, "piv_plan_create: null argument");
    *out = nullptr;
    Context *ctx = ctx_of(h);
    qlp_piv_plan *p = new qlp_piv_plan();
    p->ctx = ctx;
    size_t need = 0;
    int rc = plan_layout(H, W, pairs, cfg, p, need);
    if (rc) {
        delete p;
        return rc;
    }
    const bool multi = cfg->peak_kind == QLP_PEAK_MULTI;
    const bool separate = cfg->peak_kind != QLP_PEAK_FROM_CENTER;
    const bool valid = cfg->valid.n_stages > 0;
    const int P = p->P;
    if (ctx->hbm_budget && ctx->hbm_in_use + need > ctx->hbm_budget) {
        delete p;
        return fail(QLP_ERR_BUDGET, "piv plan exceeds the context's HBM budget");
    }
    for (size_t l = 0; l < p->levels.size() && !rc; l++) {
        Level &L = p->levels[l];
        const size_t n = L.n, pn = (size_t)pairs * n, D = 2 * L.N - 1;
        rc = check_hip(hipMalloc(&L.top, n * 4), "hipMalloc plan");
        if (!rc) rc = check_hip(hipMalloc(&L.left, n * 4), "hipMalloc plan");
        if (!rc) rc = check_hip(hipMalloc(&L.u, pn * 4), "hipMalloc plan");
        if (!rc) rc = check_hip(hipMalloc(&L.v, pn * 4), "hipMalloc plan");
        if (!rc) rc = check_hip(hipMalloc(&L.peaks, pn * 12 * P), "hipMalloc plan");
        if (!rc) rc = check_hip(hipMalloc(&L.mf, pn * 4), "hipMalloc plan");
        if (!rc) rc = check_hip(hipMalloc(&L.nbr, pn * 4 * 2 * NBR_L), "hipMalloc plan");
        if (!rc) rc = check_hip(hipMalloc(&L.sub, pn * 16 * P), "hipMalloc plan");
        if (!rc && (separate || cfg->keep_planes >= 2 || (cfg->keep_planes == 1 && l + 1 == p->levels.size())))
            rc = check_hip(hipMalloc(&L.planes, pn * D * D * 4), "hipMalloc plan");
        if (!rc && multi) rc = check_hip(hipMalloc(&L.count, pn * 4), "hipMalloc plan");
        if (!rc && valid) {
            rc = check_hip(hipMalloc(&L.rank, n * 4), "hipMalloc plan");
            if (!rc) rc = check_hip(hipMalloc(&L.vwork, valid_work_bytes(pairs, (int)n)), "hipMalloc plan");
            if (!rc) {  // the active counts start at 0 and every sweep leaves them at 0
                ValidLaunch z;
                valid_work_carve(L.vwork, pairs, (int)n, z);
                rc = check_hip(hipMemset(z.act_count, 0, (size_t)pairs * 4), "plan validation counts");
                if (!rc) rc = check_hip(hipDeviceSynchronize(), "plan validation counts");
            }
            if (!rc) rc = check_hip(hipMalloc(&L.invalid, pn), "hipMalloc plan");
            if (!rc) rc = check_hip(hipMalloc(&L.stats, (size_t)pairs * 16), "hipMalloc plan");
            if (!rc) {
                std::vector<int> rank(n);
                spiral_rank(L.nI, L.nJ, rank.data());
                rc = check_hip(hipMemcpy(L.rank, rank.data(), n * 4, hipMemcpyHostToDevice), "upload grid");
            }
        }
        if (!rc)
            rc = check_hip(hipMemcpy(L.top, L.top_h.data(), n * 4, hipMemcpyHostToDevice), "upload grid");
        if (!rc)
            rc = check_hip(hipMemcpy(L.left, L.left_h.data(), n * 4, hipMemcpyHostToDevice), "upload grid");
        if (!rc && l > 0 && cfg->inherit_mode != QLP_INHERIT_DIRECT)
            rc = prepare_inherit(cfg->inherit_mode, p->levels[l - 1], L, need);
    }
    const bool f32 = cfg->frame_type != QLP_FRAME_U8;
    if (!rc && ((tiles_route(*cfg) && p->tiles_bytes) || f32)) {
        size_t mx = 0;
        for (const Level &L : p->levels) mx = std::max(mx, (size_t)pairs * L.n);
        if (p->tiles_bytes) {
            rc = check_hip(hipMalloc(&p->tA, p->tiles_bytes), "hipMalloc plan");
            if (!rc) rc = check_hip(hipMalloc(&p->tB, p->tiles_bytes), "hipMalloc plan");
        }
        if (!rc && f32) {
            rc = check_hip(hipMalloc(&p->fA, p->ftiles_bytes), "hipMalloc plan");
            if (!rc) rc = check_hip(hipMalloc(&p->fB, p->ftiles_bytes), "hipMalloc plan");
        }
        if (!rc) rc = check_hip(hipMalloc(&p->mA, mx * 4), "hipMalloc plan");
        if (!rc) rc = check_hip(hipMalloc(&p->mB, mx * 4), "hipMalloc plan");
    }
    if (!rc && cfg->frame_type == QLP_FRAME_U16)
        rc = check_hip(hipMalloc(&p->fframes, 2 * sizeof(float) * (size_t)pairs * H * W), "hipMalloc plan");
    if (!rc && cfg->warp_mode == QLP_WARP_FULL && p->fwork_bytes) {
        rc = check_hip(hipMalloc(&p->wimg, (size_t)full_frames(*cfg) * pairs * H * W *
                                               (f32 ? sizeof(float) : 1)),
                       "hipMalloc plan");
        if (!rc) rc = check_hip(hipMalloc(&p->fwork, p->fwork_bytes), "hipMalloc plan");
        if (!rc && f32) rc = check_hip(hipMalloc(&p->fmaxp, sizeof(float) * pairs), "hipMalloc plan");
    }
    // the plan's IA-128 passes read the images and find peaks
    if (!rc && p->big_ias) rc = prepare_big(ctx, p->big_ias, cfg->norm_mode, !f32, true, SHAPE_AUTO);
    p->bytes = need;
    ctx->hbm_in_use += need;
    // the MultiPeaks workspace at its largest level's size, so that runs do not
    // grow it, when it fits the budget beside the plan (else the segment form)
    if (!rc && p->mp_work && mpwork_fits(ctx, p->mp_work)) rc = ensure_mpwork(ctx, p->mp_work);
    if (rc) {
        plan_free(p);
        delete p;
        return rc;
    }
    *out = p;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_piv_plan_destroy(qlp_piv_plan *p)
try {
    if (!p) return QLP_OK;
    if (p->ctx) (void)hipStreamSynchronize(p->ctx->stream);
    plan_free(p);
    delete p;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_piv_plan_levels(qlp_piv_plan *p, int *n_levels)
try {
    if (!p || !n_levels) return fail(QLP_ERR_ARG, "null argument");
    *n_levels = (int)p->levels.size();
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_piv_plan_level(qlp_piv_plan *p, int level, int *ia, int *nI, int *nJ, int *top, int *left)
try {
    if (!p) return fail(QLP_ERR_ARG, "null plan");
    if (level < 0 || level >= (int)p->levels.size()) return fail(QLP_ERR_ARG, "level out of range");
    const Level &L = p->levels[level];
    if (ia) *ia = L.N;
    if (nI) *nI = L.nI;
    if (nJ) *nJ = L.nJ;
    if (top)
        for (int k = 0; k < L.n; k++) top[k] = L.top_h[k];
    if (left)
        for (int k = 0; k < L.n; k++) left[k] = L.left_h[k];
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

// Float tiles of one level of a float-frame plan (ImageFloat frames, or
// ImageInt16 frames converted to float): A and B clipped (B shifted by the
// rounded displacement with no warp), or the Micro or Mini warp of the order's
// sides, for every pair.
static int float_tiles(qlp_piv_plan *p, const Level &L, size_t l, const float *imgsA,
                       const float *imgsB)
{
    Context *ctx = p->ctx;
    const qlp_piv_config &c = p->cfg;
    const int order = c.warp_order;
    // ImageInt16 rules for 10/12/16-bit frames, ImageFloat ones otherwise
    const bool i16 = c.frame_type == QLP_FRAME_U16;
    const int clamp16 = i16 ? 1 : 0;
    const float wmax = i16 ? (float)((1 << c.frame_bits) - 1) : 65535.0f;
    const bool warp = l > 0 && (c.warp_mode == QLP_WARP_MICRO || c.warp_mode == QLP_WARP_MINI);
    const bool warpA = warp && order != QLP_WARP_ORDER_SECOND;
    const bool warpB = warp && order != QLP_WARP_ORDER_FIRST;
    const int sa = order == QLP_WARP_ORDER_FIRST ? SIDE_A_FIRST : SIDE_A_BOTH;
    const int sb = order == QLP_WARP_ORDER_SECOND ? SIDE_B_SECOND : SIDE_B_BOTH;
    int rc = QLP_OK;
    if (l > 0 && c.warp_mode == QLP_WARP_FULL) {
        // warp whole frames, cut the tiles from them with the source frame's maximum
        // (clipImageMatrix(0, 0, H, W) then copyMatrixRegion, BiLinearImageWarpingStrategy
        // .warpAndClipImage :50-131, MatrixFloat.copyMatrixRegion :437-455)
        const size_t frames = (size_t)p->pairs * p->H * p->W;
        float *wA = (float *)p->wimg, *wB = wA + (order == QLP_WARP_ORDER_BOTH ? frames : 0);
        const float *src[2];
        float *dst[2];
        int sides[2], ns = 0;
        if (order != QLP_WARP_ORDER_SECOND) {
            src[ns] = imgsA;
            dst[ns] = wA;
            sides[ns++] = sa;
        }
        if (order != QLP_WARP_ORDER_FIRST) {
            src[ns] = imgsB;
            dst[ns] = wB;
            sides[ns++] = sb;
        }
        rc = launch_full_warp_f32(ctx, p->pairs, src, p->H, p->W, L.nI, L.nJ, L.N, L.N,
                                  L.top_h.data(), L.left_h.data(), L.u, L.v, p->fwork, dst, sides, ns);
        const size_t tN = (size_t)L.N * L.N, frame = (size_t)p->H * p->W;
        for (int k = 0; k < p->pairs && !rc; k++) {
            const size_t o = (size_t)k * L.n;
            const float *a = (order != QLP_WARP_ORDER_SECOND ? wA : imgsA) + k * frame;
            const float *b = (order != QLP_WARP_ORDER_FIRST ? wB : imgsB) + k * frame;
            rc = launch_clip_f32(ctx, a, p->H, p->W, L.n, L.N, L.top, L.left, nullptr, nullptr,
                                 p->fA + o * tN, p->mA + o, clamp16);
            if (!rc)
                rc = launch_clip_f32(ctx, b, p->H, p->W, L.n, L.N, L.top, L.left, nullptr, nullptr,
                                     p->fB + o * tN, p->mB + o, clamp16);
        }
        if (!rc && order != QLP_WARP_ORDER_SECOND)
            rc = launch_frame_max_spread(ctx, imgsA, p->pairs, p->H, p->W, L.n, p->fmaxp, p->mA);
        if (!rc && order != QLP_WARP_ORDER_FIRST)
            rc = launch_frame_max_spread(ctx, imgsB, p->pairs, p->H, p->W, L.n, p->fmaxp, p->mB);
        return rc;
    }
    if (l > 0 && c.warp_mode == QLP_WARP_MINI) {  // one launch covers every pair
        if (warpA)
            rc = launch_warp_mini_f32(ctx, p->pairs, imgsA, p->H, p->W, L.nI, L.nJ, L.N, L.top,
                                      L.left, L.u, L.v, p->fA, p->mA, sa, wmax);
        if (!rc && warpB)
            rc = launch_warp_mini_f32(ctx, p->pairs, imgsB, p->H, p->W, L.nI, L.nJ, L.N, L.top,
                                      L.left, L.u, L.v, p->fB, p->mB, sb, wmax);
        if (rc) return rc;
    }
    const bool micro = l > 0 && c.warp_mode == QLP_WARP_MICRO;
    const size_t tN = (size_t)L.N * L.N, frame = (size_t)p->H * p->W;
    for (int k = 0; k < p->pairs && !rc; k++) {
        const float *a = imgsA + k * frame, *b = imgsB + k * frame;
        const size_t o = (size_t)k * L.n;
        float *ta = p->fA + o * tN, *tb = p->fB + o * tN, *ma = p->mA + o, *mb = p->mB + o;
        const float *u = L.u + o, *v = L.v + o;
        if (micro && warpA)
            rc = launch_warp_micro_f32(ctx, a, p->H, p->W, L.n, L.N, L.top, L.left, u, v, ta, ma, sa,
                                       wmax);
        else if (!warpA)
            rc = launch_clip_f32(ctx, a, p->H, p->W, L.n, L.N, L.top, L.left, nullptr, nullptr, ta, ma,
                                 clamp16);
        if (rc) break;
        if (micro && warpB)
            rc = launch_warp_micro_f32(ctx, b, p->H, p->W, L.n, L.N, L.top, L.left, u, v, tb, mb, sb,
                                       wmax);
        else if (!warpB && l > 0 && c.warp_mode == QLP_WARP_NONE)
            rc = launch_clip_f32(ctx, b, p->H, p->W, L.n, L.N, L.top, L.left, u, v, tb, mb, clamp16);
        else if (!warpB)
            rc = launch_clip_f32(ctx, b, p->H, p->W, L.n, L.N, L.top, L.left, nullptr, nullptr, tb, mb,
                                 clamp16);
    }
    return rc;
}

static int run_levels(qlp_piv_plan *p, const uint8_t *imgsA, const uint8_t *imgsB,
                      const float *fimgsA, const float *fimgsB);

int qlp_piv_run_dev(qlp_piv_plan *p, const uint8_t *imgsA, const uint8_t *imgsB)
try {
    if (!p || !imgsA || !imgsB) return fail(QLP_ERR_ARG, "piv_run: null argument");
    if (p->cfg.frame_type != QLP_FRAME_U8)
        return fail(QLP_ERR_ARG, "piv_run: the plan was not made for 8-bit frames");
    return run_levels(p, imgsA, imgsB, nullptr, nullptr);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_piv_run_f32_dev(qlp_piv_plan *p, const float *imgsA, const float *imgsB)
try {
    if (!p || !imgsA || !imgsB) return fail(QLP_ERR_ARG, "piv_run: null argument");
    if (p->cfg.frame_type != QLP_FRAME_F32)
        return fail(QLP_ERR_ARG, "piv_run: the plan was not made for float frames");
    return run_levels(p, nullptr, nullptr, imgsA, imgsB);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_piv_run_u16_dev(qlp_piv_plan *p, const uint16_t *imgsA, const uint16_t *imgsB)
try {
    if (!p || !imgsA || !imgsB) return fail(QLP_ERR_ARG, "piv_run: null argument");
    if (p->cfg.frame_type != QLP_FRAME_U16)
        return fail(QLP_ERR_ARG, "piv_run: the plan was not made for 10/12/16-bit frames");
    const size_t n = (size_t)p->pairs * p->H * p->W;
    int rc = launch_u16_to_f32(p->ctx, imgsA, p->fframes, n);
    if (!rc) rc = launch_u16_to_f32(p->ctx, imgsB, p->fframes + n, n);
    if (rc) return rc;
    return run_levels(p, nullptr, nullptr, p->fframes, p->fframes + n);
} catch (...) {
    return qlpiv::fail_exception();
}

static int run_levels(qlp_piv_plan *p, const uint8_t *imgsA, const uint8_t *imgsB,
                      const float *fimgsA, const float *fimgsB)
{
    Context *ctx = p->ctx;
    const qlp_piv_config &c = p->cfg;
    const bool multi = c.peak_kind == QLP_PEAK_MULTI;
    const bool separate = c.peak_kind != QLP_PEAK_FROM_CENTER;
    for (size_t l = 0; l < p->levels.size(); l++) {
        Level &L = p->levels[l];
        const int pn = p->pairs * L.n;
        const int D = 2 * L.N - 1;
        if (p->timing) (void)hipEventRecord(p->ev[4 * l], ctx->stream);
        if (l == 0) {
            const size_t blocks = ((size_t)pn + 255) / 256;
            hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, ctx->stream, L.u, 0.0f, (size_t)pn);
            hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, ctx->stream, L.v, 0.0f, (size_t)pn);
        }
        int rc = check_hip(hipGetLastError(), "state kernel launch");
        if (rc) return rc;
        if (l > 0) {
            const Level &P = p->levels[l - 1];
            const int round_it = c.warp_mode == QLP_WARP_NONE ? 1 : 0;
            if (c.inherit_mode == QLP_INHERIT_DIRECT)
                rc = launch_inherit_direct(ctx, p->pairs, L.nI, L.nJ, P.nI, P.nJ, round_it, P.u,
                                           P.v, L.u, L.v);
            else if (c.inherit_mode == QLP_INHERIT_BICUBIC)
                rc = launch_inherit_bicubic(ctx, p->pairs, P.nI, P.nJ, L.n, L.bys, L.bxs, L.bcy,
                                            L.bcx, round_it, P.u, P.v, L.u, L.v);
            else
                rc = launch_inherit_stencil(ctx, p->pairs, L.n, P.n, L.K, L.sidx, L.sw, round_it,
                                            P.u, P.v, L.u, L.v);
            if (rc) return rc;
        }
        PassParams q;
        q.n = pn;
        q.N = L.N;
        q.mode = c.norm_mode;
        q.imgA = imgsA;
        q.imgB = imgsB;
        q.H = p->H;
        q.W = p->W;
        q.per_pair = L.n;
        q.top = L.top;
        q.left = L.left;
        q.du = (l == 0) ? nullptr : L.u;
        q.dv = (l == 0) ? nullptr : L.v;
        q.bmode = c.warp_mode == QLP_WARP_MICRO ? B_WARP_MICRO : B_CLIP_SHIFT;
        const int order = c.warp_order;
        if (fimgsA) {
            rc = float_tiles(p, L, l, fimgsA, fimgsB);
            if (rc) return rc;
            q.imgA = q.imgB = nullptr;
            q.tAf = p->fA;
            q.tBf = p->fB;
            q.maxA = p->mA;
            q.maxB = p->mB;
            q.du = q.dv = nullptr;
        } else if (l > 0 && c.warp_mode == QLP_WARP_FULL) {
            // warp the whole frames of the order (A, B or both) with the grid's field
            const size_t frames = (size_t)p->pairs * p->H * p->W;
            uint8_t *wA = p->wimg, *wB = p->wimg + (order == QLP_WARP_ORDER_BOTH ? frames : 0);
            const uint8_t *src[2];
            uint8_t *dst[2];
            int sides[2], ns = 0;
            if (order != QLP_WARP_ORDER_SECOND) {
                src[ns] = imgsA;
                dst[ns] = wA;
                sides[ns++] = order == QLP_WARP_ORDER_FIRST ? SIDE_A_FIRST : SIDE_A_BOTH;
            }
            if (order != QLP_WARP_ORDER_FIRST) {
                src[ns] = imgsB;
                dst[ns] = wB;
                sides[ns++] = order == QLP_WARP_ORDER_SECOND ? SIDE_B_SECOND : SIDE_B_BOTH;
            }
            rc = launch_full_warp(ctx, p->pairs, src, p->H, p->W, L.nI, L.nJ, L.N, L.N,
                                  L.top_h.data(), L.left_h.data(), L.u, L.v, p->fwork, dst, sides, ns);
            if (rc) return rc;
            q.du = q.dv = nullptr;
            q.bmode = B_CLIP;
            if (order == QLP_WARP_ORDER_SECOND) {
                // the pass cuts B from the warped frames at (top, left), maximum 255
                q.imgB = wB;
                q.bmax255 = true;
            } else {
                // tiles cut from the warped frames carry the maximum 255; a B tile
                // clipped from the plain frame (FirstImage) its own
                const size_t tN = (size_t)L.N * L.N;
                const uint8_t *cutA = wA, *cutB = order == QLP_WARP_ORDER_BOTH ? wB : imgsB;
                for (int k = 0; k < p->pairs && !rc; k++) {
                    rc = launch_clip(ctx, cutA + (size_t)k * p->H * p->W, p->H, p->W, L.n, L.N, L.top,
                                     L.left, p->tA + (size_t)k * L.n * tN, p->mA + (size_t)k * L.n);
                    if (!rc)
                        rc = launch_clip(ctx, cutB + (size_t)k * p->H * p->W, p->H, p->W, L.n, L.N,
                                         L.top, L.left, p->tB + (size_t)k * L.n * tN,
                                         p->mB + (size_t)k * L.n);
                }
                if (rc) return rc;
                const unsigned blocks = (unsigned)((pn + 255) / 256);
                hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, ctx->stream, p->mA,
                                   255.0f, (size_t)pn);
                if (order == QLP_WARP_ORDER_BOTH)
                    hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, ctx->stream, p->mB,
                                       255.0f, (size_t)pn);
                q.imgA = q.imgB = nullptr;
                q.tA = p->tA;
                q.tB = p->tB;
                q.maxA = p->mA;
                q.maxB = p->mB;
            }
        } else if (l > 0 && tiles_route(c)) {
            // Micro / Mini warps of the order into tiles, then correlate from the tiles
            rc = warp_clip_tiles(ctx, c.warp_mode, order, imgsA, imgsB, p->H, p->W, p->pairs, L.nI,
                                 L.nJ, L.N, L.top, L.left, L.u, L.v, p->tA, p->mA, p->tB, p->mB);
            if (rc) return rc;
            q.imgA = q.imgB = nullptr;
            q.tA = p->tA;
            q.tB = p->tB;
            q.maxA = p->mA;
            q.maxB = p->mB;
            q.du = q.dv = nullptr;
        }
        q.planes = L.planes;
        q.arith = c.xcorr_form;
        if (!separate) {  // FromCenter and the fit neighbourhood fused into the pass
            q.peaks = L.peaks;
            q.min_floor = L.mf;
            q.nbr = L.nbr;
            q.nbr_half = NBR_HALF;
        }
        // MultiPeaks: the exact LDS pass writes each plane's block maxima (the
        // entries) and only their replay runs after it
        uint2 *ment = nullptr;
        float *mcm = nullptr;
        const size_t mp_need = multipeak_work_bytes((int)pn, (int)D, p->P, c.peak_kernel_width);
        const bool entries = multi && multipeak_entries_form((int)D, p->P, c.peak_kernel_width) &&
                             pass_writes_entries(ctx, q) && mpwork_fits(ctx, mp_need);
        if (entries) {
            rc = ensure_mpwork(ctx, mp_need);
            if (rc) return rc;
            multipeak_work_carve(ctx->mpwork, (int)pn, (int)D, ment, mcm);
            q.mp_ent = ment;
            q.mp_cmax = mcm;
            q.mp_min = L.mf;
        }
        rc = launch_pass(ctx, q);
        if (rc) return rc;
        if (p->timing) (void)hipEventRecord(p->ev[4 * l + 1], ctx->stream);
        const int kind = ((int)l >= c.interpolator_start_step) ? c.subpix_kind : -1;
        if (separate) {
            if (entries)
                rc = launch_multipeak_replay(ctx, (int)pn, (int)D, p->P, ment, mcm, L.peaks, L.count);
            else
                rc = launch_find_peaks(ctx, c.peak_kind, pn, D, L.planes, p->P, c.peak_kernel_width,
                                       L.peaks, L.mf, multi ? L.count : nullptr);
            // Guo applies the displacement update itself; the other fits after it
            const bool guo = kind == QLP_SUBPIX_GUO;
            if (!rc)
                rc = launch_subpixel(ctx, kind, c.subpix_pixels, c.subpix_iterations, pn, D, p->P,
                                     L.planes, L.mf, L.peaks, L.sub, guo ? L.u : nullptr,
                                     guo ? L.v : nullptr);
            if (!rc && !guo) rc = launch_accumulate(ctx, pn, p->P, D, L.sub, L.u, L.v);
        } else {
            rc = launch_subpixel_nbr(ctx, kind, c.subpix_pixels, c.subpix_iterations, pn, D,
                                     NBR_HALF, L.nbr, L.mf, L.peaks, L.sub, L.u, L.v);
        }
        if (rc) return rc;
        if (p->timing) (void)hipEventRecord(p->ev[4 * l + 2], ctx->stream);
        if (c.valid.n_stages > 0) {
            ValidLaunch vl;
            vl.pairs = p->pairs;
            vl.nI = L.nI;
            vl.nJ = L.nJ;
            vl.P = p->P;
            vl.top = L.top;
            vl.left = L.left;
            vl.rank = L.rank;
            vl.u = L.u;
            vl.v = L.v;
            vl.pk = L.sub;  // (i', j', value, status) per peak: u_k = i' - D/2
            vl.pk_stride = 4;
            vl.pk_off = (float)(D / 2);
            vl.pcount = L.count;
            vl.invalid = L.invalid;
            vl.stats = L.stats;
            valid_work_carve(L.vwork, p->pairs, L.n, vl);
            vl.counts_zero = true;
            rc = launch_validate(ctx, c.valid, vl);
            if (rc) return rc;
        }
        if (p->timing) (void)hipEventRecord(p->ev[4 * l + 3], ctx->stream);
    }
    return QLP_OK;
}

int qlp_piv_plan_set_timing(qlp_piv_plan *p, int enable)
try {
    if (!p) return fail(QLP_ERR_ARG, "null plan");
    if (enable && p->ev.empty()) {
        p->ev.resize(4 * p->levels.size());
        for (auto &e : p->ev) {
            int rc = check_hip(hipEventCreate(&e), "hipEventCreate");
            if (rc) return rc;
        }
    }
    p->timing = enable != 0;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_piv_plan_timing(qlp_piv_plan *p, int level, float *state_and_pass_ms, float *fit_ms,
                        float *valid_ms)
try {
    if (!p) return fail(QLP_ERR_ARG, "null plan");
    if (!p->timing) return fail(QLP_ERR_ARG, "timing is not enabled on this plan");
    if (level < 0 || level >= (int)p->levels.size()) return fail(QLP_ERR_ARG, "level out of range");
    hipEvent_t *e = &p->ev[4 * level];
    int rc = check_hip(hipEventSynchronize(e[3]), "hipEventSynchronize");
    if (rc) return rc;
    float a = 0.0f, b = 0.0f, c = 0.0f;
    rc = check_hip(hipEventElapsedTime(&a, e[0], e[1]), "hipEventElapsedTime");
    if (!rc) rc = check_hip(hipEventElapsedTime(&b, e[1], e[2]), "hipEventElapsedTime");
    if (!rc) rc = check_hip(hipEventElapsedTime(&c, e[2], e[3]), "hipEventElapsedTime");
    if (rc) return rc;
    if (state_and_pass_ms) *state_and_pass_ms = a;
    if (fit_ms) *fit_ms = b;
    if (valid_ms) *valid_ms = c;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_piv_result_dev(qlp_piv_plan *p, int level, const float **u, const float **v,
                       const float **peaks, const float **planes)
try {
    if (!p) return fail(QLP_ERR_ARG, "null plan");
    if (level < 0 || level >= (int)p->levels.size()) return fail(QLP_ERR_ARG, "level out of range");
    const Level &L = p->levels[level];
    if (u) *u = L.u;
    if (v) *v = L.v;
    if (peaks) *peaks = L.sub;
    if (planes) *planes = L.planes;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_piv_validation_dev(qlp_piv_plan *p, int level, const uint8_t **invalid, const int **stats)
try {
    if (!p) return fail(QLP_ERR_ARG, "null plan");
    if (level < 0 || level >= (int)p->levels.size()) return fail(QLP_ERR_ARG, "level out of range");
    const Level &L = p->levels[level];
    if (invalid) *invalid = L.invalid;
    if (stats) *stats = L.stats;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

}  // extern "C"
