// Batches of frame pairs of mixed frame and IA sizes, run by the step engine.
//
// Every pair runs the adaptive levels of the batch configuration (Direct
// inheritance, the micro warp or the rounded shift of B, FromCenter, the
// sub-pixel fit from the interpolator start step on) from its own first IA
// size down to its last; step k of a run holds level k of every pair that has
// one, all IA sizes at once, as one launch of the engine's graph (persistent
// kernels reading the step table, qlpiv_engine.hip).  A run of L levels is L
// graph launches; the frames of a run may live anywhere on the device.  Per
// IA the results equal those of a multi-pass plan of that pair alone with the
// same configuration (qlp_piv_run_dev).
//
// Autotuner: on the first run the engine times every kernel shape of each IA
// class on the class's real IAs, keeps the fastest (cached in the context by
// IA size, count, form and normalisation) and times the graph with the
// classes in sequence and forked onto one stream each, keeping the faster.
#include <string.h>

#include <algorithm>
#include <vector>

#include "qlpiv_internal.h"

using namespace qlpiv;

namespace {

struct BatchLevel {
    int N = 0, nI = 0, nJ = 0;
    int *top = nullptr, *left = nullptr;  // device, nI * nJ
    float *u = nullptr, *v = nullptr;     // device, nI * nJ
    float *sub = nullptr;                 // device, nI * nJ x 4 (last level only)
};

}  // namespace

struct qlp_batch {
    Context *ctx = nullptr;
    qlp_piv_config cfg{};
    int npairs = 0;
    std::vector<qlp_batch_pair> geo;
    std::vector<std::vector<BatchLevel>> lv;  // per pair, per level
    int max_levels = 0;
    size_t total = 0;  // IAs over every pair and level
    Engine *eng = nullptr;
    std::vector<void *> bufs;
    size_t bytes = 0;
};

namespace {

void batch_free(qlp_batch *b)
{
    if (b->eng) engine_destroy(b->eng);
    b->eng = nullptr;
    for (void *p : b->bufs) (void)hipFree(p);
    b->bufs.clear();
    if (b->ctx) b->ctx->hbm_in_use -= b->bytes;
    b->bytes = 0;
}

template <typename T>
int balloc(qlp_batch *b, T *&p, size_t count)
{
    void *q = nullptr;
    const size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 16;
    int rc = check_hip(hipMalloc(&q, bytes), "hipMalloc batch");
    if (rc) return rc;
    b->bufs.push_back(q);
    b->bytes += bytes;
    p = (T *)q;
    return QLP_OK;
}

// The configurations the engine runs: the per-frame loop with FromCenter, Direct
// inheritance, the micro warp (SecondImage) or no warp, no validation, 8-bit
// frames.
int check_engine_config(const qlp_piv_config &c, const char *who)
{
    const std::string w(who);
    if (c.overlap_factor < 0.1f || c.overlap_factor > 0.9f)
        return fail(QLP_ERR_ARG, "Overlap factor must greater than 1/10 and less than 9/10");
    if (c.norm_mode != QLP_NORM_PER_TILE && c.norm_mode != QLP_NORM_A_MAX)
        return fail(QLP_ERR_ARG, w + ": unknown normalisation mode");
    if (c.xcorr_form < QLP_XCORR_EXACT || c.xcorr_form > QLP_XCORR_MFMA_BF16_SINGLE)
        return fail(QLP_ERR_ARG, w + ": unknown correlation form");
    if (c.subpix_kind >= 0 && c.subpix_kind != QLP_SUBPIX_POLY &&
        (c.subpix_pixels < 3 || c.subpix_pixels % 2 == 0 || c.subpix_pixels / 2 > 7))
        return fail(QLP_ERR_ARG, w + ": sub-pixel window must be odd and <= 15");
    if (c.subpix_kind > QLP_SUBPIX_POLY) return fail(QLP_ERR_ARG, w + ": unknown interpolator");
    if (c.warp_mode != QLP_WARP_NONE && c.warp_mode != QLP_WARP_MICRO)
        return fail(QLP_ERR_UNSUPPORTED, w + ": batches take no warp or the micro warp");
    if (c.warp_mode == QLP_WARP_MICRO && c.warp_order != QLP_WARP_ORDER_SECOND)
        return fail(QLP_ERR_UNSUPPORTED, w + ": batches warp the second image only");
    if (c.peak_kind != QLP_PEAK_FROM_CENTER || c.n_peaks > 1)
        return fail(QLP_ERR_UNSUPPORTED, w + ": batches find the FromCenter peak");
    if (c.valid.n_stages != 0) return fail(QLP_ERR_UNSUPPORTED, w + ": batches run no validation");
    if (c.inherit_mode != QLP_INHERIT_DIRECT)
        return fail(QLP_ERR_UNSUPPORTED, w + ": batches take the Direct inheritance");
    if (c.frame_type != QLP_FRAME_U8) return fail(QLP_ERR_UNSUPPORTED, w + ": batches take 8-bit frames");
    return QLP_OK;
}

}  // namespace

extern "C" {

int qlp_batch_create(qlp_ctx *h, int n_pairs, const qlp_batch_pair *pairs,
                     const qlp_piv_config *cfg, qlp_batch **out)
try {
    if (!h || !cfg || !out || (n_pairs > 0 && !pairs)) return fail(QLP_ERR_ARG, "batch_create: null argument");
    *out = nullptr;
    if (n_pairs < 1) return fail(QLP_ERR_ARG, "batch_create: no pairs");
    int rc = check_engine_config(*cfg, "batch_create");
    if (rc) return rc;
    Context *ctx = ctx_of(h);
    const int m[4] = {cfg->margin_top, cfg->margin_bottom, cfg->margin_left, cfg->margin_right};
    // geometry of every pair level, and the IAs of each class at each step
    std::vector<std::vector<std::vector<int>>> tops(n_pairs), lefts(n_pairs);
    std::vector<std::vector<BatchLevel>> lv(n_pairs);
    std::vector<std::vector<size_t>> step_count;  // [step][class]
    size_t total = 0, state = 0;
    int max_levels = 0;
    for (int p = 0; p < n_pairs; p++) {
        const qlp_batch_pair &g = pairs[p];
        const int L = g.levels > 0 ? g.levels : 1;
        if (g.H < 1 || g.W < 1 || eng_class_of(g.ia) < 0 || L > 4 || eng_class_of(g.ia << (L - 1)) < 0)
            return fail(QLP_ERR_ARG, "batch_create: pair with a bad frame size, IA size (16, 32, 64, "
                                     "128) or level count");
        max_levels = std::max(max_levels, L);
        if ((int)step_count.size() < L) step_count.resize(L, std::vector<size_t>(ENG_CLASSES, 0));
        for (int l = 0; l < L; l++) {
            BatchLevel B;
            B.N = g.ia << (L - 1 - l);
            std::vector<int> top, left;
            level_origins(g.H, g.W, B.N, cfg->overlap_factor, m, l == 0, B.nI, B.nJ, top, left);
            if (B.nI < 1 || B.nJ < 1) return fail(QLP_ERR_ARG, "batch_create: no IA fits a pair's frame");
            const size_t n = (size_t)B.nI * B.nJ;
            step_count[l][eng_class_of(B.N)] += n;
            total += n;
            state += n * 16 + (l == L - 1 ? n * 16 : 0);
            tops[p].push_back(std::move(top));
            lefts[p].push_back(std::move(left));
            lv[p].push_back(B);
        }
    }
    size_t cap[ENG_CLASSES] = {0, 0, 0, 0};
    for (const auto &sc : step_count)
        for (int c = 0; c < ENG_CLASSES; c++) cap[c] = std::max(cap[c], sc[c]);
    EngineConfig ec{cfg->norm_mode, cfg->subpix_kind, cfg->subpix_pixels, cfg->subpix_iterations,
                    cfg->xcorr_form};
    if (ctx->hbm_budget && ctx->hbm_in_use + engine_bytes(ec, cap) + state > ctx->hbm_budget)
        return fail(QLP_ERR_BUDGET, "batch exceeds the context's HBM budget");
    rc = check_hip(hipSetDevice(ctx->device), "hipSetDevice");
    if (rc) return rc;
    qlp_batch *b = new qlp_batch();
    b->ctx = ctx;
    b->cfg = *cfg;
    b->npairs = n_pairs;
    b->geo.assign(pairs, pairs + n_pairs);
    b->max_levels = max_levels;
    b->total = total;
    for (int p = 0; p < n_pairs && !rc; p++) {
        const int L = (int)lv[p].size();
        for (int l = 0; l < L && !rc; l++) {
            BatchLevel &B = lv[p][l];
            const size_t n = (size_t)B.nI * B.nJ;
            rc = balloc(b, B.top, n);
            if (!rc) rc = balloc(b, B.left, n);
            if (!rc) rc = balloc(b, B.u, n);
            if (!rc) rc = balloc(b, B.v, n);
            if (!rc && l == L - 1) rc = balloc(b, B.sub, n * 4);
            if (!rc) rc = check_hip(hipMemcpy(B.top, tops[p][l].data(), n * 4, hipMemcpyHostToDevice), "upload batch");
            if (!rc) rc = check_hip(hipMemcpy(B.left, lefts[p][l].data(), n * 4, hipMemcpyHostToDevice), "upload batch");
        }
    }
    b->lv = std::move(lv);
    ctx->hbm_in_use += b->bytes;
    if (!rc) rc = engine_create(ctx, ec, cap, &b->eng);
    if (rc) {
        batch_free(b);
        delete b;
        return rc;
    }
    *out = b;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_batch_fix(qlp_batch *b, int ia, int shape, int schedule)
try {
    if (!b) return fail(QLP_ERR_ARG, "null batch");
    return engine_fix(b->eng, ia, shape, schedule);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_batch_run_dev(qlp_batch *b, const uint8_t *const *imgsA, const uint8_t *const *imgsB)
try {
    if (!b || !imgsA || !imgsB) return fail(QLP_ERR_ARG, "batch_run: null argument");
    for (int p = 0; p < b->npairs; p++)
        if (!imgsA[p] || !imgsB[p]) return fail(QLP_ERR_ARG, "batch_run: null frame");
    int rc = check_hip(hipSetDevice(b->ctx->device), "hipSetDevice");
    if (rc) return rc;
    const int warp = b->cfg.warp_mode == QLP_WARP_MICRO ? STEP_MICRO : STEP_SHIFT;
    std::vector<StepSeg> segs[ENG_CLASSES];
    for (int s = 0; s < b->max_levels && !rc; s++) {
        for (auto &v : segs) v.clear();
        for (int p = 0; p < b->npairs; p++) {
            const std::vector<BatchLevel> &L = b->lv[p];
            if (s >= (int)L.size()) continue;
            const BatchLevel &B = L[s];
            StepSeg g{};
            g.a = imgsA[p];
            g.b = imgsB[p];
            g.top = B.top;
            g.left = B.left;
            if (s > 0) {
                g.pu = L[s - 1].u;
                g.pv = L[s - 1].v;
                g.pnI = L[s - 1].nI;
                g.pnJ = L[s - 1].nJ;
            }
            g.u = B.u;
            g.v = B.v;
            g.sub = B.sub;
            g.H = b->geo[p].H;
            g.W = b->geo[p].W;
            g.nI = B.nI;
            g.nJ = B.nJ;
            g.warp = warp;
            g.fit = s >= b->cfg.interpolator_start_step ? 1 : 0;
            segs[eng_class_of(B.N)].push_back(g);
        }
        rc = engine_step(b->eng, segs);
    }
    return rc;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_batch_grid(qlp_batch *b, int pair, int *nI, int *nJ, size_t *total_ias)
try {
    if (!b) return fail(QLP_ERR_ARG, "null batch");
    if (pair < 0 || pair >= b->npairs) return fail(QLP_ERR_ARG, "batch_grid: pair index out of range");
    const BatchLevel &B = b->lv[pair].back();
    if (nI) *nI = B.nI;
    if (nJ) *nJ = B.nJ;
    if (total_ias) *total_ias = b->total;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_batch_result_dev(qlp_batch *b, int pair, const float **u, const float **v,
                         const float **sub)
try {
    if (!b) return fail(QLP_ERR_ARG, "null batch");
    if (pair < 0 || pair >= b->npairs) return fail(QLP_ERR_ARG, "batch_result: pair index out of range");
    const BatchLevel &B = b->lv[pair].back();
    if (u) *u = B.u;
    if (v) *v = B.v;
    if (sub) *sub = B.sub;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_batch_report_get(qlp_batch *b, qlp_batch_report *report)
try {
    if (!b || !report) return fail(QLP_ERR_ARG, "batch_report: null argument");
    return engine_report(b->eng, report);
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_batch_destroy(qlp_batch *b)
try {
    if (!b) return QLP_OK;
    (void)hipStreamSynchronize(b->ctx->stream);
    batch_free(b);
    delete b;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

}  // extern "C"
