// Batches of frame pairs of mixed frame and IA sizes, run as one HIP graph.
//
// Every pair contributes the IAs of its first-level grid to the class of its
// IA size (16, 32, 64, 128).  Per class, a fixed chain of kernels:
//   clip (A and B windows of every IA from its own pair's frames, through a
//   device table of frame pointers, with the >= 16 clamped maxima) ->
//   zero the displacements -> correlation pass with the fused FromCenter peak
//   and fit neighbourhood -> sub-pixel fit (displacement = fit - centre),
// the single-level frame job of jobs/LocalPIVOpenCLGpuJob.call :188-409 for
// many frames at once.  The frame pointers of a run are copied into the table
// ahead of the graph launch, so one instantiated graph serves any frames.
//
// Autotuner: the correlation of a class can run with several kernel shapes
// (PassShape); the exact shapes compute the same values.  The first run times
// each candidate on the class's real IAs (best of 3 after a warm-up, HIP
// events), keeps the fastest, and caches the choice in the context.  It then
// captures the classes twice, in sequence on one stream and forked onto one
// stream per class, times both graphs and keeps the faster.
#include <string.h>

#include <algorithm>
#include <vector>

#include "qlpiv_internal.h"

namespace qlpiv {

struct FramePair {
    const uint8_t *a, *b;
    int H, W;
};

// One work-group per IA: the two N x N windows of the IA's pair at (top, left),
// zero outside the frame, and max(tile) clamped to >= 16 for A and B
// (Image.clipImageMatrix :386-494, MatrixByte :16-39) -- clip_kernel's values
// for IAs spread over many frames.
__global__ __launch_bounds__(256) void batch_clip_kernel(
    const FramePair *__restrict__ frames, const int *__restrict__ pair_of,
    const int *__restrict__ top, const int *__restrict__ left, int n, int N,
    uint8_t *__restrict__ tA, uint8_t *__restrict__ tB, float *__restrict__ mA,
    float *__restrict__ mB)
{
    __shared__ int s_max[2][4];
    const int t = blockIdx.x;
    if (t >= n) return;
    const FramePair f = frames[pair_of[t]];
    const int y0 = top[t], x0 = left[t];
    int ma = 0, mb = 0;
    for (int idx = threadIdx.x; idx < N * N; idx += 256) {
        const int i = idx / N, j = idx - (idx / N) * N;
        const int y = y0 + i, x = x0 + j;
        uint8_t va = 0, vb = 0;
        if (y >= 0 && y < f.H && x >= 0 && x < f.W) {
            const size_t o = (size_t)y * f.W + x;
            va = f.a[o];
            vb = f.b[o];
            ma = max(ma, (int)va);
            mb = max(mb, (int)vb);
        }
        tA[(size_t)t * N * N + idx] = va;
        tB[(size_t)t * N * N + idx] = vb;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        ma = max(ma, __shfl_xor(ma, o, 64));
        mb = max(mb, __shfl_xor(mb, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        s_max[0][threadIdx.x >> 6] = ma;
        s_max[1][threadIdx.x >> 6] = mb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) {
            ma = max(ma, s_max[0][w]);
            mb = max(mb, s_max[1][w]);
        }
        mA[t] = (float)(ma < 16 ? 16 : ma);
        mB[t] = (float)(mb < 16 ? 16 : mb);
    }
}

// All IAs of one IA size.
struct BatchClass {
    int N = 0, n = 0;
    int shape = SHAPE_AUTO;
    bool fixed = false;  // shape set by qlp_batch_fix
    int n_cand = 0, cand_shape[4] = {};
    float cand_ms[4] = {};
    float cost_ms = 0.0f;  // measured time of the chosen shape (schedule order)
    std::vector<int> pair_h, top_h, left_h;
    int *pair_of = nullptr, *top = nullptr, *left = nullptr;
    uint8_t *tA = nullptr, *tB = nullptr;
    float *mA = nullptr, *mB = nullptr, *peaks = nullptr, *mf = nullptr, *nbr = nullptr;
    float *sub = nullptr, *u = nullptr, *v = nullptr;
};

struct PairSlot {
    int cls = 0, first = 0, nI = 0, nJ = 0;
};

static int subpix_half(const qlp_piv_config &c)
{
    if (c.subpix_kind == QLP_SUBPIX_POLY) return 1;
    return c.subpix_kind >= 0 && c.subpix_pixels / 2 > 1 ? c.subpix_pixels / 2 : 1;
}

// Candidate shapes of a class (the first is the library default).
static int shape_candidates(int N, int arith, int *out)
{
    if (arith != QLP_XCORR_EXACT && N <= 64) {
        out[0] = SHAPE_AUTO;  // the matrix-core kernel has one shape
        return 1;
    }
    if (N == 16) {
        out[0] = SHAPE_LANE;
        out[1] = SHAPE_LDS;
        out[2] = SHAPE_LDS_HALF;
        out[3] = SHAPE_LDS_WIDE;
        return 4;
    }
    out[0] = SHAPE_LDS;
    out[1] = SHAPE_LDS_HALF;
    if (N == 128) return 2;
    out[2] = SHAPE_LDS_WIDE;
    return 3;
}

// Key of the context's shape cache: IA size, IA count rounded up to a power of
// two, arithmetic form and normalisation.
static uint64_t tune_key(int N, int n, int arith, int mode)
{
    int lg = 0;
    while ((1 << lg) < n && lg < 40) lg++;
    return ((uint64_t)N << 32) | ((uint64_t)lg << 16) | ((uint64_t)arith << 8) | (uint64_t)mode;
}

}  // namespace qlpiv

using namespace qlpiv;

struct qlp_batch {
    Context *ctx = nullptr;
    qlp_piv_config cfg{};
    int npairs = 0;
    std::vector<qlp_batch_pair> geo;
    std::vector<PairSlot> slot;
    std::vector<BatchClass> cls;
    size_t total = 0;
    // the 128 px class's own tile workspace: a graph must not hold the
    // context's, which a later pass may reallocate
    void *bigwork = nullptr;
    size_t bigwork_bytes = 0;
    FramePair *frames_d = nullptr;  // device table of the run's frame pointers
    FramePair *frames_h = nullptr;  // pinned staging of that table
    hipEvent_t frames_copied = nullptr;
    std::vector<void *> bufs;
    size_t bytes = 0;
    int schedule = QLP_SCHEDULE_TUNED;
    bool schedule_fixed = false;
    float serial_ms = 0.0f, concurrent_ms = 0.0f;
    bool built = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t> ev;
};

static void batch_free(qlp_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (hipStream_t s : b->branch) (void)hipStreamDestroy(s);
    for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
    if (b->frames_copied) (void)hipEventDestroy(b->frames_copied);
    for (void *p : b->bufs) (void)hipFree(p);
    if (b->frames_h) (void)hipHostFree(b->frames_h);
    if (b->ctx) b->ctx->hbm_in_use -= b->bytes;
}

template <typename T>
static int balloc(qlp_batch *b, T *&p, size_t count)
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

// Enqueue the kernels of class c on stream s, with shape 'shape'.
static int enqueue_class(qlp_batch *b, const BatchClass &c, int shape, hipStream_t s)
{
    if (c.n == 0) return QLP_OK;
    Context sub = *b->ctx;  // the same tables, another stream, the batch's workspace
    sub.stream = s;
    sub.bigwork = b->bigwork;
    sub.bigwork_bytes = b->bigwork_bytes;
    const int N = c.N, D = 2 * N - 1;
    hipLaunchKernelGGL(batch_clip_kernel, dim3(c.n), dim3(256), 0, s, b->frames_d, c.pair_of, c.top,
                       c.left, c.n, N, c.tA, c.tB, c.mA, c.mB);
    const unsigned blocks = (unsigned)((c.n + 255) / 256);
    hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, s, c.u, 0.0f, (size_t)c.n);
    hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, s, c.v, 0.0f, (size_t)c.n);
    int rc = check_hip(hipGetLastError(), "batch clip launch");
    if (rc) return rc;
    PassParams q;
    q.n = c.n;
    q.N = N;
    q.mode = b->cfg.norm_mode;
    q.tA = c.tA;
    q.tB = c.tB;
    q.maxA = c.mA;
    q.maxB = c.mB;
    q.peaks = c.peaks;
    q.min_floor = c.mf;
    q.nbr = c.nbr;
    q.nbr_half = subpix_half(b->cfg);
    q.arith = b->cfg.xcorr_form;
    q.shape = shape;
    rc = launch_pass(&sub, q);
    if (rc) return rc;
    const int kind = b->cfg.interpolator_start_step <= 0 ? b->cfg.subpix_kind : -1;
    rc = launch_subpixel_nbr(&sub, kind, b->cfg.subpix_pixels, b->cfg.subpix_iterations, c.n, D,
                             q.nbr_half, c.nbr, c.mf, c.peaks, c.sub, c.u, c.v);
    if (rc) return rc;
    // every buffer was sized before the first enqueue: nothing may have grown
    bool grew = sub.bigwork != b->bigwork;
    for (int a = 0; a < 3; a++)
        for (int f = 0; f < 2; f++) grew = grew || sub.mfma_tab[a][f] != b->ctx->mfma_tab[a][f];
    if (grew) return fail(QLP_ERR_DEVICE, "batch: a workspace grew during an enqueue");
    return QLP_OK;
}

// Upload the tables class c needs (outside any capture).
static int prepare_class(qlp_batch *b, const BatchClass &c)
{
    if (c.n == 0) return QLP_OK;
    if (b->cfg.xcorr_form != QLP_XCORR_EXACT && c.N <= 64) {
        // the MFMA tables are uploaded on the form's first launch: one IA
        Context *ctx = b->ctx;
        PassParams q;
        q.n = 1;
        q.N = c.N;
        q.mode = b->cfg.norm_mode;
        q.tA = c.tA;
        q.tB = c.tB;
        q.maxA = c.mA;
        q.maxB = c.mB;
        q.peaks = c.peaks;
        q.min_floor = c.mf;
        q.nbr = c.nbr;
        q.nbr_half = subpix_half(b->cfg);
        q.arith = b->cfg.xcorr_form;
        hipLaunchKernelGGL(fill_kernel, dim3(1), dim3(256), 0, ctx->stream, c.mA, 16.0f, (size_t)1);
        hipLaunchKernelGGL(fill_kernel, dim3(1), dim3(256), 0, ctx->stream, c.mB, 16.0f, (size_t)1);
        int rc = launch_pass(ctx, q);
        if (rc) return rc;
        rc = check_hip(hipStreamSynchronize(ctx->stream), "batch prepare");
        if (rc) return rc;
    }
    return QLP_OK;
}

// Best of 3 timed enqueues (after one warm-up) of class c with 'shape'.
static int time_class(qlp_batch *b, const BatchClass &c, int shape, float &ms)
{
    hipStream_t s = b->ctx->stream;
    hipEvent_t e0 = b->ev[0], e1 = b->ev[1];
    int rc = enqueue_class(b, c, shape, s);
    ms = 1e30f;
    for (int r = 0; r < 3 && !rc; r++) {
        (void)hipEventRecord(e0, s);
        rc = enqueue_class(b, c, shape, s);
        (void)hipEventRecord(e1, s);
        if (!rc) rc = check_hip(hipEventSynchronize(e1), "batch tune");
        float t = 0.0f;
        if (!rc) rc = check_hip(hipEventElapsedTime(&t, e0, e1), "batch tune");
        if (!rc) ms = std::min(ms, t);
    }
    return rc;
}

static int capture(qlp_batch *b, bool concurrent, hipGraph_t &graph, hipGraphExec_t &exec)
{
    hipStream_t s = b->ctx->stream;
    graph = nullptr;
    exec = nullptr;
    std::vector<int> order;
    for (int k = 0; k < (int)b->cls.size(); k++)
        if (b->cls[k].n) order.push_back(k);
    // longest class first, so that the short ones fill in behind it
    std::sort(order.begin(), order.end(),
              [&](int x, int y) { return b->cls[x].cost_ms > b->cls[y].cost_ms; });
    int rc = check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "batch capture");
    if (rc) return rc;
    hipEvent_t fork = b->ev[2];
    if (concurrent) rc = check_hip(hipEventRecord(fork, s), "batch fork");
    for (size_t i = 0; i < order.size() && !rc; i++) {
        const BatchClass &c = b->cls[order[i]];
        if (!concurrent) {
            rc = enqueue_class(b, c, c.shape, s);
            continue;
        }
        hipStream_t br = b->branch[i];
        rc = check_hip(hipStreamWaitEvent(br, fork, 0), "batch fork");
        if (!rc) rc = enqueue_class(b, c, c.shape, br);
        if (!rc) rc = check_hip(hipEventRecord(b->ev[3 + i], br), "batch join");
        if (!rc) rc = check_hip(hipStreamWaitEvent(s, b->ev[3 + i], 0), "batch join");
    }
    hipGraph_t g = nullptr;
    const int rc_end = check_hip(hipStreamEndCapture(s, &g), "batch capture");
    if (!rc) rc = rc_end;
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    rc = check_hip(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0), "batch graph instantiate");
    if (rc) {
        (void)hipGraphDestroy(g);
        return rc;
    }
    graph = g;
    return QLP_OK;
}

static int time_graph(qlp_batch *b, hipGraphExec_t exec, float &ms)
{
    hipStream_t s = b->ctx->stream;
    int rc = check_hip(hipGraphLaunch(exec, s), "batch graph launch");
    ms = 1e30f;
    for (int r = 0; r < 3 && !rc; r++) {
        (void)hipEventRecord(b->ev[0], s);
        rc = check_hip(hipGraphLaunch(exec, s), "batch graph launch");
        (void)hipEventRecord(b->ev[1], s);
        if (!rc) rc = check_hip(hipEventSynchronize(b->ev[1]), "batch tune");
        float t = 0.0f;
        if (!rc) rc = check_hip(hipEventElapsedTime(&t, b->ev[0], b->ev[1]), "batch tune");
        if (!rc) ms = std::min(ms, t);
    }
    return rc;
}

// First run: pick the shapes and the schedule, build the graph.
static int build(qlp_batch *b)
{
    Context *ctx = b->ctx;
    for (BatchClass &c : b->cls) {
        int rc = prepare_class(b, c);
        if (rc) return rc;
    }
    for (BatchClass &c : b->cls) {
        if (c.n == 0) continue;
        int cand[4];
        const int nc = shape_candidates(c.N, b->cfg.xcorr_form, cand);
        const uint64_t key = tune_key(c.N, c.n, b->cfg.xcorr_form, b->cfg.norm_mode);
        c.n_cand = 0;
        if (!c.fixed) {
            auto hit = ctx->tuned_shape.find(key);
            if (hit != ctx->tuned_shape.end()) {
                c.shape = hit->second;
            } else if (nc == 1) {
                c.shape = cand[0];
            } else {
                float best = 1e30f;
                for (int k = 0; k < nc; k++) {
                    float ms = 0.0f;
                    int rc = time_class(b, c, cand[k], ms);
                    if (rc) return rc;
                    c.cand_shape[c.n_cand] = cand[k];
                    c.cand_ms[c.n_cand++] = ms;
                    if (ms < best) {
                        best = ms;
                        c.shape = cand[k];
                    }
                }
                ctx->tuned_shape[key] = c.shape;
            }
        }
        float ms = 0.0f;
        int rc = time_class(b, c, c.shape, ms);
        if (rc) return rc;
        c.cost_ms = ms;
    }
    hipGraph_t g[2] = {nullptr, nullptr};
    hipGraphExec_t x[2] = {nullptr, nullptr};
    float t[2] = {0.0f, 0.0f};
    int rc = QLP_OK;
    for (int conc = 0; conc < 2 && !rc; conc++) {
        if (b->schedule_fixed && b->schedule != conc) continue;
        rc = capture(b, conc == 1, g[conc], x[conc]);
        if (!rc && !b->schedule_fixed) rc = time_graph(b, x[conc], t[conc]);
    }
    int pick = b->schedule_fixed ? b->schedule : (t[1] < t[0] ? 1 : 0);
    if (!rc) {
        b->serial_ms = b->schedule_fixed ? 0.0f : t[0];
        b->concurrent_ms = b->schedule_fixed ? 0.0f : t[1];
        b->schedule = pick;
        b->graph = g[pick];
        b->exec = x[pick];
        b->built = true;
    }
    for (int k = 0; k < 2; k++) {
        if (!rc && k == pick) continue;
        if (x[k]) (void)hipGraphExecDestroy(x[k]);
        if (g[k]) (void)hipGraphDestroy(g[k]);
    }
    return rc;
}

extern "C" {

int qlp_batch_create(qlp_ctx *h, int n_pairs, const qlp_batch_pair *pairs,
                     const qlp_piv_config *cfg, qlp_batch **out)
try {
    if (!h || !cfg || !out || (n_pairs > 0 && !pairs)) return fail(QLP_ERR_ARG, "batch_create: null argument");
    *out = nullptr;
    if (n_pairs < 1) return fail(QLP_ERR_ARG, "batch_create: no pairs");
    if (cfg->overlap_factor < 0.1f || cfg->overlap_factor > 0.9f)
        return fail(QLP_ERR_ARG, "Overlap factor must greater than 1/10 and less than 9/10");
    if (cfg->norm_mode != QLP_NORM_PER_TILE && cfg->norm_mode != QLP_NORM_A_MAX)
        return fail(QLP_ERR_ARG, "batch_create: unknown normalisation mode");
    if (cfg->xcorr_form < QLP_XCORR_EXACT || cfg->xcorr_form > QLP_XCORR_MFMA_BF16_SINGLE)
        return fail(QLP_ERR_ARG, "batch_create: unknown correlation form");
    if (cfg->subpix_kind >= 0) {
        if (cfg->subpix_kind != QLP_SUBPIX_POLY &&
            (cfg->subpix_pixels % 2 == 0 || cfg->subpix_pixels / 2 > 7))
            return fail(QLP_ERR_ARG, "batch_create: sub-pixel window must be odd and <= 15");
    }
    Context *ctx = ctx_of(h);
    qlp_batch *b = new qlp_batch();
    b->ctx = ctx;
    b->cfg = *cfg;
    b->npairs = n_pairs;
    b->geo.assign(pairs, pairs + n_pairs);
    b->slot.resize(n_pairs);
    const int sizes[4] = {16, 32, 64, 128};
    b->cls.resize(4);
    for (int k = 0; k < 4; k++) b->cls[k].N = sizes[k];
    const int m[4] = {cfg->margin_top, cfg->margin_bottom, cfg->margin_left, cfg->margin_right};
    for (int p = 0; p < n_pairs; p++) {
        const qlp_batch_pair &g = pairs[p];
        int k = -1;
        for (int s = 0; s < 4; s++)
            if (g.ia == sizes[s]) k = s;
        if (k < 0 || g.H < 1 || g.W < 1) {
            delete b;
            return fail(QLP_ERR_ARG, "batch_create: pair with a bad frame or IA size (16, 32, 64, 128)");
        }
        int nI = 0, nJ = 0;
        std::vector<int> top, left;
        grid_origins(g.H, g.W, g.ia, cfg->overlap_factor, m, nI, nJ, top, left);
        if (nI < 1 || nJ < 1) {
            delete b;
            return fail(QLP_ERR_ARG, "batch_create: no IA fits a pair's frame");
        }
        BatchClass &c = b->cls[k];
        b->slot[p] = {k, c.n, nI, nJ};
        for (int t = 0; t < nI * nJ; t++) {
            c.pair_h.push_back(p);
            c.top_h.push_back(top[t]);
            c.left_h.push_back(left[t]);
        }
        c.n += nI * nJ;
        b->total += (size_t)nI * nJ;
    }
    // device state
    size_t need = (size_t)n_pairs * sizeof(FramePair);
    for (const BatchClass &c : b->cls) {
        const size_t n = c.n, L = 2 * subpix_half(*cfg) + 1;
        need += n * (12 + 2 * (size_t)c.N * c.N + 8 + 12 + 4 + 8 * L + 16 + 8);
    }
    if (b->cls[3].n) {
        // the largest workspace any candidate shape of the 128 px class needs
        int cand[4];
        const int nc = shape_candidates(128, cfg->xcorr_form, cand);
        for (int k = 0; k < nc; k++)
            b->bigwork_bytes = std::max(
                b->bigwork_bytes, big_work_need(b->cls[3].n, cfg->norm_mode, false, true, cand[k]));
        need += b->bigwork_bytes;
    }
    if (ctx->hbm_budget && ctx->hbm_in_use + need > ctx->hbm_budget) {
        delete b;
        return fail(QLP_ERR_BUDGET, "batch exceeds the context's HBM budget");
    }
    int rc = check_hip(hipSetDevice(ctx->device), "hipSetDevice");
    if (!rc) rc = balloc(b, b->frames_d, n_pairs);
    if (!rc) rc = check_hip(hipHostMalloc((void **)&b->frames_h, sizeof(FramePair) * n_pairs, 0),
                            "hipHostMalloc batch");
    for (BatchClass &c : b->cls) {
        if (rc || c.n == 0) continue;
        const size_t n = c.n, L = 2 * subpix_half(*cfg) + 1;
        rc = balloc(b, c.pair_of, n);
        if (!rc) rc = balloc(b, c.top, n);
        if (!rc) rc = balloc(b, c.left, n);
        if (!rc) rc = balloc(b, c.tA, n * c.N * c.N);
        if (!rc) rc = balloc(b, c.tB, n * c.N * c.N);
        if (!rc) rc = balloc(b, c.mA, n);
        if (!rc) rc = balloc(b, c.mB, n);
        if (!rc) rc = balloc(b, c.peaks, n * 3);
        if (!rc) rc = balloc(b, c.mf, n);
        if (!rc) rc = balloc(b, c.nbr, n * 2 * L);
        if (!rc) rc = balloc(b, c.sub, n * 4);
        if (!rc) rc = balloc(b, c.u, n);
        if (!rc) rc = balloc(b, c.v, n);
        if (!rc) rc = check_hip(hipMemcpy(c.pair_of, c.pair_h.data(), n * 4, hipMemcpyHostToDevice), "upload batch");
        if (!rc) rc = check_hip(hipMemcpy(c.top, c.top_h.data(), n * 4, hipMemcpyHostToDevice), "upload batch");
        if (!rc) rc = check_hip(hipMemcpy(c.left, c.left_h.data(), n * 4, hipMemcpyHostToDevice), "upload batch");
    }
    if (!rc && b->bigwork_bytes) {
        uint8_t *w = nullptr;
        rc = balloc(b, w, b->bigwork_bytes);
        b->bigwork = w;
    }
    if (!rc) rc = check_hip(hipEventCreateWithFlags(&b->frames_copied, hipEventDisableTiming), "batch event");
    for (int k = 0; k < 3 + 4 && !rc; k++) {
        hipEvent_t e = nullptr;
        rc = check_hip(k < 2 ? hipEventCreate(&e) : hipEventCreateWithFlags(&e, hipEventDisableTiming),
                       "batch event");
        if (!rc) b->ev.push_back(e);
    }
    for (int k = 0; k < 4 && !rc; k++) {
        hipStream_t s = nullptr;
        rc = check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "batch stream");
        if (!rc) b->branch.push_back(s);
    }
    ctx->hbm_in_use += b->bytes;
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
    if (b->built) return fail(QLP_ERR_ARG, "batch_fix: the batch has already run");
    if (shape < QLP_SHAPE_AUTO || shape > QLP_SHAPE_LDS_WIDE)
        return fail(QLP_ERR_ARG, "batch_fix: unknown shape");
    if (schedule < QLP_SCHEDULE_TUNED || schedule > QLP_SCHEDULE_CONCURRENT)
        return fail(QLP_ERR_ARG, "batch_fix: unknown schedule");
    if (ia != 0 && ia != 16 && ia != 32 && ia != 64 && ia != 128)
        return fail(QLP_ERR_ARG, "batch_fix: IA size must be 0, 16, 32, 64 or 128");
    if (shape == QLP_SHAPE_LANE && ia != 16 && ia != 0)
        return fail(QLP_ERR_ARG, "batch_fix: the lane shape takes 16 px IAs");
    if (shape == QLP_SHAPE_LDS_WIDE && ia == 128)
        return fail(QLP_ERR_ARG, "batch_fix: the wide shape takes 16, 32 and 64 px IAs");
    for (BatchClass &c : b->cls) {
        if (ia != 0 && c.N != ia) continue;
        // shapes a class cannot take (lane above 16 px, wide at 128 px) fall back
        // to the class's default when every class is fixed at once
        const bool fits = !(shape == QLP_SHAPE_LANE && c.N != 16) &&
                          !(shape == QLP_SHAPE_LDS_WIDE && c.N == 128);
        c.shape = fits ? shape : SHAPE_AUTO;
        c.fixed = true;
    }
    if (schedule != QLP_SCHEDULE_TUNED) {
        b->schedule = schedule;
        b->schedule_fixed = true;
    }
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_batch_run_dev(qlp_batch *b, const uint8_t *const *imgsA, const uint8_t *const *imgsB)
try {
    if (!b || !imgsA || !imgsB) return fail(QLP_ERR_ARG, "batch_run: null argument");
    for (int p = 0; p < b->npairs; p++)
        if (!imgsA[p] || !imgsB[p]) return fail(QLP_ERR_ARG, "batch_run: null frame");
    Context *ctx = b->ctx;
    // the staging table is rewritten only after the previous run's copy read it
    int rc = check_hip(hipEventSynchronize(b->frames_copied), "batch run");
    if (rc) return rc;
    for (int p = 0; p < b->npairs; p++)
        b->frames_h[p] = {imgsA[p], imgsB[p], b->geo[p].H, b->geo[p].W};
    rc = check_hip(hipMemcpyAsync(b->frames_d, b->frames_h, sizeof(FramePair) * b->npairs,
                                  hipMemcpyHostToDevice, ctx->stream),
                   "batch frame table");
    if (!rc) rc = check_hip(hipEventRecord(b->frames_copied, ctx->stream), "batch run");
    if (rc) return rc;
    if (!b->built) {
        rc = build(b);
        if (rc) return rc;
    }
    return check_hip(hipGraphLaunch(b->exec, ctx->stream), "batch graph launch");
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_batch_grid(qlp_batch *b, int pair, int *nI, int *nJ, size_t *total_ias)
try {
    if (!b) return fail(QLP_ERR_ARG, "null batch");
    if (pair < 0 || pair >= b->npairs) return fail(QLP_ERR_ARG, "batch: pair out of range");
    if (nI) *nI = b->slot[pair].nI;
    if (nJ) *nJ = b->slot[pair].nJ;
    if (total_ias) *total_ias = b->total;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_batch_result_dev(qlp_batch *b, int pair, const float **u, const float **v,
                         const float **sub)
try {
    if (!b) return fail(QLP_ERR_ARG, "null batch");
    if (pair < 0 || pair >= b->npairs) return fail(QLP_ERR_ARG, "batch: pair out of range");
    const PairSlot &s = b->slot[pair];
    const BatchClass &c = b->cls[s.cls];
    if (u) *u = c.u + s.first;
    if (v) *v = c.v + s.first;
    if (sub) *sub = c.sub + (size_t)s.first * 4;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_batch_report_get(qlp_batch *b, qlp_batch_report *r)
try {
    if (!b || !r) return fail(QLP_ERR_ARG, "batch_report: null argument");
    memset(r, 0, sizeof(*r));
    for (const BatchClass &c : b->cls) {
        if (c.n == 0) continue;
        qlp_batch_class_report &o = r->cls[r->n_classes++];
        o.ia = c.N;
        o.n_ias = c.n;
        o.shape = c.shape;
        o.n_candidates = c.n_cand;
        for (int k = 0; k < c.n_cand; k++) {
            o.cand_shape[k] = c.cand_shape[k];
            o.cand_ms[k] = c.cand_ms[k];
        }
    }
    r->schedule = b->schedule;
    r->serial_ms = b->serial_ms;
    r->concurrent_ms = b->concurrent_ms;
    return QLP_OK;
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
