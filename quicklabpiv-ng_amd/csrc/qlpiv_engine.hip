// The step engine: continuous batching of frame-pair levels.
//
// A step runs one adaptive level of many frame pairs at once, each pair with
// its own frame size, IA size and level.  The IAs of a step are grouped by IA
// size into four classes (16, 32, 64, 128 px); every class runs a fixed chain
// of kernels:
//   tiles  : per IA, the pair's displacement state for this level (zero at the
//            first level, else the Direct inheritance from the parent level,
//            AdaptiveInterVelocityInheritanceStrategyDirect), the A window
//            clipped and the B window clipped, shifted by the rounded
//            displacement (NoImageWarpingStrategy :79-81) or micro-warped
//            (BiLinearImageMicroAndMiniWarpingStrategy, SecondImage), with the
//            tile maxima (>= 16 clamp; 255 for a warped tile);
//   pass   : the correlation of the tiles with the fused FromCenter peak and
//            fit neighbourhood (any kernel shape or arithmetic form);
//   fit    : the sub-pixel fit and the displacement update u += i' - D/2
//            (Tile.accumulateDisplacement), or the integer peak for a level
//            before the interpolator start step, written with the fit record
//            straight into the IA's pair level arrays.
// The per-level values are those of a one-pair multi-pass plan
// (qlp_piv_run_dev) with the same configuration.
//
// Every kernel is persistent: its grid is the work-groups the device holds at
// once (or fewer), and it reads the step's IA count of its class from device
// memory, so the chain of the classes a step holds is captured once into a HIP
// graph and replayed for every later step that holds the same classes,
// whatever their IA counts (one graph per set of classes, at most 15).  Per
// step the host writes a step table (per class: IA count and the segments, one
// per pair level: frames, grid, parent and output arrays, first IA) into
// pinned memory, copies it to the device and launches the graph of the step's
// classes.  The classes run on forked streams inside the graph or in sequence,
// whichever the first step with that set measures faster; each class's kernel
// shape is picked by timing the candidates on the first step that holds the
// class (cached in the context).
#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "qlpiv_device.h"
#include "qlpiv_internal.h"

namespace qlpiv {

static const int kEngSizes[ENG_CLASSES] = {16, 32, 64, 128};
constexpr int ENG_NBR_HALF = 7;

int eng_class_of(int ia)
{
    for (int c = 0; c < ENG_CLASSES; c++)
        if (kEngSizes[c] == ia) return c;
    return -1;
}

// The tiles of one class (persistent).  A wavefront forms 64 / LPI tiles at
// once, LPI = min(N, 64) lanes per tile, a lane per column (so each row is
// read with consecutive lanes on consecutive bytes), more tiles per wavefront
// for small IAs so that their load chains overlap.  The class's segment starts
// are staged in LDS once; an IA's segment (the last one starting at or before
// it) is then a binary search in LDS.
template <int N>
__global__ __launch_bounds__(256) void step_tiles_kernel(
    const StepHeader *__restrict__ h, const StepSeg *__restrict__ segs, int cls,
    uint8_t *__restrict__ tA, uint8_t *__restrict__ tB, float *__restrict__ mA,
    float *__restrict__ mB, float *__restrict__ cu, float *__restrict__ cv,
    int *__restrict__ seg_of, uint8_t *__restrict__ fit_on)
{
    constexpr int LPI = N < 64 ? N : 64;  // lanes per tile
    constexpr int IPW = 64 / LPI;         // tiles per wavefront
    constexpr int PER_GROUP = 4 * IPW;    // tiles per work-group and round
    __shared__ int s_first[ENG_MAX_SEGS];
    const int n = h->n[cls], nseg = h->nseg[cls], base = h->base[cls];
    if ((int)blockIdx.x * PER_GROUP >= n) return;
    for (int k = threadIdx.x; k < nseg; k += 256) s_first[k] = segs[base + k].first;
    __syncthreads();
    const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const int sub = lane / LPI, c0 = lane - sub * LPI;
    for (int b0 = (int)blockIdx.x * PER_GROUP; b0 < n; b0 += (int)gridDim.x * PER_GROUP) {
        const int ia = b0 + wave * IPW + sub;
        const bool valid = ia < n;
        int si = base, t = 0, y0 = 0, x0 = 0, by0 = 0, bx0 = 0, H = 0, W = 0, ma = 0, mb = 0;
        float u = 0.0f, v = 0.0f, fu = 0.0f, fv = 0.0f;
        bool warp = false;
        const uint8_t *fa = nullptr, *fb = nullptr;
        if (valid) {
            int lo = 0, hi = nseg - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_first[mid] <= ia)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            si = base + lo;
            const StepSeg &S = segs[si];
            t = ia - s_first[lo];
            y0 = S.top[t];
            x0 = S.left[t];
            // displacement state of this level (inherit_direct_kernel's arithmetic)
            if (S.pu) {
                const int ci = t / S.nJ, cj = t - (t / S.nJ) * S.nJ;
                const int qi = ci >> 1, qj = cj >> 1;
                if (qi < S.pnI && qj < S.pnJ) {
                    const float a = S.pu[qi * S.pnJ + qj], b = S.pv[qi * S.pnJ + qj];
                    const bool rnd = S.warp == STEP_SHIFT;
                    u = 0.0f + (rnd ? (float)(int)floorf(a + 0.5f) : a);
                    v = 0.0f + (rnd ? (float)(int)floorf(b + 0.5f) : b);
                }
            }
            by0 = y0;
            bx0 = x0;
            warp = S.warp == STEP_MICRO && S.pu;
            if (S.warp == STEP_SHIFT && S.pu) {
                by0 = fm_round((float)y0 + u);
                bx0 = fm_round((float)x0 + v);
            } else if (warp) {
                const float flu = floorf(u), flv = floorf(v);
                by0 = (short)(y0 + (int)flu);
                bx0 = (short)(x0 + (int)flv);
                fu = u - flu;  // createMicroDisplacements :195-196
                fv = v - flv;
            }
            fa = S.a;
            fb = S.b;
            H = S.H;
            W = S.W;
            uint8_t *oa = tA + (size_t)ia * N * N, *ob = tB + (size_t)ia * N * N;
            // Loads are unconditional, from addresses clamped into the frame
            // (32-bit offsets: frames below 2^31 pixels), and masked afterwards,
            // one loop per B form, so that the loads of several rows are in
            // flight together instead of one branch and wait per pixel.
            const unsigned Hu = (unsigned)H, Wu = (unsigned)W;
            auto rows = [&](auto warp_c) {
#pragma unroll 8
                for (int r = 0; r < N; r++) {
#pragma unroll
                    for (int c = c0; c < N; c += LPI) {
                        const unsigned ya = (unsigned)(y0 + r), xa = (unsigned)(x0 + c);
                        const bool ina = ya < Hu && xa < Wu;
                        int a = fa[ina ? ya * Wu + xa : 0u];
                        a = ina ? a : 0;
                        int b;
                        if constexpr (decltype(warp_c)::value) {
                            b = (int)warp_pixel<uint8_t, unsigned>(fb, H, W, fu, fv, by0 + r, bx0 + c);
                        } else {
                            const unsigned yb = (unsigned)(by0 + r), xb = (unsigned)(bx0 + c);
                            const bool inb = yb < Hu && xb < Wu;
                            b = fb[inb ? yb * Wu + xb : 0u];
                            b = inb ? b : 0;
                            mb = max(mb, b);
                        }
                        ma = max(ma, a);
                        oa[r * N + c] = (uint8_t)a;
                        ob[r * N + c] = (uint8_t)b;
                    }
                }
            };
            if (warp)
                rows(std::true_type{});
            else
                rows(std::false_type{});
        }
        // maxima over the LPI lanes of each tile
#pragma unroll
        for (int o = LPI / 2; o >= 1; o >>= 1) {
            ma = max(ma, __shfl_xor(ma, o, 64));
            mb = max(mb, __shfl_xor(mb, o, 64));
        }
        if (valid && c0 == 0) {
            mA[ia] = (float)(ma < 16 ? 16 : ma);
            mB[ia] = warp ? 255.0f : (float)(mb < 16 ? 16 : mb);  // doBiLinearWarp :298
            cu[ia] = u;
            cv[ia] = v;
            seg_of[ia] = si;
            fit_on[ia] = (uint8_t)segs[si].fit;
        }
    }
}

struct EngineClass {
    int N = 0;
    size_t cap = 0;
    int shape = SHAPE_AUTO;
    bool fixed = false, tuned = false;
    int n_cand = 0, cand_shape[4] = {};
    float cand_ms[4] = {};
    float cost_ms = 0.0f;
    int last_n = 0;
    uint8_t *tA = nullptr, *tB = nullptr, *fit_on = nullptr;
    float *mA = nullptr, *mB = nullptr, *cu = nullptr, *cv = nullptr;
    float *peaks = nullptr, *mf = nullptr, *nbr = nullptr;
    int *seg_of = nullptr;
};

struct Engine {
    Context *ctx = nullptr;
    EngineConfig cfg{};
    EngineClass cls[ENG_CLASSES];
    std::vector<void *> bufs;
    size_t bytes = 0;
    void *bigwork = nullptr;
    size_t bigwork_bytes = 0;
    // the step table: device copy and two pinned staging buffers
    uint8_t *table_d = nullptr;
    // the step table: pinned staging buffers in turn, each rewritten once its
    // previous copy to the device has run (ENG_STAGES steps may be queued)
    uint8_t *stage_h[ENG_STAGES] = {};
    hipEvent_t staged[ENG_STAGES] = {};
    int next_stage = 0;
    size_t table_bytes = 0;
    int schedule = QLP_SCHEDULE_TUNED;
    bool schedule_fixed = false;
    bool prepared = false, built = false;
    // one captured graph per set of classes (bit c: class c holds IAs)
    struct Graph {
        int mask = 0;
        int schedule = 0;
        float serial_ms = 0.0f, concurrent_ms = 0.0f;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Graph> graphs;
    int last_graph = -1;
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t> ev;
    uint64_t steps = 0;
};

static const StepHeader *hdr_of(const Engine *e) { return (const StepHeader *)e->table_d; }
static const StepSeg *segs_of(const Engine *e)
{
    return (const StepSeg *)(e->table_d + sizeof(StepHeader));
}

static size_t nbr_len() { return 2 * ENG_NBR_HALF + 1; }

// The 128 px class's tile slots: the most any of its kernel shapes needs.
static size_t big_work_bytes(const EngineConfig &cfg, size_t cap)
{
    return std::max(big_work_need((int)cap, cfg.norm_mode, false, true, SHAPE_LDS),
                    big_work_need((int)cap, cfg.norm_mode, false, true, SHAPE_LDS_HALF));
}

size_t engine_bytes(const EngineConfig &cfg, const size_t cap[ENG_CLASSES])
{
    size_t b = sizeof(StepHeader) + (size_t)ENG_MAX_SEGS * sizeof(StepSeg);
    for (int c = 0; c < ENG_CLASSES; c++) {
        const size_t n = cap[c], N = kEngSizes[c];
        if (!n) continue;
        b += n * (2 * N * N + 1 + 4 * 6 + 12 + 8 * nbr_len() + 4) + 16 * 256;
    }
    if (cap[3]) b += big_work_bytes(cfg, cap[3]);
    return b;
}

template <typename T>
static int ealloc(Engine *e, T *&p, size_t count)
{
    void *q = nullptr;
    const size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 16;
    int rc = check_hip(hipMalloc(&q, bytes), "hipMalloc engine");
    if (rc) return rc;
    e->bufs.push_back(q);
    e->bytes += bytes;
    p = (T *)q;
    return QLP_OK;
}

void engine_destroy(Engine *e)
{
    if (!e) return;
    if (e->ctx) (void)hipStreamSynchronize(e->ctx->stream);
    for (Engine::Graph &g : e->graphs) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    for (hipStream_t s : e->branch) (void)hipStreamDestroy(s);
    for (hipEvent_t x : e->ev) (void)hipEventDestroy(x);
    for (int k = 0; k < ENG_STAGES; k++) {
        if (e->staged[k]) (void)hipEventDestroy(e->staged[k]);
        if (e->stage_h[k]) (void)hipHostFree(e->stage_h[k]);
    }
    for (void *p : e->bufs) (void)hipFree(p);
    if (e->ctx) e->ctx->hbm_in_use -= e->bytes;
    delete e;
}

int engine_create(Context *ctx, const EngineConfig &cfg, const size_t cap[ENG_CLASSES], Engine **out)
{
    *out = nullptr;
    if (cfg.norm_mode != QLP_NORM_PER_TILE && cfg.norm_mode != QLP_NORM_A_MAX)
        return fail(QLP_ERR_ARG, "engine: unknown normalisation mode");
    if (cfg.xcorr_form < QLP_XCORR_EXACT || cfg.xcorr_form > QLP_XCORR_MFMA_BF16_SINGLE)
        return fail(QLP_ERR_ARG, "engine: unknown correlation form");
    if (cfg.subpix_kind >= 0 && cfg.subpix_kind != QLP_SUBPIX_POLY &&
        (cfg.subpix_pixels < 3 || cfg.subpix_pixels % 2 == 0 || cfg.subpix_pixels / 2 > ENG_NBR_HALF))
        return fail(QLP_ERR_ARG, "engine: sub-pixel window must be odd, 3 .. 15");
    for (int c = 0; c < ENG_CLASSES; c++)
        if (cap[c] > (size_t)0x7fffffff / 4) return fail(QLP_ERR_ARG, "engine: class capacity too large");
    const size_t need = engine_bytes(cfg, cap);
    if (ctx->hbm_budget && ctx->hbm_in_use + need > ctx->hbm_budget)
        return fail(QLP_ERR_BUDGET, "engine exceeds the context's HBM budget");
    Engine *e = new Engine();
    e->ctx = ctx;
    e->cfg = cfg;
    e->table_bytes = sizeof(StepHeader) + (size_t)ENG_MAX_SEGS * sizeof(StepSeg);
    int rc = check_hip(hipSetDevice(ctx->device), "hipSetDevice");
    if (!rc) rc = ealloc(e, e->table_d, e->table_bytes);
    for (int k = 0; k < ENG_STAGES && !rc; k++) {
        rc = check_hip(hipHostMalloc((void **)&e->stage_h[k], e->table_bytes, 0), "hipHostMalloc engine");
        if (!rc) rc = check_hip(hipEventCreateWithFlags(&e->staged[k], hipEventDisableTiming), "engine event");
    }
    const size_t L = nbr_len();
    for (int c = 0; c < ENG_CLASSES && !rc; c++) {
        EngineClass &C = e->cls[c];
        C.N = kEngSizes[c];
        C.cap = cap[c];
        const size_t n = cap[c];
        if (!n) continue;
        rc = ealloc(e, C.tA, n * C.N * C.N);
        if (!rc) rc = ealloc(e, C.tB, n * C.N * C.N);
        if (!rc) rc = ealloc(e, C.fit_on, n);
        if (!rc) rc = ealloc(e, C.mA, n);
        if (!rc) rc = ealloc(e, C.mB, n);
        if (!rc) rc = ealloc(e, C.cu, n);
        if (!rc) rc = ealloc(e, C.cv, n);
        if (!rc) rc = ealloc(e, C.peaks, n * 3);
        if (!rc) rc = ealloc(e, C.mf, n);
        if (!rc) rc = ealloc(e, C.nbr, n * 2 * L);
        if (!rc) rc = ealloc(e, C.seg_of, n);
    }
    if (!rc && cap[3]) {
        const size_t w = big_work_bytes(cfg, cap[3]);
        uint8_t *p = nullptr;
        rc = ealloc(e, p, w);
        e->bigwork = p;
        e->bigwork_bytes = w;
    }
    for (int k = 0; k < 3 + ENG_CLASSES && !rc; k++) {
        hipEvent_t x = nullptr;
        rc = check_hip(k < 2 ? hipEventCreate(&x) : hipEventCreateWithFlags(&x, hipEventDisableTiming),
                       "engine event");
        if (!rc) e->ev.push_back(x);
    }
    for (int k = 0; k < ENG_CLASSES && !rc; k++) {
        hipStream_t s = nullptr;
        rc = check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "engine stream");
        if (!rc) e->branch.push_back(s);
    }
    ctx->hbm_in_use += e->bytes;
    if (rc) {
        engine_destroy(e);
        return rc;
    }
    *out = e;
    return QLP_OK;
}

int engine_fix(Engine *e, int ia, int shape, int schedule)
{
    if (e->built) return fail(QLP_ERR_ARG, "batch_fix: the batch has already run");
    if (shape < QLP_SHAPE_AUTO || shape > QLP_SHAPE_LDS_WIDE)
        return fail(QLP_ERR_ARG, "batch_fix: unknown shape");
    if (schedule < QLP_SCHEDULE_TUNED || schedule > QLP_SCHEDULE_CONCURRENT)
        return fail(QLP_ERR_ARG, "batch_fix: unknown schedule");
    if (ia != 0 && eng_class_of(ia) < 0)
        return fail(QLP_ERR_ARG, "batch_fix: IA size must be 0, 16, 32, 64 or 128");
    if (shape == QLP_SHAPE_LANE && ia != 16 && ia != 0)
        return fail(QLP_ERR_ARG, "batch_fix: the lane shape takes 16 px IAs");
    if (shape == QLP_SHAPE_LDS_WIDE && ia == 128)
        return fail(QLP_ERR_ARG, "batch_fix: the wide shape takes 16, 32 and 64 px IAs");
    for (EngineClass &C : e->cls) {
        if (ia != 0 && C.N != ia) continue;
        // shapes a class cannot take (lane above 16 px, wide at 128 px) fall back
        // to the class's default when every class is fixed at once
        const bool fits = !(shape == QLP_SHAPE_LANE && C.N != 16) &&
                          !(shape == QLP_SHAPE_LDS_WIDE && C.N == 128);
        C.shape = fits ? shape : SHAPE_AUTO;
        C.fixed = true;
    }
    if (schedule != QLP_SCHEDULE_TUNED) {
        e->schedule = schedule;
        e->schedule_fixed = true;
    }
    return QLP_OK;
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

template <int N>
static void launch_tiles(Engine *e, const EngineClass &C, int cls, hipStream_t s)
{
    auto k = step_tiles_kernel<N>;
    static const int resident = resident_groups(k, 256);
    constexpr int per_group = 4 * (64 / (N < 64 ? N : 64));
    const int grid = (int)std::min<size_t>((C.cap + per_group - 1) / per_group, (size_t)resident);
    hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, s, hdr_of(e), segs_of(e), cls, C.tA, C.tB,
                       C.mA, C.mB, C.cu, C.cv, C.seg_of, C.fit_on);
}

// Enqueue class cls's chain on stream s with kernel shape 'shape'.
static int enqueue_class(Engine *e, int cls, int shape, hipStream_t s)
{
    const EngineClass &C = e->cls[cls];
    if (!C.cap) return QLP_OK;
    Context sub = *e->ctx;  // the same tables, another stream, the engine's workspace
    sub.stream = s;
    sub.bigwork = e->bigwork;
    sub.bigwork_bytes = e->bigwork_bytes;
    switch (C.N) {
    case 16: launch_tiles<16>(e, C, cls, s); break;
    case 32: launch_tiles<32>(e, C, cls, s); break;
    case 64: launch_tiles<64>(e, C, cls, s); break;
    default: launch_tiles<128>(e, C, cls, s); break;
    }
    int rc = check_hip(hipGetLastError(), "engine tiles launch");
    if (rc) return rc;
    const int *n_dev = (const int *)(e->table_d + offsetof(StepHeader, n)) + cls;
    PassParams q;
    q.n = (int)C.cap;
    q.n_dev = n_dev;
    q.N = C.N;
    q.mode = e->cfg.norm_mode;
    q.tA = C.tA;
    q.tB = C.tB;
    q.maxA = C.mA;
    q.maxB = C.mB;
    q.peaks = C.peaks;
    q.min_floor = C.mf;
    q.nbr = C.nbr;
    q.nbr_half = ENG_NBR_HALF;
    q.arith = e->cfg.xcorr_form;
    q.shape = shape;
    rc = launch_pass(&sub, q);
    if (rc) return rc;
    // the fit writes each IA's field value and fit record into its pair level
    const int D = 2 * C.N - 1;
    FitScatter sc;
    sc.segs = segs_of(e);
    sc.seg_of = C.seg_of;
    rc = launch_subpixel_nbr(&sub, e->cfg.subpix_kind, e->cfg.subpix_pixels, e->cfg.subpix_iterations,
                             (int)C.cap, D, ENG_NBR_HALF, C.nbr, C.mf, C.peaks, nullptr, C.cu, C.cv,
                             n_dev, C.fit_on, &sc);
    if (rc) return rc;
    // every workspace was sized before: nothing may have grown during the enqueue
    bool grew = sub.bigwork != e->bigwork;
    for (int a = 0; a < 3; a++)
        for (int f = 0; f < 2; f++) grew = grew || sub.mfma_tab[a][f] != e->ctx->mfma_tab[a][f];
    if (grew) return fail(QLP_ERR_DEVICE, "engine: a workspace grew during an enqueue");
    return QLP_OK;
}

// The MFMA tables are uploaded on the form's first launch: one IA, outside any
// capture.
static int prepare_tables(Engine *e)
{
    if (e->cfg.xcorr_form == QLP_XCORR_EXACT) return QLP_OK;
    Context *ctx = e->ctx;
    for (int c = 0; c < 3; c++) {
        const EngineClass &C = e->cls[c];
        if (!C.cap) continue;
        PassParams q;
        q.n = 1;
        q.N = C.N;
        q.mode = e->cfg.norm_mode;
        q.tA = C.tA;
        q.tB = C.tB;
        q.maxA = C.mA;
        q.maxB = C.mB;
        q.peaks = C.peaks;
        q.min_floor = C.mf;
        q.nbr = C.nbr;
        q.nbr_half = ENG_NBR_HALF;
        q.arith = e->cfg.xcorr_form;
        hipLaunchKernelGGL(fill_kernel, dim3(1), dim3(256), 0, ctx->stream, C.mA, 16.0f, (size_t)1);
        hipLaunchKernelGGL(fill_kernel, dim3(1), dim3(256), 0, ctx->stream, C.mB, 16.0f, (size_t)1);
        int rc = launch_pass(ctx, q);
        if (rc) return rc;
    }
    return check_hip(hipStreamSynchronize(ctx->stream), "engine prepare");
}

static int time_class(Engine *e, int cls, int shape, float &ms)
{
    hipStream_t s = e->ctx->stream;
    int rc = enqueue_class(e, cls, shape, s);
    ms = 1e30f;
    for (int r = 0; r < 3 && !rc; r++) {
        (void)hipEventRecord(e->ev[0], s);
        rc = enqueue_class(e, cls, shape, s);
        (void)hipEventRecord(e->ev[1], s);
        if (!rc) rc = check_hip(hipEventSynchronize(e->ev[1]), "engine tune");
        float t = 0.0f;
        if (!rc) rc = check_hip(hipEventElapsedTime(&t, e->ev[0], e->ev[1]), "engine tune");
        if (!rc) ms = std::min(ms, t);
    }
    return rc;
}

static int capture(Engine *e, int mask, bool concurrent, hipGraph_t &graph, hipGraphExec_t &exec)
{
    hipStream_t s = e->ctx->stream;
    graph = nullptr;
    exec = nullptr;
    std::vector<int> order;
    for (int k = 0; k < ENG_CLASSES; k++)
        if (e->cls[k].cap && (mask >> k & 1)) order.push_back(k);
    // longest class first, so that the short ones fill in behind it
    std::sort(order.begin(), order.end(),
              [&](int x, int y) { return e->cls[x].cost_ms > e->cls[y].cost_ms; });
    int rc = check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "engine capture");
    if (rc) return rc;
    hipEvent_t fork = e->ev[2];
    if (concurrent) rc = check_hip(hipEventRecord(fork, s), "engine fork");
    for (size_t i = 0; i < order.size() && !rc; i++) {
        const int c = order[i];
        if (!concurrent) {
            rc = enqueue_class(e, c, e->cls[c].shape, s);
            continue;
        }
        hipStream_t br = e->branch[i];
        rc = check_hip(hipStreamWaitEvent(br, fork, 0), "engine fork");
        if (!rc) rc = enqueue_class(e, c, e->cls[c].shape, br);
        if (!rc) rc = check_hip(hipEventRecord(e->ev[3 + i], br), "engine join");
        if (!rc) rc = check_hip(hipStreamWaitEvent(s, e->ev[3 + i], 0), "engine join");
    }
    hipGraph_t g = nullptr;
    const int rc_end = check_hip(hipStreamEndCapture(s, &g), "engine capture");
    if (!rc) rc = rc_end;
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    rc = check_hip(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0), "engine graph instantiate");
    if (rc) {
        (void)hipGraphDestroy(g);
        return rc;
    }
    graph = g;
    return QLP_OK;
}

static int time_graph(Engine *e, hipGraphExec_t exec, float &ms)
{
    hipStream_t s = e->ctx->stream;
    int rc = check_hip(hipGraphLaunch(exec, s), "engine graph launch");
    ms = 1e30f;
    for (int r = 0; r < 3 && !rc; r++) {
        (void)hipEventRecord(e->ev[0], s);
        rc = check_hip(hipGraphLaunch(exec, s), "engine graph launch");
        (void)hipEventRecord(e->ev[1], s);
        if (!rc) rc = check_hip(hipEventSynchronize(e->ev[1]), "engine tune");
        float t = 0.0f;
        if (!rc) rc = check_hip(hipEventElapsedTime(&t, e->ev[0], e->ev[1]), "engine tune");
        if (!rc) ms = std::min(ms, t);
    }
    return rc;
}

// The kernel shape of each class this step holds for the first time: the
// context's cached pick for its size, or the fastest candidate on this step.
// The step table is already on the device; every tuning run recomputes the
// same values (a step only reads its frames and the parent levels).
static int tune_classes(Engine *e, const int counts[ENG_CLASSES])
{
    Context *ctx = e->ctx;
    for (int c = 0; c < ENG_CLASSES; c++) {
        EngineClass &C = e->cls[c];
        if (!C.cap || !counts[c] || C.tuned) continue;
        int cand[4];
        const int nc = shape_candidates(C.N, e->cfg.xcorr_form, cand);
        const uint64_t key = tune_key(C.N, counts[c], e->cfg.xcorr_form, e->cfg.norm_mode);
        C.n_cand = 0;
        if (!C.fixed) {
            auto hit = ctx->tuned_shape.find(key);
            if (hit != ctx->tuned_shape.end()) {
                C.shape = hit->second;
            } else if (nc == 1) {
                C.shape = cand[0];
            } else {
                float best = 1e30f;
                for (int k = 0; k < nc; k++) {
                    float ms = 0.0f;
                    if (int rc = time_class(e, c, cand[k], ms)) return rc;
                    C.cand_shape[C.n_cand] = cand[k];
                    C.cand_ms[C.n_cand++] = ms;
                    if (ms < best) {
                        best = ms;
                        C.shape = cand[k];
                    }
                }
                ctx->tuned_shape[key] = C.shape;
            }
        }
        float ms = 0.0f;
        if (int rc = time_class(e, c, C.shape, ms)) return rc;
        C.cost_ms = ms;
        C.tuned = true;
    }
    return QLP_OK;
}

// The schedule is timed on the first step that holds more than one class and
// at least kScheduleTuneIAs IAs; graphs captured before are serial.  On small
// steps running the classes side by side looks faster (their short kernels
// overlap), which larger steps do not repay: on a saturated stream of mixed
// requests (steps of a few hundred thousand IAs) serial steps ran 1.3x the
// concurrent ones, and a choice timed on a small first step was the wrong one.
constexpr int kScheduleTuneIAs = 1 << 18;

static void drop_graph(Engine::Graph &G)
{
    if (G.exec) (void)hipGraphExecDestroy(G.exec);
    if (G.graph) (void)hipGraphDestroy(G.graph);
    G.exec = nullptr;
    G.graph = nullptr;
}

// The graph of the classes in `mask`: captured on the first step that holds
// exactly these classes (serial until the schedule is decided, then in the
// decided schedule).  With `decide` the step captures both schedules, times
// them and keeps the faster (concurrent only when 10 % faster) for every later
// graph; graphs captured earlier are serial and are kept if serial won,
// recaptured on their next step if not.
static int graph_for(Engine *e, int mask, bool decide, int &index)
{
    if (!decide)
        for (size_t k = 0; k < e->graphs.size(); k++)
            if (e->graphs[k].mask == mask) {
                index = (int)k;
                return QLP_OK;
            }
    const bool decided = e->schedule_fixed || e->schedule != QLP_SCHEDULE_TUNED;
    const int only = decided ? e->schedule : QLP_SCHEDULE_SERIAL;
    hipGraph_t g[2] = {nullptr, nullptr};
    hipGraphExec_t x[2] = {nullptr, nullptr};
    float t[2] = {0.0f, 0.0f};
    int rc = QLP_OK;
    for (int conc = 0; conc < 2 && !rc; conc++) {
        if (!decide && conc != only) continue;
        rc = capture(e, mask, conc == 1, g[conc], x[conc]);
        if (!rc && decide) rc = time_graph(e, x[conc], t[conc]);
    }
    // concurrent only by a clear margin: one step's timing does not cover the
    // mix of step sizes that follows
    const int pick = decide ? (t[1] < 0.9f * t[0] ? QLP_SCHEDULE_CONCURRENT : QLP_SCHEDULE_SERIAL) : only;
    for (int k = 0; k < 2; k++) {
        if (!rc && k == pick) continue;
        if (x[k]) (void)hipGraphExecDestroy(x[k]);
        if (g[k]) (void)hipGraphDestroy(g[k]);
    }
    if (rc) return rc;
    if (decide) {
        e->schedule = pick;
        // this mask's earlier graph, and every serial one if concurrent won
        for (size_t k = 0; k < e->graphs.size();) {
            if (e->graphs[k].mask == mask || e->graphs[k].schedule != pick) {
                drop_graph(e->graphs[k]);
                e->graphs.erase(e->graphs.begin() + (long)k);
            } else {
                k++;
            }
        }
    }
    Engine::Graph G;
    G.mask = mask;
    G.schedule = pick;
    G.serial_ms = decide ? t[0] : 0.0f;
    G.concurrent_ms = decide ? t[1] : 0.0f;
    G.graph = g[pick];
    G.exec = x[pick];
    e->graphs.push_back(G);
    index = (int)e->graphs.size() - 1;
    return QLP_OK;
}

int engine_step(Engine *e, const std::vector<StepSeg> segs[ENG_CLASSES])
{
    Context *ctx = e->ctx;
    int counts[ENG_CLASSES] = {0, 0, 0, 0}, total_segs = 0;
    for (int c = 0; c < ENG_CLASSES; c++) {
        for (const StepSeg &s : segs[c]) {
            // the tile kernel indexes frames with 32-bit offsets
            if ((long long)s.H * s.W > 0x7fffffffLL)
                return fail(QLP_ERR_UNSUPPORTED, "engine step: frames of 2^31 pixels or more");
            counts[c] += s.nI * s.nJ;
        }
        total_segs += (int)segs[c].size();
        if ((size_t)counts[c] > e->cls[c].cap)
            return fail(QLP_ERR_ARG, "engine step: a class exceeds its IA capacity");
    }
    if (total_segs > ENG_MAX_SEGS) return fail(QLP_ERR_ARG, "engine step: too many pair levels in one step");
    if (total_segs == 0) return QLP_OK;
    // the staging buffer is rewritten once its previous copy has run
    const int k = e->next_stage;
    e->next_stage = (k + 1) % ENG_STAGES;
    int rc = check_hip(hipEventSynchronize(e->staged[k]), "engine staging");
    if (rc) return rc;
    StepHeader *h = (StepHeader *)e->stage_h[k];
    StepSeg *out = (StepSeg *)(e->stage_h[k] + sizeof(StepHeader));
    memset(h, 0, sizeof(StepHeader));
    int at = 0;
    for (int c = 0; c < ENG_CLASSES; c++) {
        h->base[c] = at;
        h->nseg[c] = (int)segs[c].size();
        h->n[c] = counts[c];
        int first = 0;
        for (const StepSeg &s : segs[c]) {
            out[at] = s;
            out[at].first = first;
            first += s.nI * s.nJ;
            at++;
        }
        e->cls[c].last_n = counts[c];
    }
    const size_t bytes = sizeof(StepHeader) + (size_t)at * sizeof(StepSeg);
    rc = check_hip(hipMemcpyAsync(e->table_d, e->stage_h[k], bytes, hipMemcpyHostToDevice, ctx->stream),
                   "engine step table");
    if (!rc) rc = check_hip(hipEventRecord(e->staged[k], ctx->stream), "engine staging");
    if (rc) return rc;
    int mask = 0;
    bool fresh = false;
    for (int c = 0; c < ENG_CLASSES; c++)
        if (counts[c]) {
            mask |= 1 << c;
            fresh = fresh || !e->cls[c].tuned;
        }
    int gi = -1;
    for (size_t g = 0; g < e->graphs.size(); g++)
        if (e->graphs[g].mask == mask) gi = (int)g;
    int classes = 0, total = 0;
    for (int c = 0; c < ENG_CLASSES; c++) {
        classes += mask >> c & 1;
        total += counts[c];
    }
    const bool decide = !e->schedule_fixed && e->schedule == QLP_SCHEDULE_TUNED && classes > 1 &&
                        total >= kScheduleTuneIAs;
    if (gi < 0 || fresh || decide) {
        // a new set of classes: tables, shapes and the graph, on this step's table
        rc = check_hip(hipStreamSynchronize(ctx->stream), "engine build");
        if (!rc && !e->prepared) rc = prepare_tables(e);
        e->prepared = e->prepared || !rc;
        if (!rc) rc = tune_classes(e, counts);
        if (!rc) rc = graph_for(e, mask, decide, gi);
        if (rc) return rc;
        e->built = true;
    }
    e->last_graph = gi;
    e->steps++;
    return check_hip(hipGraphLaunch(e->graphs[gi].exec, ctx->stream), "engine graph launch");
}

size_t engine_cap(const Engine *e, int cls) { return e->cls[cls].cap; }

int engine_report(const Engine *e, qlp_batch_report *r)
{
    memset(r, 0, sizeof(*r));
    for (int c = 0; c < ENG_CLASSES; c++) {
        const EngineClass &C = e->cls[c];
        if (!C.cap) continue;
        qlp_batch_class_report &o = r->cls[r->n_classes++];
        o.ia = C.N;
        o.n_ias = C.last_n;
        o.shape = C.shape;
        o.n_candidates = C.n_cand;
        for (int k = 0; k < C.n_cand; k++) {
            o.cand_shape[k] = C.cand_shape[k];
            o.cand_ms[k] = C.cand_ms[k];
        }
    }
    if (e->last_graph >= 0) {  // the graph of the last step
        const Engine::Graph &G = e->graphs[e->last_graph];
        r->schedule = G.schedule;
        r->serial_ms = G.serial_ms;
        r->concurrent_ms = G.concurrent_ms;
    } else {
        r->schedule = e->schedule;
    }
    return QLP_OK;
}

}  // namespace qlpiv
