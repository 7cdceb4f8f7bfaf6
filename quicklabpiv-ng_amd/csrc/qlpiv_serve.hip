// The serving form: a request loop over frame pairs of mixed sizes.
//
// The reference processes frame pairs with a pool of worker threads, one
// correlation job (and device queue) per thread (InputFiles.java:172-224,
// OpenClGpuManager.java:124-142).  Here one process owns the GPU and a server
// object runs the requests with continuous batching:
//   - batched path (the configurations the step engine runs: FromCenter,
//     Direct inheritance, no warp or the micro warp, no validation, 8-bit
//     frames): every admitted request holds its frames, its level geometry
//     (shared by the requests of one frame size and configuration) and the
//     displacement field of each level on the device; each step of the worker
//     advances EVERY admitted request by one adaptive level, all of them, of
//     every frame and IA size, in one launch of the step engine's captured
//     graph of persistent kernels (qlpiv_engine.hip).  Requests that arrive
//     while others are on the GPU are admitted between steps as soon as their
//     memory fits the budget, so their levels join the running steps; a request
//     whose last level has run has its field copied to pinned host memory and
//     retires;
//   - plan path (every other configuration): slots, each a library context
//     (its own HIP stream) with device frame buffers and a cache of multi-pass
//     plans keyed by frame size and configuration; requests on different slots
//     overlap on the device;
//   - admission in arrival order against the server's HBM budget; cached plans
//     of idle slots and unused level geometries are evicted, least recently
//     used first, to make room.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "qlpiv_internal.h"

namespace qlpiv {

using Clock = std::chrono::steady_clock;
constexpr int ADMIT_PER_ROUND = 32;  // admissions per round of the worker loop

// The IA grids of every level of one frame size and configuration (device).
struct GeoLevel {
    int N = 0, nI = 0, nJ = 0;
    int *top = nullptr, *left = nullptr;
    std::vector<int> top_h, left_h;  // kept until the server ends (upload sources)
};
struct Geo {
    std::vector<GeoLevel> lv;
    size_t bytes = 0;
    int users = 0;
    uint64_t last_use = 0;
};

struct ServeRequest {
    uint64_t ticket = 0;
    qlp_request req{};
    int status = 0;  // 0 pending, 1 running, 2 done, 3 failed
    int rc = QLP_OK;
    std::string err;
    Clock::time_point t_submit, t_admit, t_done;
    // result: the last level's field
    int nI = 0, nJ = 0, ia = 0;
    std::vector<float> u, v;
    size_t ias = 0;
    // batched path state
    bool batched = false;
    int engine = -1;                // index of the server's engine for this configuration
    Geo *geo = nullptr;
    int next_level = 0;             // the level the next step runs
    const uint8_t *fa = nullptr, *fb = nullptr;  // device frames (own copies or the caller's)
    uint8_t *own = nullptr;         // server-owned frame copies (2 H W), or null
    std::vector<float *> lu, lv;    // each level's field (device), parts of `fields`
    float *fields = nullptr;
    float *hu = nullptr, *hv = nullptr;  // pinned copies of the last level's field
    size_t hucap = 0, hvcap = 0;         // their capacities (floats)
    size_t bytes = 0;               // device bytes held (frames and fields)
};

struct PlanEntry {
    qlp_piv_plan *plan = nullptr;
    size_t bytes = 0;
    uint64_t last_use = 0;
};

struct Slot {
    qlp_ctx *ctx = nullptr;
    std::map<std::string, PlanEntry> plans;  // key: frame size + configuration bytes
    uint8_t *dA = nullptr, *dB = nullptr;    // device frames
    size_t frame_bytes = 0;
    float *hu = nullptr, *hv = nullptr;      // pinned host copies of the last level
    size_t field_cap = 0;
    hipEvent_t done = nullptr;
    std::shared_ptr<ServeRequest> cur;
    PlanEntry *cur_plan = nullptr;
};

}  // namespace qlpiv

using namespace qlpiv;

struct qlp_server {
    int device = 0;
    size_t budget = 0, in_use = 0, peak = 0;
    std::vector<Slot> slots;
    // batched path: one context (stream) for every step, engines by
    // configuration, level geometries by frame size and configuration
    qlp_ctx *ectx = nullptr;
    struct EngineEntry {
        EngineConfig cfg;
        Engine *eng = nullptr;
        size_t bytes = 0;
    };
    std::vector<EngineEntry> engines;
    std::map<std::string, std::unique_ptr<Geo>> geos;
    std::deque<std::shared_ptr<ServeRequest>> active;  // admitted, in admission order
    int max_active = 512;
    std::multimap<size_t, float *> pinned_free;  // pinned result buffers by capacity (floats)
    std::vector<void *> pinned_slabs;            // their pinned allocations
    uint8_t *slab_at = nullptr;                  // unused part of the newest slab
    size_t slab_left = 0;
    // steps in flight, oldest first: the event recorded after each step (and
    // the field copies of the requests whose last level it ran) and those
    // requests, which retire when it completes
    struct StepRec {
        hipEvent_t ev = nullptr;
        std::vector<std::shared_ptr<ServeRequest>> done;
    };
    std::deque<StepRec> inflight;
    std::vector<hipEvent_t> step_events_free;
    uint64_t steps = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<ServeRequest>> pending;
    std::map<uint64_t, std::shared_ptr<ServeRequest>> all;
    uint64_t next_ticket = 1, use_clock = 0;
    bool stop = false;
    std::thread worker;
    Clock::time_point t0;
    // statistics
    uint64_t completed = 0, failed = 0, evictions = 0;
    size_t ias_done = 0;
};

namespace {

std::string plan_key(const qlp_request &r)
{
    std::string k(reinterpret_cast<const char *>(&r.H), sizeof(int) * 2);
    k.append(reinterpret_cast<const char *>(&r.cfg), sizeof(qlp_piv_config));
    return k;
}

// Frees cached plans of idle slots (least recently used first) until `need`
// more bytes fit the budget.  Called with the lock held.
bool make_room(qlp_server *s, size_t need, const Slot *keep_slot, const std::string *keep_key)
{
    while (s->in_use + need > s->budget) {
        Slot *bs = nullptr;
        std::map<std::string, PlanEntry>::iterator best;
        for (Slot &sl : s->slots) {
            if (sl.cur) continue;  // busy slots keep their plans
            for (auto it = sl.plans.begin(); it != sl.plans.end(); ++it) {
                if (&sl == keep_slot && keep_key && it->first == *keep_key) continue;
                if (!bs || it->second.last_use < best->second.last_use) {
                    bs = &sl;
                    best = it;
                }
            }
        }
        if (!bs) return false;
        qlp_piv_plan_destroy(best->second.plan);
        s->in_use -= best->second.bytes;
        bs->plans.erase(best);
        s->evictions++;
    }
    return true;
}

void finish(qlp_server *s, Slot &sl)
{
    auto r = sl.cur;
    const size_t n = (size_t)r->nI * r->nJ;
    r->u.assign(sl.hu, sl.hu + n);
    r->v.assign(sl.hv, sl.hv + n);
    r->t_done = Clock::now();
    r->status = 2;
    s->completed++;
    s->ias_done += r->ias;
    sl.cur.reset();
    sl.cur_plan = nullptr;
}

void fail_request(qlp_server *s, const std::shared_ptr<ServeRequest> &r, int rc, const char *what)
{
    r->rc = rc;
    r->err = std::string(what) + ": " + qlp_last_error();
    r->t_done = Clock::now();
    r->status = 3;
    s->failed++;
}

// Launches request r on slot sl (lock held; the calls are asynchronous).
int launch(qlp_server *s, Slot &sl, const std::shared_ptr<ServeRequest> &r, const std::string &key,
           size_t frame_need, size_t plan_need)
{
    const qlp_request &q = r->req;
    const size_t fb = (size_t)q.H * q.W;
    if (frame_need) {
        if (sl.dA) (void)hipFree(sl.dA);
        if (sl.dB) (void)hipFree(sl.dB);
        s->in_use -= 2 * sl.frame_bytes;
        sl.dA = sl.dB = nullptr;
        sl.frame_bytes = 0;
        if (hipMalloc(&sl.dA, fb) != hipSuccess || hipMalloc(&sl.dB, fb) != hipSuccess)
            return fail(QLP_ERR_DEVICE, "server: frame allocation failed");
        sl.frame_bytes = fb;
        s->in_use += 2 * fb;
    }
    auto it = sl.plans.find(key);
    if (it == sl.plans.end()) {
        PlanEntry e;
        if (int rc = qlp_piv_plan_create(sl.ctx, q.H, q.W, 1, &q.cfg, &e.plan)) return rc;
        e.bytes = plan_need;
        s->in_use += plan_need;
        it = sl.plans.emplace(key, e).first;
    }
    it->second.last_use = ++s->use_clock;
    sl.cur_plan = &it->second;
    if (s->in_use > s->peak) s->peak = s->in_use;
    qlp_piv_plan *p = it->second.plan;
    int nlev = 0;
    qlp_piv_plan_levels(p, &nlev);
    int ia = 0, nI = 0, nJ = 0;
    size_t ias = 0;
    for (int l = 0; l < nlev; l++) {
        qlp_piv_plan_level(p, l, &ia, &nI, &nJ, nullptr, nullptr);
        ias += (size_t)nI * nJ;
    }
    r->ia = ia;
    r->nI = nI;
    r->nJ = nJ;
    r->ias = ias;
    const size_t n = (size_t)nI * nJ;
    if (n > sl.field_cap) {
        if (sl.hu) (void)hipHostFree(sl.hu);
        if (sl.hv) (void)hipHostFree(sl.hv);
        sl.hu = sl.hv = nullptr;
        if (hipHostMalloc((void **)&sl.hu, n * 4, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc((void **)&sl.hv, n * 4, hipHostMallocDefault) != hipSuccess) {
            sl.field_cap = 0;
            return fail(QLP_ERR_DEVICE, "server: pinned result buffer allocation failed");
        }
        sl.field_cap = n;
    }
    hipStream_t st = (hipStream_t)qlp_ctx_stream(sl.ctx);
    const hipMemcpyKind k = q.frames_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (hipMemcpyAsync(sl.dA, q.imgA, fb, k, st) != hipSuccess ||
        hipMemcpyAsync(sl.dB, q.imgB, fb, k, st) != hipSuccess)
        return fail(QLP_ERR_DEVICE, "server: frame copy failed");
    if (int rc = qlp_piv_run_dev(p, sl.dA, sl.dB)) return rc;
    const float *du = nullptr, *dv = nullptr, *pk = nullptr, *pl = nullptr;
    qlp_piv_result_dev(p, nlev - 1, &du, &dv, &pk, &pl);
    if (hipMemcpyAsync(sl.hu, du, n * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(sl.hv, dv, n * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipEventRecord(sl.done, st) != hipSuccess)
        return fail(QLP_ERR_DEVICE, "server: result copy failed");
    sl.cur = r;
    r->status = 1;
    r->t_admit = Clock::now();
    return QLP_OK;
}

// ---- batched path -----------------------------------------------------------

// The configurations the step engine runs (every other one takes a plan slot).
bool engine_ok(const qlp_piv_config &c)
{
    return c.peak_kind == QLP_PEAK_FROM_CENTER && c.n_peaks <= 1 && c.valid.n_stages == 0 &&
           c.inherit_mode == QLP_INHERIT_DIRECT && c.frame_type == QLP_FRAME_U8 &&
           (c.warp_mode == QLP_WARP_NONE ||
            (c.warp_mode == QLP_WARP_MICRO && c.warp_order == QLP_WARP_ORDER_SECOND)) &&
           c.keep_planes == 0 && eng_class_of(c.ia_end) >= 0 && eng_class_of(c.ia_start) >= 0;
}

std::string geo_key(const qlp_request &r)
{
    const qlp_piv_config &c = r.cfg;
    const int v[9] = {r.H, r.W, c.ia_start, c.ia_end, c.margin_top, c.margin_bottom,
                      c.margin_left, c.margin_right, 0};
    std::string k(reinterpret_cast<const char *>(v), sizeof v);
    k.append(reinterpret_cast<const char *>(&c.overlap_factor), sizeof(float));
    return k;
}

// Level grids of a request (host side), and the device bytes they need.
int build_geo(const qlp_request &r, Geo &g)
{
    const qlp_piv_config &c = r.cfg;
    const int m[4] = {c.margin_top, c.margin_bottom, c.margin_left, c.margin_right};
    g.lv.clear();
    g.bytes = 0;
    for (int l = 0, tile = c.ia_start; tile >= c.ia_end; l++, tile /= 2) {
        GeoLevel L;
        L.N = tile;
        level_origins(r.H, r.W, tile, c.overlap_factor, m, l == 0, L.nI, L.nJ, L.top_h, L.left_h);
        if (L.nI < 1 || L.nJ < 1) return fail(QLP_ERR_ARG, "server: no IA fits the frame");
        g.bytes += (size_t)L.nI * L.nJ * 8;
        g.lv.push_back(std::move(L));
    }
    return QLP_OK;
}

// Step capacity of each IA class (IAs), scaled so that the engine takes at most
// a quarter of the budget.
void engine_caps(const EngineConfig &ec, size_t budget, size_t cap[ENG_CLASSES])
{
    const size_t base[ENG_CLASSES] = {1u << 20, 1u << 18, 1u << 16, 1u << 14};
    for (int shift = 0;; shift++) {
        for (int c = 0; c < ENG_CLASSES; c++) cap[c] = std::max<size_t>(base[c] >> shift, 64);
        if (engine_bytes(ec, cap) <= budget / 4 || shift >= 10) return;
    }
}

int engine_for(qlp_server *s, const qlp_piv_config &c, int &index)
{
    const EngineConfig ec{c.norm_mode, c.subpix_kind, c.subpix_pixels, c.subpix_iterations,
                          c.xcorr_form};
    for (size_t k = 0; k < s->engines.size(); k++) {
        const EngineConfig &o = s->engines[k].cfg;
        if (o.norm_mode == ec.norm_mode && o.subpix_kind == ec.subpix_kind &&
            o.subpix_pixels == ec.subpix_pixels && o.subpix_iterations == ec.subpix_iterations &&
            o.xcorr_form == ec.xcorr_form) {
            index = (int)k;
            return QLP_OK;
        }
    }
    size_t cap[ENG_CLASSES];
    engine_caps(ec, s->budget, cap);
    qlp_server::EngineEntry e;
    e.cfg = ec;
    e.bytes = engine_bytes(ec, cap);
    if (s->in_use + e.bytes > s->budget) return fail(QLP_ERR_BUDGET, "server: no room for a batch engine");
    if (int rc = engine_create(ctx_of(s->ectx), ec, cap, &e.eng)) return rc;
    s->in_use += e.bytes;
    s->engines.push_back(e);
    index = (int)s->engines.size() - 1;
    return QLP_OK;
}

hipStream_t estream(qlp_server *s) { return (hipStream_t)qlp_ctx_stream(s->ectx); }

// Frees unused level geometries (least recently used first) until `need` more
// bytes fit the budget.  Called with the lock held.
bool free_geos(qlp_server *s, size_t need)
{
    while (s->in_use + need > s->budget) {
        auto best = s->geos.end();
        for (auto it = s->geos.begin(); it != s->geos.end(); ++it)
            if (it->second->users == 0 && (best == s->geos.end() || it->second->last_use < best->second->last_use))
                best = it;
        if (best == s->geos.end()) return false;
        for (GeoLevel &L : best->second->lv) {
            (void)hipFreeAsync(L.top, estream(s));
            (void)hipFreeAsync(L.left, estream(s));
        }
        s->in_use -= best->second->bytes;
        s->geos.erase(best);
        s->evictions++;
    }
    return true;
}

// Admits request r on the batched path if its memory fits the budget now.
// Returns 1 when admitted, 0 when it must wait, -1 when it failed (rc set).
int admit_batched(qlp_server *s, const std::shared_ptr<ServeRequest> &r, int &rc)
{
    rc = QLP_OK;
    if ((int)s->active.size() >= s->max_active) return 0;
    const qlp_request &q = r->req;
    int ei = -1;
    if ((rc = engine_for(s, q.cfg, ei))) return -1;
    const std::string gk = geo_key(q);
    auto git = s->geos.find(gk);
    Geo fresh;
    Geo *g = nullptr;
    if (git != s->geos.end()) {
        g = git->second.get();
    } else {
        if ((rc = build_geo(q, fresh))) return -1;
        g = &fresh;
    }
    size_t state = 0;
    for (const GeoLevel &L : g->lv) {
        const size_t n = (size_t)L.nI * L.nJ;
        state += ((n + 63) & ~(size_t)63) * 8;  // u and v, as allocated below
        if (n > engine_cap(s->engines[ei].eng, eng_class_of(L.N))) {
            rc = fail(QLP_ERR_BUDGET, "server: a level holds more IAs than a batch step (raise the budget)");
            return -1;
        }
    }
    const size_t frames = q.frames_on_device ? 0 : 2 * (size_t)q.H * q.W;
    const size_t need = state + frames + (git == s->geos.end() ? g->bytes : 0);
    if (s->in_use + need > s->budget) {
        if (!free_geos(s, need) && !make_room(s, need, nullptr, nullptr)) return 0;
        if (s->in_use + need > s->budget) return 0;
    }
    hipStream_t st = estream(s);
    if (git == s->geos.end()) {
        for (GeoLevel &L : fresh.lv) {
            const size_t n = (size_t)L.nI * L.nJ;
            if ((rc = check_hip(hipMallocAsync((void **)&L.top, n * 4, st), "server geometry")) ||
                (rc = check_hip(hipMallocAsync((void **)&L.left, n * 4, st), "server geometry")) ||
                (rc = check_hip(hipMemcpyAsync(L.top, L.top_h.data(), n * 4, hipMemcpyHostToDevice, st),
                                "server geometry")) ||
                (rc = check_hip(hipMemcpyAsync(L.left, L.left_h.data(), n * 4, hipMemcpyHostToDevice, st),
                                "server geometry")))
                return -1;
        }
        auto up = std::make_unique<Geo>(std::move(fresh));
        g = up.get();
        s->in_use += g->bytes;
        s->geos.emplace(gk, std::move(up));
    }
    g->users++;
    g->last_use = ++s->use_clock;
    r->batched = true;
    r->engine = ei;
    r->geo = g;
    r->next_level = 0;
    r->bytes = state + frames;
    s->in_use += r->bytes;
    if (frames) {
        if ((rc = check_hip(hipMallocAsync((void **)&r->own, frames, st), "server frames"))) return -1;
        const size_t fb = (size_t)q.H * q.W;
        if ((rc = check_hip(hipMemcpyAsync(r->own, q.imgA, fb, hipMemcpyHostToDevice, st), "server frames")) ||
            (rc = check_hip(hipMemcpyAsync(r->own + fb, q.imgB, fb, hipMemcpyHostToDevice, st), "server frames")))
            return -1;
        r->fa = r->own;
        r->fb = r->own + fb;
    } else {
        r->fa = q.imgA;
        r->fb = q.imgB;
    }
    // every level's u and v in one stream-ordered allocation
    r->ias = 0;
    size_t total = 0;
    for (const GeoLevel &L : g->lv) total += 2 * (((size_t)L.nI * L.nJ + 63) & ~(size_t)63);
    if ((rc = check_hip(hipMallocAsync((void **)&r->fields, total * 4, st), "server field"))) return -1;
    float *at = r->fields;
    for (const GeoLevel &L : g->lv) {
        const size_t n = (size_t)L.nI * L.nJ, padded = (n + 63) & ~(size_t)63;
        r->lu.push_back(at);
        r->lv.push_back(at + padded);
        at += 2 * padded;
        r->ias += n;
    }
    const GeoLevel &last = g->lv.back();
    r->ia = last.N;
    r->nI = last.nI;
    r->nJ = last.nJ;
    if (s->in_use > s->peak) s->peak = s->in_use;
    r->status = 1;
    r->t_admit = Clock::now();
    s->active.push_back(r);
    return 1;
}

// Returns the device state of a batched request (stream-ordered frees).
void release_batched(qlp_server *s, ServeRequest &r)
{
    hipStream_t st = estream(s);
    if (r.fields) (void)hipFreeAsync(r.fields, st);
    r.fields = nullptr;
    r.lu.clear();
    r.lv.clear();
    if (r.own) (void)hipFreeAsync(r.own, st);
    r.own = nullptr;
    if (r.hu) s->pinned_free.emplace(r.hucap, r.hu);
    if (r.hv && r.hvcap) s->pinned_free.emplace(r.hvcap, r.hv);
    r.hu = r.hv = nullptr;
    r.hucap = r.hvcap = 0;
    if (r.geo) r.geo->users--;
    r.geo = nullptr;
    s->in_use -= r.bytes;
    r.bytes = 0;
}

// A pinned result buffer of at least n floats: a free one, else carved from a
// pinned slab (8 MiB or the buffer's size), so that host allocations, which
// wait for the device, stay rare.  Capacities are powers of two.
int pinned_get(qlp_server *s, size_t n, float **out, size_t &cap)
{
    auto it = s->pinned_free.lower_bound(n);
    if (it != s->pinned_free.end()) {
        *out = it->second;
        cap = it->first;
        s->pinned_free.erase(it);
        return QLP_OK;
    }
    size_t c = 1024;
    while (c < n) c <<= 1;
    const size_t bytes = c * 4;
    if (s->slab_left < bytes) {
        const size_t slab = std::max<size_t>(bytes, (size_t)8 << 20);
        void *p = nullptr;
        if (int rc = check_hip(hipHostMalloc(&p, slab, hipHostMallocDefault), "server pinned buffer")) return rc;
        s->pinned_slabs.push_back(p);
        s->slab_at = (uint8_t *)p;
        s->slab_left = slab;
    }
    *out = (float *)s->slab_at;
    cap = c;
    s->slab_at += bytes;
    s->slab_left -= bytes;
    return QLP_OK;
}

// Retires the requests of the completed steps (oldest first): their last
// level and field copy have run.
bool retire_batched(qlp_server *s)
{
    bool progress = false;
    while (!s->inflight.empty()) {
        qlp_server::StepRec &rec = s->inflight.front();
        if (rec.ev && hipEventQuery(rec.ev) != hipSuccess) break;
        for (const std::shared_ptr<ServeRequest> &r : rec.done) {
            try {
                const size_t n = (size_t)r->nI * r->nJ;
                r->u.assign(r->hu, r->hu + n);
                r->v.assign(r->hv, r->hv + n);
                r->t_done = Clock::now();
                r->status = 2;
                s->completed++;
                s->ias_done += r->ias;
            } catch (...) {
                fail_request(s, r, fail_exception(), "server: result");
            }
            release_batched(s, *r);
        }
        if (!rec.done.empty()) {
            s->active.erase(std::remove_if(s->active.begin(), s->active.end(),
                                           [](const std::shared_ptr<ServeRequest> &r) {
                                               return r->status != 1;
                                           }),
                            s->active.end());
            progress = true;
        }
        if (rec.ev) s->step_events_free.push_back(rec.ev);
        s->inflight.pop_front();
    }
    return progress;
}

// One step: the next level of every admitted request that has one, as far as
// the engines' class capacities allow, then the field copy of every request
// whose last level ran.  At most ENG_STAGES - 1 steps are in flight, so the
// host prepares the next steps while the device runs the queued ones.
bool step_batched(qlp_server *s)
{
    if ((int)s->inflight.size() >= ENG_STAGES - 1) return false;
    const size_t ne = s->engines.size();
    std::vector<std::array<std::vector<StepSeg>, ENG_CLASSES>> segs(ne);
    std::vector<std::array<size_t, ENG_CLASSES>> count(ne);
    for (auto &c : count) c.fill(0);
    std::vector<std::shared_ptr<ServeRequest>> stepped;
    int nseg = 0;
    for (const auto &r : s->active) {
        if (r->next_level >= (int)r->lu.size() || nseg >= ENG_MAX_SEGS) continue;
        const int l = r->next_level;
        const GeoLevel &L = r->geo->lv[l];
        const int c = eng_class_of(L.N);
        const size_t n = (size_t)L.nI * L.nJ;
        if (count[r->engine][c] + n > engine_cap(s->engines[r->engine].eng, c)) continue;
        StepSeg g{};
        g.a = r->fa;
        g.b = r->fb;
        g.top = L.top;
        g.left = L.left;
        if (l > 0) {
            const GeoLevel &P = r->geo->lv[l - 1];
            g.pu = r->lu[l - 1];
            g.pv = r->lv[l - 1];
            g.pnI = P.nI;
            g.pnJ = P.nJ;
        }
        g.u = r->lu[l];
        g.v = r->lv[l];
        g.H = r->req.H;
        g.W = r->req.W;
        g.nI = L.nI;
        g.nJ = L.nJ;
        g.warp = r->req.cfg.warp_mode == QLP_WARP_MICRO ? STEP_MICRO : STEP_SHIFT;
        g.fit = l >= r->req.cfg.interpolator_start_step ? 1 : 0;
        segs[r->engine][c].push_back(g);
        count[r->engine][c] += n;
        nseg++;
        stepped.push_back(r);
    }
    if (stepped.empty()) return false;
    hipStream_t st = estream(s);
    for (size_t e = 0; e < ne; e++) {
        bool any = false;
        for (int c = 0; c < ENG_CLASSES; c++) any = any || !segs[e][c].empty();
        if (!any) continue;
        if (int rc = engine_step(s->engines[e].eng, segs[e].data())) {
            // the requests of a failed step fail; their state is released at once
            for (auto &r : stepped)
                if (r->engine == (int)e) r->next_level = -1;
            for (auto it = s->active.begin(); it != s->active.end();) {
                if ((*it)->next_level == -1) {
                    fail_request(s, *it, rc, "server: step");
                    (void)hipStreamSynchronize(st);
                    release_batched(s, **it);
                    it = s->active.erase(it);
                } else {
                    ++it;
                }
            }
        }
    }
    qlp_server::StepRec rec;
    for (auto &r : stepped) {
        if (r->next_level < 0) continue;
        r->next_level++;
        if (r->next_level < (int)r->lu.size()) continue;
        // last level: the field to pinned host memory; the request retires with
        // this step's event
        // u and v of a level are adjacent in the request's field allocation
        // (v at u + padded): one copy into one pinned buffer
        const size_t n = (size_t)r->nI * r->nJ, padded = (n + 63) & ~(size_t)63;
        int rc = pinned_get(s, 2 * padded, &r->hu, r->hucap);
        if (!rc) {
            r->hv = r->hu + padded;
            r->hvcap = 0;  // part of hu's buffer
            rc = check_hip(hipMemcpyAsync(r->hu, r->lu.back(), 2 * padded * 4, hipMemcpyDeviceToHost, st),
                           "server result");
        }
        if (rc) {
            fail_request(s, r, rc, "server: result copy");
            (void)hipStreamSynchronize(st);
            release_batched(s, *r);
            s->active.erase(std::find(s->active.begin(), s->active.end(), r));
            continue;
        }
        rec.done.push_back(r);
    }
    if (!s->step_events_free.empty()) {
        rec.ev = s->step_events_free.back();
        s->step_events_free.pop_back();
    } else if (hipEventCreateWithFlags(&rec.ev, hipEventDisableTiming) != hipSuccess) {
        rec.ev = nullptr;
    }
    if (!rec.ev || hipEventRecord(rec.ev, st) != hipSuccess) {
        // no event: wait here, the step then counts as complete
        (void)hipStreamSynchronize(st);
        if (rec.ev) s->step_events_free.push_back(rec.ev);
        rec.ev = nullptr;
    }
    s->inflight.push_back(std::move(rec));
    s->steps++;
    return true;
}

void worker_loop(qlp_server *s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        bool progress = false;
        // retire completed requests
        for (Slot &sl : s->slots) {
            if (sl.cur && hipEventQuery(sl.done) == hipSuccess) {
                try {
                    finish(s, sl);
                } catch (...) {  // no exception may end the worker thread
                    const auto r = sl.cur;
                    sl.cur.reset();
                    sl.cur_plan = nullptr;
                    fail_request(s, r, fail_exception(), "server: result");
                }
                progress = true;
            }
        }
        if (retire_batched(s)) progress = true;
        if (progress) s->cv_done.notify_all();
        // admit in arrival order while the head request fits: the batched path
        // (memory and the in-flight limit) or a free plan slot.  At most
        // ADMIT_PER_ROUND requests per round, so that a burst of arrivals does
        // not hold back the next step
        int admitted = 0;
        while (!s->pending.empty() && admitted < ADMIT_PER_ROUND) {
            admitted++;
            const std::shared_ptr<ServeRequest> r = s->pending.front();
            if (engine_ok(r->req.cfg)) {
                int rc = QLP_OK, got;
                try {
                    got = admit_batched(s, r, rc);
                } catch (...) {
                    rc = fail_exception();
                    got = -1;
                }
                if (got == 0) break;  // wait for memory or a batch place to free up
                s->pending.pop_front();
                if (got < 0) {
                    (void)hipStreamSynchronize(estream(s));
                    release_batched(s, *r);
                    fail_request(s, r, rc, "server: admission");
                    s->cv_done.notify_all();
                }
                progress = true;
                continue;
            }
            Slot *free_slot = nullptr;
            const std::string key = plan_key(r->req);
            // prefer a free slot that already holds this plan
            for (Slot &sl : s->slots)
                if (!sl.cur && sl.plans.count(key)) {
                    free_slot = &sl;
                    break;
                }
            if (!free_slot)
                for (Slot &sl : s->slots)
                    if (!sl.cur) {
                        free_slot = &sl;
                        break;
                    }
            if (!free_slot) break;
            size_t plan_need = 0;
            if (!free_slot->plans.count(key)) {
                int rc = qlp_piv_plan_estimate(r->req.H, r->req.W, 1, &r->req.cfg, &plan_need);
                if (rc) {
                    s->pending.pop_front();
                    fail_request(s, r, rc, "server: plan");
                    progress = true;
                    s->cv_done.notify_all();
                    continue;
                }
            }
            const size_t fb = (size_t)r->req.H * r->req.W;
            const size_t frame_need = fb > free_slot->frame_bytes ? 2 * fb : 0;
            const size_t need = plan_need + (frame_need ? frame_need - 2 * free_slot->frame_bytes : 0);
            if (!make_room(s, need, free_slot, &key)) break;  // wait for memory to free up
            s->pending.pop_front();
            int rc;
            try {
                rc = launch(s, *free_slot, r, key, frame_need, plan_need);
            } catch (...) {
                rc = fail_exception();
            }
            if (rc) {
                free_slot->cur.reset();
                free_slot->cur_plan = nullptr;
                fail_request(s, r, rc, "server: launch");
                s->cv_done.notify_all();
            }
            progress = true;
        }
        // the next level of every admitted batched request
        try {
            if (step_batched(s)) progress = true;
        } catch (...) {
            (void)fail_exception();
        }
        if (s->stop && s->pending.empty()) {
            bool busy = !s->active.empty();
            for (Slot &sl : s->slots) busy |= (bool)sl.cur;
            if (!busy) break;
        }
        if (!progress) {
            bool busy = !s->active.empty();
            for (Slot &sl : s->slots) busy |= (bool)sl.cur;
            // poll the device while work is in flight, else sleep until a submit
            if (busy)
                s->cv_work.wait_for(lk, std::chrono::microseconds(20));
            else
                s->cv_work.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
}

double ms_since(const Clock::time_point &t0, const Clock::time_point &t)
{
    return std::chrono::duration<double, std::milli>(t - t0).count();
}

}  // namespace

extern "C" {

int qlp_server_create(int hip_device, const qlp_server_config *cfg, qlp_server **out)
try {
    if (!cfg || !out) return fail(QLP_ERR_ARG, "server_create: null argument");
    *out = nullptr;
    if (cfg->slots < 1 || cfg->slots > QLP_SERVER_MAX_SLOTS)
        return fail(QLP_ERR_ARG, "server_create: slots must be 1..16");
    if (cfg->hbm_budget == 0) return fail(QLP_ERR_ARG, "server_create: the HBM budget must be set");
    qlp_server *s = new qlp_server();
    s->device = hip_device;
    s->budget = cfg->hbm_budget;
    s->t0 = Clock::now();
    s->max_active = cfg->batch_requests > 0 ? cfg->batch_requests : 512;
    if (int rc = qlp_ctx_create(hip_device, 0, &s->ectx)) {
        delete s;
        return rc;
    }
    s->slots.resize(cfg->slots);
    for (Slot &sl : s->slots) {
        int rc = qlp_ctx_create(hip_device, 0, &sl.ctx);
        if (!rc) rc = check_hip(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "server event");
        if (rc) {
            for (Slot &o : s->slots) {
                if (o.done) (void)hipEventDestroy(o.done);
                if (o.ctx) qlp_ctx_destroy(o.ctx);
            }
            qlp_ctx_destroy(s->ectx);
            delete s;
            return rc;
        }
    }
    s->worker = std::thread(worker_loop, s);
    *out = s;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_server_destroy(qlp_server *s)
try {
    if (!s) return QLP_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (Slot &sl : s->slots) {
        for (auto &e : sl.plans) qlp_piv_plan_destroy(e.second.plan);
        if (sl.dA) (void)hipFree(sl.dA);
        if (sl.dB) (void)hipFree(sl.dB);
        if (sl.hu) (void)hipHostFree(sl.hu);
        if (sl.hv) (void)hipHostFree(sl.hv);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.ctx) qlp_ctx_destroy(sl.ctx);
    }
    (void)hipStreamSynchronize(estream(s));
    for (auto &e : s->engines) engine_destroy(e.eng);
    for (auto &g : s->geos)
        for (GeoLevel &L : g.second->lv) {
            (void)hipFree(L.top);
            (void)hipFree(L.left);
        }
    for (void *p : s->pinned_slabs) (void)hipHostFree(p);
    for (auto &rec : s->inflight)
        if (rec.ev) (void)hipEventDestroy(rec.ev);
    for (hipEvent_t e : s->step_events_free) (void)hipEventDestroy(e);
    qlp_ctx_destroy(s->ectx);
    delete s;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_server_submit(qlp_server *s, const qlp_request *req, uint64_t *ticket)
try {
    if (!s || !req || !ticket) return fail(QLP_ERR_ARG, "server_submit: null argument");
    if (!req->imgA || !req->imgB || req->H < 1 || req->W < 1)
        return fail(QLP_ERR_ARG, "server_submit: bad frames");
    size_t plan_need = 0;
    if (int rc = qlp_piv_plan_estimate(req->H, req->W, 1, &req->cfg, &plan_need)) return rc;
    if (plan_need + 2 * (size_t)req->H * req->W > s->budget)
        return fail(QLP_ERR_BUDGET, "server_submit: the request alone exceeds the HBM budget");
    auto r = std::make_shared<ServeRequest>();
    r->req = *req;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return fail(QLP_ERR_ARG, "server_submit: server is stopping");
        r->ticket = s->next_ticket++;
        s->pending.push_back(r);
        s->all[r->ticket] = r;
    }
    s->cv_work.notify_one();
    *ticket = r->ticket;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_server_wait(qlp_server *s, uint64_t ticket, int timeout_ms, int *done)
try {
    if (!s || !done) return fail(QLP_ERR_ARG, "server_wait: null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(QLP_ERR_ARG, "server_wait: unknown ticket");
    auto r = it->second;
    auto finished = [&] { return r->status >= 2; };
    if (timeout_ms < 0)
        s->cv_done.wait(lk, finished);
    else
        s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), finished);
    *done = finished() ? 1 : 0;
    if (r->status == 3) return fail(r->rc, r->err);
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_server_result(qlp_server *s, uint64_t ticket, const float **u, const float **v, int *nI,
                      int *nJ, int *ia, qlp_request_times *times)
try {
    if (!s) return fail(QLP_ERR_ARG, "server_result: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(QLP_ERR_ARG, "server_result: unknown ticket");
    ServeRequest &r = *it->second;
    if (r.status == 3) return fail(r.rc, r.err);
    if (r.status != 2) return fail(QLP_ERR_ARG, "server_result: request not finished");
    if (u) *u = r.u.data();
    if (v) *v = r.v.data();
    if (nI) *nI = r.nI;
    if (nJ) *nJ = r.nJ;
    if (ia) *ia = r.ia;
    if (times) {
        times->submit_ms = ms_since(s->t0, r.t_submit);
        times->admit_ms = ms_since(s->t0, r.t_admit);
        times->done_ms = ms_since(s->t0, r.t_done);
        times->ias = (double)r.ias;
    }
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_server_release(qlp_server *s, uint64_t ticket)
try {
    if (!s) return fail(QLP_ERR_ARG, "server_release: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(QLP_ERR_ARG, "server_release: unknown ticket");
    if (it->second->status < 2) return fail(QLP_ERR_ARG, "server_release: request not finished");
    s->all.erase(it);
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

int qlp_server_get_stats(qlp_server *s, qlp_server_stats *st)
try {
    if (!s || !st) return fail(QLP_ERR_ARG, "server_stats: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    st->completed = s->completed;
    st->failed = s->failed;
    st->pending = s->pending.size();
    st->evictions = s->evictions;
    st->hbm_in_use = s->in_use;
    st->hbm_peak = s->peak;
    st->ias_done = (double)s->ias_done;
    int running = (int)s->active.size();
    for (const Slot &sl : s->slots) running += sl.cur ? 1 : 0;
    st->running = running;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

}  // extern "C"
