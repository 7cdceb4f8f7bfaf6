// The serving form: a request loop over frame pairs of mixed sizes.
//
// The reference processes frame pairs with a pool of worker threads, one
// correlation job (and device queue) per thread (InputFiles.java:172-224,
// OpenClGpuManager.java:124-142).  Here one process owns the GPU and a server
// object runs the requests:
//   - slots: each slot is a library context (its own HIP stream, twiddle tables
//     and workspaces) plus device frame buffers and a cache of multi-pass plans
//     keyed by the request's frame size and configuration; requests on
//     different slots overlap on the device;
//   - admission: requests wait in arrival order and take a free slot as soon as
//     the HBM they need (frames, plus the plan unless the slot holds it
//     already) fits the server's budget; idle slots' cached plans are evicted,
//     least recently used first, to make room;
//   - a worker thread admits, launches (frames to the device, the multi-pass
//     run, the last level's field back to pinned host memory, one event) and
//     retires completed requests; callers submit, wait and read results.
// Frames, twiddles and the displacement state of every level stay on the device
// for the whole multi-pass run of a request.
#include <hip/hip_runtime.h>
#include <string.h>

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
        if (progress) s->cv_done.notify_all();
        // admit in arrival order while a slot is free and the head request fits
        while (!s->pending.empty()) {
            Slot *free_slot = nullptr;
            const std::shared_ptr<ServeRequest> r = s->pending.front();
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
        if (s->stop && s->pending.empty()) {
            bool busy = false;
            for (Slot &sl : s->slots) busy |= (bool)sl.cur;
            if (!busy) break;
        }
        if (!progress) {
            bool busy = false;
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
    s->slots.resize(cfg->slots);
    for (Slot &sl : s->slots) {
        int rc = qlp_ctx_create(hip_device, 0, &sl.ctx);
        if (!rc) rc = check_hip(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "server event");
        if (rc) {
            for (Slot &o : s->slots) {
                if (o.done) (void)hipEventDestroy(o.done);
                if (o.ctx) qlp_ctx_destroy(o.ctx);
            }
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
    int running = 0;
    for (const Slot &sl : s->slots) running += sl.cur ? 1 : 0;
    st->running = running;
    return QLP_OK;
} catch (...) {
    return qlpiv::fail_exception();
}

}  // extern "C"
