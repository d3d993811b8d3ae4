// mw_server.hip -- request server for likelihood evaluations.
//
// Wedges (geometry tables and star lists) are made resident once.  Requests
// (a wedge, a parameter vector, a precision) are queued as they arrive; a
// dispatcher thread repeatedly forms a batch of one precision class from the
// pending requests (the oldest request's class, oldest first within it),
// admitted while the in-flight scratch fits the HBM budget and a runner (a
// stream with its own buffers) is free, and launches it.  Up to max_inflight
// batches run at once on separate streams, so requests that arrive while
// others are on the GPU are admitted as runners and memory free up, and
// their work overlaps on the device.  FP64 batches, far slower per
// evaluation, are bounded by their estimated work (kFp64BatchPoints
// integrand evaluations, about a tenth of a second, and at least one
// request), so no batch holds the device for long.  An FP64 batch runs its
// persistent grid at no more than 2 workgroups per CU, which leaves wave
// slots on every CU to concurrent fast batches (the VALU-bound integral
// kernel loses about 5% at 2 workgroups per CU against 5).  While fast
// requests are waiting or running, FP64 batches hold at most half the
// runners; with none, they take every runner, so that an FP64 backlog under
// overload keeps the device full.  Each request's result is bitwise the
// same alone and in any mix (per-request fixed-order reductions).
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>

#include "mw_batch.h"

using namespace mw;

namespace {

using Clock = std::chrono::steady_clock;

// Integrand evaluations per FP64 batch (cells x convolution points summed
// over its requests): about 0.1 s of the FP64 kernel at 3 streams.
constexpr double kFp64BatchPoints = 4e10;

double now_s()
{
    return std::chrono::duration<double>(Clock::now().time_since_epoch()).count();
}

struct Request {
    long id;
    int wedge;
    int precision;
    EvalConst k;
    double t_submit;
};

struct Done {
    mw_result result;
    double t_submit, t_dispatch, t_done;
    int status;
};

struct InFlight {
    std::unique_ptr<BatchRunner> runner;
    std::vector<Request> reqs;
    size_t scratch = 0;
    double t_dispatch = 0;
    bool busy = false;
};

}  // namespace

struct mw_server {
    mw_server_config cfg{};
    int device = 0;
    std::vector<std::unique_ptr<mw_ctx>> wedges;
    size_t resident = 0, inflight_scratch = 0;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request> pending;
    std::unordered_map<long, Done> done;
    std::unordered_set<long> outstanding;    // submitted and not yet collected
    std::vector<InFlight> slots;
    long next_id = 1;
    long n_batches = 0, n_completed = 0, max_batch_seen = 0;
    int error = MW_OK;
    std::string error_msg;
    bool stop = false;
    std::thread dispatcher;

    void run();
    bool launch_locked(InFlight& f, std::unique_lock<std::mutex>& lk);
};

bool mw_server::launch_locked(InFlight& f, std::unique_lock<std::mutex>& lk)
{
    // Admission.  The batch takes one precision class: that of the oldest
    // pending request, except that FP64 batches (tens to hundreds of times
    // slower per evaluation) may hold at most half the runners (at least
    // one), so that faster requests always find one.  Within the class, requests go
    // in arrival order, up to max_batch vectors, while the batch's scratch
    // fits the budget; requests of other classes keep their places.
    int fp64_busy = 0;
    bool fast_running = false, fast_waiting = false;
    for (const InFlight& g : slots) {
        if (!g.busy || g.reqs.empty()) continue;
        if (g.reqs.front().precision == MW_PREC_FP64) fp64_busy++;
        else fast_running = true;
    }
    for (const Request& r : pending) fast_waiting |= r.precision != MW_PREC_FP64;
    const bool fast = fast_running || fast_waiting;
    const bool allow_fp64 = cfg.max_inflight == 1 || !fast || fp64_busy < std::max(1, cfg.max_inflight / 2);
    int cls = -1;
    for (const Request& r : pending) {
        if (r.precision == MW_PREC_FP64 && !allow_fp64) continue;
        cls = r.precision;
        break;
    }
    if (cls < 0) return false;
    std::vector<Request> take;
    std::vector<Slot> slots_v;
    std::deque<Request> rest;
    size_t scratch = 0;
    const size_t budget = cfg.hbm_budget_bytes;
    bool full = false;
    double points = 0.0;
    for (const Request& r : pending) {
        if (full || r.precision != cls || (int)take.size() >= cfg.max_batch_vectors) {
            rest.push_back(r);
            continue;
        }
        if (cls == MW_PREC_FP64) {
            double pts = 0.0;
            for (const auto& rd : wedges[r.wedge]->regions)
                pts += (double)rd->geom.r_steps * rd->geom.mu_steps * rd->geom.nu_steps * rd->n_conv;
            if (!take.empty() && points + pts > kFp64BatchPoints) {
                full = true;   // the rest of the class waits for the next batch
                rest.push_back(r);
                continue;
            }
            points += pts;
        }
        std::vector<Slot> trial = slots_v;
        trial.push_back(Slot{ wedges[r.wedge].get(), r.k });
        const size_t need = BatchRunner::scratch_bytes_for(trial);
        if (budget && resident + inflight_scratch + need > budget) {
            full = true;   // later requests of the class wait behind this one
            rest.push_back(r);
            continue;
        }
        slots_v.swap(trial);
        scratch = need;
        take.push_back(r);
    }
    if (take.empty()) return false;
    pending.swap(rest);
    f.reqs = std::move(take);
    f.scratch = scratch;
    f.busy = true;
    f.t_dispatch = now_s();
    inflight_scratch += scratch;
    n_batches++;
    max_batch_seen = std::max<long>(max_batch_seen, (long)f.reqs.size());
    const int precision = f.reqs.front().precision;
    lk.unlock();
    // An FP64 batch's persistent grid takes at most 2 workgroups per CU,
    // leaving wave slots on every CU to fast batches.
    const int bpc = cfg.blocks_per_cu > 0 ? cfg.blocks_per_cu : 2;
    f.runner->set_variant(cfg.variant, precision == MW_PREC_FP64 ? std::min(bpc, 2) : bpc);
    int rc = hipSetDevice(device) == hipSuccess ? MW_OK : MW_E_HIP;
    if (!rc) rc = f.runner->plan(slots_v, precision);
    if (!rc) rc = f.runner->enqueue();
    lk.lock();
    if (rc) {
        error = rc;
        error_msg = mw_last_error();
        for (const Request& r : f.reqs) {
            Done d{};
            d.status = rc;
            d.t_submit = r.t_submit;
            d.t_dispatch = f.t_dispatch;
            d.t_done = now_s();
            done[r.id] = d;
        }
        inflight_scratch -= f.scratch;
        f.busy = false;
        f.reqs.clear();
        cv_done.notify_all();
    }
    return true;
}

void mw_server::run()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    std::vector<mw_result> res;
    for (;;) {
        bool progressed = false;
        // Complete finished batches.
        for (InFlight& f : slots) {
            if (!f.busy) continue;
            lk.unlock();
            const bool fin = f.runner->query_done();
            lk.lock();
            if (!fin) continue;
            res.assign(f.reqs.size(), mw_result{});
            lk.unlock();
            const int rc = f.runner->finish(res.data());
            const double t = now_s();
            lk.lock();
            for (size_t i = 0; i < f.reqs.size(); i++) {
                Done d;
                d.result = res[i];
                d.status = rc;
                d.t_submit = f.reqs[i].t_submit;
                d.t_dispatch = f.t_dispatch;
                d.t_done = t;
                done[f.reqs[i].id] = d;
            }
            n_completed += (long)f.reqs.size();
            inflight_scratch -= f.scratch;
            f.busy = false;
            f.reqs.clear();
            progressed = true;
            cv_done.notify_all();
        }
        // Admit pending requests onto free runners.
        for (InFlight& f : slots) {
            if (f.busy || pending.empty()) continue;
            if (launch_locked(f, lk)) progressed = true;
        }
        bool any_busy = false;
        for (const InFlight& f : slots) any_busy |= f.busy;
        if (stop && !any_busy && pending.empty()) break;
        if (!progressed) {
            if (any_busy)
                cv_work.wait_for(lk, std::chrono::microseconds(50));
            else
                cv_work.wait(lk, [&] { return stop || !pending.empty(); });
        }
    }
}

extern "C" {

int mw_server_create(const mw_server_config* cfg, mw_server** out)
{
    if (!cfg || !out) return fail(MW_E_INVALID, "null argument");
    *out = nullptr;
    if (mw_device_count() <= 0) return fail(MW_E_NODEVICE, "no GPU visible");
    if (cfg->max_inflight < 1 || cfg->max_inflight > 16 || cfg->max_batch_vectors < 1)
        return fail(MW_E_INVALID, "max_inflight must be 1-16 and max_batch_vectors >= 1");
    if (cfg->variant < 0 || cfg->variant >= persistent_variants() || cfg->blocks_per_cu < 0 ||
        cfg->blocks_per_cu > 16)
        return fail(MW_E_INVALID, "bad kernel variant");
    auto s = std::make_unique<mw_server>();
    s->cfg = *cfg;
    MW_HIP_CHECK(hipGetDevice(&s->device));
    s->slots.resize(cfg->max_inflight);
    for (InFlight& f : s->slots) {
        f.runner = std::make_unique<BatchRunner>();
        int rc = f.runner->init(s->device, 0);
        if (rc) return rc;
        f.runner->set_variant(cfg->variant, cfg->blocks_per_cu > 0 ? cfg->blocks_per_cu : 2);
    }
    s->dispatcher = std::thread([p = s.get()] { p->run(); });
    *out = s.release();
    return MW_OK;
}

int mw_server_add_wedge(mw_server* s, const ASTRONOMY_PARAMETERS* ap, const STAR_POINTS* sp,
                        int* wedge_id)
{
    if (!s || !ap || !wedge_id) return fail(MW_E_INVALID, "null argument");
    if (ap->number_streams > kMaxStreams)
        return fail(MW_E_INVALID, "the request server takes models of at most 4 streams");
    mw_ctx* c = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->cfg.hbm_budget_bytes && s->resident >= s->cfg.hbm_budget_bytes)
            return fail(MW_E_NOMEM, "HBM budget exhausted by resident wedges");
    }
    int rc = mw_ctx_create(ap, sp, 0, &c);
    if (rc) return rc;
    const size_t bytes = mw_ctx_resident_bytes(c);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->cfg.hbm_budget_bytes && s->resident + bytes > s->cfg.hbm_budget_bytes) {
        mw_ctx_destroy(c);
        return fail(MW_E_NOMEM, "wedge does not fit the HBM budget");
    }
    s->resident += bytes;
    s->wedges.emplace_back(c);
    *wedge_id = (int)s->wedges.size() - 1;
    return MW_OK;
}

int mw_server_submit(mw_server* s, int wedge_id, const double* params, int precision, long* request_id)
{
    if (!s || !request_id) return fail(MW_E_INVALID, "null argument");
    if (precision != MW_PREC_FP64 && precision != MW_PREC_MIXED && precision != MW_PREC_RESAMPLED)
        return fail(MW_E_INVALID, "unknown precision");
    mw_ctx* c;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (wedge_id < 0 || wedge_id >= (int)s->wedges.size()) return fail(MW_E_INVALID, "unknown wedge");
        c = s->wedges[wedge_id].get();
    }
    if (c->n_params > 0 && !params) return fail(MW_E_INVALID, "null parameters");
    Request r;
    r.wedge = wedge_id;
    r.precision = precision;
    int rc = slot_constants(c, params, &r.k);
    if (rc) return rc;
    // A request that could never fit the budget is refused now.
    std::vector<Slot> one{ Slot{ c, r.k } };
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->cfg.hbm_budget_bytes && s->resident + BatchRunner::scratch_bytes_for(one) > s->cfg.hbm_budget_bytes)
        return fail(MW_E_NOMEM, "request does not fit the HBM budget");
    r.id = s->next_id++;
    r.t_submit = now_s();
    s->pending.push_back(r);
    s->outstanding.insert(r.id);
    *request_id = r.id;
    s->cv_work.notify_one();
    return MW_OK;
}

static int take_done(mw_server* s, long id, mw_result* out, mw_request_times* times)
{
    auto it = s->done.find(id);
    const Done& d = it->second;
    const int rc = d.status;
    if (out) *out = d.result;
    if (times) {
        times->submit_s = d.t_submit;
        times->dispatch_s = d.t_dispatch;
        times->done_s = d.t_done;
    }
    s->done.erase(it);
    s->outstanding.erase(id);
    return rc ? fail(rc, "request failed: " + s->error_msg) : MW_OK;
}

int mw_server_poll(mw_server* s, long request_id, mw_result* out, mw_request_times* times, int* finished)
{
    if (!s || !finished) return fail(MW_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *finished = 0;
    if (!s->outstanding.count(request_id)) return fail(MW_E_INVALID, "unknown or already collected request");
    if (s->done.find(request_id) == s->done.end()) return MW_OK;
    *finished = 1;
    return take_done(s, request_id, out, times);
}

int mw_server_wait(mw_server* s, long request_id, mw_result* out, mw_request_times* times)
{
    if (!s) return fail(MW_E_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (!s->outstanding.count(request_id)) return fail(MW_E_INVALID, "unknown or already collected request");
    s->cv_done.wait(lk, [&] { return s->done.count(request_id) > 0; });
    return take_done(s, request_id, out, times);
}

int mw_server_get_stats(mw_server* s, mw_server_stats* out)
{
    if (!s || !out) return fail(MW_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out->resident_bytes = s->resident;
    out->inflight_scratch_bytes = s->inflight_scratch;
    out->pending = (long)s->pending.size();
    out->batches = s->n_batches;
    out->completed = s->n_completed;
    out->largest_batch = s->max_batch_seen;
    int busy = 0;
    for (const InFlight& f : s->slots) busy += f.busy ? 1 : 0;
    out->inflight_batches = busy;
    return MW_OK;
}

void mw_server_destroy(mw_server* s)
{
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->dispatcher.joinable()) s->dispatcher.join();
    delete s;
}

}  // extern "C"
