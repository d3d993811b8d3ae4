// mw_batch.hip -- many parameter vectors over several resident wedges of
// uneven size in one persistent launch, optionally captured once as a HIP
// graph and replayed with new parameter values.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "mw_batch.h"

using namespace mw;

namespace mw {

BatchRunner::~BatchRunner()
{
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
}

int BatchRunner::init(int device, int flags)
{
    device_ = device;
    flags_ = flags;
    MW_HIP_CHECK(hipSetDevice(device));
    MW_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    for (auto& e : ev_) MW_HIP_CHECK(hipEventCreate(&e));
    MW_HIP_CHECK(hipDeviceGetAttribute(&n_cu_, hipDeviceAttributeMultiprocessorCount, device));
    if (n_cu_ <= 0) n_cu_ = 256;
    int rc;
    if ((rc = counters_.reserve(8))) return rc;
    return MW_OK;
}

// Quadrature tables of region rd under the model k, as one number: level,
// layout and background table.
int BatchRunner::table_choice(const RegionDev& rd, const EvalConst& k) const
{
    const GridPlan g = plan_grid(rd, k, resample_);
    int key = ((g.level + 1) * kTableLayouts + g.layout) * (kBgLevels + 1) + g.bg_level + 1;
    for (int s = 0; s < kMaxStreams; s++) key = key * (kQuadLevels + 1) + g.st_level[s] + 1;
    return key;
}

// The star pass's background table under the model k (resampled mixed
// batches of the Hernquist profile), or -1 for the exact convolution.
int BatchRunner::star_level(const mw_ctx& c, const EvalConst& k) const
{
    return (resample_ && mixed_ && k.bg_mode == kBgHernquist) ? star_bg_level(c, k.q, k.r0) : -1;
}

bool BatchRunner::same_modes(const std::vector<Slot>& slots) const
{
    if (slots.size() != slot_bg_.size()) return false;
    size_t li = 0;
    for (size_t s = 0; s < slots.size(); s++) {
        if (slots[s].k.bg_mode != slot_bg_[s]) return false;
        // The planned work items point at one quadrature table per region and
        // carry its layout (recurrence or not, background stride), all of
        // which follow from the model.
        for (const auto& rd : slots[s].ctx->regions)
            if (li >= levels_.size() || table_choice(*rd, slots[s].k) != levels_[li++]) return false;
        // and the star pass's background table
        if (li >= levels_.size() || star_level(*slots[s].ctx, slots[s].k) != levels_[li++]) return false;
    }
    return li == levels_.size();
}

size_t BatchRunner::scratch_bytes_for(const std::vector<Slot>& slots)
{
    const size_t NI = integral_comps(), NS = star_comps();
    size_t b = 0;
    for (const Slot& s : slots) {
        b += sizeof(EvalConst) + sizeof(StarLaunch) + 2 * sizeof(ReduceItem) + 2 * sizeof(int);
        for (const auto& r : s.ctx->regions)
            b += sizeof(RegionLaunch) + sizeof(ReduceItem) + (size_t)r->blocks * r->r_chunks * NI * 2 * 8 +
                 NI * 8;
        b += (size_t)((s.ctx->n_stars + kBlock - 1) / kBlock) * NS * 2 * 8 + NS * 8;
    }
    return b;
}

size_t BatchRunner::scratch_bytes() const
{
    return consts_.held + ritems_.held + sitems_.held + rred_.held + sred_.held + ipart_.held +
           spart_.held + irows_.held + srows_.held + first_.held + count_.held + counters_.held;
}

int BatchRunner::plan(const std::vector<Slot>& slots, int precision)
{
    const int NI = integral_comps(), NS = star_comps();
    const int n = (int)slots.size();
    if (n < 1) return fail(MW_E_INVALID, "empty batch");
    mixed_ = precision != MW_PREC_FP64;
    resample_ = precision == MW_PREC_RESAMPLED;
    levels_.clear();
    points_ = 0;
    points_rec_ = 0;
    points_bg_ = 0;
    terms_ = 0;
    slot_bg_.assign(n, 0);
    slot_n_regions_.assign(n, 0);
    slot_first_row_.assign(n, 0);
    slot_S_.assign(n, 0);
    slot_stars_.assign(n, 0);

    // Rows of region results: slot-major, region order within a slot.
    struct It {
        RegionLaunch L;
        double cost;
    };
    std::vector<It> items;
    int row = 0;
    for (int s = 0; s < n; s++) {
        mw_ctx* c = slots[s].ctx;
        if (c->device != device_) return fail(MW_E_INVALID, "batch mixes devices");
        slot_bg_[s] = slots[s].k.bg_mode;
        slot_S_[s] = slots[s].k.n_streams;
        slot_first_row_[s] = row;
        slot_n_regions_[s] = (int)c->regions.size();
        slot_stars_[s] = c->n_stars;
        for (const auto& rd : c->regions) {
            TableRef tr;
            const GridPlan gp = plan_grid(*rd, slots[s].k, resample_);
            const int level = gp.level, layout = gp.layout;
            int rc0 = region_tables(*rd, gp, &tr);
            if (rc0) return rc0;
            levels_.push_back(table_choice(*rd, slots[s].k));
            points_ += rd->n_cols * tr.nodes;
            RegionLaunch L;
            std::memset(&L, 0, sizeof(L));
            L.rq = tr.rq;
            L.rstep = tr.rstep;
            L.col = rd->col.p;
            L.n_cols = rd->n_cols;
            L.r_steps = tr.r_steps;
            L.n_conv = tr.n_conv;
            L.calc_begin = 0;
            L.calc_end = rd->n_cols * (long)tr.r_steps;
            L.slot = s;
            L.out_index = row++;
            // r-windows on the reference's tables only; resampled grids run whole columns
            L.r_chunks = level < 0 ? rd->r_chunks : 1;
            L.block_count = rd->blocks * L.r_chunks;
            L.n_streams = slots[s].k.n_streams;
            L.r_lo = rd->r_min;
            L.r_hi = rd->r_max;
            L.grid_step = layout > 0 ? recurrence_step(level, slots[s].k) : 0.0;
            L.grid_lo = tr.lo;
            L.bg_stride = layout_bg_stride(layout);
            L.bg_count = (int)tr.bg_count;
            L.bg_rq = tr.bg_rq;
            copy_stream_tables(tr, &L);
            const long terms = layout == kLayoutSplit ? tr.stream_terms : tr.nodes * L.n_streams;
            if (L.grid_step > 0.0) {
                points_rec_ += rd->n_cols * tr.nodes;
                terms_ += rd->n_cols * terms;
            }
            points_bg_ += rd->n_cols * tr.bg_nodes;
            // cost per block: stream terms and background evaluations (about
            // 6.5 stream terms each), or per point on the other layouts
            const double cost = layout == kLayoutSplit ? terms + 6.5 * tr.bg_count
                                                       : (double)L.r_steps / L.r_chunks * L.n_conv * (2.0 + L.n_streams);
            items.push_back(It{ L, cost });
        }
        levels_.push_back(star_level(*c, slots[s].k));
    }
    n_ritems_ = (int)items.size();
    // Group by background mode, largest cost per block first within a group.
    std::stable_sort(items.begin(), items.end(), [&](const It& a, const It& b) {
        const int ba = slot_bg_[a.L.slot], bb = slot_bg_[b.L.slot];
        if (ba != bb) return ba < bb;
        return a.cost > b.cost;
    });
    int rc;
    if ((rc = ritems_.reserve(n_ritems_)) || (rc = rred_.reserve(n_ritems_)) ||
        (rc = irows_.reserve((size_t)row * NI)) || (rc = h_irows_.reserve((size_t)row * NI)) ||
        (rc = consts_.reserve(n)) || (rc = h_consts_.reserve(n)) || (rc = first_.reserve(n)) ||
        (rc = count_.reserve(n)) || (rc = sitems_.reserve(n)) || (rc = sred_.reserve(n)) ||
        (rc = srows_.reserve((size_t)n * NS)) || (rc = h_srows_.reserve((size_t)n * NS)))
        return rc;
    std::vector<RegionLaunch> hl(n_ritems_);
    std::vector<ReduceItem> hr(n_ritems_);
    igroups_.clear();
    total_iblocks_ = 0;
    for (int i = 0; i < n_ritems_; i++) {
        RegionLaunch L = items[i].L;
        const int bg = slot_bg_[L.slot];
        if (igroups_.empty() || igroups_.back().bg != bg)
            igroups_.push_back(Group{ bg, i, 0, total_iblocks_, 0 });
        Group& g = igroups_.back();
        g.split = g.split && mixed_ && L.grid_step > 0.0 && L.bg_stride == 0;
        g.exact = g.exact && !(L.grid_step > 0.0);
        L.block_begin = g.n_blocks;
        hl[i] = L;
        hr[i] = ReduceItem{ total_iblocks_, L.block_count, L.out_index, 0 };
        g.n_items++;
        g.n_blocks += L.block_count;
        total_iblocks_ += L.block_count;
    }
    if ((int)igroups_.size() > (int)counters_.cap) return fail(MW_E_INVALID, "too many groups");
    if ((rc = ipart_.reserve((size_t)std::max(total_iblocks_, 1) * NI * 2))) return rc;

    // Star items, one per slot, grouped the same way.
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (slot_bg_[a] != slot_bg_[b]) return slot_bg_[a] < slot_bg_[b];
        return slot_stars_[a] * slots[a].ctx->n_conv > slot_stars_[b] * slots[b].ctx->n_conv;
    });
    std::vector<StarLaunch> hs(n);
    std::vector<ReduceItem> hsr(n);
    sgroups_.clear();
    total_sblocks_ = 0;
    for (int oi = 0; oi < n; oi++) {
        const int s = order[oi];
        mw_ctx* c = slots[s].ctx;
        const int bg = slot_bg_[s];
        const int blocks = c->n_stars > 0 ? (int)((c->n_stars + kBlock - 1) / kBlock) : 0;
        if (sgroups_.empty() || sgroups_.back().bg != bg)
            sgroups_.push_back(Group{ bg, oi, 0, total_sblocks_, 0 });
        Group& g = sgroups_.back();
        StarLaunch L;
        std::memset(&L, 0, sizeof(L));
        L.chain_flags = 3;                 // batches take models of at most 4 streams
        L.star = c->star.p;
        L.n_stars = c->n_stars;
        L.star_begin = 0;
        L.slot = s;
        L.n_conv = c->n_conv;
        L.block_begin = g.n_blocks;
        L.block_count = blocks;
        L.out_index = s;
        L.n_streams = slots[s].k.n_streams;
        L.ratio = c->ratio.p;
        L.kw = c->kw.p;
        L.kw3 = c->kw3.p;
        L.coarse_pr = c->star_coarse_pr;
        int rc1;
        if ((rc1 = star_bg_table(*c, star_level(*c, slots[s].k), &L.bg_rq, &L.bg_count))) return rc1;
        hs[oi] = L;
        hsr[oi] = ReduceItem{ total_sblocks_, blocks, s, 0 };
        g.n_items++;
        g.n_blocks += blocks;
        total_sblocks_ += blocks;
    }
    if ((rc = spart_.reserve((size_t)std::max(total_sblocks_, 1) * NS * 2))) return rc;

    std::vector<int> fr(n), cnt(n);
    for (int s = 0; s < n; s++) {
        fr[s] = slot_first_row_[s];
        cnt[s] = slot_n_regions_[s];
    }
    // Work items go up through one pinned staging area, asynchronously on the
    // runner's stream, ahead of the launches that read them: a plan never
    // waits for work that other runners have in flight.  The staging area is
    // reused only by the next plan, after this runner's batch has finished.
    struct Part {
        void* dst;
        const void* src;
        size_t bytes;
    };
    const Part parts[] = {
        { ritems_.p, hl.data(), sizeof(RegionLaunch) * n_ritems_ },
        { rred_.p, hr.data(), sizeof(ReduceItem) * n_ritems_ },
        { sitems_.p, hs.data(), sizeof(StarLaunch) * n },
        { sred_.p, hsr.data(), sizeof(ReduceItem) * n },
        { first_.p, fr.data(), sizeof(int) * n },
        { count_.p, cnt.data(), sizeof(int) * n },
    };
    size_t total = 0;
    for (const Part& q : parts) total += (q.bytes + 255) / 256 * 256;
    if ((rc = h_plan_.reserve(total))) return rc;
    size_t off = 0;
    for (const Part& q : parts) {
        std::memcpy(h_plan_.p + off, q.src, q.bytes);
        MW_HIP_CHECK(hipMemcpyAsync(q.dst, h_plan_.p + off, q.bytes, hipMemcpyHostToDevice, stream_));
        off += (q.bytes + 255) / 256 * 256;
    }
    set_constants(slots);
    return MW_OK;
}

void BatchRunner::set_constants(const std::vector<Slot>& slots)
{
    for (size_t s = 0; s < slots.size(); s++) h_consts_.p[s] = slots[s].k;
}

int BatchRunner::variant_scratch(int variant, size_t* bytes) const
{
    size_t most = 0;
    for (const Group& g : igroups_) {
        size_t b = 0;
        MW_HIP_CHECK(persistent_scratch_bytes(g.bg, mixed_ && g.bg == kBgHernquist,
                                              g.bg == kBgHernquist ? variant : 0,
                                              g.split ? kLaunchSplitOnly : g.exact ? kLaunchExactOnly : kLaunchAny,
                                              &b));
        most = std::max(most, b);
    }
    *bytes = most;
    return MW_OK;
}

int BatchRunner::enqueue(bool kernel_events)
{
    const int NI = integral_comps(), NS = star_comps();
    const int n = n_slots();
    hipStream_t st = stream_;
    MW_HIP_CHECK(hipMemcpyAsync(consts_.p, h_consts_.p, sizeof(EvalConst) * n, hipMemcpyHostToDevice, st));
    MW_HIP_CHECK(launch_reset_counters(counters_.p, (int)igroups_.size(), st));
    const int grid = n_cu_ * std::max(1, blocks_per_cu_);
    if (kernel_events) MW_HIP_CHECK(hipEventRecord(ev_[2], st));
    for (size_t gi = 0; gi < igroups_.size(); gi++) {
        const Group& g = igroups_[gi];
        const bool mixed = mixed_ && g.bg == kBgHernquist;
        const int variant = g.bg == kBgHernquist ? variant_ : 0;
        MW_HIP_CHECK(launch_persistent(g.bg, mixed, variant, std::min(grid, g.n_blocks),
                                       ritems_.p + g.first_item, g.n_items, g.n_blocks, consts_.p,
                                       ipart_.p + (size_t)g.first_block * NI * 2,
                                       counters_.p + gi, st,
                                       g.split ? kLaunchSplitOnly : g.exact ? kLaunchExactOnly : kLaunchAny));
    }
    if (kernel_events) MW_HIP_CHECK(hipEventRecord(ev_[3], st));
    MW_HIP_CHECK(launch_reduce(false, rred_.p, n_ritems_, ipart_.p, irows_.p, st));
    MW_HIP_CHECK(launch_finalize(first_.p, count_.p, irows_.p, consts_.p, n, st));
    for (const Group& g : sgroups_) {
        MW_HIP_CHECK(launch_star_any(g.bg, mixed_ && g.bg == kBgHernquist, sitems_.p + g.first_item,
                                     g.n_items, g.n_blocks, consts_.p,
                                     spart_.p + (size_t)g.first_block * NS * 2, st));
    }
    if (total_sblocks_ > 0) MW_HIP_CHECK(launch_reduce(true, sred_.p, n, spart_.p, srows_.p, st));
    const int rows = slot_first_row_.back() + slot_n_regions_.back();
    MW_HIP_CHECK(hipMemcpyAsync(h_irows_.p, irows_.p, sizeof(double) * rows * NI, hipMemcpyDeviceToHost, st));
    if (total_sblocks_ > 0)
        MW_HIP_CHECK(hipMemcpyAsync(h_srows_.p, srows_.p, sizeof(double) * n * NS, hipMemcpyDeviceToHost, st));
    return MW_OK;
}

int BatchRunner::record_start()
{
    MW_HIP_CHECK(hipEventRecord(ev_[0], stream_));
    return MW_OK;
}

int BatchRunner::record_end()
{
    MW_HIP_CHECK(hipEventRecord(ev_[1], stream_));
    return MW_OK;
}

double BatchRunner::elapsed_ms()
{
    float ms = 0;
    if (hipEventElapsedTime(&ms, ev_[0], ev_[1]) != hipSuccess) return NAN;
    return ms;
}

double BatchRunner::integral_ms()
{
    float ms = 0;
    if (hipEventElapsedTime(&ms, ev_[2], ev_[3]) != hipSuccess) return NAN;
    return ms;
}

bool BatchRunner::query_done() { return hipStreamQuery(stream_) == hipSuccess; }

int BatchRunner::finish(mw_result* results)
{
    const int NI = integral_comps(), NS = star_comps();
    MW_HIP_CHECK(hipStreamSynchronize(stream_));
    if (!results) return MW_OK;
    for (int s = 0; s < n_slots(); s++) {
        mw_result& r = results[s];
        std::memset(&r, 0, sizeof(r));
        const int S = slot_S_[s];
        r.number_streams = S;
        const double* row = h_irows_.p + (size_t)slot_first_row_[s] * NI;
        r.background_integral = row[0];
        for (int j = 0; j < S; j++) r.stream_integrals[j] = row[1 + j];
        for (int g = 1; g < slot_n_regions_[s]; g++) {
            r.background_integral -= row[g * NI];
            for (int j = 0; j < S; j++) r.stream_integrals[j] -= row[g * NI + 1 + j];
        }
        const double N = (double)slot_stars_[s];
        if (slot_stars_[s] > 0) {
            const double* sr = h_srows_.p + (size_t)s * NS;
            r.num_zero = (int)sr[NS - 1];
            r.prob_sum = sr[0] - 238.0 * r.num_zero;
            r.likelihood = r.prob_sum / N;
            r.bg_only_likelihood = sr[1] / N;
            for (int j = 0; j < S; j++) r.stream_only_likelihood[j] = sr[2 + j] / N;
        } else {
            r.likelihood = NAN;
        }
    }
    return MW_OK;
}

// Builds the EvalConst of a parameter vector for a context.
int slot_constants(mw_ctx* c, const double* vec, EvalConst* k)
{
    ModelTemplate m;
    m.load(c->tmpl.ap);
    if (c->n_params > 0) apply_vector(m.ap, vec);
    std::string err;
    if (!make_eval_const(m.ap, k, &err)) return fail(MW_E_INVALID, err);
    return MW_OK;
}

}  // namespace mw

// ===========================================================================
// Batch C ABI
// ===========================================================================
struct mw_batch {
    int device = 0;
    int flags = 0;
    int precision = MW_PREC_FP64;
    std::vector<mw_ctx*> ctxs;
    std::vector<int> nvec;
    std::vector<Slot> slots;
    BatchRunner runner;
    bool planned = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    mw_timing timing{};

    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~mw_batch() { drop_graph(); }
};

extern "C" {

int mw_batch_variant_count(void) { return persistent_variants(); }

int mw_batch_create(mw_ctx* const* ctxs, const int* n_vectors, int n_ctxs, int precision, int flags,
                    mw_batch** out)
{
    if (!ctxs || !n_vectors || n_ctxs < 1 || !out) return fail(MW_E_INVALID, "bad arguments");
    if (precision != MW_PREC_FP64 && precision != MW_PREC_MIXED && precision != MW_PREC_RESAMPLED)
        return fail(MW_E_INVALID, "unknown precision");
    *out = nullptr;
    auto b = std::make_unique<mw_batch>();
    b->flags = flags;
    b->precision = precision;
    int total = 0;
    for (int i = 0; i < n_ctxs; i++) {
        if (!ctxs[i] || n_vectors[i] < 0) return fail(MW_E_INVALID, "bad context or vector count");
        if (ctxs[i]->device != ctxs[0]->device) return fail(MW_E_INVALID, "contexts on different devices");
        b->ctxs.push_back(ctxs[i]);
        b->nvec.push_back(n_vectors[i]);
        total += n_vectors[i];
    }
    if (total < 1) return fail(MW_E_INVALID, "empty batch");
    b->device = ctxs[0]->device;
    int rc = b->runner.init(b->device, flags);
    if (rc) return rc;
    // Slots start at each context's template parameters.
    for (int i = 0; i < n_ctxs; i++) {
        for (int v = 0; v < n_vectors[i]; v++) {
            Slot s{ ctxs[i], EvalConst{} };
            std::string err;
            if (!make_eval_const(ctxs[i]->tmpl.ap, &s.k, &err)) return fail(MW_E_INVALID, err);
            b->slots.push_back(s);
        }
    }
    *out = b.release();
    return MW_OK;
}

int mw_batch_vector_count(const mw_batch* b) { return b ? (int)b->slots.size() : -1; }

int mw_batch_set_variant(mw_batch* b, int variant, int blocks_per_cu)
{
    if (!b || variant < 0 || variant >= persistent_variants() || blocks_per_cu < 1 || blocks_per_cu > 16)
        return fail(MW_E_INVALID, "bad variant");
    b->runner.set_variant(variant, blocks_per_cu);
    b->drop_graph();
    return MW_OK;
}

int mw_batch_variant_scratch(const mw_batch* b, int variant, size_t* bytes)
{
    if (!b || !bytes || variant < 0 || variant >= persistent_variants()) return fail(MW_E_INVALID, "bad arguments");
    if (!b->planned) return fail(MW_E_INVALID, "the batch has not run yet: its kernels are chosen by its first run");
    MW_HIP_CHECK(hipSetDevice(b->device));
    return b->runner.variant_scratch(variant, bytes);
}

int mw_batch_run(mw_batch* b, const double* params, mw_result* results)
{
    if (!b || !results) return fail(MW_E_INVALID, "bad arguments");
    MW_HIP_CHECK(hipSetDevice(b->device));
    // Constants of every slot from its parameter vector.
    std::vector<Slot> slots = b->slots;
    size_t off = 0, si = 0;
    for (size_t c = 0; c < b->ctxs.size(); c++) {
        mw_ctx* ctx = b->ctxs[c];
        for (int v = 0; v < b->nvec[c]; v++, si++) {
            if (ctx->n_params > 0 && !params) return fail(MW_E_INVALID, "null parameters");
            int rc = slot_constants(ctx, params ? params + off : nullptr, &slots[si].k);
            if (rc) return rc;
            off += ctx->n_params;
        }
    }
    int rc;
    if (!b->planned || !b->runner.same_modes(slots)) {
        if ((rc = b->runner.plan(slots, b->precision))) return rc;
        b->planned = true;
        b->drop_graph();
    } else {
        b->runner.set_constants(slots);
    }
    hipStream_t st = b->runner.stream();
    if ((rc = b->runner.record_start())) return rc;
    if (b->flags & MW_BATCH_GRAPH) {
        if (!b->exec) {
            MW_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            rc = b->runner.enqueue();
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(st, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(MW_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            b->graph = g;
            MW_HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        }
        MW_HIP_CHECK(hipGraphLaunch(b->exec, st));
    } else {
        if ((rc = b->runner.enqueue(true))) return rc;
    }
    if ((rc = b->runner.record_end())) return rc;
    if ((rc = b->runner.finish(results))) return rc;
    b->timing.total_ms = b->runner.elapsed_ms();
    // Inside a replayed graph only the whole sequence is timed.
    b->timing.integral_ms = (b->flags & MW_BATCH_GRAPH) ? b->timing.total_ms : b->runner.integral_ms();
    b->timing.star_ms = 0;
    b->timing.integral_blocks = b->runner.integral_blocks();
    b->timing.integral_points = b->runner.integral_points();
    b->timing.integral_points_recurrence = b->runner.integral_points_recurrence();
    b->timing.integral_points_background = b->runner.integral_points_background();
    b->timing.integral_stream_terms = b->runner.integral_stream_terms();
    b->timing.star_blocks = b->runner.star_blocks();
    return MW_OK;
}

int mw_batch_last_timing(const mw_batch* b, mw_timing* out)
{
    if (!b || !out) return fail(MW_E_INVALID, "null argument");
    *out = b->timing;
    return MW_OK;
}

void mw_batch_destroy(mw_batch* b) { delete b; }

}  // extern "C"
