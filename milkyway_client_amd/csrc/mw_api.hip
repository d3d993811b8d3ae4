// mw_api.hip -- C ABI of the separation library: resident wedge contexts,
// batched evaluation, and the drop-in replacements of the reference
// evaluator (cpu_evaluate, calculate_integrals, calculate_likelihood).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "mw_context.h"

using namespace mw;

namespace mw {
thread_local std::string g_error;

int fail(int code, const std::string& msg)
{
    g_error = msg;
    return code;
}
}  // namespace mw

namespace {

uint64_t hash_stars(const STAR_POINTS* sp)
{
    uint64_t h = 0x9e3779b97f4a7c15ull;
    if (!sp) return h;
    for (int i = 0; i < sp->number_stars; i++) {
        for (int j = 0; j < 3; j++) {
            uint64_t b;
            std::memcpy(&b, &sp->stars[i][j], 8);
            h = (h ^ b) * 0x100000001b3ull;
            h ^= h >> 29;
        }
    }
    return h;
}

}  // namespace

namespace {

// One region's calculation range inside an evaluation plan.
struct PlanRegion {
    int slot, region;
    long begin, end;
};

struct Plan {
    std::vector<EvalConst> K;              // one per slot
    // Slots per parameter vector: a model of S streams takes stream_groups(S)
    // consecutive slots (vector v's group g is slot v * groups + g), whose
    // star passes run in group order and chain each star's mixture sum.
    int groups = 1;
    std::vector<PlanRegion> regions;       // slot-major, region order within a slot
    bool integrals = false, stars = false, mixed = false, device_finalize = false;
    bool resample = false;                 // resampled quadrature where it pays (whole regions)
    long star_begin = 0;
    double* log10p_host = nullptr;         // [n_slots / groups][n_stars]
};

int build_context(mw_ctx* c, const std::vector<RegionGeom>& geoms, int n_conv, int wedge, int sgr,
                  const STAR_POINTS* sp)
{
    if (n_conv < 1 || n_conv > MW_MAX_CONVOLVE || n_conv > tile_pairs())
        return fail(MW_E_INVALID, "convolve must be between 1 and 1024");
    if (geoms.empty()) return fail(MW_E_INVALID, "no integration regions");
    MW_HIP_CHECK(hipGetDevice(&c->device));
    for (auto& e : c->ev) MW_HIP_CHECK(hipEventCreate(&e));
    c->have_events = true;
    c->n_conv = n_conv;
    c->wedge = wedge;
    c->sgr = sgr;
    for (const RegionGeom& g : geoms) {
        if (g.r_steps <= 0 || g.mu_steps <= 0 || g.nu_steps <= 0)
            return fail(MW_E_INVALID, "region with a non-positive step count");
        auto rd = std::make_unique<RegionDev>();
        RegionTables t;
        INTEGRAL_AREA a = area_of(g);
        build_region_tables(a, n_conv, wedge, sgr, &t);
        rd->geom = g;
        rd->n_conv = n_conv;
        rd->n_cols = t.n_cols;
        rd->blocks = (int)((t.n_cols + kBlock - 1) / kBlock);
        rd->r_chunks = region_r_chunks(rd->blocks, g.r_steps);
        int rc;
        if ((rc = rd->rq.reserve(t.rq.size()))) return rc;
        if ((rc = rd->rstep.reserve(t.rstep.size()))) return rc;
        if ((rc = rd->col.reserve(t.col.size()))) return rc;
        MW_HIP_CHECK(hipMemcpy(rd->rq.p, t.rq.data(), sizeof(double) * t.rq.size(), hipMemcpyHostToDevice));
        MW_HIP_CHECK(hipMemcpy(rd->rstep.p, t.rstep.data(), sizeof(RStep) * t.rstep.size(), hipMemcpyHostToDevice));
        MW_HIP_CHECK(hipMemcpy(rd->col.p, t.col.data(), sizeof(double) * t.col.size(), hipMemcpyHostToDevice));
        // What the resampled quadrature needs: the exact tables on the host
        // and the region's extremes.
        double gmin = INFINITY, gmax = -INFINITY, rmin = INFINITY;
        for (size_t i = 0; i < t.rq.size(); i += 2) {
            const double r = t.rq[i];
            rd->r_max = std::max(rd->r_max, r);
            rmin = std::min(rmin, r);
            const double gm = 5.0 * (std::log10(r * 1000.0) - 1.0) + kAbsM;
            gmin = std::min(gmin, gm);
            gmax = std::max(gmax, gm);
        }
        rd->g_span = gmax - gmin;
        rd->r_min = rmin;
        rd->r_span = rd->r_max - rmin;
        for (long col = 0; col < t.n_cols; col++) {
            const double ux = t.col[col];
            rd->min_sin_gc = std::min(rd->min_sin_gc, std::sqrt(std::max(0.0, 1.0 - ux * ux)));
        }
        tabulate_bg_width(*rd, t.col.data(), t.col.data() + t.n_cols, t.col.data() + 2 * t.n_cols,
                          g.mu_steps, g.nu_steps);
        tabulate_sights(*rd, t.col.data(), t.col.data() + t.n_cols, t.col.data() + 2 * t.n_cols,
                        g.mu_steps, g.nu_steps);
        const char* bg_env = std::getenv("MW_HIP_BG_GRID");
        rd->bg_grid = !(bg_env && std::strcmp(bg_env, "0") == 0);
        rd->h_rq = std::move(t.rq);
        rd->h_rstep = std::move(t.rstep);
        c->regions.push_back(std::move(rd));
    }
    c->n_stars = sp ? sp->number_stars : 0;
    if (c->n_stars < 0) return fail(MW_E_INVALID, "negative star count");
    std::vector<double> soa, ratio, kw;
    if (c->n_stars > 0) {
        build_star_soa(*sp, &soa);
        int rc;
        if ((rc = c->star.reserve(soa.size()))) return rc;
        MW_HIP_CHECK(hipMemcpy(c->star.p, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice));
    }
    star_conv_tables(n_conv, &ratio, &kw);
    // The resampled star pass folds the background over the convolution's
    // magnitude offsets dx = 5 log10(ratio), weights kw ratio^3.
    c->conv_dx.resize(n_conv);
    c->conv_w3.resize(n_conv);
    for (int i = 0; i < n_conv; i++) {
        c->conv_dx[i] = 5.0 * std::log10(ratio[i]);
        c->conv_w3[i] = kw[i] * (ratio[i] * ratio[i] * ratio[i]);
    }
    double gap = 0.0;
    for (int i = 1; i < n_conv; i++) gap = std::max(gap, std::fabs(ratio[i] - ratio[i - 1]));
    c->star_coarse_pr = gap > 0.0 ? 2.0 / gap : 1e300;
    if (c->n_stars > 0) tabulate_star_theta(*c, soa.data(), c->n_stars);
    int rc;
    if ((rc = c->ratio.reserve(n_conv)) || (rc = c->kw.reserve(n_conv)) || (rc = c->kw3.reserve(n_conv)))
        return rc;
    MW_HIP_CHECK(hipMemcpy(c->kw3.p, c->conv_w3.data(), sizeof(double) * n_conv, hipMemcpyHostToDevice));
    MW_HIP_CHECK(hipMemcpy(c->ratio.p, ratio.data(), sizeof(double) * n_conv, hipMemcpyHostToDevice));
    MW_HIP_CHECK(hipMemcpy(c->kw.p, kw.data(), sizeof(double) * n_conv, hipMemcpyHostToDevice));
    return MW_OK;
}

// Runs a plan: integral pass (optional), device-side region combination
// (optional), star pass (optional).  Region rows land in c->h_irows
// ([n_plan_regions][kIntComps]) and star rows in c->h_srows ([n_slots][kStarComps]).
int run_plan(mw_ctx* c, Plan& P)
{
    const int NI = integral_comps(), NS = star_comps();
    const int n_slots = (int)P.K.size();
    const int n_pr = (int)P.regions.size();
    const bool timed = (c->flags & MW_CTX_TIMING) != 0;
    int rc;
    MW_HIP_CHECK(hipSetDevice(c->device));
    if ((rc = c->ensure_stream())) return rc;
    hipStream_t st = c->stream;
    if ((rc = c->h_consts.reserve(n_slots))) return rc;
    std::memcpy(c->h_consts.p, P.K.data(), sizeof(EvalConst) * n_slots);

    // ---- integral work items, grouped by kernel variant ----------------------
    // Launch groups: items of one background mode and one stream count (the
    // kernels' template parameters); star groups also of one stream group,
    // launched in group order.
    struct Group { int bg, S, chain; int first_item, n_items, first_block, n_blocks; };
    std::vector<Group> igroups;
    int total_iblocks = 0;
    long points = 0, points_rec = 0, points_bg = 0, stream_terms = 0;
    if (P.integrals && n_pr > 0) {
        if ((rc = c->h_ritems.reserve(n_pr)) || (rc = c->h_rred.reserve(n_pr)) ||
            (rc = c->h_irows.reserve((size_t)n_pr * NI)))
            return rc;
        std::vector<int> order(n_pr);
        for (int i = 0; i < n_pr; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            const EvalConst &ka = P.K[P.regions[a].slot], &kb = P.K[P.regions[b].slot];
            if (ka.bg_mode != kb.bg_mode) return ka.bg_mode < kb.bg_mode;
            return ka.n_streams < kb.n_streams;
        });
        int item = 0;
        for (int oi = 0; oi < n_pr; oi++) {
            const PlanRegion& pr = P.regions[order[oi]];
            const int bg = P.K[pr.slot].bg_mode;
            const int nS = P.K[pr.slot].n_streams;
            if (igroups.empty() || igroups.back().bg != bg || igroups.back().S != nS)
                igroups.push_back(Group{ bg, nS, 0, item, 0, total_iblocks, 0 });
            Group& g = igroups.back();
            RegionDev& rd = *c->regions[pr.region];
            TableRef tr;
            const GridPlan gp = plan_grid(rd, P.K[pr.slot], P.resample);
            const int level = gp.level, layout = gp.layout;
            if ((rc = region_tables(rd, gp, &tr))) return rc;
            RegionLaunch& L = c->h_ritems.p[item];
            L.rq = tr.rq;
            L.rstep = tr.rstep;
            L.col = rd.col.p;
            L.n_cols = rd.n_cols;
            L.r_steps = tr.r_steps;
            L.n_conv = tr.n_conv;
            L.calc_begin = pr.begin;
            L.calc_end = pr.end;
            if (level >= 0) {
                // A resampled region is always evaluated whole; its flat
                // index space is columns x grid rows.
                L.calc_begin = 0;
                L.calc_end = rd.n_cols * (long)tr.r_steps;
                points += rd.n_cols * tr.nodes;
            } else {
                points += (pr.end - pr.begin) * (long)rd.n_conv;
            }
            L.slot = pr.slot;
            L.out_index = order[oi];
            // r-windows on the reference's tables only; resampled grids run whole columns
            L.r_chunks = level < 0 ? rd.r_chunks : 1;
            const int blocks = rd.blocks * L.r_chunks;
            L.block_begin = g.n_blocks;          // block index within the group's launch
            L.block_count = blocks;
            L.n_streams = nS;
            L.r_lo = rd.r_min;
            L.r_hi = rd.r_max;
            L.grid_step = layout > 0 ? recurrence_step(level, P.K[pr.slot]) : 0.0;
            L.grid_lo = tr.lo;
            L.bg_stride = layout_bg_stride(layout);
            L.bg_count = (int)tr.bg_count;
            L.bg_rq = tr.bg_rq;
            copy_stream_tables(tr, &L);
            if (L.grid_step > 0.0) {
                points_rec += rd.n_cols * tr.nodes;
                stream_terms += rd.n_cols * (layout == kLayoutSplit ? tr.stream_terms : tr.nodes * nS);
            }
            points_bg += (level >= 0 ? rd.n_cols * tr.bg_nodes : (pr.end - pr.begin) * (long)rd.n_conv);
            ReduceItem& R = c->h_rred.p[item];
            R.block_begin = total_iblocks;       // global partial index
            R.block_count = blocks;
            R.out_index = order[oi];
            R.pad = 0;
            g.n_items++;
            g.n_blocks += blocks;
            total_iblocks += blocks;
            item++;
        }
        if ((rc = c->ipart.reserve((size_t)total_iblocks * NI * 2))) return rc;
    }

    // ---- star work items -------------------------------------------------
    std::vector<Group> sgroups;
    int total_sblocks = 0;
    const long n_star_pass = c->n_stars - P.star_begin;
    if (P.stars) {
        if ((rc = c->h_sitems.reserve(n_slots)) || (rc = c->h_sred.reserve(n_slots)) ||
            (rc = c->h_srows.reserve((size_t)n_slots * NS)))
            return rc;
        const int blocks = n_star_pass > 0 ? (int)((n_star_pass + kBlock - 1) / kBlock) : 0;
        const int G = P.groups;
        if (G > 1 && (rc = c->chain.reserve((size_t)(n_slots / G) * std::max<long>(c->n_stars, 1))))
            return rc;
        std::vector<int> order(n_slots);
        for (int i = 0; i < n_slots; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            if (a % G != b % G) return a % G < b % G;        // stream groups in order
            if (P.K[a].bg_mode != P.K[b].bg_mode) return P.K[a].bg_mode < P.K[b].bg_mode;
            return P.K[a].n_streams < P.K[b].n_streams;
        });
        for (int oi = 0; oi < n_slots; oi++) {
            const int s = order[oi];
            const int bg = P.K[s].bg_mode, nS = P.K[s].n_streams, cg = s % G;
            if (sgroups.empty() || sgroups.back().bg != bg || sgroups.back().S != nS ||
                sgroups.back().chain != cg)
                sgroups.push_back(Group{ bg, nS, cg, oi, 0, total_sblocks, 0 });
            Group& g = sgroups.back();
            StarLaunch& L = c->h_sitems.p[oi];
            L.star = c->star.p;
            L.n_stars = c->n_stars;
            L.star_begin = P.star_begin;
            L.slot = s;
            L.n_conv = c->n_conv;
            L.block_begin = g.n_blocks;
            L.block_count = blocks;
            L.out_index = s / G;          // the vector's per-star row
            L.n_streams = nS;
            L.chain = G > 1 ? c->chain.p + (size_t)(s / G) * c->n_stars : nullptr;
            L.chain_flags = (cg == 0 ? 1 : 0) | (cg == G - 1 ? 2 : 0);
            L.ratio = c->ratio.p;
            L.kw = c->kw.p;
            L.kw3 = c->kw3.p;
            L.coarse_pr = c->star_coarse_pr;
            const int sl = (P.resample && P.mixed && bg == kBgHernquist) ? star_bg_level(*c, P.K[s].q, P.K[s].r0) : -1;
            if ((rc = star_bg_table(*c, sl, &L.bg_rq, &L.bg_count))) return rc;
            ReduceItem& R = c->h_sred.p[oi];
            R.block_begin = total_sblocks;
            R.block_count = blocks;
            R.out_index = s;
            R.pad = 0;
            g.n_items++;
            g.n_blocks += blocks;
            total_sblocks += blocks;
        }
        if ((rc = c->spart.reserve((size_t)std::max(total_sblocks, 1) * NS * 2))) return rc;
        if (P.log10p_host && (rc = c->log10p.reserve((size_t)(n_slots / G) * std::max<long>(c->n_stars, 1))))
            return rc;
    }

    // ---- device-side combination of regions per slot -------------------------
    if (P.device_finalize && P.stars && P.integrals) {
        if ((rc = c->h_first.reserve(n_slots)) || (rc = c->h_count.reserve(n_slots)))
            return rc;
        for (int s = 0; s < n_slots; s++) {
            c->h_first.p[s] = -1;
            c->h_count.p[s] = 0;
        }
        for (int i = 0; i < n_pr; i++) {
            const int s = P.regions[i].slot;
            if (c->h_first.p[s] < 0) c->h_first.p[s] = i;
            c->h_count.p[s]++;
        }
        for (int s = 0; s < n_slots; s++)
            if (c->h_first.p[s] < 0) return fail(MW_E_INVALID, "slot without regions");
    }

    // ---- one upload, one download ----------------------------------------
    // The constants and work items go up in one copy from one pinned block,
    // and the region and star rows come back in one: each hipMemcpyAsync
    // costs several microseconds of host time, as much as a small wedge's
    // kernels.  Sections are 256-byte aligned at the same offsets on both
    // sides.  The previous evaluation on this context synchronised its
    // stream before returning, so the pinned block is free to refill.
    const bool items = P.integrals && n_pr > 0;
    const bool fin = P.device_finalize && P.stars && P.integrals;
    size_t up = 0;
    auto place = [&](size_t bytes) {
        up = (up + 255) & ~size_t(255);
        const size_t at = up;
        up += bytes;
        return at;
    };
    const size_t o_consts = place(sizeof(EvalConst) * n_slots);
    const size_t o_ritems = items ? place(sizeof(RegionLaunch) * n_pr) : 0;
    const size_t o_rred = items ? place(sizeof(ReduceItem) * n_pr) : 0;
    const size_t o_sitems = P.stars ? place(sizeof(StarLaunch) * n_slots) : 0;
    const size_t o_sred = P.stars ? place(sizeof(ReduceItem) * n_slots) : 0;
    const size_t o_first = fin ? place(sizeof(int) * n_slots) : 0;
    const size_t o_count = fin ? place(sizeof(int) * n_slots) : 0;
    if ((rc = c->h_up.reserve(up)) || (rc = c->d_up.reserve(up))) return rc;
    auto stage = [&](size_t at, const void* src, size_t bytes) { std::memcpy(c->h_up.p + at, src, bytes); };
    stage(o_consts, c->h_consts.p, sizeof(EvalConst) * n_slots);
    if (items) {
        stage(o_ritems, c->h_ritems.p, sizeof(RegionLaunch) * n_pr);
        stage(o_rred, c->h_rred.p, sizeof(ReduceItem) * n_pr);
    }
    if (P.stars) {
        stage(o_sitems, c->h_sitems.p, sizeof(StarLaunch) * n_slots);
        stage(o_sred, c->h_sred.p, sizeof(ReduceItem) * n_slots);
    }
    if (fin) {
        stage(o_first, c->h_first.p, sizeof(int) * n_slots);
        stage(o_count, c->h_count.p, sizeof(int) * n_slots);
    }
    MW_HIP_CHECK(hipMemcpyAsync(c->d_up.p, c->h_up.p, up, hipMemcpyHostToDevice, st));
    EvalConst* const d_consts = reinterpret_cast<EvalConst*>(c->d_up.p + o_consts);
    const RegionLaunch* const d_ritems = reinterpret_cast<const RegionLaunch*>(c->d_up.p + o_ritems);
    const ReduceItem* const d_rred = reinterpret_cast<const ReduceItem*>(c->d_up.p + o_rred);
    const StarLaunch* const d_sitems = reinterpret_cast<const StarLaunch*>(c->d_up.p + o_sitems);
    const ReduceItem* const d_sred = reinterpret_cast<const ReduceItem*>(c->d_up.p + o_sred);
    const int* const d_first = reinterpret_cast<const int*>(c->d_up.p + o_first);
    const int* const d_count = reinterpret_cast<const int*>(c->d_up.p + o_count);
    // rows: the regions' [n_pr][NI], then the stars' [n_slots][NS]
    const size_t n_irows = items ? (size_t)n_pr * NI : 0;
    const size_t n_srows = P.stars ? (size_t)n_slots * NS : 0;
    if ((rc = c->d_rows.reserve(std::max<size_t>(n_irows + n_srows, 1))) ||
        (rc = c->h_rows.reserve(std::max<size_t>(n_irows + n_srows, 1))))
        return rc;
    double* const d_irows = c->d_rows.p;
    double* const d_srows = c->d_rows.p + n_irows;

    if (timed) MW_HIP_CHECK(hipEventRecord(c->ev[0], st));
    for (const Group& g : igroups) {
        MW_HIP_CHECK(launch_integral(g.S, g.bg, P.mixed, d_ritems + g.first_item, g.n_items,
                                     g.n_blocks, d_consts, c->ipart.p + (size_t)g.first_block * NI * 2, st,
                                     P.resample));
    }
    if (timed) MW_HIP_CHECK(hipEventRecord(c->ev[1], st));
    if (P.integrals && n_pr > 0) {
        MW_HIP_CHECK(launch_reduce(false, d_rred, n_pr, c->ipart.p, d_irows, st));
        if (fin) MW_HIP_CHECK(launch_finalize(d_first, d_count, d_irows, d_consts, n_slots, st));
    }
    if (timed) MW_HIP_CHECK(hipEventRecord(c->ev[2], st));
    for (const Group& g : sgroups) {
        if (g.n_blocks == 0) continue;
        MW_HIP_CHECK(launch_star(g.S, g.bg, P.mixed, d_sitems + g.first_item, g.n_items, g.n_blocks,
                                 d_consts, c->ratio.p, c->kw.p,
                                 c->spart.p + (size_t)g.first_block * NS * 2,
                                 P.log10p_host ? c->log10p.p : nullptr, c->n_stars, st));
    }
    if (timed) MW_HIP_CHECK(hipEventRecord(c->ev[3], st));
    if (P.stars) {
        if (total_sblocks > 0)
            MW_HIP_CHECK(launch_reduce(true, d_sred, n_slots, c->spart.p, d_srows, st));
    }
    if (timed) MW_HIP_CHECK(hipEventRecord(c->ev[4], st));
    // the star rows only where a star pass wrote them
    const size_t n_down = n_irows + (total_sblocks > 0 ? n_srows : 0);
    if (n_down > 0)
        MW_HIP_CHECK(hipMemcpyAsync(c->h_rows.p, d_irows, sizeof(double) * n_down, hipMemcpyDeviceToHost, st));
    if (P.stars && P.log10p_host && c->n_stars > 0)
        MW_HIP_CHECK(hipMemcpyAsync(P.log10p_host, c->log10p.p, sizeof(double) * (n_slots / P.groups) * c->n_stars,
                                    hipMemcpyDeviceToHost, st));
    MW_HIP_CHECK(hipStreamSynchronize(st));
    if (items) std::memcpy(c->h_irows.p, c->h_rows.p, sizeof(double) * n_irows);
    if (P.stars) {
        if (total_sblocks > 0) std::memcpy(c->h_srows.p, c->h_rows.p + n_irows, sizeof(double) * n_srows);
        else std::memset(c->h_srows.p, 0, sizeof(double) * n_srows);
    }
    if (timed) {
        float a = 0, b = 0, t = 0;
        MW_HIP_CHECK(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
        MW_HIP_CHECK(hipEventElapsedTime(&b, c->ev[2], c->ev[3]));
        MW_HIP_CHECK(hipEventElapsedTime(&t, c->ev[0], c->ev[4]));
        c->timing.integral_ms = a;
        c->timing.star_ms = b;
        c->timing.total_ms = t;
    }
    c->timing.integral_blocks = total_iblocks;
    c->timing.integral_points = points;
    c->timing.integral_points_recurrence = points_rec;
    c->timing.integral_points_background = points_bg;
    c->timing.integral_stream_terms = stream_terms;
    c->timing.star_blocks = total_sblocks;
    return MW_OK;
}

// The constant slots of one model (any number of streams) appended to K.
int make_consts(const ASTRONOMY_PARAMETERS& ap, std::vector<EvalConst>* K)
{
    std::string err;
    std::vector<EvalConst> k;
    if (!make_eval_consts(ap, &k, &err)) return fail(MW_E_INVALID, err);
    K->insert(K->end(), k.begin(), k.end());
    return MW_OK;
}

// ---------------------------------------------------------------------------
// Drop-in state: one cached context per process, rebuilt when the wedge
// geometry or the star list changes.
// ---------------------------------------------------------------------------
std::mutex g_dropin_mutex;
std::unique_ptr<mw_ctx> g_dropin;
int g_precision = -1;

// Caller holds g_dropin_mutex.
int dropin_precision_locked()
{
    if (g_precision < 0) {
        const char* e = std::getenv("MW_HIP_PRECISION");
        g_precision = (e && (std::strcmp(e, "mixed") == 0 || std::strcmp(e, "1") == 0)) ? MW_PREC_MIXED
                                                                                          : MW_PREC_FP64;
    }
    return g_precision;
}

int dropin_context(const ASTRONOMY_PARAMETERS* ap, const EVALUATION_STATE* es, const STAR_POINTS* sp,
                   mw_ctx** out)
{
    if (!ap || !es) return fail(MW_E_INVALID, "null argument");
    if (es->number_integrals != ap->number_integrals || ap->number_integrals < 1)
        return fail(MW_E_INVALID, "evaluation state does not match the parameters");
    std::vector<RegionGeom> geoms;
    for (int g = 0; g < es->number_integrals; g++) geoms.push_back(geom_of(*es->integral[g]));
    int dev = 0;
    MW_HIP_CHECK(hipGetDevice(&dev));
    const int nst = sp ? sp->number_stars : 0;
    const uint64_t h = hash_stars(sp);
    mw_ctx* c = g_dropin.get();
    const bool same = c && c->device == dev && c->key_geom == geoms && c->n_conv == ap->convolve &&
                      c->wedge == ap->wedge && c->sgr == ap->sgr_coordinates && c->key_sp == sp &&
                      c->key_stars == nst && c->key_hash == h;
    if (!same) {
        g_dropin.reset();
        auto nc = std::make_unique<mw_ctx>();
        int rc = build_context(nc.get(), geoms, ap->convolve, ap->wedge, ap->sgr_coordinates, sp);
        if (rc) return rc;
        nc->key_geom = geoms;
        nc->key_sp = sp;
        nc->key_stars = nst;
        nc->key_hash = h;
        g_dropin = std::move(nc);
    }
    g_dropin->S = ap->number_streams;
    *out = g_dropin.get();
    return MW_OK;
}

}  // namespace

namespace mw {
int dropin_precision_value()
{
    std::lock_guard<std::mutex> lock(g_dropin_mutex);
    return dropin_precision_locked();
}
}  // namespace mw

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* mw_last_error(void) { return g_error.c_str(); }

int mw_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int mw_set_device(int device)
{
    const int n = mw_device_count();
    if (n <= 0) return fail(MW_E_NODEVICE, "no GPU visible");
    if (device < 0 || device >= n) return fail(MW_E_INVALID, "device index out of range");
    MW_HIP_CHECK(hipSetDevice(device));
    return MW_OK;
}

int mw_device_synchronize(void)
{
    MW_HIP_CHECK(hipDeviceSynchronize());
    return MW_OK;
}

int mw_hip_set_precision(int precision)
{
    if (precision != MW_PREC_FP64 && precision != MW_PREC_MIXED)
        return fail(MW_E_INVALID, "unknown precision");
    std::lock_guard<std::mutex> lock(g_dropin_mutex);
    g_precision = precision;
    return MW_OK;
}

int mw_ctx_create(const ASTRONOMY_PARAMETERS* ap, const STAR_POINTS* sp, int flags, mw_ctx** out)
{
    if (!ap || !out) return fail(MW_E_INVALID, "null argument");
    *out = nullptr;
    if (mw_device_count() <= 0) return fail(MW_E_NODEVICE, "no GPU visible");
    if (ap->number_streams < 0 || ap->number_streams > MW_MAX_STREAMS)
        return fail(MW_E_INVALID, "number_streams must be between 0 and MW_MAX_STREAMS (16)");
    if (ap->number_integrals < 1 || !ap->integral) return fail(MW_E_INVALID, "no integration regions");
    std::vector<RegionGeom> geoms;
    for (int g = 0; g < ap->number_integrals; g++) geoms.push_back(geom_of(*ap->integral[g]));
    auto c = std::make_unique<mw_ctx>();
    c->flags = flags;
    int rc = build_context(c.get(), geoms, ap->convolve, ap->wedge, ap->sgr_coordinates, sp);
    if (rc) return rc;
    c->tmpl.load(*ap);
    c->S = ap->number_streams;
    c->n_params = optimized_count(c->tmpl.ap);
    // Validate the template parameters once.
    std::vector<EvalConst> k;
    if ((rc = make_consts(c->tmpl.ap, &k))) return rc;
    *out = c.release();
    return MW_OK;
}

int mw_ctx_parameter_count(const mw_ctx* ctx) { return ctx ? ctx->n_params : -1; }

void mw_ctx_destroy(mw_ctx* ctx) { delete ctx; }

size_t mw_ctx_resident_bytes(const mw_ctx* c)
{
    if (!c) return 0;
    size_t b = 0;
    for (const auto& r : c->regions) {
        b += r->rq.cap * sizeof(double) + r->rstep.cap * sizeof(RStep) + r->col.cap * sizeof(double);
        std::lock_guard<std::mutex> lock(r->level_mu);
        for (const auto& per_level : r->levels)
            for (const auto& q : per_level)
                if (q) b += (q->rq.cap + q->w.cap) * sizeof(double) + q->rstep.cap * sizeof(RStep);
        for (const auto& q : r->bg_levels)
            if (q) b += (q->rq.cap + q->w.cap) * sizeof(double) + q->rstep.cap * sizeof(RStep);
    }
    b += (c->star.cap + c->ratio.cap + c->kw.cap + c->kw3.cap) * sizeof(double);
    {
        std::lock_guard<std::mutex> lock(c->star_bg_mu);
        for (const auto& q : c->star_bg)
            if (q) b += q->rq.cap * sizeof(double);
    }
    return b;
}

int mw_ctx_eval(mw_ctx* c, const double* params, int n_vec, int precision, mw_result* results,
                double* log10p)
{
    if (!c || n_vec < 1 || !results) return fail(MW_E_INVALID, "bad arguments");
    if (c->n_params > 0 && !params) return fail(MW_E_INVALID, "null parameter vector");
    if (precision != MW_PREC_FP64 && precision != MW_PREC_MIXED && precision != MW_PREC_RESAMPLED)
        return fail(MW_E_INVALID, "unknown precision");
    const int G = (int)c->regions.size();
    const int SG = stream_groups(c->S);
    Plan P;
    P.groups = SG;
    for (int v = 0; v < n_vec; v++) {
        ModelTemplate m;
        m.load(c->tmpl.ap);
        if (c->n_params > 0) apply_vector(m.ap, params + (size_t)v * c->n_params);
        int rc = make_consts(m.ap, &P.K);
        if (rc) return rc;
        for (int sg = 0; sg < SG; sg++) {
            for (int g = 0; g < G; g++) {
                const RegionGeom& rg = c->regions[g]->geom;
                P.regions.push_back(PlanRegion{ v * SG + sg, g, 0, (long)rg.r_steps * rg.mu_steps * rg.nu_steps });
            }
        }
    }
    P.integrals = true;
    P.stars = c->n_stars > 0;
    P.mixed = precision != MW_PREC_FP64;
    // a model of more than 4 streams runs on the exact quadrature
    P.resample = precision == MW_PREC_RESAMPLED && SG == 1;
    P.device_finalize = true;
    P.star_begin = 0;
    P.log10p_host = log10p;
    int rc = run_plan(c, P);
    if (rc) return rc;
    const int NI = integral_comps(), NS = star_comps();
    constexpr int RC = 1 + MW_MAX_STREAMS;   // components of a region row of mw_ctx_region_integrals
    c->last_regions.assign((size_t)n_vec * G * RC, 0.0);
    c->last_n_vec = n_vec;
    const double N = (double)c->n_stars;
    for (int v = 0; v < n_vec; v++) {
        // region g of stream group sg of vector v: row (v SG + sg) G + g
        double* lr = c->last_regions.data() + (size_t)v * G * RC;
        for (int sg = 0; sg < SG; sg++) {
            for (int g = 0; g < G; g++) {
                const double* row = c->h_irows.p + ((size_t)(v * SG + sg) * G + g) * NI;
                if (sg == 0) lr[g * RC] = row[0];
                const int ns = P.K[v * SG + sg].n_streams;
                for (int j = 0; j < ns; j++) lr[g * RC + 1 + kMaxStreams * sg + j] = row[1 + j];
            }
        }
        mw_result& r = results[v];
        std::memset(&r, 0, sizeof(r));
        r.number_streams = c->S;
        r.background_integral = lr[0];
        for (int s = 0; s < c->S; s++) r.stream_integrals[s] = lr[1 + s];
        for (int g = 1; g < G; g++) {
            r.background_integral -= lr[g * RC];
            for (int s = 0; s < c->S; s++) r.stream_integrals[s] -= lr[g * RC + 1 + s];
        }
        if (P.stars) {
            const double* first = c->h_srows.p + (size_t)(v * SG) * NS;
            const double* last = c->h_srows.p + (size_t)(v * SG + SG - 1) * NS;
            r.num_zero = (int)last[NS - 1];
            r.prob_sum = last[0] - 238.0 * r.num_zero;
            r.likelihood = r.prob_sum / N;
            r.bg_only_likelihood = first[1] / N;
            for (int s = 0; s < c->S; s++)
                r.stream_only_likelihood[s] = c->h_srows.p[(size_t)(v * SG + s / kMaxStreams) * NS + 2 + s % kMaxStreams] / N;
        } else {
            r.likelihood = NAN;
        }
    }
    return MW_OK;
}

int mw_ctx_region_integrals(const mw_ctx* c, double* out, size_t n)
{
    if (!c || !out) return fail(MW_E_INVALID, "null argument");
    const size_t need = c->last_regions.size();
    if (n < need) return fail(MW_E_INVALID, "output too small");
    std::memcpy(out, c->last_regions.data(), sizeof(double) * need);
    return MW_OK;
}

int mw_ctx_last_timing(const mw_ctx* c, mw_timing* out)
{
    if (!c || !out) return fail(MW_E_INVALID, "null argument");
    *out = c->timing;
    return MW_OK;
}

int mw_hip_calculate_integrals(const ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es, const STAR_POINTS* sp)
{
    std::lock_guard<std::mutex> lock(g_dropin_mutex);
    if (mw_device_count() <= 0) return fail(MW_E_NODEVICE, "no GPU visible");
    mw_ctx* c = nullptr;
    int rc = dropin_context(ap, es, sp, &c);
    if (rc) return rc;
    Plan P;
    if ((rc = make_consts(*ap, &P.K))) return rc;
    P.groups = (int)P.K.size();
    std::vector<long> ends(es->number_integrals, -1);
    for (int sg = 0; sg < P.groups; sg++) {
        for (int g = es->current_integral; g < es->number_integrals; g++) {
            const INTEGRAL_AREA& a = *es->integral[g];
            const long total = (long)a.r_steps * a.mu_steps * a.nu_steps;
            long b = a.current_calculation, e = a.max_calculation;
            if (b < 0) return fail(MW_E_INVALID, "negative current_calculation");
            if (b >= total) continue;            // nothing left in this region
            // The reference computes the point at current_calculation before it
            // tests max_calculation (evaluation_optimized.c:498-522).
            if (e <= b) e = b + 1;
            if (e > total) e = total;
            ends[g] = e;
            P.regions.push_back(PlanRegion{ sg, g, b, e });
        }
    }
    P.integrals = true;
    P.mixed = dropin_precision_locked() == MW_PREC_MIXED;
    if ((rc = run_plan(c, P))) return rc;
    const int NI = integral_comps();
    for (size_t i = 0; i < P.regions.size(); i++) {
        const int sg = P.regions[i].slot;
        INTEGRAL_AREA& a = *es->integral[P.regions[i].region];
        const double* row = c->h_irows.p + i * NI;
        if (sg == 0) a.background_integral += row[0];
        for (int j = 0; j < P.K[sg].n_streams; j++) a.stream_integrals[kMaxStreams * sg + j] += row[1 + j];
        a.current_calculation = ends[P.regions[i].region];
    }
    es->current_integral = ap->number_integrals;
    es->background_integral = es->integral[0]->background_integral;
    for (int s = 0; s < ap->number_streams; s++) es->stream_integrals[s] = es->integral[0]->stream_integrals[s];
    for (int g = 1; g < ap->number_integrals; g++) {
        es->background_integral -= es->integral[g]->background_integral;
        for (int s = 0; s < ap->number_streams; s++) es->stream_integrals[s] -= es->integral[g]->stream_integrals[s];
    }
    return MW_OK;
}

int mw_hip_calculate_likelihood(const ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es, const STAR_POINTS* sp)
{
    std::lock_guard<std::mutex> lock(g_dropin_mutex);
    if (!sp) return fail(MW_E_INVALID, "null star list");
    if (mw_device_count() <= 0) return fail(MW_E_NODEVICE, "no GPU visible");
    mw_ctx* c = nullptr;
    int rc = dropin_context(ap, es, sp, &c);
    if (rc) return rc;
    Plan P;
    if ((rc = make_consts(*ap, &P.K))) return rc;
    P.groups = (int)P.K.size();
    for (int sg = 0; sg < P.groups; sg++) {
        P.K[sg].I_bg = es->background_integral;
        for (int j = 0; j < P.K[sg].n_streams; j++) P.K[sg].I_st[j] = es->stream_integrals[kMaxStreams * sg + j];
    }
    P.stars = true;
    P.mixed = dropin_precision_locked() == MW_PREC_MIXED;
    P.star_begin = std::max(0, es->current_star_point);
    if (P.star_begin > sp->number_stars) P.star_begin = sp->number_stars;
    if ((rc = run_plan(c, P))) return rc;
    const int NS = star_comps();
    const double* sr = c->h_srows.p + (size_t)(P.groups - 1) * NS;   // the last group finishes the sum
    const int nz = (int)sr[NS - 1];
    es->prob_sum = sr[0] - 238.0 * nz;
    es->num_zero += nz;
    es->current_star_point = sp->number_stars;
    return MW_OK;
}

double mw_hip_evaluate(double* parameters, ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es, STAR_POINTS* sp)
{
    // set_astronomy_parameters + reset_evaluation_state
    apply_vector(*ap, parameters);
    reset_state(es);
    int rc = mw_hip_calculate_integrals(ap, es, sp);
    if (rc) {
        fprintf(stderr, "APP: error calculating integrals: %d (%s)\n", rc, mw_last_error());
        exit(rc);
    }
    rc = mw_hip_calculate_likelihood(ap, es, sp);
    if (rc) {
        fprintf(stderr, "APP: error calculating likelihood: %d (%s)\n", rc, mw_last_error());
        exit(rc);
    }
    return es->prob_sum / (sp->number_stars - es->bad_jacobians);
}

}  // extern "C"
