// mw_context.h -- host-side state shared by the C ABI translation units:
// device/pinned buffers, region geometry, the model template that parameter
// vectors are written into, and the resident wedge context (mw_ctx).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "mw_hip.h"
#include "mw_internal.h"

namespace mw {

// rec: some item may be on a distance-grid recurrence (grid_step > 0).
hipError_t launch_integral(int S, int bg, bool mixed, const RegionLaunch* d_items, int n_items,
                           int n_blocks, const EvalConst* d_consts, double* d_partials,
                           hipStream_t stream, bool rec = true);
hipError_t launch_star(int S, int bg, bool mixed, const StarLaunch* d_items, int n_items,
                       int n_blocks, const EvalConst* d_consts, const double* d_ratio,
                       const double* d_kw, double* d_partials, double* d_log10p, long stride,
                       hipStream_t stream);
hipError_t launch_reduce(bool star, const void* d_items, int n_items, const double* d_partials,
                         double* d_rows, hipStream_t stream);
hipError_t launch_finalize(const int* d_first, const int* d_count, const double* d_rows,
                           EvalConst* d_consts, int n_slots, hipStream_t stream);
// layouts: what the launch's items hold (kLaunchAny, kLaunchSplitOnly,
// kLaunchExactOnly in mw_internal.h), which picks the kernel compiled for them.
hipError_t launch_persistent(int bg, bool mixed, int variant, int grid, const RegionLaunch* d_items,
                             int n_items, int total_blocks, const EvalConst* d_consts,
                             double* d_partials, unsigned* d_counter, hipStream_t stream,
                             int layouts = kLaunchAny);
// Per-work-item scratch (register spill) bytes of the kernel that
// launch_persistent runs for these arguments.
hipError_t persistent_scratch_bytes(int bg, bool mixed, int variant, int layouts, size_t* bytes);
hipError_t launch_reset_counters(unsigned* d_counters, int n, hipStream_t stream);
hipError_t launch_star_any(int bg, bool mixed, const StarLaunch* d_items, int n_items, int n_blocks,
                           const EvalConst* d_consts, double* d_partials, hipStream_t stream);
// Stream separation over every star of item (1 <= S <= 4): per star the
// assignment, the transformed position [3] and the S membership probabilities.
hipError_t launch_separation(int S, int bg, bool mixed, const StarLaunch& item, const EvalConst& k,
                             const SepConst& sc, int* d_assign, double* d_xyz, double* d_sprob,
                             hipStream_t stream);
int persistent_variants();
int tile_pairs();
int integral_comps();
int star_comps();

// Records msg as this thread's last error and returns code.
int fail(int code, const std::string& msg);

// Precision of the drop-in entry points (MW_HIP_PRECISION or
// mw_hip_set_precision).
int dropin_precision_value();

#define MW_HIP_CHECK(expr)                                                                  \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(MW_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

// Buffers that grow on demand.  Freeing device or pinned memory waits for
// the whole device, including other streams' kernels, so growth never frees:
// the outgrown block is kept until the buffer itself is released, and a
// buffer that grows again takes half as much again, which bounds the kept
// blocks at twice the final size and settles after a few batches.  The first
// allocation is exact (resident tables are allocated once).
template <typename T, bool PINNED>
struct GrowBuf {
    T* p = nullptr;
    size_t cap = 0;
    size_t held = 0;   // bytes allocated, outgrown blocks included
    std::vector<T*> outgrown;
    GrowBuf() = default;
    GrowBuf(const GrowBuf&) = delete;
    GrowBuf& operator=(const GrowBuf&) = delete;
    ~GrowBuf() { release(); }
    static void free_block(T* q)
    {
        if (!q) return;
        if (PINNED) (void)hipHostFree(q);
        else (void)hipFree(q);
    }
    void release()
    {
        free_block(p);
        for (T* q : outgrown) free_block(q);
        outgrown.clear();
        p = nullptr;
        cap = 0;
        held = 0;
    }
    int reserve(size_t n)
    {
        if (n <= cap) return MW_OK;
        const size_t want = cap ? std::max(n, cap + cap / 2) : n;
        const size_t bytes = sizeof(T) * std::max<size_t>(want, 1);
        T* q = nullptr;
        const hipError_t e = PINNED ? hipHostMalloc((void**)&q, bytes, hipHostMallocDefault)
                                    : hipMalloc((void**)&q, bytes);
        if (e != hipSuccess)
            return fail(MW_E_NOMEM, PINNED ? "hipHostMalloc failed" : "hipMalloc failed");
        if (p) outgrown.push_back(p);
        p = q;
        cap = want;
        held += bytes;
        return MW_OK;
    }
};

template <typename T> using DevBuf = GrowBuf<T, false>;    // device memory
template <typename T> using HostBuf = GrowBuf<T, true>;    // pinned host memory

// Geometry of one region, as INTEGRAL_AREA holds it.
struct RegionGeom {
    int r_steps, mu_steps, nu_steps;
    double r_min, r_max, mu_min, mu_max, nu_min, nu_max;
    double r_step_size, mu_step_size, nu_step_size;
    bool operator==(const RegionGeom& o) const { return std::memcmp(this, &o, sizeof(*this)) == 0; }
};

inline RegionGeom geom_of(const INTEGRAL_AREA& a)
{
    RegionGeom g;
    std::memset(&g, 0, sizeof(g));
    g.r_steps = a.r_steps; g.mu_steps = a.mu_steps; g.nu_steps = a.nu_steps;
    g.r_min = a.r_min; g.r_max = a.r_max; g.mu_min = a.mu_min; g.mu_max = a.mu_max;
    g.nu_min = a.nu_min; g.nu_max = a.nu_max;
    g.r_step_size = a.r_step_size; g.mu_step_size = a.mu_step_size; g.nu_step_size = a.nu_step_size;
    return g;
}

// INTEGRAL -> INTEGRAL_AREA geometry as initialize_integral_area computes it
// (evaluation_state.c:97-127).
inline RegionGeom geom_of(const INTEGRAL& in)
{
    INTEGRAL_AREA a;
    std::memset(&a, 0, sizeof(a));
    a.mu_min = in.mu_min; a.mu_max = in.mu_max; a.mu_steps = in.mu_steps;
    a.nu_min = in.nu_min; a.nu_max = in.nu_max; a.nu_steps = in.nu_steps;
    a.r_min = in.r_min; a.r_max = in.r_max; a.r_steps = in.r_steps;
    a.mu_step_size = (a.mu_max - a.mu_min) / a.mu_steps;
    a.nu_step_size = (a.nu_max - a.nu_min) / a.nu_steps;
    a.r_step_size = (a.r_max - a.r_min) / a.r_steps;
    return geom_of(a);
}

inline INTEGRAL_AREA area_of(const RegionGeom& g)
{
    INTEGRAL_AREA a;
    std::memset(&a, 0, sizeof(a));
    a.r_steps = g.r_steps; a.mu_steps = g.mu_steps; a.nu_steps = g.nu_steps;
    a.r_min = g.r_min; a.r_max = g.r_max; a.mu_min = g.mu_min; a.mu_max = g.mu_max;
    a.nu_min = g.nu_min; a.nu_max = g.nu_max;
    a.r_step_size = g.r_step_size; a.mu_step_size = g.mu_step_size; a.nu_step_size = g.nu_step_size;
    return a;
}

// Deep copy of the model part of ASTRONOMY_PARAMETERS (values and optimize
// masks) that parameter vectors are written into.
struct ModelTemplate {
    ASTRONOMY_PARAMETERS ap{};
    std::vector<double> bg;
    std::vector<int> bg_opt;
    std::vector<double> sw;
    std::vector<int> sw_opt;
    std::vector<std::vector<double>> sp;
    std::vector<std::vector<int>> sp_opt;
    std::vector<double*> sp_rows;
    std::vector<int*> sp_opt_rows;

    void load(const ASTRONOMY_PARAMETERS& src)
    {
        ap = src;
        const int nb = src.number_background_parameters, S = src.number_streams;
        const int nsp = src.number_stream_parameters;
        bg.assign(src.background_parameters, src.background_parameters + nb);
        bg_opt.assign(src.background_optimize, src.background_optimize + nb);
        sw.assign(src.stream_weights, src.stream_weights + S);
        sw_opt.assign(src.stream_weight_optimize, src.stream_weight_optimize + S);
        sp.resize(S);
        sp_opt.resize(S);
        sp_rows.resize(S);
        sp_opt_rows.resize(S);
        for (int s = 0; s < S; s++) {
            sp[s].assign(src.stream_parameters[s], src.stream_parameters[s] + nsp);
            sp_opt[s].assign(src.stream_optimize[s], src.stream_optimize[s] + nsp);
            sp_rows[s] = sp[s].data();
            sp_opt_rows[s] = sp_opt[s].data();
        }
        ap.background_parameters = bg.data();
        ap.background_optimize = bg_opt.data();
        ap.stream_weights = sw.data();
        ap.stream_weight_optimize = sw_opt.data();
        ap.stream_parameters = sp_rows.data();
        ap.stream_optimize = sp_opt_rows.data();
        ap.background_step = ap.background_min = ap.background_max = nullptr;
        ap.stream_weight_step = ap.stream_weight_min = ap.stream_weight_max = nullptr;
        ap.stream_step = ap.stream_min = ap.stream_max = nullptr;
        ap.integral = nullptr;
    }
};

// get_optimized_parameter_count (parameters.c:307-330)
inline int optimized_count(const ASTRONOMY_PARAMETERS& ap)
{
    int n = 0;
    for (int i = 0; i < ap.number_background_parameters; i++) n += ap.background_optimize[i] ? 1 : 0;
    for (int s = 0; s < ap.number_streams; s++) {
        n += ap.stream_weight_optimize[s] ? 1 : 0;
        for (int j = 0; j < ap.number_stream_parameters; j++) n += ap.stream_optimize[s][j] ? 1 : 0;
    }
    return n;
}

// set_astronomy_parameters (parameters.c:332-363)
inline void apply_vector(ASTRONOMY_PARAMETERS& ap, const double* v)
{
    int k = 0;
    for (int i = 0; i < ap.number_background_parameters; i++)
        if (ap.background_optimize[i]) ap.background_parameters[i] = v[k++];
    for (int s = 0; s < ap.number_streams; s++) {
        if (ap.stream_weight_optimize[s]) ap.stream_weights[s] = v[k++];
        for (int j = 0; j < ap.number_stream_parameters; j++)
            if (ap.stream_optimize[s][j]) ap.stream_parameters[s][j] = v[k++];
    }
}

// reset_evaluation_state (evaluation_state.c:170-190); every stream
// integral is cleared (the reference's memset covers number_streams bytes,
// and the integral pass overwrites the totals anyway).
inline void reset_state(EVALUATION_STATE* es)
{
    es->current_integral = 0;
    es->background_integral = 0;
    for (int s = 0; s < es->number_streams; s++) es->stream_integrals[s] = 0;
    es->current_star_point = 0;
    es->num_zero = 0;
    es->bad_jacobians = 0;
    es->prob_sum = 0;
    for (int g = 0; g < es->number_integrals; g++) {
        es->integral[g]->background_integral = 0;
        for (int s = 0; s < es->integral[g]->number_streams; s++) es->integral[g]->stream_integrals[s] = 0;
        es->integral[g]->current_calculation = es->integral[g]->min_calculation;
    }
}

// Resampled quadrature tables of one region at one grid spacing.  Table
// index t < kQuadLevels: a grid in magnitude with spacing 0.02 * 2^(-t/2);
// t = kQuadLevels + L: a grid in distance with spacing 2 * 2^(-L/2) kpc.
constexpr int kQuadLevels = 13;
constexpr int kQuadTables = 2 * kQuadLevels;
// Layouts of a table: 0 = (r, weight) pairs for kernels that evaluate each
// node's exponential; 1, 2, 3 = the recurrence layout on distance grids with
// the background on every 1st, 2nd, 4th node; 4 = the recurrence on a
// distance grid for the streams alone, the background on a grid of its own
// uniform in magnitude (a background table, index t < kBgLevels: spacing
// 0.3 * 2^(-t/2) mag).
constexpr int kTableLayouts = 5;
constexpr int kLayoutSplit = 4;
constexpr int kBgLevels = 13;
inline int layout_bg_stride(int layout)
{
    return (layout == 0 || layout == kLayoutSplit) ? 0 : 1 << (layout - 1);
}
struct ResampledDev {
    int r_steps = 0, n_conv = 0;
    long nodes = 0;
    double lo = 0.0;
    DevBuf<double> rq;
    DevBuf<RStep> rstep;
    DevBuf<double> w;   // (r, weight) tables: the weights alone, contiguous (layout 4's streams)
};

// Windows of the r-steps per 256-column tile for a region's exact tables:
// a region of few tiles would leave most of the GPU idle in a launch of its
// own (a single evaluation of a small wedge, one GPU's share of a sharded
// one), so its tiles are split along r until the region has about
// kChunkTargetBlocks logical blocks, each window keeping at least
// kChunkMinRSteps r-steps.  It depends on the region alone, so that every
// launch of the region (single, batched, without allocation) sums the same
// partials in the same order.
constexpr int kChunkTargetBlocks = 2048;
constexpr int kChunkMinRSteps = 4;
inline int region_r_chunks(int tiles, int r_steps)
{
    if (tiles <= 0 || r_steps <= 0) return 1;
    const int want = (kChunkTargetBlocks + tiles - 1) / tiles;
    const int most = r_steps / kChunkMinRSteps;
    return std::max(1, std::min(want, most));
}

struct RegionDev {
    RegionGeom geom;
    int n_conv = 0;
    long n_cols = 0;
    int blocks = 0;          // 256-column tiles
    int r_chunks = 1;        // windows of the r-steps per tile on the exact tables (region_r_chunks)
    DevBuf<double> rq, col;
    DevBuf<RStep> rstep;
    // Host copies of the exact r-step tables, and what the choice of a
    // resampled grid needs: the largest convolution distance (kpc), the
    // distance and magnitude spans of the convolution points, and the smallest
    // sqrt(1 - u_x^2) over the region's sights (their closest approach to
    // the Galactic centre is 8.5 kpc times that for q = 1).
    std::vector<double> h_rq;
    std::vector<RStep> h_rstep;
    double r_min = 0.0, r_max = 0.0, r_span = 0.0, g_span = 0.0, min_sin_gc = 1.0;
    // Smallest width of the background's peak along the region's sights
    // (kpc), tabulated at q = 2^(j / 4 - kWidthQ0), from a grid of sights
    // that includes the region's edges (see bg_width).
    std::vector<float> bg_width_q;
    // Smallest angle (radians) at which the background's singularities sit
    // off the real axis of the magnitude variable along the region's sights,
    // at the same q (see bg_theta).
    std::vector<float> bg_theta_q;
    // Sight directions sampled over the region (up to 32 x 8, its first and
    // last rows and columns included; 3 doubles each) and the cosine of the
    // largest angle between a sight of the region and its nearest sample
    // (see streams_peak_inside).
    std::vector<double> sample_u;
    double sample_cos = 1.0;
    // Layout 4 allowed (off when the environment has MW_HIP_BG_GRID=0 as the
    // region is built: the background then stays on the distance grid).
    bool bg_grid = true;
    std::unique_ptr<ResampledDev> levels[kQuadTables][kTableLayouts];
    std::unique_ptr<ResampledDev> bg_levels[kBgLevels];
    std::mutex level_mu;
};

// Tables a launch reads for one region: the exact ones or a resampled level.
struct TableRef {
    const double* rq;
    const RStep* rstep;
    int r_steps, n_conv;
    int level;         // -1 for the exact tables
    long nodes;        // integrand evaluations per sight: r_steps x n_conv, or the grid's nodes
    int layout;        // see kTableLayouts
    double lo;         // resampled: the grid's first node
    long bg_nodes;     // background evaluations per sight (nodes / stride in the recurrence layouts)
    const double* bg_rq = nullptr;   // layout 4: the background grid's (r, weight) entries
    long bg_count = 0;               // and their number (the fold's lattice)
    long stream_terms = 0;           // layout 4: stream terms per sight (the streams' grids' nodes)
    // layout 4: each stream's own distance table (st_rows 0: switched off)
    const double* st_rq[kMaxStreams] = {};   // the weights alone
    double st_step[kMaxStreams] = {}, st_lo[kMaxStreams] = {};
    int st_rows[kMaxStreams] = {};
};

// What a region is evaluated on under one model: the quadrature table index
// (-1 = the exact tables), its layout and, for layout 4, the background
// table index.
struct GridPlan {
    int level = -1, layout = 0, bg_level = -1;
    int st_level[kMaxStreams] = { -1, -1, -1, -1 };   // layout 4: each stream's distance level
};

// Tabulate RegionDev::bg_width_q and bg_theta_q from the region's sight
// directions (ux, uy, uz arrays of n_cols over mu_steps x nu_steps, mu-major).
constexpr int kWidthQ = 33, kWidthQ0 = 4;
void tabulate_bg_width(RegionDev& rd, const double* ux, const double* uy, const double* uz, int mu_steps,
                       int nu_steps);
// A lower bound on the width of the background's peak along any sight of the
// region for flattening q (kpc).
double bg_width(const RegionDev& rd, double q);
// A lower bound on that angle over the region's sights for flattening q.
double bg_theta(const RegionDev& rd, double q);
// Background table index for flattening q (the coarsest spacing within
// 0.2 x 2.17 x theta and 0.3 mag), or -1 if none is fine enough.
int bg_grid_level(const RegionDev& rd, double q);

// The parameter range over which the resampled quadrature's folds were
// measured against the exact point sets (the reference's own bounds,
// background_min / background_max and stream_min / stream_max of its
// parameter files: q 0.3-1.0, r0 1-30 kpc, sigma 0.1-20 kpc; q up to 1.6 on
// the GPU sweep).  Outside it a region keeps the exact tables and a star
// keeps its exact convolution; wider streams only make the folds more
// accurate, so sigma has no upper limit.
constexpr double kFoldQMin = 0.3, kFoldQMax = 1.6, kFoldR0Min = 1.0, kFoldR0Max = 30.0;
constexpr double kFoldSigmaMin = 0.1;
bool in_measured_domain(double q, double r0);
// Sample the region's sight directions (RegionDev::sample_u, sample_cos).
void tabulate_sights(RegionDev& rd, const double* ux, const double* uy, const double* uz, int mu_steps,
                     int nu_steps);
// Whether every switched-on stream's Gaussian peaks inside the region's
// distances along the sights that carry its integral over the region (the
// resampled grids resolve it there; a stream that passes outside the region
// leaves only steep tails, which they do not).
bool streams_peak_inside(const RegionDev& rd, const EvalConst& k);

// Resampled table index to use for region rd under the model k, or -1 for
// the exact tables (resampling off, a background mode it does not cover, a
// model outside the measured domain, or no gain).  Pure: builds nothing.
int select_level(const RegionDev& rd, const EvalConst& k, bool resample);

// Node spacing for the kernels' stream recurrence on table `level`: the
// distance grid's spacing h when every stream's step p h <= k_s h stays
// within 1/2, else 0 (evaluate each node's exponential).
double recurrence_step(int level, const EvalConst& k);
// Layout of `level` to run under the model k (mixed kernels): 0 unless the
// recurrence applies (the Hernquist profile, whose mixed kernels carry it,
// and a distance grid fine enough for every stream); then 1, 2 or 3 for the
// widest background stride (1, 2, 4 nodes) whose spacing still meets the
// background's own bound (0.15 b and 2 kpc).
int table_layout(const RegionDev& rd, int level, const EvalConst& k);
// The grid a region runs on under the model k: select_level's table and
// table_layout's layout, or layout 4 (each stream on a distance grid spaced
// for that stream alone, the background on its own magnitude grid) where a
// count of vector instructions per sight says it is cheaper.  Pure.
GridPlan plan_grid(const RegionDev& rd, const EvalConst& k, bool resample);
// The tables of that plan, built and uploaded on first use.
int region_tables(RegionDev& rd, const GridPlan& g, TableRef* out);
// Layout 4's per-stream tables into a launch item.
inline void copy_stream_tables(const TableRef& tr, RegionLaunch* L)
{
    for (int s = 0; s < kMaxStreams; s++) {
        L->st_rq[s] = tr.st_rq[s];
        L->st_step[s] = tr.st_step[s];
        L->st_lo[s] = tr.st_lo[s];
        L->st_rows[s] = tr.st_rows[s];
    }
}

}  // namespace mw

// A resident wedge: its regions' tables and star list on one device, plus
// the buffers of its evaluations.
struct mw_ctx {
    int device = 0;
    int flags = 0;
    hipStream_t stream = nullptr;
    int S = 0, n_conv = 0, wedge = 0, sgr = 0;
    int n_params = 0;
    mw::ModelTemplate tmpl;
    std::vector<std::unique_ptr<mw::RegionDev>> regions;
    long n_stars = 0;
    mw::DevBuf<double> star, ratio, kw, kw3;
    // Resampled star pass: the convolution points' magnitude offsets and
    // weights kw ratio^3 (host), the smallest sin^2 theta over the stars at
    // the tabulated q values (see star_bg_level), and the background tables
    // folded onto grids uniform in the offset, built on first use.
    std::vector<double> conv_dx, conv_w3;
    // Mixed star pass: 2 / the widest gap between adjacent ratio_i.  A
    // stream with p r_star above it may have its Gaussian's peak between
    // two points with both terms' exponents above 1.
    double star_coarse_pr = 0.0;
    std::vector<double> star_sin2_q;
    std::unique_ptr<mw::ResampledDev> star_bg[mw::kBgLevels];
    mutable std::mutex star_bg_mu;

    // per-evaluation buffers: constants and work items staged on the host
    // (h_*) and uploaded together in one block (h_up -> d_up); region and
    // star rows downloaded together (d_rows -> h_rows, then h_irows, h_srows)
    mw::DevBuf<unsigned char> d_up;
    mw::HostBuf<unsigned char> h_up;
    mw::DevBuf<double> d_rows;
    mw::HostBuf<double> h_rows;
    mw::DevBuf<double> ipart, spart, log10p;
    mw::DevBuf<double> chain;           // per-star mixture sums between stream groups
    mw::DevBuf<int> sep_assign;         // stream separation outputs
    mw::DevBuf<double> sep_xyz, sep_prob;
    mw::HostBuf<mw::EvalConst> h_consts;
    mw::HostBuf<mw::RegionLaunch> h_ritems;
    mw::HostBuf<mw::StarLaunch> h_sitems;
    mw::HostBuf<mw::ReduceItem> h_rred, h_sred;
    mw::HostBuf<double> h_irows, h_srows;
    mw::HostBuf<int> h_first, h_count;

    hipEvent_t ev[6] = {};
    bool have_events = false;
    mw_timing timing{};
    std::vector<double> last_regions;   // [n_vec][n_regions][1 + MW_MAX_STREAMS]
    int last_n_vec = 0;

    // drop-in cache key
    std::vector<mw::RegionGeom> key_geom;
    const STAR_POINTS* key_sp = nullptr;
    int key_stars = 0;
    uint64_t key_hash = 0;

    // The context's own stream is created on its first evaluation: a context
    // that only serves batches (a server's wedge) holds no stream, and with
    // it no hardware queue that a batch runner's stream might have to share.
    int ensure_stream()
    {
        if (stream) return MW_OK;
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
            stream = nullptr;
            return mw::fail(MW_E_HIP, "hipStreamCreateWithFlags failed");
        }
        return MW_OK;
    }

    ~mw_ctx()
    {
        if (have_events)
            for (auto& e : ev) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace mw {

// Resampled star pass.  Along a star's sight the background's singularities
// sit at Im x = +-2.17 theta in the magnitude variable, as along the
// integral's sights, and the convolution window is a piece of that sight:
// its background term folds onto a grid uniform in the magnitude offset at
// spacing min(0.1 x 2.17 theta, 0.3 mag), theta the smallest over the
// context's stars.
// Tabulates mw_ctx::star_sin2_q from the star list (SoA, n stars).
void tabulate_star_theta(mw_ctx& c, const double* soa, long n);
// Background table index for the star pass under flattening q and scale r0,
// or -1 when the exact convolution is used (no stars, q outside the table
// or the measured domain, no level fine enough, or no fewer nodes than the
// convolution's points).  Pure.
int star_bg_level(const mw_ctx& c, double q, double r0);
// That table (built on first use): rq (r / r_star, weight / r_star^3) pairs
// and their number; nullptr and 0 for level -1.
int star_bg_table(mw_ctx& c, int level, const double** rq, int* count);

}  // namespace mw
