// mw_internal.h -- shared declarations of the HIP separation library.
//
// Layouts of everything the kernels read: per-wedge resident tables
// (parameter independent), per-evaluation constants (one slot per parameter
// vector), and the work descriptors of a launch.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "mw_separation_types.h"

namespace mw {

constexpr int kMaxStreams = 4;        // streams per kernel instantiation
constexpr int kBlock = 256;           // work-items per workgroup (4 wavefronts)
constexpr double kSunR0 = 8.5;        // evaluation_optimized.c:43, stCoords.c:30
constexpr double kStdev = 0.6;        // evaluation_optimized.c:41
constexpr double kAbsM = 4.2;         // evaluation_optimized.c:44

// Background profile selected per evaluation.
enum BgMode : int {
    kBgHernquist = 0,     // alpha == delta == 1, no aux profile
    kBgHernquistAux = 1,  // alpha == delta == 1 with a g^2 + b g + c
    kBgGeneral = 2,       // pow(rg, alpha) * pow(rg + r0, 3 - alpha + delta)
    kBgZeroQ = 3,         // q == 0: the background is -1 per r-step
};

// What the items of a persistent integral launch hold, which picks the
// kernel compiled for them: kLaunchSplitOnly, every item in the split layout
// (resampled, mixed, bg_stride 0); kLaunchExactOnly, no item on a
// distance-grid recurrence (grid_step 0: exact tables, or resampled tables
// with an exponential per node); kLaunchAny otherwise.
constexpr int kLaunchAny = 0, kLaunchSplitOnly = 1, kLaunchExactOnly = 2;

// Per-r-step table entry of one region (host-built, FP64).
struct RStep {
    double irv;           // (r_{k+1}^3 - r_k^3) / 3 * dmu (radians)
    double reff_xr_rp3;   // reff(r') * 3 sigma / r'^3
    double q0, q1, q2;    // sum_i qw_i, sum_i qw_i g_i, sum_i qw_i g_i^2 (aux profile)
    double pad[3];
};
static_assert(sizeof(RStep) == 64, "RStep is one 64-byte row");

// Host copy of one region's tables.
struct RegionTables {
    int r_steps = 0, mu_steps = 0, nu_steps = 0, n_conv = 0;
    long n_cols = 0;
    std::vector<double> rq;        // [r_steps][n_conv][2]: (r_point, qw_r3_N)
    std::vector<RStep> rstep;      // [r_steps]
    std::vector<double> col;       // [4][n_cols] SoA: cos b cos l, cos b sin l, sin b, ids
    double lo = 0.0;               // resampled tables: the grid's first node
};

// Per-evaluation constants, one slot per parameter vector (device resident).
struct alignas(64) EvalConst {
    // background
    double q, inv_q2, r0, alpha, alpha_delta3, bg_a, bg_b, bg_c;
    int bg_mode, n_streams, pad0, pad1;
    // streams: centre c (galactocentric kpc), unit axis a, 1 / (2 sigma^2), on/off
    double sc[kMaxStreams][3];
    double sa[kMaxStreams][3];
    double inv_2s2[kMaxStreams];
    double on[kMaxStreams];
    // Stream-axis distance in linear form.  With c' = c + (8.5, 0, 0) the
    // offset of a point r u - (8.5, 0, 0) from the axis is r u_perp - c_perp,
    // where x_perp = x - (x . a) a.  Both are scaled by
    // k = sqrt(log2(e) / (2 sigma^2)), so |r u_perp k - c_perp k|^2 is the
    // exponent |v_perp|^2 / (2 sigma^2) in base 2.
    double ks[kMaxStreams];
    double cperp[kMaxStreams][3];
    // likelihood weights: exp(w_b), exp(w_s), 0.001 * (exp(w_b) + sum exp(w_s))
    double exp_bg_w, exp_st_w[kMaxStreams], sum_exp_w;
    // integrals (filled by the integral pass, read by the star pass)
    double I_bg, I_st[kMaxStreams];
};

// One region of one evaluation inside a launch.
struct RegionLaunch {
    const double* rq;          // [R][n][2]
    const RStep* rstep;        // [R]
    const double* col;         // [4][n_cols]
    long n_cols;
    int r_steps, n_conv;
    long calc_begin, calc_end; // flat index range [begin, end), mu-major, nu, r
    int slot;                  // evaluation-constant slot
    int out_index;             // index into the per-region result rows
    int block_begin, block_count;
    int r_chunks;              // exact tables: windows of the r-steps per 256-column tile (block lb
                               // takes tile lb / r_chunks, window lb % r_chunks); 0 or 1: one window
    int n_streams;             // streams of the slot's model (batched kernels)
    double r_lo, r_hi;         // exact tables: the smallest and largest r of the region's points
                               // (the mixed kernels shift each stream's exponent by its smallest
                               // t^2 over them)
    double grid_step;          // > 0: rows of nodes uniform in distance at this spacing, in the
                               // recurrence layout (mixed kernels take the stream Gaussians by
                               // recurrence); node m of the region sits at grid_lo + m grid_step
    double grid_lo;
    int bg_stride;             // recurrence layout: background on every bg_stride-th node, or
                               // 0: on its own grid, the bg_count (r, weight) entries of bg_rq
    int bg_count;
    const double* bg_rq;
    // bg_stride 0: each stream on a distance grid of its own, st_rows rows of
    // kResampleRow weights from st_rq[s] (the weights alone), node m at
    // st_lo[s] + m st_step[s]; st_rows[s] == 0 for a switched-off stream.
    const double* st_rq[kMaxStreams];
    double st_step[kMaxStreams], st_lo[kMaxStreams];
    int st_rows[kMaxStreams];
};

// Stars resident on the device, SoA.
struct StarLaunch {
    const double* star;        // [5][n]: cos b cos l, cos b sin l, sin b, r (kpc),
                               //         reff(r) * 3 sigma / r^3
    long n_stars;
    long star_begin;           // first star of this pass
    int slot, n_conv;
    int block_begin, block_count;
    int out_index;
    int n_streams;             // streams of the slot's model (batched kernel)
    const double* ratio;       // [n_conv] r_i / r_star of the star's context
    const double* kw;          // [n_conv] W_i coeff exp(-dx_i^2 / (2 sigma^2))
    // bg_count > 0 (resampled mixed pass): the background term of each star's
    // convolution from bg_count (r / r_star, weight / r_star^3) pairs of bg_rq
    // instead of the n_conv points.
    const double* bg_rq;
    int bg_count;
    const double* kw3;         // [n_conv] kw_i ratio_i^3 (the mixed pass's stream weights / r_star^3)
    double coarse_pr;          // mixed pass: p r_star above which a stream's exponents are split exactly
    // Models of more than 4 streams run one star pass per group of 4, in
    // order; each star's mixture sum is carried from pass to pass in
    // chain[star] (bit 0 of chain_flags: the first group, which starts the
    // sum with the background; bit 1: the last, which finishes it).  A model
    // of at most 4 streams has chain_flags 3 and no chain.
    double* chain;
    int chain_flags;
};

// Constants of the stream separation pass (separation.c:37-254).
struct SepConst {
    double eps_b;                   // 1 / (1 + sum_s exp(w_s))
    double eps_s[kMaxStreams];      // exp(w_s) / (1 + sum_s exp(w_s))
    double I_bg, I_st[kMaxStreams]; // the evaluation's integrals
    double rot[3][3];               // rotation taking the wedge's normal to +z
    uint64_t x0;                    // drand48 state after srand48(seed)
};

// Second-pass item: the partials of blocks [block_begin, block_begin +
// block_count) summed in block order into row out_index.
struct ReduceItem {
    int block_begin, block_count, out_index, pad;
};

// Host-side table builders (mw_tables.cpp); all FP64 with glibc libm.
void gauss_legendre(double x1, double x2, double* x, double* w, int n);
void gc_to_lb(int wedge, int sgr, double mu, double nu, double* l, double* b);
void lbr_to_xyz(const double* lbr, double* xyz);
void build_region_tables(const INTEGRAL_AREA& ia, int n_conv, int wedge, int sgr,
                         RegionTables* out);
void build_star_soa(const STAR_POINTS& sp, std::vector<double>* out);
// Fills an EvalConst from the model parameters held in ap; returns false
// (with a message) if the parameters cannot be evaluated here.
bool make_eval_const(const ASTRONOMY_PARAMETERS& ap, EvalConst* k, std::string* err);
// Constant slots of a model with any number of streams: stream_groups(S)
// slots, slot g holding streams [4 g, 4 g + 4) with the background (slot 0
// only; later slots have bg_mode kBgZeroQ, which skips the background's
// work) and the mixture's normalisation over every stream.  One slot, equal
// to make_eval_const's, for S <= 4.
inline int stream_groups(int S) { return S <= kMaxStreams ? 1 : (S + kMaxStreams - 1) / kMaxStreams; }
bool make_eval_consts(const ASTRONOMY_PARAMETERS& ap, std::vector<EvalConst>* out, std::string* err);

// The stream separation's output frame: the wedge's normal (stripe_normal /
// sgr_stripe_normal) and the rotation taking it to +z (get_transform).
// Returns false for an unknown sgr_coordinates value.
bool separation_frame(int wedge, int sgr, double normal[3], double rot[3][3]);

// Resampled quadrature.  A region's integral per column is
// sum_p w_p f(r_p) over its r-step x convolution points p, with weights w_p
// (irv * reff * qw) and distances r_p that are the same for every column.
// Folding degree-D Lagrange interpolation on a uniform grid x_m into the
// weights gives K_m = sum_p w_p L_m(x_p), and the column sum becomes
// sum_m K_m f(r(x_m)): the same quadrature up to the interpolation error of f
// between grid nodes, from far fewer density evaluations.  The grid variable
// x is the distance r itself or the magnitude g = 5 log10(1000 r) - 5 + 4.2.
constexpr int kResampleDegree = 15;   // 16-node stencils
constexpr int kResampleRow = 32;      // grid nodes per pseudo r-step row (padding < 32)

// K over the grid g_lo + m delta, m < M; returns M (or -1 on bad input).
// The grid starts pad nodes below the lowest point and ends pad nodes above
// the highest; pad < 0: degree / 2 + 1, every stencil centred on its point's
// interval (a smaller pad shifts the stencils at the ends inwards).
long resample_quadrature(const double* g, const double* w, long n, double delta, int degree,
                         double* g_lo, std::vector<double>* K, int pad = -1);
// The same fold onto a given grid lo + m delta, m < M (the caller leaves
// room for the stencils below and above the points).
long resample_quadrature_at(const double* g, const double* w, long n, double delta, int degree,
                            double lo, long M, std::vector<double>* K);
// Tables in the layout of RegionTables (rq, rstep, lo) holding the
// resampled quadrature of an exact region's rq / rstep on a grid of spacing
// delta in r (in_r) or in g, in pseudo r-steps of kResampleRow nodes with
// irv = reff = 1 and the aux-profile moments sum_k irv reff q_i in the
// first row.  bg_stride 0 or 1: (r(x_m), K_m) pairs (1: from the origin a
// background stride would use, which is the same).  bg_stride q = 2, 4 (in
// r only; the kernels' strided recurrence layout): (K_m, Kb_m) pairs, K the
// stream weights at spacing delta and Kb the background's at spacing
// q delta on every q-th node (0 elsewhere), node m at r = lo + m delta.
// Returns M.
long build_resampled_tables(const std::vector<double>& rq, const std::vector<RStep>& rstep, int n_conv,
                            double delta, bool in_r, int bg_stride, RegionTables* out, int pad = -1);

// Star-pass constants shared by every star of a context: r_i / r_star and
// W_i * coeff * exp(-dx_i^2 / (2 sigma^2)).
void star_conv_tables(int n_conv, std::vector<double>* ratio, std::vector<double>* kw);

}  // namespace mw
