// mw_quad.hip -- choice and device copies of the resampled quadrature tables.
//
// The grid spacing is chosen per region and parameter vector from the
// narrowest feature the integrand can have along a sight:
//   stream s: a Gaussian of width >= sigma_s along the sight (in distance),
//             i.e. >= (5 / ln 10) sigma_s / r_max in magnitude;
//   background: 1 / (rg (rg + r0)^3) peaks where the sight passes closest to
//             the Galactic centre.  Along a sight rg^2 = P (r - r_c)^2 + rg_min^2
//             with P = u_x^2 + u_y^2 + u_z^2 / q^2, so the peak is
//             rg_min / sqrt(P) wide: rg_min >= d min(1, 1/q) for a sight that
//             passes the centre at distance d = 8.5 sin(GC angle), and
//             sqrt(P) <= max(1, 1/q), which gives a width of at least
//             b = 8.5 min sin(GC angle) min(q, 1/q) in distance (r <= 8.5 at
//             the peak) and (5 / ln 10) b / 8.5 in magnitude.
// A grid uniform in distance resolves the streams with the same spacing at
// every distance; a grid in magnitude is finer near the Sun, which is where a
// sight passing close to the Galactic centre needs it.  Both are sized, the
// one with fewer nodes is used.  With 16-node stencils at spacing 0.3 sigma
// (distance) or 0.3 w (magnitude) for streams and 0.15 b / 0.2 w for the
// background, the per-sight quadrature error stays below 1e-9 relative for a
// stream crossing the sight at right angles, the narrowest case (measured on
// a full-size point set: 1 kpc streams crossing out to 100 kpc, 8e-10 at
// 0.3 sigma and 9e-13 at 0.2 sigma in distance, 3e-13 at 0.005 mag in
// magnitude; a sight 0.14 kpc from the centre, 1e-12 at 0.005 mag).  Levels
// are spaced by sqrt 2, so the spacing used is 0.21-0.3 sigma.
#include <hip/hip_runtime.h>

#include <cmath>

#include "mw_context.h"

namespace mw {

namespace {

constexpr double kMagPerLn = 2.1714724095162588;   // 5 / ln 10
constexpr double kDeltaMag = 0.02;                  // coarsest magnitude spacing
constexpr double kDeltaKpc = 2.0;                   // coarsest distance spacing
constexpr double kDeltaBgMag = 0.3;                 // coarsest background-grid spacing
// Background and star-pass tables: the fold's lattice ends 4 nodes beyond
// the points (not 8, every stencil centred): the end stencils shift inwards
// over points that carry the convolution's tail weights, and the measured
// worst cases of both rules are unchanged (6.4e-10, 2.5e-10) with 8 fewer
// nodes.
constexpr int kBgPad = 4;

// Spacing of table index t (see kQuadLevels).
double table_delta(int t)
{
    return t < kQuadLevels ? kDeltaMag * std::pow(2.0, -0.5 * t)
                           : kDeltaKpc * std::pow(2.0, -0.5 * (t - kQuadLevels));
}

// Finest level no coarser than delta, or kQuadLevels if none is.
int level_for(double delta, double coarsest)
{
    if (!(delta > 0.0)) return kQuadLevels;
    const double d = std::min(delta, coarsest);
    return std::max(0, (int)std::ceil(-2.0 * std::log2(d / coarsest) - 1e-12));
}

long nodes_for(double span, double delta)
{
    const long n = (long)std::ceil(span / delta) + kResampleDegree + 4;
    return (n + kResampleRow - 1) / kResampleRow * kResampleRow;   // row padding
}

}  // namespace

// Width of the background's peak along one sight u (unit vector) for
// flattening q: rg^2 = P (r - r_c)^2 + rg_min^2 with P = u_x^2 + u_y^2 +
// u_z^2 / q^2, r_c = 8.5 u_x / P and rg_min^2 = 8.5^2 (u_y^2 + u_z^2 / q^2) / P,
// so the peak is rg_min / sqrt(P) wide.  A sight pointing away from the
// centre (u_x <= 0) has r_c <= 0: along r >= 0 it varies on the scale
// rg(0) / sqrt(P) = 8.5 / sqrt(P).
static double sight_bg_width(double ux, double uy, double uz, double q)
{
    const double side = uy * uy + uz * uz / (q * q);
    const double P = ux * ux + side;
    return ux > 0.0 ? kSunR0 * std::sqrt(side) / P : kSunR0 / std::sqrt(P);
}

// In the magnitude variable x = 2.17 ln r the same singularities,
// r = r_c +- i rg_min / sqrt(P), sit at Im x = +-2.17 theta with
// sin theta = rg_min / 8.5 (P r_c^2 + rg_min^2 = 8.5^2): theta is the
// angle, seen from the Sun, between the sight and the nearest point of the
// profile's centre in the q-scaled metric.  A sight pointing away from the
// centre has theta >= pi / 2.
static double sight_bg_theta(double ux, double uy, double uz, double q)
{
    if (!(ux > 0.0)) return 0.5 * M_PI;
    const double side = uy * uy + uz * uz / (q * q);
    return std::asin(std::min(1.0, std::sqrt(side / (ux * ux + side))));
}

void tabulate_bg_width(RegionDev& rd, const double* ux, const double* uy, const double* uz, int mu_steps,
                       int nu_steps)
{
    // 64 x 64 sights, the first and last rows and columns included: the
    // width varies by a few percent between neighbouring samples, which the
    // margin in bg_width covers.
    std::vector<long> rows, cols;
    const int nm = std::min(mu_steps, 64), nn = std::min(nu_steps, 64);
    for (int i = 0; i < nm; i++) rows.push_back(nm > 1 ? (long)i * (mu_steps - 1) / (nm - 1) : 0);
    for (int j = 0; j < nn; j++) cols.push_back(nn > 1 ? (long)j * (nu_steps - 1) / (nn - 1) : 0);
    rd.bg_width_q.assign(kWidthQ, INFINITY);
    rd.bg_theta_q.assign(kWidthQ, INFINITY);
    for (int t = 0; t < kWidthQ; t++) {
        const double q = std::exp2(t / 4.0 - kWidthQ0);
        double w = INFINITY, th = INFINITY;
        for (long r : rows)
            for (long c : cols) {
                const long i = r * nu_steps + c;
                w = std::min(w, sight_bg_width(ux[i], uy[i], uz[i], q));
                th = std::min(th, sight_bg_theta(ux[i], uy[i], uz[i], q));
            }
        rd.bg_width_q[t] = (float)w;
        rd.bg_theta_q[t] = (float)th;
    }
}

double bg_width(const RegionDev& rd, double q)
{
    // Below 0.85 of the smaller of the two tabulated q around it: the width
    // changes by under 19 % from one tabulated q to the next and by a few
    // percent between sampled sights.  Outside the table, the closed-form
    // bound 8.5 min sin(GC angle) min(q, 1/q).
    const double bound = kSunR0 * rd.min_sin_gc * std::min(q, 1.0 / q);
    if (rd.bg_width_q.size() != (size_t)kWidthQ || !(q > 0.0)) return bound;
    const double x = (std::log2(q) + kWidthQ0) * 4.0;
    if (!(x >= 0.0) || !(x <= kWidthQ - 1)) return bound;
    const int t = std::min((int)x, kWidthQ - 2);
    const double w = 0.85 * std::min(rd.bg_width_q[t], rd.bg_width_q[t + 1]);
    return std::max(w, bound);
}

double bg_theta(const RegionDev& rd, double q)
{
    // As bg_width: 0.85 of the smaller tabulated value around q, and never
    // below the closed form asin(min sin(GC angle) min(1, 1/q)) (side / P
    // grows with side, and side >= (1 - u_x^2) min(1, 1/q^2)).
    const double bound = std::asin(std::min(1.0, rd.min_sin_gc * std::min(1.0, 1.0 / q)));
    if (rd.bg_theta_q.size() != (size_t)kWidthQ || !(q > 0.0)) return bound;
    const double x = (std::log2(q) + kWidthQ0) * 4.0;
    if (!(x >= 0.0) || !(x <= kWidthQ - 1)) return bound;
    const int t = std::min((int)x, kWidthQ - 2);
    return std::max(0.85 * std::min(rd.bg_theta_q[t], rd.bg_theta_q[t + 1]), bound);
}

int bg_grid_level(const RegionDev& rd, double q)
{
    // 16-node stencils on a grid uniform in x = 2.17 ln r: spacing 0.2 of
    // the singularities' distance from the real axis (2.17 theta), and at
    // most 0.3 mag, where the profile's r^-4 tail (an exponential of rate
    // 1.84 per mag in x) sets the interpolation error instead.  Measured on
    // the fold of the full-size point set (700 r-steps x 120), every
    // (rg_min, P, r0) from (0.5-8.4 kpc, 0.5-5, 1-14 kpc), towards and away
    // from the centre, each sight at its own theta: relative error of the
    // integral below 6.5e-10, the worst at the cap (4.7e-11 at 0.15 of the
    // distance).
    const double dg = std::min(0.2 * kMagPerLn * bg_theta(rd, q), kDeltaBgMag);
    if (!(dg > 0.0)) return -1;
    const int t = level_for(dg, kDeltaBgMag);
    return t < kBgLevels ? t : -1;
}

bool in_measured_domain(double q, double r0)
{
    return q >= kFoldQMin && q <= kFoldQMax && r0 >= kFoldR0Min && r0 <= kFoldR0Max;
}

void tabulate_sights(RegionDev& rd, const double* ux, const double* uy, const double* uz, int mu_steps,
                     int nu_steps)
{
    std::vector<long> rows, cols;
    const int nm = std::min(mu_steps, 32), nn = std::min(nu_steps, 8);
    for (int i = 0; i < nm; i++) rows.push_back(nm > 1 ? (long)i * (mu_steps - 1) / (nm - 1) : 0);
    for (int j = 0; j < nn; j++) cols.push_back(nn > 1 ? (long)j * (nu_steps - 1) / (nn - 1) : 0);
    rd.sample_u.clear();
    for (long r : rows)
        for (long c : cols) {
            const long i = r * nu_steps + c;
            rd.sample_u.insert(rd.sample_u.end(), { ux[i], uy[i], uz[i] });
        }
    // A sight lies within half a sample cell's diagonal of one of its corners.
    auto u = [&](size_t a, size_t b, int j) { return rd.sample_u[3 * (a * cols.size() + b) + j]; };
    double worst = 1.0;
    for (size_t a = 0; a + 1 < rows.size() || (a == 0 && rows.size() == 1); a++)
        for (size_t b = 0; b + 1 < cols.size() || (b == 0 && cols.size() == 1); b++) {
            const size_t a1 = std::min(a + 1, rows.size() - 1), b1 = std::min(b + 1, cols.size() - 1);
            for (int d = 0; d < 2; d++) {
                const size_t bb = d ? b1 : b, bc = d ? b : b1;
                const double cd = u(a, bb, 0) * u(a1, bc, 0) + u(a, bb, 1) * u(a1, bc, 1) + u(a, bb, 2) * u(a1, bc, 2);
                worst = std::min(worst, std::cos(0.5 * std::acos(std::max(-1.0, std::min(1.0, cd)))));
            }
        }
    rd.sample_cos = worst;
}

bool streams_peak_inside(const RegionDev& rd, const EvalConst& k)
{
    const size_t N = rd.sample_u.size() / 3;
    if (N == 0) return false;
    const double r_lo = rd.r_min, r_hi = rd.r_max;
    for (int s = 0; s < k.n_streams; s++) {
        if (k.on[s] == 0.0 || !(k.inv_2s2[s] > 0.0)) continue;
        // The centre's own sight reaches the axis at the centre: a centre
        // inside the region (its distance in range, its direction within a
        // sample cell of the region's sights) puts the peak inside.
        const double cp[3] = { k.sc[s][0] + kSunR0, k.sc[s][1], k.sc[s][2] };
        const double rc = std::sqrt(cp[0] * cp[0] + cp[1] * cp[1] + cp[2] * cp[2]);
        if (rc >= r_lo && rc <= r_hi) {
            const double near = rd.sample_cos * rc;
            bool in_cell = false;
            for (size_t i = 0; i < N && !in_cell; i++) {
                const double* u = &rd.sample_u[3 * i];
                in_cell = u[0] * cp[0] + u[1] * cp[1] + u[2] * cp[2] >= near;
            }
            if (in_cell) continue;
        }
        // Otherwise the sampled sights: the exponent's smallest value over
        // the region's distances (a = e + t^2 at the nearest end, t = p r - m)
        // over all of them, and over those whose closest approach to the
        // axis lies inside the distances.  The latter must come within a
        // factor 2 of the former: the integral is then carried by sights
        // whose peak the grid resolves.
        double a_all = INFINITY, a_in = INFINITY;
        for (size_t i = 0; i < N; i++) {
            const double* u = &rd.sample_u[3 * i];
            const double d = u[0] * k.sa[s][0] + u[1] * k.sa[s][1] + u[2] * k.sa[s][2];
            const double w0 = (u[0] - d * k.sa[s][0]) * k.ks[s];
            const double w1 = (u[1] - d * k.sa[s][1]) * k.ks[s];
            const double w2 = (u[2] - d * k.sa[s][2]) * k.ks[s];
            const double ww = w0 * w0 + w1 * w1 + w2 * w2;
            const double c0 = k.cperp[s][0], c1 = k.cperp[s][1], c2 = k.cperp[s][2];
            double a;
            bool inside;
            if (ww > 1e-200) {
                const double p = std::sqrt(ww);
                const double m = (w0 * c0 + w1 * c1 + w2 * c2) / p;
                const double e0 = c0 - m * w0 / p, e1 = c1 - m * w1 / p, e2 = c2 - m * w2 / p;
                const double t_lo = p * r_lo - m, t_hi = p * r_hi - m;
                inside = t_lo <= 0.0 && t_hi >= 0.0;
                const double t = inside ? 0.0 : std::min(std::fabs(t_lo), std::fabs(t_hi));
                a = e0 * e0 + e1 * e1 + e2 * e2 + t * t;
            } else {
                // sight parallel to the axis: the same value at every distance
                inside = true;
                a = c0 * c0 + c1 * c1 + c2 * c2;
            }
            a_all = std::min(a_all, a);
            if (inside) a_in = std::min(a_in, a);
        }
        if (!(a_in <= a_all + 1.0)) return false;
    }
    return true;
}

int select_level(const RegionDev& rd, const EvalConst& k, bool resample)
{
    if (!resample || rd.h_rq.empty()) return -1;
    if (k.bg_mode != kBgHernquist && k.bg_mode != kBgHernquistAux) return -1;
    if (!in_measured_domain(k.q, k.r0)) return -1;
    if (!streams_peak_inside(rd, k)) return -1;
    double sigma = INFINITY;
    for (int s = 0; s < k.n_streams; s++) {
        if (k.on[s] == 0.0 || !(k.inv_2s2[s] > 0.0)) continue;
        sigma = std::min(sigma, std::sqrt(0.5 / k.inv_2s2[s]));
    }
    if (sigma < kFoldSigmaMin) return -1;
    const double b = bg_width(rd, k.q);
    // grid in magnitude
    const double dg = std::min(0.3 * kMagPerLn * sigma / rd.r_max, 0.2 * kMagPerLn * b / kSunR0);
    const int Lg = level_for(dg, kDeltaMag);
    // grid in distance
    const double dr = std::min(0.3 * sigma, 0.15 * b);
    const int Lr = level_for(dr, kDeltaKpc);
    long best = -1;
    int t = -1;
    if (Lg < kQuadLevels) {
        best = nodes_for(rd.g_span, table_delta(Lg));
        t = Lg;
    }
    if (Lr < kQuadLevels) {
        const long n = nodes_for(rd.r_span, table_delta(kQuadLevels + Lr));
        if (best < 0 || n < best) {
            best = n;
            t = kQuadLevels + Lr;
        }
    }
    // Keep the exact tables unless the grid saves at least a fifth.
    if (t < 0 || best * 5 >= (long)rd.geom.r_steps * rd.n_conv * 4) return -1;
    return t;
}

double recurrence_step(int level, const EvalConst& k)
{
    if (level < kQuadLevels || level >= kQuadTables) return 0.0;
    const double h = table_delta(level);
    for (int s = 0; s < k.n_streams; s++)
        if (!(k.ks[s] * h <= 0.5)) return 0.0;
    return h;
}

int table_layout(const RegionDev& rd, int level, const EvalConst& k)
{
    // The recurrence lives in the mixed kernels, which exist for the
    // Hernquist profile only; the other profiles run the FP64 kernels, which
    // read (r, weight) tables.
    if (k.bg_mode != kBgHernquist) return 0;
    const double h = recurrence_step(level, k);
    if (!(h > 0.0)) return 0;
    const double b = bg_width(rd, k.q);
    const double bg_delta = std::min(0.15 * b, kDeltaKpc);
    if (4.0 * h <= bg_delta) return 3;
    if (2.0 * h <= bg_delta) return 2;
    return 1;
}

namespace {

// Vector-instruction cycles per sight and grid node of the integral kernels
// (from their AMD ISA and the VALU counters): a node of the recurrence
// layouts, each stream's recurrence step, a background evaluation (two
// quarter-rate FP64 transcendentals among its 15.5 instructions) and each
// stream's FP32 exponential where there is no recurrence.
constexpr double kCostNode = 1.5, kCostRecStream = 3.25, kCostBg = 21.5, kCostExpStream = 7.0;

long padded_rows(long n) { return (n + kResampleRow - 1) / kResampleRow * kResampleRow; }

// Nodes of the resampled table `level` over the region (as the builder
// makes them: the fold's lattice, padded to whole rows).
long table_nodes(const RegionDev& rd, int level, int stride)
{
    const int half = kResampleDegree / 2;
    if (level < kQuadLevels) {
        const double d = table_delta(level);
        return padded_rows((long)std::ceil(rd.g_span / d) + 2 * half + 3);
    }
    const double d = table_delta(level);
    const double span = rd.r_span + (half + 1) * stride * d;
    const long M = std::max((long)std::ceil(span / d) + half + 2,
                            ((long)std::ceil(span / (stride * d)) + half + 1) * stride + 1);
    return padded_rows(M);
}

// The background grid's nodes (the kernel runs over these, not the row
// padding).
long bg_table_nodes(const RegionDev& rd, int bg_level)
{
    const double d = kDeltaBgMag * std::pow(2.0, -0.5 * bg_level);
    return std::max((long)std::ceil(rd.g_span / d) + 2 * kBgPad + 1, (long)kResampleDegree + 1);
}

}  // namespace

GridPlan plan_grid(const RegionDev& rd, const EvalConst& k, bool resample)
{
    GridPlan g;
    g.level = select_level(rd, k, resample);
    if (g.level < 0) return g;
    g.layout = table_layout(rd, g.level, k);
    if (k.bg_mode != kBgHernquist || !rd.bg_grid) return g;
    const int S = k.n_streams;
    double cost;
    if (g.layout > 0) {
        const int q = layout_bg_stride(g.layout);
        const long n = table_nodes(rd, g.level, q);
        cost = n * (kCostNode + kCostRecStream * S) + (double)((n + q - 1) / q) * kCostBg;
    } else {
        cost = table_nodes(rd, g.level, 1) * (kCostBg + kCostExpStream * S);
    }
    // The split: each switched-on stream on a distance grid spaced for it
    // alone, and the background grid.  0.4 sigma_s (the recurrence's step
    // k_s h is then 0.34 at most): measured on the fold of the full-size
    // point set, sigma 0.6-12 kpc crossing at 3-90 kpc, a stream's integral
    // is within 1.4e-9 (7.7e-10 at 0.3 sigma, 1e-8 at 0.45).
    const int bl = bg_grid_level(rd, k.q);
    if (bl < 0) return g;
    GridPlan split;
    split.layout = kLayoutSplit;
    split.bg_level = bl;
    double cost_split = bg_table_nodes(rd, bl) * kCostBg;
    int finest = 0;
    for (int s = 0; s < k.n_streams; s++) {
        if (k.on[s] == 0.0 || !(k.inv_2s2[s] > 0.0)) continue;
        const double sigma = std::sqrt(0.5 / k.inv_2s2[s]);
        const int L = level_for(0.4 * sigma, kDeltaKpc);
        if (L >= kQuadLevels || !(k.ks[s] * table_delta(kQuadLevels + L) <= 0.5)) return g;
        split.st_level[s] = L;
        finest = std::max(finest, L);
        cost_split += table_nodes(rd, kQuadLevels + L, 1) * (kCostNode + kCostRecStream);
    }
    // The plan's own table: the finest stream grid (the 2 kpc grid with no
    // stream on), which the launch's node count refers to.
    split.level = kQuadLevels + finest;
    if (!(recurrence_step(split.level, k) > 0.0)) return g;
    return cost_split < 0.9 * cost ? split : g;
}

// pairs: the table holds (r, weight) pairs, whose weights alone are also
// uploaded, contiguous.
static int upload_table(const RegionTables& t, long M, bool pairs, std::unique_ptr<ResampledDev>* slot)
{
    auto q = std::make_unique<ResampledDev>();
    q->r_steps = t.r_steps;
    q->n_conv = t.n_conv;
    q->nodes = M;
    q->lo = t.lo;
    int rc;
    if ((rc = q->rq.reserve(t.rq.size())) || (rc = q->rstep.reserve(t.rstep.size()))) return rc;
    MW_HIP_CHECK(hipMemcpy(q->rq.p, t.rq.data(), sizeof(double) * t.rq.size(), hipMemcpyHostToDevice));
    MW_HIP_CHECK(hipMemcpy(q->rstep.p, t.rstep.data(), sizeof(RStep) * t.rstep.size(), hipMemcpyHostToDevice));
    if (pairs) {
        std::vector<double> w(t.rq.size() / 2);
        for (size_t i = 0; i < w.size(); i++) w[i] = t.rq[2 * i + 1];
        if ((rc = q->w.reserve(w.size()))) return rc;
        MW_HIP_CHECK(hipMemcpy(q->w.p, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
    }
    *slot = std::move(q);
    return MW_OK;
}

int region_tables(RegionDev& rd, const GridPlan& g, TableRef* out)
{
    const int level = g.level, layout = g.layout;
    if (level < 0) {
        const long pts = (long)rd.geom.r_steps * rd.n_conv;
        *out = TableRef{ rd.rq.p, rd.rstep.p, rd.geom.r_steps, rd.n_conv, -1, pts, 0, 0.0, pts };
        return MW_OK;
    }
    if (level >= kQuadTables || layout < 0 || layout >= kTableLayouts || (layout > 0 && level < kQuadLevels) ||
        (layout == kLayoutSplit) != (g.bg_level >= 0) || g.bg_level >= kBgLevels)
        return fail(MW_E_INVALID, "no such quadrature table");
    std::lock_guard<std::mutex> lock(rd.level_mu);
    // Layouts 1 (stride 1) and 4 (the streams' table) are the (r, weight)
    // table of layout 0.
    const int stride = layout_bg_stride(layout);
    std::unique_ptr<ResampledDev>& slot = rd.levels[level][stride > 1 ? layout : 0];
    if (!slot) {
        RegionTables t;
        const long M = build_resampled_tables(rd.h_rq, rd.h_rstep, rd.n_conv, table_delta(level),
                                              level >= kQuadLevels, stride > 1 ? stride : 0, &t);
        if (M < 1) return fail(MW_E_INVALID, "resampled quadrature failed");
        int rc;
        if ((rc = upload_table(t, M, stride <= 1, &slot))) return rc;
    }
    const ResampledDev& q = *slot;
    *out = TableRef{ q.rq.p, q.rstep.p, q.r_steps, q.n_conv, level, q.nodes, layout, q.lo,
                     stride > 1 ? (q.nodes + stride - 1) / stride : q.nodes };
    if (layout == kLayoutSplit) {
        for (int s = 0; s < kMaxStreams; s++) {
            if (g.st_level[s] < 0) continue;
            const int ls = kQuadLevels + g.st_level[s];
            if (ls >= kQuadTables) return fail(MW_E_INVALID, "no such quadrature table");
            std::unique_ptr<ResampledDev>& sslot = rd.levels[ls][0];
            if (!sslot) {
                RegionTables t;
                const long M = build_resampled_tables(rd.h_rq, rd.h_rstep, rd.n_conv, table_delta(ls), true, 0, &t);
                if (M < 1) return fail(MW_E_INVALID, "resampled quadrature failed");
                int rc;
                if ((rc = upload_table(t, M, true, &sslot))) return rc;
            }
            if (sslot->n_conv != kResampleRow || !sslot->w.p) return fail(MW_E_INVALID, "stream table rows");
            out->st_rq[s] = sslot->w.p;
            out->st_step[s] = table_delta(ls);
            out->st_lo[s] = sslot->lo;
            out->st_rows[s] = sslot->r_steps;
            out->stream_terms += (long)sslot->r_steps * sslot->n_conv;
        }
        std::unique_ptr<ResampledDev>& bslot = rd.bg_levels[g.bg_level];
        if (!bslot) {
            RegionTables t;
            const long M = build_resampled_tables(rd.h_rq, rd.h_rstep, rd.n_conv,
                                                  kDeltaBgMag * std::pow(2.0, -0.5 * g.bg_level), false, 0, &t,
                                                  kBgPad);
            if (M < 1) return fail(MW_E_INVALID, "resampled quadrature failed");
            int rc;
            if ((rc = upload_table(t, M, true, &bslot))) return rc;
        }
        out->bg_rq = bslot->rq.p;
        out->bg_count = bslot->nodes;   // the fold's lattice; the row padding carries no weight
        out->bg_nodes = out->bg_count;
    }
    return MW_OK;
}

void tabulate_star_theta(mw_ctx& c, const double* soa, long n)
{
    // sin^2 theta = side / (u_x^2 + side) for a sight towards the centre,
    // side = u_y^2 + u_z^2 / q^2; a sight pointing away has theta >= pi / 2.
    c.star_sin2_q.assign(kWidthQ, 1.0);
    if (n < 1) return;
    for (int t = 0; t < kWidthQ; t++) {
        const double q = std::exp2(t / 4.0 - kWidthQ0);
        const double iq2 = 1.0 / (q * q);
        double lo = 1.0;
        for (long i = 0; i < n; i++) {
            const double ux = soa[i], uy = soa[n + i], uz = soa[2 * n + i];
            if (!(ux > 0.0)) continue;
            const double side = fma(uz * uz, iq2, uy * uy);
            lo = std::min(lo, side / fma(ux, ux, side));
        }
        c.star_sin2_q[t] = lo;
    }
}

namespace {

// Nodes of the star background table at spacing d (the fold's lattice over
// the convolution offsets).
long star_bg_nodes(const mw_ctx& c, double d)
{
    const auto mm = std::minmax_element(c.conv_dx.begin(), c.conv_dx.end());
    const double span = *mm.second - *mm.first;
    return std::max((long)std::ceil(span / d) + 2 * kBgPad + 1, (long)kResampleDegree + 1);
}

double star_bg_delta(int level) { return kDeltaBgMag * std::pow(2.0, -0.5 * level); }

}  // namespace

int star_bg_level(const mw_ctx& c, double q, double r0)
{
    if (c.n_stars < 1 || c.conv_dx.size() < 2 || c.star_sin2_q.size() != (size_t)kWidthQ || !(q > 0.0))
        return -1;
    if (!in_measured_domain(q, r0)) return -1;
    // sin^2 theta falls as q grows, so the entry at the tabulated q just
    // above q bounds it from below.
    const double x = (std::log2(q) + kWidthQ0) * 4.0;
    if (!(x >= 0.0) || !(x <= kWidthQ - 1)) return -1;
    const int t = std::min((int)std::ceil(x - 1e-12), kWidthQ - 1);
    const double theta = std::asin(std::sqrt(std::max(0.0, c.star_sin2_q[t])));
    // A star's window (3.6 mag of convolution offsets) sees the pointwise
    // interpolation error with little of the cancellation a whole sight
    // gives the integral: 0.1 of the singularities' distance, measured on
    // the convolution tables (n = 60, 120) for every (rg_min, P, r0, r_star)
    // from (0.3-8 kpc, 0.5-11, 1-14 kpc, 1-60 kpc), towards and away from
    // the centre: relative error of a star's background term below 2.5e-10
    // (7.8e-9 at 0.15, 2.4e-7 at the integral's 0.2); the same holds for r0
    // up to 30 kpc.
    const double dg = std::min(0.1 * kMagPerLn * theta, kDeltaBgMag);
    if (!(dg > 0.0)) return -1;
    const int level = level_for(dg, kDeltaBgMag);
    if (level >= kBgLevels) return -1;
    if (star_bg_nodes(c, star_bg_delta(level)) >= (long)c.conv_dx.size()) return -1;
    return level;
}

int star_bg_table(mw_ctx& c, int level, const double** rq, int* count)
{
    *rq = nullptr;
    *count = 0;
    if (level < 0) return MW_OK;
    if (level >= kBgLevels) return fail(MW_E_INVALID, "no such star table");
    std::lock_guard<std::mutex> lock(c.star_bg_mu);
    std::unique_ptr<ResampledDev>& slot = c.star_bg[level];
    if (!slot) {
        std::vector<double> K;
        double lo;
        const double d = star_bg_delta(level);
        const long M = resample_quadrature(c.conv_dx.data(), c.conv_w3.data(), (long)c.conv_dx.size(), d,
                                           kResampleDegree, &lo, &K, kBgPad);
        if (M < 1) return fail(MW_E_INVALID, "star quadrature failed");
        std::vector<double> pairs(2 * M);
        for (long m = 0; m < M; m++) {
            pairs[2 * m] = std::pow(10.0, (lo + m * d) / 5.0);   // r / r_star at offset lo + m d
            pairs[2 * m + 1] = K[m];
        }
        auto q = std::make_unique<ResampledDev>();
        q->nodes = M;
        q->lo = lo;
        int rc;
        if ((rc = q->rq.reserve(pairs.size()))) return rc;
        MW_HIP_CHECK(hipMemcpy(q->rq.p, pairs.data(), sizeof(double) * pairs.size(), hipMemcpyHostToDevice));
        slot = std::move(q);
    }
    *rq = slot->rq.p;
    *count = (int)slot->nodes;
    return MW_OK;
}

}  // namespace mw
