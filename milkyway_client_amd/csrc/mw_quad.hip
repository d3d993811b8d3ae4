// mw_quad.hip -- choice and device copies of the resampled quadrature tables.
//
// The grid spacing is chosen per region and parameter vector from the
// narrowest feature the integrand can have along a sight, in magnitudes:
//   stream s: a Gaussian of width >= sigma_s along the sight, i.e. at least
//             (5 / ln 10) sigma_s / r_max in g;
//   background: 1 / (rg (rg + r0)^3) peaks where the sight passes closest to
//             the Galactic centre, at distance b >= 8.5 min sin(GC angle)
//             min(1, 1/q) and r <= 8.5, a width of at least (5 / ln 10) b / 8.5.
// Interpolating with 16-node stencils at spacing 0.3 w (streams) or 0.2 w
// (background) keeps the per-sight quadrature error below 1e-10 relative
// (measured on these profiles: 3e-13 for a 1 kpc stream at 100 kpc with
// spacing 0.005, 1e-12 for a sight 0.14 kpc from the centre).
#include <hip/hip_runtime.h>

#include <cmath>

#include "mw_context.h"

namespace mw {

namespace {

constexpr double kMagPerLn = 2.1714724095162588;   // 5 / ln 10
constexpr double kDeltaMax = 0.02;

double level_delta(int L) { return kDeltaMax * std::pow(2.0, -0.5 * L); }

}  // namespace

int select_level(const RegionDev& rd, const EvalConst& k, bool resample)
{
    if (!resample || rd.h_rq.empty()) return -1;
    if (k.bg_mode != kBgHernquist && k.bg_mode != kBgHernquistAux) return -1;
    double w = INFINITY;
    for (int s = 0; s < k.n_streams; s++) {
        if (k.on[s] == 0.0 || !(k.inv_2s2[s] > 0.0)) continue;
        const double sigma = std::sqrt(0.5 / k.inv_2s2[s]);
        w = std::min(w, 0.3 * kMagPerLn * sigma / rd.r_max);
    }
    const double b = kSunR0 * rd.min_sin_gc * std::min(1.0, 1.0 / k.q);
    w = std::min(w, 0.2 * kMagPerLn * b / kSunR0);
    const double delta = std::min(w, kDeltaMax);
    if (!(delta > 0.0)) return -1;
    const int L = (int)std::ceil(-2.0 * std::log2(delta / kDeltaMax) - 1e-12);
    if (L >= kQuadLevels) return -1;
    // Grid nodes (with stencil margins and row padding) against exact points.
    const long nodes = (long)std::ceil(rd.g_span / level_delta(L)) + kResampleDegree + 4;
    const long padded = (nodes + kResampleRow - 1) / kResampleRow * kResampleRow;
    if (padded * 5 >= (long)rd.geom.r_steps * rd.n_conv * 4) return -1;
    return L;
}

int region_tables(RegionDev& rd, int level, TableRef* out)
{
    if (level < 0) {
        *out = TableRef{ rd.rq.p, rd.rstep.p, rd.geom.r_steps, rd.n_conv, -1,
                         (long)rd.geom.r_steps * rd.n_conv };
        return MW_OK;
    }
    if (level >= kQuadLevels) return fail(MW_E_INVALID, "no such quadrature level");
    std::lock_guard<std::mutex> lock(rd.level_mu);
    if (!rd.levels[level]) {
        RegionTables t;
        const long M = build_resampled_tables(rd.h_rq, rd.h_rstep, rd.n_conv, level_delta(level), &t);
        if (M < 1) return fail(MW_E_INVALID, "resampled quadrature failed");
        auto q = std::make_unique<ResampledDev>();
        q->r_steps = t.r_steps;
        q->n_conv = t.n_conv;
        q->nodes = M;
        int rc;
        if ((rc = q->rq.reserve(t.rq.size())) || (rc = q->rstep.reserve(t.rstep.size()))) return rc;
        MW_HIP_CHECK(hipMemcpy(q->rq.p, t.rq.data(), sizeof(double) * t.rq.size(), hipMemcpyHostToDevice));
        MW_HIP_CHECK(hipMemcpy(q->rstep.p, t.rstep.data(), sizeof(RStep) * t.rstep.size(), hipMemcpyHostToDevice));
        rd.levels[level] = std::move(q);
    }
    const ResampledDev& q = *rd.levels[level];
    *out = TableRef{ q.rq.p, q.rstep.p, q.r_steps, q.n_conv, level, q.nodes };
    return MW_OK;
}

}  // namespace mw
