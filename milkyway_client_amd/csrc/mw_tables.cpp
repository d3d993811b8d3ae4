// mw_tables.cpp -- parameter-independent per-wedge tables and per-evaluation
// constants, built on the host in FP64 with glibc libm.
//
// These tables feed the device kernels.  They are computed once per wedge
// (and cached on the device), so they follow the reference's own formulas and
// libm calls to reproduce its values:
//   survey geometry   milkyway_separation/cpu/atSurveyGeometry.c:26-240, :247-412
//   wedge node/incl.  milkyway_separation/cpu/stCoords.c:109-130
//   stream centre     milkyway_separation/cpu/stCoords.c:34-48
//   quadrature        milkyway_separation/cpu/numericalIntegration.c:67-104
//   r-step tables     milkyway_separation/cpu/evaluation_optimized.c:159-188, :360-422
//   constants         milkyway_separation/cpu/evaluation_optimized.c:54-133, :639-646
//   separation frame  milkyway_separation/cpu/stCoords.c:142-185, stVector.c:109-140,
//                     atSurveyGeometry.c:414-460
// Compiled with -ffp-contract=off so the host arithmetic rounds like the
// reference's scalar SSE2 code.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

#include "mw_internal.h"

namespace mw {
namespace {

constexpr double kPi = 3.1415926535897932384626433832795028841971693993751;
constexpr double kTwoPi = 6.2831853071795864769252867665590057683943387987502;
constexpr double kDeg = 180.0 / kPi;

// dmod / dsign macros of atSurveyGeometry.h:27-29
double fmod_floor(double a, double b)
{
    if (b == 0.0) return a;
    return (a * b > 0.0) ? a - b * std::floor(a / b) : a + b * std::floor(-a / b);
}

void wrap(double* angle, double lo, double hi)
{
    while (*angle < lo) *angle += 360.0;
    while (*angle >= hi) *angle -= 360.0;
}

// Latitude into [-90, 90] and longitude into [0, 360) (atBound2).
void wrap_lat_lon(double* lat, double* lon)
{
    wrap(lat, -180.0, 180.0);
    if (std::fabs(*lat) > 90.0) {
        *lat = 180.0 - *lat;
        *lon += 180;
    }
    wrap(lat, -180.0, 180.0);
    wrap(lon, 0.0, 360.0);
    if (std::fabs(*lat) == 90.0) *lon = 0.;
}

// Great circle (node, inclination) to equatorial; inputs and outputs in
// degrees.  Degree/radian conversions keep the reference's (x * pi) / 180 and
// (x * 180) / pi order.
void great_circle_to_equatorial(double mu, double nu, double node, double inc, double* ra,
                                double* dec)
{
    mu = mu * kPi / 180.0;
    nu = nu * kPi / 180.0;
    node = node * kPi / 180.0;
    inc = inc * kPi / 180.0;
    const double cx = std::cos(mu - node) * std::cos(nu);
    const double cy = std::sin(mu - node) * std::cos(nu);
    const double cz = std::sin(nu);
    const double ey = cy * std::cos(inc) - cz * std::sin(inc);
    const double ez = cy * std::sin(inc) + cz * std::cos(inc);
    *ra = std::atan2(ey, cx) + node;
    *dec = std::asin(ez);
    *ra = *ra * 180.0 / kPi;
    *dec = *dec * 180.0 / kPi;
    wrap_lat_lon(dec, ra);
}

// Equatorial (degrees) to galactic (degrees) through the SLALIB FK4 matrix.
void equatorial_to_galactic(double ra, double dec, double* l, double* b)
{
    static const double m[3][3] = {
        { -0.054875539726, -0.873437108010, -0.483834985808 },
        {  0.494109453312, -0.444829589425,  0.746982251810 },
        { -0.867666135858, -0.198076386122,  0.455983795705 },
    };
    ra = ra * kPi / 180.0;
    dec = dec * kPi / 180.0;
    const double cd = std::cos(dec);
    const double v[3] = { std::cos(ra) * cd, std::sin(ra) * cd, std::sin(dec) };
    double g[3];
    for (int row = 0; row < 3; row++) {
        double acc = 0.0;
        for (int c = 0; c < 3; c++) acc += m[row][c] * v[c];
        g[row] = acc;
    }
    const double rho = std::sqrt(g[0] * g[0] + g[1] * g[1]);
    double lon = (rho != 0.0) ? std::atan2(g[1], g[0]) : 0.0;
    double lat = (g[2] != 0.0) ? std::atan2(g[2], rho) : 0.0;
    double t = fmod_floor(lon, kTwoPi);
    lon = (t >= 0.0) ? t : t + kTwoPi;
    t = fmod_floor(lat, kTwoPi);
    lat = (std::fabs(t) < kPi) ? t : t - ((lat < 0.0) ? -kTwoPi : kTwoPi);
    *l = lon * 180.0 / kPi;
    *b = lat * 180.0 / kPi;
    wrap_lat_lon(b, l);
}

// Sagittarius great circles (atSurveyGeometry.c:247-272).
void gc_to_sgr(double mu, double nu, int wedge, double* lam, double* beta)
{
    const double rpd = 3.141592653589793 / 180;
    mu = mu * rpd;
    nu = nu * rpd;
    const double x = std::cos(mu) * std::cos(nu);
    const double y = -std::sin(nu);
    const double z = std::sin(mu) * std::cos(nu);
    *lam = std::atan2(y, x) / rpd + 2.5 * wedge;
    if (*lam < 0) *lam = *lam + 360;
    *beta = std::asin(z) / rpd;
}

// Sagittarius (lambda, beta) to galactic (atSurveyGeometry.c:276-412).  The
// matrix entries are evaluated with the reference's expressions so that they
// carry the same rounding.
void sgr_to_gal(double lam, double beta, double* l, double* b)
{
    const double rpd = 3.141592653589793 / 180;
    const double ph = (180 + 3.75) * rpd, th = (90 - 13.46) * rpd, ps = (180 + 14.111534) * rpd;
    const double ct = std::cos(th), st = std::sin(th), cf = std::cos(ph), sf = std::sin(ph);
    const double cp = std::cos(ps), sp = std::sin(ps);
    auto sq = [](double v) { return std::pow(v, 2); };
    const double d8 = sq(cp) * sq(cf) * sq(ct) + sq(cp) * sq(cf) * sq(st) + sq(ct) * sq(sp) * sq(sf)
                    + sq(sp) * sq(cf) * sq(ct) + sq(sp) * sq(cf) * sq(st) + sq(ct) * sq(cp) * sq(sf)
                    + sq(st) * sq(sf) * sq(cp) + sq(st) * sq(sf) * sq(sp);
    const double d4a = sq(ct) * sq(sf) + sq(cf) * sq(ct) + sq(cf) * sq(st) + sq(st) * sq(sf);
    const double d4b = sq(cp) * sq(ct) + sq(sp) * sq(st) + sq(cp) * sq(st) + sq(sp) * sq(ct);
    const double m11 = -(ct * sf * sp - sq(ct) * cf * cp - cp * sq(st) * cf) / d8;
    const double m12 = -(ct * sf * cp + sq(ct) * cf * sp + sp * sq(st) * cf) / d8;
    const double m13 = (st * sf) / d4a;
    const double m21 = (ct * cf * sp + sq(ct) * cp * sf + cp * sq(st) * sf) / d8;
    const double m22 = -(-ct * cf * cp + sq(ct) * sp * sf + sp * sq(st) * sf) / d8;
    const double m23 = -(st * cf) / d4a;
    const double m31 = (sp * st) / d4b;
    const double m32 = (cp * st) / d4b;
    const double m33 = ct / (sq(ct) + sq(st));

    if (beta > 90) {
        beta = 90 - (beta - 90);
        lam = lam + 180;
        if (lam > 360) lam = lam - 360;
    }
    if (beta < -90) {
        beta = -90 - (beta + 90);
        lam = lam + 180;
        if (lam > 360) lam = lam - 360;
    }
    if (lam < 0) lam = lam + 360;
    beta = (beta + 90) * rpd;
    const double z2 = std::cos(beta);
    double x2 = 0.0, y2 = 0.0;
    // The quadrant cases of the reference, each with its own expression.
    if (lam == 0 || lam == 360) {
        x2 = std::sin(beta);
    } else if (lam == 90) {
        y2 = std::sin(beta);
    } else if (lam == 180) {
        x2 = -std::sin(beta);
    } else if (lam == 270) {
        y2 = -std::sin(beta);
    } else if (lam < 90 || (lam > 270 && lam < 360)) {
        const double t = std::tan(lam * rpd);
        x2 = std::sqrt((1 - std::cos(beta) * std::cos(beta)) / (1 + t * t));
        y2 = x2 * t;
    } else if (lam < 180) {
        const double t = std::tan(lam * rpd);
        y2 = std::sqrt((1 - std::cos(beta) * std::cos(beta)) / (1 / (t * t) + 1));
        x2 = y2 / t;
    } else if (lam < 270) {
        const double t = std::tan(lam * rpd);
        x2 = std::sqrt((1 - std::cos(beta) * std::cos(beta)) / (1 + t * t));
        y2 = -(x2 * t);
        x2 = -x2;
    }
    const double gx = m11 * x2 + m12 * y2 + m13 * z2;
    const double gy = m21 * x2 + m22 * y2 + m23 * z2;
    const double gz = m31 * x2 + m32 * y2 + m33 * z2;
    if (gz > 1) {
        *l = 0;
        *b = 90;
    } else {
        *b = std::asin(gz) / rpd;
        *l = std::atan2(gy, gx) / rpd;
        if (*l < 0) *l = *l + 360;
    }
}

// Probability constants at one distance r' (kpc): evaluation_optimized.c:159-188.
// Writes reff * 3 sigma / r'^3 and, per convolution point, (r_point, qw_r3_N)
// and the magnitude g.
void distance_constants(double rp, int n, const double* dx, const double* w, double coeff,
                        double* reff_xr_rp3, double* rq, double* g_out)
{
    const double sig[3] = { 0.9402, 1.6171, 23.5877 };
    const double gp = 5.0 * (std::log10(rp * 1000) - 1.0) + kAbsM;
    const double reff = sig[0] / (std::exp(sig[1] * (gp - sig[2])) + 1);
    *reff_xr_rp3 = reff * 3 * kStdev / (rp * rp * rp);
    for (int i = 0; i < n; i++) {
        const double g = gp + dx[i];
        const double r = std::pow(10.0, (g - kAbsM) / 5.0 + 1.0) / 1000.0;
        const double e = (g - gp) * (g - gp) / (2 * kStdev * kStdev);
        rq[2 * i] = r;
        rq[2 * i + 1] = w[i] * (r * r * r) * (coeff * std::exp(-e));
        g_out[i] = g;
    }
}

template <typename F>
void parallel_rows(long rows, F&& body)
{
    unsigned hw = std::thread::hardware_concurrency();
    long nt = std::max<long>(1, std::min<long>(rows / 8, hw ? hw : 1));
    if (nt <= 1) {
        for (long i = 0; i < rows; i++) body(i);
        return;
    }
    std::vector<std::thread> pool;
    for (long t = 0; t < nt; t++) {
        pool.emplace_back([&, t] {
            for (long i = t; i < rows; i += nt) body(i);
        });
    }
    for (auto& th : pool) th.join();
}

}  // namespace

void gauss_legendre(double x1, double x2, double* x, double* w, int n)
{
    // Newton iteration on P_n through the three-term recurrence
    // (numericalIntegration.c:67-104; tolerance EPS = 3e-11 from milkyway.h:54).
    const int half = (n + 1) / 2;
    const double mid = 0.5 * (x2 + x1), hw = 0.5 * (x2 - x1);
    for (int i = 1; i <= half; i++) {
        double z = std::cos(kPi * (i - 0.25) / (n + 0.5));
        double zprev, dp;
        for (;;) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            dp = n * (z * p1 - p2) / (z * z - 1.0);
            zprev = z;
            z = zprev - p1 / dp;
            if (!(std::fabs(z - zprev) > 3.0e-11)) break;
        }
        x[i - 1] = mid - hw * z;
        x[n - i] = mid + hw * z;
        w[i - 1] = 2.0 * hw / ((1.0 - z * z) * dp * dp);
        w[n - i] = w[i - 1];
    }
}

void gc_to_lb(int wedge, int sgr, double mu, double nu, double* l, double* b)
{
    if (sgr == 0) {
        const double node = 185.0 - 90.0;
        const double eta = wedge * 2.5 - 57.5 - (wedge > 46 ? 180.0 : 0.0);
        double ra, dec;
        great_circle_to_equatorial(mu, nu, node, eta + 32.5, &ra, &dec);
        equatorial_to_galactic(ra, dec, l, b);
    } else {
        double lam, beta;
        gc_to_sgr(mu, nu, wedge, &lam, &beta);
        sgr_to_gal(lam, beta, l, b);
    }
}

namespace {

// Unit vector towards galactic (l, b) in degrees (lbToXyz, stCoords.c:177-185).
void lb_unit(double l, double b, double* v)
{
    l = l / kDeg;
    b = b / kDeg;
    v[0] = std::cos(l) * std::cos(b);
    v[1] = std::sin(l) * std::cos(b);
    v[2] = std::sin(b);
}

// Survey (longitude, latitude) to equatorial, degrees (atSurveyToEq,
// atSurveyGeometry.c:414-440): survey centre RA 185, Dec 32.5.
void survey_to_equatorial(double slong, double slat, double* ra, double* dec)
{
    slong = slong * kPi / 180.0;
    slat = slat * kPi / 180.0;
    const double anode = (185.0 - 90.0) * kPi / 180.0;
    const double eta_pole = 32.5 * kPi / 180.0;
    const double x1 = -std::sin(slong);
    const double y1 = std::cos(slat + eta_pole) * std::cos(slong);
    const double z1 = std::sin(slat + eta_pole) * std::cos(slong);
    *ra = std::atan2(y1, x1) + anode;
    *dec = std::asin(z1);
    *ra = *ra * 180.0 / kPi;
    *dec = *dec * 180.0 / kPi;
    wrap_lat_lon(dec, ra);
}

}  // namespace

bool separation_frame(int wedge, int sgr, double normal[3], double rot[3][3])
{
    if (sgr == 0) {
        // stripe_normal (stCoords.c:142-150): the pole of the stripe's great
        // circle, at survey latitude 90 + eta.
        const double eta = wedge * 2.5 - 57.5 - (wedge > 46 ? 180.0 : 0.0);
        double ra, dec, l, b;
        survey_to_equatorial(0.0, 90.0 + eta, &ra, &dec);
        equatorial_to_galactic(ra, dec, &l, &b);
        lb_unit(l, b, normal);
    } else if (sgr == 1) {
        // sgr_stripe_normal (stCoords.c:153-174): the cross product of the
        // directions of (lambda, 0) and (lambda, 90), lambda = 2.5 wedge.
        double l1, b1, l2, b2, a[3], c[3];
        sgr_to_gal(wedge * 2.5, 0.0, &l1, &b1);
        sgr_to_gal(wedge * 2.5, 90.0, &l2, &b2);
        lb_unit(l1, b1, a);
        lb_unit(l2, b2, c);
        normal[0] = a[1] * c[2] - a[2] * c[1];
        normal[1] = a[2] * c[0] - a[0] * c[2];
        normal[2] = a[0] * c[1] - a[1] * c[0];
    } else {
        return false;
    }
    // get_transform(normal, +z) (stVector.c:109-140): the rotation about
    // normal x z by the angle between them, through its quaternion.
    const double t[3] = { 0.0, 0.0, 1.0 };
    double axis[3] = { normal[1] * t[2] - normal[2] * t[1], normal[2] * t[0] - normal[0] * t[2],
                       normal[0] * t[1] - normal[1] * t[0] };
    const double an = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
    axis[0] /= an;
    axis[1] /= an;
    axis[2] /= an;
    const double fn = std::sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
    const double tn = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    const double dp = normal[0] * t[0] + normal[1] * t[1] + normal[2] * t[2];
    const double angle = std::acos(dp / (fn * tn));
    const double sa = std::sin(angle / 2);
    const double x = axis[0] * sa, y = axis[1] * sa, z = axis[2] * sa, w = std::cos(angle / 2);
    rot[0][0] = 1 - 2 * (y * y + z * z);
    rot[0][1] = 2 * (x * y - z * w);
    rot[0][2] = 2 * (x * z + y * w);
    rot[1][0] = 2 * (x * y + z * w);
    rot[1][1] = 1 - 2 * (x * x + z * z);
    rot[1][2] = 2 * (y * z - x * w);
    rot[2][0] = 2 * (x * z - y * w);
    rot[2][1] = 2 * (y * z + x * w);
    rot[2][2] = 1 - 2 * (x * x + y * y);
    return true;
}

void lbr_to_xyz(const double* lbr, double* xyz)
{
    // stCoords.c:34-48; x keeps the reference's cancelling form.
    const double r0 = kSunR0;
    const double bs = std::sin(lbr[1] / kDeg), ls = std::sin(lbr[0] / kDeg);
    const double bc = std::cos(lbr[1] / kDeg), lc = std::cos(lbr[0] / kDeg);
    xyz[2] = lbr[2] * bs;
    const double zp = lbr[2] * bc;
    const double d = std::sqrt(r0 * r0 + zp * zp - 2 * r0 * zp * lc);
    xyz[0] = (zp * zp - r0 * r0 - d * d) / (2 * r0);
    xyz[1] = zp * ls;
}

void build_region_tables(const INTEGRAL_AREA& ia, int n, int wedge, int sgr, RegionTables* t)
{
    const int R = ia.r_steps, M = ia.mu_steps, NV = ia.nu_steps;
    t->r_steps = R;
    t->mu_steps = M;
    t->nu_steps = NV;
    t->n_conv = n;
    t->n_cols = (long)M * NV;

    std::vector<double> x(n), w(n), dx(n), g(n);
    gauss_legendre(-1.0, 1.0, x.data(), w.data(), n);
    for (int i = 0; i < n; i++) dx[i] = 3 * kStdev * x[i];
    const double coeff = 1 / (kStdev * std::sqrt(2 * kPi));

    t->rq.assign((size_t)R * n * 2, 0.0);
    t->rstep.assign(R, RStep{});
    for (int k = 0; k < R; k++) {
        const double log_r = ia.r_min + (k * ia.r_step_size);
        const double r = std::pow(10.0, (log_r - 14.2) / 5.0);
        const double rn = std::pow(10.0, (log_r + ia.r_step_size - 14.2) / 5.0);
        RStep& s = t->rstep[k];
        s.irv = (((rn * rn * rn) - (r * r * r)) / 3.0) * ia.mu_step_size / kDeg;
        double* rq = &t->rq[(size_t)k * n * 2];
        distance_constants((rn + r) / 2.0, n, dx.data(), w.data(), coeff, &s.reff_xr_rp3, rq,
                           g.data());
        double q0 = 0, q1 = 0, q2 = 0;
        for (int i = 0; i < n; i++) {
            const double qw = rq[2 * i + 1];
            q0 += qw;
            q1 += qw * g[i];
            q2 += qw * (g[i] * g[i]);
        }
        s.q0 = q0;
        s.q1 = q1;
        s.q2 = q2;
    }

    std::vector<double> nu_mid(NV), ids(NV);
    for (int j = 0; j < NV; j++) {
        const double nu = ia.nu_min + (j * ia.nu_step_size);
        ids[j] = std::cos((90 - nu - ia.nu_step_size) / kDeg) - std::cos((90 - nu) / kDeg);
        nu_mid[j] = nu + 0.5 * ia.nu_step_size;
    }
    const long C = t->n_cols;
    t->col.assign((size_t)4 * C, 0.0);
    double* col = t->col.data();
    parallel_rows(M, [&](long m) {
        const double mu = ia.mu_min + (m * ia.mu_step_size);
        for (int j = 0; j < NV; j++) {
            double l, b;
            gc_to_lb(wedge, sgr, mu + 0.5 * ia.mu_step_size, nu_mid[j], &l, &b);
            const double bs = std::sin(b / kDeg), ls = std::sin(l / kDeg);
            const double bc = std::cos(b / kDeg), lc = std::cos(l / kDeg);
            const long c = m * NV + j;
            col[c] = bc * lc;
            col[C + c] = bc * ls;
            col[2 * C + c] = bs;
            col[3 * C + c] = ids[j];
        }
    });
}

void build_star_soa(const STAR_POINTS& sp, std::vector<double>* out)
{
    const long N = sp.number_stars;
    out->assign((size_t)5 * (N > 0 ? N : 1), 0.0);
    double* s = out->data();
    const double sig[3] = { 0.9402, 1.6171, 23.5877 };
    for (long i = 0; i < N; i++) {
        const double* st = sp.stars[i];
        const double bs = std::sin(st[1] / kDeg), ls = std::sin(st[0] / kDeg);
        const double bc = std::cos(st[1] / kDeg), lc = std::cos(st[0] / kDeg);
        const double r = st[2];
        const double gp = 5.0 * (std::log10(r * 1000) - 1.0) + kAbsM;
        const double reff = sig[0] / (std::exp(sig[1] * (gp - sig[2])) + 1);
        s[i] = bc * lc;
        s[N + i] = bc * ls;
        s[2 * N + i] = bs;
        s[3 * N + i] = r;
        s[4 * N + i] = reff * 3 * kStdev / (r * r * r);
    }
}

void star_conv_tables(int n, std::vector<double>* ratio, std::vector<double>* kw)
{
    std::vector<double> x(n), w(n);
    gauss_legendre(-1.0, 1.0, x.data(), w.data(), n);
    const double coeff = 1 / (kStdev * std::sqrt(2 * kPi));
    ratio->resize(n);
    kw->resize(n);
    for (int i = 0; i < n; i++) {
        const double dx = 3 * kStdev * x[i];
        // r_i = 10^((g' + dx - 4.2) / 5 + 1) / 1000 = r_star * 10^(dx / 5)
        (*ratio)[i] = std::pow(10.0, dx / 5.0);
        (*kw)[i] = w[i] * (coeff * std::exp(-(dx * dx) / (2 * kStdev * kStdev)));
    }
}

namespace {

// Fold points [p0, p1) into acc (one per grid node).  The stencil is
// computed in double (relative error ~1e-16 per weight) and accumulated in
// long double.
void fold_points(const double* g, const double* w, long p0, long p1, double delta, int degree,
                 double lo, long M, const std::vector<double>& bw, std::vector<long double>* acc_out)
{
    std::vector<long double>& acc = *acc_out;
    double lw[32];
    for (long p = p0; p < p1; p++) {
        // Stencil of degree + 1 nodes around g[p], centred on its interval.
        const double x = (g[p] - lo) / delta;
        long i0 = (long)std::floor(x) - (degree - 1) / 2;
        if (i0 < 0) i0 = 0;
        if (i0 > M - degree - 1) i0 = M - degree - 1;
        const double u = x - (double)i0;
        int hit = -1;
        for (int i = 0; i <= degree; i++)
            if (u == (double)i) hit = i;
        if (hit >= 0) {
            acc[i0 + hit] += w[p];
            continue;
        }
        double den = 0.0;
        for (int i = 0; i <= degree; i++) {
            lw[i] = bw[i] / (u - (double)i);
            den += lw[i];
        }
        const double scale = w[p] / den;
        for (int i = 0; i <= degree; i++) acc[i0 + i] += (long double)(lw[i] * scale);
    }
}

}  // namespace

long resample_quadrature_at(const double* g, const double* w, long n, double delta, int degree,
                            double lo, long M, std::vector<double>* K)
{
    if (n < 1 || !(delta > 0.0) || degree < 1 || degree > 31 || M < degree + 1) return -1;
    // Barycentric weights of degree + 1 equally spaced nodes: (-1)^i C(D, i)
    // (exact in double up to degree 31).
    std::vector<double> bw(degree + 1);
    double c = 1.0;
    for (int i = 0; i <= degree; i++) {
        bw[i] = (i % 2 ? -c : c);
        c = c * (degree - i) / (i + 1);
    }
    // A full-size region has ~1e5 points: the fold runs over a fixed number
    // of point chunks on up to that many threads, each chunk into its own
    // accumulator, and the chunks are added in chunk order, so the result
    // does not depend on the machine's thread count.  (A table is built the
    // first time a model needs it, on the thread that plans the launch.)
    const int chunks = n >= 4096 ? 16 : 1;
    std::vector<std::vector<long double>> acc(chunks, std::vector<long double>(M, 0.0L));
    auto run = [&](int ch) {
        fold_points(g, w, n * ch / chunks, n * (ch + 1) / chunks, delta, degree, lo, M, bw, &acc[ch]);
    };
    const int threads = std::max(1, std::min<int>(chunks, (int)std::thread::hardware_concurrency()));
    if (threads == 1) {
        for (int ch = 0; ch < chunks; ch++) run(ch);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; t++)
            pool.emplace_back([&, t] {
                for (int ch = t; ch < chunks; ch += threads) run(ch);
            });
        for (std::thread& th : pool) th.join();
    }
    K->assign(M, 0.0);
    for (long m = 0; m < M; m++) {
        long double sum = 0.0L;
        for (int ch = 0; ch < chunks; ch++) sum += acc[ch][m];
        (*K)[m] = (double)sum;
    }
    return M;
}

long resample_quadrature(const double* g, const double* w, long n, double delta, int degree,
                         double* g_lo, std::vector<double>* K, int pad)
{
    if (n < 1 || !(delta > 0.0) || degree < 1 || degree > 31) return -1;
    double gmin = g[0], gmax = g[0];
    for (long i = 1; i < n; i++) {
        gmin = std::min(gmin, g[i]);
        gmax = std::max(gmax, g[i]);
    }
    if (pad < 0) pad = degree / 2 + 1;
    const double lo = gmin - pad * delta;
    const long M = std::max((long)std::ceil((gmax - lo) / delta) + pad + 1, (long)degree + 1);
    *g_lo = lo;
    return resample_quadrature_at(g, w, n, delta, degree, lo, M, K);
}

long build_resampled_tables(const std::vector<double>& rq, const std::vector<RStep>& rstep, int n_conv,
                            double delta, bool in_r, int bg_stride, RegionTables* out, int pad)
{
    const long R = (long)rstep.size();
    const long N = R * n_conv;
    if (N < 1 || bg_stride < 0 || bg_stride > 4 || (bg_stride > 0 && !in_r)) return -1;
    std::vector<double> g(N), w(N);
    double Q0 = 0, Q1 = 0, Q2 = 0;
    for (long k = 0; k < R; k++) {
        const double f = rstep[k].irv * rstep[k].reff_xr_rp3;
        Q0 += f * rstep[k].q0;
        Q1 += f * rstep[k].q1;
        Q2 += f * rstep[k].q2;
        for (long j = 0; j < n_conv; j++) {
            const double r = rq[2 * (k * n_conv + j)];
            g[k * n_conv + j] = in_r ? r : 5.0 * (std::log10(r * 1000.0) - 1.0) + kAbsM;
            w[k * n_conv + j] = f * rq[2 * (k * n_conv + j) + 1];
        }
    }
    double lo;
    long M;
    std::vector<double> K, Kb;
    if (bg_stride == 0) {
        M = resample_quadrature(g.data(), w.data(), N, delta, kResampleDegree, &lo, &K, pad);
        if (M < 1) return -1;
    } else {
        // Two folds on one lattice: the streams at spacing delta, the
        // background at q delta, from a common origin that leaves room for
        // the coarser stencils, so background node j is stream node q j.
        const int q = bg_stride;
        const int half = kResampleDegree / 2;
        const double gmin = *std::min_element(g.begin(), g.end());
        const double gmax = *std::max_element(g.begin(), g.end());
        lo = gmin - (half + 1) * q * delta;
        const long Mb = (long)std::ceil((gmax - lo) / (q * delta)) + half + 2;
        M = std::max((long)std::ceil((gmax - lo) / delta) + half + 2, (Mb - 1) * q + 1);
        if (resample_quadrature_at(g.data(), w.data(), N, delta, kResampleDegree, lo, M, &K) < 1) return -1;
        if (q == 1) {
            Kb = K;
        } else if (resample_quadrature_at(g.data(), w.data(), N, q * delta, kResampleDegree, lo, Mb, &Kb) < 1) {
            return -1;
        }
    }
    const int row = kResampleRow;
    const long Rp = (M + row - 1) / row;
    out->r_steps = (int)Rp;
    out->n_conv = row;
    out->lo = lo;
    out->rq.assign((size_t)2 * Rp * row, 0.0);
    if (bg_stride <= 1) {
        // Node distances: the grid itself in r, or r(g) = 10^((g - 4.2) / 5 + 1) / 1000.
        auto node_r = [&](long m) {
            const double x = lo + m * delta;
            return in_r ? x : std::pow(10.0, (x - kAbsM) / 5.0 + 1.0) / 1000.0;
        };
        for (long m = 0; m < Rp * row; m++) {
            // padding entries repeat the last node and carry no weight
            out->rq[2 * m] = node_r(m < M ? m : M - 1);
            out->rq[2 * m + 1] = m < M ? K[m] : 0.0;
        }
    } else {
        // (stream weight, background weight) per node; node m sits at
        // r = lo + m delta, which the kernel computes.
        const int q = bg_stride;
        for (long m = 0; m < Rp * row; m++) {
            out->rq[2 * m] = m < M ? K[m] : 0.0;
            const long j = m / q;
            out->rq[2 * m + 1] = (m % q == 0 && j < (long)Kb.size()) ? Kb[j] : 0.0;
        }
    }
    out->rstep.assign(Rp, RStep{});
    for (long k = 0; k < Rp; k++) {
        out->rstep[k].irv = 1.0;
        out->rstep[k].reff_xr_rp3 = 1.0;
    }
    // The aux profile's per-point polynomial in g is exact as three moments.
    out->rstep[0].q0 = Q0;
    out->rstep[0].q1 = Q1;
    out->rstep[0].q2 = Q2;
    return M;
}

bool make_eval_const(const ASTRONOMY_PARAMETERS& ap, EvalConst* k, std::string* err)
{
    std::memset(k, 0, sizeof(*k));
    const int S = ap.number_streams;
    if (S < 0 || S > kMaxStreams) {
        *err = "one constant slot holds at most 4 streams (batches and the request server take at most "
               "4 per model; mw_ctx_eval and the drop-in entry points take more, in groups of 4)";
        return false;
    }
    if (ap.aux_bg_profile != 0 && ap.aux_bg_profile != 1) {
        *err = "aux_bg_profile invalid";
        return false;
    }
    if (ap.sgr_coordinates != 0 && ap.sgr_coordinates != 1) {
        *err = "sgr_coordinates not valid";
        return false;
    }
    if (ap.number_background_parameters < (ap.aux_bg_profile ? 7 : 4)) {
        *err = "too few background parameters";
        return false;
    }
    if (S > 0 && ap.number_stream_parameters < 5) {
        *err = "streams need 5 parameters";
        return false;
    }
    const double* bp = ap.background_parameters;
    k->alpha = bp[0];
    k->q = bp[1];
    k->r0 = bp[2];
    const double delta = bp[3];
    k->alpha_delta3 = 3 - k->alpha + delta;
    k->inv_q2 = 1.0 / (k->q * k->q);
    if (ap.aux_bg_profile == 1) {
        k->bg_a = bp[4];
        k->bg_b = bp[5];
        k->bg_c = bp[6];
    }
    if (k->q == 0)
        k->bg_mode = kBgZeroQ;
    else if (k->alpha == 1 && delta == 1)
        k->bg_mode = ap.aux_bg_profile == 1 ? kBgHernquistAux : kBgHernquist;
    else
        k->bg_mode = kBgGeneral;
    k->n_streams = S;
    for (int s = 0; s < S; s++) {
        const double* p = ap.stream_parameters[s];
        double l, b;
        gc_to_lb(ap.wedge, ap.sgr_coordinates, p[0], 0, &l, &b);
        const double lbr[3] = { l, b, p[1] };
        lbr_to_xyz(lbr, k->sc[s]);
        k->sa[s][0] = std::sin(p[2]) * std::cos(p[3]);
        k->sa[s][1] = std::sin(p[2]) * std::sin(p[3]);
        k->sa[s][2] = std::cos(p[2]);
        const double sigma = p[4];
        const bool off = sigma > -0.0001 && sigma < 0.0001;
        k->on[s] = off ? 0.0 : 1.0;
        k->inv_2s2[s] = off ? 0.0 : 1.0 / (2.0 * sigma * sigma);
        k->ks[s] = std::sqrt(k->inv_2s2[s] * 1.4426950408889634074);   // log2(e)
        const double cp[3] = { k->sc[s][0] + kSunR0, k->sc[s][1], k->sc[s][2] };
        const double d = cp[0] * k->sa[s][0] + cp[1] * k->sa[s][1] + cp[2] * k->sa[s][2];
        for (int j = 0; j < 3; j++) k->cperp[s][j] = (cp[j] - d * k->sa[s][j]) * k->ks[s];
    }
    k->exp_bg_w = std::exp(ap.background_weight);
    double sum = k->exp_bg_w;
    for (int s = 0; s < S; s++) {
        k->exp_st_w[s] = std::exp(ap.stream_weights[s]);
        sum += std::exp(ap.stream_weights[s]);
    }
    k->sum_exp_w = sum * 0.001;
    return true;
}

bool make_eval_consts(const ASTRONOMY_PARAMETERS& ap, std::vector<EvalConst>* out, std::string* err)
{
    const int S = ap.number_streams;
    if (S < 0) {
        *err = "negative number_streams";
        return false;
    }
    const int G = stream_groups(S);
    out->assign(G, EvalConst{});
    for (int g = 0; g < G; g++) {
        // the streams [kMaxStreams g, kMaxStreams g + n) as a model of their own
        ASTRONOMY_PARAMETERS sub = ap;
        const int first = kMaxStreams * g;
        sub.number_streams = std::min(kMaxStreams, S - first);
        if (S > 0) {
            sub.stream_weights = ap.stream_weights + first;
            sub.stream_parameters = ap.stream_parameters + first;
        }
        if (!make_eval_const(sub, &(*out)[g], err)) return false;
    }
    if (G > 1) {
        // the mixture's normalisation over every stream, summed in stream
        // order as a single group would (evaluation_optimized.c:639-646)
        double sum = std::exp(ap.background_weight);
        for (int s = 0; s < S; s++) sum += std::exp(ap.stream_weights[s]);
        for (int g = 0; g < G; g++) {
            (*out)[g].sum_exp_w = sum * 0.001;
            // groups after the first carry no background: its work is skipped
            if (g > 0) (*out)[g].bg_mode = kBgZeroQ;
        }
    }
    return true;
}

}  // namespace mw

extern "C" long mw_resample_weights_padded(const double* g, const double* w, long n, double delta,
                                           int degree, int pad, double* g_lo, double* K, long capacity)
{
    if (!g || !w || !g_lo || pad < 0) return -1;
    std::vector<double> k;
    const long M = mw::resample_quadrature(g, w, n, delta, degree, g_lo, &k, pad);
    if (M < 1) return -1;
    if (K && capacity >= M) std::memcpy(K, k.data(), sizeof(double) * M);
    return M;
}

extern "C" long mw_resample_weights(const double* g, const double* w, long n, double delta, int degree,
                                    double* g_lo, double* K, long capacity)
{
    return mw_resample_weights_padded(g, w, n, delta, degree, degree / 2 + 1, g_lo, K, capacity);
}
