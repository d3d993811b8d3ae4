/*
 * mw_oracle.c -- CPU FP64 restatement of the reference separation likelihood.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  Only tests/,
 * the smoke check and the CPU-baseline leg of bench.py load this code.
 *
 * It restates milkyway_separation/cpu/evaluation_optimized.c (DOUBLE_PRECISION
 * build) and the helpers it calls, keeping their formulas, libm calls,
 * left-to-right operation order (including the order that the reference's
 * unparenthesised macros produce), loop order and Kahan summation, so that a
 * serial run reproduces the reference's rounding.  Build with
 * -ffp-contract=off and without -ffast-math.  Differences from the reference,
 * all deliberate:
 *   - no file-scope state (the reference keeps constants and the xyz scratch
 *     in globals, evaluation_optimized.c:46-51); everything lives in an
 *     OrcConst built per call, so the code is reentrant;
 *   - invalid inputs return an error code instead of printing and carrying on;
 *   - checkpointing and BOINC progress (compiled out in the reference, since
 *     MILKYWAY is never defined) are absent.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mw_oracle.h"

/* include/milkyway.h:46-57 -- pi, deg, EPS, PI, D2PI, DPI */
#define ORC_PI 3.1415926535897932384626433832795028841971693993751
#define ORC_D2PI 6.2831853071795864769252867665590057683943387987502
#define ORC_DEG (180.0 / ORC_PI)
#define ORC_GL_EPS 3.0e-11

/* evaluation_optimized.c:41-46 */
#define ORC_STDEV 0.6
#define ORC_LBR_R 8.5
#define ORC_ABSM 4.2
static const double orc_sigmoid[3] = { 0.9402, 1.6171, 23.5877 };

/* stCoords.c:30 */
static const double orc_sun_r0 = 8.5;

#define ORC_MAX_STREAMS 64

/* ------------------------------------------------------------------------ */
/* Quadrature: numericalIntegration.c:67-104                                 */
/* ------------------------------------------------------------------------ */

void orc_gauss_legendre(double x1, double x2, double* x, double* w, int n)
{
    int m = (n + 1) / 2;
    double xm = 0.5 * (x2 + x1);
    double xl = 0.5 * (x2 - x1);
    for (int i = 1; i <= m; i++) {
        double z = cos(ORC_PI * (i - 0.25) / (n + 0.5));
        double z1, pp, p1, p2, p3;
        do {
            p1 = 1.0;
            p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            z1 = z;
            z = z1 - p1 / pp;
        } while (fabs(z - z1) > ORC_GL_EPS);
        x[i - 1] = xm - xl * z;
        x[n - i] = xm + xl * z;
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
        w[n - i] = w[i - 1];
    }
}

/* ------------------------------------------------------------------------ */
/* Survey geometry: atSurveyGeometry.c:26-240, :247-412; stCoords.c:109-130  */
/* ------------------------------------------------------------------------ */

/* atSurveyGeometry.h:27-29 */
static double orc_dmod(double a, double b)
{
    return (b != 0.0) ? ((a * b > 0.0) ? a - b * floor(a / b) : a + b * floor(-a / b)) : a;
}

static double orc_dsign(double a, double b) { return (b < 0.0) ? -a : a; }

/* atSurveyGeometry.c:87-102 */
static void orc_bound(double* angle, double min, double max)
{
    while (*angle < min) *angle += 360.0;
    while (*angle >= max) *angle -= 360.0;
}

/* atSurveyGeometry.c:104-119 */
static void orc_bound2(double* theta, double* phi)
{
    orc_bound(theta, -180.0, 180.0);
    if (fabs(*theta) > 90.0) {
        *theta = 180.0 - *theta;
        *phi += 180;
    }
    orc_bound(theta, -180.0, 180.0);
    orc_bound(phi, 0.0, 360.0);
    if (fabs(*theta) == 90.0) *phi = 0.;
}

/* atSurveyGeometry.c:26-64.  The reference's at_deg2Rad / at_rad2Deg macros
 * expand unparenthesised, so "x * at_deg2Rad" is (x * DPI) / 180.0. */
void orc_gc_to_eq(double amu, double anu, double* ra, double* dec, double anode, double ainc)
{
    amu = amu * ORC_PI / 180.0;
    anu = anu * ORC_PI / 180.0;
    anode = anode * ORC_PI / 180.0;
    ainc = ainc * ORC_PI / 180.0;
    double x2 = cos(amu - anode) * cos(anu);
    double y2 = sin(amu - anode) * cos(anu);
    double z2 = sin(anu);
    double x1 = x2;
    double y1 = y2 * cos(ainc) - z2 * sin(ainc);
    double z1 = y2 * sin(ainc) + z2 * cos(ainc);
    *ra = atan2(y1, x1) + anode;
    *dec = asin(z1);
    *ra = *ra * 180.0 / ORC_PI;
    *dec = *dec * 180.0 / ORC_PI;
    orc_bound2(dec, ra);
}

/* atSurveyGeometry.c:212-240 with slaDcs2c :135-143, slaDmxv :170-191,
 * slaDcc2s :121-132, slaDranrm :203-209, slaDrange :194-200. */
static void orc_sla_eqgal(double dr, double dd, double* dl, double* db)
{
    static const double rmat[3][3] = {
        { -0.054875539726, -0.873437108010, -0.483834985808 },
        {  0.494109453312, -0.444829589425,  0.746982251810 },
        { -0.867666135858, -0.198076386122,  0.455983795705 },
    };
    double v1[3], v2[3];
    double cosb = cos(dd);
    v1[0] = cos(dr) * cosb;
    v1[1] = sin(dr) * cosb;
    v1[2] = sin(dd);
    for (int j = 0; j < 3; j++) {
        double w = 0.0;
        for (int i = 0; i < 3; i++) w += rmat[j][i] * v1[i];
        v2[j] = w;
    }
    double x = v2[0], y = v2[1], z = v2[2];
    double r = sqrt(x * x + y * y);
    *dl = (r != 0.0) ? atan2(y, x) : 0.0;
    *db = (z != 0.0) ? atan2(z, r) : 0.0;

    double w = orc_dmod(*dl, ORC_D2PI);
    *dl = (w >= 0.0) ? w : w + ORC_D2PI;
    w = orc_dmod(*db, ORC_D2PI);
    *db = (fabs(w) < ORC_PI) ? w : w - orc_dsign(ORC_D2PI, *db);
}

/* atSurveyGeometry.c:66-83 */
void orc_eq_to_gal(double ra, double dec, double* glong, double* glat)
{
    ra = ra * ORC_PI / 180.0;
    dec = dec * ORC_PI / 180.0;
    orc_sla_eqgal(ra, dec, glong, glat);
    *glong = *glong * 180.0 / ORC_PI;
    *glat = *glat * 180.0 / ORC_PI;
    orc_bound2(glat, glong);
}

/* atSurveyGeometry.c:247-272 */
void orc_gc_to_sgr(double mu, double nu, int wedge, double* lamda, double* beta)
{
    double radpdeg = 3.141592653589793 / 180;
    mu = mu * radpdeg;
    nu = nu * radpdeg;
    double x = cos(mu) * cos(nu);
    double y = -sin(nu);
    double z = sin(mu) * cos(nu);
    *lamda = atan2(y, x);
    *lamda = *lamda / radpdeg;
    *lamda = *lamda + 2.5 * wedge;
    if (*lamda < 0) *lamda = *lamda + 360;
    *beta = asin(z);
    *beta = *beta / radpdeg;
}

/* atSurveyGeometry.c:276-412.  The rotation entries are the reference's
 * expanded expressions (inverse of the Majewski lbr->sgr rotation with
 * phi = 183.75 deg, theta = 76.54 deg, psi = 194.111534 deg), evaluated with
 * the same pow(.,2) calls and the same term order so their rounding matches. */
void orc_sgr_to_gal(double lamda, double beta, double* l, double* b)
{
    double radpdeg = 3.141592653589793 / 180;
    double phi = (180 + 3.75) * radpdeg;
    double theta = (90 - 13.46) * radpdeg;
    double psi = (180 + 14.111534) * radpdeg;
    double ct = cos(theta), st = sin(theta), cf = cos(phi), sf = sin(phi);
    double cp = cos(psi), sp = sin(psi);
    /* denominators as written in the reference (each is 1 up to rounding) */
    double den8 = pow(cp, 2) * pow(cf, 2) * pow(ct, 2) + pow(cp, 2) * pow(cf, 2) * pow(st, 2)
                + pow(ct, 2) * pow(sp, 2) * pow(sf, 2) + pow(sp, 2) * pow(cf, 2) * pow(ct, 2)
                + pow(sp, 2) * pow(cf, 2) * pow(st, 2) + pow(ct, 2) * pow(cp, 2) * pow(sf, 2)
                + pow(st, 2) * pow(sf, 2) * pow(cp, 2) + pow(st, 2) * pow(sf, 2) * pow(sp, 2);
    double den4a = pow(ct, 2) * pow(sf, 2) + pow(cf, 2) * pow(ct, 2) + pow(cf, 2) * pow(st, 2)
                 + pow(st, 2) * pow(sf, 2);
    double den4b = pow(cp, 2) * pow(ct, 2) + pow(sp, 2) * pow(st, 2) + pow(cp, 2) * pow(st, 2)
                 + pow(sp, 2) * pow(ct, 2);
    double rot11 = -(ct * sf * sp - pow(ct, 2) * cf * cp - cp * pow(st, 2) * cf) / den8;
    double rot12 = -(ct * sf * cp + pow(ct, 2) * cf * sp + sp * pow(st, 2) * cf) / den8;
    double rot13 = (st * sf) / den4a;
    double rot21 = (ct * cf * sp + pow(ct, 2) * cp * sf + cp * pow(st, 2) * sf) / den8;
    double rot22 = -(-ct * cf * cp + pow(ct, 2) * sp * sf + sp * pow(st, 2) * sf) / den8;
    double rot23 = -(st * cf) / den4a;
    double rot31 = (sp * st) / den4b;
    double rot32 = (cp * st) / den4b;
    double rot33 = ct / (pow(ct, 2) + pow(st, 2));
    double x2 = 0.0, y2 = 0.0;

    if (beta > 90) {
        beta = 90 - (beta - 90);
        lamda = lamda + 180;
        if (lamda > 360) lamda = lamda - 360;
    }
    if (beta < -90) {
        beta = -90 - (beta + 90);
        lamda = lamda + 180;
        if (lamda > 360) lamda = lamda - 360;
    }
    if (lamda < 0) lamda = lamda + 360;

    beta = beta + 90;
    beta = beta * radpdeg;
    double z2 = cos(beta);

    if (lamda == 0) {
        x2 = sin(beta);
        y2 = 0;
    } else if (lamda < 90) {
        lamda = lamda * radpdeg;
        x2 = sqrt((1 - cos(beta) * cos(beta)) / (1 + tan(lamda) * tan(lamda)));
        y2 = x2 * tan(lamda);
    } else if (lamda == 90) {
        x2 = 0;
        y2 = sin(beta);
    } else if (lamda < 180) {
        lamda = lamda * radpdeg;
        y2 = sqrt((1 - cos(beta) * cos(beta)) / (1 / (tan(lamda) * tan(lamda)) + 1));
        x2 = y2 / tan(lamda);
    } else if (lamda == 180) {
        x2 = -sin(beta);
        y2 = 0;
    } else if (lamda < 270) {
        lamda = lamda * radpdeg;
        x2 = sqrt((1 - cos(beta) * cos(beta)) / (1 + tan(lamda) * tan(lamda)));
        y2 = x2 * tan(lamda);
        x2 = -x2;
        y2 = -y2;
    } else if (lamda == 270) {
        x2 = 0;
        y2 = -sin(beta);
    } else if (lamda < 360) {
        lamda = lamda * radpdeg;
        x2 = sqrt((1 - cos(beta) * cos(beta)) / (1 + tan(lamda) * tan(lamda)));
        y2 = x2 * tan(lamda);
    } else if (lamda == 360) {
        x2 = sin(beta);
        y2 = 0;
    }

    double x1 = rot11 * x2 + rot12 * y2 + rot13 * z2;
    double y1 = rot21 * x2 + rot22 * y2 + rot23 * z2;
    double z1 = rot31 * x2 + rot32 * y2 + rot33 * z2;

    if (z1 > 1) {
        *l = 0;
        *b = 90;
    } else {
        *b = asin(z1);
        *b = *b / radpdeg;
        *l = atan2(y1, x1);
        *l = *l / radpdeg;
        if (*l < 0) *l = *l + 360;
    }
}

/* stCoords.c:109-122: wedge_eta + survey centre dec */
double orc_wedge_incl(int wedge)
{
    double d = 2.5;
    double eta = wedge * d - 57.5 - (wedge > 46 ? 180.0 : 0.0);
    return eta + 32.5;
}

/* stCoords.c:125-130 */
double orc_get_node(void) { return 185.0 - 90.0; }

void orc_gc_to_lb(int wedge, int sgr, double mu, double nu, double* l, double* b)
{
    if (sgr == 0) {
        double ra, dec;
        orc_gc_to_eq(mu, nu, &ra, &dec, orc_get_node(), orc_wedge_incl(wedge));
        orc_eq_to_gal(ra, dec, l, b);
    } else {
        double lamda, beta;
        orc_gc_to_sgr(mu, nu, wedge, &lamda, &beta);
        orc_sgr_to_gal(lamda, beta, l, b);
    }
}

/* stCoords.c:34-48 (the x formula cancels; kept as written) */
void orc_lbr2xyz(const double* lbr, double* xyz)
{
    double r0 = orc_sun_r0;
    double bsin = sin(lbr[1] / ORC_DEG);
    double lsin = sin(lbr[0] / ORC_DEG);
    double bcos = cos(lbr[1] / ORC_DEG);
    double lcos = cos(lbr[0] / ORC_DEG);
    xyz[2] = lbr[2] * bsin;
    double zp = lbr[2] * bcos;
    double d = sqrt(r0 * r0 + zp * zp - 2 * r0 * zp * lcos);
    xyz[0] = (zp * zp - r0 * r0 - d * d) / (2 * r0);
    xyz[1] = zp * lsin;
}

/* ------------------------------------------------------------------------ */
/* Per-evaluation constants: evaluation_optimized.c:54-133                   */
/* ------------------------------------------------------------------------ */

typedef struct {
    int n_streams, convolve, aux;
    double alpha, q, r0, delta, coeff, alpha_delta3;
    double bg_a, bg_b, bg_c;
    double stream_a[ORC_MAX_STREAMS][3];
    double stream_c[ORC_MAX_STREAMS][3];
    double stream_sigma[ORC_MAX_STREAMS];
    double stream_sigma_sq2[ORC_MAX_STREAMS];
    double* gl_x;   /* [convolve] */
    double* gl_w;   /* [convolve] */
    double* dx;     /* [convolve] */
    double* xyz;    /* [convolve][3] scratch of the integrand */
} OrcConst;

static int orc_init_constants(const ASTRONOMY_PARAMETERS* ap, OrcConst* k)
{
    memset(k, 0, sizeof(*k));
    if (ap->number_streams < 0 || ap->number_streams > ORC_MAX_STREAMS) return 1;
    if (ap->convolve <= 0) return 1;
    if (ap->aux_bg_profile != 0 && ap->aux_bg_profile != 1) return 1;
    if (ap->sgr_coordinates != 0 && ap->sgr_coordinates != 1) return 1;
    if (ap->number_background_parameters < (ap->aux_bg_profile ? 7 : 4)) return 1;

    k->n_streams = ap->number_streams;
    k->convolve = ap->convolve;
    k->aux = ap->aux_bg_profile;
    k->alpha = ap->background_parameters[0];
    k->q = ap->background_parameters[1];
    k->r0 = ap->background_parameters[2];
    k->delta = ap->background_parameters[3];
    if (ap->aux_bg_profile == 1) {
        k->bg_a = ap->background_parameters[4];
        k->bg_b = ap->background_parameters[5];
        k->bg_c = ap->background_parameters[6];
    }
    k->coeff = 1 / (ORC_STDEV * sqrt(2 * ORC_PI));
    k->alpha_delta3 = 3 - k->alpha + k->delta;

    for (int i = 0; i < ap->number_streams; i++) {
        const double* sp = ap->stream_parameters[i];
        double l, b, lbr[3];
        k->stream_sigma[i] = sp[4];
        k->stream_sigma_sq2[i] = 2.0 * k->stream_sigma[i] * k->stream_sigma[i];
        orc_gc_to_lb(ap->wedge, ap->sgr_coordinates, sp[0], 0, &l, &b);
        lbr[0] = l;
        lbr[1] = b;
        lbr[2] = sp[1];
        orc_lbr2xyz(lbr, k->stream_c[i]);
        k->stream_a[i][0] = sin(sp[2]) * cos(sp[3]);
        k->stream_a[i][1] = sin(sp[2]) * sin(sp[3]);
        k->stream_a[i][2] = cos(sp[2]);
    }

    int n = ap->convolve;
    k->gl_x = (double*)malloc(sizeof(double) * n);
    k->gl_w = (double*)malloc(sizeof(double) * n);
    k->dx = (double*)malloc(sizeof(double) * n);
    k->xyz = (double*)malloc(sizeof(double) * 3 * n);
    orc_gauss_legendre(-1.0, 1.0, k->gl_x, k->gl_w, n);
    /* "3 * stdev * qgaus_X[i]" */
    for (int i = 0; i < n; i++) k->dx[i] = 3 * ORC_STDEV * k->gl_x[i];
    return 0;
}

static void orc_free_constants(OrcConst* k)
{
    free(k->gl_x);
    free(k->gl_w);
    free(k->dx);
    free(k->xyz);
}

/* evaluation_optimized.c:159-188.  "reff_value * xr / rPrime3" with
 * xr = 3 * stdev expands to ((reff_value * 3) * stdev) / rPrime3. */
static void orc_prob_constants(const OrcConst* k, double coords, double* r_point,
                               double* r_in_mag, double* r_in_mag2, double* qw_r3_N,
                               double* reff_xr_rp3)
{
    double gPrime = 5.0 * (log10(coords * 1000) - 1.0) + ORC_ABSM;
    double exp_result = exp(orc_sigmoid[1] * (gPrime - orc_sigmoid[2]));
    double reff_value = orc_sigmoid[0] / (exp_result + 1);
    double rPrime3 = coords * coords * coords;
    *reff_xr_rp3 = reff_value * 3 * ORC_STDEV / rPrime3;
    for (int i = 0; i < k->convolve; i++) {
        double g = gPrime + k->dx[i];
        r_in_mag[i] = g;
        r_in_mag2[i] = g * g;
        r_point[i] = pow(10.0, (g - ORC_ABSM) / 5.0 + 1.0) / 1000.0;
        double r3 = r_point[i] * r_point[i] * r_point[i];
        double exponent = (g - gPrime) * (g - gPrime) / (2 * ORC_STDEV * ORC_STDEV);
        double N = k->coeff * exp(-exponent);
        qw_r3_N[i] = k->gl_w[i] * r3 * N;
    }
}

/* evaluation_optimized.c:190-299 with the point given as sin/cos-ready
 * (l, b) in degrees.  k->xyz is the scratch the reference keeps global. */
static void orc_probabilities(const OrcConst* k, const double* r_point, const double* r_in_mag,
                              const double* r_in_mag2, const double* qw_r3_N, double reff_xr_rp3,
                              const double* integral_point, double* bg_prob, double* st_prob)
{
    double* xyz = k->xyz;
    int n = k->convolve;
    double bsin = sin(integral_point[1] / ORC_DEG);
    double lsin = sin(integral_point[0] / ORC_DEG);
    double bcos = cos(integral_point[1] / ORC_DEG);
    double lcos = cos(integral_point[0] / ORC_DEG);

    if (k->q == 0) {
        /* The reference leaves its xyz scratch untouched here, so its streams
         * then read uninitialised or stale memory.  The oracle fills the
         * scratch with this point's positions instead (the only defined
         * choice); the HIP library does the same. */
        for (int i = 0; i < n; i++) {
            xyz[3 * i + 2] = r_point[i] * bsin;
            double zp = r_point[i] * bcos;
            xyz[3 * i + 0] = zp * lcos - ORC_LBR_R;
            xyz[3 * i + 1] = zp * lsin;
        }
        *bg_prob = -1;
    } else {
        *bg_prob = 0;
        if (k->alpha == 1 && k->delta == 1) {
            for (int i = 0; i < n; i++) {
                xyz[3 * i + 2] = r_point[i] * bsin;
                double zp = r_point[i] * bcos;
                xyz[3 * i + 0] = zp * lcos - ORC_LBR_R;
                xyz[3 * i + 1] = zp * lsin;
                double rg = sqrt(xyz[3 * i] * xyz[3 * i] + xyz[3 * i + 1] * xyz[3 * i + 1]
                                 + (xyz[3 * i + 2] * xyz[3 * i + 2]) / (k->q * k->q));
                double rs = rg + k->r0;
                if (k->aux == 1) {
                    double h_prob = qw_r3_N[i] / (rg * rs * rs * rs);
                    double aux_prob = qw_r3_N[i] * (k->bg_a * r_in_mag2[i] + k->bg_b * r_in_mag[i] + k->bg_c);
                    *bg_prob += h_prob + aux_prob;
                } else {
                    *bg_prob += qw_r3_N[i] / (rg * rs * rs * rs);
                }
            }
        } else {
            for (int i = 0; i < n; i++) {
                xyz[3 * i + 2] = r_point[i] * bsin;
                double zp = r_point[i] * bcos;
                xyz[3 * i + 0] = zp * lcos - ORC_LBR_R;
                xyz[3 * i + 1] = zp * lsin;
                double rg = sqrt(xyz[3 * i] * xyz[3 * i] + xyz[3 * i + 1] * xyz[3 * i + 1]
                                 + (xyz[3 * i + 2] * xyz[3 * i + 2]) / (k->q * k->q));
                *bg_prob += qw_r3_N[i] / (pow(rg, k->alpha) * pow(rg + k->r0, k->alpha_delta3));
            }
        }
        *bg_prob *= reff_xr_rp3;
    }

    for (int s = 0; s < k->n_streams; s++) {
        st_prob[s] = 0;
        if (k->stream_sigma[s] > -0.0001 && k->stream_sigma[s] < 0.0001) continue;
        const double* c = k->stream_c[s];
        const double* a = k->stream_a[s];
        for (int j = 0; j < n; j++) {
            double v0 = xyz[3 * j + 0] - c[0];
            double v1 = xyz[3 * j + 1] - c[1];
            double v2 = xyz[3 * j + 2] - c[2];
            double dotted = a[0] * v0 + a[1] * v1 + a[2] * v2;
            v0 = v0 - dotted * a[0];
            v1 = v1 - dotted * a[1];
            v2 = v2 - dotted * a[2];
            double xyz_norm = v0 * v0 + v1 * v1 + v2 * v2;
            st_prob[s] += qw_r3_N[j] * exp(-xyz_norm / k->stream_sigma_sq2[s]);
        }
        st_prob[s] *= reff_xr_rp3;
    }
}

/* ------------------------------------------------------------------------ */
/* Tables: evaluation_optimized.c:360-422                                    */
/* ------------------------------------------------------------------------ */

static void orc_r_constants_k(const OrcConst* k, int r_steps, double r_min, double r_step_size,
                              double mu_step_size, int nu_steps, double nu_min,
                              double nu_step_size, double* irv, double* r_point,
                              double* r_in_mag, double* r_in_mag2, double* qw_r3_N,
                              double* reff_xr_rp3, double* nus, double* ids)
{
    int n = k->convolve;
    for (int i = 0; i < r_steps; i++) {
        double log_r = r_min + (i * r_step_size);
        double r = pow(10.0, (log_r - 14.2) / 5.0);
        double next_r = pow(10.0, (log_r + r_step_size - 14.2) / 5.0);
        /* "... * mu_step_size / deg" */
        irv[i] = (((next_r * next_r * next_r) - (r * r * r)) / 3.0) * mu_step_size / ORC_DEG;
        double rPrime = (next_r + r) / 2.0;
        orc_prob_constants(k, rPrime, r_point + (size_t)i * n, r_in_mag + (size_t)i * n,
                           r_in_mag2 + (size_t)i * n, qw_r3_N + (size_t)i * n, &reff_xr_rp3[i]);
    }
    for (int i = 0; i < nu_steps; i++) {
        nus[i] = nu_min + (i * nu_step_size);
        ids[i] = cos((90 - nus[i] - nu_step_size) / ORC_DEG) - cos((90 - nus[i]) / ORC_DEG);
        nus[i] += 0.5 * nu_step_size;
    }
}

void orc_r_constants(int n_convolve, const double* gl_x, const double* gl_w,
                     int r_steps, double r_min, double r_step_size,
                     double mu_step_size,
                     int nu_steps, double nu_min, double nu_step_size,
                     double* irv, double* r_point, double* r_in_mag, double* r_in_mag2,
                     double* qw_r3_N, double* reff_xr_rp3, double* nus, double* ids)
{
    OrcConst k;
    memset(&k, 0, sizeof(k));
    k.convolve = n_convolve;
    k.coeff = 1 / (ORC_STDEV * sqrt(2 * ORC_PI));
    k.gl_x = (double*)gl_x;
    k.gl_w = (double*)gl_w;
    k.dx = (double*)malloc(sizeof(double) * n_convolve);
    for (int i = 0; i < n_convolve; i++) k.dx[i] = 3 * ORC_STDEV * gl_x[i];
    orc_r_constants_k(&k, r_steps, r_min, r_step_size, mu_step_size, nu_steps, nu_min,
                      nu_step_size, irv, r_point, r_in_mag, r_in_mag2, qw_r3_N, reff_xr_rp3,
                      nus, ids);
    free(k.dx);
}

int orc_point_probabilities(const ASTRONOMY_PARAMETERS* ap, double l, double b, double r_kpc,
                            double* bg_prob, double* st_prob)
{
    OrcConst k;
    if (orc_init_constants(ap, &k)) return 1;
    int n = k.convolve;
    double* t = (double*)malloc(sizeof(double) * 4 * n);
    double reff;
    double pt[3] = { l, b, r_kpc };
    memset(k.xyz, 0, sizeof(double) * 3 * n);
    orc_prob_constants(&k, r_kpc, t, t + n, t + 2 * n, t + 3 * n, &reff);
    orc_probabilities(&k, t, t + n, t + 2 * n, t + 3 * n, reff, pt, bg_prob, st_prob);
    free(t);
    orc_free_constants(&k);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Region integral: evaluation_optimized.c:424-562; get_steps                */
/* evaluation_state.c:129-151                                                */
/* ------------------------------------------------------------------------ */

typedef struct {
    double *irv, *r_point, *r_in_mag, *r_in_mag2, *qw_r3_N, *reff_xr_rp3, *nus, *ids;
} OrcTables;

static int orc_region_valid(const INTEGRAL_AREA* ia)
{
    return ia->r_steps > 0 && ia->nu_steps > 0 && ia->mu_steps > 0;
}

static void orc_tables_alloc(const OrcConst* k, const INTEGRAL_AREA* ia, OrcTables* t)
{
    size_t rn = (size_t)ia->r_steps * k->convolve;
    t->irv = (double*)malloc(sizeof(double) * ia->r_steps);
    t->reff_xr_rp3 = (double*)malloc(sizeof(double) * ia->r_steps);
    t->r_point = (double*)malloc(sizeof(double) * rn);
    t->r_in_mag = (double*)malloc(sizeof(double) * rn);
    t->r_in_mag2 = (double*)malloc(sizeof(double) * rn);
    t->qw_r3_N = (double*)malloc(sizeof(double) * rn);
    t->nus = (double*)malloc(sizeof(double) * ia->nu_steps);
    t->ids = (double*)malloc(sizeof(double) * ia->nu_steps);
    orc_r_constants_k(k, ia->r_steps, ia->r_min, ia->r_step_size, ia->mu_step_size,
                      ia->nu_steps, ia->nu_min, ia->nu_step_size, t->irv, t->r_point,
                      t->r_in_mag, t->r_in_mag2, t->qw_r3_N, t->reff_xr_rp3, t->nus, t->ids);
}

static void orc_tables_free(OrcTables* t)
{
    free(t->irv); free(t->reff_xr_rp3); free(t->r_point); free(t->r_in_mag);
    free(t->r_in_mag2); free(t->qw_r3_N); free(t->nus); free(t->ids);
}

static void orc_integral_region(const ASTRONOMY_PARAMETERS* ap, OrcConst* k, INTEGRAL_AREA* ia)
{
    int S = k->n_streams, n = k->convolve;
    OrcTables t;
    orc_tables_alloc(k, ia, &t);

    int r_c = (int)(ia->current_calculation % ia->r_steps);
    int nu_c = (int)((ia->current_calculation / ia->r_steps) % ia->nu_steps);
    int mu_c = (int)(ia->current_calculation / (ia->r_steps * ia->nu_steps));

    double st_probs[ORC_MAX_STREAMS], st_int[ORC_MAX_STREAMS], st_int_c[ORC_MAX_STREAMS];
    double bg_prob, V, temp, ip[3];
    double bg_int = ia->background_integral;
    double bg_int_c = 0.0;
    for (int i = 0; i < S; i++) {
        st_int[i] = ia->stream_integrals[i];
        st_int_c[i] = 0.0;
    }
    int done = 0;
    for (; mu_c < ia->mu_steps; mu_c++) {
        double mu = ia->mu_min + (mu_c * ia->mu_step_size);
        for (; nu_c < ia->nu_steps; nu_c++) {
            orc_gc_to_lb(ap->wedge, ap->sgr_coordinates, mu + 0.5 * ia->mu_step_size,
                         t.nus[nu_c], &ip[0], &ip[1]);
            for (; r_c < ia->r_steps; r_c++) {
                size_t o = (size_t)r_c * n;
                V = t.irv[r_c] * t.ids[nu_c];
                orc_probabilities(k, t.r_point + o, t.r_in_mag + o, t.r_in_mag2 + o,
                                  t.qw_r3_N + o, t.reff_xr_rp3[r_c], ip, &bg_prob, st_probs);
                bg_prob *= V;
                temp = bg_int;
                bg_int += bg_prob;
                bg_int_c += bg_prob - (bg_int - temp);
                for (int i = 0; i < S; i++) {
                    st_probs[i] *= V;
                    temp = st_int[i];
                    st_int[i] += st_probs[i];
                    st_int_c[i] += st_probs[i] - (st_int[i] - temp);
                }
                ia->current_calculation++;
                if (ia->current_calculation >= ia->max_calculation) { done = 1; break; }
            }
            if (done) break;
            r_c = 0;
        }
        if (done) break;
        nu_c = 0;
    }
    ia->background_integral = bg_int + bg_int_c;
    for (int i = 0; i < S; i++) ia->stream_integrals[i] = st_int[i] + st_int_c[i];
    orc_tables_free(&t);
}

/* evaluation_optimized.c:564-611 */
int orc_calculate_integrals(const ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es,
                            const STAR_POINTS* sp)
{
    (void)sp;
    OrcConst k;
    if (orc_init_constants(ap, &k)) return 1;
    for (int i = 0; i < ap->number_integrals; i++) {
        if (!orc_region_valid(es->integral[i])) { orc_free_constants(&k); return 1; }
    }
    memset(k.xyz, 0, sizeof(double) * 3 * k.convolve);
    for (; es->current_integral < ap->number_integrals; es->current_integral++)
        orc_integral_region(ap, &k, es->integral[es->current_integral]);

    es->background_integral = es->integral[0]->background_integral;
    for (int i = 0; i < ap->number_streams; i++)
        es->stream_integrals[i] = es->integral[0]->stream_integrals[i];
    for (int i = 1; i < ap->number_integrals; i++) {
        es->background_integral -= es->integral[i]->background_integral;
        for (int j = 0; j < ap->number_streams; j++)
            es->stream_integrals[j] -= es->integral[i]->stream_integrals[j];
    }
    orc_free_constants(&k);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Star pass: evaluation_optimized.c:613-753                                 */
/* ------------------------------------------------------------------------ */

int orc_calculate_likelihood_ex(const ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es,
                                const STAR_POINTS* sp, double* log10_p,
                                double* bg_only_mean, double* st_only_mean)
{
    OrcConst k;
    if (orc_init_constants(ap, &k)) return 1;
    int S = k.n_streams, n = k.convolve;
    double exp_stream_weights[ORC_MAX_STREAMS], st_prob[ORC_MAX_STREAMS];
    double st_only_sum[ORC_MAX_STREAMS], st_only_sum_c[ORC_MAX_STREAMS];
    double* t = (double*)malloc(sizeof(double) * 4 * n);
    double *r_point = t, *r_in_mag = t + n, *r_in_mag2 = t + 2 * n, *qw_r3_N = t + 3 * n;
    double reff_xr_rp3, bg_prob, temp;
    memset(k.xyz, 0, sizeof(double) * 3 * n);

    double exp_background_weight = exp(ap->background_weight);
    double sum_exp_weights = exp_background_weight;
    for (int i = 0; i < S; i++) {
        exp_stream_weights[i] = exp(ap->stream_weights[i]);
        sum_exp_weights += exp(ap->stream_weights[i]);
    }
    sum_exp_weights *= 0.001;

    double prob_sum = 0, prob_sum_c = 0, bg_only_sum = 0, bg_only_sum_c = 0;
    for (int i = 0; i < S; i++) st_only_sum[i] = st_only_sum_c[i] = 0;

    for (; es->current_star_point < sp->number_stars; es->current_star_point++) {
        const double* star = sp->stars[es->current_star_point];
        orc_prob_constants(&k, star[2], r_point, r_in_mag, r_in_mag2, qw_r3_N, &reff_xr_rp3);
        orc_probabilities(&k, r_point, r_in_mag, r_in_mag2, qw_r3_N, reff_xr_rp3, star,
                          &bg_prob, st_prob);

        double bg_only = (bg_prob / es->background_integral) * exp_background_weight;
        double star_prob = bg_only;
        if (bg_only == 0.0)
            bg_only = -238;
        else
            bg_only = log10(bg_only / sum_exp_weights);
        temp = bg_only_sum;
        bg_only_sum += bg_only;
        bg_only_sum_c += bg_only - (bg_only_sum - temp);

        for (int s = 0; s < S; s++) {
            double st_only = st_prob[s] / es->stream_integrals[s] * exp_stream_weights[s];
            star_prob += st_only;
            if (st_only == 0.0)
                st_only = -238;
            else
                st_only = log10(st_only / sum_exp_weights);
            temp = st_only_sum[s];
            st_only_sum[s] += st_only;
            st_only_sum_c[s] += st_only - (st_only_sum[s] - temp);
        }
        star_prob /= sum_exp_weights;

        if (star_prob != 0.0) {
            star_prob = log10(star_prob);
            temp = prob_sum;
            prob_sum += star_prob;
            prob_sum_c += star_prob - (prob_sum - temp);
            if (log10_p) log10_p[es->current_star_point] = star_prob;
        } else {
            es->num_zero++;
            prob_sum -= 238.0;
            if (log10_p) log10_p[es->current_star_point] = -238.0;
        }
    }
    es->prob_sum = prob_sum + prob_sum_c;
    if (bg_only_mean) *bg_only_mean = (bg_only_sum + bg_only_sum_c) / sp->number_stars;
    if (st_only_mean)
        for (int s = 0; s < S; s++)
            st_only_mean[s] = (st_only_sum[s] + st_only_sum_c[s]) / sp->number_stars;
    free(t);
    orc_free_constants(&k);
    return 0;
}

int orc_calculate_likelihood(const ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es,
                             const STAR_POINTS* sp)
{
    return orc_calculate_likelihood_ex(ap, es, sp, NULL, NULL, NULL);
}

/* ------------------------------------------------------------------------ */
/* Parameter vector: parameters.c:307-363                                    */
/* ------------------------------------------------------------------------ */

int orc_get_optimized_parameter_count(const ASTRONOMY_PARAMETERS* ap)
{
    int count = 0;
    for (int i = 0; i < ap->number_background_parameters; i++)
        if (ap->background_optimize[i]) ++count;
    for (int i = 0; i < ap->number_streams; i++) {
        if (ap->stream_weight_optimize[i]) ++count;
        for (int j = 0; j < ap->number_stream_parameters; j++)
            if (ap->stream_optimize[i][j]) ++count;
    }
    return count;
}

void orc_set_astronomy_parameters(ASTRONOMY_PARAMETERS* ap, const double* parameters)
{
    int current = 0;
    for (int i = 0; i < ap->number_background_parameters; i++)
        if (ap->background_optimize[i]) ap->background_parameters[i] = parameters[current++];
    for (int i = 0; i < ap->number_streams; i++) {
        if (ap->stream_weight_optimize[i]) ap->stream_weights[i] = parameters[current++];
        for (int j = 0; j < ap->number_stream_parameters; j++)
            if (ap->stream_optimize[i][j]) ap->stream_parameters[i][j] = parameters[current++];
    }
}

/* ------------------------------------------------------------------------ */
/* Evaluation state: evaluation_state.c:97-211                               */
/* ------------------------------------------------------------------------ */

void orc_initialize_state(const ASTRONOMY_PARAMETERS* ap, const STAR_POINTS* sp, EVALUATION_STATE* es)
{
    es->current_integral = 0;
    es->background_integral = 0;
    es->stream_integrals = (double*)calloc(ap->number_streams > 0 ? ap->number_streams : 1, sizeof(double));
    es->number_streams = ap->number_streams;
    es->total_stars = sp->number_stars;
    es->current_star_point = 0;
    es->num_zero = 0;
    es->bad_jacobians = 0;
    es->prob_sum = 0;
    es->number_integrals = ap->number_integrals;
    es->integral = (INTEGRAL_AREA**)malloc(sizeof(INTEGRAL_AREA*) * ap->number_integrals);
    for (int i = 0; i < ap->number_integrals; i++) {
        INTEGRAL_AREA* ia = (INTEGRAL_AREA*)malloc(sizeof(INTEGRAL_AREA));
        const INTEGRAL* g = ap->integral[i];
        ia->mu_min = g->mu_min; ia->mu_max = g->mu_max; ia->mu_steps = g->mu_steps;
        ia->nu_min = g->nu_min; ia->nu_max = g->nu_max; ia->nu_steps = g->nu_steps;
        ia->r_min = g->r_min; ia->r_max = g->r_max; ia->r_steps = g->r_steps;
        ia->min_calculation = g->min_calculation;
        ia->max_calculation = g->max_calculation;
        ia->current_calculation = g->min_calculation;
        ia->mu_step_size = (ia->mu_max - ia->mu_min) / ia->mu_steps;
        ia->nu_step_size = (ia->nu_max - ia->nu_min) / ia->nu_steps;
        ia->r_step_size = (ia->r_max - ia->r_min) / ia->r_steps;
        ia->number_streams = ap->number_streams;
        ia->background_integral = 0;
        ia->stream_integrals = (double*)calloc(ap->number_streams > 0 ? ap->number_streams : 1, sizeof(double));
        es->integral[i] = ia;
    }
}

void orc_reset_evaluation_state(EVALUATION_STATE* es)
{
    es->current_integral = 0;
    es->background_integral = 0;
    for (int i = 0; i < es->number_streams; i++) es->stream_integrals[i] = 0;
    es->current_star_point = 0;
    es->num_zero = 0;
    es->bad_jacobians = 0;
    es->prob_sum = 0;
    for (int i = 0; i < es->number_integrals; i++) {
        es->integral[i]->background_integral = 0;
        for (int j = 0; j < es->integral[i]->number_streams; j++)
            es->integral[i]->stream_integrals[j] = 0;
        es->integral[i]->current_calculation = es->integral[i]->min_calculation;
    }
}

void orc_free_state(EVALUATION_STATE* es)
{
    free(es->stream_integrals);
    for (int i = 0; i < es->number_integrals; i++) {
        free(es->integral[i]->stream_integrals);
        free(es->integral[i]);
    }
    free(es->integral);
}

/* evaluation_optimized.c:756-784 */
double orc_evaluate(double* parameters, ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es,
                    STAR_POINTS* sp)
{
    orc_set_astronomy_parameters(ap, parameters);
    orc_reset_evaluation_state(es);
    if (orc_calculate_integrals(ap, es, sp)) return NAN;
    if (orc_calculate_likelihood(ap, es, sp)) return NAN;
    return es->prob_sum / (sp->number_stars - es->bad_jacobians);
}

/* ------------------------------------------------------------------------ */
/* Multi-core CPU baseline (not a reference function): the same arithmetic  */
/* with mu rows spread over threads.                                         */
/* ------------------------------------------------------------------------ */

static void orc_kahan_add(double* s, double* c, double v)
{
    double temp = *s;
    *s += v;
    *c += v - (*s - temp);
}

int orc_calculate_integrals_omp(const ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es, int n_threads)
{
    OrcConst k0;
    if (orc_init_constants(ap, &k0)) return 1;
    int S = k0.n_streams, n = k0.convolve;
    for (int g = 0; g < ap->number_integrals; g++) {
        INTEGRAL_AREA* ia = es->integral[g];
        if (!orc_region_valid(ia)) { orc_free_constants(&k0); return 1; }
        OrcTables t;
        orc_tables_alloc(&k0, ia, &t);
        int M = ia->mu_steps;
        double* row = (double*)calloc((size_t)M * (1 + S), sizeof(double));
        #pragma omp parallel num_threads(n_threads)
        {
            OrcConst k = k0;
            k.xyz = (double*)calloc((size_t)3 * n, sizeof(double));
            double st_probs[ORC_MAX_STREAMS], s_sum[ORC_MAX_STREAMS + 1], s_c[ORC_MAX_STREAMS + 1];
            #pragma omp for schedule(dynamic, 1)
            for (int mu_c = 0; mu_c < M; mu_c++) {
                double mu = ia->mu_min + (mu_c * ia->mu_step_size);
                for (int i = 0; i <= S; i++) s_sum[i] = s_c[i] = 0;
                for (int nu_c = 0; nu_c < ia->nu_steps; nu_c++) {
                    double ip[3], bg_prob;
                    orc_gc_to_lb(ap->wedge, ap->sgr_coordinates, mu + 0.5 * ia->mu_step_size,
                                 t.nus[nu_c], &ip[0], &ip[1]);
                    for (int r_c = 0; r_c < ia->r_steps; r_c++) {
                        size_t o = (size_t)r_c * n;
                        double V = t.irv[r_c] * t.ids[nu_c];
                        orc_probabilities(&k, t.r_point + o, t.r_in_mag + o, t.r_in_mag2 + o,
                                          t.qw_r3_N + o, t.reff_xr_rp3[r_c], ip, &bg_prob,
                                          st_probs);
                        orc_kahan_add(&s_sum[0], &s_c[0], bg_prob * V);
                        for (int i = 0; i < S; i++)
                            orc_kahan_add(&s_sum[1 + i], &s_c[1 + i], st_probs[i] * V);
                    }
                }
                for (int i = 0; i <= S; i++) row[(size_t)mu_c * (1 + S) + i] = s_sum[i] + s_c[i];
            }
            free(k.xyz);
        }
        double tot[ORC_MAX_STREAMS + 1] = { 0 }, tot_c[ORC_MAX_STREAMS + 1] = { 0 };
        for (int i = 0; i <= S; i++) tot[i] = tot_c[i] = 0;
        for (int mu_c = 0; mu_c < M; mu_c++)
            for (int i = 0; i <= S; i++) orc_kahan_add(&tot[i], &tot_c[i], row[(size_t)mu_c * (1 + S) + i]);
        ia->background_integral = tot[0] + tot_c[0];
        for (int i = 0; i < S; i++) ia->stream_integrals[i] = tot[1 + i] + tot_c[1 + i];
        ia->current_calculation = ia->max_calculation;
        free(row);
        orc_tables_free(&t);
    }
    es->current_integral = ap->number_integrals;
    es->background_integral = es->integral[0]->background_integral;
    for (int i = 0; i < S; i++) es->stream_integrals[i] = es->integral[0]->stream_integrals[i];
    for (int g = 1; g < ap->number_integrals; g++) {
        es->background_integral -= es->integral[g]->background_integral;
        for (int j = 0; j < S; j++) es->stream_integrals[j] -= es->integral[g]->stream_integrals[j];
    }
    orc_free_constants(&k0);
    return 0;
}

int orc_calculate_likelihood_omp(const ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es,
                                 const STAR_POINTS* sp, int n_threads)
{
    OrcConst k0;
    if (orc_init_constants(ap, &k0)) return 1;
    int S = k0.n_streams, n = k0.convolve, N = sp->number_stars;
    double exp_w[ORC_MAX_STREAMS];
    double exp_bg = exp(ap->background_weight);
    double sum_w = exp_bg;
    for (int i = 0; i < S; i++) {
        exp_w[i] = exp(ap->stream_weights[i]);
        sum_w += exp(ap->stream_weights[i]);
    }
    sum_w *= 0.001;
    int chunks = n_threads > 0 ? n_threads * 8 : 8;
    double* part = (double*)calloc((size_t)chunks * 2, sizeof(double));
    int* zeros = (int*)calloc((size_t)chunks, sizeof(int));
    #pragma omp parallel num_threads(n_threads)
    {
        OrcConst k = k0;
        k.xyz = (double*)calloc((size_t)3 * n, sizeof(double));
        double* t = (double*)malloc(sizeof(double) * 4 * n);
        double st_prob[ORC_MAX_STREAMS], reff, bg_prob;
        #pragma omp for schedule(dynamic, 1)
        for (int c = 0; c < chunks; c++) {
            int lo = (int)((long)N * c / chunks), hi = (int)((long)N * (c + 1) / chunks);
            double s = 0, sc = 0;
            for (int i = lo; i < hi; i++) {
                const double* star = sp->stars[i];
                orc_prob_constants(&k, star[2], t, t + n, t + 2 * n, t + 3 * n, &reff);
                orc_probabilities(&k, t, t + n, t + 2 * n, t + 3 * n, reff, star, &bg_prob, st_prob);
                double p = (bg_prob / es->background_integral) * exp_bg;
                for (int j = 0; j < S; j++) p += st_prob[j] / es->stream_integrals[j] * exp_w[j];
                p /= sum_w;
                if (p != 0.0) orc_kahan_add(&s, &sc, log10(p));
                else zeros[c]++;
            }
            part[2 * c] = s;
            part[2 * c + 1] = sc;
        }
        free(t);
        free(k.xyz);
    }
    double s = 0, sc = 0;
    int nz = 0;
    for (int c = 0; c < chunks; c++) {
        orc_kahan_add(&s, &sc, part[2 * c] + part[2 * c + 1]);
        nz += zeros[c];
    }
    es->prob_sum = s + sc - 238.0 * nz;
    es->num_zero += nz;
    es->current_star_point = N;
    free(part);
    free(zeros);
    orc_free_constants(&k0);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Stream separation: separation.c:37-254 with probability.c:310-403,       */
/* stCoords.c:52-65, :142-201, stVector.c:60-169, atSurveyGeometry.c:414-460 */
/* ------------------------------------------------------------------------ */

/* atSurveyGeometry.c:444-460 */
static double orc_eta_from_stripe(int wedge)
{
    double stripe_separation = 2.5;
    if (wedge <= 46) return wedge * stripe_separation - 57.5;
    return wedge * stripe_separation - 57.5 - 180.0;
}

/* atSurveyGeometry.c:414-440; at_deg2Rad / at_rad2Deg expand to
 * x * DPI / 180.0 and x * 180.0 / DPI. */
static void orc_survey_to_eq(double slong, double slat, double* ra, double* dec)
{
    double anode, eta_pole, x1, y1, z1;
    double center_ra = 185.0, center_dec = 32.5;
    slong = slong * ORC_PI / 180.0;
    slat = slat * ORC_PI / 180.0;
    anode = center_ra - 90.0;
    anode = anode * ORC_PI / 180.0;
    eta_pole = center_dec * ORC_PI / 180.0;
    x1 = -sin(slong);
    y1 = cos(slat + eta_pole) * cos(slong);
    z1 = sin(slat + eta_pole) * cos(slong);
    *ra = atan2(y1, x1) + anode;
    *dec = asin(z1);
    *ra = *ra * 180.0 / ORC_PI;
    *dec = *dec * 180.0 / ORC_PI;
    orc_bound2(dec, ra);
}

/* stCoords.c:177-185 */
static void orc_lb_to_xyz(double l, double b, double* xyz)
{
    l = l / ORC_DEG;
    b = b / ORC_DEG;
    xyz[0] = cos(l) * cos(b);
    xyz[1] = sin(l) * cos(b);
    xyz[2] = sin(b);
}

/* stCoords.c:142-150 */
void orc_stripe_normal(int wedge, double* xyz)
{
    double eta, ra, dec, l, b;
    eta = orc_eta_from_stripe(wedge);
    orc_survey_to_eq(0, 90.0 + eta, &ra, &dec);
    orc_eq_to_gal(ra, dec, &l, &b);
    orc_lb_to_xyz(l, b, xyz);
}

/* stCoords.c:153-174 */
void orc_sgr_stripe_normal(int wedge, double* xyz)
{
    double l1, b1, l2, b2, xyz1[3], xyz2[3];
    orc_sgr_to_gal(wedge * 2.5, 0, &l1, &b1);
    orc_sgr_to_gal(wedge * 2.5, 90, &l2, &b2);
    orc_lb_to_xyz(l1, b1, xyz1);
    orc_lb_to_xyz(l2, b2, xyz2);
    xyz[0] = xyz1[1] * xyz2[2] - xyz1[2] * xyz2[1];
    xyz[1] = xyz1[2] * xyz2[0] - xyz1[0] * xyz2[2];
    xyz[2] = xyz1[0] * xyz2[1] - xyz1[1] * xyz2[0];
}

/* stVector.c:60-105: norm, dotp, crossp */
static double orc_norm(const double* v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
static double orc_dotp(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* stVector.c:109-140; mat is row-major [3][3] */
void orc_get_transform(const double* f, const double* t, double* mat)
{
    double axis[3], angle, sin_a, x, y, z, w, n;
    axis[0] = f[1] * t[2] - f[2] * t[1];
    axis[1] = f[2] * t[0] - f[0] * t[2];
    axis[2] = f[0] * t[1] - f[1] * t[0];
    n = orc_norm(axis);
    axis[0] /= n;
    axis[1] /= n;
    axis[2] /= n;
    angle = acos(orc_dotp(f, t) / (orc_norm(f) * orc_norm(t)));
    sin_a = sin(angle / 2);
    x = axis[0] * sin_a;
    y = axis[1] * sin_a;
    z = axis[2] * sin_a;
    w = cos(angle / 2);
    mat[0] = 1 - 2 * (y * y + z * z);
    mat[1] = 2 * (x * y - z * w);
    mat[2] = 2 * (x * z + y * w);
    mat[3] = 2 * (x * y + z * w);
    mat[4] = 1 - 2 * (x * x + z * z);
    mat[5] = 2 * (y * z - x * w);
    mat[6] = 2 * (x * z - y * w);
    mat[7] = 2 * (y * z + x * w);
    mat[8] = 1 - 2 * (x * x + y * y);
}

/* stCoords.c:52-65 */
static void orc_xyz2lbr(const double* xyz, double* lbr)
{
    double xsun = xyz[0] + orc_sun_r0;
    double temp = xsun * xsun + xyz[1] * xyz[1];
    lbr[0] = atan2(xyz[1], xsun) * ORC_DEG;
    lbr[1] = atan2(xyz[2], sqrt(temp)) * ORC_DEG;
    lbr[2] = sqrt(temp + xyz[2] * xyz[2]);
    if (lbr[0] < 0) lbr[0] += 360;
}

/* stVector.c:155-169 with xyz_mag / xyz2lbg (stCoords.c:188-201) and
 * do_transform (stVector.c:143-153) */
void orc_transform_point(const double* point, const double* cmat, const double* xsun, double* out)
{
    double lbg[3], p[3];
    orc_xyz2lbr(point, lbg);
    lbg[2] = 5.0 * (log(100.0 * lbg[2]) / log(10.0)) + 4.2 - 11.0;
    orc_lbr2xyz(lbg, out);
    p[0] = out[0] - xsun[0];
    p[1] = out[1] - xsun[1];
    p[2] = out[2] - xsun[2];
    out[0] = orc_dotp(cmat, p);
    out[1] = orc_dotp(cmat + 3, p);
    out[2] = orc_dotp(cmat + 6, p);
}

/* probability.c:310-396 with the random number passed in.  Returns -1 where
 * the reference leaves its result unset (n = 2 or 3 with r == p[0] or NaN
 * probabilities, n = 4 with NaN), and -2 for n outside 1-4 (the reference
 * prints an error and exits). */
int orc_prob_ok(int n, const double* p, double r)
{
    double step1, step2, step3;
    int ok = -1;
    switch (n) {
    case 1:
        ok = (r > p[0]) ? 0 : 1;
        break;
    case 2:
        step1 = p[0] + p[1];
        if (r > step1) ok = 0;
        else if (r < p[0]) ok = 1;
        else if ((r > p[0]) && (r <= step1)) ok = 2;
        break;
    case 3:
        step1 = p[0] + p[1];
        step2 = p[0] + p[1] + p[2];
        if (r > step2) ok = 0;
        else if (r < p[0]) ok = 1;
        else if ((r > p[0]) && (r <= step1)) ok = 2;
        else if ((r > step1) && (r <= step2)) ok = 3;
        break;
    case 4:
        step1 = p[0] + p[1];
        step2 = p[0] + p[1] + p[2];
        step3 = p[0] + p[1] + p[2] + p[3];
        if (r > step3) ok = 0;
        else if (r <= p[0]) ok = 1;
        else if ((r > p[0]) && (r <= step1)) ok = 2;
        else if ((r > step1) && (r <= step2)) ok = 3;
        else if ((r > step2) && (r <= step3)) ok = 4;
        break;
    default:
        ok = -2;
    }
    return ok;
}

/* separation.c:37-254 (twoPanel == 1).  The star loop's probabilities come
 * from the same set_probability_constants / calculate_probabilities pair as
 * the likelihood pass (the call at separation.c:162-164 predates their
 * current signatures).  The random numbers are glibc's drand48 sequence after
 * srand48(seed) (erand48 on a state seeded the same way).  Outputs:
 * assign[N], xyz[N][3], sprob[N][S], nstars[S], q[S]; sprob may be NULL. */
int orc_separation(const ASTRONOMY_PARAMETERS* ap, const STAR_POINTS* sp, double background_integral,
                   const double* stream_integrals, unsigned long seed, int* assign, double* xyz,
                   double* sprob, double* nstars, int* q)
{
    OrcConst k;
    int S = ap->number_streams, i, j;
    double dnormal[3], dortho[3] = { 0.0, 0.0, 1.0 }, xsun[3] = { -8.5, 0.0, 0.0 }, cmatrix[9];
    double epsilon_s[ORC_MAX_STREAMS], epsilon_b, denom, prob_b;
    double prob_s[ORC_MAX_STREAMS], psg[ORC_MAX_STREAMS], p[ORC_MAX_STREAMS];
    unsigned short xsubi[3];
    if (S < 1 || S > 4) return 1;
    if (ap->sgr_coordinates == 0)
        orc_stripe_normal(ap->wedge, dnormal);
    else if (ap->sgr_coordinates == 1)
        orc_sgr_stripe_normal(ap->wedge, dnormal);
    else
        return 1;
    orc_get_transform(dnormal, dortho, cmatrix);

    denom = 1.0;
    for (j = 0; j < S; j++) denom += exp(ap->stream_weights[j]);
    for (j = 0; j < S; j++) epsilon_s[j] = exp(ap->stream_weights[j]) / denom;
    epsilon_b = 1.0 / denom;

    if (orc_init_constants(ap, &k)) return 1;
    int n = k.convolve;
    double* t = (double*)malloc(sizeof(double) * 4 * n);
    double reff;
    memset(k.xyz, 0, sizeof(double) * 3 * n);

    /* srand48(seed): x = (seed & 0xffffffff) << 16 | 0x330e */
    xsubi[0] = 0x330e;
    xsubi[1] = (unsigned short)(seed & 0xffff);
    xsubi[2] = (unsigned short)((seed >> 16) & 0xffff);

    for (j = 0; j < S; j++) {
        nstars[j] = 0;
        q[j] = 0;
    }
    for (i = 0; i < sp->number_stars; i++) {
        const double* star = sp->stars[i];
        double starxyz[3], psg_sum = 0;
        orc_prob_constants(&k, star[2], t, t + n, t + 2 * n, t + 3 * n, &reff);
        orc_probabilities(&k, t, t + n, t + 2 * n, t + 3 * n, reff, star, &prob_b, prob_s);
        double pbx = epsilon_b * prob_b / background_integral;
        for (j = 0; j < S; j++) psg[j] = epsilon_s[j] * prob_s[j] / stream_integrals[j];
        for (j = 0; j < S; j++) psg_sum += psg[j];
        for (j = 0; j < S; j++) p[j] = psg[j] / (psg_sum + pbx);
        for (j = 0; j < S; j++) nstars[j] += p[j];
        if (sprob)
            for (j = 0; j < S; j++) sprob[(size_t)i * S + j] = p[j];
        assign[i] = orc_prob_ok(S, p, erand48(xsubi));
        if (assign[i] >= 1) q[assign[i] - 1]++;
        orc_lbr2xyz(star, starxyz);
        orc_transform_point(starxyz, cmatrix, xsun, xyz + 3 * (size_t)i);
    }
    free(t);
    orc_free_constants(&k);
    return 0;
}
