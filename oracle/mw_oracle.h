/*
 * mw_oracle.h -- CPU FP64 restatement of the reference separation likelihood.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP library is
 * compared against.  Only the tests, the smoke check and the CPU-baseline leg
 * of the benchmark may load it; the product never links or calls it.
 *
 * Every function restates a reference function (file:line given at each
 * definition in mw_oracle.c) with the same formulas, libm calls, operation
 * order, loop order and Kahan summation.  Unlike the reference it keeps no
 * file-scope state, so it is reentrant.
 */
#ifndef MW_ORACLE_H
#define MW_ORACLE_H

#include "mw_separation_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Gauss-Legendre nodes and weights on [x1, x2]. */
void orc_gauss_legendre(double x1, double x2, double* x, double* w, int n);

/* Great-circle survey coordinates (mu, nu, degrees) of a wedge to galactic
 * (l, b) in degrees, through the SDSS chain (sgr == 0) or the Sagittarius
 * chain (sgr == 1). */
void orc_gc_to_lb(int wedge, int sgr, double mu, double nu, double* l, double* b);

/* Individual steps of the chains, exposed for the invariant tests. */
void orc_gc_to_eq(double amu, double anu, double* ra, double* dec, double anode, double ainc);
void orc_eq_to_gal(double ra, double dec, double* glong, double* glat);
void orc_gc_to_sgr(double mu, double nu, int wedge, double* lamda, double* beta);
void orc_sgr_to_gal(double lamda, double beta, double* l, double* b);
double orc_wedge_incl(int wedge);
double orc_get_node(void);

/* Sun-centred (l deg, b deg, r kpc) to galactocentric xyz (kpc). */
void orc_lbr2xyz(const double* lbr, double* xyz);

/* Per-region tables.  r_point, r_in_mag, r_in_mag2 and qw_r3_N are
 * [r_steps][n_convolve] row-major; irv and reff_xr_rp3 are [r_steps];
 * nus and ids are [nu_steps]. */
void orc_r_constants(int n_convolve, const double* gl_x, const double* gl_w,
                     int r_steps, double r_min, double r_step_size,
                     double mu_step_size,
                     int nu_steps, double nu_min, double nu_step_size,
                     double* irv, double* r_point, double* r_in_mag, double* r_in_mag2,
                     double* qw_r3_N, double* reff_xr_rp3, double* nus, double* ids);

/* Integrand at one sky point (l, b degrees) and one distance (kpc): fills
 * bg_prob and st_prob[number_streams] exactly as the reference integrand
 * does for a star at (l, b, r_kpc). */
int orc_point_probabilities(const ASTRONOMY_PARAMETERS* ap, double l, double b, double r_kpc,
                            double* bg_prob, double* st_prob);

/* Reference evaluator steps.  Return 0 on success, nonzero on bad input. */
int orc_calculate_integrals(const ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es,
                            const STAR_POINTS* sp);
int orc_calculate_likelihood(const ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es,
                             const STAR_POINTS* sp);

/* Star pass that also reports per-star log10 P (log10_p[number_stars], -238
 * where P == 0), the background-only and stream-only mean log10 sums
 * (bg_only_mean, st_only_mean[number_streams]; may be NULL). */
int orc_calculate_likelihood_ex(const ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es,
                                const STAR_POINTS* sp, double* log10_p,
                                double* bg_only_mean, double* st_only_mean);

/* Full evaluation: maps the parameter vector into ap, resets es, computes the
 * integrals and the likelihood, returns prob_sum / number_stars (without the
 * client's -3.0). */
double orc_evaluate(double* parameters, ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es,
                    STAR_POINTS* sp);

/* Parameter-vector mapping (optimize masks). */
int orc_get_optimized_parameter_count(const ASTRONOMY_PARAMETERS* ap);
void orc_set_astronomy_parameters(ASTRONOMY_PARAMETERS* ap, const double* parameters);

/* State lifecycle (allocates with malloc). */
void orc_initialize_state(const ASTRONOMY_PARAMETERS* ap, const STAR_POINTS* sp, EVALUATION_STATE* es);
void orc_reset_evaluation_state(EVALUATION_STATE* es);
void orc_free_state(EVALUATION_STATE* es);

/* Multi-core CPU baseline: the same integral with mu rows spread over
 * n_threads OpenMP threads, a Kahan sum per row and a fixed-order combine.
 * Ignores min/max_calculation (always the whole region). */
int orc_calculate_integrals_omp(const ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es, int n_threads);
int orc_calculate_likelihood_omp(const ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es,
                                 const STAR_POINTS* sp, int n_threads);

/* Stream separation (separation.c:37-254) and its geometry helpers.
 * mat / cmat are row-major [3][3]. */
void orc_stripe_normal(int wedge, double* xyz);
void orc_sgr_stripe_normal(int wedge, double* xyz);
void orc_get_transform(const double* f, const double* t, double* mat);
void orc_transform_point(const double* point, const double* cmat, const double* xsun, double* out);
int orc_prob_ok(int n, const double* p, double r);
int orc_separation(const ASTRONOMY_PARAMETERS* ap, const STAR_POINTS* sp, double background_integral,
                   const double* stream_integrals, unsigned long seed, int* assign, double* xyz,
                   double* sprob, double* nstars, int* q);

#ifdef __cplusplus
}
#endif

#endif /* MW_ORACLE_H */
