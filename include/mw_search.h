/*
 * mw_search.h -- the search-side glue around the evaluator, exported by
 * libmw_separation_hip.so: the partial-evaluation plugin functions a
 * distributed search registers per rank, their compose steps and range
 * splitters, the parameter-vector accessors, and finite-difference
 * Hessians with the parameter errors derived from them.
 *
 * The reference's search driver and its Hessian live in an external search
 * library that is not in the reference tree; the functions here follow the
 * reference files that use them (cited at each declaration).
 */
#ifndef MW_SEARCH_H
#define MW_SEARCH_H

#include <stddef.h>

#include "mw_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- partial-evaluation plugins (GPU) ------------------------------------ */

/* Replaces integral_f (milkyway_separation/cpu/astronomy_worker.c:35-66):
 * maps the parameter vector into ap, resets es, computes the integrals of
 * this rank's calculation ranges, and writes results[0] = background,
 * results[1 + s] = stream s.  Exits with the error code on failure, as the
 * reference does. */
void mw_hip_integral_f(double* parameters, double* results, ASTRONOMY_PARAMETERS* ap,
                       EVALUATION_STATE* es, STAR_POINTS* sp);

/* Replaces likelihood_f (astronomy_worker.c:88-113): takes the composed
 * integrals, runs the star pass over this rank's stars from
 * es->current_star_point, and writes results[0] = prob_sum,
 * results[1] = bad_jacobians. */
void mw_hip_likelihood_f(double* integrals, double* results, ASTRONOMY_PARAMETERS* ap,
                         EVALUATION_STATE* es, STAR_POINTS* sp);

/* ---- compose and split (host only) -------------------------------------- */

/* integral_compose (astronomy_worker.c:68-86): sums num_results rank
 * results of 1 + number_streams values each. */
void mw_integral_compose(const double* integral_results, int num_results, double* results,
                         const ASTRONOMY_PARAMETERS* ap);

/* likelihood_compose (astronomy_worker.c:115-134): sum of the ranks'
 * prob_sum over (total_number_stars - sum of bad_jacobians), minus 3.0.
 * The reference reads total_number_stars from a global. */
double mw_likelihood_compose(const double* results, size_t num_results, int total_number_stars);

/* split_astronomy_parameters (parameters.c:502-522): rank's share
 * [total * rank / max_rank, total * (rank + 1) / max_rank) of every region's
 * flat calculation range, written into ap's regions. */
void mw_split_astronomy_parameters(ASTRONOMY_PARAMETERS* ap, int rank, int max_rank);

/* split_star_points (star_points.c:106-130): keeps the rank's contiguous
 * share of the stars (sp's rows are replaced; free with mw_free_star_points).
 * Returns 0, or 1 when an allocation fails. */
int mw_split_star_points(STAR_POINTS* sp, int rank, int max_rank);

/* get_search_parameters / get_step / get_min_parameters / get_max_parameters
 * (parameters.c:365-500): the optimised values, steps and bounds in vector
 * order.  Each writes mw_get_optimized_parameter_count(ap) values. */
void mw_get_search_parameters(const ASTRONOMY_PARAMETERS* ap, double* out);
void mw_get_step(const ASTRONOMY_PARAMETERS* ap, double* out);
void mw_get_min_parameters(const ASTRONOMY_PARAMETERS* ap, double* out);
void mw_get_max_parameters(const ASTRONOMY_PARAMETERS* ap, double* out);

/* ---- Hessian and parameter errors --------------------------------------- */

/* Evaluates count parameter vectors (count x n doubles, row-major) into
 * values[count]; returns 0 on success. */
typedef int (*mw_vector_eval_fn)(void* user, const double* vectors, int count, double* values);

/* Central-difference Hessian, gradient and value of f at point with steps
 * step[n], from 2 n^2 + 1 evaluations issued in one call of f:
 *   H_ii = (f(x + h_i) - 2 f(x) + f(x - h_i)) / h_i^2
 *   H_ij = (f(x + h_i + h_j) - f(x + h_i - h_j) - f(x - h_i + h_j) + f(x - h_i - h_j)) / (4 h_i h_j)
 *   g_i  = (f(x + h_i) - f(x - h_i)) / (2 h_i)
 * hessian is n x n row-major; gradient and value may be NULL.  The search
 * library's get_hessian, which errors.c:38-84 calls, is not in the
 * reference tree; this is the standard central-difference stencil. */
int mw_hessian(int n, const double* point, const double* step, mw_vector_eval_fn f, void* user,
               double* hessian, double* gradient, double* value);

/* mw_hessian of the client's likelihood (mean log10 P - 3.0) over the
 * context's wedge, every stencil point evaluated in batched launches. */
int mw_ctx_hessian(mw_ctx* ctx, const double* point, const double* step, int precision,
                   double* hessian, double* gradient, double* value);

/* Inverse of the n x n row-major matrix a by LUP decomposition with partial
 * pivoting (the method of util/matrix.c:234-323).  Returns 0, or 1 if a is
 * singular. */
int mw_matrix_invert(int n, const double* a, double* inverse);

/* errors (milkyway_separation/cpu/errors.c:38-84): sigma = H^-1 / number_stars
 * and stddev_i = sqrt(|sigma_ii|).  Returns 0, or 1 if H is singular. */
int mw_errors(int n, const double* hessian, int number_stars, double* sigma, double* stddev);

#ifdef __cplusplus
}
#endif

#endif /* MW_SEARCH_H */
