/*
 * mw_separation_types.h -- the separation evaluator's data contract.
 *
 * These are the structs that the reference client passes through its
 * evaluator function pointer.  Their layout must stay byte-identical to the
 * reference's so that a caller built against the reference headers can hand
 * its own objects to this library:
 *
 *   INTEGRAL, ASTRONOMY_PARAMETERS   milkyway_separation/cpu/parameters.h:27-71
 *   INTEGRAL_AREA, EVALUATION_STATE  milkyway_separation/cpu/evaluation_state.h:30-58
 *   STAR_POINTS                      milkyway_separation/cpu/star_points.h:27-31
 *   EvalFunc                         milkyway_separation/cpu/simple_evaluator.h:26
 *
 * A caller that already includes the reference's own headers defines
 * MW_HAVE_REFERENCE_HEADERS before including this file; the reference's
 * definitions are then used and nothing is defined twice.
 */
#ifndef MW_SEPARATION_TYPES_H
#define MW_SEPARATION_TYPES_H

#ifdef MW_HAVE_REFERENCE_HEADERS

#include "parameters.h"
#include "star_points.h"
#include "evaluation_state.h"
#include "simple_evaluator.h"

#else

#ifdef __cplusplus
extern "C" {
#endif

/* One integration region (the main wedge or a cut).  r is in g magnitudes,
 * mu and nu in degrees; step sizes are (max - min) / steps.  The flat
 * calculation index runs mu-major, then nu, then r. */
typedef struct
{
    double r_min, r_max, r_step_size;
    double nu_min, nu_max, nu_step_size;
    double mu_min, mu_max, mu_step_size;
    int r_steps, nu_steps, mu_steps;

    long min_calculation, max_calculation;
} INTEGRAL;

typedef struct
{
    double parameters_version;

    int number_background_parameters;
    double background_weight;
    double* background_parameters;   /* alpha, q, r0, delta [, a, b, c] */
    double* background_step;
    double* background_min;
    double* background_max;
    int* background_optimize;

    int number_streams;
    int number_stream_parameters;    /* 5: mu (deg), r (kpc), theta, phi (rad), sigma (kpc) */

    double* stream_weights;
    double* stream_weight_step;
    double* stream_weight_min;
    double* stream_weight_max;
    int* stream_weight_optimize;

    double** stream_parameters;
    double** stream_step;
    double** stream_min;
    double** stream_max;
    int** stream_optimize;

    int convolve;
    int sgr_coordinates;
    int aux_bg_profile;
    int wedge;

    int number_integrals;            /* 1 + number of cut regions */
    INTEGRAL** integral;
} ASTRONOMY_PARAMETERS;

/* Stars as rows of (l deg, b deg, r kpc). */
typedef struct star_points
{
    int number_stars;
    double** stars;
} STAR_POINTS;

typedef struct integral_area
{
    int mu_steps, nu_steps, r_steps;
    double mu_min, nu_min, r_min;
    double mu_max, nu_max, r_max;
    double mu_step_size, nu_step_size, r_step_size;

    long min_calculation, max_calculation, current_calculation;

    int number_streams;
    double background_integral, *stream_integrals;
} INTEGRAL_AREA;

typedef struct evaluation_state
{
    INTEGRAL_AREA** integral;
    int current_integral, number_streams, number_integrals;

    double background_integral;
    double* stream_integrals;

    int current_star_point, total_stars;
    int num_zero;
    int bad_jacobians;
    double prob_sum;
} EVALUATION_STATE;

typedef double (*EvalFunc)(double*, ASTRONOMY_PARAMETERS*, EVALUATION_STATE*, STAR_POINTS*);

#ifdef __cplusplus
}
#endif

#endif /* MW_HAVE_REFERENCE_HEADERS */

#endif /* MW_SEPARATION_TYPES_H */
