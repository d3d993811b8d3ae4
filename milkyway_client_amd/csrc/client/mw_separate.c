/*
 * mw_separate.c -- stream separation run on the MI355X evaluator.
 *
 * Does what the reference's separation program does (its main,
 * milkyway_separation/cpu/separation.c:256-322, which needs MPI and the
 * external search library and is not built by the reference's CMake): read
 * the parameter and star files, evaluate the parameter file's own values
 * (get_search_parameters), then separate the stars with the integrals of that
 * evaluation and write "s x y z" per star.
 *
 *   separation_hip [-parameters FILE] [-stars FILE] -outfile FILE [-seed N]
 *
 * The reference's single-dash long options; defaults parameters.txt and
 * stars.txt as there.  The reference seeds drand48 with time(NULL); -seed
 * fixes the seed so that a run can be repeated.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mw_hip.h"
#include "mw_io.h"

int main(int argc, char** argv)
{
    const char* parameters_file = "parameters.txt";
    const char* stars_file = "stars.txt";
    const char* outfile = NULL;
    unsigned long seed = (unsigned long)time(NULL);
    ASTRONOMY_PARAMETERS ap;
    STAR_POINTS sp;
    EVALUATION_STATE es;
    double* point;
    double likelihood;
    int i, j, n, k, rc;

    for (i = 1; i < argc; i++) {
        const int has_value = i + 1 < argc;
        if (!strcmp(argv[i], "-stars") && has_value)
            stars_file = argv[++i];
        else if (!strcmp(argv[i], "-parameters") && has_value)
            parameters_file = argv[++i];
        else if (!strcmp(argv[i], "-outfile") && has_value)
            outfile = argv[++i];
        else if (!strcmp(argv[i], "-seed") && has_value)
            seed = strtoul(argv[++i], NULL, 10);
        else {
            fprintf(stderr, "usage: %s [-parameters FILE] [-stars FILE] -outfile FILE [-seed N]\n", argv[0]);
            return EXIT_FAILURE;
        }
    }
    if (!outfile) {
        fprintf(stderr, "%s: -outfile is required\n", argv[0]);
        return EXIT_FAILURE;
    }
    if (mw_read_astronomy_parameters(parameters_file, &ap)) return EXIT_FAILURE;
    if (mw_read_star_points(stars_file, &sp)) {
        mw_free_astronomy_parameters(&ap);
        return EXIT_FAILURE;
    }
    if (mw_initialize_state(&ap, &sp, &es)) return EXIT_FAILURE;
    printf("read %d stars.\n", sp.number_stars);

    /* get_search_parameters (parameters.c:365-394): the optimised values. */
    n = mw_get_optimized_parameter_count(&ap);
    point = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!point) return EXIT_FAILURE;
    k = 0;
    for (i = 0; i < ap.number_background_parameters; i++)
        if (ap.background_optimize[i]) point[k++] = ap.background_parameters[i];
    for (i = 0; i < ap.number_streams; i++) {
        if (ap.stream_weight_optimize[i]) point[k++] = ap.stream_weights[i];
        for (j = 0; j < ap.number_stream_parameters; j++)
            if (ap.stream_optimize[i][j]) point[k++] = ap.stream_parameters[i][j];
    }

    printf("evaluating point...\n");
    likelihood = mw_hip_evaluate(point, &ap, &es, &sp) - 3.0;
    printf("likelihood: %.20f\n", likelihood);
    printf("performing separation...\n");
    rc = mw_hip_separation(outfile, es.background_integral, es.stream_integrals, &ap, &sp, seed);
    if (rc) fprintf(stderr, "APP: separation failed: %d (%s)\n", rc, mw_last_error());

    free(point);
    mw_free_state(&es);
    mw_free_star_points(&sp);
    mw_free_astronomy_parameters(&ap);
    return rc ? rc : EXIT_SUCCESS;
}
