/*
 * mw_errors.c -- parameter errors from the likelihood's Hessian on the MI355X
 * evaluator.
 *
 * Does what the reference's errors program does (milkyway_separation/cpu/
 * errors.c:86-139, which needs MPI and the external search library): at the
 * parameter file's values, with the file's steps, it computes the Hessian
 * of the likelihood (mean log10 P - 3.0), inverts it, divides by the number
 * of stars and reports sqrt(|sigma_ii|).  All 2 n^2 + 1 stencil points are
 * evaluated in batched GPU launches.
 *
 *   errors_hip OUTFILE [-parameters FILE] [-stars FILE]
 *
 * The output file and stdout carry the report in the layout of errors.c:38-84
 * (the Hessian as util/matrix.c:75-85 writes matrices).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mw_hip.h"
#include "mw_io.h"
#include "mw_search.h"

static void report(FILE* out, int n, const double* h, const double* sigma, const double* sd)
{
    int i, j;
    fprintf(out, "hessian [%d x %d]:\n", n, n);
    for (i = 0; i < n; i++) {
        for (j = 0; j < n; j++) fprintf(out, " %.20f", h[(size_t)i * n + j]);
        fprintf(out, "\n");
    }
    fprintf(out, "\nSIGMA:\n");
    for (i = 0; i < n; i++) {
        for (j = 0; j < n; j++) fprintf(out, " %f", sigma[(size_t)i * n + j]);
        fprintf(out, "\n");
    }
    fprintf(out, "\nErrors: q, r0, epsilon, mu, r, theta, phi, sigma\n");
    for (i = 0; i < n; i++) fprintf(out, " %f", sd[i]);
    fprintf(out, "\n");
}

int main(int argc, char** argv)
{
    const char* parameters_file = "parameters.txt";
    const char* stars_file = "stars.txt";
    const char* outfile = NULL;
    ASTRONOMY_PARAMETERS ap;
    STAR_POINTS sp;
    mw_ctx* ctx = NULL;
    double *point, *step, *h, *sigma, *sd, value;
    FILE* f;
    int i, n, rc;

    for (i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "-stars") && i + 1 < argc)
            stars_file = argv[++i];
        else if (!strcmp(argv[i], "-parameters") && i + 1 < argc)
            parameters_file = argv[++i];
        else if (!outfile && argv[i][0] != '-')
            outfile = argv[i];
        else {
            fprintf(stderr, "usage: %s OUTFILE [-parameters FILE] [-stars FILE]\n", argv[0]);
            return EXIT_FAILURE;
        }
    }
    if (!outfile) {
        fprintf(stderr, "usage: %s OUTFILE [-parameters FILE] [-stars FILE]\n", argv[0]);
        return EXIT_FAILURE;
    }
    if (mw_read_astronomy_parameters(parameters_file, &ap)) return EXIT_FAILURE;
    if (mw_read_star_points(stars_file, &sp)) return EXIT_FAILURE;
    n = mw_get_optimized_parameter_count(&ap);
    printf("number_parameters: %d\n", n);
    if (n < 1) {
        fprintf(stderr, "%s: no parameters are optimised in %s\n", argv[0], parameters_file);
        return EXIT_FAILURE;
    }
    point = (double*)malloc(sizeof(double) * n);
    step = (double*)malloc(sizeof(double) * n);
    h = (double*)malloc(sizeof(double) * n * n);
    sigma = (double*)malloc(sizeof(double) * n * n);
    sd = (double*)malloc(sizeof(double) * n);
    if (!point || !step || !h || !sigma || !sd) return EXIT_FAILURE;
    mw_get_search_parameters(&ap, point);
    mw_get_step(&ap, step);

    if ((rc = mw_ctx_create(&ap, &sp, 0, &ctx))) {
        fprintf(stderr, "APP: %s\n", mw_last_error());
        return rc;
    }
    printf("Calculating hessian...\n");
    if ((rc = mw_ctx_hessian(ctx, point, step, MW_PREC_FP64, h, NULL, &value))) {
        fprintf(stderr, "APP: error calculating the hessian: %d (%s)\n", rc, mw_last_error());
        return rc;
    }
    printf("likelihood at the point: %.20f\n", value);
    printf("Number of stars: %d\n", sp.number_stars);
    if (mw_errors(n, h, sp.number_stars, sigma, sd)) {
        fprintf(stderr, "APP: the hessian is singular\n");
        return EXIT_FAILURE;
    }
    f = fopen(outfile, "w");
    if (!f) {
        fprintf(stderr, "APP: cannot open %s\n", outfile);
        return EXIT_FAILURE;
    }
    report(f, n, h, sigma, sd);
    fclose(f);
    report(stdout, n, h, sigma, sd);

    mw_ctx_destroy(ctx);
    free(point);
    free(step);
    free(h);
    free(sigma);
    free(sd);
    mw_free_star_points(&sp);
    mw_free_astronomy_parameters(&ap);
    return EXIT_SUCCESS;
}
