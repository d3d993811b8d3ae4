/*
 * mw_client.c -- command-line separation client on the MI355X evaluator.
 *
 * Behaves like the reference client's worker (src/milkyway_client_app.c:197-288)
 * with the HIP evaluator selected where the reference selects cpu_evaluate
 * (:263-274):
 *
 *   milkyway_separation_hip -a <astronomy parameter file> -p <star file>
 *                           [-s <file>] [-o <file>] [-b <arg>] p1 p2 ... pN
 *
 * Options come first; everything from the first non-option argument on is the
 * parameter vector, each value parsed with strtod (POSIX-style option
 * parsing, as popt's POSIXMEHARDER context does there, :130-134).  -s, -o and
 * -b are accepted and unused, as in the reference.  The vector length must
 * equal the file's optimised-parameter count (:243-259).  The result goes to
 * stderr as
 *
 *   <search_likelihood> %0.20f </search_likelihood>
 *   <search_application> ... </search_application>
 *
 * where the likelihood is the evaluator's value minus 3.0 (:276-281).  Every
 * failure exits with EXIT_FAILURE after a message on stderr; an evaluator
 * failure exits with the evaluator's error code.
 *
 * MW_HIP_PRECISION=mixed in the environment selects the mixed-precision
 * kernels (the default is FP64).
 */
#include <errno.h>
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mw_hip.h"
#include "mw_io.h"

#define CLIENT_VERSION "milkyway_separation_hip 0.1"

static void usage(FILE* out, const char* prog)
{
    fprintf(out,
            "Usage: %s [OPTION...] parameters...\n"
            "  -b, --boinc-init-graphics=STRING     Argument to boinc_init_graphics\n"
            "  -s, --search-parameter-file=STRING   Search parameter file name\n"
            "  -p, --star-points-file=STRING        Star points files\n"
            "  -a, --astronomy-parameter-file=STRING  Astronomy parameter file\n"
            "  -o, --output=STRING                  Output file\n"
            "  -?, --help                           Show this help message\n",
            prog);
}

int main(int argc, char** argv)
{
    static const struct option longopts[] = {
        { "boinc-init-graphics", required_argument, NULL, 'b' },
        { "search-parameter-file", required_argument, NULL, 's' },
        { "star-points-file", required_argument, NULL, 'p' },
        { "astronomy-parameter-file", required_argument, NULL, 'a' },
        { "output", required_argument, NULL, 'o' },
        { "help", no_argument, NULL, 'H' },
        { NULL, 0, NULL, 0 },
    };
    const char* astronomy_parameter_file = NULL;
    const char* star_points_file = NULL;
    ASTRONOMY_PARAMETERS ap;
    STAR_POINTS sp;
    EVALUATION_STATE es;
    double* parameters;
    double likelihood;
    int number_parameters, ap_number_parameters, ret1, ret2, o, i;

    opterr = 0;
    while ((o = getopt_long(argc, argv, "+b:s:p:a:o:", longopts, NULL)) != -1) {
        switch (o) {
        case 'a': astronomy_parameter_file = optarg; break;
        case 'p': star_points_file = optarg; break;
        case 'b':
        case 's':
        case 'o': break;
        case 'H':
            usage(stdout, argv[0]);
            return EXIT_SUCCESS;
        case '?':
            if (optopt == '?') {
                usage(stdout, argv[0]);
                return EXIT_SUCCESS;
            }
            /* fall through */
        default:
            fprintf(stderr, "%s: bad option %s\n", argv[0], argv[optind - 1]);
            usage(stderr, argv[0]);
            return EXIT_FAILURE;
        }
    }

    number_parameters = argc - optind;
    if (number_parameters < 1) {
        fprintf(stderr, "Could not parse parameters from the command line\n");
        return EXIT_FAILURE;
    }
    parameters = (double*)malloc(sizeof(double) * (size_t)number_parameters);
    if (!parameters) return EXIT_FAILURE;
    errno = 0;
    for (i = 0; i < number_parameters; i++) {
        parameters[i] = strtod(argv[optind + i], NULL);
        if (errno) {
            perror("error parsing command line parameters");
            usage(stderr, argv[0]);
            free(parameters);
            return EXIT_FAILURE;
        }
    }

    if (!astronomy_parameter_file || !star_points_file) {
        fprintf(stderr, "APP: both -a <astronomy parameter file> and -p <star points file> are required\n");
        free(parameters);
        return EXIT_FAILURE;
    }

    ret1 = mw_read_astronomy_parameters(astronomy_parameter_file, &ap);
    ret2 = mw_read_star_points(star_points_file, &sp);
    if (ret1)
        fprintf(stderr, "APP: error reading astronomy parameters from file %s: %d\n",
                astronomy_parameter_file, ret1);
    if (ret2)
        fprintf(stderr, "APP: error reading star points from file %s: %d\n", star_points_file, ret2);
    if (ret1 | ret2) {
        if (!ret1) mw_free_astronomy_parameters(&ap);
        if (!ret2) mw_free_star_points(&sp);
        free(parameters);
        return EXIT_FAILURE;
    }

    if (mw_initialize_state(&ap, &sp, &es)) {
        fprintf(stderr, "APP: out of memory initialising the evaluation state\n");
        return EXIT_FAILURE;
    }

    ap_number_parameters = mw_get_optimized_parameter_count(&ap);
    if (number_parameters != ap_number_parameters) {
        fprintf(stderr,
                "Error reading parameters: number of parameters from the "
                "command line (%d) does not match the number of parameters "
                "to be optimized in %s (%d)\n",
                number_parameters, astronomy_parameter_file, ap_number_parameters);
        free(parameters);
        mw_free_state(&es);
        mw_free_astronomy_parameters(&ap);
        mw_free_star_points(&sp);
        return EXIT_FAILURE;
    }

    mw_set_astronomy_parameters(&ap, parameters);

    {
        /* The evaluator is called through the reference's EvalFunc type
         * (milkyway_separation/cpu/simple_evaluator.h:26). */
        const EvalFunc evaluate = mw_hip_evaluate;
        const char* prec = getenv("MW_HIP_PRECISION");
        const int mixed = prec && (strcmp(prec, "mixed") == 0 || strcmp(prec, "1") == 0);
        likelihood = evaluate(parameters, &ap, &es, &sp) - 3.0;
        fprintf(stderr, "<search_likelihood> %0.20f </search_likelihood>\n", likelihood);
        fprintf(stderr, "<search_application> %s %s </search_application>\n", CLIENT_VERSION,
                mixed ? "mixed" : "double");
    }

    free(parameters);
    mw_free_state(&es);
    mw_free_astronomy_parameters(&ap);
    mw_free_star_points(&sp);
    return EXIT_SUCCESS;
}
