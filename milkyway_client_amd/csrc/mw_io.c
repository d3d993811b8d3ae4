/*
 * mw_io.c -- host-side readers for astronomy parameter files and star files,
 * and the evaluation-state set-up (declared in mw_io.h).
 *
 * The text formats are the reference's:
 *   parameter file   milkyway_separation/cpu/parameters.c:107-228
 *   arrays           util/io_util.c:55-71 (doubles), :125-141 (ints):
 *                    "name[N]: v0, v1, ..." -- commas optional, as fscanf's
 *                    ", " directive there matches zero or one comma
 *   star file        milkyway_separation/cpu/star_points.c:69-81
 * Every fscanf directive here starts with a blank, so any run of whitespace
 * (newlines included) separates fields, as the reference's trailing "\n"
 * directives allow.  Unlike the reference, a field that does not parse is an
 * error rather than a silently unset value.
 */
#include "mw_io.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int io_error(const char* filename, const char* what)
{
    fprintf(stderr, "Error reading astronomy parameters file [%s]: %s\n", filename, what);
    return 1;
}

/* "name[N]:" then N doubles; util/io_util.c:55-71.  *out gets N values
 * (N >= 1) and *size N. */
static int read_double_array(FILE* f, const char* name, double** out, int* size)
{
    char fmt[96];
    int n = 0, i;
    snprintf(fmt, sizeof fmt, " %s [%%d] :", name);
    if (fscanf(f, fmt, &n) != 1 || n < 1) return 1;
    *out = (double*)calloc((size_t)n, sizeof(double));
    if (!*out) return 1;
    for (i = 0; i < n; i++) {
        if (fscanf(f, " %lf", &(*out)[i]) != 1) return 1;
        if (i < n - 1 && fscanf(f, " ,") < 0) return 1;
    }
    *size = n;
    return 0;
}

/* "name[N]:" then N ints; util/io_util.c:125-141. */
static int read_int_array(FILE* f, const char* name, int** out, int* size)
{
    char fmt[96];
    int n = 0, i;
    snprintf(fmt, sizeof fmt, " %s [%%d] :", name);
    if (fscanf(f, fmt, &n) != 1 || n < 1) return 1;
    *out = (int*)calloc((size_t)n, sizeof(int));
    if (!*out) return 1;
    for (i = 0; i < n; i++) {
        if (fscanf(f, " %d", &(*out)[i]) != 1) return 1;
        if (i < n - 1 && fscanf(f, " ,") < 0) return 1;
    }
    *size = n;
    return 0;
}

/* "key: min, max, steps" for the whole wedge, "key[k]: min, max, steps" for
 * a cut (parameters.c:168-222). */
static int read_range(FILE* f, const char* key, int indexed, double* lo, double* hi, int* steps)
{
    char fmt[96];
    int k;
    if (indexed) {
        snprintf(fmt, sizeof fmt, " %s [%%d] : %%lf , %%lf , %%d", key);
        return fscanf(f, fmt, &k, lo, hi, steps) == 4 ? 0 : 1;
    }
    snprintf(fmt, sizeof fmt, " %s : %%lf , %%lf , %%d", key);
    return fscanf(f, fmt, lo, hi, steps) == 3 ? 0 : 1;
}

/* A region's step sizes and calculation range (parameters.c:186-190). */
static void finish_region(INTEGRAL* g)
{
    g->r_step_size = (g->r_max - g->r_min) / (double)g->r_steps;
    g->mu_step_size = (g->mu_max - g->mu_min) / (double)g->mu_steps;
    g->nu_step_size = (g->nu_max - g->nu_min) / (double)g->nu_steps;
    g->min_calculation = 0;
    g->max_calculation = (long)g->r_steps * g->mu_steps * g->nu_steps;
}

static int read_region(FILE* f, int cut, INTEGRAL* g)
{
    if (read_range(f, cut ? "r_cut[min,max,steps]" : "r[min,max,steps]", cut, &g->r_min, &g->r_max,
                   &g->r_steps) ||
        read_range(f, cut ? "mu_cut[min,max,steps]" : "mu[min,max,steps]", cut, &g->mu_min, &g->mu_max,
                   &g->mu_steps) ||
        read_range(f, cut ? "nu_cut[min,max,steps]" : "nu[min,max,steps]", cut, &g->nu_min, &g->nu_max,
                   &g->nu_steps))
        return 1;
    if (g->r_steps < 1 || g->mu_steps < 1 || g->nu_steps < 1) return 1;
    finish_region(g);
    return 0;
}

static int read_stream(FILE* f, ASTRONOMY_PARAMETERS* ap, int i, const char* filename)
{
    int n;
    if (fscanf(f, " stream_weight : %lf", &ap->stream_weights[i]) != 1 ||
        fscanf(f, " stream_weight_step : %lf", &ap->stream_weight_step[i]) != 1 ||
        fscanf(f, " stream_weight_min : %lf", &ap->stream_weight_min[i]) != 1 ||
        fscanf(f, " stream_weight_max : %lf", &ap->stream_weight_max[i]) != 1 ||
        fscanf(f, " optimize_weight : %d", &ap->stream_weight_optimize[i]) != 1)
        return io_error(filename, "stream weight fields");
    if (read_double_array(f, "stream_parameters", &ap->stream_parameters[i], &n) ||
        n < ap->number_stream_parameters)
        return io_error(filename, "stream_parameters");
    if (read_double_array(f, "stream_step", &ap->stream_step[i], &n) || n < ap->number_stream_parameters)
        return io_error(filename, "stream_step");
    if (read_double_array(f, "stream_min", &ap->stream_min[i], &n) || n < ap->number_stream_parameters)
        return io_error(filename, "stream_min");
    if (read_double_array(f, "stream_max", &ap->stream_max[i], &n) || n < ap->number_stream_parameters)
        return io_error(filename, "stream_max");
    if (read_int_array(f, "optimize_parameter", &ap->stream_optimize[i], &n) ||
        n < ap->number_stream_parameters)
        return io_error(filename, "stream optimize_parameter");
    return 0;
}

static int fread_parameters(FILE* f, ASTRONOMY_PARAMETERS* ap, const char* filename)
{
    int n, i;
    long pos = ftell(f);

    /* A file without a version line is version 0.01 and has no aux_bg_profile
     * line (parameters.c:111-117, :157-160). */
    if (fscanf(f, " parameters_version : %lf", &ap->parameters_version) != 1) {
        ap->parameters_version = 0.01;
        if (fseek(f, pos, SEEK_SET) != 0) return io_error(filename, "seek");
    }
    if (fscanf(f, " number_parameters : %d", &ap->number_background_parameters) != 1 ||
        ap->number_background_parameters < 1)
        return io_error(filename, "number_parameters");
    if (fscanf(f, " background_weight : %lf", &ap->background_weight) != 1)
        return io_error(filename, "background_weight");

    /* The arrays carry their own lengths; number_parameters is the count the
     * rest of the code uses, so every array must hold at least that many. */
    if (read_double_array(f, "background_parameters", &ap->background_parameters, &n) ||
        n < ap->number_background_parameters)
        return io_error(filename, "background_parameters");
    if (read_double_array(f, "background_step", &ap->background_step, &n) ||
        n < ap->number_background_parameters)
        return io_error(filename, "background_step");
    if (read_double_array(f, "background_min", &ap->background_min, &n) ||
        n < ap->number_background_parameters)
        return io_error(filename, "background_min");
    if (read_double_array(f, "background_max", &ap->background_max, &n) ||
        n < ap->number_background_parameters)
        return io_error(filename, "background_max");
    if (read_int_array(f, "optimize_parameter", &ap->background_optimize, &n) ||
        n < ap->number_background_parameters)
        return io_error(filename, "background optimize_parameter");

    if (fscanf(f, " number_streams : %d , %d", &ap->number_streams, &ap->number_stream_parameters) != 2 ||
        ap->number_streams < 0)
        return io_error(filename, "number_streams");
    /* Stream parameters are mu, r, theta, phi, sigma (evaluation_optimized.c:83-107). */
    if (ap->number_stream_parameters != 5) return io_error(filename, "streams must have 5 parameters");

    if (ap->number_streams > 0) {
        const size_t ns = (size_t)ap->number_streams;
        ap->stream_weights = (double*)calloc(ns, sizeof(double));
        ap->stream_weight_step = (double*)calloc(ns, sizeof(double));
        ap->stream_weight_min = (double*)calloc(ns, sizeof(double));
        ap->stream_weight_max = (double*)calloc(ns, sizeof(double));
        ap->stream_weight_optimize = (int*)calloc(ns, sizeof(int));
        ap->stream_parameters = (double**)calloc(ns, sizeof(double*));
        ap->stream_step = (double**)calloc(ns, sizeof(double*));
        ap->stream_min = (double**)calloc(ns, sizeof(double*));
        ap->stream_max = (double**)calloc(ns, sizeof(double*));
        ap->stream_optimize = (int**)calloc(ns, sizeof(int*));
        if (!ap->stream_weights || !ap->stream_weight_step || !ap->stream_weight_min ||
            !ap->stream_weight_max || !ap->stream_weight_optimize || !ap->stream_parameters ||
            !ap->stream_step || !ap->stream_min || !ap->stream_max || !ap->stream_optimize)
            return io_error(filename, "out of memory");
    }
    for (i = 0; i < ap->number_streams; i++)
        if (read_stream(f, ap, i, filename)) return 1;

    if (fscanf(f, " convolve : %d", &ap->convolve) != 1 || ap->convolve < 0)
        return io_error(filename, "convolve");
    if (fscanf(f, " sgr_coordinates : %d", &ap->sgr_coordinates) != 1)
        return io_error(filename, "sgr_coordinates");
    ap->aux_bg_profile = 0;
    if (ap->parameters_version > 0.01 && fscanf(f, " aux_bg_profile : %d", &ap->aux_bg_profile) != 1)
        return io_error(filename, "aux_bg_profile");
    if (fscanf(f, " wedge : %d", &ap->wedge) != 1) return io_error(filename, "wedge");

    ap->number_integrals = 1;
    ap->integral = (INTEGRAL**)calloc(1, sizeof(INTEGRAL*));
    if (!ap->integral || !(ap->integral[0] = (INTEGRAL*)calloc(1, sizeof(INTEGRAL))))
        return io_error(filename, "out of memory");
    if (read_region(f, 0, ap->integral[0])) return io_error(filename, "r/mu/nu ranges");

    if (fscanf(f, " number_cuts : %d", &n) != 1 || n < 0) return io_error(filename, "number_cuts");
    if (n > 0) {
        INTEGRAL** grown = (INTEGRAL**)realloc(ap->integral, sizeof(INTEGRAL*) * (size_t)(n + 1));
        if (!grown) return io_error(filename, "out of memory");
        ap->integral = grown;
        for (i = 1; i <= n; i++) ap->integral[i] = NULL;
        for (i = 1; i <= n; i++) {
            ap->number_integrals = i + 1;
            if (!(ap->integral[i] = (INTEGRAL*)calloc(1, sizeof(INTEGRAL))))
                return io_error(filename, "out of memory");
            if (read_region(f, 1, ap->integral[i])) return io_error(filename, "cut ranges");
        }
    }
    return 0;
}

int mw_read_astronomy_parameters(const char* filename, ASTRONOMY_PARAMETERS* ap)
{
    FILE* f;
    int rc;
    if (!filename || !ap) return 1;
    memset(ap, 0, sizeof *ap);
    f = fopen(filename, "r");
    if (!f) {
        fprintf(stderr, "Couldn't find input file [%s] to read astronomy parameters.\n", filename);
        return 1;
    }
    rc = fread_parameters(f, ap, filename);
    fclose(f);
    if (rc) mw_free_astronomy_parameters(ap);
    return rc;
}

/* "name[N]: v0, v1, ..." with 20 decimals (util/io_util.c:30-39). */
static void write_double_array(FILE* f, const char* name, int n, const double* v)
{
    int i;
    fprintf(f, "%s[%d]: ", name, n);
    for (i = 0; i < n; i++) fprintf(f, i + 1 < n ? "%.20f, " : "%.20f", v[i]);
    fprintf(f, "\n");
}

/* "name[N]:  a, b, ..." (util/io_util.c:41-50). */
static void write_int_array(FILE* f, const char* name, int n, const int* v)
{
    int i;
    fprintf(f, "%s[%d]: ", name, n);
    for (i = 0; i < n; i++) fprintf(f, i == 0 ? " %d" : ", %d", v[i]);
    fprintf(f, "\n");
}

static void write_region(FILE* f, const INTEGRAL* g, int cut)
{
    fprintf(f, cut ? "r_cut[min,max,steps][3]: %f, %f, %d\n" : "r[min,max,steps]: %f, %f, %d\n", g->r_min,
            g->r_max, g->r_steps);
    fprintf(f, cut ? "mu_cut[min,max,steps][3]: %f, %f, %d\n" : "mu[min,max,steps]: %f, %f, %d\n", g->mu_min,
            g->mu_max, g->mu_steps);
    fprintf(f, cut ? "nu_cut[min,max,steps][3]: %f, %f, %d\n" : "nu[min,max,steps]: %f, %f, %d\n", g->nu_min,
            g->nu_max, g->nu_steps);
}

int mw_write_astronomy_parameters(const char* filename, const ASTRONOMY_PARAMETERS* ap)
{
    FILE* f;
    int i, nbg, nsp;
    if (!filename || !ap) return 1;
    f = fopen(filename, "w");
    if (!f) {
        fprintf(stderr, "Couldn't find output file [%s] to write astronomy parameters.\n", filename);
        return 1;
    }
    nbg = ap->number_background_parameters;
    nsp = ap->number_stream_parameters;
    fprintf(f, "parameters_version: %f\n", ap->parameters_version);
    fprintf(f, "number_parameters: %d\n", nbg);
    fprintf(f, "background_weight: %f\n", ap->background_weight);
    write_double_array(f, "background_parameters", nbg, ap->background_parameters);
    write_double_array(f, "background_step", nbg, ap->background_step);
    write_double_array(f, "background_min", nbg, ap->background_min);
    write_double_array(f, "background_max", nbg, ap->background_max);
    write_int_array(f, "optimize_parameter", nbg, ap->background_optimize);
    fprintf(f, "number_streams: %d, %d\n", ap->number_streams, nsp);
    for (i = 0; i < ap->number_streams; i++) {
        fprintf(f, "stream_weight: %f\n", ap->stream_weights[i]);
        fprintf(f, "stream_weight_step: %f\n", ap->stream_weight_step[i]);
        fprintf(f, "stream_weight_min: %f\n", ap->stream_weight_min[i]);
        fprintf(f, "stream_weight_max: %f\n", ap->stream_weight_max[i]);
        fprintf(f, "optimize_weight: %d\n", ap->stream_weight_optimize[i]);
        write_double_array(f, "stream_parameters", nsp, ap->stream_parameters[i]);
        write_double_array(f, "stream_step", nsp, ap->stream_step[i]);
        write_double_array(f, "stream_min", nsp, ap->stream_min[i]);
        write_double_array(f, "stream_max", nsp, ap->stream_max[i]);
        write_int_array(f, "optimize_parameter", nsp, ap->stream_optimize[i]);
    }
    fprintf(f, "convolve: %d\n", ap->convolve);
    fprintf(f, "sgr_coordinates: %d\n", ap->sgr_coordinates);
    fprintf(f, "aux_bg_profile: %d\n", ap->aux_bg_profile);
    fprintf(f, "wedge: %d\n", ap->wedge);
    write_region(f, ap->integral[0], 0);
    fprintf(f, "number_cuts: %d\n", ap->number_integrals - 1);
    for (i = 1; i < ap->number_integrals; i++) write_region(f, ap->integral[i], 1);
    return fclose(f) == 0 ? 0 : 1;
}

void mw_free_astronomy_parameters(ASTRONOMY_PARAMETERS* ap)
{
    int i;
    if (!ap) return;
    free(ap->background_parameters);
    free(ap->background_step);
    free(ap->background_min);
    free(ap->background_max);
    free(ap->background_optimize);
    for (i = 0; i < ap->number_streams; i++) {
        if (ap->stream_parameters) free(ap->stream_parameters[i]);
        if (ap->stream_step) free(ap->stream_step[i]);
        if (ap->stream_min) free(ap->stream_min[i]);
        if (ap->stream_max) free(ap->stream_max[i]);
        if (ap->stream_optimize) free(ap->stream_optimize[i]);
    }
    free(ap->stream_weights);
    free(ap->stream_weight_step);
    free(ap->stream_weight_min);
    free(ap->stream_weight_max);
    free(ap->stream_weight_optimize);
    free(ap->stream_parameters);
    free(ap->stream_step);
    free(ap->stream_min);
    free(ap->stream_max);
    free(ap->stream_optimize);
    if (ap->integral)
        for (i = 0; i < ap->number_integrals; i++) free(ap->integral[i]);
    free(ap->integral);
    memset(ap, 0, sizeof *ap);
}

int mw_read_star_points(const char* filename, STAR_POINTS* sp)
{
    FILE* f;
    int i, n = 0;
    if (!filename || !sp) return 1;
    memset(sp, 0, sizeof *sp);
    f = fopen(filename, "r");
    if (!f) {
        fprintf(stderr, "Couldn't find input file %s.\n", filename);
        return 1;
    }
    if (fscanf(f, " %d", &n) != 1 || n < 0) {
        fprintf(stderr, "Error reading star count from %s.\n", filename);
        fclose(f);
        return 1;
    }
    sp->stars = (double**)calloc(n > 0 ? (size_t)n : 1, sizeof(double*));
    if (!sp->stars) {
        fclose(f);
        return 1;
    }
    for (i = 0; i < n; i++) {
        double* s = (double*)malloc(3 * sizeof(double));
        if (!s) break;
        sp->stars[i] = s;
        sp->number_stars = i + 1;
        if (fscanf(f, " %lf %lf %lf", &s[0], &s[1], &s[2]) != 3) {
            fprintf(stderr, "Error reading star %d of %d from %s.\n", i, n, filename);
            fclose(f);
            mw_free_star_points(sp);
            return 1;
        }
    }
    fclose(f);
    if (sp->number_stars != n) {
        mw_free_star_points(sp);
        return 1;
    }
    return 0;
}

int mw_write_star_points(const char* filename, const STAR_POINTS* sp)
{
    FILE* f;
    int i;
    if (!filename || !sp) return 1;
    f = fopen(filename, "w");
    if (!f) {
        fprintf(stderr, "Couldn't find input file %s.\n", filename);
        return 1;
    }
    fprintf(f, "%d\n", sp->number_stars);
    for (i = 0; i < sp->number_stars; i++)
        fprintf(f, "%f %f %f\n", sp->stars[i][0], sp->stars[i][1], sp->stars[i][2]);
    return fclose(f) == 0 ? 0 : 1;
}

void mw_free_star_points(STAR_POINTS* sp)
{
    int i;
    if (!sp) return;
    if (sp->stars)
        for (i = 0; i < sp->number_stars; i++) free(sp->stars[i]);
    free(sp->stars);
    sp->stars = NULL;
    sp->number_stars = 0;
}

int mw_initialize_state(const ASTRONOMY_PARAMETERS* ap, const STAR_POINTS* sp, EVALUATION_STATE* es)
{
    int i;
    if (!ap || !sp || !es) return 1;
    memset(es, 0, sizeof *es);
    es->number_streams = ap->number_streams;
    es->total_stars = sp->number_stars;
    es->number_integrals = ap->number_integrals;
    es->stream_integrals = (double*)calloc(ap->number_streams > 0 ? (size_t)ap->number_streams : 1,
                                           sizeof(double));
    es->integral = (INTEGRAL_AREA**)calloc((size_t)ap->number_integrals, sizeof(INTEGRAL_AREA*));
    if (!es->stream_integrals || !es->integral) {
        mw_free_state(es);
        return 1;
    }
    for (i = 0; i < ap->number_integrals; i++) {
        const INTEGRAL* g = ap->integral[i];
        INTEGRAL_AREA* a = (INTEGRAL_AREA*)calloc(1, sizeof(INTEGRAL_AREA));
        if (!a) {
            mw_free_state(es);
            return 1;
        }
        es->integral[i] = a;
        a->mu_min = g->mu_min;
        a->mu_max = g->mu_max;
        a->mu_steps = g->mu_steps;
        a->nu_min = g->nu_min;
        a->nu_max = g->nu_max;
        a->nu_steps = g->nu_steps;
        a->r_min = g->r_min;
        a->r_max = g->r_max;
        a->r_steps = g->r_steps;
        a->min_calculation = g->min_calculation;
        a->max_calculation = g->max_calculation;
        a->current_calculation = g->min_calculation;
        a->mu_step_size = (a->mu_max - a->mu_min) / a->mu_steps;
        a->nu_step_size = (a->nu_max - a->nu_min) / a->nu_steps;
        a->r_step_size = (a->r_max - a->r_min) / a->r_steps;
        a->number_streams = ap->number_streams;
        a->stream_integrals = (double*)calloc(ap->number_streams > 0 ? (size_t)ap->number_streams : 1,
                                              sizeof(double));
        if (!a->stream_integrals) {
            mw_free_state(es);
            return 1;
        }
    }
    return 0;
}

void mw_free_state(EVALUATION_STATE* es)
{
    int i;
    if (!es) return;
    if (es->integral)
        for (i = 0; i < es->number_integrals; i++)
            if (es->integral[i]) {
                free(es->integral[i]->stream_integrals);
                free(es->integral[i]);
            }
    free(es->integral);
    free(es->stream_integrals);
    memset(es, 0, sizeof *es);
}

/* fwrite_integral_area (evaluation_state.c:33-66): geometry with three
 * decimals, the range, the partial integrals with twenty. */
static void write_area(FILE* f, const INTEGRAL_AREA* a)
{
    fprintf(f, "mu[min,max,steps]: %.3f, %.3f, %d\n", a->mu_min, a->mu_max, a->mu_steps);
    fprintf(f, "nu[min,max,steps]: %.3f, %.3f, %d\n", a->nu_min, a->nu_max, a->nu_steps);
    fprintf(f, " r[min,max,steps]: %.3f, %.3f, %d\n", a->r_min, a->r_max, a->r_steps);
    fprintf(f, "min_calculation: %ld, max_calculation: %ld, current_calculation: %ld\n", a->min_calculation,
            a->max_calculation, a->current_calculation);
    fprintf(f, "background_integral: %.20f\n", a->background_integral);
    write_double_array(f, "stream_integrals", a->number_streams, a->stream_integrals);
}

int mw_write_checkpoint(const char* filename, const EVALUATION_STATE* es)
{
    FILE* f;
    int i;
    if (!filename || !es) return 1;
    f = fopen(filename, "w+");
    if (!f) {
        fprintf(stderr, "APP: error writing checkpoint (opening checkpoint file)\n");
        return 1;
    }
    fprintf(f, "background_integral: %.20f\n", es->background_integral);
    fprintf(f, "stream_integrals[%d]: ", es->number_streams);
    for (i = 0; i < es->number_streams; i++)
        fprintf(f, i + 1 < es->number_streams ? "%.20f, " : "%.20f", es->stream_integrals[i]);
    fprintf(f, "\n");
    fprintf(f, "prob_sum: %.20f, num_zero: %d, bad_jacobians: %d\n", es->prob_sum, es->num_zero,
            es->bad_jacobians);
    fprintf(f, "current_star_point: %d\n", es->current_star_point);
    fprintf(f, "current_integral: %d\n", es->current_integral);
    fprintf(f, "number_integrals: %d\n", es->number_integrals);
    for (i = 0; i < es->number_integrals; i++) write_area(f, es->integral[i]);
    if (fclose(f)) {
        fprintf(stderr, "APP: error writing checkpoint (closing checkpoint file)\n");
        return 1;
    }
    return 0;
}

/* n values after "name[n]:" into out, which holds exactly n. */
static int read_fixed_array(FILE* f, const char* name, int n, double* out)
{
    char fmt[96];
    int got = -1, i;
    snprintf(fmt, sizeof fmt, " %s [%%d] :", name);
    if (fscanf(f, fmt, &got) != 1 || got != n) return 1;
    for (i = 0; i < n; i++) {
        if (fscanf(f, " %lf", &out[i]) != 1) return 1;
        if (i < n - 1 && fscanf(f, " ,") < 0) return 1;
    }
    return 0;
}

/* fread_integral_area (evaluation_state.c:68-95): step sizes are recomputed
 * from the geometry as read. */
static int read_area(FILE* f, INTEGRAL_AREA* a)
{
    if (fscanf(f, " mu[min,max,steps] : %lf , %lf , %d", &a->mu_min, &a->mu_max, &a->mu_steps) != 3 ||
        fscanf(f, " nu[min,max,steps] : %lf , %lf , %d", &a->nu_min, &a->nu_max, &a->nu_steps) != 3 ||
        fscanf(f, " r[min,max,steps] : %lf , %lf , %d", &a->r_min, &a->r_max, &a->r_steps) != 3)
        return 1;
    if (a->mu_steps < 1 || a->nu_steps < 1 || a->r_steps < 1) return 1;
    a->mu_step_size = (a->mu_max - a->mu_min) / a->mu_steps;
    a->nu_step_size = (a->nu_max - a->nu_min) / a->nu_steps;
    a->r_step_size = (a->r_max - a->r_min) / a->r_steps;
    if (fscanf(f, " min_calculation : %ld , max_calculation : %ld , current_calculation : %ld",
               &a->min_calculation, &a->max_calculation, &a->current_calculation) != 3)
        return 1;
    if (fscanf(f, " background_integral : %lf", &a->background_integral) != 1) return 1;
    return read_fixed_array(f, "stream_integrals", a->number_streams, a->stream_integrals);
}

int mw_read_checkpoint(const char* filename, EVALUATION_STATE* es)
{
    FILE* f;
    EVALUATION_STATE t;
    INTEGRAL_AREA* areas = NULL;
    double* totals = NULL;
    int i, n_file = -1, n_parsed = 0, rc = 1;
    if (!filename || !es) return 1;
    f = fopen(filename, "r");
    if (!f) return 0;
    /* Parse into copies, so that a bad file leaves es as it was. */
    t = *es;
    totals = (double*)calloc(es->number_streams > 0 ? (size_t)es->number_streams : 1, sizeof(double));
    areas = (INTEGRAL_AREA*)calloc(es->number_integrals > 0 ? (size_t)es->number_integrals : 1,
                                   sizeof(INTEGRAL_AREA));
    if (!totals || !areas) goto out;
    if (fscanf(f, " background_integral : %lf", &t.background_integral) != 1) goto out;
    if (read_fixed_array(f, "stream_integrals", es->number_streams, totals)) goto out;
    if (fscanf(f, " prob_sum : %lf , num_zero : %d , bad_jacobians : %d", &t.prob_sum, &t.num_zero,
               &t.bad_jacobians) != 3 ||
        fscanf(f, " current_star_point : %d", &t.current_star_point) != 1 ||
        fscanf(f, " current_integral : %d", &t.current_integral) != 1 ||
        fscanf(f, " number_integrals : %d", &n_file) != 1 || n_file != es->number_integrals)
        goto out;
    for (i = 0; i < es->number_integrals; i++) {
        areas[i] = *es->integral[i];
        areas[i].stream_integrals = (double*)calloc(
            areas[i].number_streams > 0 ? (size_t)areas[i].number_streams : 1, sizeof(double));
        if (!areas[i].stream_integrals) goto out;
        n_parsed = i + 1;
        if (read_area(f, &areas[i])) goto out;
    }
    /* Commit: values into es's own arrays. */
    es->background_integral = t.background_integral;
    for (i = 0; i < es->number_streams; i++) es->stream_integrals[i] = totals[i];
    es->prob_sum = t.prob_sum;
    es->num_zero = t.num_zero;
    es->bad_jacobians = t.bad_jacobians;
    es->current_star_point = t.current_star_point;
    es->current_integral = t.current_integral;
    for (i = 0; i < es->number_integrals; i++) {
        INTEGRAL_AREA* a = es->integral[i];
        double* parsed = areas[i].stream_integrals;
        double* keep = a->stream_integrals;
        int s;
        for (s = 0; s < a->number_streams; s++) keep[s] = parsed[s];
        *a = areas[i];
        a->stream_integrals = keep;
    }
    rc = 0;
out:
    for (i = 0; i < n_parsed; i++) free(areas[i].stream_integrals);
    free(areas);
    free(totals);
    fclose(f);
    return rc;
}

double mw_calculate_progress(const EVALUATION_STATE* es)
{
    double total = 0.0, done = 0.0;
    int i;
    for (i = 0; i < es->number_integrals; i++) {
        const INTEGRAL_AREA* a = es->integral[i];
        const double cells = (double)a->r_steps * a->mu_steps * a->nu_steps;
        total += cells;
        if (i < es->current_integral) {
            done += cells;
        } else if (i == es->current_integral) {
            /* get_steps (evaluation_state.c:129-151): r fastest, then nu, then mu */
            const long c = a->current_calculation;
            const long r = c % a->r_steps;
            const long nu = (c / a->r_steps) % a->nu_steps;
            const long mu = c / ((long)a->r_steps * a->nu_steps);
            done += r + (double)nu * a->r_steps + (double)mu * a->nu_steps * a->r_steps;
        }
    }
    total += es->total_stars;
    done += es->current_star_point;
    return done / total;
}

int mw_get_optimized_parameter_count(const ASTRONOMY_PARAMETERS* ap)
{
    int i, j, count = 0;
    for (i = 0; i < ap->number_background_parameters; i++) count += ap->background_optimize[i] ? 1 : 0;
    for (i = 0; i < ap->number_streams; i++) {
        count += ap->stream_weight_optimize[i] ? 1 : 0;
        for (j = 0; j < ap->number_stream_parameters; j++) count += ap->stream_optimize[i][j] ? 1 : 0;
    }
    return count;
}

void mw_set_astronomy_parameters(ASTRONOMY_PARAMETERS* ap, const double* parameters)
{
    int i, j, k = 0;
    for (i = 0; i < ap->number_background_parameters; i++)
        if (ap->background_optimize[i]) ap->background_parameters[i] = parameters[k++];
    for (i = 0; i < ap->number_streams; i++) {
        if (ap->stream_weight_optimize[i]) ap->stream_weights[i] = parameters[k++];
        for (j = 0; j < ap->number_stream_parameters; j++)
            if (ap->stream_optimize[i][j]) ap->stream_parameters[i][j] = parameters[k++];
    }
}
