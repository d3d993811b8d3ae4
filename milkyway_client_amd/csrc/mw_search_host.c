/*
 * mw_search_host.c -- host-only parts of the search glue (declared in
 * mw_search.h): compose steps, range and star splitters, parameter-vector
 * accessors, the central-difference Hessian driver, LUP inversion and the
 * parameter errors.
 */
#include "mw_search.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mw_io.h"

void mw_integral_compose(const double* integral_results, int num_results, double* results,
                         const ASTRONOMY_PARAMETERS* ap)
{
    const int S = ap->number_streams;
    int i, j;
    for (j = 0; j <= S; j++) results[j] = 0.0;
    for (i = 0; i < num_results; i++) {
        const double* r = integral_results + (size_t)(S + 1) * i;
        for (j = 0; j <= S; j++) results[j] += r[j];
    }
}

double mw_likelihood_compose(const double* results, size_t num_results, int total_number_stars)
{
    double prob_sum = 0.0, bad_jacobians = 0.0;
    size_t i;
    for (i = 0; i < num_results; i++) {
        prob_sum += results[2 * i];
        bad_jacobians += results[2 * i + 1];
    }
    prob_sum /= (total_number_stars - bad_jacobians);
    return prob_sum - 3.0;
}

void mw_split_astronomy_parameters(ASTRONOMY_PARAMETERS* ap, int rank, int max_rank)
{
    int i;
    for (i = 0; i < ap->number_integrals; i++) {
        INTEGRAL* g = ap->integral[i];
        const long total = (long)g->r_steps * g->nu_steps * g->mu_steps;
        g->min_calculation = (long)(total * ((double)rank / (double)max_rank));
        g->max_calculation = (long)(total * ((double)(rank + 1) / (double)max_rank));
    }
}

int mw_split_star_points(STAR_POINTS* sp, int rank, int max_rank)
{
    int first, last, n, i;
    double** rows;
    if (rank == 0 && max_rank == 0) return 0;
    first = (int)(((double)sp->number_stars) * (((double)rank) / ((double)max_rank)));
    last = (int)(((double)sp->number_stars) * (((double)rank + 1.0) / ((double)max_rank)));
    n = last - first;
    rows = (double**)calloc(n > 0 ? (size_t)n : 1, sizeof(double*));
    if (!rows) return 1;
    for (i = 0; i < n; i++) {
        rows[i] = (double*)malloc(3 * sizeof(double));
        if (!rows[i]) {
            while (i-- > 0) free(rows[i]);
            free(rows);
            return 1;
        }
        memcpy(rows[i], sp->stars[first + i], 3 * sizeof(double));
    }
    mw_free_star_points(sp);
    sp->stars = rows;
    sp->number_stars = n;
    return 0;
}

/* The optimised entries of a background array and per-stream arrays, in the
 * order of set_astronomy_parameters (parameters.c:332-363). */
static void gather(const ASTRONOMY_PARAMETERS* ap, const double* bg, const double* weight,
                   double* const* stream, double* out)
{
    int i, j, k = 0;
    for (i = 0; i < ap->number_background_parameters; i++)
        if (ap->background_optimize[i]) out[k++] = bg[i];
    for (i = 0; i < ap->number_streams; i++) {
        if (ap->stream_weight_optimize[i]) out[k++] = weight[i];
        for (j = 0; j < ap->number_stream_parameters; j++)
            if (ap->stream_optimize[i][j]) out[k++] = stream[i][j];
    }
}

void mw_get_search_parameters(const ASTRONOMY_PARAMETERS* ap, double* out)
{
    gather(ap, ap->background_parameters, ap->stream_weights, ap->stream_parameters, out);
}

void mw_get_step(const ASTRONOMY_PARAMETERS* ap, double* out)
{
    gather(ap, ap->background_step, ap->stream_weight_step, ap->stream_step, out);
}

void mw_get_min_parameters(const ASTRONOMY_PARAMETERS* ap, double* out)
{
    gather(ap, ap->background_min, ap->stream_weight_min, ap->stream_min, out);
}

void mw_get_max_parameters(const ASTRONOMY_PARAMETERS* ap, double* out)
{
    gather(ap, ap->background_max, ap->stream_weight_max, ap->stream_max, out);
}

int mw_hessian(int n, const double* point, const double* step, mw_vector_eval_fn f, void* user,
               double* hessian, double* gradient, double* value)
{
    /* Stencil rows: 0 = x; 1 + 2i = x + h_i, 2 + 2i = x - h_i; then per pair
     * i < j the four corners (+,+), (+,-), (-,+), (-,-). */
    const long n_pairs = (long)n * (n - 1) / 2;
    const long count = 1 + 2L * n + 4 * n_pairs;
    double *vec, *val;
    long row, pair;
    int i, j, rc;
    if (n < 1 || !point || !step || !f || !hessian) return 1;
    for (i = 0; i < n; i++)
        if (!(step[i] != 0.0)) return 1;
    if (count > 0x7fffffffL) return 1;
    vec = (double*)malloc(sizeof(double) * (size_t)count * n);
    val = (double*)malloc(sizeof(double) * (size_t)count);
    if (!vec || !val) {
        free(vec);
        free(val);
        return 1;
    }
    for (row = 0; row < count; row++) memcpy(vec + row * n, point, sizeof(double) * n);
    for (i = 0; i < n; i++) {
        vec[(1 + 2L * i) * n + i] += step[i];
        vec[(2 + 2L * i) * n + i] -= step[i];
    }
    row = 1 + 2L * n;
    for (i = 0; i < n; i++)
        for (j = i + 1; j < n; j++) {
            vec[row * n + i] += step[i];
            vec[row * n + j] += step[j];
            vec[(row + 1) * n + i] += step[i];
            vec[(row + 1) * n + j] -= step[j];
            vec[(row + 2) * n + i] -= step[i];
            vec[(row + 2) * n + j] += step[j];
            vec[(row + 3) * n + i] -= step[i];
            vec[(row + 3) * n + j] -= step[j];
            row += 4;
        }
    rc = f(user, vec, (int)count, val);
    if (rc == 0) {
        const double f0 = val[0];
        for (i = 0; i < n; i++) {
            const double fp = val[1 + 2 * i], fm = val[2 + 2 * i];
            hessian[(size_t)i * n + i] = (fp - 2.0 * f0 + fm) / (step[i] * step[i]);
            if (gradient) gradient[i] = (fp - fm) / (2.0 * step[i]);
        }
        pair = 0;
        for (i = 0; i < n; i++)
            for (j = i + 1; j < n; j++) {
                const double* c = val + 1 + 2L * n + 4 * pair;
                const double h = (c[0] - c[1] - c[2] + c[3]) / (4.0 * step[i] * step[j]);
                hessian[(size_t)i * n + j] = h;
                hessian[(size_t)j * n + i] = h;
                pair++;
            }
        if (value) *value = f0;
    }
    free(vec);
    free(val);
    return rc;
}

int mw_matrix_invert(int n, const double* a, double* inverse)
{
    double* lu;
    double* y;
    int* perm;
    int i, j, k, col, rc = 0;
    if (n < 1 || !a || !inverse) return 1;
    lu = (double*)malloc(sizeof(double) * (size_t)n * n);
    y = (double*)malloc(sizeof(double) * (size_t)n);
    perm = (int*)malloc(sizeof(int) * (size_t)n);
    if (!lu || !y || !perm) {
        rc = 1;
        goto done;
    }
    memcpy(lu, a, sizeof(double) * (size_t)n * n);
    for (i = 0; i < n; i++) perm[i] = i;
    /* LUP decomposition with partial pivoting: P A = L U, unit-diagonal L
     * below the diagonal of lu, U on and above it. */
    for (k = 0; k < n; k++) {
        double p = 0.0;
        int kp = k;
        for (i = k; i < n; i++)
            if (fabs(lu[(size_t)i * n + k]) > p) {
                p = fabs(lu[(size_t)i * n + k]);
                kp = i;
            }
        if (p == 0.0) {
            rc = 1;
            goto done;
        }
        if (kp != k) {
            int t = perm[k];
            perm[k] = perm[kp];
            perm[kp] = t;
            for (j = 0; j < n; j++) {
                double v = lu[(size_t)k * n + j];
                lu[(size_t)k * n + j] = lu[(size_t)kp * n + j];
                lu[(size_t)kp * n + j] = v;
            }
        }
        for (i = k + 1; i < n; i++) {
            const double m = lu[(size_t)i * n + k] / lu[(size_t)k * n + k];
            lu[(size_t)i * n + k] = m;
            for (j = k + 1; j < n; j++) lu[(size_t)i * n + j] -= m * lu[(size_t)k * n + j];
        }
    }
    /* Column col of the inverse solves A x = e_col: forward substitution on
     * P e_col, then back substitution. */
    for (col = 0; col < n; col++) {
        for (i = 0; i < n; i++) {
            double s = (perm[i] == col) ? 1.0 : 0.0;
            for (j = 0; j < i; j++) s -= lu[(size_t)i * n + j] * y[j];
            y[i] = s;
        }
        for (i = n - 1; i >= 0; i--) {
            double s = y[i];
            for (j = i + 1; j < n; j++) s -= lu[(size_t)i * n + j] * inverse[(size_t)j * n + col];
            inverse[(size_t)i * n + col] = s / lu[(size_t)i * n + i];
        }
    }
done:
    free(lu);
    free(y);
    free(perm);
    return rc;
}

int mw_errors(int n, const double* hessian, int number_stars, double* sigma, double* stddev)
{
    int i, j;
    if (mw_matrix_invert(n, hessian, sigma)) return 1;
    for (i = 0; i < n; i++)
        for (j = 0; j < n; j++) sigma[(size_t)i * n + j] /= (double)number_stars;
    if (stddev)
        for (i = 0; i < n; i++) stddev[i] = sqrt(fabs(sigma[(size_t)i * n + i]));
    return 0;
}
