// mw_search.hip -- GPU-backed parts of the search glue (declared in
// mw_search.h): the per-rank plugin functions integral_f and likelihood_f on
// the drop-in evaluator, and the batched Hessian of a resident context.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mw_context.h"
#include "mw_search.h"

using namespace mw;

namespace {

struct HessianEval {
    mw_ctx* ctx;
    int precision;
};

// Largest number of parameter vectors per launch: bounds the per-launch
// scratch (per-block partials grow with the vector count).
constexpr int kHessianChunk = 256;

int eval_vectors(void* user, const double* vectors, int count, double* values)
{
    const HessianEval* h = static_cast<const HessianEval*>(user);
    const int n = h->ctx->n_params;
    std::vector<mw_result> res(kHessianChunk);
    for (int first = 0; first < count; first += kHessianChunk) {
        const int m = std::min(kHessianChunk, count - first);
        int rc = mw_ctx_eval(h->ctx, vectors + (size_t)first * n, m, h->precision, res.data(), nullptr);
        if (rc) return rc;
        // The client's likelihood: prob_sum / number_stars - 3.0
        // (src/milkyway_client_app.c:276-278, astronomy_worker.c:115-134).
        for (int i = 0; i < m; i++) values[first + i] = res[i].likelihood - 3.0;
    }
    return MW_OK;
}

}  // namespace

extern "C" {

void mw_hip_integral_f(double* parameters, double* results, ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es,
                       STAR_POINTS* sp)
{
    apply_vector(*ap, parameters);
    reset_state(es);
    const int rc = mw_hip_calculate_integrals(ap, es, sp);
    if (rc) {
        std::fprintf(stderr, "APP: error calculating integrals: %d (%s)\n", rc, mw_last_error());
        std::exit(rc);
    }
    results[0] = es->background_integral;
    for (int s = 0; s < ap->number_streams; s++) results[s + 1] = es->stream_integrals[s];
}

void mw_hip_likelihood_f(double* integrals, double* results, ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es,
                         STAR_POINTS* sp)
{
    es->background_integral = integrals[0];
    for (int s = 0; s < ap->number_streams; s++) es->stream_integrals[s] = integrals[s + 1];
    const int rc = mw_hip_calculate_likelihood(ap, es, sp);
    if (rc) {
        std::fprintf(stderr, "APP: error calculating likelihood: %d (%s)\n", rc, mw_last_error());
        std::exit(rc);
    }
    results[0] = es->prob_sum;
    results[1] = es->bad_jacobians;
}

int mw_ctx_hessian(mw_ctx* ctx, const double* point, const double* step, int precision, double* hessian,
                   double* gradient, double* value)
{
    if (!ctx || !point || !step || !hessian) return fail(MW_E_INVALID, "null argument");
    if (ctx->n_params < 1) return fail(MW_E_INVALID, "the model has no optimised parameters");
    if (ctx->n_stars < 1) return fail(MW_E_INVALID, "the likelihood needs stars");
    if (precision != MW_PREC_FP64 && precision != MW_PREC_MIXED) return fail(MW_E_INVALID, "unknown precision");
    for (int i = 0; i < ctx->n_params; i++)
        if (!(step[i] != 0.0)) return fail(MW_E_INVALID, "every step must be nonzero");
    HessianEval h{ ctx, precision };
    fail(MW_OK, "");
    const int rc = mw_hessian(ctx->n_params, point, step, eval_vectors, &h, hessian, gradient, value);
    // Failures of the evaluations carry their own message; mw_hessian's own
    // failure can only be an allocation.
    if (rc && mw_last_error()[0] == '\0') return fail(MW_E_NOMEM, "Hessian stencil allocation failed");
    return rc;
}

}  // extern "C"
