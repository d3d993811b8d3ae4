// mw_separation.hip -- stream separation (milkyway_separation/cpu/separation.c:37-254)
// on a resident wedge context.
//
// One kernel launch computes, for every star, the membership probabilities,
// the drand48-driven assignment and the output position; the host then sums
// the probabilities per stream in star order (the reference's running sum)
// and counts the assignments.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "mw_context.h"

using namespace mw;

namespace {

// The parameter vector held in the template (get_search_parameters,
// parameters.c:365-394).
std::vector<double> template_vector(const mw_ctx* c)
{
    const ASTRONOMY_PARAMETERS& ap = c->tmpl.ap;
    std::vector<double> v;
    for (int i = 0; i < ap.number_background_parameters; i++)
        if (ap.background_optimize[i]) v.push_back(ap.background_parameters[i]);
    for (int s = 0; s < ap.number_streams; s++) {
        if (ap.stream_weight_optimize[s]) v.push_back(ap.stream_weights[s]);
        for (int j = 0; j < ap.number_stream_parameters; j++)
            if (ap.stream_optimize[s][j]) v.push_back(ap.stream_parameters[s][j]);
    }
    return v;
}

// srand48(seed): the state is the low 32 bits of the seed above 0x330E.
uint64_t drand48_seed_state(unsigned long seed)
{
    return ((uint64_t)(seed & 0xffffffffUL) << 16) | 0x330Eu;
}

}  // namespace

extern "C" {

int mw_ctx_separation(mw_ctx* c, const double* params, int precision, const double* integrals,
                      unsigned long seed, int* assignment, double* xyz, double* probability,
                      mw_separation_summary* summary)
{
    if (!c) return fail(MW_E_INVALID, "null context");
    if (precision != MW_PREC_FP64 && precision != MW_PREC_MIXED) return fail(MW_E_INVALID, "unknown precision");
    const int S = c->S;
    // prob_ok handles 1 to 4 streams (probability.c:317-396).
    if (S < 1 || S > kMaxStreams) return fail(MW_E_INVALID, "separation needs 1 to 4 streams");
    MW_HIP_CHECK(hipSetDevice(c->device));

    std::vector<double> vec;
    if (c->n_params > 0) {
        if (params)
            vec.assign(params, params + c->n_params);
        else
            vec = template_vector(c);
    }
    ModelTemplate m;
    m.load(c->tmpl.ap);
    if (c->n_params > 0) apply_vector(m.ap, vec.data());
    EvalConst K;
    std::string err;
    if (!make_eval_const(m.ap, &K, &err)) return fail(MW_E_INVALID, err);

    SepConst C;
    std::memset(&C, 0, sizeof(C));
    if (integrals) {
        C.I_bg = integrals[0];
        for (int s = 0; s < S; s++) C.I_st[s] = integrals[1 + s];
    } else {
        mw_result r;
        int rc = mw_ctx_eval(c, vec.empty() ? nullptr : vec.data(), 1, precision, &r, nullptr);
        if (rc) return rc;
        C.I_bg = r.background_integral;
        for (int s = 0; s < S; s++) C.I_st[s] = r.stream_integrals[s];
    }
    // Mixture weights (separation.c:127-140): the background's weight is
    // exp(0) here, whatever background_weight holds.
    double denom = 1.0;
    for (int s = 0; s < S; s++) denom += std::exp(m.ap.stream_weights[s]);
    for (int s = 0; s < S; s++) C.eps_s[s] = std::exp(m.ap.stream_weights[s]) / denom;
    C.eps_b = 1.0 / denom;
    double normal[3];
    if (!separation_frame(c->wedge, c->sgr, normal, C.rot))
        return fail(MW_E_INVALID, "sgr_coordinates must be 0 or 1");
    C.x0 = drand48_seed_state(seed);

    const long N = c->n_stars;
    std::vector<int> h_assign;
    std::vector<double> h_prob;
    if (N > 0) {
        int rc;
        if ((rc = c->sep_assign.reserve(N)) || (rc = c->sep_xyz.reserve(3 * N)) ||
            (rc = c->sep_prob.reserve((size_t)S * N)))
            return rc;
        StarLaunch it;
        std::memset(&it, 0, sizeof(it));
        it.chain_flags = 3;
        it.star = c->star.p;
        it.n_stars = N;
        it.n_conv = c->n_conv;
        it.n_streams = S;
        it.ratio = c->ratio.p;
        it.kw = c->kw.p;
        it.kw3 = c->kw3.p;
        it.coarse_pr = c->star_coarse_pr;
        const bool mixed = precision == MW_PREC_MIXED;
        if ((rc = c->ensure_stream())) return rc;
        MW_HIP_CHECK(launch_separation(S, K.bg_mode, mixed, it, K, C, c->sep_assign.p, c->sep_xyz.p,
                                       c->sep_prob.p, c->stream));
        int* a_out = assignment;
        if (!a_out) {
            h_assign.resize(N);
            a_out = h_assign.data();
        }
        double* p_out = probability;
        if (!p_out) {
            h_prob.resize((size_t)S * N);
            p_out = h_prob.data();
        }
        MW_HIP_CHECK(hipMemcpyAsync(a_out, c->sep_assign.p, sizeof(int) * N, hipMemcpyDeviceToHost, c->stream));
        MW_HIP_CHECK(hipMemcpyAsync(p_out, c->sep_prob.p, sizeof(double) * S * N, hipMemcpyDeviceToHost,
                                    c->stream));
        if (xyz)
            MW_HIP_CHECK(hipMemcpyAsync(xyz, c->sep_xyz.p, sizeof(double) * 3 * N, hipMemcpyDeviceToHost,
                                        c->stream));
        MW_HIP_CHECK(hipStreamSynchronize(c->stream));
        assignment = a_out;
        probability = p_out;
    }
    if (summary) {
        mw_separation_summary& o = *summary;
        std::memset(&o, 0, sizeof(o));
        o.number_streams = S;
        o.number_stars = (int)N;
        o.epsilon_b = C.eps_b;
        o.background_integral = C.I_bg;
        for (int s = 0; s < S; s++) {
            o.epsilon_s[s] = C.eps_s[s];
            o.stream_integrals[s] = C.I_st[s];
        }
        // nstars[j] += sprob[j] per star, in star order (separation.c:193-196).
        for (long i = 0; i < N; i++) {
            for (int s = 0; s < S; s++) o.expected[s] += probability[i * S + s];
            const int a = assignment[i];
            if (a >= 1 && a <= S) o.separated[a - 1]++;
        }
    }
    return MW_OK;
}

int mw_hip_separation(const char* filename, double background_integral, const double* stream_integrals,
                      const ASTRONOMY_PARAMETERS* ap, const STAR_POINTS* sp, unsigned long seed)
{
    if (!filename || !ap || !sp || (ap->number_streams > 0 && !stream_integrals))
        return fail(MW_E_INVALID, "null argument");
    mw_ctx* c = nullptr;
    int rc = mw_ctx_create(ap, sp, 0, &c);
    if (rc) return rc;
    std::unique_ptr<mw_ctx> own(c);
    const int S = ap->number_streams;
    const long N = sp->number_stars;
    std::vector<double> integrals(1 + (S > 0 ? S : 0));
    integrals[0] = background_integral;
    for (int s = 0; s < S; s++) integrals[1 + s] = stream_integrals[s];
    std::vector<int> assign(N > 0 ? N : 1);
    std::vector<double> xyz(3 * (N > 0 ? N : 1));
    mw_separation_summary sum;
    if ((rc = mw_ctx_separation(c, nullptr, dropin_precision_value(), integrals.data(), seed, assign.data(),
                                xyz.data(), nullptr, &sum)))
        return rc;

    std::FILE* f = std::fopen(filename, "w");
    if (!f) return fail(MW_E_INVALID, std::string("cannot open ") + filename);
    for (long i = 0; i < N; i++)
        std::fprintf(f, "%d %lf %lf %lf\n", assign[i], xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    if (std::fclose(f) != 0) return fail(MW_E_INVALID, std::string("cannot write ") + filename);

    // The reference's closing report (separation.c:119-140, :239-250).
    std::printf("bint: %lf", background_integral);
    for (int s = 0; s < S; s++) std::printf(", sint[%d]: %lf", s, stream_integrals[s]);
    std::printf("\n");
    for (int s = 0; s < S; s++) std::printf("epsilon_s[%d]: %lf\n", s, sum.epsilon_s[s]);
    std::printf("epsilon_b:    %lf\n", sum.epsilon_b);
    std::printf("%ld total stars\n", N);
    for (int s = 0; s < S; s++)
        std::printf("%lf in stream[%d] (%lf%%)\n", sum.expected[s], s, sum.expected[s] / (double)N * 100);
    for (int s = 0; s < S; s++) std::printf("%d stars separated into stream\n", sum.separated[s]);
    std::printf("Output written to: %s\n", filename);
    std::fflush(stdout);
    return MW_OK;
}

}  // extern "C"
