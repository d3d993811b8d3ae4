// mw_noalloc.hip -- mw_ctx_eval_noalloc: one evaluation on a resident
// context that allocates nothing.  For callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).
//
// Everything mw_ctx_eval keeps in buffers that grow on demand -- the
// evaluation constants, the work items of the integral and star passes,
// the per-block partial sums and the reduced rows -- lives here in one
// file-scope device array, `scratch`, of kScratchBytes.  Each GPU holds its
// own copy until the process ends; host code reaches it only through the
// address hipGetSymbolAddress returns for it on that device, taken once per
// device (with the device current) and kept in scratch_ptr[].  Every call on
// a GPU shares that GPU's array and overwrites what it needs (nothing in it
// has to start at zero: every partial and row is written before it is read),
// so calls that overlap on one GPU race; the caller serialises them.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "mw_context.h"

using namespace mw;

namespace {

constexpr size_t kScratchBytes = MW_NOALLOC_SCRATCH_BYTES;   // per GPU
constexpr int kMaxDevices = 64;
constexpr int kMaxRegions = 16;      // integration regions per context
constexpr int kMaxBgParams = 16;     // background parameters
constexpr int kMaxStreamParams = 16; // parameters per stream

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
unsigned char* scratch_ptr[kMaxDevices] = {};

// The device address of this device's scratch array (taken on first use).
int scratch_for(int device, unsigned char** out)
{
    if (device < 0 || device >= kMaxDevices) return fail(MW_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MW_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        scratch_ptr[device] = static_cast<unsigned char*>(p);
    }
    *out = scratch_ptr[device];
    return MW_OK;
}

// Bump allocation of 256-byte-aligned sections inside the scratch array.
struct Carve {
    unsigned char* base;
    size_t used = 0;
    template <typename T>
    T* take(size_t n)
    {
        used = (used + 255) & ~size_t(255);
        T* p = reinterpret_cast<T*>(base + used);
        used += sizeof(T) * n;
        return p;
    }
};

}  // namespace

extern "C" int mw_ctx_eval_noalloc(mw_ctx* c, const double* params, int precision, void* stream,
                                   mw_result* result)
{
    if (!c || !result) return fail(MW_E_INVALID, "null argument");
    if (c->n_params > 0 && !params) return fail(MW_E_INVALID, "null parameter vector");
    if (precision != MW_PREC_FP64 && precision != MW_PREC_MIXED)
        return fail(MW_E_INVALID, "mw_ctx_eval_noalloc takes MW_PREC_FP64 or MW_PREC_MIXED");
    const ASTRONOMY_PARAMETERS& t = c->tmpl.ap;
    const int S = c->S, G = (int)c->regions.size();
    if (S > kMaxStreams) return fail(MW_E_INVALID, "mw_ctx_eval_noalloc takes models of at most 4 streams");
    if (G < 1 || G > kMaxRegions) return fail(MW_E_INVALID, "mw_ctx_eval_noalloc takes 1 to 16 regions");
    if (t.number_background_parameters > kMaxBgParams || t.number_stream_parameters > kMaxStreamParams)
        return fail(MW_E_INVALID, "too many model parameters for mw_ctx_eval_noalloc");

    // The parameter vector applied to a copy of the template's values, held
    // on the stack (set_astronomy_parameters, parameters.c:332-363).
    ASTRONOMY_PARAMETERS ap = t;
    double bg[kMaxBgParams], sw[kMaxStreams], sp[kMaxStreams][kMaxStreamParams];
    double* sp_rows[kMaxStreams];
    std::memcpy(bg, t.background_parameters, sizeof(double) * t.number_background_parameters);
    for (int s = 0; s < S; s++) {
        sw[s] = t.stream_weights[s];
        std::memcpy(sp[s], t.stream_parameters[s], sizeof(double) * t.number_stream_parameters);
        sp_rows[s] = sp[s];
    }
    ap.background_parameters = bg;
    ap.stream_weights = sw;
    ap.stream_parameters = sp_rows;
    if (c->n_params > 0) apply_vector(ap, params);
    EvalConst K;
    {
        std::string err;
        if (!make_eval_const(ap, &K, &err)) return fail(MW_E_INVALID, err);
    }
    const bool mixed = precision == MW_PREC_MIXED;
    const int NI = integral_comps(), NS = star_comps();

    // Work items (exact tables, every region whole, one launch).
    RegionLaunch items[kMaxRegions];
    ReduceItem ired[kMaxRegions];
    int iblocks = 0;
    for (int g = 0; g < G; g++) {
        const RegionDev& rd = *c->regions[g];
        RegionLaunch& L = items[g];
        std::memset(&L, 0, sizeof(L));
        L.rq = rd.rq.p;
        L.rstep = rd.rstep.p;
        L.col = rd.col.p;
        L.n_cols = rd.n_cols;
        L.r_steps = rd.geom.r_steps;
        L.n_conv = rd.n_conv;
        L.calc_begin = 0;
        L.calc_end = rd.n_cols * (long)rd.geom.r_steps;
        L.slot = 0;
        L.out_index = g;
        L.block_begin = iblocks;
        L.r_chunks = rd.r_chunks;
        L.block_count = rd.blocks * rd.r_chunks;
        L.n_streams = S;
        L.r_lo = rd.r_min;
        L.r_hi = rd.r_max;
        ired[g] = ReduceItem{ iblocks, L.block_count, g, 0 };
        iblocks += L.block_count;
    }
    const long N = c->n_stars;
    const int sblocks = N > 0 ? (int)((N + kBlock - 1) / kBlock) : 0;
    StarLaunch sitem;
    std::memset(&sitem, 0, sizeof(sitem));
    sitem.star = c->star.p;
    sitem.n_stars = N;
    sitem.n_conv = c->n_conv;
    sitem.block_count = sblocks;
    sitem.n_streams = S;
    sitem.ratio = c->ratio.p;
    sitem.kw = c->kw.p;
    sitem.kw3 = c->kw3.p;
    sitem.coarse_pr = c->star_coarse_pr;
    sitem.chain_flags = 3;
    const ReduceItem sred{ 0, sblocks, 0, 0 };
    const int first = 0, count = G;

    // Scratch layout; sizes beyond the array are refused.
    int dev = 0;
    MW_HIP_CHECK(hipGetDevice(&dev));
    if (dev != c->device) return fail(MW_E_INVALID, "the context lives on another device");
    unsigned char* base = nullptr;
    int rc = scratch_for(dev, &base);
    if (rc) return rc;
    Carve cv{ base };
    EvalConst* d_k = cv.take<EvalConst>(1);
    RegionLaunch* d_items = cv.take<RegionLaunch>(G);
    ReduceItem* d_ired = cv.take<ReduceItem>(G);
    StarLaunch* d_sitem = cv.take<StarLaunch>(1);
    ReduceItem* d_sred = cv.take<ReduceItem>(1);
    int* d_first = cv.take<int>(2);
    double* d_ipart = cv.take<double>((size_t)iblocks * NI * 2);
    double* d_irows = cv.take<double>((size_t)G * NI);
    double* d_spart = cv.take<double>((size_t)std::max(sblocks, 1) * NS * 2);
    double* d_srows = cv.take<double>(NS);
    if (cv.used > kScratchBytes)
        return fail(MW_E_INVALID, "the evaluation needs " + std::to_string(cv.used) +
                                      " bytes of scratch; mw_ctx_eval_noalloc holds " +
                                      std::to_string(kScratchBytes));

    hipStream_t st = static_cast<hipStream_t>(stream);
    const int two[2] = { first, count };
    MW_HIP_CHECK(hipMemcpyAsync(d_k, &K, sizeof(K), hipMemcpyHostToDevice, st));
    MW_HIP_CHECK(hipMemcpyAsync(d_items, items, sizeof(RegionLaunch) * G, hipMemcpyHostToDevice, st));
    MW_HIP_CHECK(hipMemcpyAsync(d_ired, ired, sizeof(ReduceItem) * G, hipMemcpyHostToDevice, st));
    MW_HIP_CHECK(hipMemcpyAsync(d_first, two, sizeof(two), hipMemcpyHostToDevice, st));
    MW_HIP_CHECK(launch_integral(S, K.bg_mode, mixed, d_items, G, iblocks, d_k, d_ipart, st, false));
    MW_HIP_CHECK(launch_reduce(false, d_ired, G, d_ipart, d_irows, st));
    if (sblocks > 0) {
        MW_HIP_CHECK(hipMemcpyAsync(d_sitem, &sitem, sizeof(sitem), hipMemcpyHostToDevice, st));
        MW_HIP_CHECK(hipMemcpyAsync(d_sred, &sred, sizeof(sred), hipMemcpyHostToDevice, st));
        MW_HIP_CHECK(launch_finalize(d_first, d_first + 1, d_irows, d_k, 1, st));
        MW_HIP_CHECK(launch_star(S, K.bg_mode, mixed, d_sitem, 1, sblocks, d_k, c->ratio.p, c->kw.p,
                                 d_spart, nullptr, 0, st));
        MW_HIP_CHECK(launch_reduce(true, d_sred, 1, d_spart, d_srows, st));
    }
    double irows[kMaxRegions * (1 + kMaxStreams)], srows[2 + kMaxStreams + 1] = {};
    MW_HIP_CHECK(hipMemcpyAsync(irows, d_irows, sizeof(double) * G * NI, hipMemcpyDeviceToHost, st));
    if (sblocks > 0) MW_HIP_CHECK(hipMemcpyAsync(srows, d_srows, sizeof(double) * NS, hipMemcpyDeviceToHost, st));
    MW_HIP_CHECK(hipStreamSynchronize(st));

    // The result as mw_ctx_eval assembles it.
    mw_result& r = *result;
    std::memset(&r, 0, sizeof(r));
    r.number_streams = S;
    r.background_integral = irows[0];
    for (int s = 0; s < S; s++) r.stream_integrals[s] = irows[1 + s];
    for (int g = 1; g < G; g++) {
        r.background_integral -= irows[g * NI];
        for (int s = 0; s < S; s++) r.stream_integrals[s] -= irows[g * NI + 1 + s];
    }
    if (N > 0) {
        r.num_zero = (int)srows[NS - 1];
        r.prob_sum = srows[0] - 238.0 * r.num_zero;
        r.likelihood = r.prob_sum / (double)N;
        r.bg_only_likelihood = srows[1] / (double)N;
        for (int s = 0; s < S; s++) r.stream_only_likelihood[s] = srows[2 + s] / (double)N;
    } else {
        r.likelihood = NAN;
    }
    return MW_OK;
}
