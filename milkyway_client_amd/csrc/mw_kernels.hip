// mw_kernels.hip -- gfx950 kernels of the separation likelihood.
//
//   integral_kernel   the volume integral over the (mu, nu, r) wedge grid
//                     (reference: evaluation_optimized.c:424-562 around the
//                     integrand :190-299).  One work-item per (mu, nu) column
//                     streams over every r-step of its region; the r-step
//                     tables (r_point, qw_r3_N) are staged through LDS in
//                     r-tiles and read as wavefront-wide broadcasts; each
//                     column keeps Kahan sums in registers; a block reduces its
//                     256 columns with double-double (TwoSum) arithmetic over
//                     DPP/permute shuffles and LDS.
//   star_kernel       the star pass (evaluation_optimized.c:613-753): one
//                     work-item per star, SoA star records, per-star log10 P.
//   reduce_kernel     second pass: each work item's per-block partials summed
//                     in block order (double-double), so the result does not
//                     depend on scheduling.
//   finalize_kernel   I = I_main - sum I_cut per evaluation slot, written into
//                     the slot's constants for the star pass.
//
// Precision modes: MIXED == false is FP64 throughout (the Gaussian as a
// polynomial 2^x, 1/sqrt and 1/x from hardware estimates refined to full FP64
// accuracy).  MIXED == true evaluates the streams' exp in FP32 on the
// transcendental unit (v_exp_f32; in the integral pass of 2^-t^2 with the
// sight's minimum exponent applied per column in FP64, in the star pass of
// the fractional part of the exponent) and keeps FP64 for the positions
// (lbr -> xyz), the squared stream-axis distance, the background (its 1/sqrt
// and 1/x from the FP64 estimates with one Newton step: as cheap as FP32
// estimates with their conversions, and exact to FP64), the per-star log and
// every accumulation but one: in rows of at least kGroupMinConv convolution
// points, the integral pass sums each stream's terms over groups of 8 points
// in FP32 before adding them in FP64 (stream_group).
#include <hip/hip_runtime.h>

#include "mw_internal.h"

namespace mw {

constexpr int kTileD2 = 1024;           // LDS r-tile: 1024 (r_point, qw) pairs = 16 KiB
constexpr int kIntComps = 1 + kMaxStreams;
constexpr int kStarComps = 2 + kMaxStreams + 1;   // prob, bg-only, st-only[4], zero count

// ---------------------------------------------------------------------------
// double-double helpers (Knuth TwoSum, Dekker FastTwoSum)
// ---------------------------------------------------------------------------
struct DD {
    double hi, lo;
};

__device__ __forceinline__ DD dd_add(DD a, DD b)
{
    const double s = a.hi + b.hi;
    const double bb = s - a.hi;
    double e = (a.hi - (s - bb)) + (b.hi - bb);
    e += a.lo + b.lo;
    const double h = s + e;
    return DD{ h, e - (h - s) };
}

__device__ __forceinline__ DD dd_shfl_xor(DD v, int mask)
{
    return DD{ __shfl_xor(v.hi, mask, 64), __shfl_xor(v.lo, mask, 64) };
}

// Sum one DD value per work-item over the workgroup; result valid in thread 0.
template <int NC>
__device__ __forceinline__ void block_reduce_dd(DD (&v)[NC], DD (*lds)[NC])
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
#pragma unroll
        for (int c = 0; c < NC; c++) v[c] = dd_add(v[c], dd_shfl_xor(v[c], m));
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int c = 0; c < NC; c++) lds[wave][c] = v[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NC; c++) {
            DD a = lds[0][c];
            for (int w = 1; w < kBlock / 64; w++) a = dd_add(a, lds[w][c]);
            v[c] = a;
        }
    }
}

// Sum one double per work-item over the workgroup in a fixed order (an xor
// butterfly within each wave, whose partners add the same two values, then
// the waves in order); result valid in thread 0.  For the resampled
// kernels' block partials, positive terms only (a block's 256 column sums
// of a few hundred positive grid terms): every term passes through 6
// butterfly additions and at most 3 wave additions, so over positive terms
// the result errs by at most 9 u of the block's sum (u = 2^-53, 1e-15
// relative; tests/test_block_reduction.py); the blocks are then added in
// double-double by
// reduce_kernel.  Double-double here cost 14% of the resampled launch's
// instructions (its blocks are short); the exact-quadrature kernels and the
// star pass, whose blocks are long or whose terms change sign, reduce in
// double-double.
template <int NC>
__device__ __forceinline__ void block_reduce_f64(double (&v)[NC], double (*lds)[NC])
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
#pragma unroll
        for (int c = 0; c < NC; c++) v[c] += __shfl_xor(v[c], m, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int c = 0; c < NC; c++) lds[wave][c] = v[c];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NC; c++) {
            double a = lds[0][c];
            for (int w = 1; w < kBlock / 64; w++) a += lds[w][c];
            v[c] = a;
        }
    }
}

// The resampled integral's block result: components 0..S summed as above
// into prow's (hi, lo) pairs (lo 0), the unused stream components 0.
template <int S>
__device__ __forceinline__ void write_block_f64(double (&out)[1 + S], double* __restrict__ prow, void* red)
{
    block_reduce_f64<1 + S>(out, reinterpret_cast<double(*)[1 + S]>(red));
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < 2 * kIntComps; c++) prow[c] = 0.0;
#pragma unroll
        for (int c = 0; c <= S; c++) prow[2 * c] = out[c];
    }
}

__device__ __forceinline__ void kahan_add(double& s, double& c, double x)
{
    const double t = s;
    s += x;
    c += x - (s - t);
}

// Work item owning block b: the last item whose block_begin <= b.
template <typename T>
__device__ __forceinline__ int find_item(const T* __restrict__ items, int n_items, int b)
{
    int lo = 0, hi = n_items - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].block_begin <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// FP64 elementary functions used by the integrand.
// ---------------------------------------------------------------------------

// 2^-a for a >= 0 (a is the Gaussian tube exponent in base 2).  t = -a is
// split into k = rint(t) and f = t - k in [-1/2, 1/2] (exact in FP64); 2^f
// is a degree-11 polynomial interpolating 2^f at the Chebyshev nodes of
// [-1/2, 1/2] (interpolation error below 5e-18 relative, about 1 ulp after
// rounding); 2^k by ldexp.  Arguments beyond 1100 are clamped: 2^-1100
// underflows to 0, as libm's exp does for every argument below -745.2.
__device__ __forceinline__ double exp2_neg(double a)
{
    const double t = fmax(-a, -1100.0);
    const double k = rint(t);
    const double f = t - k;
    double p = 4.4558179083360645e-10;
    p = fma(p, f, 7.074194297288521e-09);
    p = fma(p, f, 1.0178057087733941e-07);
    p = fma(p, f, 1.3215432535912375e-06);
    p = fma(p, f, 1.5252733841556773e-05);
    p = fma(p, f, 0.00015403530463724353);
    p = fma(p, f, 0.001333355814640647);
    p = fma(p, f, 0.009618129107587256);
    p = fma(p, f, 0.055504108664821625);
    p = fma(p, f, 0.24022650695910158);
    p = fma(p, f, 0.6931471805599453);
    p = fma(p, f, 1.0);
    return ldexp(p, (int)k);
}

// 1/sqrt(a), a > 0: the hardware estimate (measured on gfx950: relative error
// below 5.3e-8 over [1e-3, 1e4]) refined by one Newton step
// y <- y + y (1 - a y^2) / 2, which squares the error (measured: 4.1e-15).
__device__ __forceinline__ double rsqrt_nr(double a)
{
    const double y = __builtin_amdgcn_rsq(a);
    const double e = fma(-(a * y), y, 1.0);
    return fma(0.5 * y, e, y);
}

// a b + c with c uniform in an SGPR pair, as one VOP3 v_fma_f64 (the
// compiler's own choice, the two-address v_fmac_f64, first copies c into a
// VGPR pair with two moves).  c must be the same on every lane of the wave
// (here: table entries read by scalar loads); the "s" operand holds one value
// per wave.
__device__ __forceinline__ double fma_sgpr(double a, double b, double c)
{
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
}

// a b + c with b and c uniform in SGPR pairs (the same on every lane, as
// above): one instruction reads at most one SGPR pair, so a multiply and an
// add.
__device__ __forceinline__ double mul_add_sgpr(double a, double b, double c)
{
    double r;
    asm("v_mul_f64 %0, %1, %2\n\tv_add_f64 %0, %0, %3" : "=&v"(r) : "v"(a), "s"(b), "s"(c));
    return r;
}

// 1/d, d > 0: the hardware estimate (measured: below 4.6e-8) refined by one
// Newton step y <- y + y (1 - d y) (measured: 2.1e-15).
__device__ __forceinline__ double rcp_nr(double d)
{
    const double y = __builtin_amdgcn_rcp(d);
    return fma(y, fma(-d, y, 1.0), y);
}

// ---------------------------------------------------------------------------
// One line of sight.  Along the sight a point at distance r sits at
// r u - (8.5, 0, 0), and both squared distances the integrand needs are
// quadratics in r, written per sight in the shifted form (p r - m)^2 + e:
//
//  * background: rg^2 = (r u_x - 8.5)^2 + (r u_y)^2 + (r u_z / q)^2 with
//    P = u_x^2 + u_y^2 + u_z^2 / q^2, p = sqrt(P), m = 8.5 u_x / p,
//    e = 8.5^2 (u_y^2 + u_z^2 / q^2) / P (the minimum, with no cancellation);
//  * stream s: the offset from the axis is r u_perp - c_perp, where
//    x_perp = x - (x . a) a and c' = c + (8.5, 0, 0)
//    (evaluation_optimized.c:283-293 computes the same vector as v - (v . a) a).
//    Both are scaled by k = sqrt(log2(e) / (2 sigma^2)), so that
//    |r u_perp k - c_perp k|^2 is the Gaussian exponent in base 2; with
//    p = |u_perp k|, m = (u_perp k / p) . c_perp k and e the squared distance
//    of c_perp k from the line through u_perp k (again a norm, not a
//    difference of squares).
//
// Per convolution point each quadratic is then t = p r - m, t^2 + e: two
// FMAs.  Near the minimum t is small and e carries the value, so the form
// loses nothing to cancellation.
// ---------------------------------------------------------------------------
struct Quad {
    double p, m, e;
};

template <int S>
struct Sight {
    Quad bg;
    Quad st[S > 0 ? S : 1];
};

// FAST (mixed kernels): the square roots and quotients from one
// Newton-refined reciprocal square root per quadratic (rsqrt_nr, 4e-15
// relative) instead of an IEEE square root and divisions; the FP64 kernels
// keep the correctly rounded operations.
template <int S, int BG, bool FAST = false>
__device__ __forceinline__ Sight<S> make_sight(double ux, double uy, double uz, const EvalConst& K)
{
    Sight<S> v;
    if constexpr (BG != kBgZeroQ) {
        const double side = fma(uy, uy, uz * uz * K.inv_q2);
        const double P = fma(ux, ux, side);
        if constexpr (FAST) {
            const double ip = rsqrt_nr(P);
            v.bg = Quad{ P * ip, kSunR0 * ux * ip, (kSunR0 * kSunR0) * side * (ip * ip) };
        } else {
            const double p = sqrt(P);
            v.bg = Quad{ p, kSunR0 * ux / p, (kSunR0 * kSunR0) * side / P };
        }
    } else {
        v.bg = Quad{ 0.0, 0.0, 0.0 };
    }
#pragma unroll
    for (int s = 0; s < S; s++) {
        const double d = ux * K.sa[s][0] + uy * K.sa[s][1] + uz * K.sa[s][2];
        const double w0 = (ux - d * K.sa[s][0]) * K.ks[s];
        const double w1 = (uy - d * K.sa[s][1]) * K.ks[s];
        const double w2 = (uz - d * K.sa[s][2]) * K.ks[s];
        const double ww = fma(w0, w0, fma(w1, w1, w2 * w2));
        const double c0 = K.cperp[s][0], c1 = K.cperp[s][1], c2 = K.cperp[s][2];
        if (ww > 1e-200) {
            double p, h0, h1, h2;
            if constexpr (FAST) {
                const double ip = rsqrt_nr(ww);
                p = ww * ip;
                h0 = w0 * ip, h1 = w1 * ip, h2 = w2 * ip;
            } else {
                p = sqrt(ww);
                h0 = w0 / p, h1 = w1 / p, h2 = w2 / p;
            }
            const double m = fma(h0, c0, fma(h1, c1, h2 * c2));
            const double e0 = fma(-m, h0, c0), e1 = fma(-m, h1, c1), e2 = fma(-m, h2, c2);
            v.st[s] = Quad{ p, m, fma(e0, e0, fma(e1, e1, e2 * e2)) };
        } else {
            // sight parallel to the stream axis: constant distance
            v.st[s] = Quad{ 0.0, 0.0, fma(c0, c0, fma(c1, c1, c2 * c2)) };
        }
    }
    return v;
}

__device__ __forceinline__ double quad_at(const Quad& q, double r)
{
    const double t = fma(q.p, r, -q.m);
    return fma(t, t, q.e);
}

// ---------------------------------------------------------------------------
// The integrand at one convolution point (distance r along the sight, weight
// qw = W_i r^3 N(dx_i)), accumulated into bg and st[]
// (evaluation_optimized.c:222-296 per point).
// ---------------------------------------------------------------------------
//
// SHIFTED (mixed precision, integral pass): each stream's exponent is taken
// without the sight's minimum e, a = t^2, and the caller multiplies the
// column's sums by 2^-e in FP64 once.  The FP32 exponent then rounds t^2
// rather than t^2 + e (relative error of the term |t^2| 2^-24 ln 2, unbiased
// over the points), and terms below 2^-126 of the sight's own peak flush to
// zero while the range of 2^-e stays FP64's.
template <int BG>
__device__ __forceinline__ void bg_point(double r, double qw, const Quad& q, const EvalConst& K, double& bg)
{
    if constexpr (BG != kBgZeroQ) {
        const double rg2 = quad_at(q, r);
        if constexpr (BG == kBgGeneral) {
            const double rg = sqrt(rg2);
            bg += qw / (pow(rg, K.alpha) * pow(rg + K.r0, K.alpha_delta3));
        } else {
            const double rinv = rsqrt_nr(rg2);
            const double irs = rcp_nr(fma(rg2, rinv, K.r0));
            bg = fma(qw * rinv, irs * irs * irs, bg);
        }
    }
}

// The Hernquist background term at a quarter of its value, one instruction
// cheaper than bg_point: the Newton step for 1/sqrt(rg^2) in the product form
// Y = y (3 - rg^2 y^2) gives 2 / rg (three instructions besides the estimate,
// against four for y + y (1 - rg^2 y^2) / 2), so rg^2 Y + 2 r0 = 2 (rg + r0),
// its reciprocal X = 1 / (2 (rg + r0)) and qw Y X^3 = qw / (4 rg (rg + r0)^3).
// Every factor is exact (powers of two), the rounding that of bg_point's
// (the product form adds under 1e-16 to the step's 4e-15).  The caller
// multiplies the sum by 4 once.
__device__ __forceinline__ void bg_point_quarter(double r, double qw, const Quad& q, double r0x2, double& bg)
{
    const double rg2 = quad_at(q, r);
    const double y = __builtin_amdgcn_rsq(rg2);
    const double Y = y * fma(-rg2, y * y, 3.0);
    const double X = rcp_nr(fma(rg2, Y, r0x2));
    bg = fma(qw * Y, X * X * X, bg);
}

// Mixed precision's background term: the Hernquist profile at a quarter of
// its value (bg_point_quarter), the other profiles as bg_point; a sum of them
// is restored by kMixedBgScale.
template <int BG>
__device__ __forceinline__ void bg_point_mixed(double r, double qw, const Quad& q, const EvalConst& K, double& bg)
{
    if constexpr (BG == kBgHernquist)
        bg_point_quarter(r, qw, q, 2.0 * K.r0, bg);
    else
        bg_point<BG>(r, qw, q, K, bg);
}
template <int BG>
constexpr double kMixedBgScale = BG == kBgHernquist ? 4.0 : 1.0;

// SHIFTED (mixed exact tables): stream s's exponent is t^2 - s2[s], its
// smallest value over the region's points taken out (the caller scales each
// sum by 2^-(e + s2)), so that the FP32 2^x of the terms that carry the sum
// stays far from underflow even where the sight's closest approach to the
// stream lies outside the region's distances.
template <int S, int BG, bool MIXED, bool SHIFTED = false>
__device__ __forceinline__ void point_terms(double r, double qw, const Sight<S>& v,
                                            const EvalConst& K, double& bg, double (&st)[kMaxStreams],
                                            const double* s2 = nullptr)
{
    if constexpr (MIXED && SHIFTED)
        bg_point_mixed<BG>(r, qw, v.bg, K, bg);   // the caller restores the sum by kMixedBgScale
    else
        bg_point<BG>(r, qw, v.bg, K, bg);
#pragma unroll
    for (int s = 0; s < S; s++) {
        if constexpr (MIXED && SHIFTED) {
            const double t = fma(v.st[s].p, r, -v.st[s].m);
            const float e = __builtin_amdgcn_exp2f(-(float)fma(t, t, -s2[s]));
            st[s] = fma(qw, (double)e, st[s]);
            continue;
        }
        const double a = quad_at(v.st[s], r);
        if constexpr (MIXED) {
            // 2^t with t = -a split in FP64 into n + f, f in [0, 1): only 2^f
            // goes through FP32, so the relative error stays at the FP32
            // exp2's ulp however large |t| is.
            const double t = fmax(-a, -1100.0);
            const double n = floor(t);
            const float e = __builtin_amdgcn_exp2f((float)(t - n));
            st[s] = fma(qw, ldexp((double)e, (int)n), st[s]);
        } else {
            st[s] = fma(qw, exp2_neg(a), st[s]);
        }
    }
}

// Rows of a grid uniform in distance (nodes r_j = r_0 + j h along the whole
// column), mixed precision, shifted stream exponents.  With t_j = t_0 + j d,
// d = p h, the terms G_j = 2^-(t_j^2) obey
//     G_{j+1} = G_j R_j,   R_{j+1} = R_j C,
// R_0 = 2^-(2 t_0 d + d^2) and C = 2^-(2 d^2) (one per sight): two FP64
// multiplies per stream and node instead of an exponential.  The state runs
// on from row to row and is restarted from the FP64 2^x
//   * every kRestartRows rows, which bounds the accumulated error by
//     (8 x 32)^2 / 2 ulp = 4e-12 relative, and
//   * at a row that starts before the Gaussian's peak (t_0 < 0) with
//     G < 2^-1000: such a G may have lost precision to underflow, and the
//     terms are about to grow.
// Past the peak the terms only shrink, so underflow there loses nothing.
// The terms are carried relative to the sight's largest on-grid term: with
// s^2 the smallest t^2 over the grid's nodes (0 when the Gaussian's peak
// lies on the grid), G_j = 2^-(t_j^2 - s^2) <= 1, and the sight's sum is
// scaled by 2^-(e + s^2) at the end.  Without the shift, a sight whose peak
// lies beyond the grid would start its on-grid terms below 2^-1000, in
// subnormal numbers or at 0.
// A restart can still fall on a node far below that largest term: a wave runs the union of its
// lanes' rows, so a lane may start rows before its own terms matter, and a row of 32 nodes spans up
// to 16 in t.  A G started in subnormal numbers carries its rounding (up to 100 %) to every later
// term, the peak's included, until the next restart.  So G and the sight's running sum are carried
// scaled by 2^B, B a whole number >= 0 set at each restart: B = ceil(t_0^2 - s^2 - 1000) when the
// start lies before the peak, short of the clamp below, and more than 2^-1000 below the largest
// term, else 0.  G then starts in [2^-1000, 2^-999), with full precision; the scaled G stays below
// 2^B <= 2^600 and the scaled sum below 2^610, and rescaling by a power of two is exact.  The sum
// is unscaled (2^-B) at the end.  The caller guarantees d <= 1/2.  On restart t_0 is clamped at
// -sqrt(s^2 + 1600), so R stays below 2^(sqrt(s^2 + 1600)), and a clamped start gives
// G = 2^-1600, that is 0, which stays 0 until the next restart: the terms of a row that starts there lie below
// 2^-((sqrt(s^2 + 1600) - 16)^2 - s^2), 2^-190 or less, of the largest.  A sight whose s^2 reaches
// 1100 has every on-grid term below 2^-1100 of its peak, so its final factor 2^-(e + s^2) is 0: its
// recurrence is not run (G = R = 0).
constexpr int kRestartRows = 8;
constexpr double kRestartClampT2 = 1600.0;
constexpr double kGoneT2 = 1100.0;
constexpr double kScaleT2 = 1000.0;

// Restart of the recurrence at a row's first node t0 (with spacing d), terms
// relative to 2^-(s^2) and scaled by 2^B: G = 2^(B - (t0^2 - s^2)),
// R = 2^-(2 t0 d + d^2), B and the running sum a rescaled as above, t0
// clamped as above; G = R = 0 for a sight whose s^2 >= kGoneT2.
__device__ __forceinline__ void restart_recurrence(double t0, double d, double s2, double& G, double& R,
                                                   int& B, double& a)
{
    if (s2 >= kGoneT2) {
        G = R = 0.0;
        return;
    }
    const double lim2 = s2 + kRestartClampT2;
    const double lim = lim2 * rsqrt_nr(lim2);
    const double tc = fmax(t0, -lim);
    const double D = fma(tc, tc, -s2);   // t0 is a grid node: D >= 0 up to rounding
    // Scaled only before the peak and short of the clamp: past the peak the
    // terms only shrink and a deep start is negligible as it stands; a
    // clamped start (D = 1600) keeps B = 0, so that its G = 2^-1600 is 0.
    const double over = ceil(D - kScaleT2);   // > 0 exactly when D > kScaleT2
    const int Bn = (t0 < 0.0 && t0 > -lim && over > 0.0) ? (int)over : 0;
    // both exponentials side by side, unconditionally (they share the
    // polynomial's constants)
    R = exp2_neg(fma(2.0 * tc, d, d * d));
    G = exp2_neg(D - (double)Bn);
    a = ldexp(a, Bn - B);
    B = Bn;
}

// A stream's terms in a row are skipped, by the whole wave, when every one of
// them on every lane lies below 2^-kNegligibleLog2 of that sight's largest
// term over the region's nodes (2^-(min t^2)): each dropped term is below
// 2^-60 of that term, so with under 10^3 terms per sight and weights within
// 10^4 of each other what is dropped stays below 1e-11 of the sum.
constexpr double kNegligibleLog2 = 60.0;

// Smallest t^2 over the nodes from t_first to t_last (t rises with r): 0
// when the Gaussian's peak lies between them.
__device__ __forceinline__ double min_t2(double t_first, double t_last)
{
    if (t_first <= 0.0 && t_last >= 0.0) return 0.0;
    const double t = fmin(fabs(t_first), fabs(t_last));
    return t * t;
}

template <int S, int BG, int Q>
__device__ __forceinline__ void row_by_recurrence(const double2* row, int n, double r0, double h,
                                                  bool restart, const Sight<S>& v,
                                                  const double (&C)[S > 0 ? S : 1],
                                                  const double (&s2)[S > 0 ? S : 1],
                                                  double (&G)[S > 0 ? S : 1],
                                                  double (&Rr)[S > 0 ? S : 1],
                                                  int (&B)[S > 0 ? S : 1], const EvalConst& K,
                                                  double (&acc)[S + 1])
{
    if constexpr (Q == 1) {
        // (r, weight) entries, one weight for the background and the streams
#pragma unroll 2
        for (int i = 0; i < n; i++) {
            const double2 e = row[i];
            bg_point_mixed<BG>(e.x, e.y, v.bg, K, acc[0]);
        }
    } else {
        // (stream weight, background weight) entries; node i sits at r0 + i h
        // and only every Q-th node carries a background weight
        double rb = r0;
        for (int i = 0; i < n; i += Q) {
            bg_point_mixed<BG>(rb, row[i].y, v.bg, K, acc[0]);
            rb += Q * h;
        }
    }
#pragma unroll
    for (int s = 0; s < S; s++) {
        const double p = v.st[s].p;
        const double d = p * h;
        const double t0 = fma(p, r0, -v.st[s].m);
        const double t1 = fma(d, (double)(n - 1), t0);
        if (__all(min_t2(t0, t1) > s2[s] + kNegligibleLog2)) {
            // Negligible here; a later row before the peak restarts from G = 0.
            G[s] = 0.0;
            continue;
        }
        if (restart || (t0 < 0.0 && G[s] < 0x1p-1000))
            restart_recurrence(t0, d, s2[s], G[s], Rr[s], B[s], acc[1 + s]);
        double g = G[s], rr = Rr[s], a = acc[1 + s];
        const double c = C[s];
#pragma unroll 4
        for (int i = 0; i < n; i++) {
            a = fma(Q == 1 ? row[i].y : row[i].x, g, a);
            g *= rr;
            rr *= c;
        }
        G[s] = g;
        Rr[s] = rr;
        acc[1 + s] = a;
    }
}

// One logical block of a resampled region on a distance grid (grid_step > 0;
// the region is evaluated whole, from tables in the recurrence layout): the
// streams by recurrence, the background on every Q-th node.  The resampled
// tables carry every weight in the node entries (each row's irv and reff
// are 1), so the column sums take the terms directly, plainly summed (a
// column has a few hundred nodes, not the exact quadrature's tens of
// thousands of points); each stream's 2^-e is taken from the column geometry
// again at the end rather than held through the loop.  Both keep the
// register count within the occupancy targets.
template <int S, int BG, int Q>
__device__ __forceinline__ void integral_block_recurrence(const RegionLaunch& it, long lb,
                                                          const EvalConst* __restrict__ consts,
                                                          double* __restrict__ prow, double2* tile,
                                                          DD (*red)[kIntComps])
{
    const EvalConst& K = consts[it.slot];
    const int R = it.r_steps, n = it.n_conv;
    const long C = it.n_cols;
    const long col = lb * kBlock + threadIdx.x;
    const bool valid = col < C;
    const long cc = valid ? col : C - 1;
    const double h = it.grid_step;

    double acc[1 + S];
    double s2[S > 0 ? S : 1];
    int B[S > 0 ? S : 1];   // each stream's sum carried scaled by 2^B (restart_recurrence)
#pragma unroll
    for (int c = 0; c <= S; c++) acc[c] = 0.0;
    {
        const Sight<S> sight = make_sight<S, BG, true>(it.col[cc], it.col[C + cc], it.col[2 * C + cc], K);
        double Cs[S > 0 ? S : 1], G[S > 0 ? S : 1], Rr[S > 0 ? S : 1];
        // s2: each stream's smallest t^2 over the grid's nodes (the terms are
        // carried relative to 2^-s2; rows whose t^2 all exceed s2 + 60 are
        // negligible)
        const double r_last = fma((double)R * n - 1.0, h, it.grid_lo);
#pragma unroll
        for (int s = 0; s < S; s++) {
            const double d = sight.st[s].p * h;
            Cs[s] = exp2_neg(2.0 * d * d);
            G[s] = Rr[s] = 0.0;
            B[s] = 0;
            s2[s] = min_t2(fma(sight.st[s].p, it.grid_lo, -sight.st[s].m),
                           fma(sight.st[s].p, r_last, -sight.st[s].m));
        }
        const int rt = kTileD2 / n;
        const double2* rq = reinterpret_cast<const double2*>(it.rq);
        for (int k0 = 0; k0 < R; k0 += rt) {
            const int kt = (R - k0 < rt) ? R - k0 : rt;
            __syncthreads();
            for (int e = threadIdx.x; e < kt * n; e += kBlock) tile[e] = rq[(long)k0 * n + e];
            __syncthreads();
            for (int kk = 0; kk < kt; kk++) {
                const int k = k0 + kk;
                const double r0 = fma((double)k * n, h, it.grid_lo);
                row_by_recurrence<S, BG, Q>(tile + kk * n, n, r0, h, k % kRestartRows == 0, sight, Cs,
                                            s2, G, Rr, B, K, acc);
                if constexpr (BG == kBgHernquistAux) {
                    const RStep& rs = it.rstep[k];
                    acc[0] += K.bg_a * rs.q2 + K.bg_b * rs.q1 + K.bg_c * rs.q0;
                }
            }
        }
    }
    double out[1 + S];
#pragma unroll
    for (int c = 0; c <= S; c++) out[c] = 0.0;
    if (valid) {
        const double wnu = it.col[3 * C + cc];
        const Sight<S> again = make_sight<S, BG, true>(it.col[cc], it.col[C + cc], it.col[2 * C + cc], K);
        out[0] = acc[0] * wnu * kMixedBgScale<BG>;
#pragma unroll
        for (int s = 0; s < S; s++)
            out[1 + s] = ldexp(acc[1 + s], -B[s]) * wnu * K.on[s] * exp2_neg(again.st[s].e + s2[s]);
    }
    write_block_f64<S>(out, prow, red);
}

// One logical block of a resampled region in layout 4 (grid_step > 0,
// bg_stride 0): the background on its own grid uniform in magnitude (the
// bg_count (r, weight) entries of bg_rq), then each switched-on stream by
// recurrence on a distance grid of its own (st_rows rows of kResampleRow
// weights from st_rq, node m at st_lo + m st_step), with the row skipping and
// restarts of row_by_recurrence.  Every table entry is the same for all
// lanes, so the tables are read through the constant address space, where
// uniform loads become scalar loads into SGPRs: no LDS staging and no
// barrier between the loops, whose waves skip different rows.  (Through the
// LDS tile, each group of four terms waited on its ds_read: 4% slower.)
template <int S, int BG>
__device__ __forceinline__ void integral_block_split(const RegionLaunch& it, long lb,
                                                     const EvalConst* __restrict__ consts,
                                                     double* __restrict__ prow, DD (*red)[kIntComps])
{
    typedef const __attribute__((address_space(4))) double* ConstPtr;
    const EvalConst& K = consts[it.slot];
    const long C = it.n_cols;
    const long col = lb * kBlock + threadIdx.x;
    const bool valid = col < C;
    const long cc = valid ? col : C - 1;
    constexpr int n = kResampleRow;

    double acc[1 + S];
#pragma unroll
    for (int c = 0; c <= S; c++) acc[c] = 0.0;
    {
        const Sight<S> sight = make_sight<S, BG, true>(it.col[cc], it.col[C + cc], it.col[2 * C + cc], K);
        const ConstPtr bq = (ConstPtr)(it.bg_rq);
#pragma unroll 2
        for (int i = 0; i < it.bg_count; i++) bg_point_mixed<BG>(bq[2 * i], bq[2 * i + 1], sight.bg, K, acc[0]);
#pragma unroll
        for (int s = 0; s < S; s++) {
            const int R = it.st_rows[s];
            if (R == 0) continue;   // switched off: no terms
            const double h = it.st_step[s], lo = it.st_lo[s];
            const double p = sight.st[s].p, mm = sight.st[s].m;
            const double d = p * h;
            const double c = exp2_neg(2.0 * d * d);
            // Nodes with |t| above T = sqrt(min t^2 + kNegligibleLog2) hold
            // terms below 2^-kNegligibleLog2 of the sight's largest; t0, the row's
            // first t, advances by n d per row (a rounding per row, far
            // inside the bound's margin).
            const double t_lo = fma(p, lo, -mm);
            // the smallest t^2 over the grid's nodes: terms are carried relative to 2^-s2
            const double s2 = min_t2(t_lo, fma(p, fma((double)R * n - 1.0, h, lo), -mm));
            const double floor_t2 = s2 + kNegligibleLog2;
            const double T = floor_t2 * rsqrt_nr(floor_t2);
            const double dn = d * n;
            // t = t_lo + m d at node m rises along the sight, so the rows
            // holding a node with |t| <= T form one interval [k_lo, k_hi] per
            // lane (every row when d = 0, where |t| = |m| is within T); it is
            // widened by a node at each end against rounding.  The wave runs
            // the union of its lanes' intervals, so no row needs a test.
            int k_lo = 0, k_hi = R - 1;
            if (d > 0.0) {
                const double id = rcp_nr(d);
                const double m_a = fma(-T - t_lo, id, -1.0), m_b = fma(T - t_lo, id, 1.0);
                k_lo = (int)fmin(fmax(floor(m_a / n), 0.0), (double)R);
                k_hi = (int)fmin(fmax(floor(m_b / n), -1.0), (double)(R - 1));
            }
            // wavefront min / max (DPP reductions), made scalar
            const int k_first = __builtin_amdgcn_readfirstlane(__ockl_wfred_min_i32(k_lo));
            const int k_last = __builtin_amdgcn_readfirstlane(__ockl_wfred_max_i32(k_hi));
            // Eight terms per step, in Horner form:
            //   a += G_j (w_j + R_j (w_{j+1} + R_{j+1} (... + R_{j+6} w_{j+7})))
            // with R_{j+k} = R_j c^k; then G_{j+8} = G_j P_j,
            // P_j = R_j R_{j+1} ... R_{j+7} = R_j^8 c^28, P_{j+8} = P_j c^64 and
            // R_{j+8} = R_j c^8.  18 VALU per eight terms instead of 24 (the
            // innermost w_{j+6} + R w_{j+7} takes two: one instruction reads at
            // most one SGPR pair; four terms per step took 20).  c^64 by
            // squarings carries about 63 ulp, so after m steps since a restart
            // the error stays below 63 m^2 / 2 ulp: 3.6e-12 relative at the
            // restart interval's 32.
            const double c4 = (c * c) * (c * c), c8 = c4 * c4, c16 = c8 * c8, c24 = c16 * c8,
                         c64 = (c16 * c16) * (c16 * c16);
            const ConstPtr wq = (ConstPtr)(it.st_rq[s]);
            double G = 0.0, Rr = 0.0, Pq = 0.0, a = 0.0;
            int B = 0;   // G and a carried scaled by 2^B (restart_recurrence)
            double t0 = fma(dn, (double)k_first, t_lo);
            int since = kRestartRows;   // rows since the last restart: the first row restarts
            for (int k = k_first; k <= k_last; k++, t0 += dn, since++) {
                if (since >= kRestartRows || (t0 < 0.0 && G < 0x1p-1000)) {
                    restart_recurrence(t0, d, s2, G, Rr, B, a);
                    const double q = Rr * Rr * c, q2 = q * q;
                    Pq = (q2 * q2) * c24;   // (R^2 c)^4 c^24 = R^8 c^28
                    since = 0;
                }
                const ConstPtr row = wq + (long)k * n;
#pragma unroll 1
                for (int i = 0; i < n; i += 8) {
                    const double r1 = Rr * c, r2 = r1 * c, r3 = r2 * c, r4 = r3 * c, r5 = r4 * c, r6 = r5 * c;
                    double x = mul_add_sgpr(r6, row[i + 7], row[i + 6]);
                    x = fma_sgpr(r5, x, row[i + 5]);
                    x = fma_sgpr(r4, x, row[i + 4]);
                    x = fma_sgpr(r3, x, row[i + 3]);
                    x = fma_sgpr(r2, x, row[i + 2]);
                    x = fma_sgpr(r1, x, row[i + 1]);
                    a = fma(G, fma_sgpr(Rr, x, row[i]), a);
                    G *= Pq;
                    Pq *= c64;
                    Rr *= c8;
                }
            }
            // the scales taken out of every term
            acc[1 + s] = ldexp(a, -B) * exp2_neg(sight.st[s].e + s2);
        }
    }
    double out[1 + S];
#pragma unroll
    for (int c = 0; c <= S; c++) out[c] = 0.0;
    if (valid) {
        const double wnu = it.col[3 * C + cc];
        out[0] = acc[0] * wnu * kMixedBgScale<BG>;
#pragma unroll
        for (int s = 0; s < S; s++) out[1 + s] = acc[1 + s] * wnu * K.on[s];
    }
    write_block_f64<S>(out, prow, red);
}

// ---------------------------------------------------------------------------
// Mixed precision on the exact tables: G consecutive convolution points of
// one r-step (G = 0: the m < kMaxGroup points left at a row's end).  Each
// point's background term goes to bg in FP64 (bg_point_mixed); each stream
// term qw 2^-(t^2 - s2) is formed from the FP64 exponent by v_exp_f32 and
// summed over the group in FP32, with qw rounded to FP32 (qw32), then the
// group's sum is added to st[s] in FP64.  Per stream and point that is one
// FP32 FMA, against a conversion to FP64 and an FP64 FMA, plus one
// conversion and one FP64 add per group.
// Error of a group's sum: each of the G products and G - 1 additions rounds
// to nearest (relative 2^-24).  An addition drops, always downwards, a term
// under half an ulp of its running sum, that is under 2^-24 of the group's
// largest term: a bias below (G - 1) 2^-24 of the group, and only where one
// group's terms span 24 octaves.  Neighbouring convolution points' terms
// differ by a factor 2^(2 |t| dt), dt the step of t between them, so that
// happens only in a Gaussian's tail.  qw32 carries a relative error below
// 2^-24 per weight, the same in every column and parameter vector.
// The group size is the same in every kernel variant, so that the tuner's
// choice does not change a result's bits.
constexpr int kMaxGroup = 8;
constexpr int kStreamGroup = kMaxGroup;
// Rows of fewer convolution points add each term in FP64 instead
// (point_terms).  The FP32 weights' rounding is the same in every column,
// and the groups' additions drop tail terms downwards, so in an integral
// over many sights these errors average only over the points of a row:
// at 120 points the groups leave at most ~4e-9 of an integral, at 8-64
// points up to ~1.4e-8, against ~2e-10 for terms added in FP64.  Such
// coarse rows are cheap, and the kernels hold both loops.
constexpr int kGroupMinConv = 100;

template <int S, int BG, int G>
__device__ __forceinline__ void stream_group(const double2* row, const float* row32, const Sight<S>& v,
                                             const EvalConst& K, const double (&s2)[S > 0 ? S : 1],
                                             double& bg, double (&st)[kMaxStreams], int m = G)
{
    static_assert(G <= kMaxGroup, "group larger than kMaxGroup");
    float g[S > 0 ? S : 1];
#pragma unroll
    for (int s = 0; s < S; s++) g[s] = 0.0f;
#pragma unroll
    for (int u = 0; u < (G > 0 ? G : kMaxGroup); u++) {
        if (G == 0 && u >= m) break;
        const double2 p = row[u];
        const float qw32 = row32[u];
        bg_point_mixed<BG>(p.x, p.y, v.bg, K, bg);
#pragma unroll
        for (int s = 0; s < S; s++) {
            const double t = fma(v.st[s].p, p.x, -v.st[s].m);
            const float e = __builtin_amdgcn_exp2f(-(float)fma(t, t, -s2[s]));
            g[s] = __builtin_fmaf(qw32, e, g[s]);
        }
    }
#pragma unroll
    for (int s = 0; s < S; s++) st[s] += (double)g[s];
}

// ---------------------------------------------------------------------------
// Integral: one logical block = 256 consecutive columns of one region of one
// evaluation slot; the block's double-double partials go to prow.
//
// REC: the mixed kernels' recurrence and split layouts compiled in (items
// with grid_step > 0); without them the kernel holds the exact tables' loop
// alone, and its register allocation is that loop's.
template <int S, int BG, bool MIXED, int UNROLL, bool REC = true>
__device__ __forceinline__ void integral_block(const RegionLaunch& it, long lb,
                                               const EvalConst* __restrict__ consts,
                                               double* __restrict__ prow, double2* tile,
                                               float* tile32, DD (*red)[kIntComps])
{
    if constexpr (MIXED && REC) {
        if (it.grid_step > 0.0) {
            if (it.bg_stride == 0) integral_block_split<S, BG>(it, lb, consts, prow, red);
            else if (it.bg_stride == 4) integral_block_recurrence<S, BG, 4>(it, lb, consts, prow, tile, red);
            else if (it.bg_stride == 2) integral_block_recurrence<S, BG, 2>(it, lb, consts, prow, tile, red);
            else integral_block_recurrence<S, BG, 1>(it, lb, consts, prow, tile, red);
            return;
        }
    }
    const EvalConst& K = consts[it.slot];
    const int R = it.r_steps, n = it.n_conv;
    const long C = it.n_cols;
    // Logical block lb: 256-column tile lb / chunks, its r-steps
    // [k_begin, k_end) (window lb % chunks of chunks).
    const int chunks = it.r_chunks > 1 ? it.r_chunks : 1;
    const long tile_ix = lb / chunks;
    const int window = (int)(lb - tile_ix * chunks);
    const int k_begin = (int)((long)window * R / chunks), k_end = (int)((long)(window + 1) * R / chunks);
    const long col0 = tile_ix * kBlock;
    const long col = col0 + threadIdx.x;
    const bool valid = col < C;
    const long cc = valid ? col : C - 1;

    DD out[kIntComps];
#pragma unroll
    for (int c = 0; c < kIntComps; c++) out[c] = DD{ 0.0, 0.0 };

    // Blocks whose columns lie wholly outside [calc_begin, calc_end) add nothing.
    const long last = (col0 + kBlock < C ? col0 + kBlock : C);
    const bool active = it.calc_end > col0 * R && it.calc_begin < last * R;
    if (active) {
        const Sight<S> sight = make_sight<S, BG, MIXED>(it.col[cc], it.col[C + cc], it.col[2 * C + cc], K);
        const double wnu = it.col[3 * C + cc];
        const long base = cc * (long)R;
        long lo_l = it.calc_begin - base, hi_l = it.calc_end - base;
        const int r_lo = valid ? (int)(lo_l < 0 ? 0 : (lo_l > R ? R : lo_l)) : 0;
        const int r_hi = valid ? (int)(hi_l < 0 ? 0 : (hi_l > R ? R : hi_l)) : 0;

        // The column's sums over its r-steps: Kahan-compensated in FP64
        // (the reference's order, evaluation_optimized.c:504-519); in mixed
        // precision plain FP64 sums of positive terms, within (R - 1) 2^-53
        // relative (8e-14 at 700 r-steps, far below the FP32 terms' own
        // error), which frees the compensation's registers: the kernel then
        // runs 4 waves per SIMD without spills.
        double acc[1 + S], cmp[1 + S];
#pragma unroll
        for (int c = 0; c <= S; c++) acc[c] = cmp[c] = 0.0;
        // mixed: each stream's smallest t^2 over the region's points (0 where
        // the sight's closest approach to the stream lies among them)
        double s2[S > 0 ? S : 1];
#pragma unroll
        for (int s = 0; s < S; s++)
            s2[s] = MIXED ? min_t2(fma(sight.st[s].p, it.r_lo, -sight.st[s].m),
                                   fma(sight.st[s].p, it.r_hi, -sight.st[s].m))
                          : 0.0;

        const int rt = kTileD2 / n;   // r-steps per LDS tile (host guarantees n <= kTileD2)
        const double2* rq = reinterpret_cast<const double2*>(it.rq);
        for (int k0 = k_begin; k0 < k_end; k0 += rt) {
            const int kt = (k_end - k0 < rt) ? k_end - k0 : rt;
            __syncthreads();
            for (int e = threadIdx.x; e < kt * n; e += kBlock) {
                const double2 v = rq[(long)k0 * n + e];
                tile[e] = v;
                if (MIXED && n >= kGroupMinConv) tile32[e] = (float)v.y;
            }
            __syncthreads();
            for (int kk = 0; kk < kt; kk++) {
                const int k = k0 + kk;
                double bg = 0.0;
                double st[kMaxStreams] = { 0.0, 0.0, 0.0, 0.0 };
                const double2* row = tile + kk * n;
                if constexpr (MIXED) {
                    if (n >= kGroupMinConv) {
                        const float* row32 = tile32 + kk * n;
                        int i = 0;
                        for (; i + kStreamGroup <= n; i += kStreamGroup)
                            stream_group<S, BG, kStreamGroup>(row + i, row32 + i, sight, K, s2, bg, st);
                        if (i < n) stream_group<S, BG, 0>(row + i, row32 + i, sight, K, s2, bg, st, n - i);
                    } else {
#pragma unroll UNROLL
                        for (int i = 0; i < n; i++) {
                            const double2 p = row[i];
                            point_terms<S, BG, true, true>(p.x, p.y, sight, K, bg, st, s2);
                        }
                    }
                } else {
#pragma unroll UNROLL
                    for (int i = 0; i < n; i++) {
                        const double2 p = row[i];
                        point_terms<S, BG, false>(p.x, p.y, sight, K, bg, st);
                    }
                }
                const RStep& rs = it.rstep[k];
                double f = rs.reff_xr_rp3 * rs.irv;
                double bgv;
                if constexpr (BG == kBgZeroQ) {
                    bgv = -rs.irv;
                } else {
                    if constexpr (BG == kBgHernquistAux)
                        bg += K.bg_a * rs.q2 + K.bg_b * rs.q1 + K.bg_c * rs.q0;
                    bgv = bg * f;
                }
                if (k >= r_lo && k < r_hi) {
                    if constexpr (MIXED) {
                        acc[0] += bgv;
#pragma unroll
                        for (int s = 0; s < S; s++) acc[1 + s] = fma(st[s], f, acc[1 + s]);
                    } else {
                        kahan_add(acc[0], cmp[0], bgv);
#pragma unroll
                        for (int s = 0; s < S; s++) kahan_add(acc[1 + s], cmp[1 + s], st[s] * f);
                    }
                }
            }
        }
        if (valid) {
            // mixed: point_terms summed the background by bg_point_mixed
            constexpr double bg_scale = MIXED ? kMixedBgScale<BG> : 1.0;
            out[0] = DD{ (acc[0] + cmp[0]) * wnu * bg_scale, 0.0 };
#pragma unroll
            for (int s = 0; s < S; s++) {
                double v = (acc[1 + s] + cmp[1 + s]) * wnu * K.on[s];
                if constexpr (MIXED) v *= exp2_neg(sight.st[s].e + s2[s]);   // the shift taken out per point
                out[1 + s] = DD{ v, 0.0 };
            }
        }
    }
    block_reduce_dd<kIntComps>(out, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < kIntComps; c++) {
            prow[2 * c] = out[c].hi;
            prow[2 * c + 1] = out[c].lo;
        }
    }
}

// One hardware block per logical block; the stream count is a template
// parameter shared by every item of the launch (REC as for integral_block).
template <int S, int BG, bool MIXED, bool REC = true>
__global__ __launch_bounds__(kBlock) void integral_kernel(const RegionLaunch* __restrict__ items,
                                                          int n_items,
                                                          const EvalConst* __restrict__ consts,
                                                          double* __restrict__ partials)
{
    __shared__ double2 tile[kTileD2];
    __shared__ float tile32[MIXED ? kTileD2 : 1];
    __shared__ DD red[kBlock / 64][kIntComps];
    const int b = blockIdx.x;
    const RegionLaunch& it = items[find_item(items, n_items, b)];
    integral_block<S, BG, MIXED, 2, REC>(it, b - it.block_begin, consts,
                                         partials + (long)b * kIntComps * 2, tile, tile32, red);
}

template <int BG, bool MIXED, int UNROLL, bool REC>
__device__ __forceinline__ void integral_block_any(const RegionLaunch& it, long lb,
                                                   const EvalConst* __restrict__ consts,
                                                   double* __restrict__ prow, double2* tile,
                                                   float* tile32, DD (*red)[kIntComps])
{
    switch (it.n_streams) {
    case 0: integral_block<0, BG, MIXED, UNROLL, REC>(it, lb, consts, prow, tile, tile32, red); break;
    case 1: integral_block<1, BG, MIXED, UNROLL, REC>(it, lb, consts, prow, tile, tile32, red); break;
    case 2: integral_block<2, BG, MIXED, UNROLL, REC>(it, lb, consts, prow, tile, tile32, red); break;
    case 3: integral_block<3, BG, MIXED, UNROLL, REC>(it, lb, consts, prow, tile, tile32, red); break;
    default: integral_block<4, BG, MIXED, UNROLL, REC>(it, lb, consts, prow, tile, tile32, red); break;
    }
}

// Persistent form for batches: a grid sized to the device takes logical
// blocks from a device counter in order (the host orders items by cost per
// block, largest first) until every block is taken; every wave then exits.
// Items of any stream count share the launch.  Results are written per
// logical block, so they do not depend on which hardware block ran them.
// WAVES is the amdgpu waves-per-EU occupancy target (1 = compiler default);
// REC false: every item is on exact tables or an exponential-per-node layout
// (grid_step 0), so the recurrence paths are left out.
template <int BG, bool MIXED, int UNROLL, int WAVES, bool REC = true>
__global__ __launch_bounds__(kBlock, WAVES) void integral_persistent(
    const RegionLaunch* __restrict__ items, int n_items, int total_blocks,
    const EvalConst* __restrict__ consts, double* __restrict__ partials, unsigned* counter)
{
    __shared__ double2 tile[kTileD2];
    __shared__ float tile32[MIXED ? kTileD2 : 1];
    __shared__ DD red[kBlock / 64][kIntComps];
    __shared__ int next;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) next = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int b = next;
        if (b >= total_blocks) break;
        const RegionLaunch& it = items[find_item(items, n_items, b)];
        integral_block_any<BG, MIXED, UNROLL, REC>(it, b - it.block_begin, consts,
                                                   partials + (long)b * kIntComps * 2, tile, tile32, red);
    }
}

// Persistent form for a launch whose items are all in the split layout
// (resampled, mixed, bg_stride 0): only integral_block_split is compiled in,
// so the register allocation is that path's alone (121 VGPRs without spills
// at the compiler's occupancy, 96 at 5 waves per SIMD; the general kernel
// holds every path and needs 128 with spills).
template <int BG, int WAVES>
__global__ __launch_bounds__(kBlock, WAVES) void integral_persistent_split(
    const RegionLaunch* __restrict__ items, int n_items, int total_blocks,
    const EvalConst* __restrict__ consts, double* __restrict__ partials, unsigned* counter)
{
    __shared__ DD red[kBlock / 64][kIntComps];
    __shared__ int next;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) next = (int)atomicAdd(counter, 1u);
        __syncthreads();
        const int b = next;
        if (b >= total_blocks) break;
        const RegionLaunch& it = items[find_item(items, n_items, b)];
        double* prow = partials + (long)b * kIntComps * 2;
        const long lb = b - it.block_begin;
        switch (it.n_streams) {
        case 0: integral_block_split<0, BG>(it, lb, consts, prow, red); break;
        case 1: integral_block_split<1, BG>(it, lb, consts, prow, red); break;
        case 2: integral_block_split<2, BG>(it, lb, consts, prow, red); break;
        case 3: integral_block_split<3, BG>(it, lb, consts, prow, red); break;
        default: integral_block_split<4, BG>(it, lb, consts, prow, red); break;
        }
    }
}

__global__ void reset_counters(unsigned* counters, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) counters[i] = 0u;
}

// ---------------------------------------------------------------------------
// Star pass
// ---------------------------------------------------------------------------

// The streams of one star over its n convolution points, mixed precision,
// without the factor r_star^3: st[s] = sum_j kw3_j 2^-a_j(s) with
// kw3 = kw ratio^3 (qw = r_star^3 kw3).  Along the sight t_j = (p r_star)
// ratio_j - m and a_j = t_j^2 + e; t is monotone in ratio, so the smallest
// t^2 over the window, t0^2, sits at an end (or is 0 where the Gaussian's
// peak lies inside).  Each term's exponent t_j^2 - t0^2 is taken in FP64 by
// one FMA and only then rounded to FP32, which adds a relative error below
// (t_j^2 - t0^2) 2^-24 ln 2 to the 1-ulp error of the FP32 2^x; 2^-(e + t0^2)
// multiplies each sum once, in FP64.  The terms that carry the sum then have
// exponents below 1, unless the peak lies inside the window between two
// points further apart than 2 in t (p r_star above coarse_pr): a wave with
// such a stream on any lane splits every exponent into an integer k and a
// fraction f in FP64 instead, 2^-k by ldexp and only 2^-f through FP32.
template <int S>
__device__ __forceinline__ void star_streams_shifted(const Sight<S>& sight, double rstar, int n,
                                                     const __attribute__((address_space(4))) double* ratio,
                                                     const __attribute__((address_space(4))) double* kw3,
                                                     double coarse_pr, double (&st)[kMaxStreams])
{
    double pr[S > 0 ? S : 1], t0sq[S > 0 ? S : 1];
    bool coarse = false;
#pragma unroll
    for (int s = 0; s < S; s++) {
        pr[s] = sight.st[s].p * rstar;
        const double ta = fma(pr[s], ratio[0], -sight.st[s].m);
        const double tb = fma(pr[s], ratio[n - 1], -sight.st[s].m);
        const double lo = fmin(ta, tb), hi = fmax(ta, tb);
        const double t0 = lo > 0.0 ? lo : (hi < 0.0 ? hi : 0.0);
        t0sq[s] = t0 * t0;
        coarse |= t0 == 0.0 && pr[s] > coarse_pr;
    }
    if (__any(coarse)) {
        for (int j = 0; j < n; j++) {
            const double rho = ratio[j], w = kw3[j];
#pragma unroll
            for (int s = 0; s < S; s++) {
                const double t = fma(pr[s], rho, -sight.st[s].m);
                const double a = fma(t, t, -t0sq[s]);
                const double k = floor(a);
                const float e = __builtin_amdgcn_exp2f(-(float)(a - k));
                st[s] = fma(w, ldexp((double)e, -(int)fmin(k, 2000.0)), st[s]);
            }
        }
    } else {
        for (int j = 0; j < n; j++) {
            const double rho = ratio[j], w = kw3[j];
#pragma unroll
            for (int s = 0; s < S; s++) {
                const double t = fma(pr[s], rho, -sight.st[s].m);
                const float e = __builtin_amdgcn_exp2f(-(float)fma(t, t, -t0sq[s]));
                st[s] = fma(w, (double)e, st[s]);
            }
        }
    }
#pragma unroll
    for (int s = 0; s < S; s++) st[s] *= exp2_neg(sight.st[s].e + t0sq[s]);
}

// Convolved densities of star i (evaluation_optimized.c:190-299 for one star):
// bg is the background density times reff(r) 3 sigma / r^3 (or -1 when
// q == 0); st[s] is the stream density without that factor and without the
// stream's on/off flag.
//
// Mixed precision (Hernquist profile): the streams by star_streams_shifted
// (conv_kw3 = kw ratio^3), the background at the n points or, with
// bg_count > 0 (resampled pass), from the bg_count (r / r_star,
// weight / r_star^3) pairs of bg_tab, a fold of the convolution onto a grid
// uniform in the magnitude offset.
template <int S, int BG, bool MIXED>
__device__ __forceinline__ void star_densities(const double* __restrict__ star, long N, long i, int n,
                                               const EvalConst& K,
                                               const double* __restrict__ conv_ratio,
                                               const double* __restrict__ conv_kw, double& reff,
                                               double& bg, double (&st)[kMaxStreams],
                                               const double* __restrict__ conv_kw3, double coarse_pr,
                                               const double* __restrict__ bg_tab = nullptr,
                                               int bg_count = 0)
{
    const Sight<S> sight = make_sight<S, BG, MIXED>(star[i], star[N + i], star[2 * N + i], K);
    const double rstar = star[3 * N + i];
    reff = star[4 * N + i];

    bg = 0.0;
#pragma unroll
    for (int s = 0; s < kMaxStreams; s++) st[s] = 0.0;
    // The convolution tables are the same for every lane: read through the
    // constant address space, where uniform loads become scalar loads.
    typedef const __attribute__((address_space(4))) double* ConstPtr;
    const ConstPtr ratio = (ConstPtr)conv_ratio, kw = (ConstPtr)conv_kw;
    if constexpr (MIXED && BG == kBgHernquist) {
        const double rs3 = rstar * rstar * rstar;
        star_streams_shifted<S>(sight, rstar, n, ratio, (ConstPtr)conv_kw3, coarse_pr, st);
#pragma unroll
        for (int s = 0; s < S; s++) st[s] *= rs3;
        double b = 0.0;
        if (bg_count > 0) {
            const ConstPtr bq = (ConstPtr)bg_tab;
#pragma unroll 2
            for (int m = 0; m < bg_count; m++) bg_point_mixed<BG>(rstar * bq[2 * m], bq[2 * m + 1], sight.bg, K, b);
            b *= rs3;
        } else {
            for (int j = 0; j < n; j++) {
                const double r = rstar * ratio[j];
                bg_point_mixed<BG>(r, kw[j] * (r * r * r), sight.bg, K, b);
            }
        }
        bg = b * kMixedBgScale<BG>;
    } else {
        for (int j = 0; j < n; j++) {
            const double r = rstar * ratio[j];
            const double qw = kw[j] * (r * r * r);
            point_terms<S, BG, MIXED>(r, qw, sight, K, bg, st);
        }
    }
    if constexpr (BG == kBgZeroQ) {
        bg = -1.0;
    } else {
        if constexpr (BG == kBgHernquistAux) {
            for (int j = 0; j < n; j++) {
                // aux profile term qw (a g^2 + b g + c) with g = g' + dx_j
                const double r = rstar * conv_ratio[j];
                const double qw = conv_kw[j] * (r * r * r);
                const double g = 5.0 * (log10(r * 1000.0) - 1.0) + kAbsM;
                bg += qw * (K.bg_a * (g * g) + K.bg_b * g + K.bg_c);
            }
        }
        bg *= reff;
    }
}

template <int S, int BG, bool MIXED>
__device__ __forceinline__ void star_block(const StarLaunch& it, long lb,
                                           const EvalConst* __restrict__ consts,
                                           const double* __restrict__ conv_ratio,
                                           const double* __restrict__ conv_kw,
                                           double* __restrict__ prow, double* __restrict__ log10p,
                                           long log10p_stride, DD (*red)[kStarComps])
{
    const EvalConst& K = consts[it.slot];
    const long N = it.n_stars;
    const long idx = it.star_begin + lb * kBlock + threadIdx.x;
    const bool valid = idx < N;
    const long i = valid ? idx : N - 1;

    double reff, bg;
    double st[kMaxStreams];
    star_densities<S, BG, MIXED>(it.star, N, i, it.n_conv, K, conv_ratio, conv_kw, reff, bg, st, it.kw3,
                                 it.coarse_pr, it.bg_rq, it.bg_count);

    // Mixture (evaluation_optimized.c:674-716).  A model of more than 4
    // streams is summed over passes of 4 streams in stream order: the first
    // starts with the background term, each later pass continues the sum
    // from it.chain, and the last divides and takes the logarithm.
    DD out[kStarComps];
#pragma unroll
    for (int c = 0; c < kStarComps; c++) out[c] = DD{ 0.0, 0.0 };
    const bool first = (it.chain_flags & 1) != 0, last = (it.chain_flags & 2) != 0;
    double p;
    if (first) {
        const double bg_only = (bg / K.I_bg) * K.exp_bg_w;
        p = bg_only;
        out[1].hi = (bg_only == 0.0) ? -238.0 : log10(bg_only / K.sum_exp_w);
    } else {
        p = it.chain[i];
    }
#pragma unroll
    for (int s = 0; s < S; s++) {
        const double so = (st[s] * reff * K.on[s]) / K.I_st[s] * K.exp_st_w[s];
        p += so;
        out[2 + s].hi = (so == 0.0) ? -238.0 : log10(so / K.sum_exp_w);
    }
    double lp = 0.0;
    if (last) {
        p /= K.sum_exp_w;
        if (p != 0.0) {
            lp = log10(p);
            out[0].hi = lp;
        } else {
            lp = -238.0;
            out[kStarComps - 1].hi = 1.0;
        }
    } else if (valid) {
        it.chain[idx] = p;
    }
    if (!valid) {
#pragma unroll
        for (int c = 0; c < kStarComps; c++) out[c] = DD{ 0.0, 0.0 };
    } else if (log10p && last) {
        log10p[(long)it.out_index * log10p_stride + idx] = lp;
    }
    block_reduce_dd<kStarComps>(out, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < kStarComps; c++) {
            prow[2 * c] = out[c].hi;
            prow[2 * c + 1] = out[c].lo;
        }
    }
}

template <int S, int BG, bool MIXED>
__global__ __launch_bounds__(kBlock) void star_kernel(const StarLaunch* __restrict__ items, int n_items,
                                                      const EvalConst* __restrict__ consts,
                                                      const double* __restrict__ conv_ratio,
                                                      const double* __restrict__ conv_kw,
                                                      double* __restrict__ partials,
                                                      double* __restrict__ log10p, long log10p_stride)
{
    __shared__ DD red[kBlock / 64][kStarComps];
    const int b = blockIdx.x;
    const StarLaunch& it = items[find_item(items, n_items, b)];
    star_block<S, BG, MIXED>(it, b - it.block_begin, consts, conv_ratio, conv_kw,
                             partials + (long)b * kStarComps * 2, log10p, log10p_stride, red);
}

// Batched star pass: items of any stream count (and any context: each item
// carries its own star list and convolution tables) in one launch.
template <int BG, bool MIXED>
__global__ __launch_bounds__(kBlock) void star_kernel_any(const StarLaunch* __restrict__ items,
                                                          int n_items,
                                                          const EvalConst* __restrict__ consts,
                                                          double* __restrict__ partials)
{
    __shared__ DD red[kBlock / 64][kStarComps];
    const int b = blockIdx.x;
    const StarLaunch& it = items[find_item(items, n_items, b)];
    double* prow = partials + (long)b * kStarComps * 2;
    const long lb = b - it.block_begin;
    switch (it.n_streams) {
    case 0: star_block<0, BG, MIXED>(it, lb, consts, it.ratio, it.kw, prow, nullptr, 0, red); break;
    case 1: star_block<1, BG, MIXED>(it, lb, consts, it.ratio, it.kw, prow, nullptr, 0, red); break;
    case 2: star_block<2, BG, MIXED>(it, lb, consts, it.ratio, it.kw, prow, nullptr, 0, red); break;
    case 3: star_block<3, BG, MIXED>(it, lb, consts, it.ratio, it.kw, prow, nullptr, 0, red); break;
    default: star_block<4, BG, MIXED>(it, lb, consts, it.ratio, it.kw, prow, nullptr, 0, red); break;
    }
}

// ---------------------------------------------------------------------------
// Stream separation (separation.c:37-254): one work-item per star.
// ---------------------------------------------------------------------------

// drand48's generator (POSIX): X <- (a X + c) mod 2^48 with a = 0x5DEECE66D,
// c = 0xB; each call returns X / 2^48 of the new state.  The k-th call after
// srand48 is the affine map applied k times to the seeded state, built here
// by binary powering of (a, c), so star i draws the same number as the i-th
// sequential drand48() call in the reference's star loop.
__device__ __forceinline__ double drand48_at(uint64_t x0, uint64_t k)
{
    constexpr uint64_t kMask = (1ull << 48) - 1;
    uint64_t step_a = 0x5DEECE66Dull, step_c = 0xBull;
    uint64_t acc_a = 1, acc_c = 0;
    while (k) {
        if (k & 1) {
            acc_a = (acc_a * step_a) & kMask;
            acc_c = (acc_c * step_a + step_c) & kMask;
        }
        step_c = (step_c * step_a + step_c) & kMask;
        step_a = (step_a * step_a) & kMask;
        k >>= 1;
    }
    const uint64_t x = (acc_a * x0 + acc_c) & kMask;
    return (double)x * 0x1p-48;
}

// prob_ok (probability.c:310-403): 0 for the background, j + 1 for stream j,
// from cumulative sums taken in the reference's order.  Where the reference
// leaves its result unset (r == p[0] exactly with 2 or 3 streams, or NaN
// probabilities), this returns what its 4-stream branch does: 1 and 0.
template <int S>
__device__ __forceinline__ int prob_ok(const double (&p)[kMaxStreams], double r)
{
    if constexpr (S == 1) {
        return r > p[0] ? 0 : 1;
    } else {
        double cum[S];
        cum[0] = p[0];
#pragma unroll
        for (int j = 1; j < S; j++) cum[j] = cum[j - 1] + p[j];
        if (r > cum[S - 1]) return 0;
        if (r <= p[0]) return 1;
#pragma unroll
        for (int j = 1; j < S; j++)
            if (r > cum[j - 1] && r <= cum[j]) return j + 1;
        return 0;
    }
}

template <int S, int BG, bool MIXED>
__global__ __launch_bounds__(kBlock) void separation_kernel(StarLaunch it, EvalConst K, SepConst C,
                                                            int* __restrict__ assign,
                                                            double* __restrict__ xyz,
                                                            double* __restrict__ sprob)
{
    const long N = it.n_stars;
    const long i = (long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;

    double reff, bg;
    double st[kMaxStreams];
    star_densities<S, BG, MIXED>(it.star, N, i, it.n_conv, K, it.ratio, it.kw, reff, bg, st, it.kw3, it.coarse_pr);

    // Membership probabilities (separation.c:170-189).
    const double pbx = C.eps_b * bg / C.I_bg;
    double psg[kMaxStreams], p[kMaxStreams];
    double psg_sum = 0.0;
#pragma unroll
    for (int s = 0; s < S; s++) {
        psg[s] = C.eps_s[s] * (st[s] * reff * K.on[s]) / C.I_st[s];
        psg_sum += psg[s];
    }
#pragma unroll
    for (int s = 0; s < S; s++) {
        p[s] = psg[s] / (psg_sum + pbx);
        sprob[i * S + s] = p[s];
    }
    assign[i] = prob_ok<S>(p, drand48_at(C.x0, (uint64_t)i + 1));

    // Output position (separation.c:226-229): transform_point of the star's
    // galactic xyz is rot (g u), with u the unit vector towards (l, b) and
    // g = 5 log10(100 r) + 4.2 - 11 (stCoords.c:188-201, stVector.c:155-169:
    // the point is rebuilt at distance g along the same direction, and the
    // Sun's offset it carries cancels against the xsun subtraction).
    const double g = 5.0 * (log(100.0 * it.star[3 * N + i]) / log(10.0)) + 4.2 - 11.0;
    const double v[3] = { g * it.star[i], g * it.star[N + i], g * it.star[2 * N + i] };
#pragma unroll
    for (int k = 0; k < 3; k++) xyz[3 * i + k] = C.rot[k][0] * v[0] + C.rot[k][1] * v[1] + C.rot[k][2] * v[2];
}

// ---------------------------------------------------------------------------
// Second pass: per work item, the sum of its blocks' partials in block order.
// rows[out_index][c] = hi + lo.
// ---------------------------------------------------------------------------
template <int NC>
__global__ __launch_bounds__(kBlock) void reduce_kernel(const ReduceItem* __restrict__ items,
                                                        const double* __restrict__ partials,
                                                        double* __restrict__ rows)
{
    __shared__ DD red[kBlock / 64][NC];
    const ReduceItem it = items[blockIdx.x];
    DD v[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) v[c] = DD{ 0.0, 0.0 };
    for (int j = threadIdx.x; j < it.block_count; j += kBlock) {
        const double* p = partials + (long)(it.block_begin + j) * NC * 2;
#pragma unroll
        for (int c = 0; c < NC; c++) v[c] = dd_add(v[c], DD{ p[2 * c], p[2 * c + 1] });
    }
    block_reduce_dd<NC>(v, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < NC; c++) rows[(long)it.out_index * NC + c] = v[c].hi + v[c].lo;
    }
}

// I = I_region0 - I_region1 - ... per slot (evaluation_optimized.c:584-593),
// written into the slot's constants.  slot_rows[s] = first row of slot s;
// rows of one slot are consecutive, region 0 first.
__global__ void finalize_kernel(const int* __restrict__ slot_first_row,
                                const int* __restrict__ slot_n_rows,
                                const double* __restrict__ rows, EvalConst* consts, int n_slots)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_slots) return;
    const double* r = rows + (long)slot_first_row[s] * kIntComps;
    double bg = r[0];
    double st[kMaxStreams];
    for (int j = 0; j < kMaxStreams; j++) st[j] = r[1 + j];
    for (int g = 1; g < slot_n_rows[s]; g++) {
        bg -= r[g * kIntComps];
        for (int j = 0; j < kMaxStreams; j++) st[j] -= r[g * kIntComps + 1 + j];
    }
    consts[s].I_bg = bg;
    for (int j = 0; j < kMaxStreams; j++) consts[s].I_st[j] = st[j];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
using IntegralFn = void (*)(const RegionLaunch*, int, const EvalConst*, double*);
using StarFn = void (*)(const StarLaunch*, int, const EvalConst*, const double*, const double*,
                        double*, double*, long);

template <bool MIXED, int BG, bool REC = true>
IntegralFn integral_fn(int S)
{
    switch (S) {
    case 0: return integral_kernel<0, BG, MIXED, REC>;
    case 1: return integral_kernel<1, BG, MIXED, REC>;
    case 2: return integral_kernel<2, BG, MIXED, REC>;
    case 3: return integral_kernel<3, BG, MIXED, REC>;
    default: return integral_kernel<4, BG, MIXED, REC>;
    }
}

template <bool MIXED, int BG>
StarFn star_fn(int S)
{
    switch (S) {
    case 0: return star_kernel<0, BG, MIXED>;
    case 1: return star_kernel<1, BG, MIXED>;
    case 2: return star_kernel<2, BG, MIXED>;
    case 3: return star_kernel<3, BG, MIXED>;
    default: return star_kernel<4, BG, MIXED>;
    }
}

// Mixed precision is instantiated for the Hernquist profile (with the
// recurrence layouts, rec, or without); the other background modes run the
// FP64 kernels.
static IntegralFn pick_integral(int S, int bg, bool mixed, bool rec)
{
    if (mixed && bg == kBgHernquist)
        return rec ? integral_fn<true, kBgHernquist, true>(S) : integral_fn<true, kBgHernquist, false>(S);
    switch (bg) {
    case kBgHernquist: return integral_fn<false, kBgHernquist>(S);
    case kBgHernquistAux: return integral_fn<false, kBgHernquistAux>(S);
    case kBgGeneral: return integral_fn<false, kBgGeneral>(S);
    default: return integral_fn<false, kBgZeroQ>(S);
    }
}

static StarFn pick_star(int S, int bg, bool mixed)
{
    if (mixed && bg == kBgHernquist) return star_fn<true, kBgHernquist>(S);
    switch (bg) {
    case kBgHernquist: return star_fn<false, kBgHernquist>(S);
    case kBgHernquistAux: return star_fn<false, kBgHernquistAux>(S);
    case kBgGeneral: return star_fn<false, kBgGeneral>(S);
    default: return star_fn<false, kBgZeroQ>(S);
    }
}

using SeparationFn = void (*)(StarLaunch, EvalConst, SepConst, int*, double*, double*);

template <bool MIXED, int BG>
SeparationFn separation_fn(int S)
{
    switch (S) {
    case 1: return separation_kernel<1, BG, MIXED>;
    case 2: return separation_kernel<2, BG, MIXED>;
    case 3: return separation_kernel<3, BG, MIXED>;
    default: return separation_kernel<4, BG, MIXED>;
    }
}

static SeparationFn pick_separation(int S, int bg, bool mixed)
{
    if (mixed && bg == kBgHernquist) return separation_fn<true, kBgHernquist>(S);
    switch (bg) {
    case kBgHernquist: return separation_fn<false, kBgHernquist>(S);
    case kBgHernquistAux: return separation_fn<false, kBgHernquistAux>(S);
    case kBgGeneral: return separation_fn<false, kBgGeneral>(S);
    default: return separation_fn<false, kBgZeroQ>(S);
    }
}

hipError_t launch_integral(int S, int bg, bool mixed, const RegionLaunch* d_items, int n_items,
                           int n_blocks, const EvalConst* d_consts, double* d_partials,
                           hipStream_t stream, bool rec)
{
    if (n_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(pick_integral(S, bg, mixed, rec), dim3(n_blocks), dim3(kBlock), 0, stream,
                       d_items, n_items, d_consts, d_partials);
    return hipGetLastError();
}

hipError_t launch_star(int S, int bg, bool mixed, const StarLaunch* d_items, int n_items,
                       int n_blocks, const EvalConst* d_consts, const double* d_ratio,
                       const double* d_kw, double* d_partials, double* d_log10p, long stride,
                       hipStream_t stream)
{
    if (n_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(pick_star(S, bg, mixed), dim3(n_blocks), dim3(kBlock), 0, stream, d_items,
                       n_items, d_consts, d_ratio, d_kw, d_partials, d_log10p, stride);
    return hipGetLastError();
}

hipError_t launch_separation(int S, int bg, bool mixed, const StarLaunch& item, const EvalConst& k,
                             const SepConst& sc, int* d_assign, double* d_xyz, double* d_sprob,
                             hipStream_t stream)
{
    if (S < 1 || S > kMaxStreams) return hipErrorInvalidValue;
    if (item.n_stars <= 0) return hipSuccess;
    const long n_blocks = (item.n_stars + kBlock - 1) / kBlock;
    if (n_blocks > 0x7fffffffL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(pick_separation(S, bg, mixed), dim3((unsigned)n_blocks), dim3(kBlock), 0, stream,
                       item, k, sc, d_assign, d_xyz, d_sprob);
    return hipGetLastError();
}

hipError_t launch_reduce(bool star, const void* d_items, int n_items, const double* d_partials,
                         double* d_rows, hipStream_t stream)
{
    if (n_items <= 0) return hipSuccess;
    const ReduceItem* it = static_cast<const ReduceItem*>(d_items);
    if (star)
        hipLaunchKernelGGL(reduce_kernel<kStarComps>, dim3(n_items), dim3(kBlock), 0, stream, it,
                           d_partials, d_rows);
    else
        hipLaunchKernelGGL(reduce_kernel<kIntComps>, dim3(n_items), dim3(kBlock), 0, stream, it,
                           d_partials, d_rows);
    return hipGetLastError();
}

hipError_t launch_finalize(const int* d_first, const int* d_count, const double* d_rows,
                           EvalConst* d_consts, int n_slots, hipStream_t stream)
{
    if (n_slots <= 0) return hipSuccess;
    hipLaunchKernelGGL(finalize_kernel, dim3((n_slots + 63) / 64), dim3(64), 0, stream, d_first,
                       d_count, d_rows, d_consts, n_slots);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched (persistent) launchers and their tunable variants
// ---------------------------------------------------------------------------
// Variant v: inner-loop unroll factor and waves-per-EU occupancy target.
constexpr int kVariantUnroll[] = { 2, 4, 2, 4, 2, 4 };
constexpr int kVariantWaves[] = { 1, 1, 4, 4, 5, 5 };
constexpr int kNumVariants = 6;

using PersistentFn = void (*)(const RegionLaunch*, int, int, const EvalConst*, double*, unsigned*);

template <int BG, bool MIXED, int V, bool REC = true>
PersistentFn persistent_variant()
{
    return integral_persistent<BG, MIXED, kVariantUnroll[V], kVariantWaves[V], REC>;
}

template <int BG, bool MIXED, bool REC = true>
PersistentFn persistent_fn(int v)
{
    switch (v) {
    case 1: return persistent_variant<BG, MIXED, 1, REC>();
    case 2: return persistent_variant<BG, MIXED, 2, REC>();
    case 3: return persistent_variant<BG, MIXED, 3, REC>();
    case 4: return persistent_variant<BG, MIXED, 4, REC>();
    case 5: return persistent_variant<BG, MIXED, 5, REC>();
    default: return persistent_variant<BG, MIXED, 0, REC>();
    }
}

// Every variant for the Hernquist profile in both precisions (the mixed one
// with and without the recurrence paths); the other background modes run
// variant 0 in FP64.
static PersistentFn pick_persistent(int bg, bool mixed, int v, bool rec)
{
    if (bg == kBgHernquist) return !mixed ? persistent_fn<kBgHernquist, false>(v)
                                   : rec ? persistent_fn<kBgHernquist, true, true>(v)
                                         : persistent_fn<kBgHernquist, true, false>(v);
    switch (bg) {
    case kBgHernquistAux: return persistent_variant<kBgHernquistAux, false, 0>();
    case kBgGeneral: return persistent_variant<kBgGeneral, false, 0>();
    default: return persistent_variant<kBgZeroQ, false, 0>();
    }
}

int persistent_variants() { return kNumVariants; }

// Split-layout launches: the variants' occupancy targets map onto the split
// kernel's two (5 waves, or the compiler's); the unroll is the split path's own.
static PersistentFn pick_persistent_split(int v)
{
    if (v >= 0 && v < kNumVariants && kVariantWaves[v] == 5) return integral_persistent_split<kBgHernquist, 5>;
    return integral_persistent_split<kBgHernquist, 1>;
}

static PersistentFn pick_persistent_for(int bg, bool mixed, int variant, int layouts)
{
    return (layouts == kLaunchSplitOnly && mixed && bg == kBgHernquist)
               ? pick_persistent_split(variant)
               : pick_persistent(bg, mixed, variant, layouts != kLaunchExactOnly);
}

hipError_t launch_persistent(int bg, bool mixed, int variant, int grid, const RegionLaunch* d_items,
                             int n_items, int total_blocks, const EvalConst* d_consts,
                             double* d_partials, unsigned* d_counter, hipStream_t stream, int layouts)
{
    if (total_blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(pick_persistent_for(bg, mixed, variant, layouts), dim3(grid), dim3(kBlock), 0, stream,
                       d_items, n_items, total_blocks, d_consts, d_partials, d_counter);
    return hipGetLastError();
}

hipError_t persistent_scratch_bytes(int bg, bool mixed, int variant, int layouts, size_t* bytes)
{
    hipFuncAttributes attr;
    const hipError_t e = hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(
                                                  pick_persistent_for(bg, mixed, variant, layouts)));
    if (e == hipSuccess) *bytes = attr.localSizeBytes;
    return e;
}

hipError_t launch_reset_counters(unsigned* d_counters, int n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(reset_counters, dim3((n + 63) / 64), dim3(64), 0, stream, d_counters, n);
    return hipGetLastError();
}

hipError_t launch_star_any(int bg, bool mixed, const StarLaunch* d_items, int n_items, int n_blocks,
                           const EvalConst* d_consts, double* d_partials, hipStream_t stream)
{
    if (n_blocks <= 0) return hipSuccess;
    void (*fn)(const StarLaunch*, int, const EvalConst*, double*);
    if (mixed && bg == kBgHernquist) fn = star_kernel_any<kBgHernquist, true>;
    else if (bg == kBgHernquist) fn = star_kernel_any<kBgHernquist, false>;
    else if (bg == kBgHernquistAux) fn = star_kernel_any<kBgHernquistAux, false>;
    else if (bg == kBgGeneral) fn = star_kernel_any<kBgGeneral, false>;
    else fn = star_kernel_any<kBgZeroQ, false>;
    hipLaunchKernelGGL(fn, dim3(n_blocks), dim3(kBlock), 0, stream, d_items, n_items, d_consts,
                       d_partials);
    return hipGetLastError();
}

int tile_pairs() { return kTileD2; }
int integral_comps() { return kIntComps; }
int star_comps() { return kStarComps; }

}  // namespace mw
