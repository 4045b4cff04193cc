// bs_math.hpp -- device-side Black-Scholes arithmetic for gfx950.
//
// The pricing algorithm is the reference's: Hull's cumulative normal via the
// Abramowitz & Stegun 26.2.17 polynomial (CNDF,
// parsec-ff/pkgs/apps/blackscholes/src/blackscholes.c:126-184) and the
// no-dividend European price (BlkSchlsEqEuroNoDiv, blackscholes.c:190-258).
// The reference computes in float with double literals; its result is
// sensitive to every rounding: with S up to 200 and K up to 300, one ulp of
// S*N(d1) or K*exp(-rT)*N(d2) is 1.5e-5..3e-5, so a one-ulp difference in any
// transcendental can move a price by 6e-5..8e-5 after the final roundings.
// Staying inside the reference's 1e-4 check for every option therefore means
// reproducing its rounding sequence.  Three arithmetic policies:
//
//   Ref32    (default)  float, the reference's statement order and roundings:
//            correctly rounded division, reciprocal and square root (hardware
//            seed plus one residual correction), products with the double
//            literals rounded once (a*c_hi + a*c_lo in one FMA), exp and log
//            from v_exp_f32 and v_log_f32.  Each refinement was switched off
//            in turn over five 10M-option books plus the edge and golden books
//            (50M options): with all of them the largest difference from the
//            reference is 8.39e-5 and none is above 9e-5; dropping the
//            correctly rounded reciprocal, the literal products, or division
//            and square root together takes it over 1e-4; dropping square
//            root, division or the reference's polynomial order alone takes
//            it to 9.5e-5..9.9e-5; a compensated exp (Cody-Waite argument)
//            changes nothing, so exp is the plain hardware one.
//   Fast32   float, hardware transcendentals (v_exp_f32, v_log_f32,
//            v_rcp_f32, v_sqrt_f32), Horner polynomial with FMA.  About a
//            third of Ref32's VALU work; prices differ from the reference by
//            up to ~1.2e-4 on a 10M book (mean 6e-6).
//   Ref64    fptype = double: the same statements in FP64 (its log, CNDF
//            reciprocal and CNDF polynomial are written out below).
//
// Every function here is compiled with FMA contraction off; each FMA is
// written out, so every kernel that inlines these functions computes the same
// bits (the serving loop relies on that).
//
// Greeks: bs_greeks<M>() runs the same forward pass as bs_price<M>() and then
// the reverse-mode adjoint of the price (the A&S polynomial differentiated
// exactly), returning price plus dP/dS, dP/dsigma, dP/dr and -dP/dT.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bs {

// A&S 26.2.17 / Hull constants, as the reference spells them
// (blackscholes.c:124,156,164-172).
constexpr double kInvSqrt2Pi = 0.39894228040143270286;
constexpr double kP = 0.2316419;
constexpr double kB1 = 0.319381530;
constexpr double kB2 = -0.356563782;
constexpr double kB3 = 1.781477937;
constexpr double kB4 = -1.821255978;
constexpr double kB5 = 1.330274429;
constexpr float kLog2e = 1.44269504088896340736f;
constexpr float kLn2f = 0.693147180559945309417f;

// A double constant c as c_hi + c_lo with c_hi = (float)c.
struct SplitConst {
    float hi, lo;
};
constexpr SplitConst split(double c) {
    return SplitConst{float(c), float(c - double(float(c)))};
}

// ---------------------------------------------------------------------------
// Correctly rounded float primitives for normal-range operands.
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// a / b: quotient from the hardware reciprocal, then one residual step
// (the residual a - b q is exact in an FMA).  Where the residual is not a
// number (b = 0 or +-inf, a = +-inf) the unrefined quotient already is IEEE's
// answer (+-inf, 0 or NaN), so it is returned: an option with T = 0 or
// sigma = 0 divides by zero in the reference too and prices from d = +-inf.
template <bool kChecked = true>
__device__ __forceinline__ float div_rn(float a, float b) {
    const float r = hw_rcp(b);
    const float q = a * r;
    const float e = __builtin_fmaf(-b, q, a);
    const float x = __builtin_fmaf(e, r, q);
    if constexpr (!kChecked) return x;
    return x == x ? x : q;
}

template <bool kChecked = true>
__device__ __forceinline__ float recip_rn(float b) {
    const float r = hw_rcp(b);
    const float e = __builtin_fmaf(-b, r, 1.0f);
    const float x = __builtin_fmaf(e, r, r);
    if constexpr (!kChecked) return x;
    return x == x ? x : r;
}

// sqrt(x): hardware root, then pick among it and its two neighbours by the
// sign of the exact residuals x - s' s.
__device__ __forceinline__ float sqrt_rn(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __int_as_float(__float_as_int(s) - 1);
    const float sp = __int_as_float(__float_as_int(s) + 1);
    const float rm = __builtin_fmaf(-sm, s, x);
    const float rp = __builtin_fmaf(-sp, s, x);
    // 0, +inf and NaN fall through: their residuals are NaN or 0.
    float o = rm <= 0.0f ? sm : s;
    return rp > 0.0f ? sp : o;
}

// round_f(a * c) for a double constant c = hi + lo: one rounding of
// a*hi + round(a*lo); equals round_f(round_d(a*c)) except within 2^-47 of a
// float midpoint.
__device__ __forceinline__ float mul_dc(float a, SplitConst c) {
    return __builtin_fmaf(a, c.hi, a * c.lo);
}

// exp(x) = 2^(x log2 e) on v_exp_f32.  The rounding of x log2 e costs up to
// |x| 2^-24 relative, which is large only where the result is tiny: the
// price path's exponents are -d^2/2 (the result scales a CNDF tail that is
// itself below 1e-5 once |x| > 12) and -rT (|x| <= 0.2).
__device__ __forceinline__ float exp_hw(float x) { return __builtin_amdgcn_exp2f(x * kLog2e); }

// log(x) = log2(x) ln2 from the hardware v_log_f32 (70% correctly rounded,
// the rest within two ulps on the S/K quotients of a book).
__device__ __forceinline__ float log_hw(float x) { return __builtin_amdgcn_logf(x) * kLn2f; }

// dCNDF/dx of the A&S approximation, given |x|, k = 1/(1+p|x|), the
// polynomial P(k) and n(x):  g = n (|x| P + p k^2 P'(k)).  Float, Horner.
__device__ __forceinline__ float cndf_grad(float ax, float k, float poly, float n) {
    float dp = __builtin_fmaf(k, float(5.0 * kB5), float(4.0 * kB4));
    dp = __builtin_fmaf(k, dp, float(3.0 * kB3));
    dp = __builtin_fmaf(k, dp, float(2.0 * kB2));
    dp = __builtin_fmaf(k, dp, float(kB1));
    return n * __builtin_fmaf(ax, poly, float(kP) * k * k * dp);
}

// ---------------------------------------------------------------------------
// Ref32: the reference's float arithmetic.
//
// Ref32Arith<true> carries the special-value handling: the NaN tests of the
// refined quotient and reciprocal, and CNDF's cap on |x|.  Ref32Arith<false>
// leaves them out (12 fewer VALU instructions per option).  The two give the
// same bits for every option whose unchecked price is not NaN:
//   - a refinement's NaN test changes a result only where the refined value is
//     NaN; a NaN quotient S/K or d1 reaches both CNDF values, a NaN reciprocal
//     reaches its CNDF value, and a NaN CNDF value makes the price NaN for both
//     option types (S * NaN and 1 - NaN are NaN);
//   - the cap changes nothing for finite |x| > 1e10: both forms give n = 0
//     (e^{-x^2/2} underflows, and -x^2/2 = -inf gives 0), a finite k, a tail
//     of 0 and CNDF = 1 exactly, and a gradient n * (...) = +0; only |x| = inf
//     differs, where the unchecked product inf * c_hi + inf * c_lo is NaN.
// So bs_price / bs_greeks run Ref32::Unchecked and recompute with Ref32 only
// the options whose price came out NaN (never in a book of finite positive
// inputs: the branch costs one compare per option).
template <bool kChecked>
struct Ref32Arith {
    using T = float;
    static constexpr bool kSplitCndf = true;
    __device__ static __forceinline__ float exp(float x) { return exp_hw(x); }
    __device__ static __forceinline__ float log(float x) { return log_hw(x); }
    __device__ static __forceinline__ float div(float a, float b) { return div_rn<kChecked>(a, b); }
    __device__ static __forceinline__ float sqrt(float x) { return sqrt_rn(x); }

    // blackscholes.c:126-184, statement by statement: y = 1 - tail(|x|), the
    // value for x >= 0; CNDF(x) is y for x >= 0 and 1 - y for x < 0
    // (:143-147, :179-181).
    template <bool kGrad>
    __device__ static __forceinline__ float cndf_abs(float x, float* grad) {
#pragma clang fp contract(off)
        constexpr SplitConst c0 = split(kInvSqrt2Pi), cp = split(kP), b1 = split(kB1),
                             b2 = split(kB2), b3 = split(kB3), b4 = split(kB4),
                             b5 = split(kB5);
        // |x| capped at 1e10: for every |x| >= 20 the reference returns exactly
        // 1 (e^{-x^2/2} is 0 or denormal and the tail rounds away), and so does
        // the capped value; uncapped, |x| = inf would meet inf*hi + inf*lo = NaN
        // in mul_dc where the reference's double product is inf.  NaN passes.
        float ax = __builtin_fabsf(x);   // |-0| = +0 gives the same bits below
        if constexpr (kChecked) ax = ax > 1e10f ? 1e10f : ax;
        // e^{-x^2/2} (:152): the reference's exponent (-0.5 x) x, times log2 e
        // for v_exp_f32, rounds twice; -0.5 scales exactly, so x^2 times the
        // one constant -0.5 log2 e rounds to the same float for every x (all
        // 2^32 inputs checked on the MI355X)
        float n = __builtin_amdgcn_exp2f((ax * ax) * (-0.5f * kLog2e));
        n = mul_dc(n, c0);
        float k = mul_dc(ax, cp);
        k = 1.0f + k;
        k = recip_rn<kChecked>(k);
        const float k2 = k * k;
        const float k3 = k2 * k;
        const float k4 = k3 * k;
        const float k5 = k4 * k;
        float t1 = mul_dc(k, b1);
        float acc = mul_dc(k2, b2);
        acc = acc + mul_dc(k3, b3);
        acc = acc + mul_dc(k4, b4);
        acc = acc + mul_dc(k5, b5);
        t1 = acc + t1;
        const float tail = t1 * n;
        const float y = 1.0f - tail;
        if constexpr (kGrad) *grad = cndf_grad(ax, k, t1, n);
        return y;
    }

    template <bool kGrad>
    __device__ static __forceinline__ float cndf(float x, float* grad) {
        const float y = cndf_abs<kGrad>(x, grad);
        return x < 0.0f ? 1.0f - y : y;
    }
};

struct Ref32 : Ref32Arith<true> {
    using Unchecked = Ref32Arith<false>;
};

// ---------------------------------------------------------------------------
// Fast32: hardware transcendentals.
struct Fast32 {
    using T = float;
    using Unchecked = Fast32;
    static constexpr bool kSplitCndf = false;
    __device__ static __forceinline__ float exp(float x) {
        return __builtin_amdgcn_exp2f(x * kLog2e);
    }
    __device__ static __forceinline__ float log(float x) {
        return __builtin_amdgcn_logf(x) * kLn2f;
    }
    __device__ static __forceinline__ float div(float a, float b) { return a * hw_rcp(b); }
    __device__ static __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

    template <bool kGrad>
    __device__ static __forceinline__ float cndf(float x, float* grad) {
#pragma clang fp contract(off)
        const float ax = __builtin_fabsf(x);
        const float n = exp(-0.5f * ax * ax) * float(kInvSqrt2Pi);
        const float k = hw_rcp(__builtin_fmaf(float(kP), ax, 1.0f));
        float poly = __builtin_fmaf(k, float(kB5), float(kB4));
        poly = __builtin_fmaf(k, poly, float(kB3));
        poly = __builtin_fmaf(k, poly, float(kB2));
        poly = __builtin_fmaf(k, poly, float(kB1));
        poly = poly * k;
        const float y = __builtin_fmaf(-poly, n, 1.0f);
        if constexpr (kGrad) *grad = cndf_grad(ax, k, poly, n);
        return x < 0.0f ? 1.0f - y : y;
    }
};

// ---------------------------------------------------------------------------
// FP64 primitives for the fptype = double policy.

// 1/b for finite b with |b| in [2^-1000, 2^1000]: the hardware reciprocal
// (v_rcp_f64) refined by two Newton steps y <- y + y (1 - b y), each residual
// in one FMA.  NaN propagates.
__device__ __forceinline__ double recip64(double b) {
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    return __builtin_fma(y, e, y);
}

constexpr double kSqrtHalf = 0.70710678118654752440;
// ln 2 = kLn2Hi + kLn2Lo; kLn2Hi keeps 32 significant bits, so e * kLn2Hi is
// exact for every binary exponent e of a double.
constexpr double kLn2Hi = 0.6931471803691238;        // 0x3fe62e42fee00000
constexpr double kLn2Lo = 1.9082149292705877e-10;

// Natural logarithm.  x = 2^e m with m in [sqrt(1/2), sqrt(2)); f = m - 1 is
// exact and s = f / (2 + f), |s| <= 0.1716.  Then
//   log m = 2 atanh(s) = f - h + s (h + R),   h = f^2 / 2,
//   R = sum_{j=1..10} 2 s^(2j) / (2j + 1)
// (the Taylor series of atanh; the first omitted term is below 6.3e-19 of
// log m), and log x = e kLn2Hi + (log m + e kLn2Lo).  s is the quotient from
// recip64 with one residual correction.  Zero gives -inf, +inf gives +inf,
// negative arguments and NaN give NaN, as IEEE log.
__device__ __forceinline__ double log64(double x) {
#pragma clang fp contract(off)
    double m = __builtin_amdgcn_frexp_mant(x);    // [0.5, 1), subnormals too
    int e = __builtin_amdgcn_frexp_exp(x);
    const bool lo = m < kSqrtHalf;
    m = lo ? m + m : m;
    e = lo ? e - 1 : e;
    const double f = m - 1.0;
    const double d = f + 2.0;
    const double y = recip64(d);
    double s = f * y;
    s = __builtin_fma(__builtin_fma(-d, s, f), y, s);
    const double z = s * s;
    double p = 2.0 / 21.0;
    p = __builtin_fma(p, z, 2.0 / 19.0);
    p = __builtin_fma(p, z, 2.0 / 17.0);
    p = __builtin_fma(p, z, 2.0 / 15.0);
    p = __builtin_fma(p, z, 2.0 / 13.0);
    p = __builtin_fma(p, z, 2.0 / 11.0);
    p = __builtin_fma(p, z, 2.0 / 9.0);
    p = __builtin_fma(p, z, 2.0 / 7.0);
    p = __builtin_fma(p, z, 2.0 / 5.0);
    p = __builtin_fma(p, z, 2.0 / 3.0);
    const double rr = z * p;
    const double h = 0.5 * (f * f);
    const double lm = f + __builtin_fma(s, h + rr, -h);
    const double ed = (double)e;
    const double r = __builtin_fma(ed, kLn2Hi, __builtin_fma(ed, kLn2Lo, lm));
    if (x > 0.0 && x < __builtin_inf()) return r;
    return x == 0.0 ? -__builtin_inf() : (x > 0.0 ? x : __builtin_nan(""));
}

// Ref64: fptype = double, the reference's statements with every step in FP64.
// Division and square root are IEEE (the compiler's correctly rounded
// sequences) and exp is the device library's; log is log64; CNDF's
// reciprocal 1 / (1 + p|x|) is recip64, its argument capped at 1e10 (past
// |x| = 38.6 the result is already exactly 0 or 1, and the cap keeps the
// denominator finite), and CNDF's polynomial is evaluated by Horner's rule
// with FMA rather than from the powers of k.  Against the reference compiled
// with fptype = double, prices differ by a few units in the last place.
struct Ref64 {
    using T = double;
    using Unchecked = Ref64;
    static constexpr bool kSplitCndf = false;
    __device__ static __forceinline__ double exp(double x) { return ::exp(x); }
    __device__ static __forceinline__ double log(double x) { return log64(x); }
    __device__ static __forceinline__ double div(double a, double b) { return a / b; }
    __device__ static __forceinline__ double sqrt(double x) { return ::sqrt(x); }

    template <bool kGrad>
    __device__ static __forceinline__ double cndf(double x, double* /*grad*/) {
#pragma clang fp contract(off)
        static_assert(!kGrad, "Ref64 prices only");
        const bool neg = x < 0.0;
        double ax = neg ? -x : x;
        ax = ax > 1e10 ? 1e10 : ax;
        const double n = exp(-0.5 * ax * ax) * kInvSqrt2Pi;
        const double k = recip64(__builtin_fma(kP, ax, 1.0));
        // the A&S polynomial by Horner's rule, one FMA per coefficient
        double t1 = __builtin_fma(k, kB5, kB4);
        t1 = __builtin_fma(k, t1, kB3);
        t1 = __builtin_fma(k, t1, kB2);
        t1 = __builtin_fma(k, t1, kB1);
        t1 = t1 * k;
        const double y = __builtin_fma(-t1, n, 1.0);
        return neg ? 1.0 - y : y;
    }
};

// ---------------------------------------------------------------------------
// Forward pass of BlkSchlsEqEuroNoDiv (blackscholes.c:190-258), kept as
// intermediates so the adjoint can reuse them.
template <class T>
struct Forward {
    T sqrt_t, half_var, den, d1, d2, n1, n2, g1, g2, disc, fv;
    T y1, y2;   // Ref32Arith: CNDF's non-negative-side values, n_i = d_i < 0 ? 1 - y_i : y_i
};

template <class M, bool kGrad>
__device__ __forceinline__ Forward<typename M::T> bs_forward(typename M::T s, typename M::T k,
                                                            typename M::T r, typename M::T v,
                                                            typename M::T t) {
#pragma clang fp contract(off)
    using T = typename M::T;
    Forward<T> f;
    f.sqrt_t = M::sqrt(t);                     // :224
    const T log_sk = M::log(M::div(s, k));     // :226
    T half_var = v * v;                        // :231-232
    half_var = half_var * T(0.5);
    f.half_var = half_var;
    T d1 = r + half_var;                       // :234-236
    d1 = d1 * t;
    d1 = d1 + log_sk;
    f.den = v * f.sqrt_t;                      // :238
    f.d1 = M::div(d1, f.den);                  // :239
    f.d2 = f.d1 - f.den;                       // :240
    if constexpr (M::kSplitCndf) {                  // :245-246
        f.y1 = M::template cndf_abs<kGrad>(f.d1, &f.g1);
        f.y2 = M::template cndf_abs<kGrad>(f.d2, &f.g2);
        f.n1 = f.d1 < T(0) ? T(1) - f.y1 : f.y1;
        f.n2 = f.d2 < T(0) ? T(1) - f.y2 : f.y2;
    } else {
        f.n1 = M::template cndf<kGrad>(f.d1, &f.g1);
        f.n2 = M::template cndf<kGrad>(f.d2, &f.g2);
    }
    f.disc = M::exp(-r * t);                   // :248
    f.fv = k * f.disc;
    return f;
}

// Call: S N1 - FV N2 (:250).  Put: FV (1-N2) - S (1-N1) (:252-254).
template <class T>
__device__ __forceinline__ T bs_combine(T s, const Forward<T>& f, bool put) {
#pragma clang fp contract(off)
    if (!put) return s * f.n1 - f.fv * f.n2;
    const T m1 = T(1) - f.n1;
    const T m2 = T(1) - f.n2;
    return f.fv * m2 - s * m1;
}

// The same price from the CNDF halves (M::kSplitCndf), in fewer operations.
// The put's 1 - N(d) is 1 - y for d >= 0 and 1 - (1 - y) for d < 0, and the
// latter is y exactly: 1 - y is exact for y in [0.5, 1] (Sterbenz), and for
// the few y just below 0.5 that CNDF returns near d = 0, 1 - y is exact too
// (every float y it returns was checked on the MI355X), so the rounded
// difference 1 - (1 - y) is the float y itself.  So both option types
// multiply S and FV by a_i = (d_i < 0) != put ? 1 - y_i : y_i, and
// price = put ? FV a2 - S a1 : S a1 - FV a2, the reference's products and
// difference (:250, :252-254) without a branch per option type.
template <class T>
__device__ __forceinline__ T bs_combine_split(T s, const Forward<T>& f, bool put) {
#pragma clang fp contract(off)
    const T a1 = ((f.d1 < T(0)) != put) ? T(1) - f.y1 : f.y1;
    const T a2 = ((f.d2 < T(0)) != put) ? T(1) - f.y2 : f.y2;
    const T x = s * a1;
    const T z = f.fv * a2;
    return put ? z - x : x - z;
}

// A policy with a distinct unchecked form prices with it first and recomputes
// with the checked one where the price is NaN (Ref32: identical bits).
template <class M>
constexpr bool has_unchecked() {
    return !__is_same(typename M::Unchecked, M);
}

template <class M>
__device__ __forceinline__ typename M::T bs_price(typename M::T s, typename M::T k,
                                                  typename M::T r, typename M::T v,
                                                  typename M::T t, bool put) {
    if constexpr (has_unchecked<M>()) {
        using U = typename M::Unchecked;
        typename M::T p = bs_combine_split(s, bs_forward<U, false>(s, k, r, v, t), put);
        if (__builtin_expect(p != p, 0)) p = bs_combine(s, bs_forward<M, false>(s, k, r, v, t), put);
        return p;
    } else {
        return bs_combine(s, bs_forward<M, false>(s, k, r, v, t), put);
    }
}

struct Greeks32 {
    float price, delta, vega, rho, theta;
};

// Price and its adjoint in one pass.  Forward sweep as bs_price<M>; the
// reverse sweep seeds dP = 1 and walks back through
//   P = S N1 - FV N2 (call) or FV (1-N2) - S (1-N1) (put),
//   FV = K e^{-rT},  N_i = CNDF(d_i),  d2 = d1 - den,  d1 = A / den,
//   den = sigma sqrt(T),  A = (r + sigma^2/2) T + log(S/K).
// Outputs: delta = dP/dS, vega = dP/dsigma, rho = dP/dr, theta = -dP/dT (per
// year); vega and rho are per unit change, not per percent.
template <class M, bool kSplitCombine>
__device__ __forceinline__ Greeks32 bs_greeks_of(float s, float k, float r, float v, float t,
                                                 bool put) {
#pragma clang fp contract(off)
    const Forward<float> f = bs_forward<M, true>(s, k, r, v, t);
    Greeks32 o;
    if constexpr (kSplitCombine) o.price = bs_combine_split(s, f, put);
    else o.price = bs_combine(s, f, put);
    const float rden = hw_rcp(f.den);
    // adjoints of N1 and N2 are S and -FV for both option types
    const float s_bar0 = put ? f.n1 - 1.0f : f.n1;
    const float fv_bar = put ? 1.0f - f.n2 : -f.n2;
    const float d1_bar = __builtin_fmaf(s, f.g1, -(f.fv * f.g2));   // through d1 and d2
    const float den_bar = __builtin_fmaf(-d1_bar, f.d1 * rden, f.fv * f.g2);
    const float a_bar = d1_bar * rden;
    const float sqrt_t_bar = den_bar * v;
    o.delta = __builtin_fmaf(a_bar, hw_rcp(s), s_bar0);              // d log(S/K)/dS = 1/S
    o.vega = __builtin_fmaf(den_bar, f.sqrt_t, a_bar * t * v);
    const float fv_dr = fv_bar * f.fv;                                 // dFV = -FV (T dr + r dT)
    o.rho = __builtin_fmaf(a_bar, t, -(fv_dr * t));
    const float t_bar = __builtin_fmaf(a_bar, r + f.half_var,          // 1/sqrt(T) = v / den
                                       __builtin_fmaf(0.5f * sqrt_t_bar, v * rden, -(fv_dr * r)));
    o.theta = -t_bar;
    return o;
}

// Unchecked first, checked where the price is NaN: every checked-only
// difference in the forward pass makes the price NaN (see Ref32Arith), and
// the adjoint reads nothing else.
template <class M>
__device__ __forceinline__ Greeks32 bs_greeks(float s, float k, float r, float v, float t,
                                              bool put) {
    if constexpr (has_unchecked<M>()) {
        Greeks32 o = bs_greeks_of<typename M::Unchecked, true>(s, k, r, v, t, put);
        if (__builtin_expect(o.price != o.price, 0))
            o = bs_greeks_of<M, false>(s, k, r, v, t, put);
        return o;
    } else {
        return bs_greeks_of<M, false>(s, k, r, v, t, put);
    }
}

}  // namespace bs
