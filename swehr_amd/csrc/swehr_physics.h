// swehr_physics.h - point physics of the SWEHR step for gfx950, templated on the
// storage/arithmetic type R (double for the plain solver, float for the precision form).
//
// Each function restates one piece of SWEHR.c (cited per function).  Expressions keep
// the reference's operand order so the FP64 build stays within a few ulp of it; the
// transcendental calls and divisions are strength-reduced where an identity is exact in
// real arithmetic: pow(x,.5) -> sqrt(x), pow(x,2) -> x*x, a/b -> a*(1/b) with the
// reciprocal shared, sqrt(g*d^2/(16g)) -> |d|/4, and repeated calls with equal
// arguments are made once.
#pragma once
#include <hip/hip_runtime.h>

namespace swehr {

// Thresholds of the reference's FP64 tests in the precision R.  A threshold t is formed
// in FP64 as SWEHR.c forms it (a constant, a parameter or a product of them); the FP32
// form rounds it down for x > t and up for x < t (x <= t is !(x > t)), to the float at
// which every float x compares as (double)x does with t.  Rounded to nearest, a float
// equal to that rounded value could compare on the other side of t.  In FP64 both are t.
constexpr float f32_below(double t)  // largest float <= t, for finite t >= 0
{
    const float f = (float)t;
    return (double)f > t ? __builtin_bit_cast(float, __builtin_bit_cast(unsigned, f) - 1u) : f;
}
constexpr float f32_above(double t)  // smallest float >= t, for finite t >= 0
{
    const float f = (float)t;
    return (double)f < t ? __builtin_bit_cast(float, __builtin_bit_cast(unsigned, f) + 1u) : f;
}
template <typename R>
__host__ __device__ constexpr R thr_gt(double t)
{
    if constexpr (sizeof(R) == sizeof(double)) return t;
    else return f32_below(t);
}
template <typename R>
__host__ __device__ constexpr R thr_lt(double t)
{
    if constexpr (sizeof(R) == sizeof(double)) return t;
    else return f32_above(t);
}

// Math helpers overloaded for both precisions, so FP32 code never promotes to FP64.
// sqrt for x = 0 or x a normal positive number (every use in the step: g h, squared
// speeds, settling discriminants): x rsqrt(x) refined by one Goldschmidt and one
// Markstein step, within 1 ulp, half the instructions of the library sqrt (which also
// rescales denormals).  x = 0 gives x; x < 0 gives NaN as sqrt does.
__device__ __forceinline__ double msqrt(double x)
{
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = 0.5 * r;
    const double e = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
    const double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return x == 0.0 ? x : g;
}
// FP32: the hardware square root and reciprocal square root (1 ulp), without the
// library's correction to the correctly rounded result (half of the error budget of a
// form that stores FP32 anyway, at ~15 instructions each).
__device__ __forceinline__ float msqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// The same for x > 0 (settling discriminants), without the screening of x = 0.
__device__ __forceinline__ double msqrt_pos(double x)
{
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = 0.5 * r;
    const double e = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
    const double d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}
__device__ __forceinline__ float msqrt_pos(float x) { return __builtin_amdgcn_sqrtf(x); }
// 1/sqrt(x) for x >= 1 (every use: 1 + squared slopes): the hardware estimate y
// refined to y (1 + t/2 + 3t^2/8) with t = 1 - x y^2, the library's own sequence
// minus its screening of zero and infinite x, which such x never need (the results
// are those of rsqrt).
__device__ __forceinline__ double mrsqrt(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double t = __builtin_fma(y * -x, y, 1.0);
    return __builtin_fma(y * t, __builtin_fma(t, 0.375, 0.5), y);
}
__device__ __forceinline__ float mrsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double mfabs(double x) { return __builtin_fabs(x); }
__device__ __forceinline__ float mfabs(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ double mfmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float mfmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double mfmax(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ float mfmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double mexp(double x) { return exp(x); }

// 2^(j/32), j = 0..31, correctly rounded: the table of exp_tab (staged in LDS).
__constant__ double kExp2Frac[32] = {
    0x1.0000000000000p+0, 0x1.059b0d3158574p+0, 0x1.0b5586cf9890fp+0, 0x1.11301d0125b51p+0,
    0x1.172b83c7d517bp+0, 0x1.1d4873168b9aap+0, 0x1.2387a6e756238p+0, 0x1.29e9df51fdee1p+0,
    0x1.306fe0a31b715p+0, 0x1.371a7373aa9cbp+0, 0x1.3dea64c123422p+0, 0x1.44e086061892dp+0,
    0x1.4bfdad5362a27p+0, 0x1.5342b569d4f82p+0, 0x1.5ab07dd485429p+0, 0x1.6247eb03a5585p+0,
    0x1.6a09e667f3bcdp+0, 0x1.71f75e8ec5f74p+0, 0x1.7a11473eb0187p+0, 0x1.82589994cce13p+0,
    0x1.8ace5422aa0dbp+0, 0x1.93737b0cdc5e5p+0, 0x1.9c49182a3f090p+0, 0x1.a5503b23e255dp+0,
    0x1.ae89f995ad3adp+0, 0x1.b7f76f2fb5e47p+0, 0x1.c199bdd85529cp+0, 0x1.cb720dcef9069p+0,
    0x1.d5818dcfba487p+0, 0x1.dfc97337b9b5fp+0, 0x1.ea4afa2a490dap+0, 0x1.f50765b6e4540p+0};

// FP64 exp from the 2^(j/32) table tab (kExp2Frac in LDS): x = (32 m + j) ln2/32 + r with
// |r| <= ln2/64 (ln2/32 split into a 32-bit head, so n * head is exact, and a tail);
// exp(x) = 2^m 2^(j/32) exp(r), exp(r) by its degree-6 Taylor polynomial in nested form
// 1 + r (1 + r/2 (... (1 + r/6))) (truncation < 4e-18; every addend is the inline 1).
// Within 2 ulp of the correctly rounded value (checked against libm exp on 2e4
// arguments) at about half the instructions of the library exp.  Below -746 the
// result is 0 as exp's; NaN stays NaN.
__device__ __forceinline__ double exp_tab(double x, const double *tab)
{
    x = x < -746.0 ? -746.0 : x;
    const double n = __builtin_rint(x * 46.16624130844683);
    double r = __builtin_fma(-n, 0x1.62e42fee00000p-6, x);
    r = __builtin_fma(-n, 0x1.a39ef35793c76p-38, r);
    const int ni = (int)n, j = ni & 31, m = ni >> 5;
    double p = __builtin_fma(r, 1.0 / 6, 1.0);
    p = __builtin_fma(r * 0.2, p, 1.0);
    p = __builtin_fma(r * 0.25, p, 1.0);
    p = __builtin_fma(r * (1.0 / 3), p, 1.0);
    p = __builtin_fma(r * 0.5, p, 1.0);
    p = __builtin_fma(r, p, 1.0);
    return __builtin_ldexp(tab[j] * p, m);
}
__device__ __forceinline__ float mexp(float x) { return expf(x); }
// exp(x) - 1 for x <= 0 without the cancellation of the difference at small |x| (FP32
// settling): the Taylor series to x^6 for x > -1/8 (truncation below 1e-9 relative),
// below that the hardware exp2 minus 1, where the difference is at least 0.117 and
// keeps its digits.  The lanes of a wave update the same class, so their x are alike
// and the branch is mostly uniform.
__device__ __forceinline__ float mexpm1_neg(float x)
{
    if (x > -0.125f) {
        float p = __builtin_fmaf(x, 1.0f / 720, 1.0f / 120);
        p = __builtin_fmaf(x, p, 1.0f / 24);
        p = __builtin_fmaf(x, p, 1.0f / 6);
        p = __builtin_fmaf(x, p, 0.5f);
        p = __builtin_fmaf(x, p, 1.0f);
        return x * p;
    }
    return __builtin_amdgcn_exp2f(x * 1.44269504f) - 1.0f;
}

// 32 well-mixed bits of x (the finaliser of MurmurHash3)
__device__ __forceinline__ unsigned mix32(unsigned x)
{
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

// a + b rounded to a float without bias (stochastic rounding): to the float below or
// above the exact sum with probabilities in proportion to its distance from the other
// one, decided by the 32 random bits `bits`.  The expected result is the exact sum, so
// an increment smaller than half the spacing of floats at a still accumulates over many
// steps, where rounding to nearest would drop it every time.  The exact error of the
// rounded sum comes from Knuth's TwoSum (every operation rounded once, no contraction).
__device__ __forceinline__ float add_unbiased(float a, float b, unsigned bits)
{
    const float s = a + b;
    const float bv = s - a;
    const float e = (a - (s - bv)) + (b - bv);  // a + b - s, exactly
    // s's neighbour on e's side: one unit more in magnitude where e has s's sign, one
    // less otherwise.  With e = 0 (exact sum, and any zero sum) the test below fails
    // and s is returned; a NaN or infinite s makes e NaN, which fails it as well.
    const int k = __float_as_int(s);
    const float nb = __int_as_float(k + (((__float_as_int(e) ^ k) >> 31) | 1));
    const float gap = nb - s;  // exact
    return (float)(bits >> 8) * 0x1p-24f * __builtin_fabsf(gap) < __builtin_fabsf(e) ? nb : s;
}

// For F = 1 + j/32, j = -9..13: 1/F rounded, and ln F as a rounded head plus the
// rounded rest (the table of log_tab, staged in LDS).
__constant__ double kLogTab[23][3] = {
    {0x1.642c8590b2164p+0, -0x1.522ae0738a3d8p-2, 0x1.8f7e9b38a6979p-57},
    {0x1.5555555555555p+0, -0x1.269621134db92p-2, -0x1.e0efadd9db02bp-56},
    {0x1.47ae147ae147bp+0, -0x1.f991c6cb3b379p-3, -0x1.f665066f980a2p-57},
    {0x1.3b13b13b13b14p+0, -0x1.a93ed3c8ad9e3p-3, -0x1.bcafa9de97203p-57},
    {0x1.2f684bda12f68p+0, -0x1.5bf406b543db2p-3, 0x1.1f5b44c0df7e7p-61},
    {0x1.2492492492492p+0, -0x1.1178e8227e47cp-3, 0x1.0e63a5f01c691p-58},
    {0x1.1a7b9611a7b96p+0, -0x1.9335e5d594989p-4, 0x1.478a85704ccb7p-58},
    {0x1.1111111111111p+0, -0x1.08598b59e3a07p-4, 0x1.dd7009902bf32p-58},
    {0x1.0842108421084p+0, -0x1.0415d89e74444p-5, -0x1.c05cf1d753622p-59},
    {0x1.0000000000000p+0, 0x0.0p+0, 0x0.0p+0},
    {0x1.f07c1f07c1f08p-1, 0x1.f829b0e783300p-6, 0x1.33e3f04f1ef23p-60},
    {0x1.e1e1e1e1e1e1ep-1, 0x1.f0a30c01162a6p-5, 0x1.85f325c5bbacdp-59},
    {0x1.d41d41d41d41dp-1, 0x1.6f0d28ae56b4cp-4, -0x1.906d99184b992p-58},
    {0x1.c71c71c71c71cp-1, 0x1.e27076e2af2e6p-4, -0x1.61578001e0162p-60},
    {0x1.bacf914c1bad0p-1, 0x1.29552f81ff523p-3, 0x1.301771c407dbfp-57},
    {0x1.af286bca1af28p-1, 0x1.5ff3070a793d4p-3, -0x1.bc60efafc6f6ep-58},
    {0x1.a41a41a41a41ap-1, 0x1.9525a9cf456b4p-3, 0x1.d904c1d4e2e26p-57},
    {0x1.999999999999ap-1, 0x1.c8ff7c79a9a22p-3, -0x1.4f689f8434012p-57},
    {0x1.8f9c18f9c18fap-1, 0x1.fb9186d5e3e2bp-3, -0x1.caaae64f21acbp-57},
    {0x1.8618618618618p-1, 0x1.1675cababa60ep-2, 0x1.ce63eab883717p-61},
    {0x1.7d05f417d05f4p-1, 0x1.2e8e2bae11d31p-2, -0x1.8f4cdb95ebdf9p-56},
    {0x1.745d1745d1746p-1, 0x1.4618bc21c5ec2p-2, 0x1.f42decdeccf1dp-56},
    {0x1.6c16c16c16c17p-1, 0x1.5d1bdbf5809cap-2, 0x1.4236383dc7fe1p-56},
};
constexpr double kLn2Hi = 0x1.62e42fefa3800p-1;  // ln 2 to 42 bits: m kLn2Hi is exact
constexpr double kLn2Lo = 0x1.ef35793c76730p-45;

// FP64 natural log of a positive normal x from the table tab (kLogTab in LDS): x = 2^m f
// with f in [sqrt(1/2), sqrt(2)), F = 1 + j/32 the nearest table point (f - F is exact),
// r = (f - F)/F by the tabulated 1/F, ln(1 + r) by its Taylor polynomial to r^9
// (|r| <= 1/64: truncation < 2^-54 relative), and ln x = m ln2 + ln F + ln(1 + r) summed
// small terms first.  Within 1 ulp of the C library's log (checked on 2e7 arguments
// over 1e-7..1e3) at well under half the instructions of the library log.
__device__ __forceinline__ double log_tab(double x, const double *tab)
{
    double f = 2.0 * __builtin_amdgcn_frexp_mant(x);
    int m = __builtin_amdgcn_frexp_exp(x) - 1;
    if (f >= 0x1.6a09e667f3bcdp+0) {
        f *= 0.5;
        m += 1;
    }
    const double jd = __builtin_rint((f - 1.0) * 32.0);
    const int j = (int)jd + 9;
    const double r = (f - __builtin_fma(jd, 1.0 / 32, 1.0)) * tab[3 * j];
    double q = __builtin_fma(r, 1.0 / 9, -1.0 / 8);
    q = __builtin_fma(r, q, 1.0 / 7);
    q = __builtin_fma(r, q, -1.0 / 6);
    q = __builtin_fma(r, q, 1.0 / 5);
    q = __builtin_fma(r, q, -1.0 / 4);
    q = __builtin_fma(r, q, 1.0 / 3);
    q = __builtin_fma(r, q, -0.5);
    const double p = __builtin_fma(r * r, q, r);
    const double dm = (double)m;
    return __builtin_fma(dm, kLn2Hi, tab[3 * j + 1] + (p + __builtin_fma(dm, kLn2Lo, tab[3 * j + 2])));
}
// 1/x for x finite and nonzero: the hardware estimate refined by two Newton steps in
// FP64 (within 1 ulp, half the instructions of IEEE division); the 1-ulp hardware
// reciprocal in FP32.  Where x = 0 is possible the result is NaN (not inf): callers use
// it only where that case is overwritten or compared away exactly as inf would be.
__device__ __forceinline__ double mrcp(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}
__device__ __forceinline__ float mrcp(float x) { return __builtin_amdgcn_rcpf(x); }

// constants of SWEHR.c:489-494, 632-650
constexpr double kG = 9.81;
// The film depth: the FP32 form compares depths with it rounded to float (the value
// it stores for a dry cell).  That rounded value lies below 1e-5 and the next float
// above 1e-5, so for every float h, h > float(1e-5) exactly when h > 1e-5 in FP64: the
// wet / dry tests agree with the reference's without widening h.
constexpr double kHfilm = 1e-5;
constexpr double kPhi = 0.4;
constexpr double kRhos = 2600.0;
constexpr double kRhow = 1000.0;
constexpr double kNu = 1.2e-6;
constexpr double kFrictionAngle = 0.8693;
constexpr double kCohesion = 400.0;
constexpr double kLambda = 0.65;
constexpr double kCthreshold = 0.4;
constexpr double kBetaBedload = 0.5;
constexpr double kHcFriction = 0.003;
constexpr double kDepthExp = -0.33;

// HLL flux across one face, SWEHR.c:327-409, from the face's left/right depths and
// velocities (un normal, ut tangential; the reference's qn/h and qt/h, zero where
// h <= hfilm, SWEHR.c:333-352) and the celerities cl = sqrt(g hl), cr = sqrt(g hr).
// Writes mass, normal- and tangential-momentum flux; returns the largest |wave
// speed|.  The flux branches are sequential ifs: the last true one wins (so
// hl = hr = 0 yields zero flux, not the 0/0 of the middle branch).
template <typename R>
__device__ __forceinline__ R face_flux_v(R hl, R hr, R ul, R ur, R vl, R vr, R cl, R cr, R g, R &f0, R &f1,
                                         R &f2)
{
    const R hfilm = R(kHfilm);
    const R d = ul + R(2) * cl - ur - R(2) * cr;
    // h* = d^2 / (16 g), so c* = msqrt(g h*) = |d| / 4
    const R ustar = R(.5) * (ul + ur) + cl - cr;
    const R cstar = mfabs(d) * R(0.25);
    // Sl = min(ul - cl, u* - c*), Sr = max(ur + cr, u* + c*) (SWEHR.c:358-362): the
    // compare-and-assign of the reference equals fmin/fmax, which drop a NaN the same way
    R sl = mfmin(ul - cl, ustar - cstar), sm = ustar, sr = mfmax(ur + cr, ustar + cstar);
    if (hl <= hfilm) { sl = ur - R(2) * cr; sm = sl; sr = ur + cr; }
    if (hr <= hfilm) { sl = ul - cl; sr = ul + R(2) * cl; sm = sr; }
    const R hg = R(.5) * g;
    R a0 = R(NAN), a1 = R(NAN);
    if (sl >= R(0)) {
        a0 = hl * ul;
        a1 = hl * ul * ul + hg * hl * hl;
    }
    if (sl <= R(0) && sr >= R(0)) {
        const R iw = mrcp(sr - sl);  // sr = sl only when both are 0: the last branch then wins
        a0 = (sr * (hl * ul) - sl * (hr * ur) + sl * sr * (hr - hl)) * iw;
        a1 = (sr * (hl * ul * ul + hg * hl * hl) - sl * (hr * ur * ur + hg * hr * hr) + sl * sr * (hr * ur - hl * ul)) *
             iw;
    }
    if (sr <= R(0)) {
        a0 = hr * ur;
        a1 = hr * ur * ur + hg * hr * hr;
    }
    f0 = a0;
    f1 = a1;
    f2 = a0 * ((sm >= R(0)) ? vl : vr);
    // max(0, |Sl|, |Sr|, |Sm|) ignoring NaN, as the reference's if-chain (SWEHR.c:404-408)
    return mfmax(mfmax(R(0), mfabs(sl)), mfmax(mfabs(sr), mfabs(sm)));
}

// ln(0.003) and log2(0.003): the critical depth of the depth-dependent Manning law
// (SWEHR.c:490)
constexpr double kLnHcFriction = -5.809142990314028;
constexpr double kLog2HcFriction = -8.380821783940931;

// n^2 / h^p for Manning's n of SWEHR.c:1356-1357 (n0 (h/0.003)^-0.33 below 0.003 m,
// n0 above), p = PI + 0.333 (3.333 at SWEHR.c:1360, 2.333 at SWEHR.c:1559).
// FP64: n0^2 exp(-p ln h [- 0.66 (ln h - ln 0.003)]), one log and one exp in place of
// pow(h/0.003, -0.33) and pow(h, p).
// FP32: n0^2 h^-PI 2^(-0.333 log2 h [- 0.66 (log2 h - log2 0.003)]).  Splitting off the
// integer power keeps the argument of the exponential small (|y| < 12 for h >= 1e-5),
// so its rounding costs ~1e-7 relative instead of the ~3e-6 of exp(-3.333 ln h) at
// h = 1e-3; the hardware log2/exp2 (1 ulp) and reciprocal are then accurate enough.
template <int PI, typename R>
__device__ __forceinline__ R manning2_over_hpow(R h, R n0, const double *exp_table, const double *log_table)
{
    if constexpr (sizeof(R) == sizeof(double)) {
        const double p = PI == 3 ? 3.333 : 2.333;
        const double lh = log_tab(h, log_table);
        double ex = -p * lh;
        if (h <= kHcFriction) ex += 2 * kDepthExp * (lh - kLnHcFriction);
        return n0 * n0 * exp_tab(ex, exp_table);
    } else {
        static_assert(PI == 2 || PI == 3, "friction powers 2.333 and 3.333");
        const float l2 = __builtin_amdgcn_logf(h);
        float y = -0.333f * l2;
        if (!(h > thr_gt<float>(kHcFriction))) y += float(2 * kDepthExp) * (l2 - float(kLog2HcFriction));
        const float ih = __builtin_amdgcn_rcpf(h);
        float ihp = ih * ih;
        if (PI == 3) ihp = ihp * ih;
        return n0 * n0 * ihp * __builtin_amdgcn_exp2f(y);
    }
}

// Wave-speed bound of the same face without the fluxes (SWEHR.c:333-376, 404-408),
// from the normal velocities ul, ur (0 below the film depth) and celerities
// cl = sqrt(g hl), cr = sqrt(g hr) of the two sides.
template <typename R>
__device__ __forceinline__ R face_speed_uc(R hl, R hr, R ul, R ur, R cl, R cr)
{
    const R hfilm = R(kHfilm);
    const R d = ul + R(2) * cl - ur - R(2) * cr;
    const R ustar = R(.5) * (ul + ur) + cl - cr;
    const R cstar = mfabs(d) * R(0.25);
    R sl = mfmin(ul - cl, ustar - cstar), sm = ustar, sr = mfmax(ur + cr, ustar + cstar);
    if (hl <= hfilm) { sl = ur - R(2) * cr; sm = sl; sr = ur + cr; }
    if (hr <= hfilm) { sl = ul - cl; sr = ul + R(2) * cl; sm = sr; }
    return mfmax(mfmax(R(0), mfabs(sl)), mfmax(mfabs(sr), mfabs(sm)));
}

// The same from the depths and normal momenta of the two sides and the face's GZ.
template <typename R>
__device__ __forceinline__ R face_speed(R hl, R hr, R qnl, R qnr, R g)
{
    const R hfilm = R(kHfilm);
    const R ul = (hl > hfilm) ? qnl * mrcp(hl) : R(0);
    const R ur = (hr > hfilm) ? qnr * mrcp(hr) : R(0);
    return face_speed_uc(hl, hr, ul, ur, msqrt(g * hl), msqrt(g * hr));
}

// Sediment-laden inflow state at j = ny at time t, SWEHR.c:1019-1033 (FP64 always).
struct InflowState {
    double h, hv, c;  // depth, discharge, suspended mass per area per class
};
__device__ __forceinline__ InflowState inflow_state(double t)
{
    InflowState s;
    s.hv = fmax(-0.035, t / 100 * -0.035);
    s.h = pow(fabs(s.hv) * 0.045 / sqrt(0.7), 0.6);
    s.c = fmin(500.0, t / 100 * 500) * s.h;
    return s;
}

// Rain rate for time t, SWEHR.c:813-814: entry ceil(t/rint + 0.01), capped at rnum.
__device__ __forceinline__ double rain_rate(const double *rain, int rnum, double rint, double t)
{
    int rind = (int)fmin(ceil(t / rint + 0.01), (double)rnum);
    if (rind < 1) rind = 1;
    return rain[rind - 1];
}

// Adaptive dt, SWEHR.c:1409-1412.
__device__ __forceinline__ double step_dt(double smax, double cstable, double dx, double maxstep, double t,
                                          double tend)
{
    double dt = cstable * dx / smax;
    if (dt > maxstep) dt = maxstep;
    if (t + dt > tend) dt = tend - t;
    return dt;
}

}  // namespace swehr
