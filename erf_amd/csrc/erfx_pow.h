// erfx_pow.h -- portable, reproducible x**y for positive x.
//
// The dycore's equation of state raises rho*theta to the powers gamma and
// R_d/c_v; the pressure perturbation p - p0 then cancels most of the digits,
// so a one-ulp difference in pow between two libraries shows up as a visible
// difference in the momentum after one step.  To make the GPU result
// reproducible bit for bit on any IEEE-754 host, this pow is written with the
// basic operations only (+ - * / and fused multiply-add, all correctly
// rounded) in double-double arithmetic:
//
//   log x = e ln2 + 2 atanh(s),  s = (m-1)/(m+1),  x = 2^e m,  m in [1/sqrt2, sqrt2)
//   x**y  = 2^k exp(r),          y log x = k ln2 + r,  |r| <= ln2/2
//
// with the atanh and exp series summed to beyond 2^-60 relative.  The result
// is within about 0.53 ulp of the exact value (the tests measure it against
// the C library's pow).  Compile every user with -ffp-contract=off.
//
// Domain: x > 0 finite and |y log x| < 700.  x == 0 gives 0 (for y > 0);
// x < 0, NaN or infinity gives NaN, which the callers treat as a bad state.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define ERFX_POW_FN __host__ __device__ inline
#else
#define ERFX_POW_FN inline
#endif

namespace erfx_pow_detail {

struct dd {
    double hi, lo;
};

ERFX_POW_FN dd two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return {s, e};
}

ERFX_POW_FN dd fast_two_sum(double a, double b) {   // requires |a| >= |b| or a == 0
    double s = a + b;
    double e = b - (s - a);
    return {s, e};
}

ERFX_POW_FN dd two_prod(double a, double b) {
    double p = a * b;
    double e = std::fma(a, b, -p);
    return {p, e};
}

// ln 2 to 107 bits, and ln 2 split with a 32-bit head so that k * LN2_HI32 is
// exact for |k| < 2^20.
constexpr double LN2_HI = 0x1.62e42fefa39efp-1;
constexpr double LN2_LO = 0x1.abc9e3b39803fp-56;
constexpr double LN2_HI32 = 0x1.62e42feep-1;
constexpr double LN2_LO32 = 0x1.a39ef35793c76p-33;
constexpr double INV_LN2 = 0x1.71547652b82fep0;
constexpr double SQRT2 = 0x1.6a09e667f3bcdp0;

// log x as a double-double, x > 0 finite.
ERFX_POW_FN dd log_dd(double x) {
    uint64_t b = __builtin_bit_cast(uint64_t, x);
    int e = 0;
    if ((b >> 52) == 0) {                        // subnormal: scale into the normal range
        x = x * 0x1p54;
        b = __builtin_bit_cast(uint64_t, x);
        e = -54;
    }
    e += (int)((b >> 52) & 0x7ff) - 1023;
    double m = __builtin_bit_cast(double, (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > SQRT2) {
        m = 0.5 * m;
        e += 1;
    }
    // s = (m-1)/(m+1) to double-double; m-1 is exact (Sterbenz), m+1 is split.
    double f = m - 1.0;
    dd d = two_sum(m, 1.0);
    double s_hi = f / d.hi;
    double rem = std::fma(-s_hi, d.hi, f);       // exact remainder of the division
    rem = rem - s_hi * d.lo;
    double s_lo = rem / d.hi;
    // atanh(s) = s + s^3 (1/3 + s^2/5 + s^4/7 + ...), |s| <= 0.1716, s^2 <= 0.0295
    double s2 = s_hi * s_hi;
    double p = 1.0 / 27.0;
    p = 1.0 / 25.0 + s2 * p;
    p = 1.0 / 23.0 + s2 * p;
    p = 1.0 / 21.0 + s2 * p;
    p = 1.0 / 19.0 + s2 * p;
    p = 1.0 / 17.0 + s2 * p;
    p = 1.0 / 15.0 + s2 * p;
    p = 1.0 / 13.0 + s2 * p;
    p = 1.0 / 11.0 + s2 * p;
    p = 1.0 / 9.0 + s2 * p;
    p = 1.0 / 7.0 + s2 * p;
    p = 1.0 / 5.0 + s2 * p;
    // 1/3 + s2*p with the 1/3 carried to double-double
    dd third = {0x1.5555555555555p-2, 0x1.5555555555555p-56};
    double q = s2 * p;
    dd c = two_sum(third.hi, q);
    c.lo += third.lo;
    // s^3 as double-double (s_hi^3 exactly enough, plus the s_lo term)
    dd sq = two_prod(s_hi, s_hi);
    dd cube = two_prod(sq.hi, s_hi);
    cube.lo += sq.lo * s_hi + 3.0 * s2 * s_lo;
    // tail = s^3 * c
    dd tail = two_prod(cube.hi, c.hi);
    tail.lo += cube.hi * c.lo + cube.lo * c.hi;
    // atanh = s + tail
    dd a = two_sum(s_hi, tail.hi);
    a.lo += s_lo + tail.lo;
    a = fast_two_sum(a.hi, a.lo);
    // log m = 2 atanh(s)  (exact doubling)
    dd lm = {2.0 * a.hi, 2.0 * a.lo};
    // + e ln2
    double ed = (double)e;
    dd el = two_prod(ed, LN2_HI);
    el.lo += ed * LN2_LO;
    dd r = two_sum(el.hi, lm.hi);
    r.lo += el.lo + lm.lo;
    return fast_two_sum(r.hi, r.lo);
}

// exp(t.hi + t.lo), |t| < 700.
ERFX_POW_FN double exp_dd(dd t) {
    double k = std::rint(t.hi * INV_LN2);
    double r1 = t.hi - k * LN2_HI32;             // exact (k * LN2_HI32 exact, Sterbenz)
    double r2 = t.lo - k * LN2_LO32;
    dd r = two_sum(r1, r2);
    double rh = r.hi, rl = r.lo;
    // exp(rh) = 1 + rh + rh^2/2 + rh^3 Q(rh),  Q = 1/6 + rh/24 + ... (to rh^16)
    double Q = 1.0 / 20922789888000.0;           // 1/16!
    Q = 1.0 / 1307674368000.0 + rh * Q;          // 1/15!
    Q = 1.0 / 87178291200.0 + rh * Q;            // 1/14!
    Q = 1.0 / 6227020800.0 + rh * Q;             // 1/13!
    Q = 1.0 / 479001600.0 + rh * Q;              // 1/12!
    Q = 1.0 / 39916800.0 + rh * Q;               // 1/11!
    Q = 1.0 / 3628800.0 + rh * Q;                // 1/10!
    Q = 1.0 / 362880.0 + rh * Q;                 // 1/9!
    Q = 1.0 / 40320.0 + rh * Q;                  // 1/8!
    Q = 1.0 / 5040.0 + rh * Q;                   // 1/7!
    Q = 1.0 / 720.0 + rh * Q;                    // 1/6!
    Q = 1.0 / 120.0 + rh * Q;                    // 1/5!
    Q = 1.0 / 24.0 + rh * Q;                     // 1/4!
    double Qt = rh * Q;                          // Q - 1/6
    dd sixth = {0x1.5555555555555p-3, 0x1.5555555555555p-57};
    dd c = two_sum(sixth.hi, Qt);
    c.lo += sixth.lo;
    dd sq = two_prod(rh, rh);
    dd cube = two_prod(sq.hi, rh);
    cube.lo += sq.lo * rh;
    double tail_hi = cube.hi * c.hi;
    double tail_lo = std::fma(cube.hi, c.hi, -tail_hi) + (cube.hi * c.lo + cube.lo * c.hi);
    // sum 1 + rh + sq/2 + tail + rl*(1 + rh)
    dd a = two_sum(rh, 0.5 * sq.hi);
    double lo = a.lo + 0.5 * sq.lo + tail_lo + rl * (1.0 + rh);
    dd b2 = two_sum(a.hi, tail_hi);
    lo += b2.lo;
    dd b = two_sum(1.0, b2.hi);
    lo += b.lo;
    double res = b.hi + lo;
    return std::ldexp(res, (int)k);
}

}  // namespace erfx_pow_detail

// x**y for x > 0 (see the header comment for the domain).
ERFX_POW_FN double erfx_pow(double x, double y) {
    using namespace erfx_pow_detail;
    if (!(x > 0.0) || !(x <= 1.7976931348623157e308)) {
        if (x == 0.0) return 0.0;
        return __builtin_nan("");
    }
    dd L = log_dd(x);
    dd t = two_prod(y, L.hi);
    t.lo += y * L.lo;
    t = fast_two_sum(t.hi, t.lo);
    return exp_dd(t);
}
