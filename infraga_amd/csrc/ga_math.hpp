// ga_math.hpp -- exp and exp10 of the ray engine's hot loop.
//
// The library exp / exp10 carry overflow, underflow and NaN handling that
// the medium and absorption formulas never need and cost about 30 VALU
// instructions per call; the step calls them some thirty times.  These
// versions use the Cody-Waite reduction x = n ln2 + r with the two-part ln2
// of fdlibm's e_exp.c (ln2_hi has trailing zero bits, so n * ln2_hi is exact
// for |n| < 2^20), then the Taylor series of e^r to degree 13 in Horner form
// (|r| <= ln2 / 2: the first omitted term is below 2^-61 relative), scaled
// by 2^n through ldexp (arguments beyond +-1500 give 0 or inf, NaN stays NaN).
// exp10(x) = e^(x ln10) with ln10 split into hi + lo and the product's
// rounding error recovered by an FMA, so the reduced argument keeps full
// precision.  Both stay within about 1 ulp of the correctly rounded result
// for finite arguments (tests/test_ga_math.py checks them on the host
// against long double references).
//
// Plain C++ on purpose: the same code compiles for the device (GA_HD set by
// ga_device.hpp) and for the host test.
#pragma once

#include <cmath>

#ifndef GA_HD
#define GA_HD inline
#endif

namespace ga {

GA_HD double ga_exp_reduced(double t, double t_lo)
{
    const double log2e = 1.44269504088896338700e+00;
    const double ln2_hi = 6.93147180369123816490e-01, ln2_lo = 1.90821492927058770002e-10;
    // e^t is 0 below -1500 and inf above 1500 in double precision; the
    // clamp keeps n within the exact range of n * ln2_hi and of ldexp
    const double tc = fmin(fmax(t, -1500.0), 1500.0);
    const double n = rint(tc * log2e);
    double r = fma(-n, ln2_hi, tc);
    r = fma(-n, ln2_lo, r) + t_lo;
    double p = 1.6059043836821613e-10;          // 1 / 13!
    p = fma(p, r, 2.08767569878681e-09);        // 1 / 12!
    p = fma(p, r, 2.505210838544172e-08);
    p = fma(p, r, 2.755731922398589e-07);
    p = fma(p, r, 2.7557319223985893e-06);
    p = fma(p, r, 2.48015873015873e-05);
    p = fma(p, r, 0.0001984126984126984);
    p = fma(p, r, 0.001388888888888889);
    p = fma(p, r, 0.008333333333333333);
    p = fma(p, r, 0.041666666666666664);
    p = fma(p, r, 0.16666666666666666);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return (t == t) ? ldexp(p, (int)n) : t;       // NaN in, NaN out
}

GA_HD double ga_exp(double x) { return ga_exp_reduced(x, 0.0); }

GA_HD double ga_exp10(double x)
{
    const double ln10_hi = 2.302585092994046, ln10_lo = -2.1707562233822494e-16;
    const double t = x * ln10_hi;
    const double t_lo = fma(x, ln10_hi, -t) + x * ln10_lo;
    return ga_exp_reduced(t, t_lo);
}

}  // namespace ga
