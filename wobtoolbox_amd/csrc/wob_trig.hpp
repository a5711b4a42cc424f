// Double-precision sine and cosine of small non-negative arguments (0 <= x <= 8), for the
// line directions and Box-Muller angles of the walk, and a correctly rounded FP32 natural
// logarithm for the 2D fundamental solution (log_cr, at the end).
//
// The reference's CPU build evaluates a direction as std::cos(theta * M_PI) and
// std::sin(theta * M_PI) in double and rounds each to float (scene.cuh:475-476, 785-786).
// A general double sincos carries a large-argument reduction path whose registers cost the
// walk kernel an occupancy step; the arguments here are at most 2 pi, so this one has only:
//   * Cody-Waite reduction x = q pi/2 + r, |r| <= pi/4, with pi/2 split into three parts
//     (33 + 33 + 53 significant bits, so q p1 and q p2 are exact for q <= 4 and
//     x - q p1 is exact by Sterbenz's lemma);
//   * Taylor polynomials of sin r to r^15 and cos r to r^16 (truncation below 2^-53
//     relative at |r| = pi/4), evaluated by Horner's rule with fused multiply-adds;
//   * the quadrant's sign and swap.
// The relative error is about 2^-51, so the FP32 rounding of either value equals that of the
// exactly rounded double except when the exact value lies within that distance of an FP32
// rounding boundary.  Every operation is an IEEE double operation written out (explicit
// fused multiply-adds, no contraction), so host and device compute the same bits.
#pragma once

#ifndef WOB_HD
#define WOB_HD
#endif

namespace wobtrig {

// pi/2 = kPio2_1 + kPio2_2 + kPio2_3 + O(2^-122)
constexpr double kPio2_1 = 0x1.921fb54400000p+0;
constexpr double kPio2_2 = 0x1.0b4611a600000p-34;
constexpr double kPio2_3 = 0x1.3198a2e037073p-69;
constexpr double kTwoOverPi = 0.63661977236758138;

WOB_HD inline void sincos_small(double x, double &s, double &c) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const double qd = __builtin_rint(x * kTwoOverPi);
    const int q = (int)qd;
    const double r = ((x - qd * kPio2_1) - qd * kPio2_2) - qd * kPio2_3;
    const double z = r * r;
    // sin r = r + r z (-1/3! + z (1/5! - ...)), cos r = 1 + z (-1/2! + z (1/4! - ...))
    double ps = -1.0 / 1307674368000.0;              // -1/15!
    ps = __builtin_fma(ps, z, 1.0 / 6227020800.0);   //  1/13!
    ps = __builtin_fma(ps, z, -1.0 / 39916800.0);    // -1/11!
    ps = __builtin_fma(ps, z, 1.0 / 362880.0);       //  1/9!
    ps = __builtin_fma(ps, z, -1.0 / 5040.0);        // -1/7!
    ps = __builtin_fma(ps, z, 1.0 / 120.0);          //  1/5!
    ps = __builtin_fma(ps, z, -1.0 / 6.0);           // -1/3!
    const double sr = __builtin_fma(r * z, ps, r);
    double pc = 1.0 / 20922789888000.0;              //  1/16!
    pc = __builtin_fma(pc, z, -1.0 / 87178291200.0); // -1/14!
    pc = __builtin_fma(pc, z, 1.0 / 479001600.0);    //  1/12!
    pc = __builtin_fma(pc, z, -1.0 / 3628800.0);     // -1/10!
    pc = __builtin_fma(pc, z, 1.0 / 40320.0);        //  1/8!
    pc = __builtin_fma(pc, z, -1.0 / 720.0);         // -1/6!
    pc = __builtin_fma(pc, z, 1.0 / 24.0);           //  1/4!
    pc = __builtin_fma(pc, z, -0.5);                 // -1/2!
    const double cr = __builtin_fma(pc, z, 1.0);
    const int k = q & 3;
    const double s0 = (k & 1) ? cr : sr, c0 = (k & 1) ? sr : cr;
    s = (k & 2) ? -s0 : s0;
    c = (k == 1 || k == 2) ? -c0 : c0;
}

// Natural logarithm of a positive normal float, rounded to FP32 from a double evaluation
// with relative error about 2^-47: correctly rounded except where the exact logarithm lies
// that close to a rounding boundary (two of the floats in [2^-26, 2^12), where the result is
// the neighbouring float).
//   x = 2^e m, m in [1/sqrt 2, sqrt 2) (frexp, exact);
//   log m = 2 atanh s = 2 s (1 + s^2/3 + s^4/5 + ... + s^18/19), s = (m - 1)/(m + 1), |s| < 0.172
//     (truncation below 2^-50);
//   1/(m + 1): the FP32 IEEE quotient refined by one Newton step in double (error ~2^-46);
//   log x = 2 s p + e ln 2.
// IEEE operations only, contraction off, so host and device compute the same bits.
WOB_HD inline float log_cr(float x) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    int e;
    float mf = __builtin_frexpf(x, &e);                       // mf in [0.5, 1)
    if (mf < 0.70710678f) {
        mf = 2.0f * mf;
        e -= 1;
    }
    const double m = (double)mf, t = m - 1.0, d = m + 1.0;
    double r = (double)(1.0f / (float)d);
    const double dr = d * r;
    r = r * (2.0 - dr);
    const double s = t * r, z = s * s;
    double p = 1.0 / 19.0;
    p = 1.0 / 17.0 + z * p;
    p = 1.0 / 15.0 + z * p;
    p = 1.0 / 13.0 + z * p;
    p = 1.0 / 11.0 + z * p;
    p = 1.0 / 9.0 + z * p;
    p = 1.0 / 7.0 + z * p;
    p = 1.0 / 5.0 + z * p;
    p = 1.0 / 3.0 + z * p;
    p = 1.0 + z * p;
    const double s2 = 2.0 * s;
    return (float)(s2 * p + (double)e * 0.69314718055994530942);
}

}  // namespace wobtrig
