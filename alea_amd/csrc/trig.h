// sin and cos of one double for the bounded arguments of the walker sweep
// (k.r with |k.r| <~ 1e4 for positions wrapped into the cell), without the
// large-argument reduction machinery of a general libm sincos, whose register
// footprint would cap the sweep kernel's occupancy.
//
// Method (Cody & Waite, "Software Manual for the Elementary Functions",
// 1980): x = q pi/2 + r with q = rint(x 2/pi) and pi/2 split into three
// doubles of 33 significant bits (computed in exact decimal arithmetic), so
// q*P1 and q*P2 are exact for |q| < 2^20; then the Taylor series of sin and
// cos on |r| <= pi/4 (degree 15 and 16: truncation below 5e-17 and 2e-18,
// under half an ulp) and the quadrant swap.  Accurate to a few ulp for
// |x| < 1e6; beyond that the reduction degrades (tested against libm in
// tests/test_trig.py).
//
// In device code each coefficient is pinned to scalar registers where it is
// used (kc): hoisted out of the sweep's loops, the seventeen FP64 constants
// would hold 34 vector registers for the whole kernel.
#pragma once

#if defined(__HIPCC__)
#define ALEA_HD __host__ __device__
#else
#define ALEA_HD
#endif

#include <cmath>

namespace alea_amd {

ALEA_HD inline double kc(double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(c));
#endif
    return c;
}

ALEA_HD inline void sincos_bounded(double x, double* s_out, double* c_out) {
    constexpr double kTwoOverPi = 0.6366197723675814;
    constexpr double kP1 = 1.5707963267341256;
    constexpr double kP2 = 6.077100506303966e-11;
    constexpr double kP3 = 2.0222662487959506e-21;
    const double q = rint(x * kc(kTwoOverPi));
    double r = x - q * kc(kP1);
    r = r - q * kc(kP2);
    r = r - q * kc(kP3);
    const double r2 = r * r;
    // sin r = r + r^3 P(r^2), P(z) = -1/3! + z/5! - z^2/7! + ... - z^6/15!
    double ps = kc(-1.0 / 1307674368000.0);          // -1/15!
    ps = ps * r2 + kc(1.0 / 6227020800.0);           //  1/13!
    ps = ps * r2 - kc(1.0 / 39916800.0);             // -1/11!
    ps = ps * r2 + kc(1.0 / 362880.0);               //  1/9!
    ps = ps * r2 - kc(1.0 / 5040.0);                 // -1/7!
    ps = ps * r2 + kc(1.0 / 120.0);                  //  1/5!
    ps = ps * r2 - kc(1.0 / 6.0);                    // -1/3!
    const double sr = r + r * r2 * ps;
    // cos r = 1 - r^2/2! + r^4/4! - ... + r^16/16!
    double pc = kc(1.0 / 20922789888000.0);          //  1/16!
    pc = pc * r2 - kc(1.0 / 87178291200.0);          // -1/14!
    pc = pc * r2 + kc(1.0 / 479001600.0);            //  1/12!
    pc = pc * r2 - kc(1.0 / 3628800.0);              // -1/10!
    pc = pc * r2 + kc(1.0 / 40320.0);                //  1/8!
    pc = pc * r2 - kc(1.0 / 720.0);                  // -1/6!
    pc = pc * r2 + kc(1.0 / 24.0);                   //  1/4!
    pc = pc * r2 - 0.5;                              // -1/2!
    const double cr = 1.0 + r2 * pc;
    const long long qi = static_cast<long long>(q);
    const int quad = static_cast<int>(qi & 3);
    double s = sr, c = cr;
    if (quad == 1) { s = cr; c = -sr; }
    else if (quad == 2) { s = -sr; c = -cr; }
    else if (quad == 3) { s = -cr; c = sr; }
    *s_out = s;
    *c_out = c;
}

// Single precision: sin and cos of x = k.r for single-precision orbital
// values.  The reduction x = q pi/2 + r stays in double (|x| reaches ~1e2
// and beyond), the polynomials on |r| <= pi/4 are float: sin to r^9
// (truncation 7e-10), cos to r^8 (2.5e-8), below one float ulp.
ALEA_HD inline float kcf(float c) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(c));
#endif
    return c;
}

ALEA_HD inline void sincos_bounded_f32(double x, float* s_out, float* c_out) {
    const double q = rint(x * kc(0.6366197723675814));
    const double rd = (x - q * kc(1.5707963267341256)) - q * kc(6.077100506303966e-11);
    const float r = static_cast<float>(rd);
    const float r2 = r * r;
    float ps = kcf(1.0f / 362880.0f);              //  1/9!
    ps = ps * r2 - kcf(1.0f / 5040.0f);            // -1/7!
    ps = ps * r2 + kcf(1.0f / 120.0f);             //  1/5!
    ps = ps * r2 - kcf(1.0f / 6.0f);               // -1/3!
    const float sr = r + r * r2 * ps;
    float pc = kcf(1.0f / 40320.0f);               //  1/8!
    pc = pc * r2 - kcf(1.0f / 720.0f);             // -1/6!
    pc = pc * r2 + kcf(1.0f / 24.0f);              //  1/4!
    pc = pc * r2 - 0.5f;                           // -1/2!
    const float cr = 1.0f + r2 * pc;
    const int quad = static_cast<int>(static_cast<long long>(q) & 3);
    float s = sr, c = cr;
    if (quad == 1) { s = cr; c = -sr; }
    else if (quad == 2) { s = -sr; c = -cr; }
    else if (quad == 3) { s = -cr; c = sr; }
    *s_out = s;
    *c_out = c;
}

}  // namespace alea_amd
