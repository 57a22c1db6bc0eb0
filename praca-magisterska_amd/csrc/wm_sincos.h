// sin and cos of the fill's phase arguments k R (0 <= x < kSinCosMax).
//
// Cody & Waite argument reduction (W. J. Cody, W. Waite, "Software Manual for
// the Elementary Functions", 1980): n = rint(x 2/pi), r = x - n pi/2 with
// pi/2 split into three parts, P1 and P2 of 33 significant bits (so n P1 and
// n P2 are exact for n < 2^20) and P3 the remainder, each subtracted with a
// fused multiply-add; then sin r and cos r on |r| <= pi/4 (a little beyond
// after rounding n) from their Taylor series, through r^17 and r^16 (the
// first omitted terms are below 1e-19 and 3e-18), and the quadrant n mod 4.
// Every rounding is written out, so host and device builds give the same bits.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define WM_HD __host__ __device__
#else
#define WM_HD
#endif

namespace wm {

constexpr double kSinCosMax = 1.6e6;  // below 2^20 pi/2: n P1 and n P2 stay exact

WM_HD inline void sincos_phase(double x, double &s, double &c)
{
    constexpr double kTwoOverPi = 0.6366197723675814;
    constexpr double kP1 = 1.5707963267341256, kP2 = 6.077100506303966e-11, kP3 = 2.0222662487959506e-21;
    const double n = rint(x * kTwoOverPi);
    double r = fma(-n, kP1, x);
    r = fma(-n, kP2, r);
    r = fma(-n, kP3, r);
    const double r2 = r * r;
    // sin r = r + r^3 (-1/3! + r^2 (1/5! - ...)),  cos r = 1 + r^2 (-1/2! + r^2 (1/4! - ...))
    double ps = 2.8114572543455206e-15;                 // 1/17!
    ps = fma(ps, r2, -7.647163731819816e-13);           // -1/15!
    ps = fma(ps, r2, 1.6059043836821613e-10);           // 1/13!
    ps = fma(ps, r2, -2.505210838544172e-08);           // -1/11!
    ps = fma(ps, r2, 2.7557319223985893e-06);           // 1/9!
    ps = fma(ps, r2, -0.0001984126984126984);           // -1/7!
    ps = fma(ps, r2, 0.008333333333333333);             // 1/5!
    ps = fma(ps, r2, -0.16666666666666666);             // -1/3!
    const double sr = fma(r * r2, ps, r);
    double pc = 4.779477332387385e-14;                  // 1/16!
    pc = fma(pc, r2, -1.1470745597729725e-11);          // -1/14!
    pc = fma(pc, r2, 2.08767569878681e-09);             // 1/12!
    pc = fma(pc, r2, -2.755731922398589e-07);           // -1/10!
    pc = fma(pc, r2, 2.48015873015873e-05);             // 1/8!
    pc = fma(pc, r2, -0.001388888888888889);            // -1/6!
    pc = fma(pc, r2, 0.041666666666666664);             // 1/4!
    pc = fma(pc, r2, -0.5);                             // -1/2!
    const double cr = fma(r2, pc, 1.0);
    // quadrant q: (sin, cos) = (s, c), (c, -s), (-s, -c), (-c, s) for q = 0..3
    const int q = (int)n & 3;
    const double a = (q & 1) ? cr : sr, b = (q & 1) ? sr : cr;
    s = (q & 2) ? -a : a;
    c = ((q + 1) & 2) ? -b : b;
}

}  // namespace wm
