// fp64_sincos.hpp -- sine and cosine in FP64 of a phase already wrapped to
// [-pi, pi], short enough to sit on a serial per-symbol recurrence (the RDS
// Costas loop).  Plain C++ so the same code is checked on the host against
// the C library.
//
// Method: Cody & Waite reduction by pi/2 with a two-part constant
// (q = rint(2x/pi) in {-2..2}; x - q*P1 is exact by Sterbenz's lemma and
// q*P2 is exact, so r carries one rounding), then the Taylor series of sin
// (to r^17) and cos (to r^18) on |r| <= pi/4, where the first omitted terms
// are below 1e-19, evaluated by Horner in FP64, and the quadrant rotation.
// Error about 1 ulp of FP64.  P1 = pi/2 rounded to double, P2 = the rounded
// remainder; the coefficients are 1/n! rounded to double.  Arguments with
// |x| > 4 are outside the reduction's exact range and go to the library.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define ARGUS_SC_HD __host__ __device__
#else
#define ARGUS_SC_HD
#endif

ARGUS_SC_HD inline void argus_sincos_wrapped(double x, double* sn, double* cs)
{
#pragma clang fp contract(off)
    if (!(fabs(x) <= 4.0)) {
        *sn = sin(x);
        *cs = cos(x);
        return;
    }
    const double P1 = 1.5707963267948966, P2 = 6.123233995736766e-17;
    const double q = rint(x * 0.6366197723675814);
    const double r = (x - q * P1) - q * P2;
    const double z = r * r;
    // sin r = r + r^3 (-1/3! + z (1/5! - z (1/7! - ...)))
    double ps = 2.8114572543455206e-15;
    ps = ps * z - 7.647163731819816e-13;
    ps = ps * z + 1.6059043836821613e-10;
    ps = ps * z - 2.505210838544172e-08;
    ps = ps * z + 2.7557319223985893e-06;
    ps = ps * z - 0.0001984126984126984;
    ps = ps * z + 0.008333333333333333;
    ps = ps * z - 0.16666666666666666;
    const double s = r + (r * z) * ps;
    // cos r = 1 - z/2 + z^2 (1/4! - z (1/6! - ...))
    double pc = -1.5619206968586225e-16;
    pc = pc * z + 4.779477332387385e-14;
    pc = pc * z - 1.1470745597729725e-11;
    pc = pc * z + 2.08767569878681e-09;
    pc = pc * z - 2.755731922398589e-07;
    pc = pc * z + 2.48015873015873e-05;
    pc = pc * z - 0.001388888888888889;
    pc = pc * z + 0.041666666666666664;
    const double c = (1.0 - 0.5 * z) + (z * z) * pc;
    switch (((int)q) & 3) {
    case 0: *sn = s; *cs = c; break;
    case 1: *sn = c; *cs = -s; break;
    case 2: *sn = -s; *cs = -c; break;
    default: *sn = -c; *cs = s; break;
    }
}
