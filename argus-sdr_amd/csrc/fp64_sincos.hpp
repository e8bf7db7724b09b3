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

// The constants, as a value a caller may keep in registers (a device loop
// pins them in VGPRs once rather than rebuilding them per call).
struct ArgusSinCosK {
    double inv_half_pi, p1, p2;
    double s[8];  // sin: coefficients of z^7 .. z^0 in the series after r + r^3 (...)
    double c[8];  // cos: coefficients of z^7 .. z^0 in the series after z^2 (...)
};

ARGUS_SC_HD constexpr ArgusSinCosK argus_sincos_k()
{
    return ArgusSinCosK{0.6366197723675814,
                        1.5707963267948966,
                        6.123233995736766e-17,
                        {2.8114572543455206e-15, -7.647163731819816e-13, 1.6059043836821613e-10,
                         -2.505210838544172e-08, 2.7557319223985893e-06, -0.0001984126984126984,
                         0.008333333333333333, -0.16666666666666666},
                        {-1.5619206968586225e-16, 4.779477332387385e-14, -1.1470745597729725e-11,
                         2.08767569878681e-09, -2.755731922398589e-07, 2.48015873015873e-05,
                         -0.001388888888888889, 0.041666666666666664}};
}

// Every step rounds separately (no contraction); adding a negative
// coefficient rounds as subtracting its magnitude.
ARGUS_SC_HD inline void argus_sincos_wrapped_k(double x, const ArgusSinCosK& k, double* sn, double* cs)
{
#pragma clang fp contract(off)
    if (!(fabs(x) <= 4.0)) {
        *sn = sin(x);
        *cs = cos(x);
        return;
    }
    const double q = rint(x * k.inv_half_pi);
    const double r = (x - q * k.p1) - q * k.p2;
    const double z = r * r;
    // sin r = r + r^3 (-1/3! + z (1/5! - z (1/7! - ...)))
    double ps = k.s[0];
    for (int i = 1; i < 8; i++) ps = ps * z + k.s[i];
    const double s = r + (r * z) * ps;
    // cos r = 1 - z/2 + z^2 (1/4! - z (1/6! - ...))
    double pc = k.c[0];
    for (int i = 1; i < 8; i++) pc = pc * z + k.c[i];
    const double c = (1.0 - 0.5 * z) + (z * z) * pc;
    // quadrant (q mod 4): (s, c), (c, -s), (-s, -c), (-c, s), as selects
    const int qi = (int)q;
    const double a = (qi & 1) ? c : s, b = (qi & 1) ? s : c;
    *sn = (qi & 2) ? -a : a;
    *cs = ((qi + 1) & 2) ? -b : b;
}

ARGUS_SC_HD inline void argus_sincos_wrapped(double x, double* sn, double* cs)
{
    argus_sincos_wrapped_k(x, argus_sincos_k(), sn, cs);
}
