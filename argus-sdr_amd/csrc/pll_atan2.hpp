// pll_atan2.hpp -- the FP32 atan2 of the stereo PLL's phase detector, as a
// template over the reciprocal (the device passes the hardware reciprocal,
// a host check 1/x), so the same code is checked on the host against atan2.
//
// Method: the octant ratio a = min(|x|, |y|) / max(|x|, |y|) (the
// denominator floored at the smallest normal, so 0/0 gives a = 0);
// atan(a) on [0, 1] as a + a s P(s) with s = a^2 and P of degree 6 in
// Horner form, its coefficients from a weighted least-squares fit to atan
// made for this kernel (1.2e-7 rad max error in FP32); then the octant and
// quadrant reflections and the sign of y.  3.2e-7 rad max error over all
// quadrants; atan2(+-0, +0) = +-0 and atan2(+-0, -0) = +-pi as in IEEE
// atan2.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define ARGUS_AT_HD __host__ __device__
#else
#define ARGUS_AT_HD
#endif

template <typename Rcp>
ARGUS_AT_HD inline float argus_pll_atan2f(float y, float x, Rcp rcp)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(fmaxf(ax, ay), 1.17549435e-38f), mn = fminf(ax, ay);
    const float a = mn * rcp(mx);
    const float s = a * a;
    float q = -4.355480428785086e-3f;
    q = fmaf(q, s, 2.3040393367409706e-2f);
    q = fmaf(q, s, -5.7773932814598083e-2f);
    q = fmaf(q, s, 9.794257581233978e-2f);
    q = fmaf(q, s, -1.397659033536911e-1f);
    q = fmaf(q, s, 1.9962705671787262e-1f);
    q = fmaf(q, s, -3.333165943622589e-1f);
    float r = fmaf(a * s, q, a);
    r = ay > ax ? 1.5707963267948966f - r : r;
    r = signbit(x) ? 3.1415926535897932f - r : r;
    return copysignf(r, y);
}
