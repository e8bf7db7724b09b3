// common.hpp -- shared device/host helpers for libargus_hip.so (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include "../../include/argus_gpu.h"

#define ARGUS_TWO_PI 6.283185307179586476925286766559

// ---------------------------------------------------------------------------
// NCO phase as a 64-bit fixed-point fraction of a turn (2^64 == one turn).
// Adding phase increments wraps modulo one turn exactly, so a long stream
// never drifts; radians in the C ABI are converted at the boundary only.
// ---------------------------------------------------------------------------
__host__ __device__ inline uint64_t argus_rad_to_fix(double rad)
{
    double t = rad / ARGUS_TWO_PI;
    t -= rint(t);            // (-0.5, 0.5]
    if (t >= 0.5) t -= 1.0;  // [-0.5, 0.5)
    return (uint64_t)(int64_t)(t * 18446744073709551616.0);
}

__host__ __device__ inline double argus_fix_to_rad(uint64_t fix)
{
    return (double)(int64_t)fix * (ARGUS_TWO_PI / 18446744073709551616.0);
}

// Reduce to [-pi, pi) the way the reference kernels do (exports.cu:526-529),
// then fold the rint tie (+pi) to -pi so the range is half-open.
__host__ __device__ inline double argus_reduce_phase(double phase)
{
    double r = phase - rint(phase / ARGUS_TWO_PI) * ARGUS_TWO_PI;
    if (r >= M_PI) r -= ARGUS_TWO_PI;
    return r;
}

// e^{j*2pi*turns(ph)} from the top 32 bits of a fixed-point phase.  The
// hardware sine/cosine take their argument in turns (v_sin_f32/v_cos_f32
// compute sin(2*pi*x)), so no radian scaling or range reduction is needed.
__device__ __forceinline__ void argus_nco(uint64_t ph, float& s, float& c)
{
    const float x = (float)(int32_t)(uint32_t)(ph >> 32) * 2.3283064365386963e-10f; // [-0.5, 0.5)
#if defined(__HIP_DEVICE_COMPILE__)
    s = __builtin_amdgcn_sinf(x);
    c = __builtin_amdgcn_cosf(x);
#else
    s = sinf(6.2831853f * x);
    c = cosf(6.2831853f * x);
#endif
}

__device__ __forceinline__ float2 argus_rot(float2 v, float s, float c)
{
    return make_float2(v.x * c - v.y * s, v.x * s + v.y * c);
}

// Every launch site clears the runtime's last-error slot before launching
// (argus_clear_error) and reads it right after, so it reports that launch's
// own configuration error and not an earlier failed call the caller has
// already seen.
static inline void argus_clear_error() { (void)hipGetLastError(); }

#define ARGUS_CHECK(expr)                                    \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return (int)_e;                \
    } while (0)

static inline int argus_max_i(int a, int b) { return a > b ? a : b; }
static inline int argus_min_i(int a, int b) { return a < b ? a : b; }
