// fft_core.hpp -- the in-register radix-R pieces of the R^3-point FFT
// (R^2 threads, R values each, two LDS transposes) shared by the
// surveillance transforms (surv.hip) and the fast-convolution extraction
// (fastconv.hip).
//
// Index n = a + R b + R^2 c; thread (a, b) holds the R samples c = 0..R-1:
//   pass 1  DFT_R over c, twiddle W_N^{(a + R b) kc}
//   LDS transpose -> thread (a, kc) holds b = 0..R-1
//   pass 2  DFT_R over b, twiddle W_{R^2}^{a kb}
//   LDS transpose (padded) -> thread (kb, kc) holds a = 0..R-1
//   pass 3  DFT_R over a  ->  X[kc + R kb + R^2 ka]
#pragma once
#include "common.hpp"
#include "fft_twiddles.hpp"

namespace argus_fftc {

template <int R>
__host__ __device__ constexpr int brev(int i)
{
    int r = 0;
    for (int b = 1; b < R; b <<= 1) {
        r <<= 1;
        if (i & b) r |= 1;
    }
    return r;
}

// In-register forward DFT of R points (radix-2, decimation in time).
template <int R>
__device__ __forceinline__ void dft(float2 (&v)[R])
{
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const int j = brev<R>(i);
        if (i < j) {
            const float2 t = v[i];
            v[i] = v[j];
            v[j] = t;
        }
    }
#pragma unroll
    for (int len = 2; len <= R; len <<= 1) {
        const int half = len >> 1;
#pragma unroll
        for (int i = 0; i < R; i += len) {
#pragma unroll
            for (int k = 0; k < half; ++k) {
                const float2 a = v[i + k], b = v[i + k + half];
                float2 tb;
                if (k == 0) {
                    tb = b;
                } else if (4 * k == len) {  // W = -i
                    tb = make_float2(b.y, -b.x);
                } else {
                    const int wi = k * (32 / len);
                    const float cr = argus_fft::W32_RE[wi], ci = argus_fft::W32_IM[wi];
                    tb = make_float2(b.x * cr - b.y * ci, b.x * ci + b.y * cr);
                }
                v[i + k] = make_float2(a.x + tb.x, a.y + tb.y);
                v[i + k + half] = make_float2(a.x - tb.x, a.y - tb.y);
            }
        }
    }
}

// e^{-2 pi i e / M}, e in [0, M), M a power of two: exact turn argument.
__device__ __forceinline__ float2 twiddle(int e, float inv_m)
{
    const float x = -(float)e * inv_m;
    return make_float2(__builtin_amdgcn_cosf(x), __builtin_amdgcn_sinf(x));
}

__device__ __forceinline__ float2 cmul(float2 a, float2 w)
{
    return make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}

// Transpose through LDS: thread writes its R values to wr(k), barrier, reads
// R values from rd(k); real and imaginary parts go in two rounds.
template <int R, typename WR, typename RD>
__device__ __forceinline__ void exchange(float2 (&v)[R], float* xs, WR wr, RD rd)
{
    // each half is read back into the registers it was written from, so the
    // transpose needs no extra registers
#pragma unroll
    for (int k = 0; k < R; ++k) xs[wr(k)] = v[k].x;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R; ++k) v[k].x = xs[rd(k)];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R; ++k) xs[wr(k)] = v[k].y;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R; ++k) v[k].y = xs[rd(k)];
    __syncthreads();
}

// Full N = R^3 forward transform of the values each of R^2 threads holds
// (thread t = a + R b holds c = 0..R-1; t is the thread's index among the
// R^2 that share xs, all of which must reach the same barriers).  On
// return thread t = kb + R kc holds X[kc + R kb + R^2 ka] in v[ka].
template <int R>
__device__ __forceinline__ void fft_r3(float2 (&v)[R], float* xs, int t)
{
    constexpr int TH = R * R;
    constexpr float invN = 1.0f / (float)(R * R * R);
    constexpr float invRR = 1.0f / (float)(R * R);
    // opaque to the optimiser: keeps the twiddles (a function of t only)
    // from being hoisted out of a caller's segment loop into scratch memory
    asm volatile("" : "+v"(t));
    dft<R>(v);
#pragma unroll
    for (int k = 1; k < R; ++k) v[k] = cmul(v[k], twiddle(t * k, invN));
    {
        const int a = t % R, kc = t / R;
        exchange<R>(v, xs, [&](int k) { return t + TH * k; }, [&](int b) { return a + R * b + TH * kc; });
    }
    dft<R>(v);
    {
        const int a = t % R, kc = t / R;
#pragma unroll
        for (int k = 1; k < R; ++k) v[k] = cmul(v[k], twiddle(a * k, invRR));
        // padded layout a + (R+1) kb + R(R+1) kc: odd strides both ways, so
        // neither the writes (lanes vary a) nor the reads (lanes vary kb)
        // conflict, and every address is affine in the unrolled index
        const int kb2 = t % R, kc2 = t / R;
        exchange<R>(
            v, xs, [&](int kb) { return a + (R + 1) * kb + R * (R + 1) * kc; },
            [&](int aa) { return aa + (R + 1) * kb2 + R * (R + 1) * kc2; });
    }
    dft<R>(v);
}

// Twiddle tables of the R^3-point transform, built once per workgroup in
// LDS: t1[x][k] = W_N^{x k} and t2[x][k] = W_{R^2}^{x k} (x, k < R), rows
// R + 1 apart so lanes reading different rows hit different banks.  Pass 1's
// W_N^{(a + R b) k} is t1[a][k] t2[b][k]; pass 2's W_{R^2}^{a k} is t2[a][k].
template <int R>
constexpr int fft_tab_entries()
{
    return R * (R + 1);
}

template <int R>
__device__ __forceinline__ void fft_tab_init(float2* t1, float2* t2, int tid, int nthreads)
{
    constexpr float invN = 1.0f / (float)(R * R * R);
    constexpr float invRR = 1.0f / (float)(R * R);
    for (int i = tid; i < R * R; i += nthreads) {
        const int x = i / R, k = i - x * R;
        t1[x * (R + 1) + k] = twiddle(x * k, invN);
        t2[x * (R + 1) + k] = twiddle(x * k, invRR);
    }
}

// fft_r3 with the twiddles read from the tables (fft_tab_init, after a
// barrier) instead of evaluated: two LDS reads and a complex product for
// pass 1, one read for pass 2, against a sine and a cosine (quarter-rate
// instructions) per twiddle.
template <int R>
__device__ __forceinline__ void fft_r3_tab(float2 (&v)[R], float* xs, int t, const float2* t1, const float2* t2)
{
    constexpr int TH = R * R;
    asm volatile("" : "+v"(t));  // keeps the table reads (functions of t) in the caller's loop
    dft<R>(v);
    {
        const int a = t % R, b = t / R;
        const float2* r1 = t1 + a * (R + 1);
        const float2* r2 = t2 + b * (R + 1);
#pragma unroll
        for (int k = 1; k < R; ++k) v[k] = cmul(v[k], cmul(r1[k], r2[k]));
    }
    {
        const int a = t % R, kc = t / R;
        exchange<R>(v, xs, [&](int k) { return t + TH * k; }, [&](int b) { return a + R * b + TH * kc; });
    }
    dft<R>(v);
    {
        const int a = t % R, kc = t / R;
        const float2* r2 = t2 + a * (R + 1);
#pragma unroll
        for (int k = 1; k < R; ++k) v[k] = cmul(v[k], r2[k]);
        const int kb2 = t % R, kc2 = t / R;
        exchange<R>(
            v, xs, [&](int kb) { return a + (R + 1) * kb + R * (R + 1) * kc; },
            [&](int aa) { return aa + (R + 1) * kb2 + R * (R + 1) * kc2; });
    }
    dft<R>(v);
}

// LDS floats the transforms need: the padded R x (R+1) x R image.
template <int R>
constexpr int fft_lds_floats()
{
    return R * R * (R + 1);
}

}  // namespace argus_fftc
