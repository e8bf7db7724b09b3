// fastconv.hip -- overlap-save fast convolution for the dense channels of the
// streaming service: the same filter as the direct polyphase form, computed
// through one shared transform of the input instead of per-channel staging.
//
// For a channel with NCO phase phi(n) = phase0 + inc n and taps h (T <= 512),
//     y[m] = sum_k h[k] s[e_m - k],  s[n] = x[n] e^{j 2pi phi(n)}
//          = e^{j 2pi phi(e_m)} sum_k g[k] x[e_m - k],  g[k] = h[k] e^{-j 2pi inc k},
// the convolution of the raw input with the channel's modulated taps, then
// one rotation per output (the factorisation the narrow items use too).
// The raw input is shared by every channel, so its transform is too:
//
//   stage   u = the stream's last 511 samples ++ the block ++ zeros
//   FFT     X_s = FFT_4096(u[s*3584 .. s*3584 + 4096)), s < S = ceil(n/3584)
//           (one launch of the surveillance 4096-point kernel for all s)
//   channel G = FFT_4096(g) / 4096 (host, FP64, once per channel), times the
//           decimation-phase ramp e^{j 2pi k r / 4096} (a shift by r samples);
//           per segment, decimation by Dp (the power of two dividing D,
//           at most 64) in the frequency domain (folding),
//           F[j] = sum_{i < Dp} X_s[j + i J] G'[j + i J],  J = 4096 / Dp,
//           and an inverse transform of J points gives the segment's
//           convolution at u-positions q Dp + r, q < J; the outputs are the
//           positions among them that are emission points (every D/Dp-th).
//           Positions [511, 511 + 3584) are valid for T <= 512
//           (overlap-save), so each block sample belongs to exactly one
//           segment.
//   tail    the stream's last 511 samples after the block.
//
// The inverse transform is the 512-point FFT (8^3) of the conjugated folded
// spectrum, zero-stuffed by 512 / J for D > 8 (stuffing the spectrum repeats
// the signal, so its first J outputs are the J-point transform).  Each
// workgroup takes one channel: G' in LDS (32 KiB), its four wavefronts take
// four segments at a time.
//
// The last 511 samples of the stream are the channel's own history only once
// the channel has seen T - 1 samples (a channel opened mid-stream starts
// from zero history), so the service routes a channel here from its first
// block with a full history; the direct form keeps writing its shifted
// history, so either form can follow the other.
//
// Arithmetic: FP32 transforms (about 12 rounding steps on each side), FP64
// taps and G.  The outputs differ from the direct form's in the last bits
// of the in-band signal only (relative RMS ~1e-6).
#include "fastconv.hpp"
#include "fft_core.hpp"

#include <cmath>
#include <complex>
#include <vector>

namespace {

using namespace argus_fftc;

constexpr int kN = ARGUS_FAST_N;
constexpr int kHop = ARGUS_FAST_HOP;
constexpr int kTail = ARGUS_FAST_TAIL;
constexpr int kJ = 512;  // the inverse transform: 8^3 points
constexpr int kThreads = 256;
static_assert(kHop % 64 == 0 && kTail + kHop <= kN, "segment geometry");

// u[i], i < (S-1)*hop + N: tail, then the block, then zeros.
__global__ __launch_bounds__(kThreads) void argus_fast_stage_kernel(const ArgusFastHdr* __restrict__ hdr,
                                                                    const ArgusExtractHdr* __restrict__ xh)
{
    const int S = hdr->S;
    if (S <= 0) return;
    const int len = (S - 1) * kHop + kN;
    const int n = xh->n_new;
    const float2* in = xh->in;
    const float2* tail = hdr->tail;
    float2* u = hdr->u;
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < len; i += gridDim.x * kThreads) {
        const int b = i - kTail;
        u[i] = b < 0 ? tail[i] : (b < n ? in[b] : make_float2(0.0f, 0.0f));
    }
}

// F[j] = sum_{i < DP} X[j + i J] G'[j + i J] for this lane's j = t + 64 c,
// c < J / 64: every lane folds the same number of bins, all loads of one
// i issued together.
template <int DP>
__device__ __forceinline__ void fast_fold(const float2* __restrict__ xseg, const float2* gs, int t,
                                          float2 (&f)[ARGUS_FAST_N / DP / 64])
{
    constexpr int J = ARGUS_FAST_N / DP, PER = J / 64;
    float ar[PER], ai[PER];
#pragma unroll
    for (int c = 0; c < PER; c++) ar[c] = ai[c] = 0.0f;
#pragma unroll
    for (int i = 0; i < DP; i++) {
        float2 xv[PER];
#pragma unroll
        for (int c = 0; c < PER; c++) xv[c] = xseg[i * J + t + 64 * c];
#pragma unroll
        for (int c = 0; c < PER; c++) {
            const float2 gv = gs[i * J + t + 64 * c];
            ar[c] += xv[c].x * gv.x - xv[c].y * gv.y;
            ai[c] += xv[c].x * gv.y + xv[c].y * gv.x;
        }
    }
#pragma unroll
    for (int c = 0; c < PER; c++) f[c] = make_float2(ar[c], ai[c]);
}

// One segment of one channel on one wavefront: fold, inverse transform,
// outputs.  All four wavefronts of the workgroup call this the same number
// of times (the transform's barriers are workgroup-wide); act = 0 does no
// work but takes the barriers.
template <int DP>
__device__ __forceinline__ void fast_segment(const ArgusFastDesc& d, const float2* __restrict__ X, const float2* gs,
                                             float* xs, int s, bool act, int t)
{
    constexpr int J = ARGUS_FAST_N / DP, PER = J / 64, Z = kJ / J;  // Z: zero-stuffing of the 512-point input
    float2 v[8];
    if constexpr (Z == 1) {
        fast_fold<DP>(X + (size_t)(act ? s : 0) * kN, gs, t, v);
#pragma unroll
        for (int c = 0; c < 8; c++) v[c].y = -v[c].y;  // conjugate: the forward transform inverts
    } else {
        float2 f[PER];
        fast_fold<DP>(X + (size_t)(act ? s : 0) * kN, gs, t, f);
        float2* fs = (float2*)xs;  // J <= 256 bins fit the transform's scratch
#pragma unroll
        for (int c = 0; c < PER; c++) fs[t + 64 * c] = f[c];
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int nn = t + 64 * c;
            const float2 y = (nn % Z) == 0 ? fs[nn / Z] : make_float2(0.0f, 0.0f);
            v[c] = make_float2(y.x, -y.y);
        }
        __syncthreads();
    }
    fft_r3<8>(v, xs, t);
    // thread t = kb + 8 kc holds Y[kc + 8 kb + 64 ka] in v[ka]
    const int kb = t & 7, kc = t >> 3;
#pragma unroll
    for (int ka = 0; ka < 8; ka++) {
        const int q = kc + 8 * kb + 64 * ka;
        if (!act || q >= J) continue;
        const int iseg = q * DP + d.r;
        if (iseg < kTail || iseg >= kTail + kHop) continue;
        const int e = s * kHop + iseg - kTail;  // block sample, = e0 mod Dp
        const int de = e - d.e0;
        if (de < 0 || de % d.D != 0) continue;  // not an emission point (D > DP)
        const int m = de / d.D;
        if (m >= d.n_out) continue;
        const float2 y = make_float2(v[ka].x, -v[ka].y);
        float sn, cs;
        argus_nco(d.phase0 + d.inc * (uint64_t)(int64_t)e, sn, cs);
        d.out[m] = argus_rot(y, sn, cs);
    }
}

template <int DP>
__device__ __forceinline__ void fast_channel(const ArgusFastDesc& d, const float2* __restrict__ X, const float2* gs,
                                             float* xs, int S, int w, int t)
{
    for (int s0 = 0; s0 < S; s0 += 4) fast_segment<DP>(d, X, gs, xs, s0 + w, s0 + w < S, t);
}

__global__ __launch_bounds__(kThreads) void argus_fast_chan_kernel(const ArgusFastHdr* __restrict__ hdr)
{
    __shared__ float2 gs[kN];
    __shared__ float xs[4][fft_lds_floats<8>()];
    static_assert(fft_lds_floats<8>() >= 2 * 256, "folded bins of D >= 16 fit the transform scratch");
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int S = hdr->S, n_fast = hdr->n_fast;
    const float2* X = hdr->X;
    for (int c = blockIdx.x; c < n_fast; c += gridDim.x) {
        const ArgusFastDesc d = hdr->descs[c];
        // G with the ramp e^{j 2pi k r / N}: exact turn fractions
        for (int k = threadIdx.x; k < kN; k += kThreads) {
            const float2 g = d.G[k];
            const float x = (float)((k * d.r) & (kN - 1)) * (1.0f / kN);
            const float sn = __builtin_amdgcn_sinf(x), cs = __builtin_amdgcn_cosf(x);
            gs[k] = make_float2(g.x * cs - g.y * sn, g.x * sn + g.y * cs);
        }
        __syncthreads();
        switch (d.Dp) {  // uniform over the workgroup (one channel)
        case 8: fast_channel<8>(d, X, gs, xs[w], S, w, t); break;
        case 16: fast_channel<16>(d, X, gs, xs[w], S, w, t); break;
        case 32: fast_channel<32>(d, X, gs, xs[w], S, w, t); break;
        default: fast_channel<64>(d, X, gs, xs[w], S, w, t); break;
        }
        __syncthreads();  // gs is rewritten for the next channel
    }
}

// tail' = the last 511 samples of tail ++ block.
__global__ __launch_bounds__(512) void argus_fast_tail_kernel(const ArgusFastHdr* __restrict__ hdr,
                                                              const ArgusExtractHdr* __restrict__ xh)
{
    const int i = threadIdx.x;
    const int n = xh->n_new;
    float2* tail = hdr->tail;
    float2 v = make_float2(0.0f, 0.0f);
    if (i < kTail) v = i + n >= kTail ? xh->in[n - kTail + i] : tail[i + n];
    __syncthreads();  // every read of the old tail before any write
    if (i < kTail) tail[i] = v;
}

// In-place radix-2 FFT of ARGUS_FAST_N points in FP64 (host), forward; the
// twiddles e^{-2 pi i k / N} are computed once.
void fft64(std::vector<std::complex<double>>& a)
{
    static const std::vector<std::complex<double>> tw = [] {
        std::vector<std::complex<double>> w((size_t)kN / 2);
        for (int k = 0; k < kN / 2; k++) {
            const double ang = -2.0 * M_PI * (double)k / (double)kN;
            w[(size_t)k] = std::complex<double>(std::cos(ang), std::sin(ang));
        }
        return w;
    }();
    const int n = kN;
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[(size_t)i], a[(size_t)j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < len / 2; k++) {
                const std::complex<double> u = a[(size_t)(i + k)], v = a[(size_t)(i + k + len / 2)] * tw[(size_t)(k * step)];
                a[(size_t)(i + k)] = u + v;
                a[(size_t)(i + k + len / 2)] = u - v;
            }
        }
    }
}

}  // namespace

void argus_fast_spectrum(const float* h, int T, uint64_t inc, float2* G)
{
    std::vector<std::complex<double>> a((size_t)kN, 0.0);
    for (int k = 0; k < T && k < kN; k++) {
        // turn fraction of inc * k modulo one turn, exactly, then FP64
        const double turns = (double)(int64_t)(inc * (uint64_t)k) * (1.0 / 18446744073709551616.0);
        const double ang = -2.0 * M_PI * turns;
        a[(size_t)k] = (double)h[k] * std::complex<double>(std::cos(ang), std::sin(ang));
    }
    fft64(a);
    for (int k = 0; k < kN; k++) {
        G[k].x = (float)(a[(size_t)k].real() / kN);
        G[k].y = (float)(a[(size_t)k].imag() / kN);
    }
}

int argus_launch_fast(const ArgusFastHdr* d_hdr, const ArgusExtractHdr* d_xh, int S, float2* u, float2* X,
                      hipStream_t stream)
{
    if (S <= 0) return 0;
    const int len = (S - 1) * kHop + kN;
    argus_clear_error();
    argus_fast_stage_kernel<<<(len + kThreads - 1) / kThreads, kThreads, 0, stream>>>(d_hdr, d_xh);
    ARGUS_CHECK(hipGetLastError());
    int err = argus_launch_fft4096(u, kHop, S, X, stream);
    if (err) return err;
    argus_fast_chan_kernel<<<ARGUS_FAST_GRID, kThreads, 0, stream>>>(d_hdr);
    return (int)hipGetLastError();
}

int argus_launch_fast_tail(const ArgusFastHdr* d_hdr, const ArgusExtractHdr* d_xh, hipStream_t stream)
{
    argus_clear_error();
    argus_fast_tail_kernel<<<1, 512, 0, stream>>>(d_hdr, d_xh);
    return (int)hipGetLastError();
}

// Host helper: argus_fast_spectrum for a phase increment in radians per
// sample, converted as the service converts a channel's offset.
extern "C" GPUD_API int gpud_fast_spectrum(const float* taps, int num_taps, double phase_inc, gpud_float2* G)
{
    if (!taps || !G || num_taps <= 0 || num_taps > ARGUS_FAST_TAIL + 1) return -1;
    argus_fast_spectrum(taps, num_taps, argus_rad_to_fix(phase_inc), (float2*)G);
    return 0;
}
