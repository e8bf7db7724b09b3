// audio.hip -- post-demodulation audio stage kernels (layout and semantics
// in audio.hpp).
//
//   argus_audio_lpf_kernel  tiles of the plain-WFM 15 kHz low-pass (FP32,
//                           101 taps, delay line carried, tail item per channel)
//   argus_audio_kernel      one workgroup per channel: every output of the
//                           block's resampling computed independently (FP64
//                           arms and accumulation, as the reference), in
//                           chunks held in LDS; de-emphasis as a parallel
//                           scan of affine maps y -> a y + b in FP64; gain;
//                           the next 33-sample delay line per side.
#include "audio.hpp"

#include <cmath>
#include <vector>

namespace {

// resample.go:280-292 modified Bessel I0 by its series (stops below 1e-12 of the sum)
double bessel_i0(double x)
{
    double sum = 1.0, term = 1.0;
    for (int k = 1; k < 30; k++) {
        const double h = x / (2.0 * k);
        term *= h * h;
        sum += term;
        if (term < 1e-12 * sum) break;
    }
    return sum;
}

int gcd_int(int a, int b)
{
    while (b != 0) {
        const int t = a % b;
        a = b;
        b = t;
    }
    return a < 0 ? -a : a;
}

}  // namespace

__global__ __launch_bounds__(256) void argus_audio_lpf_kernel(const ArgusAudioHdr* __restrict__ hdr)
{
    __shared__ float xs[ARGUS_AUDIO_LPF_TILE + ARGUS_AUDIO_LPF_HIST];
    __shared__ float hs[ARGUS_AUDIO_LPF_TAPS];
    const ArgusAudioHdr h = *hdr;
    for (int w = blockIdx.x; w < h.n_items; w += gridDim.x) {
        const ArgusAudioItem it = h.items[w];
        const ArgusAudioDesc& d = h.descs[it.chan];
        const int n = d.n_in;
        auto at = [&](int idx) -> float {
            return idx >= 0 ? d.lpf_in[idx] : d.lpf_hist_rd[ARGUS_AUDIO_LPF_HIST + idx];
        };
        if (it.kind == 1) {  // the next delay line: the last 100 inputs
            for (int j = threadIdx.x; j < ARGUS_AUDIO_LPF_HIST; j += blockDim.x)
                d.lpf_hist_wr[j] = at(n - ARGUS_AUDIO_LPF_HIST + j);
            continue;
        }
        const int i0 = it.start;
        const int cnt = min(ARGUS_AUDIO_LPF_TILE, n - i0);
        const int span = cnt + ARGUS_AUDIO_LPF_HIST;
        for (int j = threadIdx.x; j < ARGUS_AUDIO_LPF_TAPS; j += blockDim.x) hs[j] = d.lpf_taps[j];
        for (int j = threadIdx.x; j < span; j += blockDim.x) xs[j] = at(i0 - ARGUS_AUDIO_LPF_HIST + j);
        __syncthreads();
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
            const float* xp = xs + i + ARGUS_AUDIO_LPF_HIST;
            float a = 0.0f;
#pragma unroll 4
            for (int k = 0; k < ARGUS_AUDIO_LPF_TAPS; k++) a = fmaf(hs[k], xp[-k], a);
            d.lpf_out[i0 + i] = a;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void argus_audio_kernel(const ArgusAudioHdr* __restrict__ hdr)
{
    __shared__ float ys[2][ARGUS_AUDIO_CHUNK];
    __shared__ double sa[256], sb[256];
    const ArgusAudioHdr h = *hdr;
    const int t = threadIdx.x;
    for (int c = blockIdx.x; c < h.n_chans; c += gridDim.x) {
        const ArgusAudioDesc d = h.descs[c];
        const int nch = d.channels, n = d.n_in;
        if (n <= 0) continue;  // an empty block leaves the stage's state as it is
        auto x_at = [&](int side, int idx) -> float {  // block input, idx >= -ARGUS_AUDIO_ARM
            return idx >= 0 ? d.in[(size_t)idx * nch + side]
                            : d.hist_rd[side * ARGUS_AUDIO_ARM + ARGUS_AUDIO_ARM + idx];
        };
        double y[2] = {0.0, 0.0};
        if (d.deemph_on)
            for (int s = 0; s < nch; s++) y[s] = d.deemph[s];
        for (int k0 = 0; k0 < d.n_out; k0 += ARGUS_AUDIO_CHUNK) {
            const int cnt = min(ARGUS_AUDIO_CHUNK, d.n_out - k0);
            for (int kk = t; kk < cnt; kk += blockDim.x) {
                const int k = k0 + kk;
                if (d.L == d.M) {  // identity ratio: the reference copies
                    for (int s = 0; s < nch; s++) ys[s][kk] = x_at(s, k);
                    continue;
                }
                const long long tk = (long long)d.t0 + (long long)k * d.M;
                const int Dk = (int)(tk / d.L);
                const double* arm = d.arms + (size_t)(tk - (long long)Dk * d.L) * ARGUS_AUDIO_ARM;
                for (int s = 0; s < nch; s++) {
                    double acc = 0.0;
#pragma unroll
                    for (int i = 0; i < ARGUS_AUDIO_ARM; i++) acc = fma(arm[i], (double)x_at(s, Dk - 1 - i), acc);
                    ys[s][kk] = (float)acc;
                }
            }
            __syncthreads();
            if (d.deemph_on) {
                const double al = d.alpha, bl = 1.0 - d.alpha;
                const int seg = (cnt + (int)blockDim.x - 1) / (int)blockDim.x;
                const int b0 = min(t * seg, cnt), e0 = min(b0 + seg, cnt);
                for (int s = 0; s < nch; s++) {
                    // this thread's segment as the map y -> a y + b
                    double a = 1.0, b = 0.0;
                    for (int i = b0; i < e0; i++) {
                        a *= al;
                        b = al * b + bl * (double)ys[s][i];
                    }
                    sa[t] = a;
                    sb[t] = b;
                    __syncthreads();
                    for (int off = 1; off < (int)blockDim.x; off <<= 1) {  // inclusive scan: earlier maps first
                        double pa = 1.0, pb = 0.0;
                        const bool have = t >= off;
                        if (have) {
                            pa = sa[t - off];
                            pb = sb[t - off];
                        }
                        __syncthreads();
                        if (have) {
                            const double ca = sa[t], cb = sb[t];
                            sa[t] = ca * pa;
                            sb[t] = ca * pb + cb;
                        }
                        __syncthreads();
                    }
                    double yy = t > 0 ? sa[t - 1] * y[s] + sb[t - 1] : y[s];
                    for (int i = b0; i < e0; i++) {
                        yy = al * yy + bl * (double)ys[s][i];
                        ys[s][i] = (float)yy;
                    }
                    y[s] = sa[blockDim.x - 1] * y[s] + sb[blockDim.x - 1];
                    __syncthreads();
                }
            }
            for (int j = t; j < cnt * nch; j += blockDim.x) {
                const int kk = j / nch, s = j - kk * nch;
                d.out[(size_t)k0 * nch + j] = ys[s][kk] * d.gain;
            }
            __syncthreads();
        }
        if (t == 0 && d.deemph_on)
            for (int s = 0; s < nch; s++) d.deemph[s] = y[s];
        for (int j = t; j < nch * ARGUS_AUDIO_ARM; j += blockDim.x) {
            const int s = j / ARGUS_AUDIO_ARM, i = j - s * ARGUS_AUDIO_ARM;
            d.hist_wr[j] = x_at(s, n - ARGUS_AUDIO_ARM + i);
        }
        __syncthreads();
    }
}

// resample.go:45-93 NewResampler(in, out, 32): L/M over the gcd, the
// 32 L + 1-tap prototype sin(2 pi fc x) / x * Kaiser(6) * L with
// fc = 0.9 / max(L, M), split into L arms of 33 taps.
int argus_audio_arms(int in_rate, int out_rate, int* L, int* M, double* arms)
{
    if (in_rate <= 0 || out_rate <= 0) return -1;
    const int g = gcd_int(in_rate, out_rate);
    const int l = out_rate / g, m = in_rate / g;
    if (L) *L = l;
    if (M) *M = m;
    if (!arms) return 0;
    int len = 32 * l;
    if (len % 2 == 0) len++;
    if ((len + l - 1) / l != ARGUS_AUDIO_ARM) return -1;
    const double fc = 0.90 / (double)(l > m ? l : m);
    const double mid = (double)(len - 1) / 2.0;
    const double i0b = bessel_i0(6.0);
    std::vector<double> proto((size_t)len);
    for (int i = 0; i < len; i++) {
        const double x = (double)i - mid;
        const double s = std::fabs(x) < 1e-12 ? 2 * M_PI * fc : std::sin(2 * M_PI * fc * x) / x;
        const double u = x / mid;
        proto[(size_t)i] = s * (bessel_i0(6.0 * std::sqrt(1 - u * u)) / i0b) * (double)l;
    }
    for (int p = 0; p < l; p++)
        for (int k = 0; k < ARGUS_AUDIO_ARM; k++) {
            const int idx = p + k * l;
            arms[(size_t)p * ARGUS_AUDIO_ARM + k] = idx < len ? proto[(size_t)idx] : 0.0;
        }
    return 0;
}

// One call of n inputs from phase t0: outputs and the next phase.
void argus_audio_block(int L, int M, int t0, int n, int* n_out, int* t_next)
{
    if (n <= 0 || L == M) {
        *n_out = n > 0 ? n : 0;
        *t_next = t0;
        return;
    }
    const long long need = (long long)n * L - t0;
    const long long kn = need > 0 ? (need + M - 1) / M : 0;
    const long long D = ((long long)t0 + kn * M) / L;
    const long long K = D == n ? kn + 1 : kn;
    *n_out = (int)K;
    *t_next = (int)((long long)t0 + K * M - (long long)L * (n + (D == n ? 0 : 1)));
}

int argus_launch_audio_hdr(const ArgusAudioHdr* d_hdr, bool lpf, hipStream_t stream)
{
    argus_clear_error();
    if (lpf) {
        hipLaunchKernelGGL(argus_audio_lpf_kernel, dim3(ARGUS_AUDIO_GRID), dim3(256), 0, stream, d_hdr);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL(argus_audio_kernel, dim3(ARGUS_AUDIO_GRID), dim3(256), 0, stream, d_hdr);
    return (int)hipGetLastError();
}
