// audio.hip -- post-demodulation audio stage kernel and its host-side filter
// design (layout and semantics in audio.hpp).
//
// argus_audio_kernel: one workgroup per channel at a time (grid-stride over
// the block's audio channels).  Per chunk of outputs the inputs the chunk
// reads (and the channel's arms, once, when they fit) are staged in LDS; each
// thread evaluates whole outputs from LDS in FP32 (both sides of a stereo
// frame with one arm read); the de-emphasis runs in FP64 as a wave-level
// scan of the affine maps y -> a y + b (one wave per side); gain; the next
// delay line of the last T inputs per side.
#include "audio.hpp"

#include <cmath>
#include <vector>

int argus_lowpass_fir(double cutoff_hz, int sample_rate, int taps, double* out, bool hamming);

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

// resample.go:45-93 NewResampler(in, out, 32): the 32 L + 1-tap prototype
// sin(2 pi fc x) / x * Kaiser(6) * L with fc = 0.9 / max(L, M), split into
// L arms of 33 taps (arm p, tap k = prototype[p + k L]), FP64.
void resampler_arms(int l, int m, std::vector<double>& arms)
{
    int len = 32 * l;
    if (len % 2 == 0) len++;
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
    arms.assign((size_t)l * ARGUS_AUDIO_ARM, 0.0);
    for (int p = 0; p < l; p++)
        for (int k = 0; k < ARGUS_AUDIO_ARM; k++) {
            const int idx = p + k * l;
            if (idx < len) arms[(size_t)p * ARGUS_AUDIO_ARM + k] = proto[(size_t)idx];
        }
}

template <bool LDS_ARMS, int NCH>
__device__ __forceinline__ void audio_outputs(const ArgusAudioDesc& d, const float* __restrict__ as,
                                              const float* __restrict__ xs, float* __restrict__ ys, int xside,
                                              unsigned p0, int cnt)
{
    const int T = d.T;
    for (int kk = threadIdx.x; kk < cnt; kk += ARGUS_AUDIO_THREADS) {
        unsigned dq = (unsigned)kk, p = 0;
        if (!d.ident) {
            const unsigned u = p0 + (unsigned)kk * (unsigned)d.M;
            dq = u / (unsigned)d.L;
            p = u - dq * (unsigned)d.L;
        }
        const float* a = (LDS_ARMS ? as : d.arms) + (size_t)p * T;
        const float* x0 = xs + (T - 1 + (int)dq);  // newest input of the output's window
        float acc0 = 0.0f, acc1 = 0.0f;
#pragma unroll 4
        for (int i = 0; i < T; i++) {
            const float w = a[i];
            acc0 = fmaf(w, x0[-i], acc0);
            if (NCH == 2) acc1 = fmaf(w, x0[xside - i], acc1);
        }
        ys[kk * NCH] = acc0;
        if (NCH == 2) ys[kk * NCH + 1] = acc1;
    }
}

}  // namespace

__global__ __launch_bounds__(ARGUS_AUDIO_THREADS) void argus_audio_kernel(const ArgusAudioHdr* __restrict__ hdr)
{
    __shared__ float xs[ARGUS_AUDIO_XCAP];
    __shared__ float as[ARGUS_AUDIO_ARMCAP];
    __shared__ float ys[ARGUS_AUDIO_OUTCAP];
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int n_chans = hdr->n_chans;
    for (int c = blockIdx.x; c < n_chans; c += gridDim.x) {
        const ArgusAudioDesc d = hdr->descs[c];
        const int nch = d.channels, n = d.n_in, T = d.T;
        if (n <= 0) continue;  // an empty block leaves the stage's state as it is
        const int xside = ARGUS_AUDIO_XCAP / nch;
        // outputs per chunk: the inputs they read span at most
        // (cnt - 1) M / L + 1 + T <= xside samples per side
        int chunk = d.ident ? xside - T : (int)(((long long)(xside - T - 1) * d.L) / d.M) + 1;
        chunk = min(chunk, ARGUS_AUDIO_OUTCAP / nch);
        const bool arms_lds = d.L * T <= ARGUS_AUDIO_ARMCAP;
        if (arms_lds)
            for (int j = t; j < d.L * T; j += ARGUS_AUDIO_THREADS) as[j] = d.arms[j];
        double y = 0.0;  // de-emphasis state of side `wave`
        if (d.deemph_on && wave < nch) y = d.deemph[wave];
        for (int k0 = 0; k0 < d.n_out; k0 += chunk) {
            const int cnt = min(chunk, d.n_out - k0);
            long long D0, Dl;  // inputs consumed by the chunk's first and last outputs
            unsigned p0 = 0;
            if (d.ident) {
                D0 = k0 + 1;
                Dl = k0 + cnt;
            } else {
                const long long tk = (long long)d.t0 + (long long)k0 * d.M;
                D0 = tk / d.L;
                p0 = (unsigned)(tk - D0 * d.L);
                Dl = ((long long)d.t0 + (long long)(k0 + cnt - 1) * d.M) / d.L;
            }
            const long long lo = D0 - T;     // first staged input (>= -T)
            const int span = (int)(Dl - lo);  // staged inputs lo .. Dl - 1 per side
            __syncthreads();  // the previous chunk's readers of xs and ys are done
            for (int j = t; j < span * nch; j += ARGUS_AUDIO_THREADS) {
                const int i = nch == 2 ? j >> 1 : j, s = j - i * nch;
                const long long idx = lo + i;
                xs[s * xside + i] = idx >= 0 ? d.in[idx * nch + s]
                                             : d.hist_rd[s * ARGUS_AUDIO_TMAX + T + (int)idx];
            }
            __syncthreads();
            if (nch == 2) {
                if (arms_lds) audio_outputs<true, 2>(d, as, xs, ys, xside, p0, cnt);
                else audio_outputs<false, 2>(d, as, xs, ys, xside, p0, cnt);
            } else {
                if (arms_lds) audio_outputs<true, 1>(d, as, xs, ys, xside, p0, cnt);
                else audio_outputs<false, 1>(d, as, xs, ys, xside, p0, cnt);
            }
            __syncthreads();
            if (d.deemph_on && wave < nch) {
                // lane l takes outputs [l seg, (l + 1) seg) of side `wave`
                const double al = d.alpha, bl = 1.0 - d.alpha;
                const int seg = (cnt + 63) >> 6;
                const int b0 = min(lane * seg, cnt), e0 = min(b0 + seg, cnt);
                double a = 1.0, b = 0.0;
                for (int i = b0; i < e0; i++) {
                    a *= al;
                    b = al * b + bl * (double)ys[i * nch + wave];
                }
                for (int off = 1; off < 64; off <<= 1) {  // inclusive scan, earlier maps applied first
                    const double pa = __shfl_up(a, off, 64), pb = __shfl_up(b, off, 64);
                    if (lane >= off) {
                        b = a * pb + b;
                        a = a * pa;
                    }
                }
                double ea = __shfl_up(a, 1, 64), eb = __shfl_up(b, 1, 64);
                if (lane == 0) {
                    ea = 1.0;
                    eb = 0.0;
                }
                double yy = ea * y + eb;
                for (int i = b0; i < e0; i++) {
                    yy = al * yy + bl * (double)ys[i * nch + wave];
                    ys[i * nch + wave] = (float)yy;
                }
                y = __shfl(a, 63, 64) * y + __shfl(b, 63, 64);
            }
            __syncthreads();
            for (int j = t; j < cnt * nch; j += ARGUS_AUDIO_THREADS) d.out[(size_t)k0 * nch + j] = ys[j] * d.gain;
        }
        if (d.deemph_on && wave < nch && lane == 0) d.deemph[wave] = y;
        for (int j = t; j < nch * T; j += ARGUS_AUDIO_THREADS) {
            const int s = j / T, i = j - s * T;
            const int idx = n - T + i;
            d.hist_wr[s * ARGUS_AUDIO_TMAX + i] =
                idx >= 0 ? d.in[(size_t)idx * nch + s] : d.hist_rd[s * ARGUS_AUDIO_TMAX + T + idx];
        }
    }
}

int argus_audio_filter(int in_rate, int out_rate, bool lowpass, int* L, int* M, int* T, float* arms)
{
    if (in_rate <= 0 || out_rate <= 0) return -1;
    const int g = gcd_int(in_rate, out_rate);
    const int l = out_rate / g, m = in_rate / g;
    if (m >= ARGUS_AUDIO_MAX_M || l >= ARGUS_AUDIO_MAX_M) return -1;
    const bool ident = l == m;
    const int taps = ident ? (lowpass ? ARGUS_AUDIO_LPF_TAPS : 1) : (lowpass ? ARGUS_AUDIO_TMAX : ARGUS_AUDIO_ARM);
    if (L) *L = ident ? 1 : l;
    if (M) *M = ident ? 1 : m;
    if (T) *T = taps;
    if (!arms) return 0;
    // wfmAudioFilter's LowpassFIR(15000, rate, 101), Hann, FP64 (streamer.go:1511-1520)
    std::vector<double> h((size_t)ARGUS_AUDIO_LPF_TAPS + 1, 0.0);
    if (lowpass) argus_lowpass_fir(15000.0, in_rate, ARGUS_AUDIO_LPF_TAPS, h.data(), false);
    if (ident) {
        if (lowpass)
            for (int k = 0; k < ARGUS_AUDIO_LPF_TAPS; k++) arms[k] = (float)h[(size_t)k];
        else
            arms[0] = 1.0f;
        return 0;
    }
    std::vector<double> r;
    resampler_arms(l, m, r);
    for (int p = 0; p < l; p++) {
        const double* rp = r.data() + (size_t)p * ARGUS_AUDIO_ARM;
        float* out = arms + (size_t)p * taps;
        if (!lowpass) {
            for (int k = 0; k < ARGUS_AUDIO_ARM; k++) out[k] = (float)rp[k];
            continue;
        }
        // composite arm: c[m] = sum_{i + j = m} arm[i] h[j] (both newest-first)
        for (int mm = 0; mm < taps; mm++) {
            double acc = 0.0;
            for (int i = std::max(0, mm - (ARGUS_AUDIO_LPF_TAPS - 1)); i <= std::min(mm, ARGUS_AUDIO_ARM - 1); i++)
                acc += rp[i] * h[(size_t)(mm - i)];
            out[mm] = (float)acc;
        }
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

int argus_launch_audio_hdr(const ArgusAudioHdr* d_hdr, hipStream_t stream)
{
    argus_clear_error();
    hipLaunchKernelGGL(argus_audio_kernel, dim3(ARGUS_AUDIO_GRID), dim3(ARGUS_AUDIO_THREADS), 0, stream, d_hdr);
    return (int)hipGetLastError();
}
