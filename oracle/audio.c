/* audio.c -- CPU oracle of the post-demodulation audio stage (test
 * infrastructure only; nothing in the product links it).
 *
 *   internal/dsp/resample.go:28-170   Resampler (rational polyphase L/M)
 *   internal/dsp/resample.go:236-290  gcd, windowedSinc, kaiserWindow, bessel0
 *   internal/recorder/streamer.go:1408-1470  resample to 48 kHz, de-emphasis,
 *                                            WFM gain 0.35
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846  /* math.Pi */

/* resample.go:238-248 */
int orc_gcd(int a, int b)
{
    while (b != 0) {
        int t = a % b;
        a = b;
        b = t;
    }
    return a < 0 ? -a : a;
}

/* resample.go:280-292 bessel0: series, stops when a term drops below 1e-12 of the sum */
static double bessel0(double x)
{
    double sum = 1.0, term = 1.0;
    for (int k = 1; k < 30; k++) {
        term *= (x / (2 * (double)k)) * (x / (2 * (double)k));
        sum += term;
        if (term < 1e-12 * sum) break;
    }
    return sum;
}

/* resample.go:272-277 kaiserWindow */
static double kaiser(int n, int N, double beta)
{
    double mid = (double)(N - 1) / 2.0;
    double x = ((double)n - mid) / mid;
    return bessel0(beta * sqrt(1 - x * x)) / bessel0(beta);
}

/* resample.go:252-269 windowedSinc: sin(2 pi fc x) / x (2 pi fc at x = 0),
 * Kaiser beta = 6, times gain */
void orc_windowed_sinc(int length, double fc, double gain, double* out)
{
    double mid = (double)(length - 1) / 2.0;
    for (int n = 0; n < length; n++) {
        double x = (double)n - mid;
        double s = fabs(x) < 1e-12 ? 2 * ORC_PI * fc : sin(2 * ORC_PI * fc * x) / x;
        out[n] = s * kaiser(n, length, 6.0) * gain;
    }
}

/* resample.go:45-93 NewResampler */
int orc_resampler_init(orc_resampler* r, int in_rate, int out_rate, int taps_per_phase)
{
    memset(r, 0, sizeof *r);
    if (in_rate <= 0 || out_rate <= 0) in_rate = out_rate = 1;
    if (taps_per_phase < 4) taps_per_phase = 4;
    int g = orc_gcd(in_rate, out_rate);
    int l = out_rate / g, m = in_rate / g;
    int proto_len = l * taps_per_phase;
    if (proto_len % 2 == 0) proto_len++;
    double fc = 0.90 / (double)(l > m ? l : m);
    double* proto = (double*)malloc(sizeof(double) * (size_t)proto_len);
    orc_windowed_sinc(proto_len, fc, (double)l, proto);
    int taps = (proto_len + l - 1) / l;
    r->l = l;
    r->m = m;
    r->taps = taps;
    r->bank = (double*)calloc((size_t)l * (size_t)taps, sizeof(double));
    r->delay = (double*)calloc((size_t)taps, sizeof(double));
    for (int p = 0; p < l; p++)
        for (int t = 0; t < taps; t++) {
            int idx = p + t * l;
            if (idx < proto_len) r->bank[(size_t)p * taps + t] = proto[idx];
        }
    free(proto);
    r->out_time = 0;
    return 0;
}

void orc_resampler_free(orc_resampler* r)
{
    free(r->bank);
    free(r->delay);
    r->bank = r->delay = NULL;
}

/* resample.go:106-152 Process: returns the number of outputs written (out
 * must hold n * l / m + 4).  Identity ratio copies. */
int orc_resampler_process(orc_resampler* r, const float* in, int n, float* out)
{
    if (n <= 0) return 0;
    if (r->l == r->m) {
        memcpy(out, in, sizeof(float) * (size_t)n);
        return n;
    }
    const int L = r->l, M = r->m, taps = r->taps;
    int in_pos = 0, t = r->out_time, k = 0;
    while (in_pos < n) {
        while (t >= L) {
            t -= L;
            if (in_pos >= n) {
                r->out_time = t;
                return k;
            }
            memmove(r->delay + 1, r->delay, sizeof(double) * (size_t)(taps - 1));
            r->delay[0] = (double)in[in_pos++];
        }
        const double* arm = r->bank + (size_t)t * taps;
        double acc = 0;
        for (int i = 0; i < taps; i++) acc += r->delay[i] * arm[i];
        out[k++] = (float)acc;
        t += M;
    }
    r->out_time = t;
    return k;
}

/* streamer.go:1443-1467 de-emphasis: y = a y + (1 - a) x in FP64, samples
 * rewritten as float32; stride 1 (mono) or 2 (one channel of interleaved
 * stereo).  Returns the new state. */
double orc_deemphasis(float* x, int n, int stride, double alpha, double y)
{
    for (int i = 0; i < n; i++) {
        y = alpha * y + (1 - alpha) * (double)x[(size_t)i * stride];
        x[(size_t)i * stride] = (float)y;
    }
    return y;
}
