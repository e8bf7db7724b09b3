/*
 * dsp.c -- oracle restatement of the reference's internal/dsp primitives.
 * Test infrastructure only (see oracle.h).  Complex samples are interleaved
 * float32 pairs (Go complex64 layout).
 */
#include "oracle.h"
#include <math.h>
#include <string.h>
#include <stdlib.h>

#define ORC_PI 3.14159265358979323846

/* internal/dsp/fir.go:6-28  LowpassFIR: Hann-windowed sinc, odd length. */
int orc_lowpass_fir(double cutoff_hz, int sample_rate, int taps, double* out)
{
    if (taps % 2 == 0) taps++;
    for (int n = 0; n < taps; n++) out[n] = 0.0;
    double fc = cutoff_hz / (double)sample_rate;
    if (fc <= 0) return taps;
    double m = (double)(taps - 1) / 2.0;
    for (int n = 0; n < taps; n++) {
        double x = (double)n - m;
        double sinc = (x == 0) ? 2 * fc : sin(2 * ORC_PI * fc * x) / (ORC_PI * x);
        double w = 0.5 * (1 - cos(2 * ORC_PI * (double)n / (double)(taps - 1)));
        out[n] = sinc * w;
    }
    return taps;
}

/* Same sinc as internal/dsp/fir.go:6-28 with a Hamming window
 * (0.54 - 0.46 cos), the designer the plain 2.048 Msps WFM workload names. */
int orc_lowpass_fir_hamming(double cutoff_hz, int sample_rate, int taps, double* out)
{
    if (taps % 2 == 0) taps++;
    for (int n = 0; n < taps; n++) out[n] = 0.0;
    double fc = cutoff_hz / (double)sample_rate;
    if (fc <= 0) return taps;
    double m = (double)(taps - 1) / 2.0;
    for (int n = 0; n < taps; n++) {
        double x = (double)n - m;
        double sinc = (x == 0) ? 2 * fc : sin(2 * ORC_PI * fc * x) / (ORC_PI * x);
        double w = 0.54 - 0.46 * cos(2 * ORC_PI * (double)n / (double)(taps - 1));
        out[n] = sinc * w;
    }
    return taps;
}

/* internal/dsp/fir.go:31-52  ApplyFIR: causal FIR, zero history, FP64 accumulate. */
void orc_apply_fir(const float* iq, int n, const double* taps, int nt, float* out)
{
    if (n <= 0 || nt <= 0) return;
    for (int i = 0; i < n; i++) {
        double ar = 0, ai = 0;
        for (int k = 0; k < nt; k++) {
            int idx = i - k;
            if (idx < 0) break;
            double w = taps[k];
            ar += (double)iq[2 * idx] * w;
            ai += (double)iq[2 * idx + 1] * w;
        }
        out[2 * i] = (float)ar;
        out[2 * i + 1] = (float)ai;
    }
}

/* internal/dsp/fir.go:84-95  Decimate: keep indices 0, D, 2D, ... */
int orc_decimate(const float* iq, int n, int factor, float* out)
{
    if (factor <= 1) {
        memcpy(out, iq, sizeof(float) * 2 * (size_t)(n > 0 ? n : 0));
        return n;
    }
    int m = 0;
    for (int i = 0; i < n; i += factor) {
        out[2 * m] = iq[2 * i];
        out[2 * m + 1] = iq[2 * i + 1];
        m++;
    }
    return m;
}

/* internal/dsp/freqshift.go:6-23  FreqShift: pre-increment NCO from phase 0. */
void orc_freq_shift(const float* iq, int n, int sample_rate, double offset_hz, float* out)
{
    if (n <= 0 || sample_rate <= 0 || offset_hz == 0) {
        if (n > 0) memcpy(out, iq, sizeof(float) * 2 * (size_t)n);
        return;
    }
    double phase = 0.0;
    double inc = -2 * ORC_PI * offset_hz / (double)sample_rate;
    for (int i = 0; i < n; i++) {
        phase += inc;
        double re = cos(phase), im = sin(phase);
        double vr = iq[2 * i], vi = iq[2 * i + 1];
        out[2 * i] = (float)(vr * re - vi * im);
        out[2 * i + 1] = (float)(vr * im + vi * re);
    }
}

/* internal/dsp/fir_stateful.go:6-54  StatefulFIRReal.ProcessInto: FP64 delay
 * line of len(taps), newest sample at delay[0]. */
void orc_stateful_fir_real(const double* taps, int nt, double* delay, const float* x, int n, float* out)
{
    if (n <= 0 || nt <= 0) return;
    for (int i = 0; i < n; i++) {
        memmove(delay + 1, delay, sizeof(double) * (size_t)(nt - 1));
        delay[0] = (double)x[i];
        double acc = 0;
        for (int k = 0; k < nt; k++) acc += delay[k] * taps[k];
        out[i] = (float)acc;
    }
}

/* internal/dsp/decimating_fir.go:51-95  StatefulDecimatingFIRComplex.Process:
 * emits when phase==0 (then phase = factor-1), else phase--.  Returns count. */
int orc_stateful_decim_fir_complex(const double* taps, int nt, int factor, double* dr, double* di,
                                   int* phase_io, const float* iq, int n, float* out)
{
    if (n <= 0 || nt <= 0) return 0;
    if (factor < 1) factor = 1;
    int m = 0;
    int phase = *phase_io;
    for (int i = 0; i < n; i++) {
        memmove(dr + 1, dr, sizeof(double) * (size_t)(nt - 1));
        memmove(di + 1, di, sizeof(double) * (size_t)(nt - 1));
        dr[0] = (double)iq[2 * i];
        di[0] = (double)iq[2 * i + 1];
        if (factor <= 1 || phase == 0) {
            double ar = 0, ai = 0;
            for (int k = 0; k < nt; k++) {
                double w = taps[k];
                ar += dr[k] * w;
                ai += di[k] * w;
            }
            out[2 * m] = (float)ar;
            out[2 * m + 1] = (float)ai;
            m++;
            if (factor > 1) phase = factor - 1;
        } else {
            phase--;
        }
    }
    *phase_io = phase;
    return m;
}
