/*
 * demod.c -- oracle restatement of the reference's demodulators
 * (internal/demod) and the RDS baseband extraction.  Test infrastructure
 * only (oracle.h).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

/* internal/demod/fm.go:179-196  fmDiscrimInto: atan2 of conj(prev)*curr in
 * FP64, n-1 outputs. */
int orc_fm_discrim(const float* iq, int n, float* out)
{
    if (n < 2) return 0;
    for (int i = 1; i < n; i++) {
        double pr = iq[2 * (i - 1)], pi = iq[2 * (i - 1) + 1];
        double cr = iq[2 * i], ci = iq[2 * i + 1];
        double num = pr * ci - pi * cr;
        double den = pr * cr + pi * ci;
        out[i - 1] = (float)atan2(num, den);
    }
    return n - 1;
}

/* internal/recorder/streamer.go:1287-1302  block-bridged discriminator: the
 * previous block's last IQ sample is prepended, so a block of n samples gives
 * n outputs once a previous sample exists. */
int orc_fm_discrim_bridged(const float* iq, int n, float* last_iq, int* have_last, float* out)
{
    int m;
    if (*have_last && n > 0) {
        float* tmp = (float*)malloc(sizeof(float) * 2 * (size_t)(n + 1));
        tmp[0] = last_iq[0];
        tmp[1] = last_iq[1];
        memcpy(tmp + 2, iq, sizeof(float) * 2 * (size_t)n);
        m = orc_fm_discrim(tmp, n + 1, out);
        free(tmp);
    } else {
        m = orc_fm_discrim(iq, n, out);
    }
    if (n > 0) {
        last_iq[0] = iq[2 * (n - 1)];
        last_iq[1] = iq[2 * (n - 1) + 1];
        *have_last = 1;
    }
    return m;
}

/* internal/demod/am.go:11-27  |x| (hypot, FP64) minus the block mean. */
int orc_am_demod(const float* iq, int n, float* out)
{
    if (n <= 0) return 0;
    double mean = 0;
    for (int i = 0; i < n; i++) {
        double mag = hypot((double)iq[2 * i], (double)iq[2 * i + 1]);
        mean += mag;
        out[i] = (float)mag;
    }
    mean /= (double)n;
    float fm = (float)mean;
    for (int i = 0; i < n; i++) out[i] -= fm;
    return n;
}

/* internal/demod/ssb.go:19-39  product detector, BFO +700 Hz (USB) or
 * -700 Hz (LSB), pre-incremented phase.  The reference restarts the phase at
 * 0 on every call; *bfo_phase_io carries it so a stream of blocks is
 * continuous (the reference GPU engine carries it too,
 * gpudemod_windows.go:515).  Passing 0 every call gives the CPU behaviour. */
int orc_ssb_demod(const float* iq, int n, int sample_rate, int usb, double* bfo_phase_io, float* out)
{
    if (n <= 0) return 0;
    double bfo = usb ? 700.0 : -700.0;
    double phase = *bfo_phase_io;
    double inc = 2 * ORC_PI * bfo / (double)sample_rate;
    for (int i = 0; i < n; i++) {
        phase += inc;
        double c = cos(phase), s = sin(phase);
        out[i] = (float)((double)iq[2 * i] * c - (double)iq[2 * i + 1] * s);
    }
    *bfo_phase_io = fmod(phase, 2 * ORC_PI); /* carried modulo one turn */
    return n;
}

/* internal/demod/cw.go:11-26  Re(x) * cos(phase), BFO +700 Hz. */
int orc_cw_demod(const float* iq, int n, int sample_rate, double* bfo_phase_io, float* out)
{
    if (n <= 0) return 0;
    double phase = *bfo_phase_io;
    double inc = 2 * ORC_PI * 700.0 / (double)sample_rate;
    for (int i = 0; i < n; i++) {
        phase += inc;
        out[i] = (float)((double)iq[2 * i] * cos(phase));
    }
    *bfo_phase_io = fmod(phase, 2 * ORC_PI); /* carried modulo one turn */
    return n;
}

/* internal/demod/fm.go:234-271  RDSBasebandComplexInto: discriminate, mix by
 * +57 kHz (pre-increment from phase 0), 101-tap LowpassFIR(7500) complex FIR
 * with zero history, keep every floor(fs/19000)-th sample.  Returns the
 * sample count; *out_rate = fs / decim. */
int orc_rds_baseband(const float* iq, int n, int sample_rate, float* out, int out_cap, int* out_rate)
{
    *out_rate = 0;
    if (n < 2 || sample_rate <= 0) return 0;
    int nd = n - 1;
    float* disc = (float*)malloc(sizeof(float) * (size_t)nd);
    float* cplx = (float*)malloc(sizeof(float) * 2 * (size_t)nd);
    float* filt = (float*)malloc(sizeof(float) * 2 * (size_t)nd);
    orc_fm_discrim(iq, n, disc);
    double phase = 0.0;
    double inc = -2 * ORC_PI * (-57000.0) / (double)sample_rate;
    for (int i = 0; i < nd; i++) {
        phase += inc;
        double re = cos(phase), im = sin(phase);
        double b = disc[i];
        cplx[2 * i] = (float)(b * re);
        cplx[2 * i + 1] = (float)(b * im);
    }
    double taps[103];
    int nt = orc_lowpass_fir(7500, sample_rate, 101, taps);
    orc_apply_fir(cplx, nd, taps, nt, filt);
    int decim = sample_rate / 19000;
    if (decim < 1) decim = 1;
    int m = 0;
    for (int i = 0; i < nd; i += decim) {
        if (m >= out_cap) break;
        out[2 * m] = filt[2 * i];
        out[2 * m + 1] = filt[2 * i + 1];
        m++;
    }
    *out_rate = sample_rate / decim;
    free(disc);
    free(cplx);
    free(filt);
    return m;
}
