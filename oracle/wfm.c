/*
 * wfm.c -- oracle restatement of the WFM stereo decoder, the RDS decoder,
 * the pilot-presence detector and an RDS group encoder for synthesis.
 * Test infrastructure only (oracle.h).
 */
#include "oracle.h"
#include "wfm.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------ stereo */

/* internal/recorder/streamer.go:1496-1506 pllCoefficients */
void orc_pll_coefficients(double loop_bw, double damping, int sample_rate, double* alpha, double* beta)
{
    if (sample_rate <= 0 || loop_bw <= 0) {
        *alpha = 0;
        *beta = 0;
        return;
    }
    double bl = loop_bw / (double)sample_rate;
    double theta = bl / (damping + 0.25 / damping);
    double d = 1 + 2 * damping * theta + theta * theta;
    *alpha = (4 * damping * theta) / d;
    *beta = (4 * theta * theta) / d;
}

static void fir_state_init(orc_fir_state* f, double cutoff, int sr)
{
    f->nt = orc_lowpass_fir(cutoff, sr, 101, f->taps);
    memset(f->delay, 0, sizeof f->delay);
}

void orc_stereo_init(orc_stereo_state* st, int sample_rate)
{
    memset(st, 0, sizeof *st);
    st->rate = sample_rate;
    fir_state_init(&st->lpf, 15000, sample_rate);
    fir_state_init(&st->bp_hi, 53000, sample_rate);
    fir_state_init(&st->bp_lo, 23000, sample_rate);
    fir_state_init(&st->lr_lpf, 15000, sample_rate);
    fir_state_init(&st->pilot_hi, 21000, sample_rate);
    fir_state_init(&st->pilot_lo, 17000, sample_rate);
    st->phase = 0;
    st->freq = 2 * ORC_PI * 19000 / (double)sample_rate;
    orc_pll_coefficients(50, 0.707, sample_rate, &st->alpha, &st->beta);
    st->err_avg = 0;
    st->i_state = 0;
    st->q_state = 0;
    st->lp_alpha = 1 - exp(-2 * ORC_PI * 200 / (double)sample_rate);
}

/*
 * streamer.go:1525-1643 stereoDecodeStateful.  out receives 2n interleaved
 * L, R float32; returns the lock flag of this block.  The lock hysteresis of
 * streamer.go:1367-1380 (4 locked blocks on, 10 unlocked off) updates
 * st->enabled.
 */
int orc_stereo_decode(orc_stereo_state* st, const float* mono, int n, float* out)
{
    if (n <= 0 || st->rate <= 0) return 0;
    float* lpr = (float*)malloc(sizeof(float) * (size_t)n * 5);
    float* bpf = lpr + n;
    float* lr = lpr + 2 * n;
    float* w1 = lpr + 3 * n;
    float* w2 = lpr + 4 * n;
    orc_stateful_fir_real(st->lpf.taps, st->lpf.nt, st->lpf.delay, mono, n, lpr);
    orc_stateful_fir_real(st->bp_hi.taps, st->bp_hi.nt, st->bp_hi.delay, mono, n, w1);
    orc_stateful_fir_real(st->bp_lo.taps, st->bp_lo.nt, st->bp_lo.delay, mono, n, w2);
    for (int i = 0; i < n; i++) bpf[i] = w1[i] - w2[i];
    orc_stateful_fir_real(st->pilot_hi.taps, st->pilot_hi.nt, st->pilot_hi.delay, mono, n, w1);
    orc_stateful_fir_real(st->pilot_lo.taps, st->pilot_lo.nt, st->pilot_lo.delay, mono, n, w2);
    for (int i = 0; i < n; i++) w1[i] = w1[i] - w2[i];
    double phase = st->phase, freq = st->freq, is = st->i_state, qs = st->q_state;
    const double alpha = st->alpha, beta = st->beta, lpa = st->lp_alpha;
    const double min_f = 2 * ORC_PI * 17000 / (double)st->rate;
    const double max_f = 2 * ORC_PI * 21000 / (double)st->rate;
    double pilot_power = 0, err_sum = 0;
    for (int i = 0; i < n; i++) {
        double p = (double)w1[i];
        double sp = sin(phase), cp = cos(phase);
        double im = p * cp;
        double qm = p * -sp;
        is += lpa * (im - is);
        qs += lpa * (qm - qs);
        double err = atan2(qs, is);
        freq += beta * err;
        if (freq < min_f) freq = min_f;
        else if (freq > max_f) freq = max_f;
        phase += freq + alpha * err;
        if (phase > 2 * ORC_PI) phase -= 2 * ORC_PI;
        else if (phase < 0) phase += 2 * ORC_PI;
        pilot_power += p * p;
        err_sum += fabs(err);
        lr[i] = bpf[i] * (float)(2 * sin(2 * phase));
    }
    st->phase = phase;
    st->freq = freq;
    st->i_state = is;
    st->q_state = qs;
    double block_err = err_sum / (double)n;
    st->err_avg = 0.9 * st->err_avg + 0.1 * block_err;
    orc_stateful_fir_real(st->lr_lpf.taps, st->lr_lpf.nt, st->lr_lpf.delay, lr, n, w2);
    double coherence = 0;
    if (pilot_power > 1e-12) coherence = (is * is + qs * qs) * (double)n / pilot_power;
    double freq_hz = st->freq * (double)st->rate / (2 * ORC_PI);
    int locked = coherence > 0.15 && fabs(freq_hz - 19000) < 250 && st->err_avg < 0.35;
    for (int i = 0; i < n; i++) {
        out[2 * i] = 0.5f * (lpr[i] + w2[i]);
        out[2 * i + 1] = 0.5f * (lpr[i] - w2[i]);
    }
    /* streamer.go:1367-1380 */
    if (locked) {
        st->on_count++;
        st->off_count = 0;
        if (st->on_count >= 4) st->enabled = 1;
    } else {
        st->on_count = 0;
        st->off_count++;
        if (st->off_count >= 10) st->enabled = 0;
    }
    free(lpr);
    return locked;
}

/*
 * internal/recorder/wfm_batch_stereo.go:76-157 wfmStereoDecodeBatch: the
 * stateless decoder of event recordings.  Every filter starts from zero and
 * the PLL from 19 kHz; lock = pilotPower/totalPower > 0.003 and |f - 19 kHz|
 * < 250 Hz and the block's mean |err| < 0.35.  out receives 2n interleaved
 * L, R (the matrix output, whatever the lock); returns the lock flag.
 */
int orc_stereo_decode_batch(const float* mono, int n, int sample_rate, float* out)
{
    if (n <= 0 || sample_rate <= 0) return 0;
    orc_stereo_state st;
    orc_stereo_init(&st, sample_rate);
    float* lpr = (float*)malloc(sizeof(float) * (size_t)n * 5);
    float* bpf = lpr + n;
    float* lr = lpr + 2 * n;
    float* w1 = lpr + 3 * n;
    float* w2 = lpr + 4 * n;
    orc_stateful_fir_real(st.lpf.taps, st.lpf.nt, st.lpf.delay, mono, n, lpr);
    orc_stateful_fir_real(st.bp_hi.taps, st.bp_hi.nt, st.bp_hi.delay, mono, n, w1);
    orc_stateful_fir_real(st.bp_lo.taps, st.bp_lo.nt, st.bp_lo.delay, mono, n, w2);
    for (int i = 0; i < n; i++) bpf[i] = w1[i] - w2[i];
    orc_stateful_fir_real(st.pilot_hi.taps, st.pilot_hi.nt, st.pilot_hi.delay, mono, n, w1);
    orc_stateful_fir_real(st.pilot_lo.taps, st.pilot_lo.nt, st.pilot_lo.delay, mono, n, w2);
    for (int i = 0; i < n; i++) w1[i] = w1[i] - w2[i];
    double phase = 0, freq = st.freq, is = 0, qs = 0;
    const double alpha = st.alpha, beta = st.beta, lpa = st.lp_alpha;
    const double min_f = 2 * ORC_PI * 17000 / (double)sample_rate;
    const double max_f = 2 * ORC_PI * 21000 / (double)sample_rate;
    double pilot_power = 0, total_power = 0, err_sum = 0;
    for (int i = 0; i < n; i++) {
        double p = (double)w1[i];
        double sp = sin(phase), cp = cos(phase);
        is += lpa * (p * cp - is);
        qs += lpa * (p * -sp - qs);
        double err = atan2(qs, is);
        freq += beta * err;
        if (freq < min_f) freq = min_f;
        else if (freq > max_f) freq = max_f;
        phase += freq + alpha * err;
        if (phase > 2 * ORC_PI) phase -= 2 * ORC_PI;
        else if (phase < 0) phase += 2 * ORC_PI;
        total_power += (double)mono[i] * (double)mono[i];
        pilot_power += p * p;
        err_sum += fabs(err);
        lr[i] = bpf[i] * (float)(2 * sin(2 * phase));
    }
    orc_stateful_fir_real(st.lr_lpf.taps, st.lr_lpf.nt, st.lr_lpf.delay, lr, n, w2);
    double ratio = total_power > 0 ? pilot_power / total_power : 0;
    double freq_hz = freq * (double)sample_rate / (2 * ORC_PI);
    int locked = ratio > 0.003 && fabs(freq_hz - 19000) < 250 && err_sum / (double)n < 0.35;
    for (int i = 0; i < n; i++) {
        out[2 * i] = 0.5f * (lpr[i] + w2[i]);
        out[2 * i + 1] = 0.5f * (lpr[i] - w2[i]);
    }
    free(lpr);
    return locked;
}

/* ------------------------------------------------------------ pilot presence */

/* internal/classifier/pll.go:178-193 goertzelMagnitude (FP64 samples) */
static double goertzel(const double* x, int n, double freq, int sr)
{
    if (n == 0) return 0;
    double k = freq / ((double)sr / (double)n);
    double w = 2 * ORC_PI * k / (double)n;
    double c = 2 * cos(w);
    double s1 = 0, s2 = 0;
    for (int i = 0; i < n; i++) {
        double s0 = x[i] + c * s1 - s2;
        s2 = s1;
        s1 = s0;
    }
    return sqrt(s1 * s1 + s2 * s2 - c * s1 * s2);
}

static int cmp_d(const void* a, const void* b)
{
    double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

/* internal/classifier/pll.go:82-131 StereoPilotPresent/estimateWFMPilot: the
 * FP64 FM discriminator (fmDemod, pll.go:160-176) of the last <= fs/2
 * samples, Goertzel magnitudes at 19 kHz +- 1 kHz in 10 Hz steps, present if
 * max / median > 4 (pilotLockRatio); rates below 40 kHz are never stereo.
 * offset_out = best frequency - 19 kHz (the first maximum, pll.go:107-114);
 * mags_out receives the 201 magnitudes in frequency order. */
int orc_stereo_pilot_estimate(const float* iq, int n, int sample_rate, double* ratio_out, double* offset_out,
                              double* mags_out)
{
    if (ratio_out) *ratio_out = 0;
    if (offset_out) *offset_out = 0;
    if (mags_out) memset(mags_out, 0, sizeof(double) * 201);
    if (sample_rate <= 0) return 0;
    int maxn = sample_rate / 2;
    if (maxn > 0 && n > maxn) {
        iq += 2 * (size_t)(n - maxn);
        n = maxn;
    }
    if (sample_rate < 40000 || n < 2) return 0;
    double* d = (double*)malloc(sizeof(double) * (size_t)(n - 1));
    for (int i = 1; i < n; i++) {
        double pr = iq[2 * (i - 1)], pi = iq[2 * (i - 1) + 1];
        double cr = iq[2 * i], ci = iq[2 * i + 1];
        d[i - 1] = atan2(pr * ci - pi * cr, pr * cr + pi * ci);
    }
    double mags[201];
    double best = 0, best_f = 19000.0;
    int k = 0;
    for (double off = -1000.0; off <= 1000.0; off += 10.0, k++) {
        mags[k] = goertzel(d, n - 1, 19000.0 + off, sample_rate);
        if (mags[k] > best) {
            best = mags[k];
            best_f = 19000.0 + off;
        }
    }
    if (mags_out) memcpy(mags_out, mags, sizeof mags);
    qsort(mags, (size_t)k, sizeof(double), cmp_d);
    double med = mags[k / 2];
    free(d);
    double ratio = med > 0 ? best / med : 0;
    if (ratio_out) *ratio_out = ratio;
    if (offset_out) *offset_out = best_f - 19000.0;
    return med > 0 && ratio > 4;
}

int orc_stereo_pilot_present(const float* iq, int n, int sample_rate, double* ratio_out)
{
    return orc_stereo_pilot_estimate(iq, n, sample_rate, ratio_out, NULL, NULL);
}

/* ------------------------------------------------------------ RDS decode */

/* rds.go:590-599 crcSyndrome, generator 0x1B9 */
uint16_t orc_rds_syndrome(uint32_t raw)
{
    uint32_t reg = raw;
    for (int i = 25; i >= 10; i--)
        if ((reg >> i) & 1) reg ^= (uint32_t)0x1B9 << (i - 10);
    return (uint16_t)(reg & 0x3FF);
}

enum { OFF_A = 0x0FC, OFF_B = 0x198, OFF_C = 0x168, OFF_CP = 0x350, OFF_D = 0x1B4 }; /* rds.go:509-515 */

static int decode_raw(uint32_t raw, uint16_t* data, uint16_t* off)
{
    uint16_t s = orc_rds_syndrome(raw);
    if (s == OFF_A || s == OFF_B || s == OFF_C || s == OFF_CP || s == OFF_D) {
        *data = (uint16_t)(raw >> 10);
        *off = s;
        return 1;
    }
    return 0;
}

static int allowed(uint16_t off, const uint16_t* al, int na)
{
    for (int i = 0; i < na; i++)
        if (al[i] == off) return 1;
    return 0;
}

/* rds.go:549-579 decodeBlockExpected with the unique-1-bit-flip FEC */
static int decode_block(const uint8_t* bits, const uint16_t* al, int na, int fec, uint16_t* data)
{
    uint32_t raw = 0;
    for (int i = 0; i < 26; i++) raw = (raw << 1) | (bits[i] & 1u);
    uint16_t d, o;
    if (decode_raw(raw, &d, &o) && allowed(o, al, na)) {
        *data = d;
        return 1;
    }
    if (!fec) return 0;
    int count = 0;
    uint16_t cand = 0;
    for (int i = 0; i < 26; i++) {
        uint32_t f = raw ^ (1u << (25 - i));
        if (!decode_raw(f, &d, &o) || !allowed(o, al, na)) continue;
        cand = d;
        if (++count > 1) return 0;
    }
    if (count == 1) {
        *data = cand;
        return 1;
    }
    return 0;
}

/* rds.go:432-473 decodeGroupAt */
static int decode_group(const uint8_t* bits, int n, int start, int fec, uint16_t g[4])
{
    static const uint16_t A[1] = {OFF_A}, B[1] = {OFF_B}, Cc[2] = {OFF_C, OFF_CP}, Dd[1] = {OFF_D};
    if (start < 0 || start + 104 > n) return 0;
    if (!decode_block(bits + start, A, 1, fec, &g[0])) return 0;
    if (!decode_block(bits + start + 26, B, 1, fec, &g[1])) return 0;
    if (!decode_block(bits + start + 52, Cc, 2, fec, &g[2])) return 0;
    if (!decode_block(bits + start + 78, Dd, 1, fec, &g[3])) return 0;
    return 1;
}

static char sanitize(uint8_t b) { return (b < 32 || b > 126) ? ' ' : (char)b; }

/* rds.go:475-502 applyGroup; also appends the group to the log */
static void apply_group(orc_rds_decoder* d, const uint16_t g[4])
{
    if (g[0] != 0) d->pi = g[0];
    int type = (g[1] >> 12) & 0xF;
    int version_a = ((g[1] >> 11) & 1) == 0;
    if (type == 0 && version_a) {
        int idx = (g[1] & 3) * 2;
        d->ps[idx] = sanitize((uint8_t)(g[3] >> 8));
        d->ps[idx + 1] = sanitize((uint8_t)(g[3] & 0xFF));
    }
    if (type == 2 && version_a) {
        int idx = (g[1] & 0xF) * 4;
        uint8_t c[4] = {(uint8_t)(g[2] >> 8), (uint8_t)(g[2] & 0xFF), (uint8_t)(g[3] >> 8), (uint8_t)(g[3] & 0xFF)};
        for (int j = 0; j < 4 && idx + j < 64; j++) d->rt[idx + j] = sanitize(c[j]);
    }
    if (d->n_groups < ORC_RDS_MAX_GROUPS) {
        for (int k = 0; k < 4; k++) d->groups[d->n_groups][k] = g[k];
    }
    d->n_groups++;
}

/* rds.go:365-430 tryDecode with the +-3 bit flywheel */
static int try_decode(orc_rds_decoder* d, const uint8_t* bits, int n, int fec, int apply)
{
    int groups = 0;
    uint16_t g[4];
    for (int i = 0; i + 104 <= n;) {
        if (!decode_group(bits, n, i, fec, g)) {
            i++;
            continue;
        }
        groups++;
        if (apply) apply_group(d, g);
        int next = i + 104;
        int locked = 1;
        while (locked && next + 104 <= n) {
            if (decode_group(bits, n, next, fec, g)) {
                if (apply) apply_group(d, g);
                groups++;
                next += 104;
                continue;
            }
            int matched = 0;
            for (int delta = 1; delta <= 3 && !matched; delta++) {
                int left = next - delta;
                if (left >= 0 && decode_group(bits, n, left, fec, g)) {
                    if (apply) apply_group(d, g);
                    groups++;
                    next = left + 104;
                    matched = 1;
                    break;
                }
                int right = next + delta;
                if (right + 104 <= n && decode_group(bits, n, right, fec, g)) {
                    if (apply) apply_group(d, g);
                    groups++;
                    next = right + 104;
                    matched = 1;
                    break;
                }
            }
            if (!matched) locked = 0;
        }
        int resume = next - 3;
        if (resume <= i) resume = i + 1;
        i = resume;
    }
    return groups;
}

void orc_rds_init(orc_rds_decoder* d)
{
    memset(d, 0, sizeof *d);
    memset(d->ps, ' ', 8);
    memset(d->rt, ' ', 64);
}

/*
 * rds.go:78-264 Decoder.Decode: Mueller & Muller timing (gain 0.01, linear
 * interpolation), Costas loop (alpha 0.132, beta alpha^2/4), sign(I) hard
 * bits, differential decode carried across calls, direct then inverted
 * polarity block search.  Returns the number of groups found this call.
 * Diagnostics of the reference (analyzeStream) change no state and are not
 * restated.
 */
int orc_rds_decode(orc_rds_decoder* d, const float* samples, int n, int sample_rate)
{
    if (n < 104 || sample_rate <= 0) return 0;
    d->total_decodes++;
    if (d->last_rate != 0 && d->last_rate != sample_rate) {
        d->costas_phase = d->costas_freq = 0;
        d->mu = 0;
        d->prev[0] = d->prev[1] = 0;
        d->prev_dec = 0;
        d->has_prev = 0;
        d->last_hard = 0;
        d->has_last_hard = 0;
    }
    d->last_rate = sample_rate;
    double sps = (double)sample_rate / 1187.5;
    double mu = d->mu;
    if (mu < 0 || mu >= sps || isnan(mu) || isinf(mu)) mu = sps / 2;
    int cap = (int)(n / (int)sps) + 16;
    float* sym = (float*)malloc(sizeof(float) * 2 * (size_t)cap);
    int ns = 0;
    float prev_r = d->prev[0], prev_i = d->prev[1];
    float prev_dec = d->prev_dec;
    int have_prev = d->has_prev;
    while (mu < (double)(n - 1)) {
        int idx = (int)mu;
        double frac = mu - (double)idx;
        if (idx + 1 >= n) break;
        float sr = (float)((double)samples[2 * idx] * (1 - frac) + (double)samples[2 * (idx + 1)] * frac);
        float si = (float)((double)samples[2 * idx + 1] * (1 - frac) + (double)samples[2 * (idx + 1) + 1] * frac);
        float dec = sr >= 0 ? 1.0f : -1.0f;
        if (have_prev) {
            double err = (double)(dec - prev_dec) * (double)prev_r - (double)(sr - prev_r) * (double)prev_dec;
            mu += sps + 0.01 * err;
        } else {
            mu += sps;
            have_prev = 1;
        }
        prev_dec = dec;
        prev_r = sr;
        prev_i = si;
        if (ns == cap) {
            cap *= 2;
            sym = (float*)realloc(sym, sizeof(float) * 2 * (size_t)cap);
        }
        sym[2 * ns] = sr;
        sym[2 * ns + 1] = si;
        ns++;
    }
    double res = mu - (double)n;
    while (res < 0) res += sps;
    while (res >= sps) res -= sps;
    d->mu = res;
    d->prev[0] = prev_r;
    d->prev[1] = prev_i;
    d->prev_dec = prev_dec;
    d->has_prev = have_prev;
    if (ns < 104) {
        free(sym);
        return 0;
    }
    /* Costas loop */
    const double alpha = 0.132, beta = alpha * alpha / 4.0;
    double phase = d->costas_phase, freq = d->costas_freq;
    uint8_t* hard = (uint8_t*)malloc((size_t)ns);
    for (int i = 0; i < ns; i++) {
        float cp = (float)cos(phase), sp = (float)sin(phase);
        float re = sym[2 * i] * cp + sym[2 * i + 1] * sp;
        float im = sym[2 * i + 1] * cp - sym[2 * i] * sp;
        double err = re >= 0 ? (double)im : -(double)im;
        freq += beta * err;
        phase += freq + alpha * err;
        while (phase > ORC_PI) phase -= 2 * ORC_PI;
        while (phase < -ORC_PI) phase += 2 * ORC_PI;
        hard[i] = re >= 0 ? 1 : 0;
    }
    d->costas_phase = phase;
    d->costas_freq = freq;
    /* differential decode with carry */
    uint8_t* bits = (uint8_t*)malloc((size_t)ns + 1);
    uint8_t* inv = (uint8_t*)malloc((size_t)ns + 1);
    int nb = 0;
    if (d->has_last_hard) {
        bits[nb++] = hard[0] ^ (uint8_t)d->last_hard;
        for (int i = 1; i < ns; i++) bits[nb++] = hard[i] ^ hard[i - 1];
    } else if (ns >= 2) {
        for (int i = 1; i < ns; i++) bits[nb++] = hard[i] ^ hard[i - 1];
    }
    d->last_hard = hard[ns - 1];
    d->has_last_hard = 1;
    for (int i = 0; i < nb; i++) inv[i] = 1 - bits[i];
    int found = try_decode(d, bits, nb, 1, 1);
    if (found == 0) found = try_decode(d, inv, nb, 1, 1);
    if (found > 0) d->groups_found++;
    free(sym);
    free(hard);
    free(bits);
    free(inv);
    return found;
}

/* ------------------------------------------------------------ RDS encode */

/* Checkword of a 16-bit information word: remainder of m(x) x^10 mod
 * g(x) = 0x5B9 (x^10+x^8+x^7+x^5+x^4+x^3+1), plus the offset word (IEC 62106). */
uint32_t orc_rds_block(uint16_t data, uint16_t offset)
{
    uint32_t reg = (uint32_t)data << 10;
    for (int i = 25; i >= 10; i--)
        if ((reg >> i) & 1) reg ^= (uint32_t)0x5B9 << (i - 10);
    return ((uint32_t)data << 10) | ((reg & 0x3FF) ^ offset);
}

/* Bits (MSB first) of n_groups groups: 0A groups carrying PS and 2A groups
 * carrying RT, alternating, PI fixed.  Returns the bit count (104 per group). */
int orc_rds_encode_groups(uint16_t pi, const char* ps8, const char* rt64, int n_groups, uint8_t* bits)
{
    int nb = 0;
    for (int gi = 0; gi < n_groups; gi++) {
        uint16_t b, c, d;
        if (gi % 2 == 0) { /* 0A: PS segment */
            int addr = (gi / 2) % 4;
            b = (uint16_t)((0u << 12) | (0u << 11) | (uint16_t)addr);
            c = 0xE0CD;
            d = (uint16_t)(((uint8_t)ps8[2 * addr] << 8) | (uint8_t)ps8[2 * addr + 1]);
        } else { /* 2A: RT segment */
            int addr = (gi / 2) % 16;
            b = (uint16_t)((2u << 12) | (0u << 11) | (uint16_t)addr);
            c = (uint16_t)(((uint8_t)rt64[4 * addr] << 8) | (uint8_t)rt64[4 * addr + 1]);
            d = (uint16_t)(((uint8_t)rt64[4 * addr + 2] << 8) | (uint8_t)rt64[4 * addr + 3]);
        }
        uint32_t blocks[4] = {orc_rds_block(pi, OFF_A), orc_rds_block(b, OFF_B), orc_rds_block(c, OFF_C),
                              orc_rds_block(d, OFF_D)};
        for (int k = 0; k < 4; k++)
            for (int i = 25; i >= 0; i--) bits[nb++] = (uint8_t)((blocks[k] >> i) & 1);
    }
    return nb;
}

/* ------------------------------------------------------------ streaming RDS baseband */

/*
 * The RDS baseband of internal/demod/fm.go:234-271 (RDSBasebandComplexInto)
 * carried across blocks instead of restarted per call: the +57 kHz mixer
 * phase (pre-incremented, as FreqShift(-57000) does, but computed exactly
 * from the sample count instead of by repeated FP64 addition), the 101-tap
 * LowpassFIR(7500) complex FIR history (FP64 accumulation as ApplyFIR) and
 * the decimation phase (keep stream indices 0, D, 2D, ...; D = fs/19000).
 * Input is the FM discriminator output.  Returns the output count.
 */
void orc_rds_stream_init(orc_rds_stream* s, int sample_rate)
{
    memset(s, 0, sizeof *s);
    s->rate = sample_rate;
    s->nt = orc_lowpass_fir(7500, sample_rate, 101, s->taps);
    s->decim = sample_rate / 19000;
    if (s->decim < 1) s->decim = 1;
    s->mix_num = 0;
}

int orc_rds_stream_process(orc_rds_stream* s, const float* disc, int n, float* out)
{
    if (s->rate <= 0) return 0;
    int m = 0;
    for (int i = 0; i < n; i++) {
        /* phase += -2 pi (-57000) / fs, pre-incremented as FreqShift does */
        s->mix_num = (s->mix_num + 57000) % s->rate;
        const double phase = 2 * ORC_PI * (double)s->mix_num / (double)s->rate;
        double b = disc[i];
        float cr = (float)(b * cos(phase)), ci = (float)(b * sin(phase));
        /* delay line, newest first */
        memmove(s->hr + 1, s->hr, sizeof(double) * (size_t)(s->nt - 1));
        memmove(s->hi + 1, s->hi, sizeof(double) * (size_t)(s->nt - 1));
        s->hr[0] = cr;
        s->hi[0] = ci;
        if (s->count % s->decim == 0) {
            double ar = 0, ai = 0;
            for (int k = 0; k < s->nt; k++) {
                ar += s->hr[k] * s->taps[k];
                ai += s->hi[k] * s->taps[k];
            }
            out[2 * m] = (float)ar;
            out[2 * m + 1] = (float)ai;
            m++;
        }
        s->count++;
    }
    return m;
}
