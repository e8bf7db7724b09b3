/*
 * surv.c -- oracle restatement of the reference's surveillance path:
 * Hann window, FFT, Welch PSD, single-FFT dB spectrum, CFAR (CA, GOSCA,
 * CASO, OS) and the detection list.  Test infrastructure only (oracle.h).
 *
 * FFT: the reference calls gonum.org/v1/gonum v0.15.0 dsp/fourier
 * (go.mod:8; fft/plan.go:3-24), which is not vendored under the reference.
 * Its published definition is the unnormalised forward DFT
 * X[k] = sum_n x[n] exp(-2*pi*i*k*n/N); this file computes that definition
 * in FP64 with an iterative radix-2 transform (N a power of two) and a
 * direct O(N^2) sum otherwise.  Last-ulp agreement with gonum is not
 * claimed; the transform is pinned by brute force, Parseval and tone
 * position (tests/test_oracle_surv.py).
 */
#include "oracle.h"
#include "surv.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

/* internal/fft/fft.go:5-17 */
void orc_hann(int n, double* w)
{
    if (n <= 1) {
        if (n == 1) w[0] = 1;
        return;
    }
    for (int i = 0; i < n; i++) w[i] = 0.5 * (1 - cos(2 * ORC_PI * (double)i / (double)(n - 1)));
}

/* Forward unnormalised DFT of n complex128 values (re/im interleaved). */
void orc_fft(double* x, int n)
{
    if (n <= 1) return;
    if ((n & (n - 1)) != 0) {
        double* y = (double*)malloc(sizeof(double) * 2 * (size_t)n);
        for (int k = 0; k < n; k++) {
            double sr = 0, si = 0;
            for (int t = 0; t < n; t++) {
                double a = -2 * ORC_PI * (double)(((long long)k * t) % n) / (double)n;
                double c = cos(a), s = sin(a);
                sr += x[2 * t] * c - x[2 * t + 1] * s;
                si += x[2 * t] * s + x[2 * t + 1] * c;
            }
            y[2 * k] = sr;
            y[2 * k + 1] = si;
        }
        memcpy(x, y, sizeof(double) * 2 * (size_t)n);
        free(y);
        return;
    }
    /* bit reversal */
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = x[2 * i], ti = x[2 * i + 1];
            x[2 * i] = x[2 * j];
            x[2 * i + 1] = x[2 * j + 1];
            x[2 * j] = tr;
            x[2 * j + 1] = ti;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1;
        for (int k = 0; k < half; k++) {
            /* exact twiddle from the angle (no recurrence drift) */
            double a = -2 * ORC_PI * (double)k / (double)len;
            double wr = cos(a), wi = sin(a);
            for (int i = 0; i < n; i += len) {
                double* u = x + 2 * (i + k);
                double* v = x + 2 * (i + k + half);
                double vr = v[0] * wr - v[1] * wi;
                double vi = v[0] * wi + v[1] * wr;
                v[0] = u[0] - vr;
                v[1] = u[1] - vi;
                u[0] += vr;
                u[1] += vi;
            }
        }
    }
}

/* internal/fft/fft.go:24-55  SpectrumWithPlan: 20*log10(|X|/N + 1e-12), fftshift. */
void orc_spectrum(const float* iq, int n, const double* window, int wlen, double* out)
{
    if (n <= 0) return;
    double* x = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    for (int i = 0; i < n; i++) {
        double w = (wlen == n) ? window[i] : 1.0;
        x[2 * i] = (double)iq[2 * i] * w;
        x[2 * i + 1] = (double)iq[2 * i + 1] * w;
    }
    orc_fft(x, n);
    double invN = 1.0 / (double)n;
    for (int i = 0; i < n; i++) {
        int idx = (i + n / 2) % n;
        double mag = hypot(x[2 * idx], x[2 * idx + 1]) * invN;
        out[i] = 20 * log10(mag + 1e-12);
    }
    free(x);
}

/* internal/fft/fft.go:57-73  SpectrumFromFFT (complex64 FFT output in). */
void orc_spectrum_from_fft(const float* X, int n, double* out)
{
    double invN = 1.0 / (double)n;
    for (int i = 0; i < n; i++) {
        int idx = (i + n / 2) % n;
        double mag = hypot((double)X[2 * idx], (double)X[2 * idx + 1]) * invN;
        out[i] = 20 * log10(mag + 1e-12);
    }
}

/* internal/fft/welch.go:15-66  WelchPSD.  Returns the number of segments
 * averaged (0 -> single zero-padded periodogram fallback). */
int orc_welch_psd(const float* iq, int n, int seg, double overlap, const double* window, int wlen, double* out)
{
    if (seg <= 0) return -1;
    if (overlap < 0) overlap = 0;
    if (overlap > 0.95) overlap = 0.95;
    int step = (int)((double)seg * (1 - overlap));
    if (step < 1) step = 1;
    double* x = (double*)malloc(sizeof(double) * 2 * (size_t)seg);
    double* acc = (double*)calloc((size_t)seg, sizeof(double));
    double invN = 1.0 / (double)seg;
    int segs = 0;
    for (int start = 0; start + seg <= n; start += step) {
        for (int i = 0; i < seg; i++) {
            double w = (wlen == seg) ? window[i] : 1.0;
            x[2 * i] = (double)iq[2 * (start + i)] * w;
            x[2 * i + 1] = (double)iq[2 * (start + i) + 1] * w;
        }
        orc_fft(x, seg);
        for (int i = 0; i < seg; i++) {
            double mag = hypot(x[2 * i], x[2 * i + 1]) * invN;
            acc[i] += mag * mag;
        }
        segs++;
    }
    if (segs == 0) {
        /* not enough samples: Spectrum(iq, window) over the whole input
         * (a length-n transform; the caller's out must hold n values) */
        orc_spectrum(iq, n, window, wlen, out);
    } else {
        for (int i = 0; i < seg; i++) {
            int idx = (i + seg / 2) % seg;
            out[i] = 10 * log10(acc[idx] / (double)segs + 1e-24);
        }
    }
    free(x);
    free(acc);
    return segs;
}

/* ---------------------------------------------------------------- CFAR */

static double at_bin(const double* s, int n, int i, int wrap)
{
    if (wrap) return s[((i % n) + n) % n];
    if (i < 0) return s[0];
    if (i >= n) return s[n - 1];
    return s[i];
}

static double to_linear(double db) { return pow(10, db / 10.0); }

/* internal/cfar/ca.go:22-68 (mode CA), gosca.go:22-78 (GOSCA), caso.go:16-70
 * (CASO): running linear-power sums over the train cells either side.
 * internal/cfar/os.go:31-80 (OS): rank-th of the sorted 2*train dB window. */
static int cmp_dbl(const void* a, const void* b)
{
    double x = *(const double*)a, y = *(const double*)b;
    return (x > y) - (x < y);
}

int orc_cfar_thresholds(int mode, int guard, int train, int rank, double scale_db, int wrap, const double* s, int n,
                        double* out)
{
    if (n <= 0) return 0;
    if (train <= 0) return -1; /* cfar.go:7-9: New returns nil */
    if (guard < 0) guard = 0;
    if (scale_db <= 0) scale_db = 6;
    if (mode == ORC_CFAR_OS) {
        int total = 2 * train;
        int rk = rank - 1;
        if (rk < 0) rk = 0;
        if (rk >= total) rk = total - 1;
        double* win = (double*)malloc(sizeof(double) * (size_t)total);
        for (int i = 0; i < n; i++) {
            /* rebuilding and sorting the window gives the same multiset as the
             * reference's incremental insert/remove (os.go:63-76) */
            int m = 0;
            for (int k = 1; k <= train; k++) {
                win[m++] = at_bin(s, n, i - guard - k, wrap);
                win[m++] = at_bin(s, n, i + guard + k, wrap);
            }
            qsort(win, (size_t)total, sizeof(double), cmp_dbl);
            out[i] = win[rk] + scale_db;
        }
        free(win);
        return 0;
    }
    double left = 0, right = 0;
    for (int k = 1; k <= train; k++) {
        left += to_linear(at_bin(s, n, 0 - guard - k, wrap));
        right += to_linear(at_bin(s, n, 0 + guard + k, wrap));
    }
    for (int i = 0; i < n; i++) {
        if (i > 0) {
            left -= to_linear(at_bin(s, n, i - 1 - guard - train, wrap));
            left += to_linear(at_bin(s, n, i - guard - 1, wrap));
            right -= to_linear(at_bin(s, n, i - 1 + guard + 1, wrap));
            right += to_linear(at_bin(s, n, i + guard + train, wrap));
        }
        double noise;
        if (mode == ORC_CFAR_CA) {
            double invN = 1.0 / (double)(2 * train);
            out[i] = 10 * log10((left + right) * invN) + scale_db;
            continue;
        }
        double inv = 1.0 / (double)train;
        double lm = left * inv, rm = right * inv;
        if (mode == ORC_CFAR_GOSCA) noise = rm > lm ? rm : lm;
        else noise = rm < lm ? rm : lm;
        out[i] = 10 * log10(noise) + scale_db;
    }
    return 0;
}

/* ---------------------------------------------------------------- detection list */

static int cmp_sig_first(const void* a, const void* b)
{
    const orc_signal* x = (const orc_signal*)a;
    const orc_signal* y = (const orc_signal*)b;
    return (x->first_bin > y->first_bin) - (x->first_bin < y->first_bin);
}

/* detector.go:685-696 median (sorted copy; mean of the middle pair) */
double orc_median(const double* v, int n)
{
    if (n <= 0) return 0;
    double* c = (double*)malloc(sizeof(double) * (size_t)n);
    memcpy(c, v, sizeof(double) * (size_t)n);
    qsort(c, (size_t)n, sizeof(double), cmp_dbl);
    int mid = n / 2;
    double m = (n % 2 == 0) ? (c[mid - 1] + c[mid]) / 2 : c[mid];
    free(c);
    return m;
}

static double min_in_range(const double* s, int n, int from, int to)
{
    if (n == 0) return 0;
    if (from < 0) from = 0;
    if (to >= n) to = n - 1;
    if (from > to) {
        int i = from < 0 ? 0 : from;
        if (i > n - 1) i = n - 1;
        return s[i];
    }
    double m = s[from];
    for (int i = from + 1; i <= to; i++)
        if (s[i] < m) m = s[i];
    return m;
}

static double center_for_bin(const orc_detector_cfg* d, double bin)
{
    return d->center_hz + (bin - (double)d->nbins / 2.0) * d->bin_width;
}

/* detector.go:602-639 powerWeightedCenter */
static double power_weighted_center(const orc_detector_cfg* d, const double* s, int n, int first, int last)
{
    if (first > last || first < 0 || last >= n) return center_for_bin(d, (double)(first + last) / 2.0);
    int peak_bin = first;
    double peak_val = s[first];
    for (int i = first; i <= last; i++)
        if (s[i] > peak_val) {
            peak_val = s[i];
            peak_bin = i;
        }
    int win = 1;
    if (d->bin_width > 0) win = (int)(30000.0 / d->bin_width);
    if (win < 1) win = 1;
    int lo = peak_bin - win, hi = peak_bin + win;
    if (lo < first) lo = first;
    if (hi > last) hi = last;
    double sp = 0, sw = 0;
    for (int i = lo; i <= hi; i++) {
        double p = pow(10, s[i] / 10.0);
        sp += p;
        sw += p * (double)i;
    }
    if (sp <= 0) return center_for_bin(d, (double)peak_bin);
    return center_for_bin(d, sw / sp);
}

/* detector.go:698-713 makeSignal */
static orc_signal make_signal(const orc_detector_cfg* d, int first, int last, double peak, int peak_bin,
                              double noise)
{
    orc_signal g;
    g.first_bin = first;
    g.last_bin = last;
    g.peak_bin = peak_bin;
    g.center_hz = d->center_hz + ((double)peak_bin - (double)d->nbins / 2.0) * d->bin_width;
    g.bw_hz = (double)(last - first + 1) * d->bin_width;
    g.peak_db = peak;
    g.snr_db = peak - noise;
    g.noise_db = noise;
    return g;
}

/*
 * detector.go:429-480 (CFAR branch of detectSignals) + expandSignalEdges
 * (483-549) + mergeOverlapping (551-588) + powerWeightedCenter (602-639) on an
 * already-smoothed spectrum.  thr may be NULL (fixed ThresholdDb).  Returns
 * the number of signals (<= cap).  *noise_global_out gets median(spectrum).
 */
int orc_detect_signals(const orc_detector_cfg* d, const double* s, const double* thr, int n, orc_signal* out, int cap,
                       double* noise_global_out)
{
    if (n <= 0) return 0;
    double noise_global = orc_median(s, n);
    if (noise_global_out) *noise_global_out = noise_global;
    int ns = 0;
    int in = 0, start = 0, peak_bin = 0;
    double peak = -1e9;
    int scap = 64;
    orc_signal* sig = (orc_signal*)malloc(sizeof(orc_signal) * (size_t)scap);
    for (int i = 0; i <= n; i++) {
        int closing = 0, end = 0;
        if (i < n) {
            double v = s[i];
            double on = d->threshold_db;
            if (thr && !isnan(thr[i])) on = thr[i];
            double off = on - d->hysteresis_db;
            if (v >= on) {
                if (!in) {
                    in = 1;
                    start = i;
                    peak = v;
                    peak_bin = i;
                } else if (v > peak) {
                    peak = v;
                    peak_bin = i;
                }
            } else if (in && v < off) {
                closing = 1;
                end = i - 1;
            }
        } else if (in) {
            closing = 1;
            end = n - 1;
        }
        if (closing) {
            double noise = noise_global;
            if (thr && peak_bin >= 0 && peak_bin < n && !isnan(thr[peak_bin])) noise = thr[peak_bin] - d->cfar_scale_db;
            if (ns == scap) {
                scap *= 2;
                sig = (orc_signal*)realloc(sig, sizeof(orc_signal) * (size_t)scap);
            }
            sig[ns++] = make_signal(d, start, end, peak, peak_bin, noise);
            in = 0;
        }
    }
    /* expandSignalEdges */
    double margin = d->edge_margin_db > 0 ? d->edge_margin_db : 3.0;
    int max_exp = (int)(d->max_signal_bw_hz / d->bin_width);
    if (max_exp < 10) max_exp = 10;
    for (int i = 0; i < ns; i++) {
        orc_signal* g = &sig[i];
        double local = noise_global;
        int lp = g->first_bin - 50, rp = g->last_bin + 50;
        if (lp >= 0 && rp < n) {
            int a0 = lp > 0 ? lp : 0, a1 = g->first_bin - 5 > 0 ? g->first_bin - 5 : 0;
            int b0 = g->last_bin + 5 < n - 1 ? g->last_bin + 5 : n - 1, b1 = rp < n - 1 ? rp : n - 1;
            double ln = min_in_range(s, n, a0, a1);
            double rn = min_in_range(s, n, b0, b1);
            local = ln < rn ? ln : rn;
        }
        double edge = local + margin;
        int nf = g->first_bin;
        double prev = s[nf];
        for (int j = 0; j < max_exp; j++) {
            int nx = nf - 1;
            if (nx < 0) break;
            double val = s[nx];
            if (val <= edge) break;
            if (val > prev + 1.0 && val < g->peak_db - 6.0) break;
            prev = val;
            nf = nx;
        }
        int nl = g->last_bin;
        prev = s[nl];
        for (int j = 0; j < max_exp; j++) {
            int nx = nl + 1;
            if (nx >= n) break;
            double val = s[nx];
            if (val <= edge) break;
            if (val > prev + 1.0 && val < g->peak_db - 6.0) break;
            prev = val;
            nl = nx;
        }
        g->first_bin = nf;
        g->last_bin = nl;
        g->bw_hz = (double)(nl - nf + 1) * d->bin_width;
    }
    /* mergeOverlapping */
    if (ns > 1) {
        int gap_bins = 0;
        if (d->merge_gap_hz > 0 && d->bin_width > 0) gap_bins = (int)ceil(d->merge_gap_hz / d->bin_width);
        /* sort.Slice is not stable; the reference only ever sorts seeds with
         * distinct FirstBin unless two seeds expanded to the same first bin,
         * in which case the order of equal keys is unspecified there too.
         * A stable order (by first bin, then original index) is used here. */
        for (int i = 1; i < ns; i++) {
            orc_signal key = sig[i];
            int j = i - 1;
            while (j >= 0 && sig[j].first_bin > key.first_bin) {
                sig[j + 1] = sig[j];
                j--;
            }
            sig[j + 1] = key;
        }
        int m = 1;
        for (int i = 1; i < ns; i++) {
            orc_signal* last = &sig[m - 1];
            orc_signal cur = sig[i];
            int gap = cur.first_bin - last->last_bin - 1;
            if (gap <= gap_bins) {
                if (cur.last_bin > last->last_bin) last->last_bin = cur.last_bin;
                if (cur.peak_db > last->peak_db) last->peak_db = cur.peak_db;
                if (cur.snr_db > last->snr_db) last->snr_db = cur.snr_db;
                double cb = (double)(last->first_bin + last->last_bin) / 2.0;
                last->bw_hz = (double)(last->last_bin - last->first_bin + 1) * d->bin_width;
                last->center_hz = center_for_bin(d, cb);
                if (cur.noise_db < last->noise_db || last->noise_db == 0) last->noise_db = cur.noise_db;
            } else {
                sig[m++] = cur;
            }
        }
        ns = m;
    }
    (void)cmp_sig_first;
    int nout = ns < cap ? ns : cap;
    for (int i = 0; i < nout; i++) {
        sig[i].center_hz = power_weighted_center(d, s, n, sig[i].first_bin, sig[i].last_bin);
        sig[i].bw_hz = (double)(sig[i].last_bin - sig[i].first_bin + 1) * d->bin_width;
        out[i] = sig[i];
    }
    free(sig);
    return ns;
}
