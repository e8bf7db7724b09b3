/*
 * extract.c -- oracle restatement of the reference's streaming polyphase
 * extractor (internal/demod/gpudemod).  Test infrastructure only (oracle.h).
 *
 * Float arithmetic follows Go's rules for complex64 on amd64:
 *   complex64 * complex64 is evaluated in float64 and narrowed once to float32
 *   (cmd/compile ssagen, "Compute in Float64 to minimize cancellation error");
 *   complex64 + complex64 is two float32 adds.  A real tap times a sample is
 *   therefore fl32(tap*x) followed by a float32 add, which plain float code
 *   with -ffp-contract=off reproduces bit for bit.
 */
#include "oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define ORC_PI 3.14159265358979323846

static int imax(int a, int b) { return a > b ? a : b; }

/* polyphase.go:24-28 */
int orc_polyphase_phase_len(int base_len, int decim)
{
    if (decim <= 0 || base_len <= 0) return 0;
    return (base_len + decim - 1) / decim;
}

/* polyphase.go:6-22  out[p*L+k] = base[p + k*D], zero padded. */
int orc_build_polyphase(const float* base, int n, int decim, float* out)
{
    if (decim <= 0 || n <= 0) return 0;
    int L = (n + decim - 1) / decim;
    for (int p = 0; p < decim; p++)
        for (int k = 0; k < L; k++) {
            int src = p + k * decim;
            out[p * L + k] = src < n ? base[src] : 0.0f;
        }
    return decim * L;
}

/* cpu_oracle.go:155-163 */
int orc_exact_decim(int sample_rate, int out_rate)
{
    if (sample_rate <= 0 || out_rate <= 0) return -1;
    if (sample_rate % out_rate != 0) return -1;
    return sample_rate / out_rate;
}

/* streaming_types.go:58-72  FNV-1a-64 over "sig=%d|out=%d|taps=%d|sr=%d".
 * Offset and bandwidth are deliberately not part of the hash. */
uint64_t orc_config_hash(int64_t signal_id, double offset_hz, double bandwidth, int out_rate, int num_taps,
                         int sample_rate)
{
    (void)offset_hz;
    (void)bandwidth;
    char buf[128];
    int len = snprintf(buf, sizeof buf, "sig=%lld|out=%d|taps=%d|sr=%d", (long long)signal_id, out_rate,
                       num_taps, sample_rate);
    uint64_t h = 14695981039346656037ULL;
    for (int i = 0; i < len; i++) {
        h ^= (uint8_t)buf[i];
        h *= 1099511628211ULL;
    }
    return h;
}

static inline void wrap_phase(double* phase)
{
    if (*phase >= ORC_PI) *phase -= 2 * ORC_PI;
    else if (*phase < -ORC_PI) *phase += 2 * ORC_PI;
}

/* s = x * complex64(cos(phase), sin(phase)) with Go's complex64 product. */
static inline void nco_rotate(const float* x, double phase, float* s)
{
    float cr = (float)cos(phase), ci = (float)sin(phase);
    double xr = x[0], xi = x[1];
    s[0] = (float)(xr * (double)cr - xi * (double)ci);
    s[1] = (float)(xr * (double)ci + xi * (double)cr);
}

/* Keep the newest keep samples of a history buffer of *len samples. */
static void trim_history(float* hist, int* len, int keep)
{
    if (keep < 0) keep = 0;
    if (*len > keep) {
        memmove(hist, hist + 2 * (*len - keep), sizeof(float) * 2 * (size_t)keep);
        *len = keep;
    }
}

/*
 * streaming_host_core.go:5-64  runStreamingPolyphaseHostCore.
 * history_io must hold at least num_taps-1 (and >= 1) complex samples plus
 * room for one appended sample.  Returns n_out (or -1 if out_cap too small).
 */
int orc_host_core(const float* iq_new, int n_new, int sample_rate, double offset_hz, double* nco_phase_io,
                  int* phase_count_io, int num_taps, int decim, float* history_io, int* history_len_io,
                  const float* poly, int poly_len, float* out, int out_cap)
{
    int D = imax(1, decim);
    int phase_len = orc_polyphase_phase_len(poly_len / D * D, decim);
    if (phase_len == 0) phase_len = orc_polyphase_phase_len(poly_len, decim);
    double phase = *nco_phase_io;
    int pc = *phase_count_io;
    int hlen = *history_len_io;
    double inc = -2.0 * ORC_PI * offset_hz / (double)sample_rate;
    int keep = num_taps - 1;
    int n_out = 0;
    for (int i = 0; i < n_new; i++) {
        nco_rotate(iq_new + 2 * i, phase, history_io + 2 * hlen);
        hlen++;
        pc++;
        if (pc == decim) {
            float yr = 0.0f, yi = 0.0f;
            for (int p = 0; p < decim; p++) {
                for (int k = 0; k < phase_len; k++) {
                    int it = p * phase_len + k;
                    if (it >= poly_len) continue;
                    float tap = poly[it];
                    if (tap == 0) continue;
                    int idx = hlen - 1 - (p + k * decim);
                    if (idx < 0) continue;
                    float pr = tap * history_io[2 * idx];
                    float pi = tap * history_io[2 * idx + 1];
                    yr += pr;
                    yi += pi;
                }
            }
            if (n_out >= out_cap) return -1;
            out[2 * n_out] = yr;
            out[2 * n_out + 1] = yi;
            n_out++;
            pc = 0;
        }
        if (hlen > keep) trim_history(history_io, &hlen, keep);
        phase += inc;
        wrap_phase(&phase);
    }
    *nco_phase_io = phase;
    *phase_count_io = pc;
    *history_len_io = hlen;
    return n_out;
}

/* cpu_oracle.go:32-75  CPUOracleExtract (direct form, k = 0..T-1). */
int orc_extract_direct(const float* iq_new, int n_new, double phase_inc, double* nco_phase_io,
                       int* phase_count_io, int num_taps, int decim, const float* base, float* hist,
                       int* history_len_io, float* out, int out_cap)
{
    if (num_taps <= 0 || decim <= 0) return 0;
    double phase = *nco_phase_io;
    int pc = *phase_count_io, hlen = *history_len_io, n_out = 0;
    for (int i = 0; i < n_new; i++) {
        nco_rotate(iq_new + 2 * i, phase, hist + 2 * hlen);
        hlen++;
        pc++;
        if (pc == decim) {
            float yr = 0.0f, yi = 0.0f;
            for (int k = 0; k < num_taps; k++) {
                int idx = hlen - 1 - k;
                float sr = 0.0f, si = 0.0f;
                if (idx >= 0) {
                    sr = hist[2 * idx];
                    si = hist[2 * idx + 1];
                }
                float pr = base[k] * sr, pi = base[k] * si;
                yr += pr;
                yi += pi;
            }
            if (n_out >= out_cap) return -1;
            out[2 * n_out] = yr;
            out[2 * n_out + 1] = yi;
            n_out++;
            pc = 0;
        }
        if (hlen > num_taps - 1) trim_history(hist, &hlen, num_taps - 1);
        phase += phase_inc;
        wrap_phase(&phase);
    }
    *nco_phase_io = phase;
    *phase_count_io = pc;
    *history_len_io = hlen;
    return n_out;
}

/* cpu_oracle.go:79-132  CPUOracleExtractPolyphase. */
int orc_extract_polyphase(const float* iq_new, int n_new, double phase_inc, double* nco_phase_io,
                          int* phase_count_io, int num_taps, int decim, const float* base, const float* poly,
                          float* hist, int* history_len_io, float* out, int out_cap)
{
    (void)base;
    if (num_taps <= 0 || decim <= 0) return 0;
    int L = orc_polyphase_phase_len(num_taps, decim);
    double phase = *nco_phase_io;
    int pc = *phase_count_io, hlen = *history_len_io, n_out = 0;
    for (int i = 0; i < n_new; i++) {
        nco_rotate(iq_new + 2 * i, phase, hist + 2 * hlen);
        hlen++;
        pc++;
        if (pc == decim) {
            float yr = 0.0f, yi = 0.0f;
            for (int p = 0; p < decim; p++)
                for (int k = 0; k < L; k++) {
                    float tap = poly[p * L + k];
                    if (tap == 0) continue;
                    int idx = hlen - 1 - (p + k * decim);
                    if (idx < 0) continue;
                    float pr = tap * hist[2 * idx], pi = tap * hist[2 * idx + 1];
                    yr += pr;
                    yi += pi;
                }
            if (n_out >= out_cap) return -1;
            out[2 * n_out] = yr;
            out[2 * n_out + 1] = yi;
            n_out++;
            pc = 0;
        }
        if (hlen > num_taps - 1) trim_history(hist, &hlen, num_taps - 1);
        phase += phase_inc;
        wrap_phase(&phase);
    }
    *nco_phase_io = phase;
    *phase_count_io = pc;
    *history_len_io = hlen;
    return n_out;
}

/* ---------------- CPUOracleRunner (streaming_oracle_extract.go) ---------------- */

typedef struct orc_state {
    int64_t signal_id;
    uint64_t config_hash;
    double nco_phase;
    int decim, phase_count, num_taps;
    float* hist; /* capacity hist_cap + 1 complex */
    int hist_len, hist_cap;
    float* base;
    int base_len;
    float* poly;
    int poly_len;
    int live;
} orc_state;

struct orc_runner {
    int sample_rate;
    orc_state* st;
    int n, cap;
    int threads;
};

orc_runner* orc_runner_new(int sample_rate)
{
    orc_runner* r = (orc_runner*)calloc(1, sizeof(orc_runner));
    r->sample_rate = sample_rate;
    r->threads = 1;
    return r;
}

static void free_state(orc_state* s)
{
    free(s->hist);
    free(s->base);
    free(s->poly);
    memset(s, 0, sizeof *s);
}

void orc_runner_free(orc_runner* r)
{
    if (!r) return;
    for (int i = 0; i < r->n; i++) free_state(&r->st[i]);
    free(r->st);
    free(r);
}

int orc_runner_set_threads(orc_runner* r, int threads)
{
    r->threads = threads < 1 ? 1 : threads;
    return r->threads;
}

static orc_state* find_state(const orc_runner* r, int64_t id)
{
    for (int i = 0; i < r->n; i++)
        if (r->st[i].signal_id == id) return &r->st[i];
    return NULL;
}

static void remove_state_at(orc_runner* r, int i)
{
    free_state(&r->st[i]);
    r->st[i] = r->st[r->n - 1];
    memset(&r->st[r->n - 1], 0, sizeof(orc_state));
    r->n--;
}

/* streaming_oracle_extract.go:35-71  getOrInitState. */
static orc_state* get_or_init_state(orc_runner* r, const orc_job* job)
{
    int decim = orc_exact_decim(r->sample_rate, job->out_rate);
    if (decim < 0) return NULL;
    orc_state* s = find_state(r, job->signal_id);
    if (!s) {
        if (r->n == r->cap) {
            r->cap = r->cap ? 2 * r->cap : 16;
            r->st = (orc_state*)realloc(r->st, sizeof(orc_state) * (size_t)r->cap);
        }
        s = &r->st[r->n++];
        memset(s, 0, sizeof *s);
        s->signal_id = job->signal_id;
    }
    /* cpu_oracle.go:20-30  ResetCPUOracleStateIfConfigChanged */
    if (s->config_hash != job->config_hash) {
        s->config_hash = job->config_hash;
        s->nco_phase = 0;
        s->phase_count = 0;
        s->hist_len = 0;
    }
    s->decim = decim;
    s->num_taps = job->num_taps > 0 ? job->num_taps : 101;
    double cutoff = job->bandwidth / 2;
    if (cutoff < 200) cutoff = 200;
    double* base64 = (double*)malloc(sizeof(double) * (size_t)(s->num_taps + 1));
    int nb = orc_lowpass_fir(cutoff, r->sample_rate, s->num_taps, base64);
    free(s->base);
    s->base = (float*)malloc(sizeof(float) * (size_t)nb);
    for (int i = 0; i < nb; i++) s->base[i] = (float)base64[i];
    s->base_len = nb;
    free(base64);
    int L = orc_polyphase_phase_len(nb, decim);
    free(s->poly);
    s->poly = (float*)malloc(sizeof(float) * (size_t)(decim * L));
    s->poly_len = orc_build_polyphase(s->base, nb, decim, s->poly);
    int cap = imax(1, s->num_taps - 1);
    if (s->hist_cap < cap) {
        float* nh = (float*)calloc((size_t)(cap + 1) * 2, sizeof(float));
        if (s->hist && s->hist_len) memcpy(nh, s->hist, sizeof(float) * 2 * (size_t)s->hist_len);
        free(s->hist);
        s->hist = nh;
        s->hist_cap = cap;
    }
    if (s->hist_len > s->num_taps - 1) trim_history(s->hist, &s->hist_len, imax(0, s->num_taps - 1));
    s->live = 1;
    return s;
}

typedef struct {
    orc_runner* r;
    const float* iq;
    int n;
    const orc_job* jobs;
    orc_state** states;
    float* out;
    orc_result* results;
    int first, step, count;
    int err;
} orc_work;

static void* run_jobs(void* arg)
{
    orc_work* w = (orc_work*)arg;
    for (int i = w->first; i < w->count; i += w->step) {
        orc_state* s = w->states[i];
        const orc_job* job = &w->jobs[i];
        int cap = w->results[i].n_out;
        int n_out = orc_host_core(w->iq, w->n, w->r->sample_rate, job->offset_hz, &s->nco_phase, &s->phase_count,
                                  s->num_taps, s->decim, s->hist, &s->hist_len, s->poly, s->poly_len,
                                  w->out + 2 * w->results[i].out_offset, cap);
        if (n_out < 0) w->err = -14;
        w->results[i].n_out = n_out;
        w->results[i].phase_count = s->phase_count;
        w->results[i].history_len = s->hist_len;
        w->results[i].nco_phase = s->nco_phase;
    }
    return NULL;
}

/* streaming_oracle_extract.go:73-111  CPUOracleRunner.StreamingExtract.
 * Outputs of job i land at out + 2*results[i].out_offset (complex offsets
 * assigned in job order).  Returns 0, -1 on a non-integer decimation, -14 if
 * out_cap is too small. */
int orc_runner_streaming_extract(orc_runner* r, const float* iq, int n, const orc_job* jobs, int n_jobs,
                                 float* out, int64_t out_cap, orc_result* results)
{
    if (n < 0 || n_jobs < 0) return -11;
    orc_state** states = (orc_state**)malloc(sizeof(orc_state*) * (size_t)(n_jobs > 0 ? n_jobs : 1));
    for (int i = 0; i < r->n; i++) r->st[i].live = 0;
    int64_t off = 0;
    for (int i = 0; i < n_jobs; i++) {
        orc_state* s = get_or_init_state(r, &jobs[i]);
        if (!s) {
            free(states);
            return -1;
        }
        states[i] = NULL; /* pointers resolved after all inits (realloc may move) */
        results[i].signal_id = jobs[i].signal_id;
        results[i].rate = jobs[i].out_rate;
        results[i].n_out = (s->phase_count + n) / s->decim;
        results[i].out_offset = off;
        off += results[i].n_out;
    }
    if (off > out_cap) {
        free(states);
        return -14;
    }
    for (int i = 0; i < n_jobs; i++) states[i] = find_state(r, jobs[i].signal_id);
    int T = r->threads;
    if (T > n_jobs) T = n_jobs;
    if (T <= 1) {
        orc_work w = {r, iq, n, jobs, states, out, results, 0, 1, n_jobs, 0};
        run_jobs(&w);
        if (w.err) {
            free(states);
            return w.err;
        }
    } else {
        pthread_t th[256];
        orc_work w[256];
        if (T > 256) T = 256;
        for (int t = 0; t < T; t++) {
            orc_work wt = {r, iq, n, jobs, states, out, results, t, T, n_jobs, 0};
            w[t] = wt;
            pthread_create(&th[t], NULL, run_jobs, &w[t]);
        }
        int err = 0;
        for (int t = 0; t < T; t++) {
            pthread_join(th[t], NULL);
            if (w[t].err) err = w[t].err;
        }
        if (err) {
            free(states);
            return err;
        }
    }
    free(states);
    /* prune vanished signals */
    for (int i = r->n - 1; i >= 0; i--)
        if (!r->st[i].live) remove_state_at(r, i);
    return 0;
}

int orc_runner_num_states(const orc_runner* r) { return r->n; }

int orc_runner_state_ids(const orc_runner* r, int64_t* ids, int cap)
{
    int m = r->n < cap ? r->n : cap;
    for (int i = 0; i < m; i++) ids[i] = r->st[i].signal_id;
    return r->n;
}

int orc_runner_get_state(const orc_runner* r, int64_t id, uint64_t* config_hash, double* nco_phase, int* decim,
                         int* phase_count, int* num_taps, float* history, int hist_cap)
{
    const orc_state* s = find_state(r, id);
    if (!s) return -1;
    if (config_hash) *config_hash = s->config_hash;
    if (nco_phase) *nco_phase = s->nco_phase;
    if (decim) *decim = s->decim;
    if (phase_count) *phase_count = s->phase_count;
    if (num_taps) *num_taps = s->num_taps;
    if (history) {
        int m = s->hist_len < hist_cap ? s->hist_len : hist_cap;
        memcpy(history, s->hist, sizeof(float) * 2 * (size_t)m);
    }
    return s->hist_len;
}

void orc_runner_reset_signal(orc_runner* r, int64_t id)
{
    for (int i = 0; i < r->n; i++)
        if (r->st[i].signal_id == id) {
            remove_state_at(r, i);
            return;
        }
}

void orc_runner_reset_all(orc_runner* r)
{
    while (r->n) remove_state_at(r, r->n - 1);
}
