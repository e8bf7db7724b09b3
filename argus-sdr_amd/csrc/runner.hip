// runner.hip -- host runtime of the streaming batch runner (successor of the
// reference's gpudemod.BatchRunner streaming surface).  Owns per-signal
// device state (taps, ping-pong shifted history) and host scalars (NCO
// phase, decimator phase), builds the per-chunk channel table and work list,
// and issues one fused launch per chunk (or, in the plain per-signal mode,
// the reference's per-signal call pattern).
#include "extract.hpp"
#include "runner.hpp"
#include "outring.hpp"
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <vector>

int argus_launch_extract_one(const ArgusChanDesc& ch, const float2* in, int n_new, hipStream_t stream);

// ---------------------------------------------------------------------------
// Tap design (dsp/fir.go:6-28) and polyphase layout (polyphase.go:6-22).
// ---------------------------------------------------------------------------
int argus_lowpass_fir(double cutoff_hz, int sample_rate, int taps, double* out, bool hamming)
{
    if (taps % 2 == 0) taps++;
    for (int n = 0; n < taps; n++) out[n] = 0.0;
    const double fc = cutoff_hz / (double)sample_rate;
    if (fc <= 0) return taps;
    const double m = (double)(taps - 1) / 2.0;
    for (int n = 0; n < taps; n++) {
        const double x = (double)n - m;
        const double sinc = x == 0 ? 2 * fc : std::sin(2 * M_PI * fc * x) / (M_PI * x);
        const double ph = 2 * M_PI * (double)n / (double)(taps - 1);
        const double w = hamming ? 0.54 - 0.46 * std::cos(ph) : 0.5 * (1 - std::cos(ph));
        out[n] = sinc * w;
    }
    return taps;
}

static int build_poly(const float* base, int n, int decim, float* out)
{
    if (decim <= 0 || n <= 0) return 0;
    const int L = (n + decim - 1) / decim;
    for (int p = 0; p < decim; p++)
        for (int k = 0; k < L; k++) {
            const int src = p + k * decim;
            out[p * L + k] = src < n ? base[src] : 0.0f;
        }
    return decim * L;
}

extern "C" GPUD_API int gpud_runner_get_tune(gpud_runner* r, gpud_tune* t)
{
    if (!r || !t) return -1;
    *t = gpud_tune{r->tune.lds_entries, r->tune.narrow_outputs, r->tune.grid, r->tune.fast_grid, 0.0f, 0.0f};
    return 0;
}

extern "C" GPUD_API int gpud_runner_set_tune(gpud_runner* r, const gpud_tune* t)
{
    if (!r || !t || t->lds_entries < 1024 || t->lds_entries > ARGUS_MAX_LDS_ENTRIES || t->narrow_outputs < 1 ||
        t->narrow_outputs > 1024 || t->grid < ARGUS_XCDS || t->grid > 65536 || t->grid % ARGUS_XCDS != 0 ||
        t->fast_grid < 0 || t->fast_grid > 65536 || t->fast_grid % ARGUS_XCDS != 0)
        return -1;
    r->tune = ArgusTune{t->lds_entries, t->narrow_outputs, t->grid, t->fast_grid ? t->fast_grid : r->tune.fast_grid};
    return 0;
}

extern "C" GPUD_API int gpud_lowpass_fir(double cutoff_hz, int sample_rate, int taps, double* out)
{
    if (!out || sample_rate <= 0 || taps <= 0) return -1;
    return argus_lowpass_fir(cutoff_hz, sample_rate, taps, out, false);
}

extern "C" GPUD_API int gpud_build_polyphase_taps(const float* base, int n, int decim, float* out)
{
    if (!base || !out) return -1;
    return build_poly(base, n, decim, out);
}

extern "C" GPUD_API int gpud_exact_integer_decimation(int sample_rate, int out_rate)
{
    if (sample_rate <= 0 || out_rate <= 0 || sample_rate % out_rate != 0) return -1;
    return sample_rate / out_rate;
}

extern "C" GPUD_API uint64_t gpud_streaming_config_hash(int64_t signal_id, double offset_hz, double bandwidth,
                                                        int out_rate, int num_taps, int sample_rate)
{
    (void)offset_hz;
    (void)bandwidth;
    char buf[128];
    const int len = std::snprintf(buf, sizeof buf, "sig=%lld|out=%d|taps=%d|sr=%d", (long long)signal_id,
                                  out_rate, num_taps, sample_rate);
    uint64_t h = 14695981039346656037ULL;  // FNV-1a 64 offset basis
    for (int i = 0; i < len; i++) {
        h ^= (uint8_t)buf[i];
        h *= 1099511628211ULL;             // FNV-1a 64 prime
    }
    return h;
}

// ---------------------------------------------------------------------------
// Runner
// ---------------------------------------------------------------------------
namespace {

template <typename T>
int ensure_dev(T** p, size_t* cap, size_t need)
{
    if (*p && *cap >= need) return 0;
    if (*p) hipFree(*p);
    *p = nullptr;
    *cap = 0;
    size_t n = need < 16 ? 16 : need + need / 2;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return (int)e;
    *cap = n;
    return 0;
}

template <typename T>
int ensure_host(T** p, size_t* cap, size_t need)
{
    if (*p && *cap >= need) return 0;
    if (*p) hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    size_t n = need < 16 ? 16 : need + need / 2;
    hipError_t e = hipHostMalloc((void**)p, n * sizeof(T), hipHostMallocDefault);
    if (e != hipSuccess) return (int)e;
    *cap = n;
    return 0;
}

}  // namespace

void argus_signal_release(ArgusSignal& s)
{
    if (s.d_taps) hipFree(s.d_taps);
    if (s.d_ctaps) hipFree(s.d_ctaps);
    if (s.d_hist[0]) hipFree(s.d_hist[0]);
    if (s.d_hist[1]) hipFree(s.d_hist[1]);
    if (s.d_out) hipFree(s.d_out);
    if (s.d_fastG) hipFree(s.d_fastG);
    s.d_fastG = nullptr;
    s.fastG_valid = false;
    s.d_taps = nullptr;
    s.d_ctaps = nullptr;
    s.d_ctaps_len = 0;
    s.ctaps_valid = false;
    s.d_hist[0] = s.d_hist[1] = nullptr;
    s.d_out = nullptr;
    s.out_cap = 0;
}

ArgusSignal* argus_runner_find(gpud_runner* r, int64_t id)
{
    for (auto& s : r->signals)
        if (s.signal_id == id) return &s;
    return nullptr;
}

// stream_state.go:25-85 getOrInitExtractState, plus the device side of
// streaming_gpu_native_prepare.go:119-185 (buffers sized to the geometry;
// taps re-uploaded whenever they change).
int argus_runner_prepare_signal(gpud_runner* r, const gpud_streaming_job& job, ArgusSignal** out)
{
    const int decim = gpud_exact_integer_decimation(r->sample_rate, job.out_rate);
    if (decim < 0) return -1;
    ArgusSignal* s = argus_runner_find(r, job.signal_id);
    if (!s) {
        r->signals.emplace_back();
        s = &r->signals.back();
        s->signal_id = job.signal_id;
    }
    if (s->config_hash != job.config_hash || !s->initialized) {
        // streaming_types.go:46-56 ResetExtractStreamState
        s->config_hash = job.config_hash;
        s->nco_phase = 0.0;
        s->phase_count = 0;
        s->hist_len = 0;
        s->base.clear();
    }
    const int T = job.num_taps > 0 ? job.num_taps : 101;
    if (T > 4095) return -1;
    s->window = r->window;
    double cutoff = job.bandwidth / 2;
    if (cutoff < 200) cutoff = 200;
    const bool geometry_changed = s->decim != decim || s->num_taps != T;
    s->decim = decim;
    s->num_taps = T;
    if (s->base.empty() || cutoff != s->taps_cutoff || T != s->taps_num || decim != s->taps_decim ||
        r->sample_rate != s->taps_rate || s->window != s->taps_window) {
        std::vector<double> b64((size_t)T + 1);
        const int nb = argus_lowpass_fir(cutoff, r->sample_rate, T, b64.data(), s->window == 1);
        s->base.assign((size_t)nb, 0.0f);
        for (int i = 0; i < nb; i++) s->base[(size_t)i] = (float)b64[(size_t)i];
        // The reference applies only taps whose delay fits the T-1 sample
        // history plus the newest sample (streaming_host_core.go:44-47), i.e.
        // the first T taps even when LowpassFIR rounded an even T up.
        const int L = (T + decim - 1) / decim;
        s->poly.assign((size_t)decim * L, 0.0f);
        build_poly(s->base.data(), std::min(nb, T), decim, s->poly.data());
        s->taps_cutoff = cutoff;
        s->taps_num = T;
        s->taps_decim = decim;
        s->taps_rate = r->sample_rate;
        s->taps_window = s->window;
        s->taps_dirty = true;
    }
    if (geometry_changed || !s->d_hist[0]) {
        if (s->d_hist[0]) hipFree(s->d_hist[0]);
        if (s->d_hist[1]) hipFree(s->d_hist[1]);
        s->d_hist[0] = s->d_hist[1] = nullptr;
        const size_t cap = (size_t)std::max(1, T - 1);
        ARGUS_CHECK(hipMalloc((void**)&s->d_hist[0], cap * sizeof(float2)));
        ARGUS_CHECK(hipMalloc((void**)&s->d_hist[1], cap * sizeof(float2)));
        s->hist_sel = 0;
        if (s->hist_len > T - 1) s->hist_len = 0;
    }
    // modulated taps follow the taps and the NCO increment (the increment the
    // batched launch derives from the same offset)
    const uint64_t inc = argus_rad_to_fix(-2.0 * M_PI * job.offset_hz / (double)r->sample_rate);
    if (s->taps_dirty || !s->ctaps_valid || s->ctaps_inc != inc) {
        std::vector<float2> ct((size_t)T);
        argus_modulated_taps(s->poly.data(), decim, T, inc, ct.data());
        if (s->d_ctaps && s->d_ctaps_len != ct.size()) {
            hipFree(s->d_ctaps);
            s->d_ctaps = nullptr;
        }
        if (!s->d_ctaps) {
            ARGUS_CHECK(hipMalloc((void**)&s->d_ctaps, ct.size() * sizeof(float2)));
            s->d_ctaps_len = ct.size();
        }
        ARGUS_CHECK(hipMemcpyAsync(s->d_ctaps, ct.data(), ct.size() * sizeof(float2), hipMemcpyHostToDevice,
                                   r->stream));
        ARGUS_CHECK(hipStreamSynchronize(r->stream));  // ct is pageable host memory
        s->ctaps_inc = inc;
        s->ctaps_valid = true;
    }
    // the fast form's response follows the taps and the increment too
    if (argus_fast_eligible(decim, T)) {
        if (s->taps_dirty || !s->fastG_valid || s->fastG_inc != inc) {
            const int L = (T + decim - 1) / decim;
            std::vector<float> h((size_t)T);
            for (int k = 0; k < T; k++) h[(size_t)k] = s->poly[(size_t)(k % decim) * L + k / decim];
            std::vector<float2> G((size_t)ARGUS_FAST_N);
            argus_fast_spectrum(h.data(), T, inc, G.data());
            if (!s->d_fastG) ARGUS_CHECK(hipMalloc((void**)&s->d_fastG, (size_t)ARGUS_FAST_N * sizeof(float2)));
            ARGUS_CHECK(hipMemcpyAsync(s->d_fastG, G.data(), (size_t)ARGUS_FAST_N * sizeof(float2),
                                       hipMemcpyHostToDevice, r->stream));
            ARGUS_CHECK(hipStreamSynchronize(r->stream));  // G is pageable host memory
            s->fastG_inc = inc;
            s->fastG_valid = true;
        }
    } else if (s->d_fastG) {
        hipFree(s->d_fastG);
        s->d_fastG = nullptr;
        s->fastG_valid = false;
    }
    if (s->taps_dirty) {
        if (s->d_taps && s->d_taps_len != s->poly.size()) {
            hipFree(s->d_taps);
            s->d_taps = nullptr;
        }
        if (!s->d_taps) {
            ARGUS_CHECK(hipMalloc((void**)&s->d_taps, s->poly.size() * sizeof(float)));
            s->d_taps_len = s->poly.size();
        }
        ARGUS_CHECK(hipMemcpyAsync(s->d_taps, s->poly.data(), s->poly.size() * sizeof(float),
                                   hipMemcpyHostToDevice, r->stream));
        ARGUS_CHECK(hipStreamSynchronize(r->stream));  // poly is pageable host memory
        s->taps_dirty = false;
    }
    s->initialized = true;
    s->active = true;
    *out = s;
    return 0;
}

// streaming_gpu_native_prepare.go:211-223 syncNativeStreamingStates
void argus_runner_prune(gpud_runner* r)
{
    for (size_t i = 0; i < r->signals.size();) {
        if (!r->signals[i].active) {
            argus_signal_release(r->signals[i]);
            r->signals[i] = r->signals.back();
            r->signals.back() = ArgusSignal();
            r->signals.pop_back();
        } else {
            i++;
        }
    }
}

extern "C" {

GPUD_API int gpud_runner_create(int max_samples, int sample_rate, gpud_runner** out)
{
    if (!out || max_samples <= 0 || sample_rate <= 0) return -1;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return (int)hipErrorNoDevice;
    gpud_runner* r = new gpud_runner();
    r->max_samples = max_samples;
    r->sample_rate = sample_rate;
    hipError_t e = hipStreamCreate(&r->stream);
    if (e != hipSuccess) {
        delete r;
        return (int)e;
    }
    for (int i = 0; i < 2; i++) {
        e = hipEventCreateWithFlags(&r->stage_ev[i], hipEventDisableTiming);
        if (e != hipSuccess) return (int)e;
    }
    e = hipMalloc((void**)&r->d_in, (size_t)max_samples * sizeof(float2));
    if (e != hipSuccess) return (int)e;
    e = hipHostMalloc((void**)&r->h_in, (size_t)max_samples * sizeof(float2), hipHostMallocDefault);
    if (e != hipSuccess) return (int)e;
    *out = r;
    return 0;
}

GPUD_API int gpud_runner_destroy(gpud_runner* r)
{
    if (!r) return 0;
    hipStreamSynchronize(r->stream);
    for (auto& s : r->signals) argus_signal_release(s);
    hipFree(r->d_in);
    hipHostFree(r->h_in);
    hipFree(r->d_out);
    hipHostFree(r->h_out);
    for (int i = 0; i < 2; i++) {
        hipFree(r->d_chans[i]);
        hipFree(r->d_items[i]);
        hipHostFree(r->h_chans[i]);
        hipHostFree(r->h_items[i]);
        hipEventDestroy(r->stage_ev[i]);
    }
    for (hipEvent_t ev : r->tev) hipEventDestroy(ev);
    argus_legacy_release(r->legacy);
    argus_out_rings_free(r->rings);
    hipFree(r->d_ftail);
    hipFree(r->d_fu);
    hipFree(r->d_fX);
    for (int i = 0; i < 2; i++) {
        hipHostFree(r->h_fdesc[i]);
        hipFree(r->d_fdesc[i]);
        hipHostFree(r->h_fhdr[i]);
        hipFree(r->d_fhdr[i]);
    }
    hipStreamDestroy(r->stream);
    delete r;
    return 0;
}

GPUD_API int gpud_runner_set_exec_mode(gpud_runner* r, int mode)
{
    if (!r || (mode != GPUD_EXEC_BATCHED && mode != GPUD_EXEC_PER_SIGNAL)) return -1;
    r->exec_mode = mode;
    return 0;
}

// The fast-convolution form in the batched launch (on by default): a signal
// whose decimation D is a multiple of 8 with D < T <= 512 is filtered
// through the shared segment transforms once its history is full.
GPUD_API int gpud_runner_set_fast(gpud_runner* r, int enable)
{
    if (!r) return -1;
    r->fast = enable != 0;
    return 0;
}

GPUD_API int gpud_runner_set_window(gpud_runner* r, int window)
{
    if (!r || (window != 0 && window != 1)) return -1;
    r->window = window;
    return 0;
}

GPUD_API gpud_stream_handle gpud_runner_stream(gpud_runner* r) { return r ? (gpud_stream_handle)r->stream : nullptr; }

}  // extern "C"

// Drain recorded timing events into the running total.
static int argus_timing_drain(gpud_runner* r)
{
    for (int i = 0; i + 1 < r->tev_used; i += 2) {
        ARGUS_CHECK(hipEventSynchronize(r->tev[(size_t)i + 1]));
        float ms = 0.0f;
        ARGUS_CHECK(hipEventElapsedTime(&ms, r->tev[(size_t)i], r->tev[(size_t)i + 1]));
        r->t_ms += ms;
        r->t_launches++;
    }
    r->tev_used = 0;
    return 0;
}

// Record the start event of a timed launch; the caller records the stop
// event at tev[tev_used - 1] right after the launch.
int argus_timing_begin(gpud_runner* r)
{
    if ((size_t)r->tev_used + 2 > r->tev.size()) {
        if (r->tev.size() >= 4096) {
            int e = argus_timing_drain(r);
            if (e) return e;
        } else {
            for (int i = 0; i < 64; i++) {
                hipEvent_t ev;
                hipError_t e = hipEventCreate(&ev);
                if (e != hipSuccess) return (int)e;
                r->tev.push_back(ev);
            }
        }
    }
    hipError_t e = hipEventRecord(r->tev[(size_t)r->tev_used], r->stream);
    if (e != hipSuccess) return (int)e;
    r->tev_used += 2;
    return 0;
}

extern "C" GPUD_API int gpud_runner_enable_kernel_timing(gpud_runner* r, int enable)
{
    if (!r) return -1;
    ARGUS_CHECK(hipStreamSynchronize(r->stream));
    r->tev_used = 0;
    r->t_ms = 0.0;
    r->t_launches = 0;
    r->timing = enable != 0;
    return 0;
}

extern "C" GPUD_API int gpud_runner_kernel_time(gpud_runner* r, double* total_ms, int* launches)
{
    if (!r || !total_ms || !launches) return -1;
    int e = argus_timing_drain(r);
    if (e) return e;
    *total_ms = r->t_ms;
    *launches = r->t_launches;
    return 0;
}

// Build the channel table and work list for one chunk and launch the fused
// kernel.  Inputs must already be on the device (r->stream ordered).
static int run_batched(gpud_runner* r, const float2* d_in, int n_new, const gpud_streaming_job* jobs, int n_jobs,
                       std::vector<ArgusSignal*>& sig, gpud_streaming_result* results)
{
    int64_t total_out = 0;
    for (int i = 0; i < n_jobs; i++) total_out += results[i].n_out;
    ARGUS_CHECK((hipError_t)ensure_dev(&r->d_out, &r->d_out_cap, (size_t)std::max<int64_t>(total_out, 1)));

    const int slot = r->stage_slot;
    r->stage_slot ^= 1;
    if (r->stage_recorded[slot])  // staging slot no longer read by a pending copy
        ARGUS_CHECK(hipEventSynchronize(r->stage_ev[slot]));
    ARGUS_CHECK((hipError_t)ensure_host(&r->h_chans[slot], &r->h_chans_cap[slot], (size_t)std::max(n_jobs, 1)));
    ARGUS_CHECK((hipError_t)ensure_dev(&r->d_chans[slot], &r->d_chans_cap[slot], (size_t)std::max(n_jobs, 1)));

    r->items.clear();
    r->fdescs.clear();
    // the fast form needs the stream's own last ARGUS_FAST_TAIL samples in
    // the tail, which every batched call keeps
    const bool tail_ok = r->tail_fill >= ARGUS_FAST_TAIL;
    size_t lds = 0;
    for (int i = 0; i < n_jobs; i++) {
        ArgusSignal* s = sig[(size_t)i];
        const int T = s->num_taps, D = s->decim;
        const int keep = std::min(s->hist_len + n_new, T - 1);
        ArgusChanDesc& ch = r->h_chans[slot][i];
        const double inc = -2.0 * M_PI * jobs[i].offset_hz / (double)r->sample_rate;
        if (argus_fill_chan(&ch, s->d_hist[s->hist_sel], s->hist_len, s->d_hist[s->hist_sel ^ 1], keep, s->d_taps,
                            s->d_ctaps, r->d_out + results[i].out_offset, D, T, s->phase_count, n_new, s->nco_phase,
                            inc, r->tune) != 0)
            return -1;
        // once the signal's history is the stream's own last T-1 samples,
        // its outputs come from the fast form and its direct items are only
        // the history
        const bool fast = r->fast && tail_ok && s->fastG_valid && s->hist_len == T - 1 && n_new > 0 &&
                          n_new <= ARGUS_FAST_MAX_BLOCK;
        if (fast) {
            r->fdescs.push_back(argus_fast_desc(ch, s->d_fastG));
            ArgusChanDesc hist_only = ch;
            hist_only.n_out = 0;
            argus_append_items(hist_only, i, r->items);
        } else {
            if (ch.n_out > 0) lds = std::max(lds, argus_item_lds_bytes(ch));
            argus_append_items(ch, i, r->items);
        }
    }
    argus_order_items(r->items);
    const int n_items = (int)r->items.size();
    ARGUS_CHECK((hipError_t)ensure_host(&r->h_items[slot], &r->h_items_cap[slot], (size_t)std::max(n_items, 1)));
    ARGUS_CHECK((hipError_t)ensure_dev(&r->d_items[slot], &r->d_items_cap[slot], (size_t)std::max(n_items, 1)));
    if (n_items) std::memcpy(r->h_items[slot], r->items.data(), (size_t)n_items * sizeof(ArgusWorkItem));
    ARGUS_CHECK(hipMemcpyAsync(r->d_chans[slot], r->h_chans[slot], (size_t)n_jobs * sizeof(ArgusChanDesc),
                               hipMemcpyHostToDevice, r->stream));
    ARGUS_CHECK(hipMemcpyAsync(r->d_items[slot], r->h_items[slot], (size_t)n_items * sizeof(ArgusWorkItem),
                               hipMemcpyHostToDevice, r->stream));
    // the fast form's descriptors and headers (the tail is kept every call)
    const int n_fast = (int)r->fdescs.size();
    const int S = n_fast > 0 ? argus_fast_segments(n_new) : 0;
    if (!r->d_ftail) {
        ARGUS_CHECK(hipMalloc((void**)&r->d_ftail, (size_t)ARGUS_FAST_TAIL * sizeof(float2)));
        ARGUS_CHECK(hipMemset(r->d_ftail, 0, (size_t)ARGUS_FAST_TAIL * sizeof(float2)));
    }
    if (S > 0) {
        ARGUS_CHECK((hipError_t)ensure_dev(&r->d_fu, &r->fu_cap, argus_fast_u_len(S)));
        ARGUS_CHECK((hipError_t)ensure_dev(&r->d_fX, &r->fX_cap, (size_t)S * ARGUS_FAST_N));
    }
    ARGUS_CHECK((hipError_t)ensure_host(&r->h_fdesc[slot], &r->h_fdesc_cap[slot], (size_t)std::max(n_fast, 1)));
    ARGUS_CHECK((hipError_t)ensure_dev(&r->d_fdesc[slot], &r->d_fdesc_cap[slot], (size_t)std::max(n_fast, 1)));
    if (!r->h_fhdr[slot]) {
        ARGUS_CHECK(hipHostMalloc((void**)&r->h_fhdr[slot], sizeof(gpud_runner::FastHdrs), hipHostMallocDefault));
        ARGUS_CHECK(hipMalloc((void**)&r->d_fhdr[slot], sizeof(gpud_runner::FastHdrs)));
    }
    if (n_fast) std::memcpy(r->h_fdesc[slot], r->fdescs.data(), (size_t)n_fast * sizeof(ArgusFastDesc));
    gpud_runner::FastHdrs* fh = r->h_fhdr[slot];
    fh->f = ArgusFastHdr{r->d_fdesc[slot], r->d_ftail, r->d_fu, r->d_fX, n_fast, S};
    fh->x = ArgusExtractHdr{r->d_chans[slot], r->d_items[slot], d_in, n_items, n_new};
    ARGUS_CHECK(hipMemcpyAsync(r->d_fdesc[slot], r->h_fdesc[slot], (size_t)n_fast * sizeof(ArgusFastDesc),
                               hipMemcpyHostToDevice, r->stream));
    ARGUS_CHECK(hipMemcpyAsync(r->d_fhdr[slot], fh, sizeof(gpud_runner::FastHdrs), hipMemcpyHostToDevice,
                               r->stream));
    ARGUS_CHECK(hipEventRecord(r->stage_ev[slot], r->stream));
    r->stage_recorded[slot] = true;
    if (r->timing) ARGUS_CHECK((hipError_t)argus_timing_begin(r));
    const gpud_runner::FastHdrs* dfh = r->d_fhdr[slot];
    int err = S > 0 ? argus_launch_fast(&dfh->f, &dfh->x, S, r->d_fu, r->d_fX, r->tune.fast_grid, r->stream) : 0;
    if (!err)
        err = argus_launch_extract_batch(r->d_chans[slot], r->d_items[slot], n_items, lds, d_in, n_new, r->tune.grid,
                                         r->stream);
    if (err) return err;
    if (r->timing) ARGUS_CHECK(hipEventRecord(r->tev[(size_t)r->tev_used - 1], r->stream));
    if (n_new > 0) {
        err = argus_launch_fast_tail(&dfh->f, &dfh->x, r->stream);
        if (err) return err;
        r->tail_fill = std::min(ARGUS_FAST_TAIL, r->tail_fill + n_new);
    }
    // carry state forward (the kernel computes the same quantities)
    for (int i = 0; i < n_jobs; i++) {
        ArgusSignal* s = sig[(size_t)i];
        const double inc = -2.0 * M_PI * jobs[i].offset_hz / (double)r->sample_rate;
        const int keep = std::min(s->hist_len + n_new, s->num_taps - 1);
        s->phase_count = (s->phase_count + n_new) % s->decim;
        s->nco_phase = argus_reduce_phase(s->nco_phase + inc * (double)n_new);
        if (keep > 0) s->hist_sel ^= 1;
        s->hist_len = keep;
        results[i].phase_count = s->phase_count;
        results[i].history_len = s->hist_len;
        results[i].nco_phase = s->nco_phase;
    }
    return 0;
}

static int prepare_jobs(gpud_runner* r, int n_new, const gpud_streaming_job* jobs, int n_jobs,
                        std::vector<ArgusSignal*>& sig, gpud_streaming_result* results)
{
    if (n_new < 0 || n_jobs < 0 || (n_jobs > 0 && (!jobs || !results))) return -1;
    if (n_new > r->max_samples) return -2;
    for (auto& s : r->signals) s.active = false;
    // first pass creates every state (vector may reallocate), second resolves pointers
    for (int i = 0; i < n_jobs; i++) {
        ArgusSignal* s = nullptr;
        int err = argus_runner_prepare_signal(r, jobs[i], &s);
        if (err) return err;
    }
    argus_runner_prune(r);
    if (r->rings) {  // syncNativeStreamingStates prunes the output rings too (streaming_gpu_native_prepare.go:211-223)
        std::vector<int64_t> ids((size_t)n_jobs);
        for (int i = 0; i < n_jobs; i++) ids[(size_t)i] = jobs[i].signal_id;
        argus_out_rings_prune(r->rings, ids.data(), n_jobs);
    }
    sig.assign((size_t)n_jobs, nullptr);
    int64_t off = 0;
    for (int i = 0; i < n_jobs; i++) {
        ArgusSignal* s = argus_runner_find(r, jobs[i].signal_id);
        sig[(size_t)i] = s;
        results[i].signal_id = jobs[i].signal_id;
        results[i].rate = jobs[i].out_rate;
        results[i].n_out = (s->phase_count + n_new) / s->decim;
        results[i].out_offset = off;
        off += results[i].n_out;
    }
    return 0;
}

// The plain per-signal form: for each signal, upload the whole chunk, run the
// single-signal stateful entry point, and copy outputs and history back
// synchronously -- the reference's call pattern
// (streaming_gpu_native_prepare.go:17-108), with its kernels replaced.
static int run_per_signal(gpud_runner* r, const gpud_float2* iq_host, int n_new, const gpud_streaming_job* jobs,
                          int n_jobs, std::vector<ArgusSignal*>& sig, gpud_streaming_result* results,
                          gpud_float2* out_host)
{
    for (int i = 0; i < n_jobs; i++) {
        ArgusSignal* s = sig[(size_t)i];
        if (n_new > 0)
            ARGUS_CHECK(hipMemcpy(r->d_in, iq_host, (size_t)n_new * sizeof(float2), hipMemcpyHostToDevice));
        const int out_cap = n_new / std::max(1, s->decim) + 2;
        if (s->out_cap < out_cap) {
            if (s->d_out) hipFree(s->d_out);
            ARGUS_CHECK(hipMalloc((void**)&s->d_out, (size_t)out_cap * sizeof(float2)));
            s->out_cap = out_cap;
        }
        const double inc = -2.0 * M_PI * jobs[i].offset_hz / (double)r->sample_rate;
        int hlen = s->hist_len, pc = s->phase_count, n_out = 0;
        double ph = s->nco_phase;
        int err = gpud_launch_streaming_polyphase_stateful_hip(
            (const gpud_float2*)r->d_in, n_new, (gpud_float2*)r->d_in, s->d_taps, (int)s->poly.size(), s->decim,
            s->num_taps, (gpud_float2*)s->d_hist[s->hist_sel], (gpud_float2*)s->d_hist[s->hist_sel ^ 1],
            s->num_taps - 1, &hlen, &pc, &ph, inc, (gpud_float2*)s->d_out, out_cap, &n_out);
        if (err) return err;
        if (n_out > 0)
            ARGUS_CHECK(hipMemcpy(out_host + results[i].out_offset, s->d_out, (size_t)n_out * sizeof(float2),
                                  hipMemcpyDeviceToHost));
        if (hlen > 0) {
            r->hist_host.resize((size_t)hlen);
            ARGUS_CHECK(hipMemcpy(r->hist_host.data(), s->d_hist[s->hist_sel], (size_t)hlen * sizeof(float2),
                                  hipMemcpyDeviceToHost));
        }
        s->hist_len = hlen;
        s->phase_count = pc;
        s->nco_phase = ph;
        results[i].n_out = n_out;
        results[i].phase_count = pc;
        results[i].history_len = hlen;
        results[i].nco_phase = ph;
    }
    return 0;
}

extern "C" {

GPUD_API int gpud_runner_streaming_extract(gpud_runner* r, const gpud_float2* iq, int n_new,
                                           const gpud_streaming_job* jobs, int n_jobs, gpud_float2* out,
                                           int64_t out_cap, gpud_streaming_result* results)
{
    if (!r || (n_new > 0 && !iq)) return -1;
    std::vector<ArgusSignal*> sig;
    int err = prepare_jobs(r, n_new, jobs, n_jobs, sig, results);
    if (err) return err;
    int64_t total = 0;
    for (int i = 0; i < n_jobs; i++) total += results[i].n_out;
    if (total > out_cap || (total > 0 && !out)) return -14;
    if (r->exec_mode == GPUD_EXEC_PER_SIGNAL) {
        if (n_new > 0) r->tail_fill = 0;  // the per-signal form does not keep the fast form's tail
        return run_per_signal(r, iq, n_new, jobs, n_jobs, sig, results, out);
    }
    if (n_new > 0) {
        ARGUS_CHECK(hipStreamSynchronize(r->stream));  // h_in reuse
        std::memcpy(r->h_in, iq, (size_t)n_new * sizeof(float2));
        ARGUS_CHECK(hipMemcpyAsync(r->d_in, r->h_in, (size_t)n_new * sizeof(float2), hipMemcpyHostToDevice,
                                   r->stream));
    }
    err = run_batched(r, r->d_in, n_new, jobs, n_jobs, sig, results);
    if (err) return err;
    if (total > 0) {
        ARGUS_CHECK((hipError_t)ensure_host(&r->h_out, &r->h_out_cap, (size_t)total));
        ARGUS_CHECK(hipMemcpyAsync(r->h_out, r->d_out, (size_t)total * sizeof(float2), hipMemcpyDeviceToHost,
                                   r->stream));
    }
    ARGUS_CHECK(hipStreamSynchronize(r->stream));
    if (total > 0) std::memcpy(out, r->h_out, (size_t)total * sizeof(float2));
    return 0;
}

GPUD_API int gpud_runner_streaming_extract_ring(gpud_runner* r, const gpud_float2* iq, int n_new,
                                                const gpud_streaming_job* jobs, int n_jobs, gpud_float2** outs,
                                                gpud_streaming_result* results)
{
    if (!r || (n_new > 0 && !iq) || (n_jobs > 0 && !outs)) return -1;
    if (!r->rings) r->rings = new gpud_out_rings();
    std::vector<ArgusSignal*> sig;
    int err = prepare_jobs(r, n_new, jobs, n_jobs, sig, results);
    if (err) return err;
    int64_t total = 0;
    for (int i = 0; i < n_jobs; i++) total += results[i].n_out;
    ARGUS_CHECK((hipError_t)ensure_host(&r->h_out, &r->h_out_cap, (size_t)std::max<int64_t>(total, 1)));
    if (r->exec_mode == GPUD_EXEC_PER_SIGNAL) {
        if (n_new > 0) r->tail_fill = 0;
        err = run_per_signal(r, iq, n_new, jobs, n_jobs, sig, results, (gpud_float2*)r->h_out);
        if (err) return err;
    } else {
        if (n_new > 0) {
            ARGUS_CHECK(hipStreamSynchronize(r->stream));  // h_in reuse
            std::memcpy(r->h_in, iq, (size_t)n_new * sizeof(float2));
            ARGUS_CHECK(hipMemcpyAsync(r->d_in, r->h_in, (size_t)n_new * sizeof(float2), hipMemcpyHostToDevice,
                                       r->stream));
        }
        err = run_batched(r, r->d_in, n_new, jobs, n_jobs, sig, results);
        if (err) return err;
        if (total > 0)
            ARGUS_CHECK(hipMemcpyAsync(r->h_out, r->d_out, (size_t)total * sizeof(float2), hipMemcpyDeviceToHost,
                                       r->stream));
        ARGUS_CHECK(hipStreamSynchronize(r->stream));
    }
    // each signal's outputs into its ring's next buffer (streaming_gpu_native_prepare.go:82-88)
    for (int i = 0; i < n_jobs; i++) {
        gpud_float2* b = nullptr;
        if (argus_out_ring_next(r->rings, jobs[i].signal_id, results[i].n_out, &b) != 0) return -3;
        if (results[i].n_out > 0)
            std::memcpy(b, (const gpud_float2*)r->h_out + results[i].out_offset,
                        (size_t)results[i].n_out * sizeof(gpud_float2));
        outs[i] = results[i].n_out > 0 ? b : nullptr;
    }
    return 0;
}

GPUD_API gpud_out_rings* gpud_runner_out_rings(gpud_runner* r) { return r ? r->rings : nullptr; }

GPUD_API int gpud_runner_streaming_extract_device(gpud_runner* r, const gpud_float2* iq_dev, int n_new,
                                                  const gpud_streaming_job* jobs, int n_jobs,
                                                  gpud_float2** out_dev, gpud_streaming_result* results)
{
    if (!r || (n_new > 0 && !iq_dev)) return -1;
    if (r->exec_mode != GPUD_EXEC_BATCHED) return -1;
    std::vector<ArgusSignal*> sig;
    int err = prepare_jobs(r, n_new, jobs, n_jobs, sig, results);
    if (err) return err;
    err = run_batched(r, (const float2*)iq_dev, n_new, jobs, n_jobs, sig, results);
    if (err) return err;
    if (out_dev) *out_dev = (gpud_float2*)r->d_out;
    return 0;
}

GPUD_API int gpud_runner_get_signal_state(gpud_runner* r, int64_t signal_id, uint64_t* config_hash,
                                          double* nco_phase, int* decim, int* phase_count, int* num_taps,
                                          gpud_float2* history, int hist_cap)
{
    if (!r) return -1;
    ArgusSignal* s = argus_runner_find(r, signal_id);
    if (!s) return -1;
    if (config_hash) *config_hash = s->config_hash;
    if (nco_phase) *nco_phase = s->nco_phase;
    if (decim) *decim = s->decim;
    if (phase_count) *phase_count = s->phase_count;
    if (num_taps) *num_taps = s->num_taps;
    if (history && hist_cap > 0 && s->hist_len > 0) {
        ARGUS_CHECK(hipStreamSynchronize(r->stream));
        const int m = std::min(hist_cap, s->hist_len);
        ARGUS_CHECK(hipMemcpy(history, s->d_hist[s->hist_sel], (size_t)m * sizeof(float2), hipMemcpyDeviceToHost));
    }
    return s->hist_len;
}

GPUD_API int gpud_runner_num_states(gpud_runner* r) { return r ? (int)r->signals.size() : -1; }

GPUD_API int gpud_runner_state_ids(gpud_runner* r, int64_t* ids, int cap)
{
    if (!r) return -1;
    int i = 0;
    for (auto& s : r->signals) {
        if (i < cap && ids) ids[i] = s.signal_id;
        i++;
    }
    return i;
}

GPUD_API int gpud_runner_reset_signal_state(gpud_runner* r, int64_t signal_id)
{
    if (!r) return -1;
    hipStreamSynchronize(r->stream);
    for (size_t i = 0; i < r->signals.size(); i++)
        if (r->signals[i].signal_id == signal_id) {
            argus_signal_release(r->signals[i]);
            r->signals.erase(r->signals.begin() + (long)i);
            break;
        }
    return 0;
}

GPUD_API int gpud_runner_reset_all_signal_states(gpud_runner* r)
{
    if (!r) return -1;
    hipStreamSynchronize(r->stream);
    for (auto& s : r->signals) argus_signal_release(s);
    r->signals.clear();
    return 0;
}

}  // extern "C"
