// engine.hip -- job building of the reference's DSP loop and its legacy
// block path, on the device.
//
//  * Rate snapping and the job builders of cmd/sdrd: snapOutRate
//    (helpers.go:292-310), nearestExactDecimationRate
//    (streaming_refactor.go:105-137), buildStreamingJobs
//    (streaming_refactor.go:16-60) and the legacy ExtractJob builder of
//    extractSignalIQBatch (helpers.go:162-195).  Pure host integer/FP64 work.
//  * The legacy block path: Engine.ShiftFilterDecimate / DemodFused / Demod
//    (gpudemod_windows.go:320-654) and BatchRunner.ShiftFilterDecimateBatch
//    [WithPhase] (batch_runner.go:61-98, batch_runner_windows.go:87-182).
//    The reference shifts, filters every input sample and keeps indices 0,
//    D, 2D, ... of the filtered stream, in three kernels with full-rate
//    intermediates.  Here every job of a call is a transient channel of the
//    fused extraction kernel -- decimator phase D-1 (so input 0 is the first
//    emission point), empty history, the job's phase start -- so only the
//    kept outputs are computed, in one launch for all jobs, and the
//    demodulators run on the decimated output on the same stream.
#include "extract.hpp"
#include "runner.hpp"
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

int argus_launch_fm_discrim(const float2* in, float* out, int n, hipStream_t st);
int argus_launch_am_envelope(const float2* in, float* out, int n, hipStream_t st);
int argus_launch_ssb_product(const float2* in, float* out, int n, double phase_inc, double phase_start,
                             hipStream_t st);

// ---------------------------------------------------------------------------
// Rate snapping and job builders
// ---------------------------------------------------------------------------
namespace {

constexpr int kWfmStreamOutRate = 512000;   // helpers.go:279
constexpr double kWfmStreamMinBW = 250000;  // helpers.go:280
constexpr int kLegacyDecimTarget = 200000;  // helpers.go:168

// FM broadcast band or a WFM classification (streaming_refactor.go:29-31).
bool is_wfm(const gpud_detected_signal& s)
{
    const double mhz = s.center_hz / 1e6;
    return (mhz >= 87.5 && mhz <= 108.0) || s.mod_type == GPUD_MOD_WFM || s.mod_type == GPUD_MOD_WFM_STEREO;
}

}  // namespace

extern "C" {

GPUD_API int gpud_snap_out_rate(int sample_rate, int target)
{
    if (sample_rate <= 0 || target <= 0) return target;
    const double ideal = (double)sample_rate / (double)target;
    int best = 0;
    double best_err = DBL_MAX;
    for (int d = 1; d <= 256 && d <= sample_rate; d++) {
        if (sample_rate % d != 0) continue;
        const double e = std::fabs((double)d - ideal);
        if (e < best_err) {
            best_err = e;
            best = d;
        }
    }
    return best < 1 ? target : sample_rate / best;
}

GPUD_API int gpud_nearest_exact_decimation_rate(int sample_rate, int target_rate, int min_rate)
{
    if (sample_rate <= 0 || target_rate <= 0) return target_rate;
    if (sample_rate % target_rate == 0) return target_rate;
    const int target_decim = sample_rate / target_rate;
    int best_rate = 0;
    int best_dist = sample_rate;
    for (int d = std::max(1, target_decim - 2); d <= target_decim + 2; d++) {
        const int rate = sample_rate / d;
        if (rate < min_rate || rate <= 0) continue;
        if (sample_rate % rate != 0) continue;
        const int dist = std::abs(target_rate - rate);
        if (dist < best_dist) {
            best_dist = dist;
            best_rate = rate;
        }
    }
    return best_rate > 0 ? best_rate : target_rate;
}

GPUD_API int gpud_build_streaming_jobs(int sample_rate, double center_hz, const gpud_detected_signal* signals,
                                       int n_signals, const gpud_extraction_config* cfg, gpud_streaming_job* jobs,
                                       int* failed_index)
{
    if (sample_rate <= 0 || n_signals < 0 || (n_signals > 0 && (!signals || !jobs))) return -1;
    double bw_mult = cfg ? cfg->bw_mult : 0.0;
    if (!(bw_mult > 0)) bw_mult = 1.0;
    int fir_taps = cfg ? cfg->fir_taps : 0;
    if (fir_taps <= 0) fir_taps = 101;
    const bool snap = cfg && cfg->snap_wfm != 0;
    for (int i = 0; i < n_signals; i++) {
        const gpud_detected_signal& sig = signals[i];
        double bw = sig.bw_hz * bw_mult;
        int out_rate;
        if (is_wfm(sig)) {
            out_rate = snap ? gpud_snap_out_rate(sample_rate, kWfmStreamOutRate) : kWfmStreamOutRate;
            if (bw < kWfmStreamMinBW) bw = kWfmStreamMinBW;
        } else {
            // the nearest exact rate >= 128 kHz to 200 kHz (streaming_refactor.go:39-44)
            out_rate = gpud_nearest_exact_decimation_rate(sample_rate, 200000, 128000);
            if (bw < 20000) bw = 20000;
        }
        if (gpud_exact_integer_decimation(sample_rate, out_rate) < 0) {
            if (failed_index) *failed_index = i;
            return -2;
        }
        const double offset = sig.center_hz - center_hz;
        gpud_streaming_job& j = jobs[i];
        j.signal_id = sig.id;
        j.offset_hz = offset;
        j.bandwidth = bw;
        j.out_rate = out_rate;
        j.num_taps = fir_taps;
        j.config_hash = gpud_streaming_config_hash(sig.id, offset, bw, out_rate, fir_taps, sample_rate);
    }
    return 0;
}

GPUD_API int gpud_build_extract_jobs(int sample_rate, double center_hz, const gpud_detected_signal* signals,
                                     int n_signals, gpud_extract_job* jobs)
{
    if (sample_rate <= 0 || n_signals < 0 || (n_signals > 0 && (!signals || !jobs))) return -1;
    for (int i = 0; i < n_signals; i++) {
        const gpud_detected_signal& sig = signals[i];
        double bw = sig.bw_hz;
        int job_rate = kLegacyDecimTarget;
        if (is_wfm(sig)) {
            job_rate = kWfmStreamOutRate;
            if (bw < kWfmStreamMinBW) bw = kWfmStreamMinBW;
        } else if (bw < 20000) {
            bw = 20000;
        }
        jobs[i] = gpud_extract_job{sig.center_hz - center_hz, bw, gpud_snap_out_rate(sample_rate, job_rate), 0, 0.0};
    }
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Legacy block path on the fused extraction kernel
// ---------------------------------------------------------------------------
namespace {

template <typename T>
int grow_dev(T** p, size_t* cap, size_t need)
{
    if (*p && *cap >= need) return 0;
    if (*p) hipFree(*p);
    *p = nullptr;
    *cap = 0;
    const size_t n = need < 256 ? 256 : need + need / 2;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return (int)e;
    *cap = n;
    return 0;
}

template <typename T>
int grow_host(T** p, size_t* cap, size_t need)
{
    if (*p && *cap >= need) return 0;
    if (*p) hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    const size_t n = need < 256 ? 256 : need + need / 2;
    hipError_t e = hipHostMalloc((void**)p, n * sizeof(T), hipHostMallocDefault);
    if (e != hipSuccess) return (int)e;
    *cap = n;
    return 0;
}

size_t align16(size_t v) { return (v + 15) & ~(size_t)15; }

}  // namespace

using LegacyJob = ArgusLegacyJob;

// Device and pinned staging of legacy calls: per job the phase-major and
// modulated taps, then the channel descriptors (one region), and the work
// items (a second region, so growing it leaves the first one's pointers valid).
struct ArgusLegacyScratch {
    char* d_data = nullptr;
    char* h_data = nullptr;
    size_t d_cap = 0, h_cap = 0;
    ArgusWorkItem* d_items = nullptr;
    ArgusWorkItem* h_items = nullptr;
    size_t d_items_cap = 0, h_items_cap = 0;
    std::vector<ArgusWorkItem> items;
    std::vector<float> poly;
    std::vector<float2> ctaps;
};

void argus_legacy_release(ArgusLegacyScratch* sc)
{
    if (!sc) return;
    hipFree(sc->d_data);
    hipHostFree(sc->h_data);
    hipFree(sc->d_items);
    hipHostFree(sc->h_items);
    delete sc;
}

// Queue the transient-channel launch of a legacy call on `st` (inputs on the
// device, outputs to each job's device pointer).  The staging is reused, so
// the caller synchronises the stream before the next call.
int argus_legacy_launch(ArgusLegacyScratch*& sc, hipStream_t st, const ArgusTune& tune, const float2* d_in, int n,
                        const std::vector<LegacyJob>& jobs)
{
    if (!sc) sc = new ArgusLegacyScratch();
    const int nj = (int)jobs.size();
    // per-job taps first, then the descriptors
    std::vector<size_t> off_poly((size_t)nj), off_ct((size_t)nj);
    size_t bytes = 0;
    for (int j = 0; j < nj; j++) {
        const int L = (jobs[j].T + jobs[j].D - 1) / jobs[j].D;
        off_poly[(size_t)j] = bytes;
        bytes = align16(bytes + (size_t)jobs[j].D * L * sizeof(float));
        off_ct[(size_t)j] = bytes;
        bytes = align16(bytes + (size_t)jobs[j].T * sizeof(float2));
    }
    const size_t off_desc = bytes;
    bytes += (size_t)std::max(nj, 1) * sizeof(ArgusChanDesc);
    ARGUS_CHECK((hipError_t)grow_dev(&sc->d_data, &sc->d_cap, bytes));
    ARGUS_CHECK((hipError_t)grow_host(&sc->h_data, &sc->h_cap, bytes));
    ArgusChanDesc* descs = (ArgusChanDesc*)(sc->h_data + off_desc);
    sc->items.clear();
    size_t lds = 0;
    for (int j = 0; j < nj; j++) {
        const LegacyJob& jb = jobs[j];
        const int L = (jb.T + jb.D - 1) / jb.D;
        float* poly = (float*)(sc->h_data + off_poly[(size_t)j]);
        for (int p = 0; p < jb.D; p++)
            for (int k = 0; k < L; k++) {
                const int src = p + k * jb.D;
                poly[(size_t)p * L + k] = src < jb.T ? jb.taps[src] : 0.0f;
            }
        float2* ct = (float2*)(sc->h_data + off_ct[(size_t)j]);
        argus_modulated_taps(poly, jb.D, jb.T, argus_rad_to_fix(jb.phase_inc), ct);
        ArgusChanDesc& ch = descs[j];
        // decimator phase D-1: input 0 is an emission point; no history is
        // read (H = 0) or written (keep = 0)
        if (argus_fill_chan(&ch, nullptr, 0, nullptr, 0, (const float*)(sc->d_data + off_poly[(size_t)j]),
                            (const float2*)(sc->d_data + off_ct[(size_t)j]), jb.out, jb.D, jb.T, jb.D - 1, n,
                            jb.phase_start, jb.phase_inc, tune) != 0)
            return -1;
        ch.n_out = jb.n_out;
        if (ch.n_out > 0) lds = std::max(lds, argus_item_lds_bytes(ch));
        argus_append_items(ch, j, sc->items);
    }
    argus_order_items(sc->items);
    const int n_items = (int)sc->items.size();
    ARGUS_CHECK((hipError_t)grow_dev(&sc->d_items, &sc->d_items_cap, (size_t)std::max(n_items, 1)));
    ARGUS_CHECK((hipError_t)grow_host(&sc->h_items, &sc->h_items_cap, (size_t)std::max(n_items, 1)));
    if (n_items) std::memcpy(sc->h_items, sc->items.data(), (size_t)n_items * sizeof(ArgusWorkItem));
    ARGUS_CHECK(hipMemcpyAsync(sc->d_data, sc->h_data, bytes, hipMemcpyHostToDevice, st));
    ARGUS_CHECK(hipMemcpyAsync(sc->d_items, sc->h_items, (size_t)n_items * sizeof(ArgusWorkItem),
                               hipMemcpyHostToDevice, st));
    return argus_launch_extract_batch((const ArgusChanDesc*)(sc->d_data + off_desc), sc->d_items, n_items, lds, d_in,
                                      n, tune.grid, st);
}

// ---------------------------------------------------------------------------
// BatchRunner.ShiftFilterDecimateBatch[WithPhase]
// ---------------------------------------------------------------------------
extern "C" GPUD_API int gpud_runner_shift_filter_decimate_batch(gpud_runner* r, const gpud_float2* iq, int n,
                                                                const gpud_extract_job* jobs, int n_jobs,
                                                                gpud_float2* out, int64_t out_cap,
                                                                gpud_extract_result* results)
{
    if (!r || n < 0 || n_jobs < 0 || (n > 0 && !iq) || (n_jobs > 0 && (!jobs || !results))) return -1;
    if (n > r->max_samples) return -2;
    if (n == 0) {  // batch_runner_windows.go:93-96: empty outputs, zero rates
        for (int i = 0; i < n_jobs; i++)
            results[i] = gpud_extract_result{0, 0, 0, jobs[i].phase_start};
        return 0;
    }
    // per job (shiftFilterDecimateSlotParallel, batch_runner_windows.go:137-182):
    // 101-tap LowpassFIR at max(bw/2, 200) Hz, exact decimation, n/D outputs
    std::vector<std::vector<float>> taps((size_t)n_jobs);
    std::vector<LegacyJob> lj((size_t)n_jobs);
    int64_t total = 0;
    for (int i = 0; i < n_jobs; i++) {
        const gpud_extract_job& jb = jobs[i];
        double cutoff = jb.bw / 2;
        if (cutoff < 200) cutoff = 200;
        std::vector<double> b64(102);
        const int T = argus_lowpass_fir(cutoff, r->sample_rate, 101, b64.data(), false);
        taps[(size_t)i].assign(b64.begin(), b64.begin() + T);
        const int D = gpud_exact_integer_decimation(r->sample_rate, jb.out_rate);
        if (D < 0) return -3;                // ExactIntegerDecimation error
        if (n / D <= 0) return -4;           // not enough output samples after decimation
        const double inc = -2.0 * M_PI * jb.offset_hz / (double)r->sample_rate;
        lj[(size_t)i] = LegacyJob{taps[(size_t)i].data(), T, D, jb.phase_start, inc, n / D, nullptr};
        results[i].rate = r->sample_rate / D;
        results[i].n_out = n / D;
        results[i].out_offset = total;
        // ShiftFilterDecimateBatchWithPhase's PhaseEnd (batch_runner.go:87-90)
        results[i].phase_end = jb.phase_start + inc * (double)n;
        total += n / D;
    }
    if (total > out_cap || (total > 0 && !out)) return -14;
    ARGUS_CHECK(hipStreamSynchronize(r->stream));  // h_in and the staging are reused
    std::memcpy(r->h_in, iq, (size_t)n * sizeof(float2));
    ARGUS_CHECK(hipMemcpyAsync(r->d_in, r->h_in, (size_t)n * sizeof(float2), hipMemcpyHostToDevice, r->stream));
    // outputs packed into the runner's output buffer; with a ceil(n/D)
    // allowance per job, which the kernel's tiles never exceed
    int64_t cap = 0;
    for (int i = 0; i < n_jobs; i++) cap += n / lj[(size_t)i].D + 1;
    if (!r->d_out || r->d_out_cap < (size_t)cap) {
        if (r->d_out) hipFree(r->d_out);
        r->d_out = nullptr;
        r->d_out_cap = 0;
        ARGUS_CHECK(hipMalloc((void**)&r->d_out, (size_t)cap * sizeof(float2)));
        r->d_out_cap = (size_t)cap;
    }
    for (int i = 0; i < n_jobs; i++) lj[(size_t)i].out = r->d_out + results[i].out_offset;
    int err = argus_legacy_launch(r->legacy, r->stream, r->tune, r->d_in, n, lj);
    if (err) return err;
    if (total > 0) {
        if (!r->h_out || r->h_out_cap < (size_t)total) {
            if (r->h_out) hipHostFree(r->h_out);
            r->h_out = nullptr;
            r->h_out_cap = 0;
            ARGUS_CHECK(hipHostMalloc((void**)&r->h_out, (size_t)total * sizeof(float2), hipHostMallocDefault));
            r->h_out_cap = (size_t)total;
        }
        ARGUS_CHECK(hipMemcpyAsync(r->h_out, r->d_out, (size_t)total * sizeof(float2), hipMemcpyDeviceToHost,
                                   r->stream));
    }
    ARGUS_CHECK(hipStreamSynchronize(r->stream));
    if (total > 0) std::memcpy(out, r->h_out, (size_t)total * sizeof(float2));
    return 0;
}

// ---------------------------------------------------------------------------
// Engine (gpudemod_windows.go): carried shift phase and BFO phase, a FIR tap
// set that sticks once designed or set, the shift-filter-decimate chain and
// the demodulators on the decimated output.
// ---------------------------------------------------------------------------
struct gpud_engine {
    int max_samples = 0;
    int sample_rate = 0;
    double phase = 0.0;      // FreqShift phase, carried unreduced as the reference does
    double bfo_phase = 0.0;
    std::vector<float> fir;  // <= 256 taps (SetFIR)
    ArgusTune tune = argus_default_tune();
    hipStream_t stream = nullptr;
    float2* d_in = nullptr;
    float2* h_in = nullptr;
    float2* d_dec = nullptr;
    float* d_audio = nullptr;
    float2* h_dec = nullptr;
    float* h_audio = nullptr;
    ArgusLegacyScratch* legacy = nullptr;
};

namespace {

// The engine's taps: the set ones, else LowpassFIR(max(bw/2, 200), fs, 101)
// designed now and kept (gpudemod_windows.go:342-351, 431-440).
void engine_taps(gpud_engine* e, double bw)
{
    if (!e->fir.empty()) return;
    double cutoff = bw / 2;
    if (cutoff < 200) cutoff = 200;
    std::vector<double> b64(102);
    const int T = argus_lowpass_fir(cutoff, e->sample_rate, 101, b64.data(), false);
    e->fir.assign(b64.begin(), b64.begin() + T);
}

// decim = round(fs / out_rate), at least 1 (gpudemod_windows.go:355-358)
int engine_decim(const gpud_engine* e, int out_rate)
{
    const int d = (int)std::lround((double)e->sample_rate / (double)out_rate);
    return d < 1 ? 1 : d;
}

// Upload, then the shift-filter-decimate chain for n/D outputs into d_dec.
int engine_chain(gpud_engine* e, const gpud_float2* iq, int n, double inc, int D)
{
    ARGUS_CHECK(hipStreamSynchronize(e->stream));
    std::memcpy(e->h_in, iq, (size_t)n * sizeof(float2));
    ARGUS_CHECK(hipMemcpyAsync(e->d_in, e->h_in, (size_t)n * sizeof(float2), hipMemcpyHostToDevice, e->stream));
    std::vector<LegacyJob> jobs{LegacyJob{e->fir.data(), (int)e->fir.size(), D, e->phase, inc, n / D, e->d_dec}};
    return argus_legacy_launch(e->legacy, e->stream, e->tune, e->d_in, n, jobs);
}

}  // namespace

extern "C" {

GPUD_API int gpud_engine_create(int max_samples, int sample_rate, gpud_engine** out)
{
    if (!out || max_samples <= 0 || sample_rate <= 0) return -1;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return (int)hipErrorNoDevice;
    gpud_engine* e = new gpud_engine();
    e->max_samples = max_samples;
    e->sample_rate = sample_rate;
    const size_t n = (size_t)max_samples;
    hipError_t err = hipStreamCreate(&e->stream);
    if (err == hipSuccess) err = hipMalloc((void**)&e->d_in, n * sizeof(float2));
    if (err == hipSuccess) err = hipHostMalloc((void**)&e->h_in, n * sizeof(float2), hipHostMallocDefault);
    if (err == hipSuccess) err = hipMalloc((void**)&e->d_dec, (n + 1) * sizeof(float2));
    if (err == hipSuccess) err = hipMalloc((void**)&e->d_audio, (n + 1) * sizeof(float));
    if (err == hipSuccess) err = hipHostMalloc((void**)&e->h_dec, (n + 1) * sizeof(float2), hipHostMallocDefault);
    if (err == hipSuccess) err = hipHostMalloc((void**)&e->h_audio, (n + 1) * sizeof(float), hipHostMallocDefault);
    if (err != hipSuccess) {
        gpud_engine_destroy(e);
        return (int)err;
    }
    *out = e;
    return 0;
}

GPUD_API int gpud_engine_destroy(gpud_engine* e)
{
    if (!e) return 0;
    if (e->stream) hipStreamSynchronize(e->stream);
    hipFree(e->d_in);
    hipHostFree(e->h_in);
    hipFree(e->d_dec);
    hipFree(e->d_audio);
    hipHostFree(e->h_dec);
    hipHostFree(e->h_audio);
    argus_legacy_release(e->legacy);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
    return 0;
}

GPUD_API int gpud_engine_set_fir(gpud_engine* e, const float* taps, int n)
{
    if (!e || (n > 0 && !taps)) return -1;
    if (n <= 0) {
        e->fir.clear();
        return 0;
    }
    e->fir.assign(taps, taps + std::min(n, 256));
    return 0;
}

GPUD_API int gpud_engine_state(gpud_engine* e, double* phase, double* bfo_phase, int* fir_taps)
{
    if (!e) return -1;
    if (phase) *phase = e->phase;
    if (bfo_phase) *bfo_phase = e->bfo_phase;
    if (fir_taps) *fir_taps = (int)e->fir.size();
    return 0;
}

GPUD_API int gpud_engine_shift_filter_decimate(gpud_engine* e, const gpud_float2* iq, int n, double offset_hz,
                                               double bw, int out_rate, gpud_float2* out, int out_cap, int* n_out,
                                               int* rate)
{
    if (!e || n < 0 || (n > 0 && !iq)) return -1;
    if (n_out) *n_out = 0;
    if (rate) *rate = 0;
    if (n == 0) return 0;
    if (n > e->max_samples) return -2;
    if (out_rate <= 0) return -3;
    engine_taps(e, bw);
    const int D = engine_decim(e, out_rate);
    const int m = n / D;
    if (m <= 0) return -4;
    if (m > out_cap || !out) return -14;
    const double inc = -2.0 * M_PI * offset_hz / (double)e->sample_rate;
    int err = engine_chain(e, iq, n, inc, D);
    if (err) return err;
    ARGUS_CHECK(hipMemcpyAsync(e->h_dec, e->d_dec, (size_t)m * sizeof(float2), hipMemcpyDeviceToHost, e->stream));
    ARGUS_CHECK(hipStreamSynchronize(e->stream));
    std::memcpy(out, e->h_dec, (size_t)m * sizeof(float2));
    e->phase += inc * (double)n;
    if (n_out) *n_out = m;
    if (rate) *rate = e->sample_rate / D;
    return 0;
}

GPUD_API int gpud_engine_demod(gpud_engine* e, const gpud_float2* iq, int n, double offset_hz, double bw, int mode,
                               int fused, float* out, int out_cap, int* n_out, int* rate)
{
    if (!e || n < 0 || (n > 0 && !iq)) return -1;
    if (n_out) *n_out = 0;
    if (rate) *rate = 0;
    if (n == 0) return 0;
    if (n > e->max_samples) return -2;
    int out_rate;
    switch (mode) {
    case GPUD_ENGINE_NFM:
    case GPUD_ENGINE_AM:
    case GPUD_ENGINE_USB:
    case GPUD_ENGINE_LSB:
    case GPUD_ENGINE_CW: out_rate = 48000; break;
    case GPUD_ENGINE_WFM: out_rate = 192000; break;
    default: return -5;  // unsupported demod type
    }
    engine_taps(e, bw);
    const int D = engine_decim(e, out_rate);
    const int m = n / D;
    // DemodFused needs two outputs (gpudemod_windows.go:443-446); Demod's GPU
    // decimation needs one (tryCUDADecimate)
    if (m <= (fused ? 1 : 0)) return -4;
    const bool fm = mode == GPUD_ENGINE_NFM || mode == GPUD_ENGINE_WFM;
    const int na = fm ? m - 1 : m;
    if (na > out_cap || (na > 0 && !out)) return -14;
    const double inc = -2.0 * M_PI * offset_hz / (double)e->sample_rate;
    int err = engine_chain(e, iq, n, inc, D);
    if (err) return err;
    double bfo_inc = 0.0;
    if (fm) {
        err = argus_launch_fm_discrim(e->d_dec, e->d_audio, m, e->stream);
    } else if (mode == GPUD_ENGINE_AM) {
        err = argus_launch_am_envelope(e->d_dec, e->d_audio, m, e->stream);
    } else {
        // BFO increment over the input rate, as the reference computes it
        // (gpudemod_windows.go:304, 508)
        const double bfo_hz = mode == GPUD_ENGINE_LSB ? -700.0 : 700.0;
        bfo_inc = 2.0 * M_PI * bfo_hz / (double)e->sample_rate;
        err = argus_launch_ssb_product(e->d_dec, e->d_audio, m, bfo_inc, e->bfo_phase, e->stream);
    }
    if (err) return err;
    if (na > 0)
        ARGUS_CHECK(hipMemcpyAsync(e->h_audio, e->d_audio, (size_t)na * sizeof(float), hipMemcpyDeviceToHost,
                                   e->stream));
    ARGUS_CHECK(hipStreamSynchronize(e->stream));
    if (na > 0) std::memcpy(out, e->h_audio, (size_t)na * sizeof(float));
    e->phase += inc * (double)n;
    e->bfo_phase += bfo_inc * (double)m;
    if (n_out) *n_out = na;
    if (rate) *rate = e->sample_rate / D;
    return 0;
}

}  // extern "C"
