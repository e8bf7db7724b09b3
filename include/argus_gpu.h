/*
 * argus_gpu.h -- C ABI of libargus_hip.so, the MI355X (gfx950) implementation
 * of Argus's per-signal extraction and surveillance hot path.
 *
 * Conventions (kept from the reference's native export surface,
 * internal/demod/gpudemod/native/exports.cu:4-12):
 *   - every function returns int: 0 = OK, negative = argument/capacity error
 *     (same numbering as the entry point it succeeds), positive = hipError_t;
 *   - a launch with n <= 0 is a no-op returning 0;
 *   - complex samples are gpud_float2 {float x, y} == Go complex64;
 *   - device pointers are plain pointers from gpud_malloc; the library never
 *     aborts the process.
 * Names: a successor of a reference export keeps its name with the "_cuda"
 * suffix replaced by "_hip"; entry points with no reference counterpart use
 * the gpud_runner_ / gpud_surv_ / gpud_serve_ / gpud_wfm_ prefixes.
 */
#ifndef ARGUS_GPU_H
#define ARGUS_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPUD_API __attribute__((visibility("default")))

typedef void* gpud_stream_handle;             /* exports.cu:12 */
typedef struct gpud_float2 { float x, y; } gpud_float2; /* windows_bridge.go:12 */

/* ------------------------------------------------------------------ */
/* Device, memory and stream helpers.  The reference's cgo preamble     */
/* calls cudart for these directly (gpudemod.go:13-31,                  */
/* windows_bridge.go:50-54); here they sit behind the C ABI so Go never */
/* includes a HIP header.                                               */
/* ------------------------------------------------------------------ */
GPUD_API int gpud_device_count(int* count);                        /* gpudemod.go:112-118 Available() */
GPUD_API int gpud_malloc(void** ptr, size_t bytes);                /* gpudemod.go:13-16 gpud_cuda_malloc */
GPUD_API int gpud_free(void* ptr);                                 /* gpudemod.go:17-19 gpud_cuda_free */
GPUD_API int gpud_host_alloc(void** ptr, size_t bytes);            /* pinned host memory (no reference counterpart) */
GPUD_API int gpud_host_free(void* ptr);
GPUD_API int gpud_memcpy_h2d(void* dst, const void* src, size_t bytes); /* gpudemod.go:20-22 */
GPUD_API int gpud_memcpy_d2h(void* dst, const void* src, size_t bytes); /* gpudemod.go:23-25 */
GPUD_API int gpud_memcpy_d2d(void* dst, const void* src, size_t bytes);
GPUD_API int gpud_memset_zero(void* dst, size_t bytes);
GPUD_API int gpud_device_sync(void);                               /* gpudemod.go:26-28 */
GPUD_API int gpud_stream_create(gpud_stream_handle* out);          /* exports.cu:18-25 */
GPUD_API int gpud_stream_destroy(gpud_stream_handle stream);       /* exports.cu:27-30 */
GPUD_API int gpud_stream_sync(gpud_stream_handle stream);          /* exports.cu:32-35 */
GPUD_API const char* gpud_version(void);
/* HIP events for timing on a given stream (no reference counterpart: the
 * reference has no GPU timing). */
typedef void* gpud_event_handle;
GPUD_API int gpud_event_create(gpud_event_handle* ev);
GPUD_API int gpud_event_destroy(gpud_event_handle ev);
GPUD_API int gpud_event_record(gpud_event_handle ev, gpud_stream_handle stream);
GPUD_API int gpud_event_elapsed_ms(gpud_event_handle start, gpud_event_handle stop, float* ms);

/* ------------------------------------------------------------------ */
/* Primitive launchers: successors of exports.cu's kernels.             */
/* ------------------------------------------------------------------ */
/* exports.cu:62-75  out[i] = in[i] * e^{j(phase_start + phase_inc*i)} */
GPUD_API int gpud_launch_freq_shift_stream_hip(const gpud_float2* in, gpud_float2* out, int n, double phase_inc,
                                               double phase_start, gpud_stream_handle stream);
/* exports.cu:77-85 */
GPUD_API int gpud_launch_freq_shift_hip(const gpud_float2* in, gpud_float2* out, int n, double phase_inc,
                                        double phase_start);
/* exports.cu:102-112  n-1 outputs: atan2(Im(conj(x_i) x_{i+1}), Re(...)) */
GPUD_API int gpud_launch_fm_discrim_hip(const gpud_float2* in, float* out, int n);
/* exports.cu:163-167  taps for gpud_launch_fir[_stream]_hip, n <= 256, else -1 */
GPUD_API int gpud_upload_fir_taps_hip(const float* taps, int n);
/* exports.cu:169-181  causal FIR with the uploaded taps, zero history */
GPUD_API int gpud_launch_fir_hip(const gpud_float2* in, gpud_float2* out, int n, int num_taps);
/* exports.cu:183-196 */
GPUD_API int gpud_launch_fir_stream_hip(const gpud_float2* in, gpud_float2* out, int n, int num_taps,
                                        gpud_stream_handle stream);
/* exports.cu:230-244  causal FIR, taps from a device pointer */
GPUD_API int gpud_launch_fir_v2_stream_hip(const gpud_float2* in, gpud_float2* out, const float* taps, int n,
                                           int num_taps, gpud_stream_handle stream);
/* exports.cu:246-257  out[i] = in[i*factor] */
GPUD_API int gpud_launch_decimate_hip(const gpud_float2* in, gpud_float2* out, int n_out, int factor);
/* exports.cu:259-271 */
GPUD_API int gpud_launch_decimate_stream_hip(const gpud_float2* in, gpud_float2* out, int n_out, int factor,
                                             gpud_stream_handle stream);
/* exports.cu:284-294  |x| */
GPUD_API int gpud_launch_am_envelope_hip(const gpud_float2* in, float* out, int n);
/* exports.cu:314-326  Re(x * e^{j(phase_start + phase_inc*i)}) */
GPUD_API int gpud_launch_ssb_product_hip(const gpud_float2* in, float* out, int n, double phase_inc,
                                         double phase_start);

/* exports.cu:356-460  legacy one-shot streaming polyphase (history in/out
 * arrays).  Errors: -1 bad args, -2 polyphase_len < decim*phase_len,
 * -3 null in_new, -4 null out, -5 null history_in.  Unlike the reference it
 * allocates nothing per call. */
GPUD_API int gpud_launch_streaming_polyphase_prepare_hip(
    const gpud_float2* in_new, int n_new, const gpud_float2* history_in, int history_len,
    const float* polyphase_taps, int polyphase_len, int decim, int num_taps, int phase_count_in,
    double phase_start, double phase_inc, gpud_float2* out, int* n_out, int* phase_count_out,
    double* phase_end_out, gpud_float2* history_out);

/* exports.cu:533-631  production stateful streaming extractor for one signal:
 * NCO shift + phase-major polyphase FIR/decimate + history update, one fused
 * kernel.  history_state holds the shifted history (<= history_cap samples);
 * the new history is written to history_scratch and copied back on the same
 * stream.  Errors: -10 missing pointer/arg, -11 negative size, -12
 * polyphase_len < decim*ceil(num_taps/decim), -13 null in_new or
 * shifted_new_tmp with n_new > 0, -14 out_cap too small, -15 null out,
 * -16 null history buffers.  shifted_new_tmp is accepted for ABI
 * compatibility and not written (the shift is fused).  Runs on the default
 * stream and returns after the history copy has been queued. */
GPUD_API int gpud_launch_streaming_polyphase_stateful_hip(
    const gpud_float2* in_new, int n_new, gpud_float2* shifted_new_tmp, const float* polyphase_taps,
    int polyphase_len, int decim, int num_taps, gpud_float2* history_state, gpud_float2* history_scratch,
    int history_cap, int* history_len_io, int* phase_count_state, double* phase_state, double phase_inc,
    gpud_float2* out, int out_cap, int* n_out);

/* ------------------------------------------------------------------ */
/* Streaming batch runner: successor of gpudemod.BatchRunner's          */
/* streaming surface (batch_runner.go:23-34, streaming_gpu_stub.go:21,  */
/* stream_state.go:9-24).  State (NCO phase, decimator phase, shifted   */
/* history, taps) stays on the device between calls.                    */
/* ------------------------------------------------------------------ */
typedef struct gpud_runner gpud_runner;

/* streaming_types.go:8-15  StreamingExtractJob */
typedef struct gpud_streaming_job {
    int64_t  signal_id;
    double   offset_hz;
    double   bandwidth;
    int32_t  out_rate;
    int32_t  num_taps;     /* <= 0 -> 101 (stream_state.go:50-53) */
    uint64_t config_hash;  /* a change resets the signal's state (streaming_types.go:46-56) */
} gpud_streaming_job;

/* streaming_types.go:17-24  StreamingExtractResult (+ where its IQ landed) */
typedef struct gpud_streaming_result {
    int64_t signal_id;
    int32_t rate;
    int32_t n_out;
    int32_t phase_count;
    int32_t history_len;
    int64_t out_offset;   /* complex-sample offset of this signal's IQ in the output buffer */
    double  nco_phase;    /* NCO phase carried to the next call, radians in [-pi, pi) */
} gpud_streaming_result;

enum {
    GPUD_EXEC_BATCHED = 0,     /* one fused launch for all signals per chunk (default) */
    GPUD_EXEC_PER_SIGNAL = 1   /* plain form: per-signal upload, launch and sync copies,
                                  the reference's call pattern (streaming_gpu_native_prepare.go:17-108) */
};

/* Tile and launch shape of the fused extraction launch. */
typedef struct gpud_tune {
    int32_t lds_entries;     /* complex samples of LDS one dense tile may stage (1024..16384; default 6144) */
    int32_t narrow_outputs;  /* outputs per narrow work item (default 32) */
    int32_t grid;            /* persistent workgroups, a multiple of 8 (default 2048) */
    int32_t fast_grid;       /* workgroups of the fast form's per-channel kernel, a multiple of 8
                              * (default 2048; 0 in a set call keeps the current value) */
    float   kernel_ms;       /* autotune: best launch time found */
    float   pad2;
} gpud_tune;
/* batch_runner.go:23-34 NewBatchRunner(maxSamples, sampleRate).  -1 bad args. */
GPUD_API int gpud_runner_create(int max_samples, int sample_rate, gpud_runner** out);
/* batch_runner.go:36-50 Close */
GPUD_API int gpud_runner_destroy(gpud_runner* r);
GPUD_API int gpud_runner_set_exec_mode(gpud_runner* r, int mode);
/* The fast-convolution form in the batched launch (on by default; no
 * reference counterpart): a signal whose decimation D is a multiple of 8
 * with D < T <= 512 is filtered through shared 4096-point transforms of
 * the chunk (overlap-save, decimation by spectral folding) once its history
 * is full -- the same filter, outputs within FP32 rounding of the direct
 * form.  The runner keeps the stream's last 511 samples for it; a call in
 * the per-signal mode restarts that count. */
GPUD_API int gpud_runner_set_fast(gpud_runner* r, int enable);
/* Tap window for the low-pass design of the runner's signals: 0 = Hann
 * (LowpassFIR, dsp/fir.go:6-28, the default), 1 = Hamming (same sinc). */
GPUD_API int gpud_runner_set_window(gpud_runner* r, int window);
/* Kernel timing: while enabled, the runner brackets every fused extraction
 * launch with HIP events on its stream; gpud_runner_kernel_time waits for
 * them and returns the summed kernel milliseconds and the launch count since
 * it was enabled. */
GPUD_API int gpud_runner_enable_kernel_timing(gpud_runner* r, int enable);
GPUD_API int gpud_runner_kernel_time(gpud_runner* r, double* total_ms, int* launches);
/* Tile and launch shape of the runner's fused launches (no reference
 * counterpart; see gpud_serve_autotune).  -1 if out of range. */
GPUD_API int gpud_runner_get_tune(gpud_runner* r, gpud_tune* t);
GPUD_API int gpud_runner_set_tune(gpud_runner* r, const gpud_tune* t);
/* Stream the runner launches on (created by the runner, default priority). */
GPUD_API gpud_stream_handle gpud_runner_stream(gpud_runner* r);

/* streaming_gpu_stub.go:21-26 StreamingExtractGPU.  iq is host memory.
 * Outputs of job i are written to out + results[i].out_offset (assigned in
 * job order).  Signals absent from jobs are pruned (streaming_gpu_prepare.go:
 * 40-44).  Errors: -1 bad args / non-integer decimation
 * (cpu_oracle.go:155-163), -2 n_new > max_samples, -14 out_cap too small. */
GPUD_API int gpud_runner_streaming_extract(gpud_runner* r, const gpud_float2* iq, int n_new,
                                           const gpud_streaming_job* jobs, int n_jobs, gpud_float2* out,
                                           int64_t out_cap, gpud_streaming_result* results);
/* Same with device-resident input; on return *out_dev points at the packed
 * device outputs (valid until the next call on this runner).  No host
 * copies, no synchronisation: completion is ordered on gpud_runner_stream. */
GPUD_API int gpud_runner_streaming_extract_device(gpud_runner* r, const gpud_float2* iq_dev, int n_new,
                                                  const gpud_streaming_job* jobs, int n_jobs,
                                                  gpud_float2** out_dev, gpud_streaming_result* results);
/* Inspect a signal's carried state (ExtractStreamState, streaming_types.go:26-44).
 * Returns the history length, -1 if the signal has no state. */
GPUD_API int gpud_runner_get_signal_state(gpud_runner* r, int64_t signal_id, uint64_t* config_hash,
                                          double* nco_phase, int* decim, int* phase_count, int* num_taps,
                                          gpud_float2* history, int hist_cap);
GPUD_API int gpud_runner_num_states(gpud_runner* r);
GPUD_API int gpud_runner_state_ids(gpud_runner* r, int64_t* ids, int cap);
/* stream_state.go:9-16 ResetSignalState */
GPUD_API int gpud_runner_reset_signal_state(gpud_runner* r, int64_t signal_id);
/* stream_state.go:18-24 ResetAllSignalStates */
GPUD_API int gpud_runner_reset_all_signal_states(gpud_runner* r);
/* ------------------------------------------------------------------ */
/* Output rings: successor of streamOutRing / outBufPool                 */
/* (internal/demod/gpudemod/streaming_out_ring.go:22-112).  Each signal  */
/* owns a ring of GPUD_OUT_RING_DEPTH host buffers; a call hands its     */
/* outputs out in the ring's next buffer, which stays untouched for the  */
/* following DEPTH - 1 calls.  A slot too small for a call draws a buffer */
/* from a free pool shared by all signal IDs (first free buffer that     */
/* fits, else the largest, released if still too small, else a new one  */
/* of exactly n samples); rings of pruned signals return their buffers   */
/* to the pool, which keeps at most 4 x 32 of them.  Host memory only:   */
/* these calls need no GPU.  Not thread-safe.                            */
/* ------------------------------------------------------------------ */
#define GPUD_OUT_RING_DEPTH 4
typedef struct gpud_out_rings gpud_out_rings;
GPUD_API int gpud_out_rings_create(gpud_out_rings** out);
GPUD_API int gpud_out_rings_destroy(gpud_out_rings* o);
/* streaming_out_ring.go:32-45 next (with outRingFor 91-101): the buffer
 * for signal_id's next n samples.  -3 when host memory runs out. */
GPUD_API int gpud_out_rings_next(gpud_out_rings* o, int64_t signal_id, int64_t n, gpud_float2** buf);
/* streaming_out_ring.go:105-112 pruneOutRings: rings of signals not in
 * active[0..n_active) return their buffers to the pool. */
GPUD_API int gpud_out_rings_prune(gpud_out_rings* o, const int64_t* active, int n_active);
/* streaming_out_ring.go:56-78 outBufPool.get and 81-89 put.  A buffer put
 * must come from this object (get or next); its ownership moves to the pool. */
GPUD_API int gpud_out_rings_pool_get(gpud_out_rings* o, int64_t n, gpud_float2** buf, int64_t* cap);
GPUD_API int gpud_out_rings_pool_put(gpud_out_rings* o, gpud_float2* buf, int64_t cap);
/* Live rings, pooled buffers, and buffers allocated / freed so far. */
GPUD_API int gpud_out_rings_stats(const gpud_out_rings* o, int* n_rings, int* n_free, int64_t* allocated,
                                  int64_t* released);
/* A signal's ring: the slot its next call takes and the capacity of each
 * of its GPUD_OUT_RING_DEPTH buffers (caps may be NULL).  -1 if it has none. */
GPUD_API int gpud_out_rings_slot(const gpud_out_rings* o, int64_t signal_id, int* next_idx, int64_t* caps);
/* StreamingExtractGPU with the reference's output ownership
 * (streaming_gpu_native_prepare.go:78-104): as gpud_runner_streaming_extract,
 * but job i's outputs land in outs[i], the next buffer of that signal's
 * ring in the runner's own rings (NULL when it has no outputs).  Each call
 * prunes the rings of signals absent from jobs together with their states. */
GPUD_API int gpud_runner_streaming_extract_ring(gpud_runner* r, const gpud_float2* iq, int n_new,
                                                const gpud_streaming_job* jobs, int n_jobs, gpud_float2** outs,
                                                gpud_streaming_result* results);
/* The runner's rings (created on its first ring call; NULL before). */
GPUD_API gpud_out_rings* gpud_runner_out_rings(gpud_runner* r);
/* streaming_types.go:58-72 StreamingConfigHash (FNV-1a-64) */
GPUD_API uint64_t gpud_streaming_config_hash(int64_t signal_id, double offset_hz, double bandwidth, int out_rate,
                                             int num_taps, int sample_rate);
/* cpu_oracle.go:155-163 ExactIntegerDecimation: decimation or -1 */
GPUD_API int gpud_exact_integer_decimation(int sample_rate, int out_rate);
/* dsp/fir.go:6-28 LowpassFIR (Hann); returns the (odd) tap count */
GPUD_API int gpud_lowpass_fir(double cutoff_hz, int sample_rate, int taps, double* out);
/* polyphase.go:6-22 BuildPolyphaseTapsPhaseMajor; returns decim*ceil(n/decim) */
GPUD_API int gpud_build_polyphase_taps(const float* base, int n, int decim, float* out);

/* ------------------------------------------------------------------ */
/* Job building of the DSP loop (cmd/sdrd): rate snapping and the job   */
/* builders that turn detected signals into extraction jobs.  Host      */
/* integer/FP64 work.                                                   */
/* ------------------------------------------------------------------ */
enum { GPUD_MOD_OTHER = 0, GPUD_MOD_WFM = 1, GPUD_MOD_WFM_STEREO = 2 };
/* detector.Signal fields the builders read (detector.go:82-94): mod_type is
 * Signal.Class.ModType ("WFM" / "WFM_STEREO", else other or unclassified). */
typedef struct gpud_detected_signal {
    int64_t id;
    double  center_hz;
    double  bw_hz;
    int32_t mod_type;
    int32_t pad;
} gpud_detected_signal;
/* extractionConfig (helpers.go:270-274).  snap_wfm = 0 keeps
 * buildStreamingJobs as written (WFM out-rate 512000, which must divide the
 * sample rate); 1 snaps it with snapOutRate, as the legacy builders do
 * (helpers.go:194, 478). */
typedef struct gpud_extraction_config {
    int32_t fir_taps;  /* <= 0 -> 101 */
    int32_t snap_wfm;
    double  bw_mult;   /* <= 0 -> 1.0 */
} gpud_extraction_config;
/* batch.go:5-10 ExtractJob (legacy block path) */
typedef struct gpud_extract_job {
    double  offset_hz;
    double  bw;
    int32_t out_rate;
    int32_t pad;
    double  phase_start;
} gpud_extract_job;
/* batch.go:14-18 ExtractResult, plus where its IQ landed */
typedef struct gpud_extract_result {
    int32_t rate;
    int32_t n_out;
    int64_t out_offset;   /* complex-sample offset of the job's IQ in the output buffer */
    double  phase_end;    /* phase_start + phase_inc * n (batch_runner.go:87-90) */
} gpud_extract_result;
/* helpers.go:292-310 snapOutRate: sample_rate / d for the divisor d <= 256 of
 * sample_rate nearest sample_rate / target (first on ties). */
GPUD_API int gpud_snap_out_rate(int sample_rate, int target);
/* streaming_refactor.go:105-137 nearestExactDecimationRate */
GPUD_API int gpud_nearest_exact_decimation_rate(int sample_rate, int target_rate, int min_rate);
/* streaming_refactor.go:16-60 buildStreamingJobs: 0, -1 bad arguments, -2
 * when a job's out-rate does not divide the sample rate (the reference's
 * ExactIntegerDecimation error; *failed_index = that signal). */
GPUD_API int gpud_build_streaming_jobs(int sample_rate, double center_hz, const gpud_detected_signal* signals,
                                       int n_signals, const gpud_extraction_config* cfg, gpud_streaming_job* jobs,
                                       int* failed_index);
/* helpers.go:175-195 the ExtractJob list of extractSignalIQBatch (BW floors
 * 250 kHz WFM / 20 kHz others, out-rate snapOutRate(512 k WFM, 200 k)). */
GPUD_API int gpud_build_extract_jobs(int sample_rate, double center_hz, const gpud_detected_signal* signals,
                                     int n_signals, gpud_extract_job* jobs);

/* ------------------------------------------------------------------ */
/* Legacy block path: successors of BatchRunner.ShiftFilterDecimateBatch */
/* [WithPhase] (batch_runner.go:61-98, batch_runner_windows.go:87-182)   */
/* and of the demod Engine (gpudemod_windows.go:94-654).  Shift from the */
/* phase start, causal FIR from zero history, keep filtered samples 0,   */
/* D, 2D, ...; only the kept outputs are computed (one fused launch).    */
/* ------------------------------------------------------------------ */
/* Per job: 101-tap LowpassFIR at max(bw/2, 200) Hz, decimation
 * out_rate -> fs/out_rate (exact), n/D outputs at results[i].out_offset.
 * n = 0 gives empty results with rate 0.  Errors: -1 bad args, -2 n >
 * max_samples, -3 non-integer decimation, -4 fewer than D samples, -14
 * out_cap too small.  Synchronous, like the reference. */
GPUD_API int gpud_runner_shift_filter_decimate_batch(gpud_runner* r, const gpud_float2* iq, int n,
                                                     const gpud_extract_job* jobs, int n_jobs, gpud_float2* out,
                                                     int64_t out_cap, gpud_extract_result* results);

typedef struct gpud_engine gpud_engine;
/* gpudemod.go:81-90 DemodType */
enum { GPUD_ENGINE_NFM = 0, GPUD_ENGINE_WFM = 1, GPUD_ENGINE_AM = 2, GPUD_ENGINE_USB = 3, GPUD_ENGINE_LSB = 4,
       GPUD_ENGINE_CW = 5 };
/* gpudemod_windows.go:112-162 New(maxSamples, sampleRate) */
GPUD_API int gpud_engine_create(int max_samples, int sample_rate, gpud_engine** out);
GPUD_API int gpud_engine_destroy(gpud_engine* e);
/* gpudemod_windows.go:164-176 SetFIR: n <= 0 clears, longer than 256 is cut */
GPUD_API int gpud_engine_set_fir(gpud_engine* e, const float* taps, int n);
/* The carried shift and BFO phases and the current tap count. */
GPUD_API int gpud_engine_state(gpud_engine* e, double* phase, double* bfo_phase, int* fir_taps);
/* gpudemod_windows.go:320-393 ShiftFilterDecimate: taps = the engine's (or
 * LowpassFIR(max(bw/2, 200), fs, 101), designed and kept), D =
 * round(fs/out_rate), shift from the carried phase, n/D outputs, rate fs/D.
 * Errors: -1, -2 n > max_samples, -3 out_rate <= 0, -4 n < D, -14. */
GPUD_API int gpud_engine_shift_filter_decimate(gpud_engine* e, const gpud_float2* iq, int n, double offset_hz,
                                               double bw, int out_rate, gpud_float2* out, int out_cap, int* n_out,
                                               int* rate);
/* gpudemod_windows.go:395-521 DemodFused (fused = 1) and 523-654 Demod
 * (fused = 0): the chain to 48 kHz (192 kHz WFM), then the discriminator
 * (n_out = m - 1), |x| (AM) or Re(x e^{j bfo}) with the carried BFO phase at
 * +-700 Hz over the input rate (USB/CW +700, LSB -700).  -5 unsupported mode;
 * -4 when the chain gives fewer than 2 (fused) or 1 (Demod) samples. */
GPUD_API int gpud_engine_demod(gpud_engine* e, const gpud_float2* iq, int n, double offset_hz, double bw, int mode,
                               int fused, float* out, int out_cap, int* n_out, int* rate);

/* ------------------------------------------------------------------ */
/* Streaming service (no reference counterpart: the reference drives one */
/* runner per DSP tick from a goroutine, cmd/sdrd/dsp_loop.go:54-172).  */
/* Channels are admitted against a fixed HBM budget and retired while   */
/* blocks flow; each block is one fused extraction launch plus one      */
/* demodulation launch over every live channel, pipelined against the   */
/* block's upload and the previous block's download.                    */
/* ------------------------------------------------------------------ */
typedef struct gpud_serve gpud_serve;

enum {
    GPUD_DEMOD_NONE = 0, /* baseband IQ only */
    GPUD_DEMOD_NFM = 1,  /* fm.go:179-196, bridged (streamer.go:1287-1302) */
    GPUD_DEMOD_WFM = 2,  /* same discriminator */
    GPUD_DEMOD_AM = 3,   /* am.go:11-27 */
    GPUD_DEMOD_USB = 4,  /* ssb.go:19-39, BFO +700 Hz, carried phase */
    GPUD_DEMOD_LSB = 5,  /* ssb.go:19-39, BFO -700 Hz */
    GPUD_DEMOD_CW = 6,   /* cw.go:11-26 */
    /* WFM multiplex: the discriminator as WFM, then the stateful stereo
     * decoder with lock hysteresis and mono fallback (streamer.go:1367-1395,
     * 1496-1643) and the RDS baseband (fm.go:234-271) carried across
     * blocks.  audio = 2 x samples, L/R interleaved. */
    GPUD_DEMOD_WFM_STEREO = 7
};

#define GPUD_ERR_BUDGET (-20)    /* admission refused: HBM budget exhausted */
#define GPUD_ERR_DUPLICATE (-21) /* channel id already live */

typedef struct gpud_channel_spec {
    int64_t channel_id;
    double  offset_hz;
    double  bandwidth;    /* low-pass cutoff = max(bandwidth/2, 200) (stream_state.go:54-57) */
    int32_t out_rate;     /* must divide the sample rate */
    int32_t num_taps;     /* <= 0 -> 101 */
    int32_t demod;        /* GPUD_DEMOD_* */
    int32_t window;       /* 0 = Hann (LowpassFIR), 1 = Hamming */
} gpud_channel_spec;

typedef struct gpud_serve_block_info {
    int64_t ticket;
    int32_t n_channels;
    int32_t n_samples;
    int64_t iq_total;
    int64_t audio_total;
    float   gpu_ms;       /* event time: upload start -> download end */
    float   kernel_ms;    /* event time: extraction + demodulation kernels */
    float   extract_ms;   /* event time: the extraction launches (fast form, the demodulation of the
                           * channels it completes, direct form) */
    float   pad;
} gpud_serve_block_info;

/* hbm_budget_bytes bounds the channels admitted (their state plus their share
 * of the pipelined output buffers).  Channel state lives in an arena grown
 * in regions -- 256 MiB first, each next one as large as all before, up to
 * min(budget, 16 GiB) -- kept until the service is destroyed: opening a
 * channel allocates on the device only when the arena grows, retiring one
 * frees nothing (its block returns to the arena), and a channel's initial
 * state is copied asynchronously on the service's compute stream, ahead of
 * its first block. */
GPUD_API int gpud_serve_create(int sample_rate, int block_samples, uint64_t hbm_budget_bytes, gpud_serve** out);
GPUD_API int gpud_serve_destroy(gpud_serve* s);
/* 0, -1 bad args / non-integer decimation, GPUD_ERR_DUPLICATE, GPUD_ERR_BUDGET */
GPUD_API int gpud_serve_open(gpud_serve* s, const gpud_channel_spec* spec);
/* Per-channel options beyond the spec. */
typedef struct gpud_channel_opts {
    /* > 0: the live recorder's pre-demod stage (streamer.go:1098-1161): a
     * stateful decimating FIR (dsp.StatefulDecimatingFIRComplex,
     * decimating_fir.go:51-95; 101 FP64 taps, LowpassFIR at 90 kHz for WFM,
     * else 0.8 of the demod Nyquist) by round(out_rate / demod_rate), forced
     * to 2 for WFM when out_rate / 2 >= 200 kHz; the demodulator, stereo,
     * RDS and audio stages then run at out_rate / factor.  0: none. */
    int32_t demod_rate;
    int32_t pad[7];
} gpud_channel_opts;
/* gpud_serve_open with options (opts may be NULL). */
GPUD_API int gpud_serve_open_ex(gpud_serve* s, const gpud_channel_spec* spec, const gpud_channel_opts* opts);
GPUD_API int gpud_serve_close(gpud_serve* s, int64_t channel_id);
GPUD_API int gpud_serve_info(gpud_serve* s, int* live_channels, uint64_t* bytes_in_use, uint64_t* budget);
/* Enqueue a block (host memory, copied before return) and return at once. */
GPUD_API int gpud_serve_submit(gpud_serve* s, const gpud_float2* iq_host, int n, int64_t* ticket);
/* Enqueue a block already resident in device memory (must stay valid until
 * the block completes). */
GPUD_API int gpud_serve_submit_device(gpud_serve* s, const gpud_float2* iq_dev, int n, int64_t* ticket);
/* Disable (0) / enable (1, default) the download of outputs to host memory. */
GPUD_API int gpud_serve_set_download(gpud_serve* s, int enable);
/* Which outputs of each block come to host memory (the others stay on the
 * device; their host views keep older contents). */
enum {
    GPUD_DL_IQ = 1,        /* baseband at out_rate */
    GPUD_DL_DEMOD = 2,     /* demodulator output (L/R frames for stereo) */
    GPUD_DL_WFM = 4,       /* RDS baseband and stereo lock flags */
    GPUD_DL_AUDIO = 8,     /* audio-stage output */
    GPUD_DL_PREDEMOD = 16, /* pre-demod stage output */
    GPUD_DL_ALL = 31
};
GPUD_API int gpud_serve_set_download_parts(gpud_serve* s, int mask);
/* No counterpart in the reference (its extraction returns outputs per call).
 * The fast-convolution channels whose outputs are complete after their own
 * kernels (no pre-demod, stereo or audio stage) are computed in up to
 * max_parts (1..4) launches over consecutive channel ranges, one launch per
 * min_part_bytes of such outputs the block downloads (GPUD_DL_IQ and
 * GPUD_DL_DEMOD), and each range's outputs start down to the host as soon as
 * its launch ends.  Defaults 4 and 16 MiB; min_part_bytes 0 always uses
 * max_parts.  Outputs are the same whatever the split.  -1 on bad args. */
GPUD_API int gpud_serve_set_download_split(gpud_serve* s, int max_parts, int64_t min_part_bytes);
/* Non-blocking completion test of a submitted block: 1 when its outputs
 * (those downloaded) are on the host, 0 while it runs, -2 unknown ticket. */
GPUD_API int gpud_serve_poll(gpud_serve* s, int64_t ticket);
/* Graph instantiations so far (a block whose captured graphs still fit is
 * replayed, not re-captured) and the free bytes of the channel-state arena. */
GPUD_API int gpud_serve_counters(gpud_serve* s, int64_t* graph_captures, uint64_t* arena_free);
/* Wait for a block.  -2 if the ticket is unknown or its slot was reused. */
GPUD_API int gpud_serve_wait(gpud_serve* s, int64_t ticket, gpud_serve_block_info* info);
/* A channel's outputs in a completed block (pinned host memory, valid until
 * its slot is reused, i.e. for the next two submits). */
GPUD_API int gpud_serve_channel_output(gpud_serve* s, int64_t ticket, int64_t channel_id, const gpud_float2** iq,
                                       int* n_iq, const float** audio, int* n_audio, int* rate);
/* The pre-demod stage's output of a channel in a completed block (the
 * demodulator's input, at *rate; pinned host memory, valid as
 * gpud_serve_channel_output's).  -1 if the channel was not part of the
 * block, -3 if it has no pre-demod stage. */
GPUD_API int gpud_serve_channel_predemod(gpud_serve* s, int64_t ticket, int64_t channel_id, const gpud_float2** iq,
                                         int* n, int* rate);
/* Replay (1, default) or not (0) a captured per-slot hipGraph of each
 * block's launches; waits for blocks in flight. */
GPUD_API int gpud_serve_set_graph(gpud_serve* s, int enable);
GPUD_API int gpud_serve_get_tune(gpud_serve* s, gpud_tune* t);
GPUD_API int gpud_serve_set_tune(gpud_serve* s, const gpud_tune* t);
/* Autotune the tile and launch shape on the live channel set with a block
 * at iq_dev (device memory, n samples): times the extraction stage (best of
 * reps) over grid, LDS budget and narrow item size, one parameter at a
 * time, keeps the fastest and returns it.  Channel state does not advance.
 * Tune after the channels' first block, so channels that take the fast
 * form once their history is full are timed in it. */
/* Event-timed extraction stage of the live channel set at the current tune
 * on the block at iq_dev (n samples, device memory): the fast form's
 * transforms and per-channel pass, then the fused direct launch, as a block
 * runs them, without the stream-tail update; reps launches on a scratch slot
 * after one warm-up (blocks in flight are completed first; channel state
 * does not advance); average and minimum milliseconds (either pointer may
 * be NULL).  No reference counterpart: the measurement behind a kernel
 * roofline. */
GPUD_API int gpud_serve_extract_time(gpud_serve* s, const gpud_float2* iq_dev, int n, int reps, float* avg_ms,
                                     float* min_ms);
GPUD_API int gpud_serve_autotune(gpud_serve* s, const gpud_float2* iq_dev, int n, int reps, gpud_tune* best);
/* RDS baseband samples (rate = out_rate / floor(out_rate / 19000)) and the
 * stereo lock of a GPUD_DEMOD_WFM_STEREO channel in a completed block:
 * *locked = this block's lock test (streamer.go:1634), *stereo = 1 when the
 * block's audio is the L/R matrix, 0 when it is the mono fallback.  Valid
 * while downloads are enabled.  -3 if the channel is not a stereo channel. */
GPUD_API int gpud_serve_channel_wfm(gpud_serve* s, int64_t ticket, int64_t channel_id, const gpud_float2** rds,
                                    int* n_rds, int* rds_rate, int* locked, int* stereo);

/* Post-demodulation audio stage of one live channel (streamer.go:1382-1470
 * with internal/dsp/resample.go:28-170): plain WFM is first low-passed at
 * 15 kHz (wfmAudioFilter, 101 taps, streamer.go:1511-1520); every mode is
 * resampled from its out_rate to audio_rate by the rational polyphase
 * resampler (NewResampler(in, out, 32); two of them for the L/R frames of a
 * GPUD_DEMOD_WFM_STEREO channel); WFM and WFM stereo are then de-emphasised
 * (time constant deemphasis_us, 0 = off; FP64 state) and scaled by 0.35.
 * audio_rate 0 turns the stage off.  Resets the stage's state (a new
 * resampler, filter and de-emphasis state) and waits for blocks in flight;
 * applies from the next submitted block.  0, -1 bad arguments or unknown
 * channel, GPUD_ERR_BUDGET if the stage's device memory does not fit. */
GPUD_API int gpud_serve_set_audio(gpud_serve* s, int64_t channel_id, int audio_rate, double deemphasis_us);
/* The audio stage's output of a channel in a completed block: n_frames
 * frames of `channels` (1, or 2 interleaved L/R) float samples at `rate`
 * (pinned host memory, valid as gpud_serve_channel_output's).  -1 if the
 * channel was not part of the block, -3 if its audio stage was off. */
GPUD_API int gpud_serve_channel_audio(gpud_serve* s, int64_t ticket, int64_t channel_id, const float** audio,
                                      int* n_frames, int* channels, int* rate);

/* Device-side RDS group decoding of a live GPUD_DEMOD_WFM_STEREO channel
 * (the decode of pipeline_runtime.go:1132-1142 with rds.Decoder.Decode,
 * rds.go:78-264, on the GPU): the channel's streaming RDS baseband fills
 * windows of window_samples samples on the device, and each full window is
 * decoded there by one Decode call of the channel's decoder (the reference
 * decodes the latest 4 s every 4 s: window_samples = 4 * rds_rate).
 * window_samples 0 turns decoding off; else it must hold a group (>= 104)
 * and a whole block's RDS samples.  Resets the decoder and waits for blocks
 * in flight; applies from the next submitted block.  0, -1 bad arguments,
 * unknown or non-stereo channel, GPUD_ERR_BUDGET if it does not fit. */
GPUD_API int gpud_serve_set_rds(gpud_serve* s, int64_t channel_id, int window_samples);
/* The fast-convolution form of the extraction (on by default): a channel
 * whose decimation D is a multiple of 8 and whose tap count T satisfies
 * D < T <= 512, once its
 * history is full, is filtered through one shared 4096-point transform of each input
 * segment (overlap-save, decimation by spectral folding) instead of the
 * direct polyphase form; same filter, outputs within FP32 rounding of the
 * direct form.  Applies from the next submitted block. */
GPUD_API int gpud_serve_set_fast(gpud_serve* s, int enable);
/* Demodulation fused into the extraction kernels (no counterpart in the
 * reference, whose demodulators run per signal after extraction,
 * streamer.go:1287-1302): on by default.  For a channel whose demodulator
 * reads its baseband (NFM/WFM, USB/LSB, CW; no pre-demod or stereo stage)
 * and which runs on the fast form or as the direct form's narrow items, the
 * kernel that writes the baseband writes the demodulated outputs too, so
 * the baseband is not read back; FM's first output of each 3584-sample
 * segment or narrow item (and the carried sample) is added by a short pass
 * over those seams.  Off: the separate demodulation pass, the same outputs
 * bit for bit.  Applies from the next submitted block. */
GPUD_API int gpud_serve_set_demod_fusion(gpud_serve* s, int enable);
/* Host helper of the fast form: the 4096 bins of FFT(g) / 4096 for
 * g[k] = taps[k] e^{-j phase_inc k} (k < num_taps <= 512, natural order;
 * FP64, rounded once), the spectrum the service keeps per channel. */
GPUD_API int gpud_fast_spectrum(const float* taps, int num_taps, double phase_inc, gpud_float2* G);
/* RDS decoding of a channel in a completed block: *decoded = 1 if a window
 * completed in this block and was decoded, with *found groups found and the
 * *n_groups groups it applied (4 x 16-bit blocks each, pinned host memory,
 * valid as gpud_serve_channel_output's; whether or not downloads are on).
 * PI, PS (8 + NUL) and RT (64 + NUL) are the decoder's after this block's
 * decode, or after the latest completed one when *decoded = 0.  -1 if the
 * channel was not part of the block, -3 if its decoding was off. */
GPUD_API int gpud_serve_channel_rds(gpud_serve* s, int64_t ticket, int64_t channel_id, int* decoded, int* found,
                                    uint16_t* pi, char* ps9, char* rt65, const uint16_t** groups, int* n_groups);

/* ------------------------------------------------------------------ */
/* SDRIQ001 capture files: successors of iqfile.Write / Read / Source    */
/* (internal/iqfile/iqfile.go:5-166).  Little-endian "SDRIQ001", int64   */
/* sample rate, float64 centre Hz, int64 count, complex64 samples.       */
/* ------------------------------------------------------------------ */
typedef struct gpud_iqfile gpud_iqfile;
/* iqfile.Write (iqfile.go:37-63).  -2 cannot create, -3 write failed. */
GPUD_API int gpud_iqfile_write(const char* path, const gpud_float2* iq, int64_t n, int64_t sample_rate,
                               double center_hz);
/* Open a capture for replay and report its header (iqfile.Read's header,
 * iqfile.go:66-90).  Samples are read from the file as they are asked for,
 * never loaded whole.  -2 cannot open, -3 bad magic or header, -4 file
 * shorter than its declared samples. */
GPUD_API int gpud_iqfile_open(const char* path, gpud_iqfile** out, int64_t* sample_rate, double* center_hz,
                              int64_t* n_samples);
/* Source.ReadIQ (iqfile.go:121-142): n samples from the current position,
 * looping to the start at the end; an empty capture gives zeros. */
GPUD_API int gpud_iqfile_read(gpud_iqfile* f, gpud_float2* dst, int n);
GPUD_API int gpud_iqfile_seek(gpud_iqfile* f, int64_t pos);
GPUD_API int gpud_iqfile_close(gpud_iqfile* f);

/* ------------------------------------------------------------------ */
/* Ingest: a ring of depth pinned host blocks and their device copies,   */
/* filled by a producer ahead of the consumer with asynchronous H2D      */
/* copies on the ring's own stream (the step before the path:            */
/* captureSpectrum's srcMgr.ReadIQ, pipeline_runtime.go:484-556).  No     */
/* reference counterpart as an API: the reference hands host slices to   */
/* each GPU call, which copies them synchronously.                       */
/* ------------------------------------------------------------------ */
typedef struct gpud_ingest gpud_ingest;
/* block_samples per slot, 2 <= depth <= 64 slots, on the current device. */
GPUD_API int gpud_ingest_create(int block_samples, int depth, gpud_ingest** out);
GPUD_API int gpud_ingest_destroy(gpud_ingest* g);
/* Replay a capture (looping, unpaced) from a producer thread; the file must
 * stay open until the ingest is destroyed. */
GPUD_API int gpud_ingest_start_file(gpud_ingest* g, gpud_iqfile* f);
/* Push one block of n <= block_samples samples from the caller's source;
 * waits while every slot is queued or acquired (so the caller must release
 * what it acquires).  Not with a replaying capture. */
GPUD_API int gpud_ingest_push(gpud_ingest* g, const gpud_float2* iq, int n);
/* The next block in order, resident on the device (*dev valid until its
 * release), with its length and sequence number; waits for it and for its
 * copy.  timeout_ms < 0 waits without bound; returns 1 if it timed out. */
GPUD_API int gpud_ingest_acquire(gpud_ingest* g, const gpud_float2** dev, int* n, int64_t* seq, int timeout_ms);
/* Give a block's slot back once the work that reads it has completed. */
GPUD_API int gpud_ingest_release(gpud_ingest* g, int64_t seq);
GPUD_API int gpud_ingest_stats(gpud_ingest* g, int64_t* produced, int64_t* consumed, int* queued);

/* ------------------------------------------------------------------ */
/* RDS group decoder (host): successor of rds.Decoder.Decode             */
/* (internal/rds/rds.go:78-264, tryDecode 365-430, decodeGroupAt 432-473, */
/* applyGroup 475-502, decodeBlockExpected 549-579, crcSyndrome 590-599): */
/* Mueller & Muller timing, Costas loop, differential bits carried across  */
/* calls, CRC block sync with 1-bit FEC and the +-3-bit group flywheel.   */
/* ------------------------------------------------------------------ */
typedef struct gpud_rds_decoder gpud_rds_decoder;
GPUD_API int gpud_rds_decoder_create(gpud_rds_decoder** out);
GPUD_API int gpud_rds_decoder_destroy(gpud_rds_decoder* d);
/* One Decode call on n complex baseband samples.  Returns the groups found in
 * this call (>= 0), -1 on bad arguments. */
GPUD_API int gpud_rds_decode(gpud_rds_decoder* d, const gpud_float2* samples, int n, int sample_rate);
/* PI, PS (8 chars + NUL) and RT (64 chars + NUL) as the decoder holds them
 * (rds.go:604-636: unset characters read as spaces, trailing spaces trimmed). */
GPUD_API int gpud_rds_decoder_state(gpud_rds_decoder* d, uint16_t* pi, char* ps9, char* rt65);
/* Up to cap groups applied so far (4 x 16-bit blocks each, in order);
 * returns the total number applied. */
GPUD_API int gpud_rds_decoder_groups(gpud_rds_decoder* d, uint16_t* groups4, int cap);

/* ------------------------------------------------------------------ */
/* RDS group decoder on the device: n_stations decoders with the         */
/* semantics of rds.Decoder.Decode (rds.go:78-264, tryDecode 365-430,    */
/* applyGroup 475-502, decodeBlockExpected 549-579), their state held in */
/* device memory, decoding many stations' windows in one call.  Where    */
/* pipeline_runtime.go:1132-1137 decodes each station's 4 s window on    */
/* its own goroutine, one call here takes every station due that tick.   */
/* ------------------------------------------------------------------ */
typedef struct gpud_rds_batch gpud_rds_batch;
GPUD_API int gpud_rds_batch_create(gpud_rds_batch** out, int n_stations);
GPUD_API int gpud_rds_batch_destroy(gpud_rds_batch* b);
/* Back to a fresh decoder (state zero, group log empty). */
GPUD_API int gpud_rds_batch_reset(gpud_rds_batch* b, int station);
/* One Decode call for each of n_calls stations (at most 65535, each at
 * most once): station stations[i] decodes the complex baseband
 * samples_dev[offsets[i] .. offsets[i+1]) (device memory) at rates[i];
 * found[i] = groups found in that call.  Kernels run on `stream` after
 * the work already queued there; returns when they have finished and the
 * groups and state are on the host.  -1 bad arguments. */
GPUD_API int gpud_rds_batch_decode_device(gpud_rds_batch* b, int n_calls, const int32_t* stations,
                                          const gpud_float2* samples_dev, const int64_t* offsets,
                                          const int32_t* rates, int32_t* found, gpud_stream_handle stream);
/* PI, PS and RT of a station, as gpud_rds_decoder_state reports them. */
GPUD_API int gpud_rds_batch_state(gpud_rds_batch* b, int station, uint16_t* pi, char* ps9, char* rt65);
/* Up to cap of the station's applied groups in order; returns the total. */
GPUD_API int gpud_rds_batch_groups(gpud_rds_batch* b, int station, uint16_t* groups4, int cap);

/* ------------------------------------------------------------------ */
/* The daemon's RDS caller flow: successor of updateRDS                  */
/* (cmd/sdrd/pipeline_runtime.go:1044-1156) with extractRDS              */
/* (cmd/sdrd/helpers.go:117-160), its fallback chain (pipeline_runtime.go */
/* 1104-1124) and the ID prune of maintenance (1158-1176).  One state per */
/* tracked signal ID, each with a device RDS decoder; a signal decodes    */
/* every 4 s of the caller's clock, a state unseen for more than 30 s is  */
/* dropped.  Every signal due in one update call is decoded in the same   */
/* batched device passes (extraction, stereo pilot, RDS baseband, decoder) */
/* over the call's window of recent IQ, where the reference starts one    */
/* goroutine per signal.  Not thread-safe.                                */
/* ------------------------------------------------------------------ */
typedef struct gpud_rds_runtime gpud_rds_runtime;
/* max_stations: live states at once (decoder slots), at most 65535. */
GPUD_API int gpud_rds_runtime_create(int max_stations, gpud_rds_runtime** out);
GPUD_API int gpud_rds_runtime_destroy(gpud_rds_runtime* rt);
/* updateRDS for n_signals signals in order at time now_s (seconds, the
 * caller's monotonic clock).  window: the recorder ring's recent IQ
 * (SliceRecentInto(4.0), recorder.go:354-370) -- n_window samples at
 * window_rate, tuned to window_center_hz; host memory, or device memory
 * when window_on_device is 1.  A window shorter than one second decodes
 * nothing (the turns are still taken).  Per signal: stereo[i] = the 19 kHz
 * pilot flag of its last decode; station9 + 9 i = its station name (the
 * last decoded PS with spaces trimmed, "" if none yet); turn[i] = 1 when
 * this call took its decode turn, 0 when not due, -1 for ID 0 (skipped).
 * Outputs may be NULL.  Results are final on return.  -1 bad arguments,
 * -8 more live states than max_stations. */
GPUD_API int gpud_rds_runtime_update(gpud_rds_runtime* rt, double now_s, const int64_t* ids, const double* center_hz,
                                     int n_signals, const gpud_float2* window, int64_t n_window, int window_on_device,
                                     int window_rate, double window_center_hz, int32_t* stereo, char* station9,
                                     int32_t* turn);
/* maintenance: drop the states of IDs not in active_ids. */
GPUD_API int gpud_rds_runtime_maintain(gpud_rds_runtime* rt, const int64_t* active_ids, int n_active);
GPUD_API int gpud_rds_runtime_num_states(const gpud_rds_runtime* rt);
/* A state's clock stamps, stereo flag and kept result (PI, PS, RT; empty
 * until a decode reported a PS).  Returns 1 with a result, 0 without, -1
 * if the ID has no state.  last_decode is -inf before the first turn. */
GPUD_API int gpud_rds_runtime_state(const gpud_rds_runtime* rt, int64_t id, double* last_seen, double* last_decode,
                                    int32_t* stereo, uint16_t* pi, char* ps9, char* rt65);
/* The decimated baseband and RDS baseband the last decoding update made for
 * the ID (up to the caps; counts and rates always).  -2 if that update did
 * not decode it. */
GPUD_API int gpud_rds_runtime_last(gpud_rds_runtime* rt, int64_t id, gpud_float2* base, int64_t base_cap,
                                   int64_t* n_base, int32_t* base_rate, gpud_float2* rds, int64_t rds_cap,
                                   int64_t* n_rds, int32_t* rds_rate);
/* Up to cap of the ID's decoded groups in order; returns the total. */
GPUD_API int gpud_rds_runtime_groups(gpud_rds_runtime* rt, int64_t id, uint16_t* groups4, int cap);

/* ------------------------------------------------------------------ */
/* Batch WFM stereo: wfmStereoDecodeBatch (internal/recorder/           */
/* wfm_batch_stereo.go:76-157) over n_stations whole recordings at once. */
/* offsets holds n_stations + 1 sample offsets: station i's             */
/* discriminator output is mono_dev[offsets[i] .. offsets[i+1]) at       */
/* rates[i]; its L/R interleaved audio goes to                           */
/* out_dev[2*offsets[i] .. 2*offsets[i+1]).  Each station is decoded     */
/* from zero filter state and a fresh PLL, and locked[i] is the          */
/* recording's lock test (pilot/total power > 0.003, |f - 19 kHz| <      */
/* 250 Hz, mean |err| < 0.35).  mono_fallback = 0 writes the L/R matrix  */
/* whatever the lock, as wfmStereoDecodeBatch does; mono_fallback = 1    */
/* writes the raw discriminator on both channels for unlocked stations,  */
/* as its caller demodWFMStereoBatchAudio does (lines 42-49).  A station */
/* with no samples or rate <= 0 gets locked = 0 and no output.  Device   */
/* pointers; workspace is taken from the stream-ordered pool; returns    */
/* after the stream has finished.  -1 bad arguments.                     */
/* ------------------------------------------------------------------ */
GPUD_API int gpud_wfm_stereo_batch(const float* mono_dev, const int64_t* offsets, const int32_t* rates,
                                   int n_stations, int mono_fallback, float* out_dev, int32_t* locked,
                                   gpud_stream_handle stream);

/* ------------------------------------------------------------------ */
/* Stereo pilot presence: StereoPilotPresent / estimateWFMPilot         */
/* (internal/classifier/pll.go:82-131) for n_signals WFM baseband IQ    */
/* slices at once.  offsets holds n_signals + 1 sample offsets into      */
/* iq_dev (device); signal i is analysed on its last rates[i]/2 samples: */
/* FP64 FM discriminator, |DFT| at 19 kHz +- 1 kHz in 10 Hz steps (the   */
/* 201 goertzelMagnitude bins), present[i] = median > 0 and max/median > */
/* 4.  ratio[i] = max/median (0 if the median is 0), offset_hz[i] = the  */
/* first maximum's frequency - 19 kHz, mags201 the 201 magnitudes per    */
/* signal in frequency order; each of the three may be NULL.  Signals    */
/* below 40 kHz or with fewer than 2 samples report 0s.  Workspace from  */
/* the stream-ordered pool; returns after the stream has finished.       */
/* -1 bad arguments.                                                     */
/* ------------------------------------------------------------------ */
GPUD_API int gpud_stereo_pilot_present(const gpud_float2* iq_dev, const int64_t* offsets, const int32_t* rates,
                                       int n_signals, int32_t* present, double* ratio, double* offset_hz,
                                       double* mags201, gpud_stream_handle stream);
/* No reference counterpart: the same call for callers that cannot allocate
 * device memory (tight serving loops, code that must not call hipMalloc).
 * Results identical to gpud_stereo_pilot_present; its workspace (the chunk
 * table, partial DFT sums and magnitudes: about 3.2 KiB per 1024
 * discriminator samples analysed) lives in a file-scope device array of
 * 16 MiB per GPU instead of the stream-ordered pool, so it allocates
 * nothing; a batch needing more returns -7 before any launch.  Every call
 * on one GPU shares that array: calls that overlap on one GPU (another
 * thread, another stream) race, so serialise them. */
GPUD_API int gpud_stereo_pilot_present_noalloc(const gpud_float2* iq_dev, const int64_t* offsets,
                                               const int32_t* rates, int n_signals, int32_t* present, double* ratio,
                                               double* offset_hz, double* mags201, gpud_stream_handle stream);

/* ------------------------------------------------------------------ */
/* Surveillance: Welch PSD, FFT, CFAR and the detection list.           */
/* Successors of gpufft.Engine (internal/fft/gpufft/gpu.go:33-91), the  */
/* CPU WelchPSD (internal/fft/welch.go:15-66), internal/cfar and the    */
/* threshold path of detector.detectSignals (detector.go:408-713).      */
/* ------------------------------------------------------------------ */
typedef struct gpud_surv gpud_surv;

enum { GPUD_CFAR_CA = 0, GPUD_CFAR_GOSCA = 1, GPUD_CFAR_CASO = 2, GPUD_CFAR_OS = 3 /* os.go:31-80 */ };

typedef struct gpud_cfar_params {
    int32_t mode;          /* GPUD_CFAR_* (cfar/types.go:6-14) */
    int32_t guard;         /* guard cells per side */
    int32_t train;         /* training cells per side */
    int32_t wrap;          /* 1 = wrap-around indexing, 0 = clamp to edge */
    double  scale_db;      /* <= 0 -> 6 dB (cfar.go:13-15) */
    double  hysteresis_db; /* run closes below threshold - hysteresis (detector.go:445) */
    int32_t rank;          /* OS only: 1-based rank in the 2*train cells, clamped (os.go:14-22); train <= 128 */
    int32_t pad;
} gpud_cfar_params;

/* One threshold run of a frame (detector.go:436-480, before edge expansion). */
typedef struct gpud_surv_run {
    int32_t first_bin;
    int32_t last_bin;
    int32_t peak_bin;
    int32_t pad;
    double  peak_db;
    double  noise_db;      /* threshold[peak] - scale_db */
} gpud_surv_run;

/* detector.go:82-94 Signal fields of the detection list. */
typedef struct gpud_surv_signal {
    int32_t first_bin;
    int32_t last_bin;
    int32_t peak_bin;
    int32_t pad;
    double  center_hz;
    double  bw_hz;
    double  peak_db;
    double  snr_db;
    double  noise_db;
} gpud_surv_signal;

/* detector.go:101-261 Detector fields the detection list reads. */
typedef struct gpud_detector_params {
    double threshold_db;      /* used where a threshold is NaN */
    double hysteresis_db;
    double cfar_scale_db;
    double edge_margin_db;    /* <= 0 -> 3 */
    double max_signal_bw_hz;
    double merge_gap_hz;
    double center_hz;
} gpud_detector_params;

/* config.DetectorConfig fields the detection path reads (config.go:30-48). */
typedef struct gpud_detector_config {
    double  threshold_db;
    double  ema_alpha;        /* <= 0 or > 1 -> 0.2 */
    double  hysteresis_db;    /* <= 0 -> 3 */
    double  cfar_guard_hz;
    double  cfar_train_hz;
    double  cfar_scale_db;    /* <= 0 -> 6 */
    double  edge_margin_db;   /* <= 0 -> 3 */
    double  max_signal_bw_hz; /* <= 0 -> 150 kHz */
    double  merge_gap_hz;     /* <= 0 -> 5 kHz */
    int32_t cfar_mode;        /* GPUD_CFAR_* */
    int32_t cfar_rank;        /* OS; outside 1..2*train -> round(0.75 * 2 * train) */
    int32_t cfar_wrap;
    int32_t pad;
} gpud_detector_config;
/* detector.go:101-261 New(cfg, sampleRate, fftSize), the detection settings:
 * guard = max(0, ceil(guard_hz / binWidth)), train = max(1, ceil(train_hz /
 * binWidth)), binWidth = sample_rate / fft (fft outside (0, 4 Mi] -> 65536,
 * returned in *fft_bins), the rank and the defaults above.  Any output
 * pointer may be NULL.  Host code; -1 bad arguments. */
GPUD_API int gpud_detector_settings(const gpud_detector_config* cfg, int sample_rate, int fft_size,
                                    gpud_cfar_params* cfar, gpud_detector_params* det, double* ema_alpha,
                                    int* fft_bins);
/* detector.go:263-278: alpha = 1 - (1 - ema_alpha)^(dt / (1/15 s)) clamped to
 * [0.01, 0.99]; dt <= 0 or > 1 s (and the first frame: pass 0) -> 1/15 s. */
GPUD_API double gpud_detector_adaptive_alpha(double ema_alpha, double dt_seconds);

/* gpu.go:33-51 New(n): transforms at n_fft in {4096, 32768}.  Any other
 * power of two from 1024 to 32768 makes a detection-only engine (CFAR runs
 * and detection lists of spectra another engine computed, e.g. a decimated
 * level's: detector.New(cfg, level.SampleRate, level.FFTSize),
 * pipeline_runtime.go:791-817), whose transform entry points return -1.
 * -1 otherwise. */
GPUD_API int gpud_surv_create(int n_fft, int sample_rate, gpud_surv** out);
GPUD_API int gpud_surv_destroy(gpud_surv* s);
GPUD_API gpud_stream_handle gpud_surv_stream(gpud_surv* s);
/* welch.go:15-66 per output frame: dB, fftshifted, float32 [n_frames][n_fft]. */
GPUD_API int gpud_surv_welch_device(gpud_surv* s, const gpud_float2* iq_dev, int64_t frame_step, int seg_hop,
                                    int n_segs, int n_frames, float* psd_dev);
/* gpu.go:67-91 Exec: forward unnormalised C2C, natural order; window 1 =
 * Hann applied on the device. */
GPUD_API int gpud_surv_fft_device(gpud_surv* s, const gpud_float2* in_dev, int64_t frame_step, int n_frames,
                                  int window, gpud_float2* out_dev);
/* pipeline_runtime.go:315-341 spectrumFromIQWithPlan (GPU branch) +
 * fft.go:58-73 SpectrumFromFFT + pipeline_runtime.go:391-397 sanitizeSpectrum:
 * window (1 = Hann), FFT, 20 log10(|X|/N + 1e-12), fftshift, NaN/Inf ->
 * -200 dB; float32 [n_frames][n_fft].  (The Welch output is sanitised the
 * same way.) */
GPUD_API int gpud_surv_spectrum_device(gpud_surv* s, const gpud_float2* in_dev, int64_t frame_step, int n_frames,
                                       int window, float* db_dev);
/* The spectrum of a decimated surveillance level (the multi-resolution
 * derived level: plan pipeline_runtime.go:1274-1293, spectrum 624-645 with
 * decimateSurveillanceIQ 399-426): per frame the n_fft samples at in_dev +
 * f*frame_step are low-passed by LowpassFIR((sample_rate/decim)*0.5*0.8,
 * sample_rate, 101) from zero history (ApplyFIR), decimated by decim,
 * Hann-windowed over the level's n_fft/decim samples and transformed;
 * dB, fftshifted, NaN/Inf -> -200: float32 [n_frames][n_fft/decim] on the
 * device.  decim a power of two with n_fft/decim >= 256 (the reference
 * derives decim 2).  Queued on the engine's stream. */
GPUD_API int gpud_surv_level_spectrum_device(gpud_surv* s, const gpud_float2* in_dev, int64_t frame_step,
                                             int n_frames, int decim, float* db_dev);
/* detector.go:806-817 smoothSpectrum over n_frames consecutive frames (psd_dev
 * [n_frames][n_fft] dB in, out_dev the smoothed frames, float32): per bin
 * ema = alphas[f] * v + (1 - alphas[f]) * ema in FP64, the state ema_dev
 * (n_fft doubles, device) carried across calls; first = 1 for a detector
 * with no EMA yet (its first frame is copied, detector.go:807-810).  alphas
 * is host memory (gpud_detector_adaptive_alpha per frame).  Queued on the
 * engine's stream. */
GPUD_API int gpud_surv_smooth_device(gpud_surv* s, const float* psd_dev, int n_frames, const double* alphas,
                                     int first, double* ema_dev, float* out_dev);
/* ca.go / gosca.go / caso.go / os.go + detector.go:436-480 per frame. */
GPUD_API int gpud_surv_cfar_runs_device(gpud_surv* s, const float* psd_dev, int n_frames,
                                        const gpud_cfar_params* params, gpud_surv_run* runs_dev,
                                        int* run_counts_dev, int max_runs);
/* The same, with ambiguous_dev[f] = the number of frame f's bins whose CFAR
 * decision the reference's running sums (ca.go:61-66, slid along the frame
 * in FP64) could take the other way: bins whose dB value lies within the
 * bound of that summation's rounding (2 u (i + t) t lmax per side at bin i,
 * lmax the frame's largest linear cell), plus the rounding of this
 * library's prefix-sum window sums, of the threshold or of the threshold
 * less the hysteresis.  Zero means the frame's runs are those the reference
 * computes from the same spectrum.  OS-CFAR ranks without arithmetic: always
 * zero.  ambiguous_dev may be NULL.  The engine keeps a workspace of
 * n_frames * (n_fft / 4 + 4) bytes, grown on demand. */
GPUD_API int gpud_surv_cfar_runs_device_ex(gpud_surv* s, const float* psd_dev, int n_frames,
                                           const gpud_cfar_params* params, gpud_surv_run* runs_dev,
                                           int* run_counts_dev, int max_runs, int* ambiguous_dev);
/* detector.go:483-639 expandSignalEdges + mergeOverlapping +
 * powerWeightedCenter on one frame's runs (host memory).  Returns the signal
 * count (<= cap) or -1. */
GPUD_API int gpud_surv_finish_frame(const float* psd, int n_bins, int sample_rate, const gpud_surv_run* runs,
                                    int n_runs, const gpud_detector_params* params, gpud_surv_signal* out, int cap);
/* The same for n_frames frames on the device (one workgroup per frame):
 * psd_dev [n_frames][n_fft] and the CFAR runs (runs_dev [n_frames][max_runs],
 * run_counts_dev) of gpud_surv_cfar_runs_device in; signals_dev
 * [n_frames][cap_per_frame] and counts_dev[n_frames] (the signal count,
 * which may exceed cap_per_frame) out.  max_runs <= 2048.  Queued on the
 * engine's stream.  Bins identical to the host version; centres to the
 * last bits of pow(). */
GPUD_API int gpud_surv_finish_device(gpud_surv* s, const float* psd_dev, int n_frames, const gpud_surv_run* runs_dev,
                                     const int* run_counts_dev, int max_runs, const gpud_detector_params* params,
                                     gpud_surv_signal* signals_dev, int cap_per_frame, int* counts_dev);
/* Whole detection list for n_frames frames of device-resident IQ: Welch,
 * CFAR, runs, expansion, merge and centroids, all on the device; only the
 * lists (and the spectra if asked) come back.  signals: n_frames *
 * cap_per_frame records; counts per frame.  psd_host (optional, n_frames *
 * n_fft floats) receives the spectra. */
GPUD_API int gpud_surv_detect(gpud_surv* s, const gpud_float2* iq_dev, int64_t frame_step, int seg_hop, int n_segs,
                              int n_frames, const gpud_cfar_params* cfar, const gpud_detector_params* det,
                              gpud_surv_signal* signals, int cap_per_frame, int* counts, float* psd_host);

#ifdef __cplusplus
}
#endif
#endif /* ARGUS_GPU_H */
