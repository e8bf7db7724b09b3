/*
 * oracle.h -- CPU restatement of the reference's per-signal extraction,
 * demodulation, surveillance (Welch/CFAR/detection list), WFM stereo and
 * RDS paths.
 *
 * THIS IS TEST INFRASTRUCTURE.  It is the checker the HIP library is compared
 * against, and the CPU baseline the benchmark times beside the GPU.  Nothing
 * in the product (libargus_hip.so, the Python binding, the Go binding) links,
 * loads or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it.
 *
 * Every function cites the reference file:line whose semantics it restates
 * (paths relative to the reference repository root).  Compile with
 * -ffp-contract=off so float32 products and sums round exactly like the
 * reference Go code on amd64 (Go does not fuse a*b+c there).
 *
 * Parity pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_build_polyphase      known-answer vector of polyphase_test.go:5-22
 *   orc_exact_decim          cpu_oracle_test.go:60-67 known answers
 *   orc_lowpass_fir          closed form + symmetry + DC gain ~1
 *   orc_host_core            direct form == polyphase; monolithic == chunked
 *                            (cpu_oracle_test.go:39-89); FP64 brute force
 *   orc_runner_*             lifecycle sequences of streaming_gpu_exec_test.go
 *   orc_config_hash          FNV-1a-64 published test vectors + format
 *   demods                   closed forms (tone in -> constant out)
 *   fft / welch              brute-force DFT, Parseval, tone bin position
 *   cfar                     cfar_test.go:65-94 closed-form thresholds
 */
#ifndef ARGUS_ORACLE_H
#define ARGUS_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Job/result layouts are identical to gpud_streaming_job/_result in
 * include/argus_gpu.h so the same host structs feed both. */
typedef struct orc_job {
    int64_t  signal_id;
    double   offset_hz;
    double   bandwidth;
    int32_t  out_rate;
    int32_t  num_taps;
    uint64_t config_hash;
} orc_job;

typedef struct orc_result {
    int64_t signal_id;
    int32_t rate;
    int32_t n_out;
    int32_t phase_count;
    int32_t history_len;
    int64_t out_offset;
    double  nco_phase;
} orc_result;

/* ---- dsp primitives (internal/dsp) ---- */
int  orc_lowpass_fir(double cutoff_hz, int sample_rate, int taps, double* out); /* returns tap count (odd) */
int  orc_lowpass_fir_hamming(double cutoff_hz, int sample_rate, int taps, double* out);
void orc_apply_fir(const float* iq, int n, const double* taps, int nt, float* out);
int  orc_decimate(const float* iq, int n, int factor, float* out);
void orc_freq_shift(const float* iq, int n, int sample_rate, double offset_hz, float* out);
void orc_stateful_fir_real(const double* taps, int nt, double* delay, const float* x, int n, float* out);
int  orc_stateful_decim_fir_complex(const double* taps, int nt, int factor, double* delay_r, double* delay_i,
                                    int* phase_io, const float* iq, int n, float* out);

/* ---- streaming extractor (internal/demod/gpudemod) ---- */
int  orc_polyphase_phase_len(int base_len, int decim);
int  orc_build_polyphase(const float* base, int n, int decim, float* out); /* returns decim*L */
int  orc_exact_decim(int sample_rate, int out_rate); /* -1 on error */
uint64_t orc_config_hash(int64_t signal_id, double offset_hz, double bandwidth, int out_rate, int num_taps, int sample_rate);
int  orc_host_core(const float* iq_new, int n_new, int sample_rate, double offset_hz,
                   double* nco_phase_io, int* phase_count_io, int num_taps, int decim,
                   float* history_io, int* history_len_io,
                   const float* poly_taps, int poly_len, float* out, int out_cap);
int  orc_extract_direct(const float* iq_new, int n_new, double phase_inc, double* nco_phase_io,
                        int* phase_count_io, int num_taps, int decim, const float* base_taps,
                        float* history_io, int* history_len_io, float* out, int out_cap);
int  orc_extract_polyphase(const float* iq_new, int n_new, double phase_inc, double* nco_phase_io,
                           int* phase_count_io, int num_taps, int decim, const float* base_taps,
                           const float* poly_taps, float* history_io, int* history_len_io,
                           float* out, int out_cap);

typedef struct orc_runner orc_runner;
orc_runner* orc_runner_new(int sample_rate);
void orc_runner_free(orc_runner* r);
int  orc_runner_streaming_extract(orc_runner* r, const float* iq, int n, const orc_job* jobs, int n_jobs,
                                  float* out, int64_t out_cap, orc_result* results);
int  orc_runner_num_states(const orc_runner* r);
int  orc_runner_state_ids(const orc_runner* r, int64_t* ids, int cap);
int  orc_runner_get_state(const orc_runner* r, int64_t signal_id, uint64_t* config_hash, double* nco_phase,
                          int* decim, int* phase_count, int* num_taps, float* history, int hist_cap);
void orc_runner_reset_signal(orc_runner* r, int64_t signal_id);
void orc_runner_reset_all(orc_runner* r);
/* Per-channel demod state for the streamed demod epilogue. */
int  orc_runner_set_threads(orc_runner* r, int threads);

/* ---- demodulators (internal/demod) ---- */
int  orc_fm_discrim(const float* iq, int n, float* out);               /* n-1 outputs */
int  orc_fm_discrim_bridged(const float* iq, int n, float* last_iq_io, int* have_last_io, float* out);
int  orc_am_demod(const float* iq, int n, float* out);
int  orc_ssb_demod(const float* iq, int n, int sample_rate, int usb, double* bfo_phase_io, float* out);
int  orc_cw_demod(const float* iq, int n, int sample_rate, double* bfo_phase_io, float* out);
int  orc_rds_baseband(const float* iq, int n, int sample_rate, float* out, int out_cap, int* out_rate);

/* ---- post-demodulation audio (internal/dsp/resample.go, streamer.go:1408-1470) ---- */
typedef struct orc_resampler {
    int l, m, taps, out_time;
    double* bank;   /* l x taps */
    double* delay;  /* taps, newest first */
} orc_resampler;
int    orc_gcd(int a, int b);
void   orc_windowed_sinc(int length, double fc, double gain, double* out);
int    orc_resampler_init(orc_resampler* r, int in_rate, int out_rate, int taps_per_phase);
void   orc_resampler_free(orc_resampler* r);
int    orc_resampler_process(orc_resampler* r, const float* in, int n, float* out);
double orc_deemphasis(float* x, int n, int stride, double alpha, double y);

#ifdef __cplusplus
}
#endif
#endif
