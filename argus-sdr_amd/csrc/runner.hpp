// runner.hpp -- host-side state of the streaming batch runner.
#pragma once
#include "extract.hpp"
#include "fastconv.hpp"
#include <vector>

// Per-signal carried state (streaming_types.go:26-44 ExtractStreamState and
// streaming_gpu_native_state.go:5-28 nativeStreamingSignalState in one).
struct ArgusSignal {
    int64_t signal_id = 0;
    uint64_t config_hash = 0;
    bool initialized = false;
    bool active = false;
    double nco_phase = 0.0;   // radians, [-pi, pi)
    int decim = 0;
    int phase_count = 0;
    int num_taps = 0;
    int window = 0;           // 0 = Hann (LowpassFIR), 1 = Hamming
    std::vector<float> base;  // designed taps (f32)
    std::vector<float> poly;  // phase-major polyphase layout of the first num_taps taps
    double taps_cutoff = 0;
    int taps_num = 0, taps_decim = 0, taps_rate = 0, taps_window = -1;
    bool taps_dirty = false;
    float* d_taps = nullptr;
    size_t d_taps_len = 0;
    float2* d_ctaps = nullptr;  // modulated taps for the NCO increment ctaps_inc (T)
    size_t d_ctaps_len = 0;
    uint64_t ctaps_inc = 0;
    bool ctaps_valid = false;
    float2* d_hist[2] = {nullptr, nullptr};  // ping-pong shifted history, cap num_taps-1
    int hist_sel = 0;
    int hist_len = 0;
    float2* d_out = nullptr;   // per-signal output (plain per-signal form only)
    int out_cap = 0;
    float2* d_fastG = nullptr; // fast form: FFT_4096 of the modulated taps / 4096 (eligible signals)
    uint64_t fastG_inc = 0;
    bool fastG_valid = false;
};

// One job of a legacy block call (engine.hip): taps in natural order
// (float32), decimation, NCO start and increment (radians), outputs kept
// (input m*D feeds output m; at most ceil(n / D)), device output.
struct ArgusLegacyJob {
    const float* taps;
    int T;
    int D;
    double phase_start;
    double phase_inc;
    int n_out;
    float2* out;
};

struct ArgusLegacyScratch;
void argus_legacy_release(ArgusLegacyScratch* sc);
// Queue one fused launch of the jobs over d_in[0..n) on st (engine.hip).
// The staging in *sc is reused: synchronise st before the next call on it.
int argus_legacy_launch(ArgusLegacyScratch*& sc, hipStream_t st, const ArgusTune& tune, const float2* d_in, int n,
                        const std::vector<ArgusLegacyJob>& jobs);
struct gpud_out_rings;

struct gpud_runner {
    ArgusTune tune = argus_default_tune();
    int max_samples = 0;
    int sample_rate = 0;
    int exec_mode = GPUD_EXEC_BATCHED;
    int window = 0;  // tap window for signals: 0 = Hann, 1 = Hamming
    hipStream_t stream = nullptr;
    std::vector<ArgusSignal> signals;
    float2* d_in = nullptr;
    float2* h_in = nullptr;
    float2* d_out = nullptr;
    size_t d_out_cap = 0;
    float2* h_out = nullptr;
    size_t h_out_cap = 0;
    // double-buffered staging of the channel table and work list
    int stage_slot = 0;
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    bool stage_recorded[2] = {false, false};
    ArgusChanDesc* h_chans[2] = {nullptr, nullptr};
    ArgusChanDesc* d_chans[2] = {nullptr, nullptr};
    size_t h_chans_cap[2] = {0, 0}, d_chans_cap[2] = {0, 0};
    ArgusWorkItem* h_items[2] = {nullptr, nullptr};
    ArgusWorkItem* d_items[2] = {nullptr, nullptr};
    size_t h_items_cap[2] = {0, 0}, d_items_cap[2] = {0, 0};
    std::vector<ArgusWorkItem> items;
    // kernel timing (event pair per fused launch)
    bool timing = false;
    std::vector<hipEvent_t> tev;  // 2 per launch
    int tev_used = 0;
    double t_ms = 0.0;
    int t_launches = 0;
    std::vector<gpud_float2> hist_host;
    ArgusLegacyScratch* legacy = nullptr;  // staging of the legacy block path (engine.hip)
    // fast-convolution form of the batched launch (fastconv.hip): the
    // stream's last ARGUS_FAST_TAIL raw samples (tail_fill of them known to
    // be the stream's), the staged segments and their spectra, and per
    // staging slot the fast descriptors and the two headers the kernels read
    bool fast = true;
    int tail_fill = 0;
    float2* d_ftail = nullptr;
    float2* d_fu = nullptr;
    size_t fu_cap = 0;
    float2* d_fX = nullptr;
    size_t fX_cap = 0;
    ArgusFastDesc* h_fdesc[2] = {nullptr, nullptr};
    ArgusFastDesc* d_fdesc[2] = {nullptr, nullptr};
    size_t h_fdesc_cap[2] = {0, 0}, d_fdesc_cap[2] = {0, 0};
    struct FastHdrs {
        ArgusFastHdr f;
        ArgusExtractHdr x;
    };
    FastHdrs* h_fhdr[2] = {nullptr, nullptr};
    FastHdrs* d_fhdr[2] = {nullptr, nullptr};
    std::vector<ArgusFastDesc> fdescs;
    gpud_out_rings* rings = nullptr;       // per-signal output rings (outring.hip), on first ring call
};

int argus_lowpass_fir(double cutoff_hz, int sample_rate, int taps, double* out, bool hamming);
ArgusSignal* argus_runner_find(gpud_runner* r, int64_t id);
void argus_signal_release(ArgusSignal& s);
