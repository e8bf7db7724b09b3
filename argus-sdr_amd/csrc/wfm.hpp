// wfm.hpp -- WFM multiplex stage of the streaming service: stateful stereo
// decoding (internal/recorder/streamer.go:1525-1643 stereoDecodeStateful with
// the lock hysteresis of 1367-1380 and the mono fallback filter of
// 1511-1520) and the streaming RDS baseband (internal/demod/fm.go:234-271
// carried across blocks), run on the FM discriminator output of a channel.
//
// Three launches per block, each over every stereo channel:
//   pre  (tiles of ARGUS_WFM_TILE samples): the three 101-tap real FIRs that
//        feed the PLL and the matrix (L+R low-pass, L-R band-pass, pilot
//        band-pass; each band-pass is one FIR whose taps are the difference
//        of two low-pass designs) and the RDS chain (mix by +57 kHz, 101-tap
//        complex low-pass, keep every rds_decim-th sample);
//   pll  (one wavefront per channel): the type-II pilot PLL, a serial
//        recurrence; stores 2*phase per sample (as a turn fraction) and
//        decides lock, hysteresis and the block's output form;
//   post (tiles): L-R product and its 101-tap low-pass, then the L/R matrix,
//        or the mono fallback low-pass when stereo is not enabled.
// Filter delay lines are carried as the last ARGUS_WFM_HIST inputs of each
// filter in ping-pong pairs; a tail item per channel writes the new ones.
//
// Batch mode (ArgusWfmDesc::batch != 0) is the one-shot decoder of a whole
// recording (internal/recorder/wfm_batch_stereo.go:76-157): zero delay lines,
// a fresh PLL, no tail items, the lock test pilot power / total power >
// 0.003 (no hysteresis), and always the L/R matrix output -- or, with
// ARGUS_WFM_BATCH_FALLBACK, the raw discriminator duplicated on both channels
// when the block does not lock (demodWFMStereoBatchAudio, 42-49).
#pragma once
#include "common.hpp"

#define ARGUS_WFM_TAPS 101
#define ARGUS_WFM_HIST (ARGUS_WFM_TAPS - 1)
#define ARGUS_WFM_TILE 1024

enum { ARGUS_WFM_ITEM_TILE = 0, ARGUS_WFM_ITEM_TAIL = 1 };

// Index of each tap set in ArgusWfmDesc::taps (ARGUS_WFM_TAPS floats each).
enum { ARGUS_WFM_LPF = 0, ARGUS_WFM_BPF = 1, ARGUS_WFM_PILOT = 2, ARGUS_WFM_RDS = 3, ARGUS_WFM_NTAPSETS = 4 };

// Carried PLL and lock state of one channel (device-resident).
struct ArgusWfmState {
    double phase, freq, i_state, q_state, err_avg;
    int on_count, off_count, enabled, fb_sel;
};

// What the PLL decided for this block (device-written, read by post and
// downloaded with the outputs).
struct ArgusWfmBlock {
    int locked;   // lock test of this block
    int stereo;   // 1: L/R matrix output; 0: mono fallback output
    int fb_rd;    // fallback delay line read this block
    int fb_wr;    // fallback delay line written this block (fallback only)
};

struct ArgusWfmDesc {
    const float* mono;         // n discriminator outputs of this block
    float* audio;              // 2n: L, R interleaved
    float* scratch;            // 4n: lpr, bpf, pilot, t2 (frac(phase / pi))
    float2* rds;               // n_rds RDS baseband samples
    const float* mono_hist_rd; // last ARGUS_WFM_HIST inputs of the stereo FIRs
    float* mono_hist_wr;
    const float* lr_hist_rd;   // last ARGUS_WFM_HIST inputs of the L-R low-pass
    float* lr_hist_wr;
    float* fb_hist;            // 2 x ARGUS_WFM_HIST: fallback low-pass delay lines
    const float* taps;         // ARGUS_WFM_NTAPSETS x ARGUS_WFM_TAPS
    ArgusWfmState* st;
    ArgusWfmBlock* blk;
    uint64_t mix0;             // RDS mixer phase of sample 0, pre-increment applied (turns, 2^64 == 1)
    uint64_t mix_inc;          // RDS mixer increment per sample (turns)
    double alpha, beta, lpa, min_f, max_f, rate;
    int n;
    int rds_first;             // first sample index of this block the RDS decimator keeps
    int rds_decim;
    int n_rds;
    int batch;                 // 0: streaming; ARGUS_WFM_BATCH_MATRIX or ARGUS_WFM_BATCH_FALLBACK
    int pad;
};

enum { ARGUS_WFM_STREAM = 0, ARGUS_WFM_BATCH_MATRIX = 1, ARGUS_WFM_BATCH_FALLBACK = 2 };

struct ArgusWfmItem {
    int chan;
    int start;
    int kind;
    int pad;
};

struct ArgusWfmHdr {
    const ArgusWfmDesc* descs;
    const ArgusWfmItem* items;
    int n_chans;
    int n_items;
};

#define ARGUS_WFM_GRID 256

// Host: the four tap sets at `rate` (ARGUS_WFM_NTAPSETS x ARGUS_WFM_TAPS
// floats) and the PLL loop constants and clamps of a descriptor.
void argus_wfm_design_taps(int rate, float* taps);
void argus_wfm_loop_constants(int rate, ArgusWfmDesc& wd);

int argus_launch_wfm(const ArgusWfmDesc* d_desc, int n_chans, const ArgusWfmItem* d_items, int n_items,
                     hipStream_t stream);
// Graph-mode launch of the three stages: counts and pointers from *hdr.
int argus_launch_wfm_hdr(const ArgusWfmHdr* d_hdr, hipStream_t stream);
