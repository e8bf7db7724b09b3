// audio.hpp -- post-demodulation audio stage of the streaming service
// (internal/recorder/streamer.go:1380-1470 with internal/dsp/resample.go):
//
//   plain WFM   15 kHz stateful low-pass of the discriminator output
//               (wfmAudioFilter, LowpassFIR(15000, rate, 101));
//   all modes   rational polyphase resampling to the audio rate
//               (Resampler, L/M = out/in over their gcd, 32 taps per phase
//               -> 33-tap arms of a Kaiser-windowed sinc), two resamplers
//               for stereo;
//   WFM         de-emphasis y = a y + (1 - a) x (FP64 state) and gain 0.35.
//
// Resampler semantics per call (block) of n inputs, from the carried phase
// t0: output k consumes D_k = floor((t0 + k M) / L) samples of the block and
// uses arm t0 + k M - L D_k over the last 33 samples; with k_n the first k
// whose D_k reaches n, the call emits k_n + 1 outputs if D_{k_n} == n, else
// k_n and the next call starts one phase step further on
// (resample.go:124-134 decrements the phase before finding the input
// exhausted).  The host computes the counts and phases; the kernel computes
// every output independently.
#pragma once
#include "common.hpp"

#define ARGUS_AUDIO_ARM 33          // taps per arm (32 per phase, odd prototype)
#define ARGUS_AUDIO_LPF_TAPS 101
#define ARGUS_AUDIO_LPF_HIST (ARGUS_AUDIO_LPF_TAPS - 1)
#define ARGUS_AUDIO_LPF_TILE 1024
#define ARGUS_AUDIO_CHUNK 2048      // output frames per scan chunk (LDS)

struct ArgusAudioDesc {
    const float* in;         // block input: n_in mono samples or n_in L/R frames
    float* out;              // n_out frames x channels, interleaved
    const double* arms;      // L x ARGUS_AUDIO_ARM
    const float* hist_rd;    // channels x ARGUS_AUDIO_ARM previous inputs (oldest first)
    float* hist_wr;
    double* deemph;          // channels FP64 de-emphasis states (read, then written)
    // 15 kHz low-pass (plain WFM): lpf_in -> in, delay lines ping-ponged
    const float* lpf_in;
    float* lpf_out;          // == in when lpf
    const float* lpf_taps;   // ARGUS_AUDIO_LPF_TAPS
    const float* lpf_hist_rd;  // ARGUS_AUDIO_LPF_HIST previous inputs
    float* lpf_hist_wr;
    double alpha;            // de-emphasis pole (deemph_on)
    int L, M, t0;            // ratio and phase at the block start (L == M: copy)
    int n_in, n_out, channels;
    int deemph_on;           // WFM with a de-emphasis time constant
    float gain;              // 0.35 for WFM, else 1
    int lpf;                 // 1: in is the low-pass of lpf_in
    int pad;
};

struct ArgusAudioItem {  // low-pass tiles: chan, first sample (or tail)
    int chan;
    int start;
    int kind;  // 0 tile, 1 tail (writes the next delay line)
    int pad;
};

struct ArgusAudioHdr {
    const ArgusAudioDesc* descs;
    const ArgusAudioItem* items;
    int n_chans;
    int n_items;
};

#define ARGUS_AUDIO_GRID 1024

// Host: the resampler arms for in_rate -> out_rate (FP64, L x ARGUS_AUDIO_ARM)
// and L, M; the block's output count and next phase.
int argus_audio_arms(int in_rate, int out_rate, int* L, int* M, double* arms /* may be null */);
void argus_audio_block(int L, int M, int t0, int n, int* n_out, int* t_next);

int argus_launch_audio_hdr(const ArgusAudioHdr* d_hdr, bool lpf, hipStream_t stream);
