// audio.hpp -- post-demodulation audio stage of the streaming service
// (internal/recorder/streamer.go:1382-1470 with internal/dsp/resample.go):
//
//   plain WFM   15 kHz stateful low-pass of the discriminator output
//               (wfmAudioFilter, LowpassFIR(15000, rate, 101));
//   all modes   rational polyphase resampling to the audio rate
//               (Resampler, L/M = out/in over their gcd, 32 taps per phase
//               -> 33-tap arms of a Kaiser-windowed sinc), two resamplers
//               for stereo, none when the rates are equal;
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
//
// One filter per channel.  The low-pass and the resampler are both linear
// and time-invariant per phase, so for plain WFM the host composes them: arm
// p of the composite is the 33-tap arm convolved with the 101 low-pass taps
// (133 taps, in FP64, rounded once), applied to the discriminator output
// directly.  Equal rates use arm [1] (a copy) or the low-pass taps alone,
// with D_k = k + 1.  The kernel therefore only evaluates
//   y_k = sum_{i < T} arm[p_k][i] * x[D_k - 1 - i],   T in {1, 33, 101, 133},
// over a delay line of the last T inputs carried between blocks.
#pragma once
#include "common.hpp"

#define ARGUS_AUDIO_ARM 33          // taps per resampler arm (32 per phase, odd prototype)
#define ARGUS_AUDIO_LPF_TAPS 101    // plain-WFM low-pass
#define ARGUS_AUDIO_TMAX (ARGUS_AUDIO_ARM + ARGUS_AUDIO_LPF_TAPS - 1)  // composite arm
#define ARGUS_AUDIO_THREADS 256
#define ARGUS_AUDIO_XCAP 4096       // staged input samples per chunk (all sides), LDS
#define ARGUS_AUDIO_ARMCAP 3328     // arm coefficients held in LDS (L x T), else read from memory
#define ARGUS_AUDIO_OUTCAP 2048     // output samples per chunk (all sides), LDS
#define ARGUS_AUDIO_MAX_M (1 << 20) // phase arithmetic is 32-bit within a chunk
// workgroups of the audio launch: one per channel up to this many (those past
// the block's channel count exit at once), the hardware balancing their lengths
#define ARGUS_AUDIO_GRID 8192

struct ArgusAudioDesc {
    const float* in;         // block input: n_in mono samples or n_in L/R frames
    float* out;              // n_out frames x channels, interleaved
    const float* arms;       // L x T filter (see above)
    const float* hist_rd;    // channels x ARGUS_AUDIO_TMAX: the last T inputs, oldest first
    float* hist_wr;
    double* deemph;          // channels FP64 de-emphasis states (read, then written)
    double alpha;            // de-emphasis pole (deemph_on)
    int L, M, t0, T;         // ratio, phase at the block start, taps per arm
    int ident;               // equal rates: D_k = k + 1, arm 0
    int n_in, n_out, channels;
    int deemph_on;           // WFM with a de-emphasis time constant
    float gain;              // 0.35 for WFM, else 1
};

struct ArgusAudioHdr {
    const ArgusAudioDesc* descs;
    int n_chans;
    int pad;
};

// The filter of one (input rate, output rate, plain-WFM low-pass) stage:
// L, M, T and, if arms is not null, the L x T coefficients (FP32).
// Returns -1 for rates the stage does not take (non-positive, M >= 2^20).
int argus_audio_filter(int in_rate, int out_rate, bool lowpass, int* L, int* M, int* T, float* arms);
// One block of n inputs from phase t0: output count and next phase.
void argus_audio_block(int L, int M, int t0, int n, int* n_out, int* t_next);

int argus_launch_audio_hdr(const ArgusAudioHdr* d_hdr, hipStream_t stream);
