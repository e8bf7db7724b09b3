// predemod.hpp -- the stateful decimating FIR between extraction and
// demodulation (the pre-demod stage of the live recorder,
// internal/recorder/streamer.go:1098-1161 with
// dsp.StatefulDecimatingFIRComplex, internal/dsp/decimating_fir.go:51-95):
// y_j = sum_k h[k] x[i_j - k] with FP64 taps and accumulation, emitted at
// i_j = first + j * factor, over the channel's baseband with its last T-1
// inputs carried between blocks.
#pragma once
#include "common.hpp"

struct ArgusPreDesc {
    const float2* in;       // the block's baseband at the channel rate, n_in samples
    float2* out;            // n_out outputs
    const double* taps;     // T FP64 taps; taps[0] weighs the newest sample
    const float2* hist_rd;  // the T-1 inputs before the block, oldest first (zeros at the start)
    float2* hist_wr;        // the T-1 inputs that end the block
    int n_in;
    int n_out;
    int first;              // block index of output 0
    int factor;
    int T;
    int pad;
};

struct ArgusPreItem {
    int chan;
    int start;   // first output (kind 0)
    int count;
    int kind;    // 0 = outputs, 1 = history
};

struct ArgusPreHdr {
    const ArgusPreDesc* descs;
    const ArgusPreItem* items;
    int n_items;
    int pad;
};

#define ARGUS_PRE_TILE 256
#define ARGUS_PRE_TAPS 101
#define ARGUS_PRE_GRID 1024

// Outputs of a block of n inputs and the phase after it, for a stage whose
// phase p counts the inputs until the next output (0: the next input emits;
// decimating_fir.go:81-92).
static inline void argus_pre_block(int p, int factor, int n, int* n_out, int* p_next)
{
    const int m = n > p ? (n - 1 - p) / factor + 1 : 0;
    if (n_out) *n_out = m;
    if (p_next) *p_next = m > 0 ? factor - 1 - (n - 1 - (p + (m - 1) * factor)) : p - n;
}

// Graph-mode launch: counts and pointers come from *hdr (device memory).
int argus_launch_predemod_hdr(const ArgusPreHdr* d_hdr, hipStream_t stream);
