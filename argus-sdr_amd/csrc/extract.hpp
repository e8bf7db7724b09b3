// extract.hpp -- fused NCO + polyphase FIR/decimate work items.
//
// Two shapes of work item, picked per channel by the host:
//
//  dense tile (D < T): a tile of M consecutive outputs.  The NCO-shifted
//    input span the tile needs is staged in LDS grouped by polyphase residue
//    ("planes"): plane p holds s[e_m0 - p + j*D], j = -(L-1) .. M-1, so for
//    a fixed (p, k) consecutive outputs read consecutive LDS words -- no bank
//    conflicts whatever D is.  Each output is
//        y[m] = sum_{p < P} sum_{k < L_p} h[p + k*D] * s[e_m - p - k*D]
//    (the reference's phase-major order, streaming_host_core.go:29-52).
//    When M < 256 the planes of one output are split over S lanes and the
//    partial sums are combined with lane shuffles, so every thread works.
//
//  narrow group (D >= T): each output needs only the T samples before its
//    emission point and windows do not overlap, so a wavefront computes one
//    output at a time: lanes take taps k = lane, lane+64, ..., shift their
//    sample with the NCO straight from global memory (coalesced) and the
//    wave reduces.  No LDS staging, no redundant rotations.
//
// A third kind writes the channel's new shifted history (the newest `keep`
// samples of history ++ shifted input) into the other half of a ping-pong
// pair, so no item reads history another item is overwriting.
#pragma once
#include "common.hpp"

struct ArgusChanDesc {
    const float2* hist_rd;  // old shifted history, H samples, oldest first
    float2* hist_wr;        // new shifted history, keep samples
    const float* taps;      // phase-major polyphase taps (D planes of L floats)
    float2* out;            // n_out baseband outputs of this chunk
    uint64_t phase0;        // NCO phase of new sample 0 (turns, 2^64 == 1)
    uint64_t inc;           // NCO increment per sample (turns, 2^64 == 1)
    int D, T, L, P;         // decimation, taps, phase length ceil(T/D), planes min(D,T)
    int H;                  // old history length
    int e0;                 // new-sample index of output 0: D - phase_count - 1
    int n_out;              // outputs this chunk
    int keep;               // new history length
    int M;                  // outputs per item
    int PS;                 // LDS plane stride (complex samples), dense only
    int S;                  // lanes per output in the dense FIR phase (power of 2)
    int narrow;             // 1 = narrow (wave-per-output) items
};

struct ArgusWorkItem {
    int chan;
    int m0;    // first output of the item
    int kind;  // 0 = dense tile, 1 = history tail, 2 = narrow group
    int pad;   // host: cost estimate for ordering
};

enum { ARGUS_ITEM_DENSE = 0, ARGUS_ITEM_HISTORY = 1, ARGUS_ITEM_NARROW = 2, ARGUS_ITEM_NOP = 3 };

// Number of XCDs workgroups are dealt to round-robin (MI355X: 8).
#define ARGUS_XCDS 8

#define ARGUS_EXTRACT_THREADS 256
#define ARGUS_LDS_ENTRIES 6144     // complex samples of LDS per dense tile (48 KiB)
#define ARGUS_MAX_TAPS 4095
#define ARGUS_NARROW_OUTPUTS 32    // outputs per narrow item (8 per wave)

// Host-side planning for one channel.  Fills M, PS, S, narrow.  Returns 0,
// or -1 if the geometry cannot be staged.
static inline int argus_plan_chan(ArgusChanDesc* ch)
{
    const int D = ch->D, T = ch->T, L = ch->L, P = ch->P;
    if (D >= T) {
        ch->narrow = 1;
        ch->M = ARGUS_NARROW_OUTPUTS;
        ch->PS = 0;
        ch->S = 1;
        return 0;
    }
    ch->narrow = 0;
    int M = ARGUS_LDS_ENTRIES / P - (L - 1);
    if (M > ARGUS_EXTRACT_THREADS) M = ARGUS_EXTRACT_THREADS;
    if (ch->n_out > 0 && M > ch->n_out) M = ch->n_out;
    while (M > 0 && P * ((M + L - 1) | 1) > ARGUS_LDS_ENTRIES) M--;
    if (M <= 0) return -1;
    ch->M = M;
    ch->PS = (M + L - 1) | 1;  // odd stride spreads planes over banks
    int S = 1;
    while (S * 2 * M <= ARGUS_EXTRACT_THREADS && S * 2 <= P && S < 64) S *= 2;
    ch->S = S;
    return 0;
}

// LDS bytes a dense tile of this channel needs (samples + taps).
static inline size_t argus_dense_lds_bytes(const ArgusChanDesc& ch)
{
    return (size_t)ch.P * ch.PS * 8 + (size_t)ch.P * ch.L * 4;
}
