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
//    With register blocking (R = 4) a thread computes R consecutive outputs
//    from a sliding window of R samples, so each LDS read feeds R complex
//    MACs; each plane is then stored as R interleaved sub-planes (position x
//    at sub-plane x mod R, index x div R) so the lanes of a read still touch
//    consecutive words.
//
//  narrow group (D >= T): each output needs only the T samples before its
//    emission point and windows do not overlap, so no input is staged.
//    With modulated taps (ctaps, g[k] = h[k] e^{-j 2pi inc k}) the raw input
//    is filtered directly, since
//        y[m] = sum_k h[k] s[e - k] = e^{j 2pi (phase0 + inc e)} sum_k g[k] x[e - k]
//    for s[n] = x[n] e^{j 2pi (phase0 + inc n)}: one rotation per output
//    instead of one per tap.  A wavefront computes four outputs, one per
//    16-lane row (lanes over taps, coalesced loads, row reduction by DPP).
//    Without ctaps (the single-signal entry point, whose caller supplies only
//    the polyphase taps) a wavefront computes one output and rotates every
//    sample.  Also the fallback when a dense tile does not fit the LDS budget
//    (very large D with T > D).
//
// A third kind writes the channel's new shifted history (the newest `keep`
// samples of history ++ shifted input) into the other half of a ping-pong
// pair, so no item reads history another item is overwriting.
#pragma once
#include "common.hpp"
#include <vector>

struct ArgusChanDesc {
    const float2* hist_rd;  // old shifted history, H samples, oldest first
    float2* hist_wr;        // new shifted history, keep samples
    const float* taps;      // phase-major polyphase taps (D planes of L floats)
    const float2* ctaps;    // modulated taps, natural order (T), or null
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
    int SP;                 // LDS sub-plane stride (complex samples), dense only
    int R;                  // outputs per thread in the dense FIR phase (1 or 4)
    int S;                  // lanes per output group in the dense FIR phase (power of 2)
    int narrow;             // 1 = narrow (wave-per-output) items
};

// Everything the graph-mode extraction kernel reads for one block (lives in
// device memory so a captured launch can be replayed with new contents).
struct ArgusWorkItem;
struct ArgusExtractHdr {
    const ArgusChanDesc* chans;
    const ArgusWorkItem* items;
    const float2* in;
    int n_items;
    int n_new;
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
#define ARGUS_LDS_ENTRIES 6144     // default complex samples of LDS per dense tile (48 KiB)
#define ARGUS_MAX_LDS_ENTRIES 16384
#define ARGUS_MAX_TAPS 4095
#define ARGUS_NARROW_OUTPUTS 32    // default outputs per narrow item (8 per wave)
#define ARGUS_GRID 2048            // default persistent workgroups (a multiple of ARGUS_XCDS)

// Tile and launch shape of the fused extraction launch (picked by the
// autotuner, gpud_*_autotune).
struct ArgusTune {
    int lds_entries;     // complex samples of LDS a dense tile may stage
    int narrow_outputs;  // outputs per narrow item
    int grid;            // persistent workgroups (multiple of ARGUS_XCDS)
};

static inline ArgusTune argus_default_tune() { return ArgusTune{ARGUS_LDS_ENTRIES, ARGUS_NARROW_OUTPUTS, ARGUS_GRID}; }

#define ARGUS_DENSE_R 4             // register blocking of the dense FIR

// Plane stride: at least R*SP, and congruent mod 16 (complex words) to 16/S
// so the 16 lanes of one LDS read phase (S planes x 16/S output groups)
// touch distinct bank pairs.
static inline int argus_plane_stride(int R, int SP, int S)
{
    const int want = S >= 16 || S == 1 ? 1 : 16 / S;
    int PS = R * SP;
    while ((PS & 15) != want) PS++;
    return PS;
}

// Dense geometry with R outputs per thread within `entries` complex words of
// LDS.  Returns 0, or -1 if not even one output group fits.
static inline int argus_plan_dense(ArgusChanDesc* ch, int R, int entries)
{
    const int P = ch->P, L = ch->L;
    const int tail = (L + R - 2) / R;  // sub-plane words beyond the G groups
    int G = (entries / P) / R - tail;
    const int need = ch->n_out > 0 ? (ch->n_out + R - 1) / R : 1;
    if (G > ARGUS_EXTRACT_THREADS) G = ARGUS_EXTRACT_THREADS;
    if (G > need) G = need;
    for (; G > 0; G--) {
        int S = 1;
        while (S * 2 * G <= ARGUS_EXTRACT_THREADS && S * 2 <= P && S < 64) S *= 2;
        int PS = argus_plane_stride(R, G + tail, S);
        if (P * PS > entries) PS = R * (G + tail);  // no room for bank padding
        if (P * PS <= entries) {
            ch->narrow = 0;
            ch->R = R;
            ch->M = G * R;
            ch->SP = G + tail;
            ch->PS = PS;
            ch->S = S;
            return 0;
        }
    }
    return -1;
}

// Host-side planning for one channel.  Fills M, PS, SP, R, S, narrow.
// Dense tiles when the windows of consecutive outputs overlap (D < T) and a
// tile fits the LDS budget, else narrow items (any D and T).
static inline int argus_plan_chan(ArgusChanDesc* ch, const ArgusTune& tune)
{
    const int D = ch->D, T = ch->T;
    if (D < T) {
        ArgusChanDesc c = *ch;
        if (argus_plan_dense(&c, ARGUS_DENSE_R, tune.lds_entries) == 0 && c.M / ARGUS_DENSE_R >= 4) {
            *ch = c;
            return 0;
        }
        c = *ch;
        if (argus_plan_dense(&c, 1, tune.lds_entries) == 0) {
            *ch = c;
            return 0;
        }
    }
    ch->narrow = 1;
    ch->M = tune.narrow_outputs;
    ch->PS = ch->SP = 0;
    ch->R = 1;
    ch->S = 1;
    return 0;
}

// LDS bytes a dense tile of this channel needs (samples + taps).
static inline size_t argus_dense_lds_bytes(const ArgusChanDesc& ch)
{
    return (size_t)ch.P * ch.PS * 8 + (size_t)ch.P * ch.L * 4;
}

// Modulated taps g[k] = h[k] e^{-j 2pi inc k} (inc in turns, 2^64 == 1) in
// natural tap order, from the phase-major layout; FP64, rounded once.
void argus_modulated_taps(const float* poly, int D, int T, uint64_t inc, float2* out);

// argus_fill_chan leaves ctaps null; callers that keep modulated taps for
// the channel set it afterwards.
int argus_fill_chan(ArgusChanDesc* ch, const float2* hist_rd, int H, float2* hist_wr, int keep, const float* taps,
                    float2* out, int D, int T, int phase_count, int n_new, double phase_start, double phase_inc,
                    const ArgusTune& tune);
size_t argus_item_lds_bytes(const ArgusChanDesc& ch);
int argus_launch_extract_batch(const ArgusChanDesc* d_chans, const ArgusWorkItem* d_items, int n_items,
                               size_t lds_bytes, const float2* in, int n_new, int grid, hipStream_t stream);
// Allow the extraction kernels dynamic LDS of `bytes` (needed above 64 KiB;
// not a stream operation, so call it outside graph capture).
int argus_extract_allow_lds(size_t bytes);
// Graph-mode launch: counts and pointers come from *hdr (device memory).
int argus_launch_extract_hdr(const ArgusExtractHdr* d_hdr, size_t lds_bytes, int grid, hipStream_t stream);
void argus_append_items(const ArgusChanDesc& ch, int chan_index, int n_new, std::vector<ArgusWorkItem>& items,
                        std::vector<int>& pos);
void argus_order_items(std::vector<ArgusWorkItem>& items, const std::vector<int>& pos, int n_new);
// LDS bytes of the largest dense tile or narrow item a tune allows for taps <= max_taps.
size_t argus_tune_max_lds(const ArgusTune& tune, int max_taps);
