// extract.hpp -- fused NCO + polyphase FIR/decimate work items.
//
// Two shapes of work item, picked per channel by the host:
//
//  dense tile (D < T): a tile of M = 4G consecutive outputs.  The
//    NCO-shifted input span the tile needs is staged in LDS grouped by
//    polyphase residue ("planes"): plane p holds s[e_m0 - p + j*D] for
//    j = -(Lp-1) .. M-1 (position x = j + Lp - 1), so each output is
//        y[m] = sum_{p < D} sum_{k < Lp} h[p + k*D] * s[e_m - p - k*D]
//    (the reference's phase-major order, streaming_host_core.go:29-52) with
//    the L = ceil(T/D) taps of a plane zero-padded to Lp = roundup(L, 4).
//    Sample i of the span (input n_lo + i) lands in plane D-1-(i mod D) at
//    position i div D.  A thread computes the 4 consecutive outputs of one
//    output group over its share of the planes (S lanes per group, combined
//    by lane shuffles), 4 taps at a time: a block needs 7 consecutive
//    positions.  Positions are stored reversed, 5 words per 4 positions
//    (word of reversed position y = y + y div 4), so the 7 samples of a
//    block sit at fixed offsets {0,1,2,3,5,6,7} from one pointer that
//    advances by 5 words per block, and the output groups of a read are 5
//    words apart (odd: distinct banks).  The plane stride is chosen so the S
//    planes of one read land in disjoint bank ranges.
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
#include "demod.hpp"
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
    int PS;                 // LDS plane stride (complex words), dense only
    int SP;                 // LDS word offset of the padded taps, dense only
    int R;                  // outputs per thread in the dense FIR phase (4; 1 for narrow)
    int S;                  // lanes per output group in the dense FIR phase (power of 2)
    int narrow;             // 1 = narrow (wave-per-output) items
    // demodulation in the epilogue of the narrow items with modulated taps
    // (streaming service): ARGUS_DEMOD_NFM/WFM (every output but each item's
    // first: the demodulation pass adds those and the carry), USB/LSB or CW
    // (every output); ARGUS_DEMOD_NONE (0): baseband only
    int demod;
    int skip;               // FM: 1 when output 0 has no predecessor (no carry)
    float* audio;           // demodulated outputs (output m at audio[m - skip])
    uint64_t bfo0;          // SSB/CW: BFO phase of output 0, pre-increment applied (turns)
    uint64_t bfo_inc;       // SSB/CW: BFO increment per output (turns)
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
    int fast_grid;       // workgroups of the fast form's per-channel kernel (multiple of ARGUS_XCDS)
};

#define ARGUS_FAST_GRID_DEFAULT 2048

static inline ArgusTune argus_default_tune()
{
    return ArgusTune{ARGUS_LDS_ENTRIES, ARGUS_NARROW_OUTPUTS, ARGUS_GRID, ARGUS_FAST_GRID_DEFAULT};
}

#define ARGUS_DENSE_R 4             // outputs per thread in the dense FIR

// Dense tile geometry: taps per plane padded to a multiple of 4, positions
// per plane for G output groups, and words per plane (5 per 4 positions).
static inline int argus_dense_lp(int L) { return (L + 3) & ~3; }
static inline int argus_dense_words(int G, int Lp)
{
    const int y = 4 * G + Lp - 2;  // largest reversed position
    return y + (y >> 2) + 1;
}

// Plane stride for S lanes per output group: at least `words`, and
// congruent mod 32 to -5*32/S, so the S planes x 32/S output groups of a
// 32-lane read (addresses g*PS - 5q) cover 32 distinct banks.  S = 1 or
// S >= 32: odd (then consecutive planes start in distinct banks, which
// keeps the staging writes apart too).
static inline int argus_plane_stride(int words, int S)
{
    int PS = words;
    if (S >= 2 && S < 32) {
        const int want = (32 - (5 * 32 / S) % 32) & 31;
        while ((PS & 31) != want) PS++;
    } else if (!(PS & 1)) {
        PS++;
    }
    return PS;
}

// Complex words of LDS a dense tile takes: planes, then the padded taps
// (P*Lp floats) from an even word so they can be read 16 bytes at a time.
static inline int argus_dense_entries(int P, int PS, int Lp)
{
    const int samples = (P * PS + 1) & ~1;
    return samples + (P * Lp + 1) / 2;
}

// Dense geometry within `entries` complex words of LDS.  Returns 0, or -1 if
// fewer than 4 output groups fit.  The bank-padded plane stride is taken if
// it keeps at least 7/8 of the largest tile the budget holds unpadded.
static inline int argus_plan_dense(ArgusChanDesc* ch, int entries)
{
    const int P = ch->P, Lp = argus_dense_lp(ch->L);
    const int need = ch->n_out > 0 ? (ch->n_out + 3) / 4 : 1;
    int Gmax = entries / P / 5;  // upper bound: a plane holds 4G + Lp - 1 positions at 5 words per 4
    if (Gmax > ARGUS_EXTRACT_THREADS) Gmax = ARGUS_EXTRACT_THREADS;
    if (Gmax > need) Gmax = need;
    auto lanes = [&](int G) {
        int S = 1;
        while (S * 2 * G <= ARGUS_EXTRACT_THREADS && S * 2 <= P && S < 64) S *= 2;
        return S;
    };
    // largest G with the padded stride, and with the plain one
    int Gpad = 0, Graw = 0;
    for (int G = Gmax; G > 0 && (G >= 4 || G == need) && (!Gpad || !Graw); G--) {
        const int words = argus_dense_words(G, Lp);
        if (!Gpad && argus_dense_entries(P, argus_plane_stride(words, lanes(G)), Lp) <= entries) Gpad = G;
        if (!Graw && argus_dense_entries(P, words, Lp) <= entries) Graw = G;
    }
    if (!Gpad && !Graw) return -1;
    const bool pad = Gpad && 8 * Gpad >= 7 * Graw;
    const int G = pad ? Gpad : Graw;
    const int S = lanes(G);
    const int words = argus_dense_words(G, Lp);
    const int PS = pad ? argus_plane_stride(words, S) : words;
    ch->narrow = 0;
    ch->R = ARGUS_DENSE_R;
    ch->M = 4 * G;
    ch->SP = (P * PS + 1) & ~1;  // word offset of the taps (even)
    ch->PS = PS;
    ch->S = S;
    return 0;
}

// Host-side planning for one channel.  Fills M, PS, SP, R, S, narrow.
// Dense tiles when the windows of consecutive outputs overlap (D < T), a
// tile fits the LDS budget and, with modulated taps available, the dense
// form is the cheaper one (staging costs ~10 lane operations per input
// sample, a dense tap ~2.2 per output, a narrow complex tap ~5.5 per
// output); else narrow items (any D and T).
static inline int argus_plan_chan(ArgusChanDesc* ch, const ArgusTune& tune, bool mod_taps)
{
    const int D = ch->D, T = ch->T;
    bool dense = D < T;
    if (dense && mod_taps) {
        const long dense_cost = 22L * argus_dense_lp(ch->L) * D + 100L * D;
        const long narrow_cost = 55L * ((T + 63) & ~63);
        dense = dense_cost < narrow_cost;
    }
    if (dense) {
        ArgusChanDesc c = *ch;
        if (argus_plan_dense(&c, tune.lds_entries) == 0) {
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
    return (size_t)argus_dense_entries(ch.P, ch.PS, argus_dense_lp(ch.L)) * 8;
}

// Modulated taps g[k] = h[k] e^{-j 2pi inc k} (inc in turns, 2^64 == 1) in
// natural tap order, from the phase-major layout; FP64, rounded once.
void argus_modulated_taps(const float* poly, int D, int T, uint64_t inc, float2* out);

// ctaps: the channel's modulated taps (argus_modulated_taps for the same
// increment), or null when the caller keeps none.
int argus_fill_chan(ArgusChanDesc* ch, const float2* hist_rd, int H, float2* hist_wr, int keep, const float* taps,
                    const float2* ctaps, float2* out, int D, int T, int phase_count, int n_new, double phase_start,
                    double phase_inc, const ArgusTune& tune);
size_t argus_item_lds_bytes(const ArgusChanDesc& ch);
int argus_launch_extract_batch(const ArgusChanDesc* d_chans, const ArgusWorkItem* d_items, int n_items,
                               size_t lds_bytes, const float2* in, int n_new, int grid, hipStream_t stream);
// Allow the extraction kernels dynamic LDS of `bytes` (needed above 64 KiB;
// not a stream operation, so call it outside graph capture).
int argus_extract_allow_lds(size_t bytes);
// Graph-mode launch: counts and pointers come from *hdr (device memory).
int argus_launch_extract_hdr(const ArgusExtractHdr* d_hdr, size_t lds_bytes, int grid, hipStream_t stream);
void argus_append_items(const ArgusChanDesc& ch, int chan_index, std::vector<ArgusWorkItem>& items);
void argus_order_items(std::vector<ArgusWorkItem>& items);
// LDS bytes of the largest dense tile or narrow item a tune allows for taps <= max_taps.
size_t argus_tune_max_lds(const ArgusTune& tune, int max_taps);
