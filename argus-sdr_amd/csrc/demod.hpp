// demod.hpp -- batched demodulation of extracted baseband (FM discriminator
// with block bridging, AM envelope minus block mean, SSB/CW product
// detector with a carried BFO phase).
#pragma once
#include "common.hpp"

enum {
    ARGUS_DEMOD_NONE = 0,
    ARGUS_DEMOD_NFM = 1,
    ARGUS_DEMOD_WFM = 2,
    ARGUS_DEMOD_AM = 3,
    ARGUS_DEMOD_USB = 4,
    ARGUS_DEMOD_LSB = 5,
    ARGUS_DEMOD_CW = 6,
};

struct ArgusDemodDesc {
    const float2* iq;     // this block's baseband, n samples
    float* audio;         // n - skip outputs
    const float2* carry_rd;  // previous block's last baseband sample (FM)
    float2* carry_wr;        // this block's last baseband sample (FM)
    uint64_t bfo0;        // SSB/CW: NCO phase of sample 0, pre-increment applied (turns)
    uint64_t bfo_inc;     // SSB/CW: increment per sample (turns)
    int n;                // baseband samples this block
    int mode;             // ARGUS_DEMOD_*
    int have_carry;       // FM: carry_rd holds a sample
    int skip;             // outputs not produced at the head (FM without carry: 1)
};

// An item: outputs [start, start + count) of channel chan, or (kind
// ARGUS_DEMOD_ITEM_SEAMS) only the FM outputs the fast-convolution kernel
// leaves to this pass when it demodulates in its epilogue: the first output
// of each segment (whose predecessor another wavefront produced), with the
// channel's output grid e0 = start, decimation D = count, and the FM carry;
// or (kind ARGUS_DEMOD_ITEM_SEAMS_EVERY) those the direct form's narrow
// items leave: every count-th output (each item's first) and the FM carry.
struct ArgusDemodItem {
    int chan;
    int start;
    int count;
    int kind;
};

enum {
    ARGUS_DEMOD_ITEM_OUT = 0,
    ARGUS_DEMOD_ITEM_SEAMS = 1,
    ARGUS_DEMOD_ITEM_SEAMS_EVERY = 2,
};

// Per-output arithmetic of the demodulators, shared by the demodulation pass
// and the fast-convolution kernel's epilogue: separate roundings (no
// contraction), so both give the same bits for the same baseband.
__device__ __forceinline__ float argus_fm_step(float2 p, float2 c)
{
#pragma clang fp contract(off)
    return atan2f(p.x * c.y - p.y * c.x, p.x * c.x + p.y * c.y);
}

__device__ __forceinline__ float argus_ssb_out(float2 v, float s, float c)
{
#pragma clang fp contract(off)
    return v.x * c - v.y * s;
}

#define ARGUS_DEMOD_TILE 4096

// Audio samples a channel produces from n baseband samples.
static inline int argus_demod_count(int mode, int n, int have_carry)
{
    if (mode == ARGUS_DEMOD_NONE) return 0;
    if (mode == ARGUS_DEMOD_NFM || mode == ARGUS_DEMOD_WFM) return have_carry ? n : (n > 0 ? n - 1 : 0);
    return n;
}

struct ArgusDemodHdr {
    const ArgusDemodDesc* descs;
    const ArgusDemodItem* items;
    int n_items;
    int pad;
};

#define ARGUS_DEMOD_GRID 2048

int argus_launch_demod(const ArgusDemodDesc* d_desc, const ArgusDemodItem* d_items, int n_items,
                       hipStream_t stream);
// Graph-mode launch: counts and pointers come from *hdr (device memory).
int argus_launch_demod_hdr(const ArgusDemodHdr* d_hdr, hipStream_t stream);
