// extract.hpp -- fused NCO + phase-major polyphase FIR/decimate work items.
//
// One work item is a tile of M consecutive outputs of one channel.  The item
// stages the NCO-shifted input samples its outputs need into LDS, grouped by
// polyphase residue ("planes"), then each thread computes one output as
//   y[m] = sum_{p < P} sum_{k < L_p} h[p + k*D] * s[e_m - p - k*D]
// with p outer and k inner, the order of the reference's accumulation
// (streaming_host_core.go:29-52).  Plane p holds s[e_m0 - p + j*D] for
// j = -(L-1) .. M-1, so for a fixed (p, k) consecutive threads read
// consecutive LDS words: no bank conflicts, whatever D is.
//   dense  (D <  T): the tile's input span is contiguous, every residue used;
//   sparse (D >= T): only the T samples before each emission point are
//                    shifted and staged (P = T planes, one slot per output).
// A second item kind writes the channel's new shifted history (the newest
// `keep` samples of history ++ shifted input) into the other half of a
// ping-pong pair, so no item reads history another item is overwriting.
#pragma once
#include "common.hpp"

struct ArgusChanDesc {
    const float2* hist_rd;  // old shifted history, H samples, oldest first
    float2* hist_wr;        // new shifted history, keep samples
    const float* taps;      // phase-major polyphase taps, P planes of L floats
    float2* out;            // n_out baseband outputs of this chunk
    uint64_t phase0;        // NCO phase of new sample 0 (turns, 2^64 == 1)
    uint64_t inc;           // NCO increment per sample (turns, 2^64 == 1)
    int D, T, L, P;         // decimation, taps, phase length ceil(T/D), planes min(D,T)
    int H;                  // old history length
    int e0;                 // new-sample index of output 0: D - phase_count - 1
    int n_out;              // outputs this chunk
    int keep;               // new history length
    int M;                  // outputs per tile
    int PS;                 // LDS plane stride (complex samples)
};

struct ArgusWorkItem {
    int chan;
    int m0;    // first output of the tile (kind 0)
    int kind;  // 0 = output tile, 1 = history tail
    int pad;
};

#define ARGUS_EXTRACT_THREADS 256
#define ARGUS_LDS_ENTRIES 4096   // complex samples of LDS per work item (32 KiB)

// Host-side tile planning shared by the batched and per-signal launchers.
// Returns M (outputs per tile) and the plane stride for a channel; M == 0 if
// the channel cannot fit (T too large).
static inline int argus_plan_tile(int D, int T, int n_out, int* ps_out)
{
    const int L = (T + D - 1) / D;
    const int P = D < T ? D : T;
    int M = ARGUS_LDS_ENTRIES / P - (L - 1);
    if (M > ARGUS_EXTRACT_THREADS) M = ARGUS_EXTRACT_THREADS;
    if (n_out > 0 && M > n_out) M = n_out;
    while (M > 0) {
        int ps = (M + L - 1) | 1;   // odd stride spreads planes over banks
        if (P * ps <= ARGUS_LDS_ENTRIES) {
            *ps_out = ps;
            return M;
        }
        M--;
    }
    *ps_out = 0;
    return 0;
}
