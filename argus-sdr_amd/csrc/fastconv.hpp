// fastconv.hpp -- overlap-save fast convolution for the dense channels of
// the streaming service (layout and semantics in fastconv.hip).
#pragma once
#include "common.hpp"
#include "demod.hpp"
#include "extract.hpp"

#define ARGUS_FAST_N 4096      // segment transform size
#define ARGUS_FAST_TAIL 511    // input samples kept before each block (taps <= 512)
#define ARGUS_FAST_HOP 3584    // new samples per segment: <= N - TAIL, a multiple of 64
// Longest block the fast path takes: its output indexing divides block
// positions by float reciprocals, exact to one correction step below 2^22.
#define ARGUS_FAST_MAX_BLOCK (1 << 22)
// The streaming service runs the per-channel pass in this many launches over
// consecutive ranges of its fast channels, so the first range's outputs can
// start down to the host while the others are computed.
#define ARGUS_FAST_PARTS 4

// One channel on the fast path for one block.
struct ArgusFastDesc {
    const float2* G;   // ARGUS_FAST_N bins: FFT of the modulated taps, scaled 1/N
    float2* out;       // n_out outputs of this block
    uint64_t phase0;   // NCO phase of block sample 0 (turns, 2^64 == 1)
    uint64_t inc;      // NCO increment per sample
    int D;             // decimation, a multiple of 8
    int Dp;            // folding factor: the power of two dividing D, at most 64
    int e0;            // block sample of output 0
    int n_out;         // outputs this block
    int r;             // (e0 + ARGUS_FAST_TAIL) mod Dp
    // demodulation in the epilogue: ARGUS_DEMOD_NFM/WFM (every output but
    // the first of each segment: the demodulation pass adds those), USB/LSB
    // or CW (every output); ARGUS_DEMOD_NONE: baseband only
    int demod;
    float* audio;      // demodulated outputs (output m at audio[m - skip])
    uint64_t bfo0;     // SSB/CW: BFO phase of output 0, pre-increment applied (turns)
    uint64_t bfo_inc;  // SSB/CW: BFO increment per output (turns)
    int skip;          // FM: 1 when output 0 has no predecessor (no carry)
    int pad;
};

// What the fast-path kernels of one block read (device memory, so a
// captured launch replays with new contents).
struct ArgusFastHdr {
    const ArgusFastDesc* descs;
    float2* tail;      // the stream's last ARGUS_FAST_TAIL input samples
    float2* u;         // staged input: tail ++ block ++ zeros
    float2* X;         // segment spectra, S x ARGUS_FAST_N
    int n_fast;        // channels on the fast path this block
    int S;             // segments this block (ceil(n / ARGUS_FAST_HOP))
    // part p of the per-channel pass takes descs [part_end[p-1], part_end[p])
    // (part_end[-1] = 0); non-decreasing, the last equal to n_fast
    int part_end[ARGUS_FAST_PARTS];
};

// Whether a channel of decimation D and T taps takes the fast path: a
// multiple of 8 (so the fold factor, D's power-of-two part capped at 64,
// leaves at most 512 bins), at most 512 taps, and overlapping windows
// (D < T).  Measured at 8192 channels of the serving mix, the fast form
// costs about 0.28 us of kernel time per channel and 20 ms block whatever D
// is, the direct form about 0.68 ps per tap and output: the two meet near
// T = D.
static inline bool argus_fast_eligible(int D, int T)
{
    return D >= 8 && D % 8 == 0 && D < T && T <= ARGUS_FAST_TAIL + 1;
}
static inline int argus_fast_fold(int D)
{
    const int p = D & -D;
    return p > 64 ? 64 : p;
}
// The fast-path descriptor of a channel described for the direct form
// (same outputs, NCO and grid), baseband only.
static inline ArgusFastDesc argus_fast_desc(const ArgusChanDesc& ch, const float2* G)
{
    ArgusFastDesc fd{};
    fd.G = G;
    fd.out = ch.out;
    fd.phase0 = ch.phase0;
    fd.inc = ch.inc;
    fd.D = ch.D;
    fd.Dp = argus_fast_fold(ch.D);
    fd.e0 = ch.e0;
    fd.n_out = ch.n_out;
    fd.r = (ch.e0 + ARGUS_FAST_TAIL) % fd.Dp;
    fd.demod = ARGUS_DEMOD_NONE;
    fd.audio = nullptr;
    return fd;
}
static inline int argus_fast_segments(int n) { return n > 0 ? (n + ARGUS_FAST_HOP - 1) / ARGUS_FAST_HOP : 0; }
static inline size_t argus_fast_u_len(int segments)
{
    return segments > 0 ? (size_t)(segments - 1) * ARGUS_FAST_HOP + ARGUS_FAST_N : 0;
}

// G from the natural-order taps h (T floats) and the NCO increment: the
// N-point transform of g[k] = h[k] e^{-j 2pi inc k}, scaled by 1/N, in FP64,
// rounded once.
void argus_fast_spectrum(const float* h, int T, uint64_t inc, float2* G);
// Segment spectra and the per-channel pass (when S > 0), reading the block
// input and its length from *xh; u and X are the buffers d_hdr names (fixed
// per service), passed again for the transform launch.
int argus_launch_fast(const ArgusFastHdr* d_hdr, const ArgusExtractHdr* d_xh, int S, float2* u, float2* X,
                      int grid, hipStream_t stream);
// The two halves of argus_launch_fast: the staged input and its segment
// spectra; the per-channel pass over every fast channel (part -1) or over
// part p's range of them (d_hdr->part_end).
int argus_launch_fast_input(const ArgusFastHdr* d_hdr, const ArgusExtractHdr* d_xh, int S, float2* u, float2* X,
                            hipStream_t stream);
int argus_launch_fast_chans(const ArgusFastHdr* d_hdr, int S, int grid, int part, hipStream_t stream);
// The stream's tail after the block (every block, fast channels or not).
int argus_launch_fast_tail(const ArgusFastHdr* d_hdr, const ArgusExtractHdr* d_xh, hipStream_t stream);
// surv.hip: forward 4096-point transforms of n_frames frames at in + f * frame_step.
int argus_launch_fft4096(const float2* in, int64_t frame_step, int n_frames, float2* out, hipStream_t stream);
