// rds.hpp -- device-side RDS group decoder shared by the batch API
// (gpud_rds_batch_*) and the streaming service (gpud_serve_set_rds).
// Kernels and semantics are in rds_batch.hip.
#pragma once
#include "common.hpp"

// One decoder's state (rds.go:10-31 Decoder fields the decode reads), device
// resident; 128 bytes, zero = a fresh decoder.
struct ArgusRdsState {
    double costas_phase, costas_freq, mu;
    float prev_re, prev_im, prev_dec;
    int32_t have_prev, last_rate, last_hard, have_last_hard;
    uint32_t pi;
    char ps[8];
    char rt[64];
};
static_assert(sizeof(ArgusRdsState) == 128, "state layout");

// What one Decode call returns: differential bits produced (-1: the call
// stopped before its group scan), groups found and written to the call's
// list, and the decoder's PI, PS and RT after the call.
struct ArgusRdsOut {
    int32_t nb, found, written;
    uint32_t pi;
    char ps[8];
    char rt[64];
};

// A Decode call in progress: the symbol loop's registers between the blocks
// that bring its window's samples (zero = not started).
struct ArgusRdsProgress {
    double m, phase, freq;
    float prev_re, prev_im, prev_dec;
    int32_t have_prev, ns, nb, prev_hard, active, loop_done;
};

// One Decode call over a window of n samples at x (device) for the decoder
// at st.  The symbol loop runs over the first `avail` samples and pauses
// where it needs more; `final` = the window is complete (avail == n): the
// call finishes and, if it reached 104 symbols, its group scan runs.  The
// call's bits, match flags (bit_cap bytes each), next-hit table (2 x
// bit_cap ints) and group list (group_cap groups) come from
// argus_rds_caps; results in *out.
struct ArgusRdsCall {
    const float2* x;
    uint8_t* bits;
    uint8_t* flags;
    int32_t* next;
    uint16_t* groups;   // 4 words per group
    ArgusRdsState* st;
    ArgusRdsProgress* pr;
    ArgusRdsOut* out;
    int32_t n, avail, rate, bit_cap, group_cap, final;
};

// Copy of a block's RDS baseband into a station's decode windows: n0
// samples to dst0, then n1 to dst1 (the next window).
struct ArgusRdsAppend {
    const float2* src;
    float2* dst0;
    float2* dst1;
    int32_t n0, n1;
};

// Bits and groups one call of n samples at `rate` may produce (0, 0 for a
// call that cannot reach its group scan).
void argus_rds_caps(int n, int rate, int* bit_cap, int* group_cap);
// The decode launches for n_calls calls (table in device memory): the
// symbol loop of every call, then the group scan of the first n_final
// calls, which must be the final ones; max_bit_cap = the largest bit_cap
// among those.  At most 65535 calls.
int argus_launch_rds(const ArgusRdsCall* d_calls, int n_calls, int n_final, int max_bit_cap, hipStream_t stream);
int argus_launch_rds_append(const ArgusRdsAppend* d_appends, int n, hipStream_t stream);
// PS or RT as the reference's Result reports it (rds.go:604-636): unset
// characters read as spaces, trailing spaces trimmed, NUL-terminated.
void argus_rds_text(const char* src, int len, char* dst);
