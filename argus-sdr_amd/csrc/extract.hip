// extract.hip -- fused streaming extraction kernels and the stateful
// single-signal entry points (successors of exports.cu:356-631).
#include "extract.hpp"
#include <algorithm>

// NCO-shifted sample n of the chunk (n < 0 reads the old shifted history).
__device__ __forceinline__ float2 argus_shifted(const ArgusChanDesc& ch, const float2* __restrict__ in, int n)
{
    if (n >= 0) {
        const float2 x = in[n];
        float s, c;
        argus_nco(ch.phase0 + ch.inc * (uint64_t)(uint32_t)n, s, c);
        return argus_rot(x, s, c);
    }
    const int h = ch.H + n;
    return h >= 0 ? ch.hist_rd[h] : make_float2(0.0f, 0.0f);
}

__device__ __forceinline__ float argus_wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Dense tile: M outputs starting at m0, samples staged in LDS by residue.
__device__ __forceinline__ void argus_dense_tile(const ArgusChanDesc& ch, const float2* __restrict__ in, int m0,
                                                 float2* lds_s)
{
    const int D = ch.D, T = ch.T, L = ch.L, P = ch.P, PS = ch.PS;
    const int M = min(ch.M, ch.n_out - m0);
    const int em0 = ch.e0 + m0 * D;
    float* lds_t = (float*)(lds_s + P * PS);
    for (int i = threadIdx.x; i < P * L; i += blockDim.x) lds_t[i] = ch.taps[i];

    // Staging: sample n = n_lo + i for i in [0, span); d = em0 - n = T-1-i,
    // residue p = d mod D, plane index j = (p - d) / D.  One division per
    // thread, then stepped by blockDim without dividing again.
    const int span = (M - 1) * D + T;
    const int n_lo = em0 - (T - 1);
    int i = threadIdx.x;
    if (i < span) {
        const int B = blockDim.x;
        const int stepq = B / D, stepr = B - stepq * D;
        int d = (T - 1) - i;
        int p = d % D;
        if (p < 0) p += D;
        int j = (p - d) / D;
        uint64_t ph = ch.phase0 + ch.inc * (uint64_t)(int64_t)(n_lo + i);
        const uint64_t phstep = ch.inc * (uint64_t)B;
        for (; i < span; i += B) {
            const int n = n_lo + i;
            float2 v;
            if (n >= 0) {
                float sn, cs;
                argus_nco(ph, sn, cs);
                v = argus_rot(in[n], sn, cs);
            } else {
                const int h = ch.H + n;
                v = h >= 0 ? ch.hist_rd[h] : make_float2(0.0f, 0.0f);
            }
            lds_s[p * PS + j + L - 1] = v;
            ph += phstep;
            p -= stepr;
            j += stepq;
            if (p < 0) {
                p += D;
                j += 1;
            }
        }
    }
    __syncthreads();

    // FIR: output m = t / S, lanes g = t % S share its planes p = g, g+S, ...
    const int S = ch.S;
    const int t = threadIdx.x;
    const int m = t / S;
    const int g = t & (S - 1);
    float ar0 = 0.0f, ai0 = 0.0f, ar1 = 0.0f, ai1 = 0.0f;
    if (m < M) {
        const int pfull = T - (L - 1) * D;  // planes p < pfull carry L taps, the rest L-1
        for (int p = g; p < P; p += S) {
            const int kmax = p < pfull ? L : L - 1;
            const float* tp = lds_t + p * L;
            const float2* sp = lds_s + p * PS + m + L - 1;
            int k = 0;
            for (; k + 4 <= kmax; k += 4) {
                const float h0 = tp[k], h1 = tp[k + 1], h2 = tp[k + 2], h3 = tp[k + 3];
                const float2 v0 = sp[-k], v1 = sp[-k - 1], v2 = sp[-k - 2], v3 = sp[-k - 3];
                ar0 = fmaf(h0, v0.x, ar0);
                ai0 = fmaf(h0, v0.y, ai0);
                ar1 = fmaf(h1, v1.x, ar1);
                ai1 = fmaf(h1, v1.y, ai1);
                ar0 = fmaf(h2, v2.x, ar0);
                ai0 = fmaf(h2, v2.y, ai0);
                ar1 = fmaf(h3, v3.x, ar1);
                ai1 = fmaf(h3, v3.y, ai1);
            }
            for (; k < kmax; ++k) {
                const float h = tp[k];
                const float2 v = sp[-k];
                ar0 = fmaf(h, v.x, ar0);
                ai0 = fmaf(h, v.y, ai0);
            }
        }
    }
    float ar = ar0 + ar1, ai = ai0 + ai1;
    for (int o = S >> 1; o > 0; o >>= 1) {
        ar += __shfl_xor(ar, o);
        ai += __shfl_xor(ai, o);
    }
    if (m < M && g == 0) ch.out[m0 + m] = make_float2(ar, ai);
    __syncthreads();
}

// Narrow group: one wavefront per output, lanes over taps, no staging.
__device__ __forceinline__ void argus_narrow_group(const ArgusChanDesc& ch, const float2* __restrict__ in, int m0,
                                                   float* lds_t)
{
    const int T = ch.T, D = ch.D;
    for (int i = threadIdx.x; i < T; i += blockDim.x) lds_t[i] = ch.taps[i];  // L == 1: taps[k] = h[k]
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const int mend = min(m0 + ch.M, ch.n_out);
    for (int m = m0 + wave; m < mend; m += nw) {
        const int e = ch.e0 + m * D;
        float ar = 0.0f, ai = 0.0f;
        for (int k = lane; k < T; k += 64) {
            const float2 s = argus_shifted(ch, in, e - k);
            const float h = lds_t[k];
            ar = fmaf(h, s.x, ar);
            ai = fmaf(h, s.y, ai);
        }
        ar = argus_wave_sum(ar);
        ai = argus_wave_sum(ai);
        if (lane == 0) ch.out[m] = make_float2(ar, ai);
    }
    __syncthreads();
}

__device__ __forceinline__ void argus_history_tail(const ArgusChanDesc& ch, const float2* __restrict__ in, int n_new)
{
    for (int i = threadIdx.x; i < ch.keep; i += blockDim.x)
        ch.hist_wr[i] = argus_shifted(ch, in, n_new - ch.keep + i);
}

__device__ __forceinline__ void argus_run_item(const ArgusChanDesc& ch, int kind, int m0,
                                               const float2* __restrict__ in, int n_new, float2* lds)
{
    if (kind == ARGUS_ITEM_DENSE)
        argus_dense_tile(ch, in, m0, lds);
    else if (kind == ARGUS_ITEM_NARROW)
        argus_narrow_group(ch, in, m0, (float*)lds);
    else if (kind == ARGUS_ITEM_HISTORY)
        argus_history_tail(ch, in, n_new);
}

// Batched, persistent: every work item of every channel of one chunk.
__global__ __launch_bounds__(ARGUS_EXTRACT_THREADS) void argus_extract_batch_kernel(
    const ArgusChanDesc* __restrict__ chans, const ArgusWorkItem* __restrict__ items, int n_items,
    const float2* __restrict__ in, int n_new)
{
    extern __shared__ float2 lds[];
    for (int w = blockIdx.x; w < n_items; w += gridDim.x) {
        const ArgusWorkItem it = items[w];
        const ArgusChanDesc ch = chans[it.chan];
        argus_run_item(ch, it.kind, it.m0, in, n_new, lds);
    }
}

// One signal: blocks [0, n_items) are output items, block n_items (if
// launched) writes the new history.
__global__ __launch_bounds__(ARGUS_EXTRACT_THREADS) void argus_extract_one_kernel(ArgusChanDesc ch,
                                                                                   const float2* __restrict__ in,
                                                                                   int n_new, int n_items)
{
    extern __shared__ float2 lds[];
    const int kind = (int)blockIdx.x < n_items ? (ch.narrow ? ARGUS_ITEM_NARROW : ARGUS_ITEM_DENSE)
                                               : ARGUS_ITEM_HISTORY;
    argus_run_item(ch, kind, blockIdx.x * ch.M, in, n_new, lds);
}

size_t argus_item_lds_bytes(const ArgusChanDesc& ch)
{
    return ch.narrow ? (size_t)ch.T * sizeof(float) : argus_dense_lds_bytes(ch);
}

// Host: launch the single-signal kernel for a fully described channel.
int argus_launch_extract_one(const ArgusChanDesc& ch, const float2* in, int n_new, hipStream_t stream)
{
    const int n_items = ch.n_out > 0 ? (ch.n_out + ch.M - 1) / ch.M : 0;
    const int blocks = n_items + (ch.keep > 0 ? 1 : 0);
    if (blocks == 0) return 0;
    const size_t lds = n_items > 0 ? argus_item_lds_bytes(ch) : 0;
    argus_clear_error();
    hipLaunchKernelGGL(argus_extract_one_kernel, dim3(blocks), dim3(ARGUS_EXTRACT_THREADS), lds, stream, ch, in,
                       n_new, n_items);
    return (int)hipGetLastError();
}

int argus_launch_extract_batch(const ArgusChanDesc* d_chans, const ArgusWorkItem* d_items, int n_items,
                               size_t lds_bytes, const float2* in, int n_new, hipStream_t stream)
{
    if (n_items <= 0) return 0;
    // Persistent grid: enough blocks to fill every CU several times over; the
    // item list is ordered heaviest first so the tail is short.
    const int max_grid = 256 * 8;
    const int grid = n_items < max_grid ? n_items : max_grid;
    argus_clear_error();
    hipLaunchKernelGGL(argus_extract_batch_kernel, dim3(grid), dim3(ARGUS_EXTRACT_THREADS), lds_bytes, stream,
                       d_chans, d_items, n_items, in, n_new);
    return (int)hipGetLastError();
}

// Append the work items of one described channel; pos receives each item's
// first input sample (history reads count as sample 0).
void argus_append_items(const ArgusChanDesc& ch, int chan_index, int n_new, std::vector<ArgusWorkItem>& items,
                        std::vector<int>& pos)
{
    const int tiles = ch.n_out > 0 ? (ch.n_out + ch.M - 1) / ch.M : 0;
    // cost in lane-operations: NCO rotations (~24 each) plus taps (~2 each)
    const long rot = ch.narrow ? (long)ch.M * ch.T : (long)(ch.M - 1) * ch.D + ch.T;
    const long cost = rot * 24 + (long)ch.M * ch.T * 2;
    const int kind = ch.narrow ? ARGUS_ITEM_NARROW : ARGUS_ITEM_DENSE;
    for (int t = 0; t < tiles; t++) {
        items.push_back({chan_index, t * ch.M, kind, (int)std::min<long>(cost, 1L << 30)});
        pos.push_back(std::max(0, ch.e0 + t * ch.M * ch.D - (ch.T - 1)));
    }
    if (ch.keep > 0) {
        items.push_back({chan_index, 0, ARGUS_ITEM_HISTORY, ch.keep * 24});
        pos.push_back(std::max(0, n_new - ch.keep));
    }
}

// Order the items of one launch for the persistent grid.  Workgroups are
// dealt to the XCDs round-robin, and the grid is a multiple of ARGUS_XCDS, so
// item w always runs on XCD w % ARGUS_XCDS.  Items are bucketed by where
// their input lies in the block (one eighth per XCD), heaviest first within a
// bucket, and interleaved: each XCD's L2 then holds one eighth of the block
// instead of all of it.  Short buckets are padded with no-op items.
//
// "Heaviest first" is by cost class (half-octaves of the cost estimate), a
// stable counting sort: O(items) host time per block, which matters at
// thousands of channels (a comparison sort of ~10^5 items costs milliseconds).
void argus_order_items(std::vector<ArgusWorkItem>& items, const std::vector<int>& pos, int n_new)
{
    constexpr int kClasses = 64;
    constexpr int kKeys = ARGUS_XCDS * kClasses;
    thread_local std::vector<int> key;
    thread_local std::vector<ArgusWorkItem> sorted;
    int count[kKeys + 1] = {0};
    const size_t n = items.size();
    key.resize(n);
    const long span = std::max(1, n_new);
    for (size_t i = 0; i < n; i++) {
        int b = (int)((long)pos[i] * ARGUS_XCDS / span);
        b = std::min(std::max(b, 0), ARGUS_XCDS - 1);
        const unsigned long c = (unsigned long)std::max(items[i].pad, 1);
        const int msb = 63 - __builtin_clzl(c);
        const int half = msb > 0 ? (int)((c >> (msb - 1)) & 1) : 0;
        const int cls = std::min(kClasses - 1, 2 * msb + half);
        key[i] = b * kClasses + (kClasses - 1 - cls);  // heavier -> smaller key
        count[key[i] + 1]++;
    }
    for (int k = 0; k < kKeys; k++) count[k + 1] += count[k];
    int bstart[ARGUS_XCDS + 1];
    for (int b = 0; b <= ARGUS_XCDS; b++) bstart[b] = count[b * kClasses];
    sorted.resize(n);
    for (size_t i = 0; i < n; i++) sorted[(size_t)count[key[i]]++] = items[i];
    int longest = 0;
    for (int b = 0; b < ARGUS_XCDS; b++) longest = std::max(longest, bstart[b + 1] - bstart[b]);
    items.resize((size_t)longest * ARGUS_XCDS);
    const ArgusWorkItem nop{0, 0, ARGUS_ITEM_NOP, 0};
    for (int j = 0; j < longest; j++)
        for (int x = 0; x < ARGUS_XCDS; x++)
            items[(size_t)j * ARGUS_XCDS + x] = j < bstart[x + 1] - bstart[x] ? sorted[(size_t)(bstart[x] + j)] : nop;
}

// Describe one signal for the kernels.  Returns 0 or -1 if the geometry
// cannot be staged (T too large).
int argus_fill_chan(ArgusChanDesc* ch, const float2* hist_rd, int H, float2* hist_wr, int keep, const float* taps,
                    float2* out, int D, int T, int phase_count, int n_new, double phase_start, double phase_inc)
{
    if (T > ARGUS_MAX_TAPS) return -1;
    ch->hist_rd = hist_rd;
    ch->hist_wr = hist_wr;
    ch->taps = taps;
    ch->out = out;
    ch->phase0 = argus_rad_to_fix(phase_start);
    ch->inc = argus_rad_to_fix(phase_inc);
    ch->D = D;
    ch->T = T;
    ch->L = (T + D - 1) / D;
    ch->P = D < T ? D : T;
    ch->H = H;
    ch->e0 = D - phase_count - 1;
    ch->n_out = (phase_count + n_new) / D;
    ch->keep = keep;
    return argus_plan_chan(ch);
}

extern "C" {

// exports.cu:533-631
GPUD_API int gpud_launch_streaming_polyphase_stateful_hip(
    const gpud_float2* in_new, int n_new, gpud_float2* shifted_new_tmp, const float* polyphase_taps,
    int polyphase_len, int decim, int num_taps, gpud_float2* history_state, gpud_float2* history_scratch,
    int history_cap, int* history_len_io, int* phase_count_state, double* phase_state, double phase_inc,
    gpud_float2* out, int out_cap, int* n_out)
{
    if (!polyphase_taps || decim <= 0 || num_taps <= 0 || !history_len_io || !phase_count_state || !phase_state ||
        !n_out)
        return -10;
    if (n_new < 0 || out_cap < 0 || history_cap < 0) return -11;
    const int L = (num_taps + decim - 1) / decim;
    if (polyphase_len < decim * L) return -12;
    int H = *history_len_io;
    if (H < 0) H = 0;
    if (H > history_cap) H = history_cap;
    int pc = *phase_count_state;
    if (pc < 0) pc = 0;
    if (pc >= decim) pc %= decim;
    const double phase_start = *phase_state;
    if (n_new > 0 && (!in_new || !shifted_new_tmp)) return -13;
    const int total = pc + n_new;
    const int out_count = total / decim;
    if (out_count > out_cap) return -14;
    if (out_count > 0 && !out) return -15;
    const int new_h = history_cap > 0 ? std::min(H + n_new, history_cap) : 0;
    if (new_h > 0 && (!history_state || !history_scratch)) return -16;

    ArgusChanDesc ch;
    if (argus_fill_chan(&ch, (const float2*)history_state, H, (float2*)history_scratch, new_h, polyphase_taps,
                        (float2*)out, decim, num_taps, pc, n_new, phase_start, phase_inc) != 0)
        return -12;
    int err = argus_launch_extract_one(ch, (const float2*)in_new, n_new, 0);
    if (err) return err;
    if (new_h > 0)
        ARGUS_CHECK(hipMemcpyAsync(history_state, history_scratch, (size_t)new_h * sizeof(float2),
                                   hipMemcpyDeviceToDevice, 0));
    *history_len_io = new_h;
    *phase_count_state = total % decim;
    *phase_state = argus_reduce_phase(phase_start + phase_inc * (double)n_new);
    *n_out = out_count;
    return 0;
}

// exports.cu:356-460 (no per-call allocation here: the shift is fused).
GPUD_API int gpud_launch_streaming_polyphase_prepare_hip(
    const gpud_float2* in_new, int n_new, const gpud_float2* history_in, int history_len,
    const float* polyphase_taps, int polyphase_len, int decim, int num_taps, int phase_count_in,
    double phase_start, double phase_inc, gpud_float2* out, int* n_out, int* phase_count_out,
    double* phase_end_out, gpud_float2* history_out)
{
    if (n_new < 0 || !polyphase_taps || polyphase_len <= 0 || decim <= 0 || num_taps <= 0) return -1;
    const int L = (num_taps + decim - 1) / decim;
    if (polyphase_len < decim * L) return -2;
    const int keep = num_taps > 1 ? num_taps - 1 : 0;
    int H = history_len;
    if (H < 0) H = 0;
    if (H > keep) H = keep;
    if (H > 0 && !history_in) return -5;
    if (n_new > 0 && !in_new) return -3;
    int pc = phase_count_in;
    if (pc < 0) pc = 0;
    if (pc >= decim) pc %= decim;
    const int total = pc + n_new;
    const int out_count = total / decim;
    if (out_count > 0 && !out) return -4;
    const int new_h = (history_out && keep > 0) ? std::min(H + n_new, keep) : 0;
    ArgusChanDesc ch;
    if (argus_fill_chan(&ch, (const float2*)history_in, H, (float2*)history_out, new_h, polyphase_taps,
                        (float2*)out, decim, num_taps, pc, n_new, phase_start, phase_inc) != 0)
        return -2;
    int err = argus_launch_extract_one(ch, (const float2*)in_new, n_new, 0);
    if (err) return err;
    if (n_out) *n_out = out_count;
    if (phase_count_out) *phase_count_out = total % decim;
    if (phase_end_out) *phase_end_out = argus_reduce_phase(phase_start + phase_inc * (double)n_new);
    return 0;
}

}  // extern "C"
