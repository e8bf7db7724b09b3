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

__device__ __forceinline__ void argus_extract_tile(const ArgusChanDesc& ch, const float2* __restrict__ in, int m0,
                                                   float2* lds)
{
    const int D = ch.D, T = ch.T, L = ch.L, P = ch.P, PS = ch.PS;
    const int M = min(ch.M, ch.n_out - m0);
    const int em0 = ch.e0 + m0 * D;
    if (D < T) {
        const int n_lo = em0 - (T - 1);
        const int span = (M - 1) * D + T;
        for (int i = threadIdx.x; i < span; i += blockDim.x) {
            const int n = n_lo + i;
            const int d = em0 - n;               // in [-(M-1)D, T-1]
            int p = d % D;
            if (p < 0) p += D;
            const int j = (p - d) / D;           // exact: p - d is a multiple of D
            lds[p * PS + j + L - 1] = argus_shifted(ch, in, n);
        }
    } else {
        const int tot = M * T;
        for (int i = threadIdx.x; i < tot; i += blockDim.x) {
            const int mr = i / T;
            const int q = i - mr * T;
            const int n = em0 + mr * D - (T - 1) + q;
            lds[(T - 1 - q) * PS + mr] = argus_shifted(ch, in, n);
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < M; t += blockDim.x) {
        float ar = 0.0f, ai = 0.0f;
        for (int p = 0; p < P; ++p) {
            const float* __restrict__ tp = ch.taps + p * L;
            const float2* sp = lds + p * PS + t + L - 1;
            const int kmax = (T - p + D - 1) / D;   // taps with p + k*D < T
            for (int k = 0; k < kmax; ++k) {
                const float tap = tp[k];
                const float2 v = sp[-k];
                ar = fmaf(tap, v.x, ar);
                ai = fmaf(tap, v.y, ai);
            }
        }
        ch.out[m0 + t] = make_float2(ar, ai);
    }
    __syncthreads();
}

__device__ __forceinline__ void argus_history_tail(const ArgusChanDesc& ch, const float2* __restrict__ in, int n_new)
{
    for (int i = threadIdx.x; i < ch.keep; i += blockDim.x)
        ch.hist_wr[i] = argus_shifted(ch, in, n_new - ch.keep + i);
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
        if (it.kind == 0)
            argus_extract_tile(ch, in, it.m0, lds);
        else
            argus_history_tail(ch, in, n_new);
    }
}

// One signal: blocks [0, n_tiles) are output tiles, block n_tiles (if
// launched) writes the new history.
__global__ __launch_bounds__(ARGUS_EXTRACT_THREADS) void argus_extract_one_kernel(ArgusChanDesc ch,
                                                                                   const float2* __restrict__ in,
                                                                                   int n_new, int n_tiles)
{
    extern __shared__ float2 lds[];
    if ((int)blockIdx.x < n_tiles)
        argus_extract_tile(ch, in, blockIdx.x * ch.M, lds);
    else
        argus_history_tail(ch, in, n_new);
}

// Host: launch the single-signal kernel for a fully described channel.
int argus_launch_extract_one(const ArgusChanDesc& ch, const float2* in, int n_new, hipStream_t stream)
{
    const int n_tiles = ch.n_out > 0 ? (ch.n_out + ch.M - 1) / ch.M : 0;
    const int blocks = n_tiles + (ch.keep > 0 ? 1 : 0);
    if (blocks == 0) return 0;
    const size_t lds = n_tiles > 0 ? (size_t)ch.P * ch.PS * sizeof(float2) : 0;
    hipLaunchKernelGGL(argus_extract_one_kernel, dim3(blocks), dim3(ARGUS_EXTRACT_THREADS), lds, stream, ch, in,
                       n_new, n_tiles);
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
    hipLaunchKernelGGL(argus_extract_batch_kernel, dim3(grid), dim3(ARGUS_EXTRACT_THREADS), lds_bytes, stream,
                       d_chans, d_items, n_items, in, n_new);
    return (int)hipGetLastError();
}

// Describe one signal for the kernels.  Returns 0 or -1 if T is too large.
int argus_fill_chan(ArgusChanDesc* ch, const float2* hist_rd, int H, float2* hist_wr, int keep, const float* taps,
                    float2* out, int D, int T, int phase_count, int n_new, double phase_start, double phase_inc)
{
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
    int ps = 0;
    ch->M = argus_plan_tile(D, T, ch->n_out, &ps);
    ch->PS = ps;
    return ch->M > 0 ? 0 : -1;
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
