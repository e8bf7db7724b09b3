// wfm.hip -- WFM multiplex kernels: stereo decoding and the streaming RDS
// baseband over every stereo channel of a block (layout and stages in
// wfm.hpp).
//
// Semantics (internal/recorder/streamer.go:1525-1643):
//   lpr   = LPF15k(mono), bpf = LPF53k(mono) - LPF23k(mono),
//   pilot = LPF21k(mono) - LPF17k(mono)      (101-tap Hann designs, fir.go:6-28)
//   PLL per sample: I/Q mix with (cos, -sin)(phase), one-pole 200 Hz low-pass,
//   err = atan2(Q, I), freq += beta*err clamped to 17-21 kHz,
//   phase += freq + alpha*err wrapped into [0, 2pi]
//   lr    = bpf * 2 sin(2 phase), filtered by LPF15k
//   L, R  = (lpr + lr)/2, (lpr - lr)/2
//   lock  = coherence > 0.15 and |f - 19 kHz| < 250 Hz and errAvg < 0.35,
//           enabled after 4 locked blocks, disabled after 10 unlocked
//           (streamer.go:1367-1380); when not enabled the output is the mono
//           signal through its own stateful LPF15k, duplicated (1385-1395).
// RDS baseband (fm.go:234-271, carried across blocks): mono * e^{j phi_c}
//   with phi_c = 2pi ((c+1) 57000 mod fs)/fs for stream sample c, 101-tap
//   LPF7.5k (real taps on I and Q), keep stream samples c = 0 mod decim.
//
// Arithmetic: FIRs accumulate in FP32 (the reference accumulates in FP64),
// the PLL state and loop arithmetic are FP64 as in the reference, the
// oscillators use the hardware turn-based sine/cosine, and the phase
// detector's atan2 is the FP32 polynomial of pll_atan2.hpp (3.2e-7 rad).
#include "wfm.hpp"
#include "pll_atan2.hpp"
#include "runner.hpp"

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kInvTwoPi = 0.15915494309189533576888376337251;
constexpr double kInvPi = 0.31830988618379067153776752674503;

__device__ __forceinline__ float mono_at(const ArgusWfmDesc& d, int idx)  // idx >= -ARGUS_WFM_HIST
{
    return idx >= 0 ? d.mono[idx] : d.mono_hist_rd[ARGUS_WFM_HIST + idx];
}

// Input of the L-R low-pass at block sample idx: bpf * 2 sin(2 phase).
__device__ __forceinline__ float lr_at(const ArgusWfmDesc& d, int idx)
{
    if (idx < 0) return d.lr_hist_rd[ARGUS_WFM_HIST + idx];
    const float* bpf = d.scratch + d.n;
    const float* t2 = d.scratch + 3 * d.n;
    return bpf[idx] * (2.0f * __builtin_amdgcn_sinf(t2[idx]));
}

// Input of the mono fallback low-pass at block sample idx.
__device__ __forceinline__ float fb_at(const ArgusWfmDesc& d, int fb_rd, int idx)
{
    return idx >= 0 ? d.mono[idx] : d.fb_hist[fb_rd * ARGUS_WFM_HIST + ARGUS_WFM_HIST + idx];
}

// atan2 of the PLL's phase detector (pll_atan2.hpp) on the hardware
// reciprocal.
__device__ __forceinline__ float pll_atan2f(float y, float x)
{
    return argus_pll_atan2f(y, x, [](float v) { return __builtin_amdgcn_rcpf(v); });
}

}  // namespace

__device__ __forceinline__ void wfm_pre(const ArgusWfmDesc* __restrict__ descs, const ArgusWfmItem* __restrict__ items,
                                        int n_items)
{
    __shared__ float xs[ARGUS_WFM_TILE + ARGUS_WFM_HIST];
    __shared__ float2 ms[ARGUS_WFM_TILE + ARGUS_WFM_HIST];
    __shared__ float hs[ARGUS_WFM_NTAPSETS * ARGUS_WFM_TAPS];
    for (int w = blockIdx.x; w < n_items; w += gridDim.x) {
        const ArgusWfmItem it = items[w];
        const ArgusWfmDesc& d = descs[it.chan];
        const int n = d.n;
        if (it.kind == ARGUS_WFM_ITEM_TAIL) {
            for (int j = threadIdx.x; j < ARGUS_WFM_HIST; j += blockDim.x)
                d.mono_hist_wr[j] = mono_at(d, n - ARGUS_WFM_HIST + j);
            continue;
        }
        const int i0 = it.start;
        const int cnt = min(ARGUS_WFM_TILE, n - i0);
        const int span = cnt + ARGUS_WFM_HIST;
        for (int j = threadIdx.x; j < ARGUS_WFM_NTAPSETS * ARGUS_WFM_TAPS; j += blockDim.x) hs[j] = d.taps[j];
        for (int j = threadIdx.x; j < span; j += blockDim.x) {
            const int idx = i0 - ARGUS_WFM_HIST + j;
            const float v = mono_at(d, idx);
            xs[j] = v;
            float s, c;
            argus_nco(d.mix0 + d.mix_inc * (uint64_t)(int64_t)idx, s, c);
            ms[j] = make_float2(v * c, v * s);
        }
        __syncthreads();

        float* lpr = d.scratch;
        float* bpf = d.scratch + n;
        float* pil = d.scratch + 2 * n;
        const float* h0 = hs + ARGUS_WFM_LPF * ARGUS_WFM_TAPS;
        const float* h1 = hs + ARGUS_WFM_BPF * ARGUS_WFM_TAPS;
        const float* h2 = hs + ARGUS_WFM_PILOT * ARGUS_WFM_TAPS;
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
            const float* xp = xs + i + ARGUS_WFM_HIST;
            float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
#pragma unroll 4
            for (int k = 0; k < ARGUS_WFM_TAPS; k++) {
                const float x = xp[-k];
                a0 = fmaf(h0[k], x, a0);
                a1 = fmaf(h1[k], x, a1);
                a2 = fmaf(h2[k], x, a2);
            }
            lpr[i0 + i] = a0;
            bpf[i0 + i] = a1;
            pil[i0 + i] = a2;
        }

        // RDS outputs whose sample index c = rds_first + r * rds_decim falls in this tile
        const float* h3 = hs + ARGUS_WFM_RDS * ARGUS_WFM_TAPS;
        const int D = d.rds_decim;
        const int r0 = i0 <= d.rds_first ? 0 : (i0 - d.rds_first + D - 1) / D;
        for (int r = r0 + (int)threadIdx.x; r < d.n_rds; r += blockDim.x) {
            const int c = d.rds_first + r * D;
            if (c >= i0 + cnt) break;
            const float2* mp = ms + (c - i0) + ARGUS_WFM_HIST;
            float ar = 0.0f, ai = 0.0f;
#pragma unroll 4
            for (int k = 0; k < ARGUS_WFM_TAPS; k++) {
                const float2 v = mp[-k];
                ar = fmaf(h3[k], v.x, ar);
                ai = fmaf(h3[k], v.y, ai);
            }
            d.rds[r] = make_float2(ar, ai);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void wfm_pll_one(const ArgusWfmDesc& d)
{
    // One wavefront (the whole 64-thread workgroup) per station. Every lane
    // runs the same serial recurrence: the pilot sample of step j comes from
    // lane j of a 64-sample register tile (readlane, no LDS round trip), lane
    // j keeps the phase after step j, and the lanes turn their phases into
    // the L-R carrier argument together after each tile.
    __shared__ double tp[64];
    const int n = d.n;
    const int lane = threadIdx.x;
    ArgusWfmState* sp = d.st;
    if (n <= 0) {  // no samples: state and lock counters unchanged (stereoDecodeStateful returns early)
        if (lane == 0) {
            ArgusWfmBlock b;
            b.locked = 0;
            b.stereo = sp->enabled;
            b.fb_rd = sp->fb_sel;
            b.fb_wr = sp->fb_sel;
            *d.blk = b;
        }
        return;
    }
    double phase = sp->phase, freq = sp->freq, is = sp->i_state, qs = sp->q_state;
    const double alpha = d.alpha, beta = d.beta, lpa = d.lpa, min_f = d.min_f, max_f = d.max_f;
    double es = 0.0;    // sum of |err|, in sample order (every lane holds it)
    double pp_l = 0.0;  // this lane's share of the pilot power
    if (d.batch) {  // total power of the discriminator output, for the batch lock test
        double acc = 0.0;
        for (int j = lane; j < n; j += 64) {
            const double m = (double)d.mono[j];
            acc += m * m;
        }
        tp[lane] = acc;
    }
    typedef __attribute__((address_space(1))) const float GlobalF;  // global loads: no LDS counter to wait on
    const GlobalF* pilot = (const GlobalF*)(d.scratch + 2 * n);
    float* t2 = d.scratch + 3 * n;
    float pv = 0.0f;
    double keep = 0.0;
    // One PLL step (streamer.go:1588-1611) on the pilot sample in lane j of pv.
    auto step = [&](int j) {
        const double p = (double)__int_as_float(__builtin_amdgcn_readlane(__float_as_int(pv), j));
        const float ph_t = (float)(phase * kInvTwoPi);  // phase in [0, 2pi] -> [0, 1] turns
        const double sn = (double)__builtin_amdgcn_sinf(ph_t);
        const double cs = (double)__builtin_amdgcn_cosf(ph_t);
        is += lpa * (p * cs - is);
        qs += lpa * (p * -sn - qs);
        const double err = (double)pll_atan2f((float)qs, (float)is);
        freq += beta * err;
        // clamp to [min_f, max_f] with the bare max/min instructions (fmax and
        // fmin would add canonicalising copies of the bounds to every step)
        asm("v_max_f64 %0, %1, %2" : "=v"(freq) : "v"(freq), "v"(min_f));
        asm("v_min_f64 %0, %1, %2" : "=v"(freq) : "v"(freq), "v"(max_f));
        // phase += step, then -2pi if above 2pi, else +2pi if below 0: one
        // FMA with k in {1, -1, 0} rounds as the subtraction or addition does
        const double ph = phase + (freq + alpha * err);
        const double k = ph > kTwoPi ? 1.0 : (ph < 0.0 ? -1.0 : 0.0);
        phase = fma(-kTwoPi, k, ph);
        es += fabs(err);
        keep = lane == j ? phase : keep;
    };
    float pv_next = lane < n ? pilot[lane] : 0.0f;
    // Land the first tile here: the inner loop then enters with only the
    // next tile's prefetch in flight, which it does not read.
    __builtin_amdgcn_s_waitcnt(0);
    for (int i0 = 0; i0 < n; i0 += 64) {
        const int cnt = min(64, n - i0);
        pv = pv_next;  // 0 beyond the end
        pv_next = i0 + 64 + lane < n ? pilot[i0 + 64 + lane] : 0.0f;
        pp_l += (double)pv * (double)pv;
        if (cnt == 64) {
#pragma unroll 8
            for (int j = 0; j < 64; j++) step(j);
        } else {
            for (int j = 0; j < cnt; j++) step(j);
        }
        double u = keep * kInvPi;  // sin(2 phase) = sin(2 pi frac(phase / pi))
        u -= floor(u);
        if (lane < cnt) t2[i0 + lane] = (float)u;
    }
    double pp = pp_l;
    for (int o = 32; o > 0; o >>= 1) pp += __shfl_xor(pp, o);
    __syncthreads();  // tp
    if (lane == 0 && d.batch) {
        // wfm_batch_stereo.go:139-145: one decision for the whole recording
        double total = 0.0;
        for (int j = 0; j < 64; j++) total += tp[j];
        const double ratio = total > 0.0 ? pp / total : 0.0;
        const double f_hz = freq * d.rate / kTwoPi;
        const int locked = ratio > 0.003 && fabs(f_hz - 19000.0) < 250.0 && es / (double)n < 0.35;
        ArgusWfmBlock b;
        b.locked = locked;
        b.stereo = d.batch == ARGUS_WFM_BATCH_MATRIX ? 1 : locked;
        b.fb_rd = 0;
        b.fb_wr = 0;
        *d.blk = b;
    } else if (lane == 0) {
        ArgusWfmState st = *sp;
        st.phase = phase;
        st.freq = freq;
        st.i_state = is;
        st.q_state = qs;
        st.err_avg = 0.9 * st.err_avg + 0.1 * (es / (double)n);
        const double coherence = pp > 1e-12 ? (is * is + qs * qs) * (double)n / pp : 0.0;
        const double f_hz = freq * d.rate / kTwoPi;
        const int locked = coherence > 0.15 && fabs(f_hz - 19000.0) < 250.0 && st.err_avg < 0.35;
        if (locked) {
            st.on_count++;
            st.off_count = 0;
            if (st.on_count >= 4) st.enabled = 1;
        } else {
            st.on_count = 0;
            st.off_count++;
            if (st.off_count >= 10) st.enabled = 0;
        }
        ArgusWfmBlock b;
        b.locked = locked;
        b.stereo = st.enabled;
        b.fb_rd = st.fb_sel;
        b.fb_wr = st.fb_sel ^ 1;
        if (!st.enabled) st.fb_sel ^= 1;  // the fallback filter ran: its new delay line is the other half
        *d.blk = b;
        *sp = st;
    }
}

__device__ __forceinline__ void wfm_post(const ArgusWfmDesc* __restrict__ descs, const ArgusWfmItem* __restrict__ items,
                                         int n_items)
{
    __shared__ float xs[ARGUS_WFM_TILE + ARGUS_WFM_HIST];
    __shared__ float hs[ARGUS_WFM_TAPS];
    for (int w = blockIdx.x; w < n_items; w += gridDim.x) {
        const ArgusWfmItem it = items[w];
        const ArgusWfmDesc& d = descs[it.chan];
        const ArgusWfmBlock b = *d.blk;
        const int n = d.n;
        if (it.kind == ARGUS_WFM_ITEM_TAIL) {
            for (int j = threadIdx.x; j < ARGUS_WFM_HIST; j += blockDim.x) {
                const int idx = n - ARGUS_WFM_HIST + j;
                d.lr_hist_wr[j] = lr_at(d, idx);
                if (!b.stereo) d.fb_hist[b.fb_wr * ARGUS_WFM_HIST + j] = fb_at(d, b.fb_rd, idx);
            }
            continue;
        }
        const int i0 = it.start;
        const int cnt = min(ARGUS_WFM_TILE, n - i0);
        if (d.batch && !b.stereo) {  // unlocked recording: dual mono of the raw discriminator
            for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
                const float m = d.mono[i0 + i];
                d.audio[2 * (i0 + i)] = m;
                d.audio[2 * (i0 + i) + 1] = m;
            }
            continue;
        }
        const int span = cnt + ARGUS_WFM_HIST;
        // the L-R low-pass and the fallback low-pass are both LowpassFIR(15000)
        for (int j = threadIdx.x; j < ARGUS_WFM_TAPS; j += blockDim.x) hs[j] = d.taps[ARGUS_WFM_LPF * ARGUS_WFM_TAPS + j];
        for (int j = threadIdx.x; j < span; j += blockDim.x) {
            const int idx = i0 - ARGUS_WFM_HIST + j;
            xs[j] = b.stereo ? lr_at(d, idx) : fb_at(d, b.fb_rd, idx);
        }
        __syncthreads();
        const float* lpr = d.scratch;
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
            const float* xp = xs + i + ARGUS_WFM_HIST;
            float a = 0.0f;
#pragma unroll 4
            for (int k = 0; k < ARGUS_WFM_TAPS; k++) a = fmaf(hs[k], xp[-k], a);
            float left = a, right = a;
            if (b.stereo) {
                const float l = lpr[i0 + i];
                left = 0.5f * (l + a);
                right = 0.5f * (l - a);
            }
            d.audio[2 * (i0 + i)] = left;
            d.audio[2 * (i0 + i) + 1] = right;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void argus_wfm_pre_kernel(const ArgusWfmDesc* __restrict__ descs,
                                                            const ArgusWfmItem* __restrict__ items, int n_items)
{
    wfm_pre(descs, items, n_items);
}

__global__ __launch_bounds__(64) void argus_wfm_pll_kernel(const ArgusWfmDesc* __restrict__ descs, int n_chans)
{
    for (int c = blockIdx.x; c < n_chans; c += gridDim.x) {
        wfm_pll_one(descs[c]);
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void argus_wfm_post_kernel(const ArgusWfmDesc* __restrict__ descs,
                                                             const ArgusWfmItem* __restrict__ items, int n_items)
{
    wfm_post(descs, items, n_items);
}

__global__ __launch_bounds__(256) void argus_wfm_pre_hdr_kernel(const ArgusWfmHdr* __restrict__ hdr)
{
    const ArgusWfmHdr h = *hdr;
    wfm_pre(h.descs, h.items, h.n_items);
}

__global__ __launch_bounds__(64) void argus_wfm_pll_hdr_kernel(const ArgusWfmHdr* __restrict__ hdr)
{
    const ArgusWfmHdr h = *hdr;
    for (int c = blockIdx.x; c < h.n_chans; c += gridDim.x) {
        wfm_pll_one(h.descs[c]);
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void argus_wfm_post_hdr_kernel(const ArgusWfmHdr* __restrict__ hdr)
{
    const ArgusWfmHdr h = *hdr;
    wfm_post(h.descs, h.items, h.n_items);
}

// L+R low-pass LowpassFIR(15k); L-R band-pass LowpassFIR(53k) - LowpassFIR(23k);
// pilot band-pass LowpassFIR(21k) - LowpassFIR(17k) (streamer.go:1532-1549,
// wfm_batch_stereo.go:81-98); RDS low-pass LowpassFIR(7.5k) (fm.go:254).
void argus_wfm_design_taps(int rate, float* taps)
{
    const double cut[6] = {15000.0, 53000.0, 23000.0, 21000.0, 17000.0, 7500.0};
    double h[6][ARGUS_WFM_TAPS + 1];
    for (int i = 0; i < 6; i++) argus_lowpass_fir(cut[i], rate, ARGUS_WFM_TAPS, h[i], false);
    for (int k = 0; k < ARGUS_WFM_TAPS; k++) {
        taps[ARGUS_WFM_LPF * ARGUS_WFM_TAPS + k] = (float)h[0][k];
        taps[ARGUS_WFM_BPF * ARGUS_WFM_TAPS + k] = (float)(h[1][k] - h[2][k]);
        taps[ARGUS_WFM_PILOT * ARGUS_WFM_TAPS + k] = (float)(h[3][k] - h[4][k]);
        taps[ARGUS_WFM_RDS * ARGUS_WFM_TAPS + k] = (float)h[5][k];
    }
}

// pllCoefficients(50, 0.707) (streamer.go:1496-1506; the batch decoder's
// pllCoefficientsBatch, wfm_batch_stereo.go:159-169, is the same), the 200 Hz
// arm low-pass and the 17-21 kHz frequency clamp.
void argus_wfm_loop_constants(int rate, ArgusWfmDesc& wd)
{
    const double loop_bw = 50.0, damping = 0.707;
    const double bl = loop_bw / (double)rate;
    const double theta = bl / (damping + 0.25 / damping);
    const double dd = 1 + 2 * damping * theta + theta * theta;
    wd.alpha = (4 * damping * theta) / dd;
    wd.beta = (4 * theta * theta) / dd;
    wd.lpa = 1.0 - std::exp(-2.0 * M_PI * 200.0 / (double)rate);
    wd.min_f = 2.0 * M_PI * 17000.0 / (double)rate;
    wd.max_f = 2.0 * M_PI * 21000.0 / (double)rate;
    wd.rate = (double)rate;
}

int argus_launch_wfm_hdr(const ArgusWfmHdr* d_hdr, hipStream_t stream)
{
    argus_clear_error();
    hipLaunchKernelGGL(argus_wfm_pre_hdr_kernel, dim3(ARGUS_WFM_GRID), dim3(256), 0, stream, d_hdr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(argus_wfm_pll_hdr_kernel, dim3(ARGUS_WFM_GRID), dim3(64), 0, stream, d_hdr);
    e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    hipLaunchKernelGGL(argus_wfm_post_hdr_kernel, dim3(ARGUS_WFM_GRID), dim3(256), 0, stream, d_hdr);
    return (int)hipGetLastError();
}

int argus_launch_wfm(const ArgusWfmDesc* d_desc, int n_chans, const ArgusWfmItem* d_items, int n_items,
                     hipStream_t stream)
{
    if (n_chans <= 0) return 0;
    const int grid = n_items < 2048 ? n_items : 2048;
    argus_clear_error();
    if (n_items > 0) {
        hipLaunchKernelGGL(argus_wfm_pre_kernel, dim3(grid), dim3(256), 0, stream, d_desc, d_items, n_items);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL(argus_wfm_pll_kernel, dim3(n_chans), dim3(64), 0, stream, d_desc, n_chans);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    if (n_items > 0) {
        hipLaunchKernelGGL(argus_wfm_post_kernel, dim3(grid), dim3(256), 0, stream, d_desc, d_items, n_items);
        e = hipGetLastError();
        if (e != hipSuccess) return (int)e;
    }
    return 0;
}
