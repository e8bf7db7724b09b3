// surv.hip -- wideband surveillance: windowed FFT / Welch PSD, CFAR and the
// detection list (successors of internal/fft, internal/cfar and the
// threshold path of internal/detector).
//
// FFT of N = R^3 points (R = 32: 32768; R = 16: 4096) by one workgroup of
// R^2 threads.  Index n = a + R b + R^2 c; thread (a, b) holds the R samples
// c = 0..R-1 in registers:
//   pass 1  DFT_R over c, twiddle W_N^{(a + R b) kc}
//   LDS transpose -> thread (a, kc) holds b = 0..R-1
//   pass 2  DFT_R over b, twiddle W_{R^2}^{a kb}
//   LDS transpose (XOR-swizzled) -> thread (kb, kc) holds a = 0..R-1
//   pass 3  DFT_R over a  ->  X[kc + R kb + R^2 ka]
// Each transpose moves the real parts, then the imaginary parts, through a
// 4N-byte LDS array, so a 32768-point transform needs 128 KiB of LDS and no
// HBM round trip.  Welch segments of one output frame are transformed back to
// back by the same workgroup and their |X|^2 accumulates in registers.
#include "common.hpp"
#include "fft_core.hpp"
#include "surv.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

int argus_lowpass_fir(double cutoff_hz, int sample_rate, int taps, double* out, bool hamming);

namespace {

using namespace argus_fftc;


template <int R>
__global__ __launch_bounds__(R* R) void argus_welch_kernel(const float2* __restrict__ in, int64_t frame_step,
                                                           int seg_hop, int n_segs, const float* __restrict__ win,
                                                           float* __restrict__ out_db)
{
    constexpr int N = R * R * R, TH = R * R, PL = R * (R + 1);
    __shared__ float xs[fft_lds_floats<R>()];
    __shared__ float2 tw1[fft_tab_entries<R>()], tw2[fft_tab_entries<R>()];
    const int t = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * frame_step;
    const float invN = 1.0f / (float)N;
    fft_tab_init<R>(tw1, tw2, t, TH);
    __syncthreads();
    float acc[R];
#pragma unroll
    for (int k = 0; k < R; ++k) acc[k] = 0.0f;
    for (int seg = 0; seg < n_segs; ++seg) {
        const float2* x = in + base + (int64_t)seg * seg_hop;
        float2 v[R];
        int tl = t;
        asm volatile("" : "+v"(tl));  // per segment, so window loads are not hoisted into scratch
#pragma unroll
        for (int c = 0; c < R; ++c) {
            const int idx = tl + TH * c;
            const float2 s = x[idx];
            const float w = win[idx];
            v[c] = make_float2(s.x * w, s.y * w);
        }
        fft_r3_tab<R>(v, xs, (int)threadIdx.x, tw1, tw2);
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const float mr = v[k].x * invN, mi = v[k].y * invN;
            acc[k] = fmaf(mr, mr, fmaf(mi, mi, acc[k]));
        }
    }
    // thread t = kb + R kc holds bin k = kc + R kb + R^2 ka; fftshift puts it
    // at i = kc + R kb + R^2 ((ka + R/2) mod R).  Stage through the padded
    // LDS image (odd stride R+1 between lanes) for coalesced stores.
    const int kb = t % R, kc = t / R;
    const float inv_segs = 1.0f / (float)(n_segs > 0 ? n_segs : 1);
#pragma unroll
    for (int ka = 0; ka < R; ++ka) {
        const int q = (ka + R / 2) % R;
        const float db = 10.0f * log10f(acc[ka] * inv_segs + 1e-24f);
        // pipeline_runtime.go:391-397 sanitizeSpectrum: NaN / Inf -> -200 dB
        xs[kc + (R + 1) * kb + PL * q] = isfinite(db) ? db : -200.0f;
    }
    __syncthreads();
    float* o = out_db + (int64_t)blockIdx.x * N;
    const int jc = t % R, jb = t / R;
#pragma unroll
    for (int q = 0; q < R; ++q) o[t + TH * q] = xs[jc + (R + 1) * jb + PL * q];
}

// Plain forward transform of one frame, natural-order complex output
// (the successor of cufftExecC2C(CUFFT_FORWARD), gpufft/gpu.go:67-91),
// optionally windowed.
template <int R>
__global__ __launch_bounds__(R* R) void argus_fft_kernel(const float2* __restrict__ in, const float* __restrict__ win,
                                                         float2* __restrict__ out, int64_t frame_step)
{
    constexpr int N = R * R * R, TH = R * R, PL = R * (R + 1);
    __shared__ float xs[fft_lds_floats<R>()];
    __shared__ float2 tw1[fft_tab_entries<R>()], tw2[fft_tab_entries<R>()];
    const int t = threadIdx.x;
    const float2* x = in + (int64_t)blockIdx.x * frame_step;
    fft_tab_init<R>(tw1, tw2, t, TH);
    float2 v[R];
#pragma unroll
    for (int c = 0; c < R; ++c) {
        const int idx = t + TH * c;
        float2 s = x[idx];
        if (win) {
            const float w = win[idx];
            s = make_float2(s.x * w, s.y * w);
        }
        v[c] = s;
    }
    __syncthreads();  // the twiddle tables are complete
    fft_r3_tab<R>(v, xs, (int)threadIdx.x, tw1, tw2);
    const int kb = t % R, kc = t / R;
    float2* o = out + (int64_t)blockIdx.x * N;
    const int jc = t % R, jb = t / R;
    // real parts then imaginary parts through the padded LDS image, each
    // read back into the registers it came from
#pragma unroll
    for (int ka = 0; ka < R; ++ka) xs[kc + (R + 1) * kb + PL * ka] = v[ka].x;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R; ++q) v[q].x = xs[jc + (R + 1) * jb + PL * q];
    __syncthreads();
#pragma unroll
    for (int ka = 0; ka < R; ++ka) xs[kc + (R + 1) * kb + PL * ka] = v[ka].y;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R; ++q) o[t + TH * q] = make_float2(v[q].x, xs[jc + (R + 1) * jb + PL * q]);
}

// Single-FFT spectrum in dB (fft.go:58-73 SpectrumFromFFT): bin i of the
// output is FFT bin (i + n/2) mod n, 20 log10(|X| / n + 1e-12), with NaN /
// Inf mapped to -200 dB (pipeline_runtime.go:391-397).  Frame f's n bins
// are the first n of X + f * x_stride.
__global__ __launch_bounds__(256) void argus_spectrum_db_kernel(const float2* __restrict__ X, int n, int64_t x_stride,
                                                                int64_t total, float* __restrict__ out)
{
    const double inv_n = 1.0 / (double)n;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = g / n;
        const int i = (int)(g - f * n);
        const float2 v = X[f * x_stride + (i + n / 2) % n];
        const double mag = hypot((double)v.x, (double)v.y) * inv_n;
        const double db = 20.0 * log10(mag + 1e-12);
        out[g] = isfinite(db) ? (float)db : -200.0f;
    }
}

// A decimated surveillance level (pipeline_runtime.go:399-426, 624-645):
// output m of frame f is ApplyFIR's y[m * decim] over the frame's window
// (real FP64 taps, FP64 accumulation from zero history, fir.go:30-53, then
// Decimate, fir.go:84-95), Hann-windowed over the level's n_d samples.  It
// is written at z[f * N + m * decim] with zeros between, N = n_d * decim:
// the N-point transform of that zero-stuffed frame is the level's
// n_d-point transform repeated decim times, so the engine's own N-point FFT
// serves every level.
//
// One workgroup per 256 outputs of a frame: the input span they read
// (256 * decim + taps - 1 samples, zero before the window's start, which
// adds nothing to an accumulation exactly as ApplyFIR's early break),
// converted to FP64 once, and the taps are staged in LDS (dynamic, sized to
// the span: level_fir_lds_bytes), then each thread accumulates its output
// over the taps.
constexpr int kLevelOut = 256;
constexpr int kLevelMaxTaps = 128;
constexpr int kLevelMaxDecim = 16;

// Positions per residue plane of the span, and the kernel's LDS bytes.
static inline int level_fir_planes(int decim, int n_taps)
{
    return (kLevelOut * decim + n_taps - 1 + decim - 1) / decim;
}
static inline size_t level_fir_lds_bytes(int decim, int n_taps)
{
    return (size_t)level_fir_planes(decim, n_taps) * decim * sizeof(double2) + (size_t)n_taps * sizeof(double);
}

__global__ __launch_bounds__(kLevelOut) void argus_level_fir_kernel(const float2* __restrict__ in, int64_t frame_step,
                                                                    int decim, int n_d,
                                                                    const double* __restrict__ taps, int n_taps,
                                                                    const float* __restrict__ win,
                                                                    float2* __restrict__ z)
{
#pragma clang fp contract(off)
    // the span by polyphase residue, in FP64: sample j (input base + j) at
    // plane j mod decim, position j / decim, planes P positions apart; then
    // the taps
    extern __shared__ double2 level_lds[];
    const int f = blockIdx.y;
    const int m0 = blockIdx.x * kLevelOut;
    const float2* x = in + (int64_t)f * frame_step;
    const int base = m0 * decim - (n_taps - 1);  // input index of span sample 0
    const int span = kLevelOut * decim + n_taps - 1;
    const int ld = __builtin_ctz(decim);  // decim is a power of two
    const int P = (span + decim - 1) >> ld;
    double2* xs = level_lds;
    double* hs = (double*)(level_lds + (size_t)P * decim);
    for (int j = threadIdx.x; j < span; j += kLevelOut) {
        const int idx = base + j;
        const int lim = n_d * decim;
        const float2 v = (idx >= 0 && idx < lim) ? x[idx] : make_float2(0.0f, 0.0f);
        xs[(j & (decim - 1)) * P + (j >> ld)] = make_double2((double)v.x, (double)v.y);
    }
    // the taps grouped the same way: residue rho's taps k = k0(rho) + i decim
    // (k0 = (T - 1 - rho) mod decim) at hs[off(rho) + i], off the count of
    // taps of the residues before it
    for (int k = threadIdx.x; k < n_taps; k += kLevelOut) {
        const int rho = (n_taps - 1 - k) & (decim - 1);
        int off = 0;
        for (int r = 0; r < rho; r++) {
            const int k0 = (n_taps - 1 - r) & (decim - 1);
            off += k0 < n_taps ? (n_taps - 1 - k0) / decim + 1 : 0;
        }
        hs[off + (k >> ld)] = taps[k];
    }
    __syncthreads();
    const int m = m0 + threadIdx.x;
    if (m >= n_d) return;
    // y[i] = sum_k h[k] x[i - k], i = m * decim: span sample t decim + jj,
    // jj = T - 1 - k.  The taps are summed grouped by the sample's residue
    // rho = jj mod decim: those read plane rho at positions t + (jj - rho) /
    // decim, consecutive over the lanes (no bank conflicts) and a fixed
    // offset per tap, and their coefficients in a row.  FP64 throughout; the
    // grouping changes only the last bits of the FP64 sum, far below the
    // FP32 transform that follows.
    double ar = 0.0, ai = 0.0;
    const double* hr = hs;
    for (int rho = 0; rho < decim; rho++) {
        const int k0 = (n_taps - 1 - rho) & (decim - 1);  // first tap of this residue
        if (k0 >= n_taps) continue;
        const int cnt = (n_taps - 1 - k0) / decim + 1;
        const double2* xp = xs + rho * P + (int)threadIdx.x + ((n_taps - 1 - k0 - rho) >> ld);
#pragma unroll 4
        for (int i = 0; i < cnt; i++) {
            const double2 v = xp[-i];
            const double w = hr[i];
            ar += v.x * w;
            ai += v.y * w;
        }
        hr += cnt;
    }
    const float wd = win[m];
    float2* o = z + (int64_t)f * n_d * decim + (int64_t)m * decim;
    o[0] = make_float2((float)ar * wd, (float)ai * wd);
    for (int j = 1; j < decim; j++) o[j] = make_float2(0.0f, 0.0f);
}

// ---------------------------------------------------------------------------
// CFAR in three launches over all frames of a call:
//   argus_cfar_prep_kernel    per frame: its largest dB value (which bounds
//                             the reference's running sums, for the near-tie
//                             count) and the count cleared;
//   argus_cfar_status_kernel  per (frame, tile of kCfarTile bins): the tile's
//                             cells and halo in linear power, their FP64
//                             prefix sum (wavefront scans), each bin's window
//                             sums as two differences of it, and the bin's
//                             decision as bits of 64-bin masks;
//   argus_cfar_runs_kernel    per frame: the hysteresis state across the
//                             masks (a scan of per-thread transfer functions),
//                             the runs, their peaks and noise estimates.
// ---------------------------------------------------------------------------
constexpr int kCfarThreads = 256;
constexpr int kCfarTile = 1024;  // bins per workgroup of the status kernel
constexpr double kLn10Over10 = 0.23025850929940458;  // 10^(x/10) = exp(x ln10 / 10)

__device__ __forceinline__ int wrap_bin(int i, int n, int wrap)
{
    if (wrap) return ((i % n) + n) % n;
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// The noise estimate in the reference's operation order: the sums times the
// reciprocal of the cell count (ca.go:57-58, gosca.go:55-57, caso.go).
__device__ __forceinline__ double cfar_combine(int mode, double left, double right, int train)
{
    if (mode == GPUD_CFAR_CA) return (left + right) * (1.0 / (double)(2 * train));
    const double inv = 1.0 / (double)train;
    const double lm = left * inv, rm = right * inv;
    if (mode == GPUD_CFAR_GOSCA) return lm > rm ? lm : rm;
    return lm < rm ? lm : rm;  // CASO
}

// Near-tie test against the reference's running sums.  The reference slides
// its window sums along the frame (ca.go:61-66: per bin one subtraction and
// one addition per side), so at bin i each side carries at most 2 (i + t)
// roundings of partial sums no larger than t * lmax (lmax the frame's largest
// linear cell): |error| <= 2 u (i + t) t lmax per side, u = 2^-53.  Here each
// side is a difference of two prefix sums, within `ours` of its exact value.
// Returns the relative uncertainty of the noise estimate, plus a slack for
// last-bit differences between the reference's pow / log10 / dB threshold
// and this kernel's exp and comparison in linear power.
__device__ __forceinline__ double cfar_tie_rel(int mode, double left, double right, int train, int i, double lmax,
                                               double ours)
{
    const double u = 1.1102230246251565e-16;
    const double side = 2.0 * u * (double)(i + 2 * train + 2) * (double)train * lmax + ours;
    double ref;
    if (mode == GPUD_CFAR_CA) ref = left + right;
    else if (mode == GPUD_CFAR_GOSCA) ref = left > right ? left : right;
    else ref = left < right ? left : right;
    return (mode == GPUD_CFAR_CA ? 2.0 * side : side) / ref + 2.5e-13;
}

// OS-CFAR (os.go:31-80): the rank-th smallest (0-based rk) of the 2*train
// training cells in dB.  The window is a multiset: cell c is the answer when
// fewer than rk+1 cells are below it and at least rk+1 are at or below it,
// which is the sorted window's element rk whatever the order of ties.
template <typename Cell>
__device__ __forceinline__ double os_select(Cell cell, int train, int rk)
{
    const int total = 2 * train;
    double best = 0.0;
    for (int c = 0; c < total; ++c) {
        const double v = cell(c);
        int less = 0, leq = 0;
        for (int m = 0; m < total; ++m) {
            const double u = cell(m);
            less += u < v;
            leq += u <= v;
        }
        if (less <= rk && rk < leq) best = v;
    }
    return best;
}

__device__ __forceinline__ int os_rank(const gpud_cfar_params& p)
{
    int rk = p.rank - 1;
    if (rk < 0) rk = 0;
    if (rk >= 2 * p.train) rk = 2 * p.train - 1;
    return rk;
}

// Threshold of one bin straight from the dB spectrum (used for peak bins).
__device__ double cfar_threshold_at(const float* __restrict__ db, int n, int i, const gpud_cfar_params& p)
{
    if (p.mode == GPUD_CFAR_OS) {
        auto cell = [&](int m) {
            const int k = m / 2 + 1;
            return (double)db[wrap_bin((m & 1) ? i + p.guard + k : i - p.guard - k, n, p.wrap)];
        };
        return os_select(cell, p.train, os_rank(p)) + p.scale_db;
    }
    double left = 0.0, right = 0.0;
    for (int k = 1; k <= p.train; ++k) {
        left += pow(10.0, (double)db[wrap_bin(i - p.guard - k, n, p.wrap)] / 10.0);
        right += pow(10.0, (double)db[wrap_bin(i + p.guard + k, n, p.wrap)] / 10.0);
    }
    return 10.0 * log10(cfar_combine(p.mode, left, right, p.train)) + p.scale_db;
}

// Inclusive prefix sum over the 64 lanes of a wavefront.
__device__ __forceinline__ double wave_scan_add(double x, int lane)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

__device__ __forceinline__ int wave_scan_add(int x, int lane)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

__global__ __launch_bounds__(256) void argus_cfar_prep_kernel(const float* __restrict__ psd, int n,
                                                              float* __restrict__ frame_max,
                                                              int* __restrict__ ambiguous)
{
    __shared__ float red[4];
    const float* db = psd + (int64_t)blockIdx.x * n;
    float m = -INFINITY;
    for (int b = threadIdx.x; b < n; b += 256) m = fmaxf(m, db[b]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        frame_max[blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        ambiguous[blockIdx.x] = 0;
    }
}

// Bin b's decision, as the reference takes it from its threshold thr_b =
// 10 log10(noise_b) + scale (ca.go:57-66, gosca.go, caso.go) and the
// crossing with hysteresis (detector.go:436-480): 2 if v >= thr_b, 0 if v <
// thr_b - hysteresis, else 1.  In linear power these are 10^(v/10) >=
// noise_b * k_on and 10^(v/10) < noise_b * k_off, k_on = 10^(scale/10),
// k_off = 10^((scale - hysteresis)/10), so no bin needs a logarithm.  Masks:
// for each 64-bin group (x: code 2, y: code 0).
__global__ __launch_bounds__(kCfarThreads) void argus_cfar_status_kernel(const float* __restrict__ psd, int n,
                                                                         gpud_cfar_params p, double k_on, double k_off,
                                                                         const float* __restrict__ frame_max,
                                                                         int* __restrict__ ambiguous,
                                                                         ulonglong2* __restrict__ masks)
{
    extern __shared__ double cf_lds[];
    __shared__ double wtot[kCfarThreads / 64];
    __shared__ int wties[kCfarThreads / 64];
    const int halo = p.guard + p.train;
    const int E = kCfarTile + 2 * halo;  // cells: the tile and its halo on both sides
    double* cell = cf_lds;               // E cells, linear power (dB for OS-CFAR)
    double* pre = cf_lds + E;            // pre[e] = cell[0] + ... + cell[e - 1], e = 0..E
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int tiles = n / kCfarTile, f = blockIdx.x / tiles, t0 = (blockIdx.x % tiles) * kCfarTile;
    const float* db = psd + (int64_t)f * n;
    const bool os = p.mode == GPUD_CFAR_OS;  // OS ranks the dB values themselves
    const bool ties = ambiguous != nullptr && !os;

    for (int e = t; e < E; e += kCfarThreads) {  // cell e is bin t0 - halo + e
        const double v = (double)db[wrap_bin(t0 - halo + e, n, p.wrap)];
        cell[e] = os ? v : exp(v * kLn10Over10);
    }
    __syncthreads();
    // Prefix sum: wave w scans its rows of 64 consecutive cells in order,
    // carrying each row's total into the next; the waves' totals are added
    // afterwards.  Every partial sum is of non-negative terms no larger than
    // the tile's total, and each entry takes at most rpw + 11 roundings.
    const int rows = (E + 63) >> 6, rpw = (rows + kCfarThreads / 64 - 1) / (kCfarThreads / 64);
    const int r0 = wave * rpw, r1 = min(r0 + rpw, rows);
    if (!os) {
        double carry = 0.0;
        for (int r = r0; r < r1; ++r) {
            const int e = (r << 6) + lane;
            const double v = wave_scan_add(e < E ? cell[e] : 0.0, lane) + carry;
            if (e < E) pre[e + 1] = v;
            carry = __shfl(v, 63, 64);
        }
        if (lane == 0) wtot[wave] = carry;
        __syncthreads();
        double off = 0.0;
        for (int w = 0; w < wave; ++w) off += wtot[w];
        for (int r = r0; r < r1 && wave > 0; ++r) {
            const int e = (r << 6) + lane;
            if (e < E) pre[e + 1] += off;
        }
        if (t == 0) pre[0] = 0.0;
        __syncthreads();
    }
    const double u = 1.1102230246251565e-16;
    const double lmax = ties ? exp((double)frame_max[f] * kLn10Over10) : 0.0;
    // a side is pre[a] - pre[b]: at most 2 (rpw + 11) + 1 roundings of the total
    const double ours = os ? 0.0 : (double)(2 * rpw + 24) * u * pre[E];
    int nt = 0;
#pragma unroll
    for (int c = 0; c < kCfarTile / kCfarThreads; ++c) {
        const int j = c * kCfarThreads + t, e = j + halo, b = t0 + j;
        int code;
        if (os) {
            auto cellf = [&](int m) {
                const int k = m / 2 + 1;
                return cell[(m & 1) ? e + p.guard + k : e - p.guard - k];
            };
            const double thr = os_select(cellf, p.train, os_rank(p)) + p.scale_db;
            const double v = cell[e];
            code = v >= thr ? 2 : (v < thr - p.hysteresis_db ? 0 : 1);
        } else {
            const double left = pre[e - p.guard] - pre[e - p.guard - p.train];
            const double right = pre[e + p.guard + p.train + 1] - pre[e + p.guard + 1];
            const double noise = cfar_combine(p.mode, left, right, p.train);
            const double on = noise * k_on, off = noise * k_off, v = cell[e];
            code = v >= on ? 2 : (v < off ? 0 : 1);
            if (ties) {
                const double rel = cfar_tie_rel(p.mode, left, right, p.train, b, lmax, ours);
                nt += (fabs(v - on) <= rel * on || fabs(v - off) <= rel * off) ? 1 : 0;
            }
        }
        const uint64_t hi = __ballot(code == 2), lo = __ballot(code == 0);
        if (lane == 0)
            masks[((int64_t)f * n + t0 + c * kCfarThreads + (wave << 6)) >> 6] = make_ulonglong2(hi, lo);
    }
    if (ties) {
        for (int o = 32; o > 0; o >>= 1) nt += __shfl_xor(nt, o);
        if (lane == 0) wties[wave] = nt;
        __syncthreads();
        if (t == 0) {
            int sum = 0;
            for (int w = 0; w < kCfarThreads / 64; ++w) sum += wties[w];
            if (sum) atomicAdd(&ambiguous[f], sum);
        }
    }
}

// Runs of one frame from its masks.  The "in a run" state changes to 1 at a
// code-2 bin and to 0 at a code-0 bin; thread t owns a contiguous range of
// 64-bin groups, whose effect on the state is a transfer function (keep, or
// set to the last decisive bin's value).  An exclusive scan of the transfers
// gives each thread its entering state; starts and ends are then found with
// bit scans, counted, placed by a second scan and written in order.
__global__ __launch_bounds__(kCfarThreads) void argus_cfar_runs_kernel(const float* __restrict__ psd, int n,
                                                                       gpud_cfar_params p,
                                                                       const ulonglong2* __restrict__ masks_all,
                                                                       gpud_surv_run* __restrict__ runs_all,
                                                                       int* __restrict__ run_counts, int max_runs)
{
    constexpr int NW = kCfarThreads / 64;
    __shared__ int wsc[3][NW];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int f = blockIdx.x;
    const int G = n >> 6, gpt = (G + kCfarThreads - 1) / kCfarThreads;
    const ulonglong2* masks = masks_all + (int64_t)f * G;
    const float* db = psd + (int64_t)f * n;
    gpud_surv_run* runs = runs_all + (int64_t)f * max_runs;
    const int g0 = min(t * gpt, G), g1 = min(g0 + gpt, G);

    int tf = -1;  // -1 keeps the state, 0/1 sets it
    for (int g = g0; g < g1; ++g) {
        const ulonglong2 m = masks[g];
        const uint64_t d = m.x | m.y;
        if (d) tf = (int)((m.x >> (63 - __clzll((long long)d))) & 1ull);
    }
    int s = tf;  // inclusive scan, the later constant winning
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(s, o, 64);
        if (lane >= o && s == -1) s = y;
    }
    if (lane == 63) wsc[0][wave] = s;
    int prev = __shfl_up(s, 1, 64);
    __syncthreads();
    int before = -1, last = -1;
    for (int w = 0; w < NW; ++w) {
        if (w < wave && wsc[0][w] != -1) before = wsc[0][w];
        if (wsc[0][w] != -1) last = wsc[0][w];
    }
    if (lane == 0) prev = -1;
    int in_state = prev != -1 ? prev : before;
    if (in_state < 0) in_state = 0;
    const bool open_at_end = last == 1;  // a run still open at the last bin ends there

    int ns = 0, ne = 0;
    {
        int st = in_state;
        for (int g = g0; g < g1; ++g) {
            const ulonglong2 m = masks[g];
            int pos = 0;
            while (pos < 64) {
                const uint64_t want = (st ? m.y : m.x) & (~0ull << pos);
                if (!want) break;
                const int bit = __builtin_ctzll(want);
                if (st) ne++;
                else ns++;
                st ^= 1;
                pos = bit + 1;
            }
        }
        if (t == kCfarThreads - 1 && open_at_end) ne++;
    }
    const int is = wave_scan_add(ns, lane), ie = wave_scan_add(ne, lane);
    if (lane == 63) {
        wsc[1][wave] = is;
        wsc[2][wave] = ie;
    }
    __syncthreads();
    int ps = is - ns, pe = ie - ne, total = 0;
    for (int w = 0; w < NW; ++w) {
        if (w < wave) {
            ps += wsc[1][w];
            pe += wsc[2][w];
        }
        total += wsc[1][w];
    }
    {
        int st = in_state;
        for (int g = g0; g < g1; ++g) {
            const ulonglong2 m = masks[g];
            int pos = 0;
            while (pos < 64) {
                const uint64_t want = (st ? m.y : m.x) & (~0ull << pos);
                if (!want) break;
                const int bit = __builtin_ctzll(want);
                const int b = (g << 6) + bit;
                if (st) {
                    if (pe < max_runs) runs[pe].last_bin = b - 1;
                    pe++;
                } else {
                    if (ps < max_runs) runs[ps].first_bin = b;
                    ps++;
                }
                st ^= 1;
                pos = bit + 1;
            }
        }
        if (t == kCfarThreads - 1 && open_at_end && pe < max_runs) runs[pe].last_bin = n - 1;
    }
    __syncthreads();  // every run's first and last bin written (workgroup-scope fence)
    const int nr = total < max_runs ? total : max_runs;
    for (int r = t; r < nr; r += kCfarThreads) {
        const int s0 = runs[r].first_bin, e0 = runs[r].last_bin;
        int pk = s0;
        double pv = -1e300;
        for (int g = s0 >> 6; g <= (e0 >> 6); ++g) {
            uint64_t hi = masks[g].x;
            if (g == (s0 >> 6)) hi &= ~0ull << (s0 & 63);
            if (g == (e0 >> 6) && (e0 & 63) != 63) hi &= (1ull << ((e0 & 63) + 1)) - 1ull;
            while (hi) {
                const int b = (g << 6) + __builtin_ctzll(hi);
                hi &= hi - 1ull;
                const double v = (double)db[b];
                if (v > pv) {
                    pv = v;
                    pk = b;
                }
            }
        }
        gpud_surv_run rr;
        rr.first_bin = s0;
        rr.last_bin = e0;
        rr.peak_bin = pk;
        rr.pad = 0;
        rr.peak_db = pv;
        rr.noise_db = cfar_threshold_at(db, n, pk, p) - p.scale_db;
        runs[r] = rr;
    }
    if (t == 0) run_counts[f] = total;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static int fft_radix(int n)
{
    if (n == 32768) return 32;
    if (n == 4096) return 16;
    return 0;
}

static size_t argus_cfar_lds_bytes(const gpud_cfar_params& p)
{
    return (size_t)(2 * (kCfarTile + 2 * (p.guard + p.train)) + 1) * sizeof(double);
}

extern "C" {

GPUD_API int gpud_surv_create(int n_fft, int sample_rate, gpud_surv** out)
{
    // transforms at 4096 and 32768 points; any other power of two from 1024
    // to 32768 makes a detection-only engine (CFAR and the detection list of
    // a decimated level whose spectra another engine computes)
    if (!out || sample_rate <= 0 || n_fft < 1024 || n_fft > 32768 || (n_fft & (n_fft - 1)) != 0) return -1;
    gpud_surv* s = new gpud_surv();
    s->n = n_fft;
    s->sample_rate = sample_rate;
    ARGUS_CHECK(hipStreamCreate(&s->stream));
    // symmetric Hann (internal/fft/fft.go:5-17), FP64 then rounded
    std::vector<float> w((size_t)n_fft);
    for (int i = 0; i < n_fft; i++) w[(size_t)i] = (float)(0.5 * (1 - std::cos(2 * M_PI * (double)i / (double)(n_fft - 1))));
    ARGUS_CHECK(hipMalloc((void**)&s->d_win, sizeof(float) * (size_t)n_fft));
    ARGUS_CHECK(hipMemcpy(s->d_win, w.data(), sizeof(float) * (size_t)n_fft, hipMemcpyHostToDevice));
    *out = s;
    return 0;
}

GPUD_API int gpud_surv_destroy(gpud_surv* s)
{
    if (!s) return 0;
    hipStreamSynchronize(s->stream);
    hipFree(s->d_win);
    if (s->d_scratch) hipFree(s->d_scratch);
    if (s->d_cfar) hipFree(s->d_cfar);
    if (s->d_zstuff) hipFree(s->d_zstuff);
    if (s->d_ltaps) hipFree(s->d_ltaps);
    if (s->d_lwin) hipFree(s->d_lwin);
    if (s->d_alpha) hipFree(s->d_alpha);
    if (s->h_alpha) hipHostFree(s->h_alpha);
    hipStreamDestroy(s->stream);
    delete s;
    return 0;
}

GPUD_API gpud_stream_handle gpud_surv_stream(gpud_surv* s) { return s ? (gpud_stream_handle)s->stream : nullptr; }

// Welch PSD (internal/fft/welch.go:15-66) of n_frames output frames: frame f
// averages n_segs Hann-windowed segments starting at f*frame_step +
// s*seg_hop.  Output: dB, fftshifted, float32 [n_frames][n_fft] on the
// device.  The caller guarantees (n_frames-1)*frame_step +
// (n_segs-1)*seg_hop + n_fft samples are readable.
GPUD_API int gpud_surv_welch_device(gpud_surv* s, const gpud_float2* iq_dev, int64_t frame_step, int seg_hop,
                                    int n_segs, int n_frames, float* psd_dev)
{
    if (!s || !iq_dev || !psd_dev || n_segs <= 0 || seg_hop < 0 || frame_step < 0) return -1;
    if (fft_radix(s->n) == 0) return -1;  // a detection-only engine
    if (n_frames <= 0) return 0;
    const int R = fft_radix(s->n);
    argus_clear_error();
    if (R == 32)
        hipLaunchKernelGGL(argus_welch_kernel<32>, dim3(n_frames), dim3(1024), 0, s->stream, (const float2*)iq_dev,
                           frame_step, seg_hop, n_segs, s->d_win, psd_dev);
    else
        hipLaunchKernelGGL(argus_welch_kernel<16>, dim3(n_frames), dim3(256), 0, s->stream, (const float2*)iq_dev,
                           frame_step, seg_hop, n_segs, s->d_win, psd_dev);
    return (int)hipGetLastError();
}

// Forward unnormalised C2C FFT of n_frames frames (frame f at f*frame_step),
// natural-order output [n_frames][n_fft]; window 1 applies the Hann window
// (the host multiply of pipeline_runtime.go:315-341 fused in).
GPUD_API int gpud_surv_fft_device(gpud_surv* s, const gpud_float2* in_dev, int64_t frame_step, int n_frames,
                                  int window, gpud_float2* out_dev)
{
    if (!s || !in_dev || !out_dev || fft_radix(s->n) == 0) return -1;
    if (n_frames <= 0) return 0;
    const float* w = window ? s->d_win : nullptr;
    argus_clear_error();
    if (fft_radix(s->n) == 32)
        hipLaunchKernelGGL(argus_fft_kernel<32>, dim3(n_frames), dim3(1024), 0, s->stream, (const float2*)in_dev, w,
                           (float2*)out_dev, frame_step);
    else
        hipLaunchKernelGGL(argus_fft_kernel<16>, dim3(n_frames), dim3(256), 0, s->stream, (const float2*)in_dev, w,
                           (float2*)out_dev, frame_step);
    return (int)hipGetLastError();
}

}  // extern "C"

int argus_launch_fft4096(const float2* in, int64_t frame_step, int n_frames, float2* out, hipStream_t stream)
{
    if (n_frames <= 0) return 0;
    argus_clear_error();
    hipLaunchKernelGGL(argus_fft_kernel<16>, dim3(n_frames), dim3(256), 0, stream, in, (const float*)nullptr, out,
                       frame_step);
    return (int)hipGetLastError();
}

extern "C" {

// Single-FFT dB spectra of n_frames frames (frame f at f*frame_step): the
// GPU path of spectrumFromIQWithPlan (pipeline_runtime.go:315-341) --
// window (1 = Hann on the device), forward FFT, SpectrumFromFFT and the
// NaN/Inf sanitising -- as float32 [n_frames][n_fft], fftshifted.
GPUD_API int gpud_surv_spectrum_device(gpud_surv* s, const gpud_float2* in_dev, int64_t frame_step, int n_frames,
                                       int window, float* db_dev)
{
    if (!s || !in_dev || !db_dev) return -1;
    if (n_frames <= 0) return 0;
    const size_t need = (size_t)n_frames * (size_t)s->n;
    if (s->scratch_cap < need) {
        if (s->d_scratch) hipFree(s->d_scratch);
        s->d_scratch = nullptr;
        s->scratch_cap = 0;
        ARGUS_CHECK(hipMalloc((void**)&s->d_scratch, need * sizeof(float2)));
        s->scratch_cap = need;
    }
    int err = gpud_surv_fft_device(s, in_dev, frame_step, n_frames, window, (gpud_float2*)s->d_scratch);
    if (err) return err;
    const int64_t total = (int64_t)need;
    const int grid = (int)std::min<int64_t>((total + 255) / 256, 4096);
    argus_clear_error();
    hipLaunchKernelGGL(argus_spectrum_db_kernel, dim3(grid), dim3(256), 0, s->stream, (const float2*)s->d_scratch,
                       s->n, (int64_t)s->n, total, db_dev);
    return (int)hipGetLastError();
}

static int grow_scratch(float2** p, size_t* cap, size_t need)
{
    if (*cap >= need) return 0;
    if (*p) hipFree(*p);
    *p = nullptr;
    *cap = 0;
    ARGUS_CHECK(hipMalloc((void**)p, need * sizeof(float2)));
    *cap = need;
    return 0;
}

// The spectrum of a decimated surveillance level (the multi-resolution
// derived level, pipeline_runtime.go:1274-1293 plan, 624-645 use): per frame
// the n_fft samples at in_dev + f * frame_step are low-passed with
// LowpassFIR((sample_rate / decim) * 0.5 * 0.8, sample_rate, 101)
// (decimationTaps, 399-412), decimated by decim, Hann-windowed over the
// level's n_fft / decim samples and transformed; dB, fftshifted, NaN/Inf
// sanitised: float32 [n_frames][n_fft / decim].
GPUD_API int gpud_surv_level_spectrum_device(gpud_surv* s, const gpud_float2* in_dev, int64_t frame_step,
                                             int n_frames, int decim, float* db_dev)
{
    if (!s || !in_dev || !db_dev || frame_step < 0 || fft_radix(s->n) == 0) return -1;
    if (decim < 2 || decim > kLevelMaxDecim || (decim & (decim - 1)) != 0 || s->n % decim != 0 || s->n / decim < 256)
        return -1;
    if (n_frames <= 0) return 0;
    const int n_d = s->n / decim;
    if (s->level_decim != decim) {  // taps and window of this level, once
        std::vector<double> taps(102);
        const int nt = argus_lowpass_fir((double)(s->sample_rate / decim) * 0.5 * 0.8, s->sample_rate, 101,
                                         taps.data(), false);
        if (nt <= 0 || nt > kLevelMaxTaps) return -1;
        std::vector<float> w((size_t)n_d);
        for (int i = 0; i < n_d; i++)
            w[(size_t)i] = (float)(0.5 * (1 - std::cos(2 * M_PI * (double)i / (double)(n_d - 1))));
        if (s->d_ltaps) hipFree(s->d_ltaps);
        if (s->d_lwin) hipFree(s->d_lwin);
        s->d_ltaps = nullptr;
        s->d_lwin = nullptr;
        s->level_decim = 0;
        ARGUS_CHECK(hipMalloc((void**)&s->d_ltaps, sizeof(double) * (size_t)nt));
        ARGUS_CHECK(hipMalloc((void**)&s->d_lwin, sizeof(float) * (size_t)n_d));
        ARGUS_CHECK(hipMemcpy(s->d_ltaps, taps.data(), sizeof(double) * (size_t)nt, hipMemcpyHostToDevice));
        ARGUS_CHECK(hipMemcpy(s->d_lwin, w.data(), sizeof(float) * (size_t)n_d, hipMemcpyHostToDevice));
        s->level_taps = nt;
        s->level_decim = decim;
    }
    const size_t need = (size_t)n_frames * (size_t)s->n;
    int err = grow_scratch(&s->d_zstuff, &s->zstuff_cap, need);
    if (!err) err = grow_scratch(&s->d_scratch, &s->scratch_cap, need);
    if (err) return err;
    argus_clear_error();
    hipLaunchKernelGGL(argus_level_fir_kernel, dim3((n_d + kLevelOut - 1) / kLevelOut, n_frames), dim3(kLevelOut),
                       level_fir_lds_bytes(decim, s->level_taps), s->stream,
                       (const float2*)in_dev, frame_step, decim, n_d, (const double*)s->d_ltaps, s->level_taps,
                       (const float*)s->d_lwin, s->d_zstuff);
    err = (int)hipGetLastError();
    if (!err)
        err = gpud_surv_fft_device(s, (const gpud_float2*)s->d_zstuff, s->n, n_frames, 0,
                                   (gpud_float2*)s->d_scratch);
    if (err) return err;
    const int64_t total = (int64_t)n_frames * n_d;
    const int grid = (int)std::min<int64_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(argus_spectrum_db_kernel, dim3(grid), dim3(256), 0, s->stream, (const float2*)s->d_scratch,
                       n_d, (int64_t)s->n, total, db_dev);
    return (int)hipGetLastError();
}

// CA / GOSCA / CASO CFAR (internal/cfar/ca.go:22-68, gosca.go:22-78,
// caso.go:16-70) on each frame of a dB spectrum, followed by the threshold
// crossing with hysteresis of detector.go:436-480: runs (first, last, peak
// bin, peak dB, noise = threshold[peak] - scale).  runs_dev holds n_frames *
// max_runs records; run_counts_dev the true run count per frame (may exceed
// max_runs, in which case the list is truncated).
GPUD_API int gpud_surv_cfar_runs_device(gpud_surv* s, const float* psd_dev, int n_frames,
                                        const gpud_cfar_params* params, gpud_surv_run* runs_dev,
                                        int* run_counts_dev, int max_runs)
{
    return gpud_surv_cfar_runs_device_ex(s, psd_dev, n_frames, params, runs_dev, run_counts_dev, max_runs, nullptr);
}

GPUD_API int gpud_surv_cfar_runs_device_ex(gpud_surv* s, const float* psd_dev, int n_frames,
                                           const gpud_cfar_params* params, gpud_surv_run* runs_dev,
                                           int* run_counts_dev, int max_runs, int* ambiguous_dev)
{
    if (!s || !psd_dev || !params || !runs_dev || !run_counts_dev || max_runs <= 0) return -1;
    const gpud_cfar_params p = *params;
    if (p.train <= 0 || p.guard < 0 || p.guard + p.train > 2048) return -1;
    if (p.mode != GPUD_CFAR_CA && p.mode != GPUD_CFAR_GOSCA && p.mode != GPUD_CFAR_CASO && p.mode != GPUD_CFAR_OS)
        return -1;
    if (p.mode == GPUD_CFAR_OS && p.train > 128) return -1;
    if (s->n % kCfarTile != 0) return -1;
    if (n_frames <= 0) return 0;
    const size_t lds = argus_cfar_lds_bytes(p);  // at most 82 KB (guard + train <= 2048)
    gpud_cfar_params pp = p;
    if (pp.scale_db <= 0) pp.scale_db = 6;  // cfar.go:13-15
    // per frame n/64 masks of 16 bytes, then one float (the frame's largest dB value)
    const size_t mask_bytes = (size_t)n_frames * (size_t)(s->n / 64) * sizeof(ulonglong2);
    const size_t need = mask_bytes + (size_t)n_frames * sizeof(float);
    if (s->cfar_cap < need) {
        if (s->d_cfar) hipFree(s->d_cfar);
        s->d_cfar = nullptr;
        s->cfar_cap = 0;
        ARGUS_CHECK(hipMalloc(&s->d_cfar, need));
        s->cfar_cap = need;
    }
    ulonglong2* masks = (ulonglong2*)s->d_cfar;
    float* fmax = (float*)((char*)s->d_cfar + mask_bytes);
    const double k_on = std::pow(10.0, pp.scale_db / 10.0);
    const double k_off = std::pow(10.0, (pp.scale_db - pp.hysteresis_db) / 10.0);
    if (lds > 64 * 1024)
        ARGUS_CHECK(hipFuncSetAttribute((const void*)argus_cfar_status_kernel,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    argus_clear_error();
    if (ambiguous_dev)
        hipLaunchKernelGGL(argus_cfar_prep_kernel, dim3(n_frames), dim3(256), 0, s->stream, psd_dev, s->n, fmax,
                           ambiguous_dev);
    hipLaunchKernelGGL(argus_cfar_status_kernel, dim3((s->n / kCfarTile) * n_frames), dim3(kCfarThreads), lds, s->stream,
                       psd_dev, s->n, pp, k_on, k_off, (const float*)fmax, ambiguous_dev, masks);
    hipLaunchKernelGGL(argus_cfar_runs_kernel, dim3(n_frames), dim3(kCfarThreads), 0, s->stream, psd_dev, s->n, pp,
                       (const ulonglong2*)masks, runs_dev, run_counts_dev, max_runs);
    return (int)hipGetLastError();
}

}  // extern "C"
