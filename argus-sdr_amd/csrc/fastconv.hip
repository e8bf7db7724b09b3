// fastconv.hip -- overlap-save fast convolution for the dense channels of the
// streaming service: the same filter as the direct polyphase form, computed
// through one shared transform of the input instead of per-channel staging.
//
// For a channel with NCO phase phi(n) = phase0 + inc n and taps h (T <= 512),
//     y[m] = sum_k h[k] s[e_m - k],  s[n] = x[n] e^{j 2pi phi(n)}
//          = e^{j 2pi phi(e_m)} sum_k g[k] x[e_m - k],  g[k] = h[k] e^{-j 2pi inc k},
// the convolution of the raw input with the channel's modulated taps, then
// one rotation per output (the factorisation the narrow items use too).
// The raw input is shared by every channel, so its transform is too:
//
//   stage   u = the stream's last 511 samples ++ the block ++ zeros
//   FFT     X_s = FFT_4096(u[s*3584 .. s*3584 + 4096)), s < S = ceil(n/3584)
//           (one launch of the surveillance 4096-point kernel for all s)
//   channel G = FFT_4096(g) / 4096 (host, FP64, once per channel), times the
//           decimation-phase ramp e^{j 2pi k r / 4096} (a shift by r samples);
//           per segment, decimation by Dp (the power of two dividing D,
//           at most 64) in the frequency domain (folding),
//           F[j] = sum_{i < Dp} X_s[j + i J] G'[j + i J],  J = 4096 / Dp,
//           and an inverse transform of J points gives the segment's
//           convolution at u-positions q Dp + r, q < J; the outputs are the
//           positions among them that are emission points (every D/Dp-th).
//           Positions [511, 511 + 3584) are valid for T <= 512
//           (overlap-save), so each block sample belongs to exactly one
//           segment.
//   tail    the stream's last 511 samples after the block.
//
// The inverse transform is the 512-point FFT (8^3) of the conjugated folded
// spectrum: for Dp = 16 two segments' 256 bins packed into one transform
// and separated after it, for Dp = 32 and 64 zero-stuffed by 512 / J
// (stuffing the spectrum repeats the signal, so its first J outputs are the
// J-point transform).  A work item is one channel and a block of segments:
// G' in LDS (32 KiB), each wavefront of the workgroup runs its own segments
// with transposes synchronised within the wave.  An output's rotation is
// one NCO evaluation per lane and segment, then a complex step per 64 fold
// positions.
//
// The last 511 samples of the stream are the channel's own history only once
// the channel has seen T - 1 samples (a channel opened mid-stream starts
// from zero history), so the service routes a channel here from its first
// block with a full history; the direct form keeps writing its shifted
// history, so either form can follow the other.
//
// Arithmetic: FP32 transforms (about 12 rounding steps on each side), FP64
// taps and G.  The outputs differ from the direct form's in the last bits
// of the in-band signal only (relative RMS ~1e-6).
#include "fastconv.hpp"
#include "fft_core.hpp"

#include <cmath>
#include <complex>
#include <vector>

namespace {

using namespace argus_fftc;

constexpr int kN = ARGUS_FAST_N;
constexpr int kHop = ARGUS_FAST_HOP;
constexpr int kTail = ARGUS_FAST_TAIL;
constexpr int kJ = 512;  // the inverse transform: 8^3 points
constexpr int kThreads = 256;
static_assert(kHop % 64 == 0 && kTail + kHop <= kN, "segment geometry");

// u[i], i < (S-1)*hop + N: tail, then the block, then zeros.
__global__ __launch_bounds__(kThreads) void argus_fast_stage_kernel(const ArgusFastHdr* __restrict__ hdr,
                                                                    const ArgusExtractHdr* __restrict__ xh)
{
    const int S = hdr->S;
    if (S <= 0) return;
    const int len = (S - 1) * kHop + kN;
    const int n = xh->n_new;
    const float2* in = xh->in;
    const float2* tail = hdr->tail;
    float2* u = hdr->u;
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < len; i += gridDim.x * kThreads) {
        const int b = i - kTail;
        u[i] = b < 0 ? tail[i] : (b < n ? in[b] : make_float2(0.0f, 0.0f));
    }
}

// ---------------------------------------------------------------------------
// The per-channel pass.  Work items are (channel, block of kSegItem
// segments); each of a workgroup's kWaves wavefronts takes every kWaves-th
// segment of the block on its own: fold, a 512-point transform with its
// LDS transposes synchronised within the wave only, and the outputs.  The
// workgroup synchronises only to stage the channel's response G' in LDS.
// Items are numbered channel-major, so with a grid that is a multiple of
// the eight XCDs and two blocks per channel, each XCD works on one half of
// the segments and its L2 holds that half of the segment spectra.
// ---------------------------------------------------------------------------
constexpr int kWaves = 8;                      // wavefronts per workgroup
constexpr int kChanThreads = 64 * kWaves;
constexpr int kSegWave = 7;                    // segments per wavefront per item
constexpr int kSegItem = kWaves * kSegWave;    // 56: two items cover a 20 ms block at 20 Msps
constexpr int kXStride = 72;                   // first transpose: 64 values + 8 pad per row (conflict-free reads)
constexpr int kPadC = 32;                      // complex words of padding between a group's packed segments
constexpr int kXsFloats = 2 * (512 + kPadC);   // the packed folded bins; also the transposes (8 x 72, 8 x 9 x 8 floats)

// The wave's LDS operations complete in issue order, so data a wave passes
// between its own lanes through LDS needs no hardware barrier: only the
// compiler must not move the accesses across this point.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Global (not generic) pointers for the kernels' HBM accesses: a flat access
// counts on the LDS counter too, so the waves' LDS waits (fold, transposes)
// would also wait for segment-spectrum loads and output stores in flight.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(1))) const float2 GlobalCF2;
typedef __attribute__((address_space(1))) const float4 GlobalCF4;
typedef __attribute__((address_space(1))) float2 GlobalF2;
typedef __attribute__((address_space(1))) float GlobalF;
typedef const __attribute__((address_space(4))) ArgusFastDesc CFastDesc;
#else  // the host pass only needs the kernels to compile
typedef const float2 GlobalCF2;
typedef const float4 GlobalCF4;
typedef float2 GlobalF2;
typedef float GlobalF;
typedef const ArgusFastDesc CFastDesc;
#endif

// Per-lane twiddles of the 512-point transform (functions of the lane only):
// pass 1 W_512^{t k}, pass 2 W_64^{(t mod 8) k}, k = 1..7.
struct Tw512 {
    float2 w1[7], w2[7];
};

__device__ __forceinline__ void tw512_init(Tw512& tw, int t)
{
#pragma unroll
    for (int k = 1; k < 8; k++) {
        tw.w1[k - 1] = twiddle((t * k) & 511, 1.0f / 512.0f);
        tw.w2[k - 1] = twiddle(((t & 7) * k) & 63, 1.0f / 64.0f);
    }
}

// Transpose through the wave's LDS region: write v[k] at wr(k), read v[k]
// from rd(k); real parts then imaginary parts through the same region.
template <typename WR, typename RD>
__device__ __forceinline__ void wave_exchange(float2 (&v)[8], float* xs, WR wr, RD rd)
{
#pragma unroll
    for (int k = 0; k < 8; ++k) xs[wr(k)] = v[k].x;
    wave_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k].x = xs[rd(k)];
    wave_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) xs[wr(k)] = v[k].y;
    wave_sync();
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k].y = xs[rd(k)];
    wave_sync();
}

// 512-point forward transform on one wavefront (8^3: lane t = a + 8 b holds
// n = t + 64 c, c = 0..7).  On return lane t = kb + 8 kc holds
// X[kc + 8 kb + 64 ka] in v[ka].
__device__ __forceinline__ void wave_fft512(float2 (&v)[8], float* xs, int t, const Tw512& tw)
{
    dft<8>(v);
#pragma unroll
    for (int k = 1; k < 8; ++k) v[k] = cmul(v[k], tw.w1[k - 1]);
    {
        // element (t, kc) at t + 72 kc: the writes are consecutive per kc and
        // the reads (lanes vary a and kc) fall on distinct banks per half-wave
        const int a = t & 7, kc = t >> 3;
        wave_exchange(v, xs, [&](int k) { return t + kXStride * k; },
                      [&](int b) { return a + 8 * b + kXStride * kc; });
    }
    dft<8>(v);
    {
        const int a = t & 7, kc = t >> 3;
#pragma unroll
        for (int k = 1; k < 8; ++k) v[k] = cmul(v[k], tw.w2[k - 1]);
        const int kb2 = t & 7, kc2 = t >> 3;
        wave_exchange(v, xs, [&](int kb) { return a + 9 * kb + 72 * kc; },
                      [&](int aa) { return aa + 9 * kb2 + 72 * kc2; });
    }
    dft<8>(v);
}

// F[j] = sum_{i < DP} X[j + i J] G'[j + i J] for this lane's j = t + 64 c,
// c < J / 64: every lane folds the same number of bins, all loads of one
// i issued together.
template <int DP>
__device__ __forceinline__ void fast_fold(const GlobalCF2* __restrict__ xseg, const float2* gs, int t,
                                          float2 (&f)[ARGUS_FAST_N / DP / 64])
{
    constexpr int J = ARGUS_FAST_N / DP, PER = J / 64;
    float ar[PER], ai[PER];
#pragma unroll
    for (int c = 0; c < PER; c++) ar[c] = ai[c] = 0.0f;
#pragma unroll 4
    for (int i = 0; i < DP; i++) {
        float2 xv[PER];
#pragma unroll
        for (int c = 0; c < PER; c++) xv[c] = xseg[i * J + t + 64 * c];
#pragma unroll
        for (int c = 0; c < PER; c++) {
            const float2 gv = gs[i * J + t + 64 * c];
            ar[c] += xv[c].x * gv.x - xv[c].y * gv.y;
            ai[c] += xv[c].x * gv.y + xv[c].y * gv.x;
        }
    }
#pragma unroll
    for (int c = 0; c < PER; c++) f[c] = make_float2(ar[c], ai[c]);
}

// The same fold with two adjacent bins per lane, j = 2 t + 128 c and j + 1
// (c < J / 128), both read with one 16-byte load: half the vector memory
// instructions, which bound this kernel.  f[c] holds F[j] in .xy and
// F[j + 1] in .zw.
template <int DP>
__device__ __forceinline__ void fast_fold_pairs(const GlobalCF2* __restrict__ xseg, const float2* gs, int t,
                                                float4 (&f)[ARGUS_FAST_N / DP / 128])
{
    constexpr int J = ARGUS_FAST_N / DP, P = J / 128;
    float ar[P], ai[P], br[P], bi[P];
#pragma unroll
    for (int c = 0; c < P; c++) ar[c] = ai[c] = br[c] = bi[c] = 0.0f;
#pragma unroll 4
    for (int i = 0; i < DP; i++) {
        float4 xv[P];
#pragma unroll
        for (int c = 0; c < P; c++) xv[c] = *(const GlobalCF4*)(xseg + i * J + 2 * t + 128 * c);
#pragma unroll
        for (int c = 0; c < P; c++) {
            const float4 gv = *(const float4*)(gs + i * J + 2 * t + 128 * c);
            ar[c] += xv[c].x * gv.x - xv[c].y * gv.y;
            ai[c] += xv[c].x * gv.y + xv[c].y * gv.x;
            br[c] += xv[c].z * gv.z - xv[c].w * gv.w;
            bi[c] += xv[c].z * gv.w + xv[c].w * gv.z;
        }
    }
#pragma unroll
    for (int c = 0; c < P; c++) f[c] = make_float4(ar[c], ai[c], br[c], bi[c]);
}

// Segments per 512-point transform for a fold factor (see fast_group).
template <int DP>
__host__ __device__ constexpr int fast_pack()
{
    return DP == 16 ? 2 : 1;
}

// The segment's outputs m_lo + i, i < cnt, from os[i]: baseband, and the
// demodulated value when the channel demodulates here.  FM takes its
// predecessor from the same segment, so a segment's first output (i = 0) is
// left to the demodulation pass.
__device__ __forceinline__ void fast_emit(const ArgusFastDesc& d, const float2* os, int m_lo, int cnt, int t)
{
    GlobalF2* out = (GlobalF2*)d.out;
    GlobalF* audio = (GlobalF*)d.audio;
    for (int i = t; i < cnt; i += 64) {
        const int m = m_lo + i;
        const float2 y = os[i];
        out[m] = y;
        switch (d.demod) {  // uniform: one channel per workgroup
        case ARGUS_DEMOD_NFM:
        case ARGUS_DEMOD_WFM:
            if (i > 0) audio[m - d.skip] = argus_fm_step(os[i - 1], y);
            break;
        case ARGUS_DEMOD_USB:
        case ARGUS_DEMOD_LSB: {
            float sn, cn;
            argus_nco(d.bfo0 + d.bfo_inc * (uint64_t)(uint32_t)m, sn, cn);
            audio[m] = argus_ssb_out(y, sn, cn);
            break;
        }
        case ARGUS_DEMOD_CW: {
            float sn, cn;
            argus_nco(d.bfo0 + d.bfo_inc * (uint64_t)(uint32_t)m, sn, cn);
            audio[m] = y.x * cn;
            break;
        }
        default:
            break;
        }
    }
}

// Per channel, the constants of the output stage.
struct FastOut {
    int E;          // D / DP: emission every E-th fold position
    float invE;     // 1 / E (rounded up: floor by multiply, then corrected)
    float invD;     // 1 / D, the same way
    float2 step;    // rotation between a lane's consecutive values: e^{j 2pi inc 64 DP}
};

// 1 / d rounded up (d >= 1): for 0 <= a < 2^22, (int)(a * recip) is
// floor(a / d) or one off, which fast_div corrects.
__device__ __forceinline__ float fast_recip(int d)
{
    return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, 1.0f / (float)d) + 1u);
}

// floor(a / d) for 0 <= a < 2^22 without an integer division.
__device__ __forceinline__ int fast_div(int a, int d, float recip)
{
    int q = (int)((float)a * recip);
    const int r = a - q * d;
    if (r < 0) q--;
    else if (r >= d) q++;
    return q;
}

// One group of segments of one channel on one wavefront: fold each, one
// 512-point transform for all of them, outputs.  A J-point inverse (J =
// 4096 / DP) fills 512 points by zero-stuffing (ZS) or by packing ZP
// segments (ZS ZP J = 512): DP = 8 takes one segment, DP = 16 packs two,
// DP = 32 and 64 take one segment stuffed 4 or 8 times (packing more would
// need more registers than four waves per SIMD leave).  The group's
// segments are s_z = sb + kWaves z, z < ZP (those at or past s_end are
// absent: zero input, no outputs).
//
// Input x[ZS (ZP n + z)] = conj(F_z[n]) (n < J), zero elsewhere.  Its
// forward transform at k = q + J j (q < J, j < ZP) is
// sum_z W_ZP^{zj} W_{ZP J}^{zq} Yc_z[q], Yc_z the J-point transform of
// conj(F_z); the ZP values of one q sit in one lane (v[ka + KA j], KA =
// J / 64 values of q per lane).  So with c_j = conj(V[q + J j]), the
// segment's inverse transform out_z[q] = conj(Yc_z[q]) is
// W_{ZP J}^{zq} (sum_j W_ZP^{zj} c_j) / ZP: a ZP-point DFT and a twiddle
// per lane and value (ZP = 1: the forward transform of the conjugate).
// Folded segment z sits at fs + z RS; RS = J + 32 / ZP keeps the
// transform's strided reads of the regions in distinct banks (64 banks of
// 4 bytes, 32 lanes per cycle).
template <int DP>
__device__ __forceinline__ void fast_group(const ArgusFastDesc& d, const FastOut& fo, const GlobalCF2* __restrict__ X,
                                           const float2* gs, float* xs, int sb, int s_end, int t, const Tw512& tw)
{
    constexpr int J = ARGUS_FAST_N / DP, PER = J / 64, Z = fast_pack<DP>(), ZS = kJ / J / Z, KA = J / 64;
    constexpr int RS = J + kPadC / Z;
    static_assert(Z * RS * 2 <= kXsFloats, "the group's regions fit the wave's scratch");
    static_assert(ZS * Z * J == kJ, "stuffing and packing fill the transform");
    // opaque to the optimiser: keeps the lane's LDS addresses (functions of
    // t only) from being hoisted out of the segment loop into registers
    asm volatile("" : "+v"(t));
    float2* fs = (float2*)xs;
#pragma unroll
    for (int z = 0; z < Z; z++) {
        const int sz = sb + kWaves * z;  // uniform over the wave
        float2* fz = fs + z * RS;
        if constexpr (PER >= 2) {
            // folded in bin pairs, then laid out as the transform takes them
            float4 f[PER / 2];
            if (sz < s_end) {
                fast_fold_pairs<DP>(X + (size_t)sz * kN, gs, t, f);
            } else {
#pragma unroll
                for (int c = 0; c < PER / 2; c++) f[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            }
#pragma unroll
            for (int c = 0; c < PER / 2; c++) *(float4*)(fz + 2 * t + 128 * c) = f[c];
        } else {
            float2 f[PER];
            if (sz < s_end) {
                fast_fold<DP>(X + (size_t)sz * kN, gs, t, f);
            } else {
#pragma unroll
                for (int c = 0; c < PER; c++) f[c] = make_float2(0.0f, 0.0f);
            }
#pragma unroll
            for (int c = 0; c < PER; c++) fz[t + 64 * c] = f[c];
        }
    }
    wave_sync();
    float2 v[8];
    {
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const int nn = t + 64 * c;
            const int mm = nn / ZS;
            const float2 y = (nn % ZS) == 0 ? fs[(mm % Z) * RS + mm / Z] : make_float2(0.0f, 0.0f);
            v[c] = make_float2(y.x, -y.y);
        }
        wave_sync();
    }
    wave_fft512(v, xs, t, tw);
    // lane t = kb + 8 kc holds V[q + J j], q = kc + 8 kb + 64 ka (ka < KA),
    // in v[ka + KA j] (the rest, for ZS > 1, repeats)
    const int kb = t & 7, kc = t >> 3;
    const int q0 = kc + 8 * kb;
    float2 y[Z][KA];  // out_z[q0 + 64 ka]
#pragma unroll
    for (int ka = 0; ka < KA; ka++) {
        float2 c[Z];
#pragma unroll
        for (int j = 0; j < Z; j++) c[j] = make_float2(v[ka + KA * j].x, -v[ka + KA * j].y);
        if constexpr (Z > 1) {
            dft<Z>(c);
            const int q = q0 + 64 * ka;
            constexpr float inv = 1.0f / (float)Z;  // exact
            y[0][ka] = make_float2(c[0].x * inv, c[0].y * inv);
#pragma unroll
            for (int z = 1; z < Z; z++) {
                const float2 w = cmul(c[z], twiddle((ZS * z * q) & (kJ - 1), 1.0f / kJ));
                y[z][ka] = make_float2(w.x * inv, w.y * inv);
            }
        } else {
            y[0][ka] = c[0];
        }
    }
    // fold position q of segment s is block sample e = s hop + q DP + r -
    // tail, and e - e0 = DP (q + cs) with cs uniform over the segment; it is
    // an output when q + cs >= 0 is a multiple of E, output (q + cs) / E.
    // Segment s's outputs are m in [m_lo, m_hi) (block samples e0 + m D in
    // [s hop, (s + 1) hop)), at most hop / DP + 1 <= J of them; they are
    // laid out in order in region z of the scratch (free after the
    // transform), then written (and demodulated) by consecutive lanes.
    float2* os = fs;
    int m_lo[Z], m_hi[Z];
#pragma unroll
    for (int z = 0; z < Z; z++) {
        const int sz = sb + kWaves * z;
        m_lo[z] = m_hi[z] = 0;
        if (sz >= s_end) continue;
        const int cs = (sz * kHop + d.r - kTail - d.e0) / DP;  // exact: e - e0 is a multiple of DP
        // e0 < D: the numerators are >= 0 (and below 2^22: blocks are shorter)
        m_lo[z] = fast_div(sz * kHop - d.e0 + d.D - 1, d.D, fo.invD);
        m_hi[z] = min(d.n_out, fast_div(sz * kHop + kHop - d.e0 + d.D - 1, d.D, fo.invD));
        // rotation of the lane's first position, then one step per 64 positions
        float sn, cn;
        argus_nco(d.phase0 + d.inc * (uint64_t)(int64_t)(sz * kHop + q0 * DP + d.r - kTail), sn, cn);
        float2 rot = make_float2(cn, sn);
        float2* oz = os + z * RS;
#pragma unroll
        for (int ka = 0; ka < KA; ka++) {
            const int q = q0 + 64 * ka;
            const int iseg = q * DP + d.r;
            const int u = q + cs;
            if (iseg >= kTail && iseg < kTail + kHop && u >= 0) {
                int m = (int)((float)u * fo.invE);
                int rem = u - m * fo.E;
                if (rem < 0) {
                    m--;
                    rem += fo.E;
                } else if (rem >= fo.E) {
                    m++;
                    rem -= fo.E;
                }
                if (rem == 0 && m < d.n_out) {
                    const float2 yy = y[z][ka];
                    oz[m - m_lo[z]] = make_float2(yy.x * rot.x - yy.y * rot.y, yy.x * rot.y + yy.y * rot.x);
                }
            }
            rot = make_float2(rot.x * fo.step.x - rot.y * fo.step.y, rot.x * fo.step.y + rot.y * fo.step.x);
        }
    }
    wave_sync();
#pragma unroll
    for (int z = 0; z < Z; z++) fast_emit(d, os + z * RS, m_lo[z], m_hi[z] - m_lo[z], t);
    wave_sync();  // the scratch holds the next group's folded bins
}

template <int DP>
__device__ __forceinline__ void fast_block(const ArgusFastDesc& d, const GlobalCF2* __restrict__ X, const float2* gs,
                                           float* xs, int s_begin, int s_end, int w, int t, const Tw512& tw)
{
    constexpr int Z = fast_pack<DP>();
    FastOut fo;
    fo.E = d.D / DP;
    fo.invE = fast_recip(fo.E);
    fo.invD = fast_recip(d.D);
    float sn, cn;
    argus_nco(d.inc * (uint64_t)(64 * DP), sn, cn);
    fo.step = make_float2(cn, sn);
    for (int sb = s_begin + w; sb < s_end; sb += kWaves * Z) fast_group<DP>(d, fo, X, gs, xs, sb, s_end, t, tw);
}

__global__ __launch_bounds__(kChanThreads) void argus_fast_chan_kernel(const ArgusFastHdr* __restrict__ hdr, int part)
{
    __shared__ __attribute__((aligned(16))) float2 gs[kN];
    __shared__ __attribute__((aligned(16))) float xs[kWaves][kXsFloats];
    static_assert(kXsFloats >= 2 * kJ, "the folded bins fit the transform scratch");
    static_assert(kXsFloats >= 9 * 8 * 8 && kXsFloats >= 8 * kXStride, "the transposes fit the scratch");
    const int w = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int S = hdr->S;
    int c_begin = 0, c_end = hdr->n_fast;  // the channels of this launch
    if (part >= 0) {
        c_begin = part > 0 ? hdr->part_end[part - 1] : 0;
        c_end = hdr->part_end[part];
    }
    const int blocks = (S + kSegItem - 1) / kSegItem;
    const int n_items = (c_end - c_begin) * blocks;
    const GlobalCF2* X = (const GlobalCF2*)hdr->X;
    Tw512 tw;
    tw512_init(tw, t);
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        const int ci = it / blocks, b = it - ci * blocks;
        // scalar loads (constant address space): the descriptor is uniform and
        // read-only here, and its fields then live in SGPRs
        const ArgusFastDesc d = ((const CFastDesc*)hdr->descs)[c_begin + ci];
        __syncthreads();  // the previous item's waves are done with gs
        // G with the ramp e^{j 2pi k r / N}: exact turn fractions
        for (int k = threadIdx.x; k < kN; k += kChanThreads) {
            const float2 g = *(GlobalCF2*)(d.G + k);
            const float x = (float)((k * d.r) & (kN - 1)) * (1.0f / kN);
            const float sn = __builtin_amdgcn_sinf(x), cs = __builtin_amdgcn_cosf(x);
            gs[k] = make_float2(g.x * cs - g.y * sn, g.x * sn + g.y * cs);
        }
        __syncthreads();
        const int s0 = b * kSegItem, s1 = min(S, s0 + kSegItem);
        switch (d.Dp) {  // uniform over the workgroup (one channel)
        case 8: fast_block<8>(d, X, gs, xs[w], s0, s1, w, t, tw); break;
        case 16: fast_block<16>(d, X, gs, xs[w], s0, s1, w, t, tw); break;
        case 32: fast_block<32>(d, X, gs, xs[w], s0, s1, w, t, tw); break;
        default: fast_block<64>(d, X, gs, xs[w], s0, s1, w, t, tw); break;
        }
    }
}

// tail' = the last 511 samples of tail ++ block.
__global__ __launch_bounds__(512) void argus_fast_tail_kernel(const ArgusFastHdr* __restrict__ hdr,
                                                              const ArgusExtractHdr* __restrict__ xh)
{
    const int i = threadIdx.x;
    const int n = xh->n_new;
    float2* tail = hdr->tail;
    float2 v = make_float2(0.0f, 0.0f);
    if (i < kTail) v = i + n >= kTail ? xh->in[n - kTail + i] : tail[i + n];
    __syncthreads();  // every read of the old tail before any write
    if (i < kTail) tail[i] = v;
}

// In-place radix-2 FFT of ARGUS_FAST_N points in FP64 (host), forward; the
// twiddles e^{-2 pi i k / N} are computed once.
void fft64(std::vector<std::complex<double>>& a)
{
    static const std::vector<std::complex<double>> tw = [] {
        std::vector<std::complex<double>> w((size_t)kN / 2);
        for (int k = 0; k < kN / 2; k++) {
            const double ang = -2.0 * M_PI * (double)k / (double)kN;
            w[(size_t)k] = std::complex<double>(std::cos(ang), std::sin(ang));
        }
        return w;
    }();
    const int n = kN;
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[(size_t)i], a[(size_t)j]);
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < len / 2; k++) {
                const std::complex<double> u = a[(size_t)(i + k)], v = a[(size_t)(i + k + len / 2)] * tw[(size_t)(k * step)];
                a[(size_t)(i + k)] = u + v;
                a[(size_t)(i + k + len / 2)] = u - v;
            }
        }
    }
}

}  // namespace

void argus_fast_spectrum(const float* h, int T, uint64_t inc, float2* G)
{
    std::vector<std::complex<double>> a((size_t)kN, 0.0);
    for (int k = 0; k < T && k < kN; k++) {
        // turn fraction of inc * k modulo one turn, exactly, then FP64
        const double turns = (double)(int64_t)(inc * (uint64_t)k) * (1.0 / 18446744073709551616.0);
        const double ang = -2.0 * M_PI * turns;
        a[(size_t)k] = (double)h[k] * std::complex<double>(std::cos(ang), std::sin(ang));
    }
    fft64(a);
    for (int k = 0; k < kN; k++) {
        G[k].x = (float)(a[(size_t)k].real() / kN);
        G[k].y = (float)(a[(size_t)k].imag() / kN);
    }
}

int argus_launch_fast_input(const ArgusFastHdr* d_hdr, const ArgusExtractHdr* d_xh, int S, float2* u, float2* X,
                            hipStream_t stream)
{
    if (S <= 0) return 0;
    const int len = (S - 1) * kHop + kN;
    argus_clear_error();
    argus_fast_stage_kernel<<<(len + kThreads - 1) / kThreads, kThreads, 0, stream>>>(d_hdr, d_xh);
    ARGUS_CHECK(hipGetLastError());
    return argus_launch_fft4096(u, kHop, S, X, stream);
}

int argus_launch_fast_chans(const ArgusFastHdr* d_hdr, int S, int grid, int part, hipStream_t stream)
{
    if (S <= 0) return 0;
    if (part >= ARGUS_FAST_PARTS) return -1;
    argus_clear_error();
    argus_fast_chan_kernel<<<grid > 0 ? grid : ARGUS_FAST_GRID_DEFAULT, kChanThreads, 0, stream>>>(d_hdr, part);
    return (int)hipGetLastError();
}

int argus_launch_fast(const ArgusFastHdr* d_hdr, const ArgusExtractHdr* d_xh, int S, float2* u, float2* X,
                      int grid, hipStream_t stream)
{
    const int err = argus_launch_fast_input(d_hdr, d_xh, S, u, X, stream);
    return err ? err : argus_launch_fast_chans(d_hdr, S, grid, -1, stream);
}

int argus_launch_fast_tail(const ArgusFastHdr* d_hdr, const ArgusExtractHdr* d_xh, hipStream_t stream)
{
    argus_clear_error();
    argus_fast_tail_kernel<<<1, 512, 0, stream>>>(d_hdr, d_xh);
    return (int)hipGetLastError();
}

// Host helper: argus_fast_spectrum for a phase increment in radians per
// sample, converted as the service converts a channel's offset.
extern "C" GPUD_API int gpud_fast_spectrum(const float* taps, int num_taps, double phase_inc, gpud_float2* G)
{
    if (!taps || !G || num_taps <= 0 || num_taps > ARGUS_FAST_TAIL + 1) return -1;
    argus_fast_spectrum(taps, num_taps, argus_rad_to_fix(phase_inc), (float2*)G);
    return 0;
}
