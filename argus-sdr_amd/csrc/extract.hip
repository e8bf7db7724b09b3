// extract.hip -- fused streaming extraction kernels and the stateful
// single-signal entry points (successors of exports.cu:356-631).
#include "extract.hpp"
#include <algorithm>
#include <type_traits>

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

// v + (v moved by a DPP lane pattern); rows outside row_mask add 0.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ float argus_dpp_add(float v)
{
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, false));
}

// Sum over aligned groups of S lanes (S = 1, 2, 4, 8, 16) with DPP
// (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror);
// every lane of a group receives the group's sum.
template <int S>
__device__ __forceinline__ float argus_group_sum(float v)
{
    if constexpr (S >= 2) v = argus_dpp_add<0xB1>(v);
    if constexpr (S >= 4) v = argus_dpp_add<0x4E>(v);
    if constexpr (S >= 8) v = argus_dpp_add<0x141>(v);
    if constexpr (S >= 16) v = argus_dpp_add<0x140>(v);
    return v;
}

// Sum over the wavefront; the total is valid in lane 63 (row_bcast15 into
// rows 1 and 3, then row_bcast31 into rows 2 and 3).
__device__ __forceinline__ float argus_wave_sum63(float v)
{
    v = argus_group_sum<16>(v);
    v = argus_dpp_add<0x142, 0xA>(v);
    v = argus_dpp_add<0x143, 0xC>(v);
    return v;
}

// Sum over aligned groups of S lanes (power of two <= 64); the sum is valid
// in the first lane of each group for S <= 16 (every lane), and in lane 63 of
// the wave when S == 64.
__device__ __forceinline__ float argus_lanes_sum(float v, int S)
{
    switch (S) {
    case 1: return v;
    case 2: return argus_group_sum<2>(v);
    case 4: return argus_group_sum<4>(v);
    case 8: return argus_group_sum<8>(v);
    case 16: return argus_group_sum<16>(v);
    default:
        for (int o = S >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o);
        return v;
    }
}

// Dense tile: the outputs m0 .. m0+M-1 of one channel, staged and filtered
// in LDS (layout in extract.hpp).
__device__ __forceinline__ void argus_dense_tile(const ArgusChanDesc& ch, const float2* __restrict__ in, int m0,
                                                 float2* lds_s)
{
    const int D = ch.D, L = ch.L, PS = ch.PS;
    const int Lp = (L + 3) & ~3;
    const int Gt = ch.M >> 2;                // output groups of a full tile: fixes the layout
    const int M = min(ch.M, ch.n_out - m0);  // outputs of this tile
    const int ytop = 4 * Gt + Lp - 2;        // reversed position of position 0
    const int em0 = ch.e0 + m0 * D;
    float* lds_t = (float*)(lds_s + ch.SP);
    // Padded taps: plane p, slot k < Lp holds taps[p*L + k] (k < L) or 0.
    // With Lp <= blockDim each thread keeps one slot k and steps over planes
    // (one division per thread); otherwise threads step over the slots of
    // each plane.
    auto load_taps = [&]() {
        const int B = blockDim.x;
        if (Lp <= B) {
            const int per = B / Lp;
            const int p0 = (int)threadIdx.x / Lp, k = (int)threadIdx.x - p0 * Lp;
            if (p0 < per)
                for (int p = p0; p < D; p += per) lds_t[p * Lp + k] = k < L ? ch.taps[p * L + k] : 0.0f;
        } else {
            for (int p = 0; p < D; p++)
                for (int k = threadIdx.x; k < Lp; k += B) lds_t[p * Lp + k] = k < L ? ch.taps[p * L + k] : 0.0f;
        }
    };

    // Staging: positions 0 .. M+Lp-2 of every plane, i.e. the span of
    // (M + Lp - 1) * D samples from n_lo; sample i goes to plane D-1-(i mod D)
    // at reversed position ytop - i div D, word y + y div 4.  Loads are
    // issued kBatch at a time; the NCO is evaluated exactly at the first
    // sample of each batch and advanced across the batch by complex
    // multiplication (at most kBatch-1 steps, rounding drift below 1e-6).
    // Tiles wholly in the new input skip the history checks.
    const int span = (M + Lp - 1) * D;
    const int n_lo = em0 - (Lp * D - 1);
    const int B = blockDim.x;
    constexpr int kBatch = 8;
    const bool all_new = n_lo >= 0;
    if (all_new && D <= B) {
        // Threads [0, Bp) with Bp a multiple of D: each stays in one plane
        // and its reversed position drops by sq = Bp / D per sample, by
        // kBatch*sq (a multiple of 4) per batch -- so y mod 4, and with it the
        // word offsets of a batch's samples from its first, are the same for
        // every batch, and the first word drops by 10*sq per batch.  The input
        // is read through a buffer resource (lane offset in a vector register,
        // batch and sample offsets in scalar registers; reads past the span
        // return 0 and are not stored).
        static_assert(kBatch == 8, "word step per batch assumes 8 samples");
        const int Bp = (B / D) * D;
        if ((int)threadIdx.x < Bp) {
            const int t = threadIdx.x;
            const int sq = B / D;
            const int x0 = t / D;
            const int y0 = ytop - x0;
            const int w0 = y0 + (y0 >> 2);
            int off[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; u++) {
                const int y = y0 - u * sq;
                off[u] = y + (y >> 2) - w0;
            }
            float2* w = lds_s + (D - 1 - (t - x0 * D)) * PS + w0;
            const __amdgpu_buffer_rsrc_t rs =
                __builtin_amdgcn_make_buffer_rsrc((void*)(in + n_lo), (short)0, span * (int)sizeof(float2), 0x00020000);
            const int voff = t * (int)sizeof(float2);
            uint64_t ph = ch.phase0 + ch.inc * (uint64_t)(int64_t)(n_lo + t);
            const uint64_t phstep = ch.inc * (uint64_t)Bp;
            float sb_s, sb_c;
            argus_nco(phstep, sb_s, sb_c);
            auto load = [&](float2* v, int first) {  // first: span index of the batch's sample 0 (thread 0)
#pragma unroll
                for (int u = 0; u < kBatch; u++)
                    v[u] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(
                                                          rs, voff, (first + u * Bp) * (int)sizeof(float2), 0));
            };
            auto store = [&](const float2* v, int first, auto checked) {
                float sn, cs;
                argus_nco(ph, sn, cs);
#pragma unroll
                for (int u = 0; u < kBatch; u++) {
                    if (!decltype(checked)::value || t + first + u * Bp < span) w[off[u]] = argus_rot(v[u], sn, cs);
                    const float c2 = cs * sb_c - sn * sb_s;
                    sn = fmaf(sn, sb_c, cs * sb_s);
                    cs = c2;
                }
                w -= 10 * sq;
                ph += (uint64_t)kBatch * phstep;
            };
            // The first two batches and the taps are in flight together, so a
            // typical tile (<= 16 samples per thread) waits on memory once.
            const int nfull = span / (kBatch * Bp);  // batches wholly inside the span for every thread
            const int nb_all = (span + kBatch * Bp - 1) / (kBatch * Bp);
            float2 va[kBatch], vb[kBatch];
            load(va, 0);
            if (nb_all > 1) load(vb, kBatch * Bp);
            load_taps();
            if (nfull > 0)
                store(va, 0, std::false_type{});
            else
                store(va, 0, std::true_type{});
            if (nb_all > 1) {
                if (nfull > 1)
                    store(vb, kBatch * Bp, std::false_type{});
                else
                    store(vb, kBatch * Bp, std::true_type{});
            }
            for (int bt = 2; bt < nb_all; bt++) {
                load(va, bt * kBatch * Bp);
                if (bt < nfull)
                    store(va, bt * kBatch * Bp, std::false_type{});
                else
                    store(va, bt * kBatch * Bp, std::true_type{});
            }
        } else {
            load_taps();
        }
    } else {
        load_taps();
    }
    if (!(all_new && D <= B) && (int)threadIdx.x < span) {
        // General order: thread t takes samples t, t+B, ...; plane residue r
        // and reversed position y are stepped without dividing again.
        int i = threadIdx.x;
        const int stepq = B / D, stepr = B - stepq * D;
        const int x0 = i / D;
        int r = i - x0 * D;
        int y = ytop - x0;
        uint64_t ph = ch.phase0 + ch.inc * (uint64_t)(int64_t)(n_lo + i);
        const uint64_t phstep = ch.inc * (uint64_t)B;
        auto put = [&](float2 v) {
            lds_s[(D - 1 - r) * PS + y + (y >> 2)] = v;
            ph += phstep;
            r += stepr;
            y -= stepq;
            if (r >= D) {
                r -= D;
                y -= 1;
            }
            i += B;
        };
        if (all_new) {
            float st_s, st_c;
            argus_nco(phstep, st_s, st_c);
            const float2* src = in + n_lo;
            while (i + (kBatch - 1) * B < span) {
                float2 v[kBatch];
#pragma unroll
                for (int u = 0; u < kBatch; u++) v[u] = src[i + u * B];
                float sn, cs;
                argus_nco(ph, sn, cs);
#pragma unroll
                for (int u = 0; u < kBatch; u++) {
                    put(argus_rot(v[u], sn, cs));
                    const float c2 = cs * st_c - sn * st_s;
                    sn = fmaf(sn, st_c, cs * st_s);
                    cs = c2;
                }
            }
            while (i < span) {
                float sn, cs;
                argus_nco(ph, sn, cs);
                put(argus_rot(src[i], sn, cs));
            }
        } else {
            while (i < span) {
                const int n = n_lo + i;
                float2 x = make_float2(0.0f, 0.0f);
                if (n >= 0) {
                    float sn, cs;
                    argus_nco(ph, sn, cs);
                    x = argus_rot(in[n], sn, cs);
                } else if (ch.H + n >= 0) {
                    x = ch.hist_rd[ch.H + n];  // history samples are stored already shifted
                }
                put(x);
            }
        }
    }
    __syncthreads();

    // FIR: output group q = t / S (outputs 4q .. 4q+3), lanes g = t % S
    // share its planes p = g, g+S, ...  Block kc (taps 4kc .. 4kc+3) reads
    // the samples v_t at words 5(Gt-1-q+kc) + {0,1,2,3,5,6,7}; output r at
    // tap 4kc+i uses v_{3-r+i}.
    const int S = ch.S;
    const int t = threadIdx.x;
    const int q = t >> (__ffs(S) - 1);
    const int g = t & (S - 1);
    const int Gm = (M + 3) >> 2;
    float ar[4], ai[4];
#pragma unroll
    for (int r = 0; r < 4; r++) ar[r] = ai[r] = 0.0f;
    if (q < Gm) {
        const int nb = Lp >> 2;
        const float2* lane = lds_s + 5 * (Gt - 1 - q);
        for (int p = g; p < D; p += S) {
            const float2* b = lane + p * PS;
            const float4* h4 = (const float4*)(lds_t + p * Lp);
#pragma unroll 2
            for (int kc = 0; kc < nb; kc++) {
                const float2 v0 = b[0], v1 = b[1], v2 = b[2], v3 = b[3], v4 = b[5], v5 = b[6], v6 = b[7];
                const float4 H = h4[kc];
#define ARGUS_TAP(HC, A, B_, C, E)                                                    \
    ar[3] = fmaf(HC, A.x, ar[3]); ai[3] = fmaf(HC, A.y, ai[3]);                       \
    ar[2] = fmaf(HC, B_.x, ar[2]); ai[2] = fmaf(HC, B_.y, ai[2]);                     \
    ar[1] = fmaf(HC, C.x, ar[1]); ai[1] = fmaf(HC, C.y, ai[1]);                       \
    ar[0] = fmaf(HC, E.x, ar[0]); ai[0] = fmaf(HC, E.y, ai[0]);
                ARGUS_TAP(H.x, v0, v1, v2, v3)
                ARGUS_TAP(H.y, v1, v2, v3, v4)
                ARGUS_TAP(H.z, v2, v3, v4, v5)
                ARGUS_TAP(H.w, v3, v4, v5, v6)
#undef ARGUS_TAP
                b += 5;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        ar[r] = argus_lanes_sum(ar[r], S);
        ai[r] = argus_lanes_sum(ai[r], S);
    }
    if (q < Gm && g == 0) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int m = 4 * q + r;
            if (m < M) ch.out[m0 + m] = make_float2(ar[r], ai[r]);
        }
    }
    __syncthreads();
}

// Narrow group: one wavefront per output, lanes over taps, no staging.  Each
// lane issues its loads four at a time (taps k, k+64, k+128, k+192) before
// rotating them; items whose windows lie wholly in the new input skip the
// history checks.  The wave sum lands in lane 63, which writes the output.
__device__ __forceinline__ void argus_narrow_group(const ArgusChanDesc& ch, const float2* __restrict__ in, int m0,
                                                   float* lds_t)
{
    const int T = ch.T, D = ch.D, L = ch.L;
    // natural tap order from the phase-major layout: h[k] = taps[(k mod D) L + k div D]
    for (int i = threadIdx.x; i < T; i += blockDim.x) lds_t[i] = L == 1 ? ch.taps[i] : ch.taps[(i % D) * L + i / D];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    const int mend = min(m0 + ch.M, ch.n_out);
    const uint64_t inc64 = ch.inc * 64u;
    const bool all_new = ch.e0 + m0 * D - (T - 1) >= 0;
    for (int m = m0 + wave; m < mend; m += nw) {
        const int e = ch.e0 + m * D;
        float ar = 0.0f, ai = 0.0f;
        for (int k0 = 0; k0 < T; k0 += 256) {
            float2 x[4];
            if (all_new) {
                const float2* src = in + (e - k0 - lane);
#pragma unroll
                for (int u = 0; u < 4; u++) x[u] = k0 + lane + 64 * u < T ? src[-64 * u] : make_float2(0.0f, 0.0f);
            } else {
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int k = k0 + lane + 64 * u;
                    const int n = e - k;
                    float2 v = make_float2(0.0f, 0.0f);
                    if (k < T) {
                        if (n >= 0) {
                            v = in[n];
                        } else {
                            const int h = ch.H + n;
                            if (h >= 0) v = ch.hist_rd[h];
                        }
                    }
                    x[u] = v;
                }
            }
            uint64_t ph = ch.phase0 + ch.inc * (uint64_t)(int64_t)(e - k0 - lane);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int k = k0 + lane + 64 * u;
                if (k < T) {
                    float2 v = x[u];
                    if (all_new || e - k >= 0) {  // history samples are stored already shifted
                        float sn, cs;
                        argus_nco(ph, sn, cs);
                        v = argus_rot(v, sn, cs);
                    }
                    const float h = lds_t[k];
                    ar = fmaf(h, v.x, ar);
                    ai = fmaf(h, v.y, ai);
                }
                ph -= inc64;
            }
        }
        ar = argus_wave_sum63(ar);
        ai = argus_wave_sum63(ai);
        if (lane == 63) ch.out[m] = make_float2(ar, ai);
    }
    __syncthreads();
}

// Narrow group with modulated taps: four outputs per wavefront, one per
// 16-lane row.  Waves whose padded windows (T rounded up to 128, zero taps
// beyond T) lie wholly in the new input run a branch-free loop in which lane
// r of a row takes the tap pairs (k, k + 1), k = 2 r + 32 i: each load
// brings two consecutive samples (16 bytes), which halves the vector memory
// instructions per tap -- the address unit is what bounds these items.
// Otherwise each tap (k = r + 16 i) checks its sample, and samples before
// the chunk (the history, stored already shifted) are filtered with the real
// taps and added after the rotation.  The row sum lands in every lane of the
// row; lane 15 writes.
struct ArgusPair {
    float2 a, b;  // two consecutive complex samples (taps)
};
// The same 16 bytes as one vector, aligned as a complex sample (8 bytes).
typedef float ArgusPairV4 __attribute__((ext_vector_type(4)));
typedef ArgusPairV4 ArgusPairV __attribute__((aligned(8)));

template <typename V>
__device__ __forceinline__ ArgusPair argus_pair(const V& v)
{
    return ArgusPair{make_float2(v.x, v.y), make_float2(v.z, v.w)};
}

__device__ __forceinline__ void argus_narrow_mod(const ArgusChanDesc& ch, const float2* __restrict__ in, int m0,
                                                 float2* lds)
{
    const int T = ch.T, D = ch.D, L = ch.L;
    const int TP = (T + 127) & ~127;
    float2* gs = lds;
    float* hs = (float*)(lds + TP);
    float2* ys = lds + TP + (T + 1) / 2;  // the item's outputs, for the demodulation epilogue
    const bool hist = ch.e0 + m0 * D - (T - 1) < 0;  // a window of this item reaches into the history
    for (int i = threadIdx.x; i < TP; i += blockDim.x) gs[i] = i < T ? ch.ctaps[i] : make_float2(0.0f, 0.0f);
    if (hist)
        for (int i = threadIdx.x; i < T; i += blockDim.x) hs[i] = L == 1 ? ch.taps[i] : ch.taps[(i % D) * L + i / D];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int row = lane >> 4, rl = lane & 15;
    const int nw = blockDim.x >> 6;
    const int mend = min(m0 + ch.M, ch.n_out);
#define ARGUS_CMAC(G, X)                                                        \
    ar = fmaf(G.x, X.x, ar); ar = fmaf(-G.y, X.y, ar);                          \
    ai = fmaf(G.x, X.y, ai); ai = fmaf(G.y, X.x, ai);
    for (int mb = m0 + 4 * (int)(threadIdx.x >> 6); mb < mend; mb += 4 * nw) {
        const int m = mb + row;
        const int e = ch.e0 + m * D;
        float ar = 0.0f, ai = 0.0f, br = 0.0f, bi = 0.0f;
        if (ch.e0 + mb * D - (TP - 1) >= 0) {  // same for the whole wave: its lowest output decides
            if (m < mend) {
                // pair u of this lane: taps k, k + 1 with k = k0 + 2 rl + 32 u,
                // samples x[e - k - 1], x[e - k] in one 16-byte load (the
                // lowest pair is the base, so the others sit at immediate
                // offsets above it); tap pairs as 16-byte LDS reads
                // (global address space: loads with immediate offsets)
                typedef __attribute__((address_space(1))) const ArgusPairV GlobalPair;
                GlobalPair* src = (GlobalPair*)(in + (e - 2 * rl - 1)) - 48;
                const float4* gp = (const float4*)(gs + 2 * rl);
                for (int k = 0; k < TP; k += 128) {
                    const int q = k >> 1;  // in pairs
                    GlobalPair* b = src - q;
                    const ArgusPairV v3 = b[0], v2 = b[16], v1 = b[32], v0 = b[48];
                    const ArgusPair x0 = argus_pair(v0), x1 = argus_pair(v1), x2 = argus_pair(v2),
                                    x3 = argus_pair(v3);
                    const float4 f0 = gp[q], f1 = gp[q + 16], f2 = gp[q + 32], f3 = gp[q + 48];
                    const ArgusPair g0 = argus_pair(f0), g1 = argus_pair(f1), g2 = argus_pair(f2),
                                    g3 = argus_pair(f3);
                    ARGUS_CMAC(g0.a, x0.b) ARGUS_CMAC(g0.b, x0.a)
                    ARGUS_CMAC(g1.a, x1.b) ARGUS_CMAC(g1.b, x1.a)
                    ARGUS_CMAC(g2.a, x2.b) ARGUS_CMAC(g2.b, x2.a)
                    ARGUS_CMAC(g3.a, x3.b) ARGUS_CMAC(g3.b, x3.a)
                }
            }
        } else if (m < mend) {
            for (int k = rl; k < T; k += 16) {
                const int n = e - k;
                if (n >= 0) {
                    const float2 x = in[n];
                    const float2 g = gs[k];
                    ARGUS_CMAC(g, x)
                } else if (ch.H + n >= 0) {
                    const float2 v = ch.hist_rd[ch.H + n];
                    br = fmaf(hs[k], v.x, br);
                    bi = fmaf(hs[k], v.y, bi);
                }
            }
        }
        ar = argus_group_sum<16>(ar);
        ai = argus_group_sum<16>(ai);
        if (hist) {
            br = argus_group_sum<16>(br);
            bi = argus_group_sum<16>(bi);
        }
        if (rl == 15 && m < mend) {
            float sn, cs;
            argus_nco(ch.phase0 + ch.inc * (uint64_t)(int64_t)e, sn, cs);
            const float2 y = argus_rot(make_float2(ar, ai), sn, cs);
            const float2 o = make_float2(y.x + br, y.y + bi);
            ch.out[m] = o;
            if (ch.demod) ys[m - m0] = o;
        }
    }
#undef ARGUS_CMAC
    if (ch.demod) {  // uniform: one channel per workgroup
        __syncthreads();
        // the same arithmetic as the demodulation pass (demod.hip), so the
        // same bits; an FM item's first output needs the previous item's
        // last, and is left to that pass with the carry
        const bool fm = ch.demod == ARGUS_DEMOD_NFM || ch.demod == ARGUS_DEMOD_WFM;
        for (int j = threadIdx.x + (fm ? 1 : 0); j < mend - m0; j += blockDim.x) {
            const int m = m0 + j;
            const float2 v = ys[j];
            if (fm) {
                ch.audio[m - ch.skip] = argus_fm_step(ys[j - 1], v);
            } else {
                float s, c;
                argus_nco(ch.bfo0 + ch.bfo_inc * (uint64_t)(uint32_t)m, s, c);
                ch.audio[m] = ch.demod == ARGUS_DEMOD_CW ? v.x * c : argus_ssb_out(v, s, c);
            }
        }
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
    // a work list built for the most outputs a chunk can give may hold one
    // tile past this chunk's last output
    if ((kind == ARGUS_ITEM_DENSE || kind == ARGUS_ITEM_NARROW) && m0 >= ch.n_out) return;
    if (kind == ARGUS_ITEM_DENSE)
        argus_dense_tile(ch, in, m0, lds);
    else if (kind == ARGUS_ITEM_NARROW) {
        if (ch.ctaps)
            argus_narrow_mod(ch, in, m0, lds);
        else
            argus_narrow_group(ch, in, m0, (float*)lds);
    }
    else if (kind == ARGUS_ITEM_HISTORY)
        argus_history_tail(ch, in, n_new);
}

// Batched, persistent: every work item of every channel of one chunk.
__device__ __forceinline__ void argus_extract_items(const ArgusChanDesc* __restrict__ chans,
                                                    const ArgusWorkItem* __restrict__ items, int n_items,
                                                    const float2* __restrict__ in, int n_new, float2* lds)
{
    // The tables are written before the launch and only read during it, so
    // they are read through the constant address space: scalar loads even
    // when the table pointers themselves come from device memory.
#if defined(__HIP_DEVICE_COMPILE__)
    typedef const __attribute__((address_space(4))) ArgusChanDesc CChan;
    typedef const __attribute__((address_space(4))) ArgusWorkItem CItem;
    CChan* cchans = (CChan*)chans;
    CItem* citems = (CItem*)items;
#else  // the host pass only needs the function to compile
    const ArgusChanDesc* cchans = chans;
    const ArgusWorkItem* citems = items;
#endif
    for (int w = blockIdx.x; w < n_items; w += gridDim.x) {
        const ArgusWorkItem it = citems[w];
        const ArgusChanDesc ch = cchans[it.chan];
        argus_run_item(ch, it.kind, it.m0, in, n_new, lds);
    }
}

__global__ __launch_bounds__(ARGUS_EXTRACT_THREADS) void argus_extract_batch_kernel(
    const ArgusChanDesc* __restrict__ chans, const ArgusWorkItem* __restrict__ items, int n_items,
    const float2* __restrict__ in, int n_new)
{
    extern __shared__ __attribute__((aligned(16))) float2 lds[];
    argus_extract_items(chans, items, n_items, in, n_new, lds);
}

// The same launch with its arguments read from device memory, so a captured
// graph replays it unchanged while the block's tables change.
__global__ __launch_bounds__(ARGUS_EXTRACT_THREADS) void argus_extract_hdr_kernel(const ArgusExtractHdr* __restrict__ hdr)
{
    extern __shared__ __attribute__((aligned(16))) float2 lds[];
    const ArgusExtractHdr h = *hdr;
    argus_extract_items(h.chans, h.items, h.n_items, h.in, h.n_new, lds);
}

// One signal: blocks [0, n_items) are output items, block n_items (if
// launched) writes the new history.
__global__ __launch_bounds__(ARGUS_EXTRACT_THREADS) void argus_extract_one_kernel(ArgusChanDesc ch,
                                                                                   const float2* __restrict__ in,
                                                                                   int n_new, int n_items)
{
    extern __shared__ __attribute__((aligned(16))) float2 lds[];
    const int kind = (int)blockIdx.x < n_items ? (ch.narrow ? ARGUS_ITEM_NARROW : ARGUS_ITEM_DENSE)
                                               : ARGUS_ITEM_HISTORY;
    argus_run_item(ch, kind, blockIdx.x * ch.M, in, n_new, lds);
}

size_t argus_item_lds_bytes(const ArgusChanDesc& ch)
{
    if (!ch.narrow) return argus_dense_lds_bytes(ch);
    // modulated taps (T rounded up to 128), real taps, the item's outputs
    return ch.ctaps ? (size_t)(((ch.T + 127) & ~127) + (ch.T + 1) / 2 + ch.M) * sizeof(float2)
                    : (size_t)ch.T * sizeof(float);
}

// Dynamic LDS above 64 KiB must be allowed per kernel once.
static int argus_allow_big_lds()
{
    static int done = 0;
    if (done) return 0;
    const int bytes = 160 * 1024;
    hipError_t e = hipFuncSetAttribute((const void*)argus_extract_batch_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)argus_extract_hdr_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                bytes);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void*)argus_extract_one_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                bytes);
    if (e != hipSuccess) return (int)e;
    done = 1;
    return 0;
}

// Host: launch the single-signal kernel for a fully described channel.
int argus_launch_extract_one(const ArgusChanDesc& ch, const float2* in, int n_new, hipStream_t stream)
{
    const int n_items = ch.n_out > 0 ? (ch.n_out + ch.M - 1) / ch.M : 0;
    const int blocks = n_items + (ch.keep > 0 ? 1 : 0);
    if (blocks == 0) return 0;
    const size_t lds = n_items > 0 ? argus_item_lds_bytes(ch) : 0;
    if (lds > 64 * 1024) {
        int e = argus_allow_big_lds();
        if (e) return e;
    }
    argus_clear_error();
    hipLaunchKernelGGL(argus_extract_one_kernel, dim3(blocks), dim3(ARGUS_EXTRACT_THREADS), lds, stream, ch, in,
                       n_new, n_items);
    return (int)hipGetLastError();
}

int argus_extract_allow_lds(size_t bytes) { return bytes > 64 * 1024 ? argus_allow_big_lds() : 0; }

size_t argus_tune_max_lds(const ArgusTune& tune, int max_taps)
{
    // dense: P*PS <= lds_entries words of samples plus P*L <= T + P - 1 < 2*T taps;
    // narrow: T (+ 127 padding) complex and T real taps, and its outputs
    const size_t dense = (size_t)tune.lds_entries * sizeof(float2) + (size_t)2 * max_taps * sizeof(float);
    const size_t narrow = (size_t)(max_taps + 127 + (max_taps + 1) / 2 + tune.narrow_outputs) * sizeof(float2);
    return std::max(dense, narrow);
}

int argus_launch_extract_hdr(const ArgusExtractHdr* d_hdr, size_t lds_bytes, int grid, hipStream_t stream)
{
    if (lds_bytes > 64 * 1024) {
        int e = argus_allow_big_lds();
        if (e) return e;
    }
    argus_clear_error();
    hipLaunchKernelGGL(argus_extract_hdr_kernel, dim3(grid), dim3(ARGUS_EXTRACT_THREADS), lds_bytes, stream, d_hdr);
    return (int)hipGetLastError();
}

int argus_launch_extract_batch(const ArgusChanDesc* d_chans, const ArgusWorkItem* d_items, int n_items,
                               size_t lds_bytes, const float2* in, int n_new, int grid_max, hipStream_t stream)
{
    if (n_items <= 0) return 0;
    // Persistent grid: enough blocks to fill every CU several times over; the
    // item list is ordered heaviest first so the tail is short.
    const int grid = n_items < grid_max ? n_items : grid_max;
    if (lds_bytes > 64 * 1024) {
        int e = argus_allow_big_lds();
        if (e) return e;
    }
    argus_clear_error();
    hipLaunchKernelGGL(argus_extract_batch_kernel, dim3(grid), dim3(ARGUS_EXTRACT_THREADS), lds_bytes, stream,
                       d_chans, d_items, n_items, in, n_new);
    return (int)hipGetLastError();
}

// Append the work items of one described channel, each with its cost
// estimate (in `pad`) for argus_order_items.
void argus_append_items(const ArgusChanDesc& ch, int chan_index, std::vector<ArgusWorkItem>& items)
{
    const int tiles = ch.n_out > 0 ? (ch.n_out + ch.M - 1) / ch.M : 0;
    // cost in lane-operations: NCO rotations (~24 each) plus taps (~2 each,
    // ~6 for a complex tap with its load)
    long cost;
    if (!ch.narrow) {
        const int Lp = argus_dense_lp(ch.L);
        cost = (long)(ch.M + Lp - 1) * ch.D * 10 + (long)ch.M * Lp * ch.D * 2;
    } else if (ch.ctaps) {
        cost = (long)ch.M * (((ch.T + 63) & ~63) * 6 + 64);
    } else {
        cost = (long)ch.M * ch.T * 26;
    }
    const int kind = ch.narrow ? ARGUS_ITEM_NARROW : ARGUS_ITEM_DENSE;
    for (int t = 0; t < tiles; t++) items.push_back({chan_index, t * ch.M, kind, (int)std::min<long>(cost, 1L << 30)});
    if (ch.keep > 0) items.push_back({chan_index, 0, ARGUS_ITEM_HISTORY, ch.keep * 24});
}

// Order the items of one launch for the persistent grid.  Workgroups are
// dealt to the XCDs round-robin, and the grid is a multiple of ARGUS_XCDS, so
// item w always runs on XCD w % ARGUS_XCDS.  Each channel's items go to one
// XCD (channels packed heaviest first onto the least loaded XCD), heaviest
// items first within an XCD, and the XCDs' lists are interleaved: a channel's
// taps are then fetched into one L2, not eight (the 3.2 MB block, which every
// XCD reads, stays in each L2).  Short lists are padded with no-op items.
//
// "Heaviest first" is by cost class (half-octaves of the cost estimate), a
// stable counting sort: O(items) host time per work-list build, which
// matters at thousands of channels (a comparison sort of ~10^5 items costs
// milliseconds); the channel packing sorts channels, not items.
void argus_order_items(std::vector<ArgusWorkItem>& items)
{
    constexpr int kClasses = 64;
    constexpr int kKeys = ARGUS_XCDS * kClasses;
    thread_local std::vector<int> key;
    thread_local std::vector<ArgusWorkItem> sorted;
    int count[kKeys + 1] = {0};
    const size_t n = items.size();
    key.resize(n);
    // each channel's items on one XCD (its taps are then fetched into one L2,
    // not eight): channels packed onto the XCDs heaviest first, each to the
    // least loaded one
    thread_local std::vector<long> ccost;
    thread_local std::vector<int> corder, cbin;
    int n_ch = 0;
    for (size_t i = 0; i < n; i++) n_ch = std::max(n_ch, items[i].chan + 1);
    ccost.assign((size_t)n_ch, 0);
    for (size_t i = 0; i < n; i++) ccost[(size_t)items[i].chan] += std::max(items[i].pad, 1);
    corder.resize((size_t)n_ch);
    for (int c = 0; c < n_ch; c++) corder[(size_t)c] = c;
    std::sort(corder.begin(), corder.end(), [&](int a, int b) { return ccost[(size_t)a] > ccost[(size_t)b]; });
    cbin.assign((size_t)n_ch, 0);
    long load[ARGUS_XCDS] = {0};
    for (int c : corder) {
        int best = 0;
        for (int x = 1; x < ARGUS_XCDS; x++)
            if (load[x] < load[best]) best = x;
        cbin[(size_t)c] = best;
        load[best] += ccost[(size_t)c];
    }
    for (size_t i = 0; i < n; i++) {
        const int b = cbin[(size_t)items[i].chan];
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
                    const float2* ctaps, float2* out, int D, int T, int phase_count, int n_new, double phase_start,
                    double phase_inc, const ArgusTune& tune)
{
    if (T > ARGUS_MAX_TAPS) return -1;
    ch->hist_rd = hist_rd;
    ch->hist_wr = hist_wr;
    ch->taps = taps;
    ch->ctaps = ctaps;
    ch->out = out;
    ch->phase0 = argus_rad_to_fix(phase_start);
    ch->inc = argus_rad_to_fix(phase_inc);
    ch->D = D;
    ch->T = T;
    ch->L = (T + D - 1) / D;
    ch->P = D < T ? D : T;
    ch->H = H;
    ch->e0 = D - phase_count - 1;
    ch->keep = keep;
    ch->demod = 0;  // baseband only unless the caller asks for the epilogue
    ch->skip = 0;
    ch->audio = nullptr;
    ch->bfo0 = ch->bfo_inc = 0;
    // the tile shape depends on the chunk length only (planned for the most
    // outputs a chunk of n_new samples can give), so a work list built for
    // one chunk stays valid for the next of the same length
    ch->n_out = (D - 1 + n_new) / D;
    const int rc = argus_plan_chan(ch, tune, ctaps != nullptr);
    ch->n_out = (phase_count + n_new) / D;
    return rc;
}

void argus_modulated_taps(const float* poly, int D, int T, uint64_t inc, float2* out)
{
    const int L = (T + D - 1) / D;
    for (int k = 0; k < T; k++) {
        const double h = poly[(size_t)(k % D) * L + k / D];
        const double a = argus_fix_to_rad(inc * (uint64_t)k);
        out[k] = make_float2((float)(h * std::cos(a)), (float)(-h * std::sin(a)));
    }
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
                        nullptr, (float2*)out, decim, num_taps, pc, n_new, phase_start, phase_inc, argus_default_tune()) != 0)
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
                        nullptr, (float2*)out, decim, num_taps, pc, n_new, phase_start, phase_inc, argus_default_tune()) != 0)
        return -2;
    int err = argus_launch_extract_one(ch, (const float2*)in_new, n_new, 0);
    if (err) return err;
    if (n_out) *n_out = out_count;
    if (phase_count_out) *phase_count_out = total % decim;
    if (phase_end_out) *phase_end_out = argus_reduce_phase(phase_start + phase_inc * (double)n_new);
    return 0;
}

}  // extern "C"
