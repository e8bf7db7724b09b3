// rds_batch.hip -- device-side RDS group decoder: one decoder per station,
// many stations per call (gpud_rds_batch_*), and the kernels the streaming
// service runs on each block (gpud_serve_set_rds).  Same semantics as the
// host decoder in rds.hip, which follows internal/rds/rds.go Decoder.Decode
// (78-264), tryDecode (365-430), decodeGroupAt (432-473), applyGroup
// (475-502), decodeBlockExpected (549-579) and crcSyndrome (590-599).
//
// Launches, all on the caller's stream:
//
//   argus_rds_symbols_kernel  one wavefront per call; one call per
//       workgroup up to 256 calls, then up to four per workgroup (one per
//       SIMD of a CU).  Mueller & Muller
//       timing with linear interpolation, the Costas loop on each symbol as
//       it is produced, sign(I) hard bits and the differential bits (carried
//       across calls).  The loop is a serial recurrence, run by every lane
//       alike (uniform values, scalar branches) over samples the wavefront
//       stages in LDS 1024 at a time (the next sample
//       index depends on the previous symbol, so a global load per symbol
//       would put HBM latency on the recurrence).  A call may cover a window
//       whose samples arrive over several launches: the loop pauses where it
//       needs a sample not yet there and resumes from its saved registers
//       (ArgusRdsProgress), which gives exactly the symbols of one call over
//       the whole window.  Arithmetic follows the reference's
//       float32/float64 mix operation by operation, with contraction off, so
//       the bits equal the host decoder's.  The Costas oscillator's FP64
//       sine and cosine (fp64_sincos.hpp) round to the same float32 as the C
//       library's on the wrapped phase.  The Costas state and the last
//       hard bit are committed only when the call produced at least 104
//       symbols, as the reference returns before its Costas loop otherwise.
//   argus_rds_match_kernel    one thread per (final call, bit position).
//       Whether a whole group (blocks A, B, C/C', D) starts there, in the
//       direct and in the inverted bit stream, with the unique-single-flip
//       correction.  The CRC is linear over GF(2), so the syndrome of a word
//       with bit i flipped is syndrome(word) ^ syndrome(1 << i): one division
//       per block and 26 table lookups.
//   argus_rds_next_kernel     one workgroup per final call: for every
//       position, the next position at or after it where a group starts
//       (a suffix minimum), per stream.
//   argus_rds_scan_kernel     one lane per final call.  The serial block-sync
//       scan with the +-3-bit flywheel (direct stream, then the inverted one
//       if the direct found nothing).  Where the reference steps one bit at
//       a time while it searches, the scan jumps to the next position with a
//       group, which visits the same groups in the same order.  Each group
//       taken is decoded again at its position and applied to the station's
//       PI, PS and RT on the device, and written to the call's group list.
//
// Deviation: the symbol loop of a call stops after 2 n / sps + 16 symbols
// (the reference has no bound; a symbol step shorter than half a symbol
// period needs baseband amplitudes in the hundreds), and a step that takes
// the timing position below -1 ends the call's symbol loop, where the
// reference would index before the window and panic.
#include "rds.hpp"
#include "fp64_sincos.hpp"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr int kBlockBits = 26;
constexpr int kGroupBits = 4 * kBlockBits;
constexpr int kFlywheel = 3;
constexpr int kThreads = 256;

// Offset words (rds.go:509-515) and the allowed set per block slot.
constexpr uint32_t kOffA = 0x0FC, kOffB = 0x198, kOffC = 0x168, kOffCp = 0x350, kOffD = 0x1B4;

__host__ __device__ inline uint32_t rds_syndrome(uint32_t word)
{
    for (int bit = 25; bit >= 10; --bit)
        if (word & (1u << bit)) word ^= 0x1B9u << (bit - 10);
    return word & 0x3FFu;
}

// Syndromes of the 26 single-bit errors, index i = bit (25 - i) of the word.
struct FlipTable {
    uint16_t e[kBlockBits];
};

FlipTable make_flip_table()
{
    FlipTable t;
    for (int i = 0; i < kBlockBits; i++) t.e[i] = (uint16_t)rds_syndrome(1u << (25 - i));
    return t;
}

__device__ __forceinline__ bool rds_allowed(uint32_t s, int slot)
{
    switch (slot) {
    case 0: return s == kOffA;
    case 1: return s == kOffB;
    case 2: return s == kOffC || s == kOffCp;
    default: return s == kOffD;
    }
}

// rds.go:549-579 with the FEC on: the block's data if its syndrome is an
// allowed offset word, or if exactly one single-bit flip makes it one.
__device__ bool rds_block(uint32_t word, int slot, const FlipTable& ft, uint16_t* data)
{
    const uint32_t s = rds_syndrome(word);
    if (rds_allowed(s, slot)) {
        *data = (uint16_t)(word >> 10);
        return true;
    }
    int found = 0, at = 0;
    for (int i = 0; i < kBlockBits; i++) {
        if (rds_allowed(s ^ ft.e[i], slot)) {
            found++;
            at = i;
        }
    }
    if (found != 1) return false;
    *data = (uint16_t)((word ^ (1u << (25 - at))) >> 10);
    return true;
}

__device__ uint32_t rds_word(const uint8_t* b, uint32_t inv)
{
    uint32_t w = 0;
#pragma unroll
    for (int i = 0; i < kBlockBits; i++) w = (w << 1) | (b[i] & 1u);
    return w ^ inv;
}

__device__ bool rds_group(const uint8_t* b, uint32_t inv, const FlipTable& ft, uint16_t g[4])
{
#pragma unroll 1
    for (int j = 0; j < 4; j++)
        if (!rds_block(rds_word(b + j * kBlockBits, inv), j, ft, &g[j])) return false;
    return true;
}

__device__ char rds_printable(unsigned v) { return (v < 32 || v > 126) ? ' ' : (char)v; }

// ---------------------------------------------------------------- kernels

constexpr int kStage = 1024;     // samples staged in LDS per pass of the symbol loop
constexpr int kSymWaves = 4;     // at most this many calls per workgroup, one wavefront each
constexpr int kSymCUs = 256;     // MI355X compute units

// The staging of a wave's own LDS region is read by the same wave only: a
// wavefront-scope fence and a wave barrier order it (no workgroup barrier,
// so the waves of a workgroup run their calls independently).
__device__ __forceinline__ void rds_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane 0's value of a double that every lane holds alike, as a uniform
// value (SGPRs), so that what follows from it branches on scalars.
__device__ __forceinline__ double rds_uniform(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// blockDim.x / 64 calls per workgroup (1 to kSymWaves). The loop is bound
// by its CU's instruction issue (much of it scalar), so up to one call per CU
// the launch gives each call a workgroup of its own (the dispatcher spreads
// them over the CUs). Beyond that it packs calls into workgroups of up to
// four waves: one workgroup's waves go to the four SIMDs of one CU, which
// keeps the CUs evenly loaded, where many single-wave workgroups stacked
// several calls on some SIMDs.
__global__ __launch_bounds__(64 * kSymWaves) void argus_rds_symbols_kernel(const ArgusRdsCall* __restrict__ calls,
                                                                          int n_calls)
{
#pragma clang fp contract(off)
    // What lane 0 works out at the call's start, for the whole wave.
    struct Start {
        ArgusRdsProgress pr;
        double sps;
        int next_stage;  // first sample of the first pass, or -1
        int carry, carried;
    };
    __shared__ float2 xs_all[kSymWaves][kStage];
    __shared__ Start start_all[kSymWaves];
    // wave-uniform (readfirstlane): the call, its LDS and the whole
    // recurrence are then uniform values and the loop's branches scalar
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ci = blockIdx.x * (int)(blockDim.x >> 6) + wave;
    if (ci >= n_calls) return;  // whole waves only; no workgroup barrier follows
    float2* xs = xs_all[wave];
    // by value: the loop's byte stores to the bits could alias a reference
    // into the table and force a reload per symbol
    const ArgusRdsCall call = calls[ci];
    const int lane = threadIdx.x & 63;
    const int n = call.n, rate = call.rate, avail = call.avail, bit_cap = call.bit_cap;
    if (lane == 0) {
        ArgusRdsProgress pr = *call.pr;
        ArgusRdsState& st = *call.st;
        double sps = 0.0;
        int next_stage = -1;
        if (n >= kGroupBits && rate > 0) {  // rds.go:79-81
            sps = (double)rate / 1187.5;
            if (!pr.active) {  // the call starts: rds.go:84-110
                if (st.last_rate != 0 && st.last_rate != rate) {
                    st.costas_phase = st.costas_freq = 0;
                    st.mu = 0;
                    st.prev_re = st.prev_im = st.prev_dec = 0;
                    st.have_prev = 0;
                    st.last_hard = 0;
                    st.have_last_hard = 0;
                }
                st.last_rate = rate;
                double m = st.mu;
                if (m < 0 || m >= sps || isnan(m) || isinf(m)) m = sps / 2;
                pr.m = m;
                pr.prev_re = st.prev_re;
                pr.prev_im = st.prev_im;
                pr.prev_dec = st.prev_dec;
                pr.have_prev = st.have_prev;
                pr.phase = st.costas_phase;
                pr.freq = st.costas_freq;
                pr.ns = pr.nb = pr.prev_hard = 0;
                pr.loop_done = 0;
                pr.active = 1;
            }
            if (!pr.loop_done && pr.m < (double)(n - 1)) {
                const int k = (int)pr.m;
                if (k >= 0 && k + 1 < avail) next_stage = k;
            }
        }
        // the bit carried into this call (unchanged until the call ends)
        start_all[wave] = Start{pr, sps, next_stage, st.have_last_hard != 0, st.last_hard};
    }
    rds_wave_sync();
    const Start& s0 = start_all[wave];
    ArgusRdsProgress pr = s0.pr;
    const double sps = s0.sps;
    int next_stage = s0.next_stage;
    const bool carry = s0.carry != 0;
    const int carried = s0.carried;
    const double alpha = 0.132, beta = alpha * alpha / 4.0;
    // The sine/cosine constants held in VGPRs for the whole loop (an empty asm
    // hides their values, so the compiler cannot rebuild them in SGPR pairs
    // at every symbol).
    ArgusSinCosK sck = argus_sincos_k();
    asm volatile("" : "+v"(sck.inv_half_pi), "+v"(sck.p1), "+v"(sck.p2));
#pragma unroll
    for (int i = 0; i < 8; i++) asm volatile("" : "+v"(sck.s[i]), "+v"(sck.c[i]));
    // a global (not generic) pointer: a flat store would count on the LDS
    // counter, and the next symbol's LDS read would wait for it to land
    typedef __attribute__((address_space(1))) uint8_t GlobalU8;
    GlobalU8* bits = (GlobalU8*)call.bits;
    // Every lane runs the same recurrence on the same values; the bit stores
    // are the same byte from every lane.
    while (next_stage >= 0) {
        const int k0 = next_stage;
        const int cnt = min(kStage, avail - k0);
        rds_wave_sync();  // the previous pass is done reading xs
        for (int i = lane; i < cnt; i += 64) xs[i] = call.x[k0 + i];
        rds_wave_sync();
        double m = pr.m, phase = pr.phase, freq = pr.freq;
        float prev_re = pr.prev_re, prev_im = pr.prev_im, prev_dec = pr.prev_dec;
        bool have_prev = pr.have_prev != 0;
        int ns = pr.ns, nb = pr.nb, prev_hard = pr.prev_hard;
        int restage = -1;
        bool done = true;  // the loop condition ended the call
        while (m < (double)(n - 1)) {
            const int k = (int)m;
            const double f = m - (double)k;
            if (k + 1 >= n || k < 0 || ns >= bit_cap) break;
            if (k + 1 >= avail) {  // wait for the window's next samples
                done = false;
                break;
            }
            if (k < k0 || k + 1 >= k0 + cnt) {  // outside this pass
                restage = k;
                done = false;
                break;
            }
            const float2 a = xs[k - k0], b = xs[k + 1 - k0];
            const float re = (float)((double)a.x * (1 - f) + (double)b.x * f);
            const float im = (float)((double)a.y * (1 - f) + (double)b.y * f);
            const float dec = re >= 0 ? 1.0f : -1.0f;
            if (have_prev) {
                const double e =
                    (double)(dec - prev_dec) * (double)prev_re - (double)(re - prev_re) * (double)prev_dec;
                m += sps + 0.01 * e;
            } else {
                m += sps;
                have_prev = true;
            }
            prev_dec = dec;
            prev_re = re;
            prev_im = im;
            // Costas de-rotation of this symbol (rds.go:157-183)
            double sd, cd;
            argus_sincos_wrapped_k(phase, sck, &sd, &cd);  // phase stays in [-pi, pi]
            sd = rds_uniform(sd);  // computed from the pinned constants: uniform again
            cd = rds_uniform(cd);
            const float cp = (float)cd, sp = (float)sd;
            const float sre = re * cp + im * sp;
            const float sim = im * cp - re * sp;
            const double err = sre >= 0 ? (double)sim : -(double)sim;
            freq += beta * err;
            phase += freq + alpha * err;
            while (phase > M_PI) phase -= 2 * M_PI;
            while (phase < -M_PI) phase += 2 * M_PI;
            const int hard = sre >= 0 ? 1 : 0;
            // differential bits, carried across calls (rds.go:214-233)
            if (ns == 0) {
                if (carry) bits[nb++] = (uint8_t)(hard ^ carried);
            } else {
                bits[nb++] = (uint8_t)(hard ^ prev_hard);
            }
            prev_hard = hard;
            ns++;
        }
        pr.m = m;
        pr.phase = phase;
        pr.freq = freq;
        pr.prev_re = prev_re;
        pr.prev_im = prev_im;
        pr.prev_dec = prev_dec;
        pr.have_prev = have_prev ? 1 : 0;
        pr.ns = ns;
        pr.nb = nb;
        pr.prev_hard = prev_hard;
        if (done) pr.loop_done = 1;
        next_stage = restage;
    }
    if (lane != 0) return;
    ArgusRdsState& st = *call.st;
    ArgusRdsOut& out = *call.out;
    if (!call.final) {  // more of the window comes with later launches
        *call.pr = pr;
        return;
    }
    out.found = 0;
    out.written = 0;
    out.nb = -1;
    if (n >= kGroupBits && rate > 0) {  // rds.go:136-146: the call's end
        double res = pr.m - (double)n;
        while (res < 0) res += sps;
        while (res >= sps) res -= sps;
        st.mu = res;
        st.prev_re = pr.prev_re;
        st.prev_im = pr.prev_im;
        st.prev_dec = pr.prev_dec;
        st.have_prev = pr.have_prev;
        if (pr.ns >= kGroupBits) {  // else Costas and bit carry untouched
            st.costas_phase = pr.phase;
            st.costas_freq = pr.freq;
            st.last_hard = pr.prev_hard;
            st.have_last_hard = 1;
            out.nb = pr.nb;
        }
    }
    pr.active = 0;
    *call.pr = pr;
    // the decoder's text as it stands (the scan kernel updates it if it runs)
    out.pi = st.pi;
    for (int i = 0; i < 8; i++) out.ps[i] = st.ps[i];
    for (int i = 0; i < 64; i++) out.rt[i] = st.rt[i];
}

__global__ __launch_bounds__(kThreads) void argus_rds_match_kernel(const ArgusRdsCall* __restrict__ calls,
                                                                   FlipTable ft)
{
    const ArgusRdsCall& call = calls[blockIdx.y];
    const int nb = call.out->nb;
    const int p = blockIdx.x * kThreads + threadIdx.x;
    if (nb < kGroupBits || p + kGroupBits > nb) return;
    uint16_t g[4];
    const uint8_t* b = call.bits + p;
    const uint8_t f = (uint8_t)(rds_group(b, 0u, ft, g) ? 1 : 0) | (uint8_t)(rds_group(b, 0x3FFFFFFu, ft, g) ? 2 : 0);
    call.flags[p] = f;
}

// next[2p + pol] = the first q >= p with a group in stream pol, or nb.
__global__ __launch_bounds__(kThreads) void argus_rds_next_kernel(const ArgusRdsCall* __restrict__ calls)
{
    __shared__ int first[2][kThreads];
    const ArgusRdsCall& call = calls[blockIdx.x];
    const int nb = call.out->nb;
    if (nb < kGroupBits) return;
    const int npos = nb - kGroupBits + 1;
    const int seg = (npos + kThreads - 1) / kThreads;
    const int t = threadIdx.x;
    const int lo = min(npos, t * seg), hi = min(npos, lo + seg);
    int nd = nb, ni = nb;
    for (int p = hi - 1; p >= lo; p--) {
        const uint8_t f = call.flags[p];
        if (f & 1) nd = p;
        if (f & 2) ni = p;
    }
    first[0][t] = nd;
    first[1][t] = ni;
    __syncthreads();
    for (int d = 1; d < kThreads; d <<= 1) {  // suffix minimum over the segments
        int a = nb, b = nb;
        if (t + d < kThreads) {
            a = first[0][t + d];
            b = first[1][t + d];
        }
        __syncthreads();
        first[0][t] = min(first[0][t], a);
        first[1][t] = min(first[1][t], b);
        __syncthreads();
    }
    nd = t + 1 < kThreads ? first[0][t + 1] : nb;
    ni = t + 1 < kThreads ? first[1][t + 1] : nb;
    for (int p = hi - 1; p >= lo; p--) {
        const uint8_t f = call.flags[p];
        if (f & 1) nd = p;
        if (f & 2) ni = p;
        call.next[2 * p] = nd;
        call.next[2 * p + 1] = ni;
    }
}

struct ScanCtx {
    const uint8_t* bits;
    const uint8_t* flags;
    const int32_t* next;
    int nb;
    uint32_t inv;
    int pol;
};

__device__ __forceinline__ bool scan_hit(const ScanCtx& s, int p)
{
    return p >= 0 && p + kGroupBits <= s.nb && ((s.flags[p] >> s.pol) & 1);
}

// rds.go:475-502 applyGroup, plus the call's group list.
__device__ void scan_apply(const ScanCtx& s, int p, const FlipTable& ft, ArgusRdsState& st, const ArgusRdsCall& call,
                           ArgusRdsOut& out)
{
    uint16_t g[4] = {0, 0, 0, 0};
    rds_group(s.bits + p, s.inv, ft, g);
    if (g[0] != 0) st.pi = g[0];
    const int type = (g[1] >> 12) & 0xF;
    const bool version_a = ((g[1] >> 11) & 1) == 0;
    if (type == 0 && version_a) {
        const int at = (g[1] & 3) * 2;
        st.ps[at] = rds_printable(g[3] >> 8);
        st.ps[at + 1] = rds_printable(g[3] & 0xFF);
    }
    if (type == 2 && version_a) {
        const int at = (g[1] & 0xF) * 4;
        const unsigned ch[4] = {(unsigned)(g[2] >> 8), (unsigned)(g[2] & 0xFF), (unsigned)(g[3] >> 8),
                                (unsigned)(g[3] & 0xFF)};
        for (int j = 0; j < 4 && at + j < 64; j++) st.rt[at + j] = rds_printable(ch[j]);
    }
    if (out.written < call.group_cap) {
        uint16_t* dst = call.groups + 4 * out.written;
        dst[0] = g[0];
        dst[1] = g[1];
        dst[2] = g[2];
        dst[3] = g[3];
        out.written++;
    }
}

// rds.go:365-430 tryDecode over the precomputed group flags.
__device__ int scan_stream(const ScanCtx& s, const FlipTable& ft, ArgusRdsState& st, const ArgusRdsCall& call,
                           ArgusRdsOut& out)
{
    int groups = 0;
    for (int i = 0; i + kGroupBits <= s.nb;) {
        if (!scan_hit(s, i)) {  // the reference steps i++ until a group starts
            i = s.next[2 * i + s.pol];
            continue;
        }
        groups++;
        scan_apply(s, i, ft, st, call, out);
        int next = i + kGroupBits;
        bool locked = true;
        while (locked && next + kGroupBits <= s.nb) {
            if (scan_hit(s, next)) {
                scan_apply(s, next, ft, st, call, out);
                groups++;
                next += kGroupBits;
                continue;
            }
            bool hit = false;
            for (int dl = 1; dl <= kFlywheel && !hit; dl++) {
                if (scan_hit(s, next - dl)) {
                    scan_apply(s, next - dl, ft, st, call, out);
                    groups++;
                    next = next - dl + kGroupBits;
                    hit = true;
                } else if (scan_hit(s, next + dl)) {
                    scan_apply(s, next + dl, ft, st, call, out);
                    groups++;
                    next = next + dl + kGroupBits;
                    hit = true;
                }
            }
            locked = hit;
        }
        i = next - kFlywheel > i ? next - kFlywheel : i + 1;
    }
    return groups;
}

__global__ __launch_bounds__(64) void argus_rds_scan_kernel(const ArgusRdsCall* __restrict__ calls, int n_calls,
                                                            FlipTable ft)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_calls) return;
    const ArgusRdsCall call = calls[c];
    ArgusRdsOut& out = *call.out;
    if (out.nb < kGroupBits) return;
    ArgusRdsState& st = *call.st;
    ScanCtx s{call.bits, call.flags, call.next, out.nb, 0u, 0};
    int found = scan_stream(s, ft, st, call, out);
    if (found == 0) {  // rds.go:242-250: the inverted stream
        s.inv = 0x3FFFFFFu;
        s.pol = 1;
        found = scan_stream(s, ft, st, call, out);
    }
    out.found = found;
    out.pi = st.pi;
    for (int i = 0; i < 8; i++) out.ps[i] = st.ps[i];
    for (int i = 0; i < 64; i++) out.rt[i] = st.rt[i];
}

// One workgroup per station: a block's RDS baseband into its decode windows.
__global__ __launch_bounds__(kThreads) void argus_rds_append_kernel(const ArgusRdsAppend* __restrict__ ap)
{
    const ArgusRdsAppend a = ap[blockIdx.x];
    for (int i = threadIdx.x; i < a.n0 + a.n1; i += kThreads) {
        const float2 v = a.src[i];
        if (i < a.n0)
            a.dst0[i] = v;
        else
            a.dst1[i - a.n0] = v;
    }
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

const FlipTable kFlips = make_flip_table();

}  // namespace

void argus_rds_caps(int n, int rate, int* bit_cap, int* group_cap)
{
    int bc = 0, gc = 0;
    if (n >= kGroupBits && rate > 0) {
        const double cap = 2.0 * (double)n / ((double)rate / 1187.5) + 16.0;
        bc = cap > (double)0x3fffffff ? 0x3fffffff : (int)cap;
        gc = bc / (kGroupBits - kFlywheel) + 2;  // groups start >= 101 bits apart
    }
    *bit_cap = bc;
    *group_cap = gc;
}

int argus_launch_rds(const ArgusRdsCall* d_calls, int n_calls, int n_final, int max_bit_cap, hipStream_t stream)
{
    if (n_calls <= 0) return 0;
    if (n_calls > 65535 || n_final < 0 || n_final > n_calls) return -1;
    argus_clear_error();
    const int waves = std::min(kSymWaves, (n_calls + kSymCUs - 1) / kSymCUs);
    argus_rds_symbols_kernel<<<(n_calls + waves - 1) / waves, 64 * waves, 0, stream>>>(d_calls, n_calls);
    ARGUS_CHECK(hipGetLastError());
    if (n_final == 0 || max_bit_cap < kGroupBits) return 0;
    dim3 grid((unsigned)((max_bit_cap + kThreads - 1) / kThreads), (unsigned)n_final);
    argus_rds_match_kernel<<<grid, kThreads, 0, stream>>>(d_calls, kFlips);
    ARGUS_CHECK(hipGetLastError());
    argus_rds_next_kernel<<<n_final, kThreads, 0, stream>>>(d_calls);
    ARGUS_CHECK(hipGetLastError());
    argus_rds_scan_kernel<<<(n_final + 63) / 64, 64, 0, stream>>>(d_calls, n_final, kFlips);
    return (int)hipGetLastError();
}

int argus_launch_rds_append(const ArgusRdsAppend* d_appends, int n, hipStream_t stream)
{
    if (n <= 0) return 0;
    argus_clear_error();
    argus_rds_append_kernel<<<n, kThreads, 0, stream>>>(d_appends);
    return (int)hipGetLastError();
}

void argus_rds_text(const char* src, int len, char* dst)  // rds.go:604-636
{
    int end = len;
    while (end > 0 && (src[end - 1] == ' ' || src[end - 1] == 0)) end--;
    for (int i = 0; i < end; i++) dst[i] = src[i] ? src[i] : ' ';
    dst[end] = 0;
}

struct gpud_rds_batch {
    int n_stations = 0;
    ArgusRdsState* d_states = nullptr;
    std::vector<ArgusRdsOut> last;           // each station's state after its last call
    std::vector<std::vector<uint16_t>> log;  // applied groups per station
};

extern "C" {

GPUD_API int gpud_rds_batch_create(gpud_rds_batch** out, int n_stations)
{
    if (!out || n_stations <= 0) return -1;
    auto* b = new gpud_rds_batch();
    b->n_stations = n_stations;
    const size_t bytes = (size_t)n_stations * sizeof(ArgusRdsState);
    hipError_t e = hipMalloc((void**)&b->d_states, bytes);
    if (e == hipSuccess) e = hipMemset(b->d_states, 0, bytes);
    if (e != hipSuccess) {
        if (b->d_states) (void)hipFree(b->d_states);
        delete b;
        return (int)e;
    }
    b->last.assign((size_t)n_stations, ArgusRdsOut{});
    std::memset(b->last.data(), 0, b->last.size() * sizeof(ArgusRdsOut));
    b->log.resize((size_t)n_stations);
    *out = b;
    return 0;
}

GPUD_API int gpud_rds_batch_destroy(gpud_rds_batch* b)
{
    if (!b) return 0;
    if (b->d_states) (void)hipFree(b->d_states);
    delete b;
    return 0;
}

GPUD_API int gpud_rds_batch_reset(gpud_rds_batch* b, int station)
{
    if (!b || station < 0 || station >= b->n_stations) return -1;
    ARGUS_CHECK(hipMemset(b->d_states + station, 0, sizeof(ArgusRdsState)));
    std::memset(&b->last[(size_t)station], 0, sizeof(ArgusRdsOut));
    b->log[(size_t)station].clear();
    return 0;
}

GPUD_API int gpud_rds_batch_decode_device(gpud_rds_batch* b, int n_calls, const int32_t* stations,
                                          const gpud_float2* samples_dev, const int64_t* offsets,
                                          const int32_t* rates, int32_t* found, gpud_stream_handle stream)
{
    if (!b || n_calls < 0 || n_calls > 65535) return -1;
    if (n_calls == 0) return 0;
    if (!stations || !offsets || !rates || !found) return -1;
    std::vector<char> seen((size_t)b->n_stations, 0);
    if (offsets[0] < 0) return -1;
    for (int c = 0; c < n_calls; c++) {
        const int s = stations[c];
        if (s < 0 || s >= b->n_stations || seen[(size_t)s]) return -1;  // one call per station
        seen[(size_t)s] = 1;
        const int64_t n = offsets[c + 1] - offsets[c];
        if (n < 0 || n > (int64_t)0x3fffffff) return -1;
        found[c] = 0;
    }
    if (offsets[n_calls] > offsets[0] && !samples_dev) return -1;
    hipStream_t st = (hipStream_t)stream;

    // Per-call workspace: [calls | progress | outs | bits | flags | next | groups];
    // calls and (zero) progress are uploaded.
    std::vector<ArgusRdsCall> calls((size_t)n_calls);
    size_t bit_total = 0, group_total = 0;
    int max_cap = 0;
    for (int c = 0; c < n_calls; c++) {
        ArgusRdsCall& k = calls[(size_t)c];
        k.n = (int)(offsets[c + 1] - offsets[c]);
        k.avail = k.n;
        k.final = 1;
        k.rate = rates[c];
        argus_rds_caps(k.n, k.rate, &k.bit_cap, &k.group_cap);
        bit_total += align256((size_t)k.bit_cap);
        group_total += (size_t)k.group_cap * 4;
        max_cap = std::max(max_cap, k.bit_cap);
    }
    const size_t off_prog = align256((size_t)n_calls * sizeof(ArgusRdsCall));
    const size_t off_outs = off_prog + align256((size_t)n_calls * sizeof(ArgusRdsProgress));
    const size_t off_bits = off_outs + align256((size_t)n_calls * sizeof(ArgusRdsOut));
    const size_t off_flags = off_bits + bit_total;
    const size_t off_next = off_flags + bit_total;
    const size_t off_groups = off_next + 8 * bit_total;
    const size_t bytes = off_groups + group_total * sizeof(uint16_t);
    char* d = nullptr;
    ARGUS_CHECK(hipMallocAsync((void**)&d, bytes, st));
    ArgusRdsOut* d_outs = (ArgusRdsOut*)(d + off_outs);
    size_t bo = 0, go = 0;
    for (int c = 0; c < n_calls; c++) {
        ArgusRdsCall& k = calls[(size_t)c];
        k.x = (const float2*)(samples_dev + offsets[c]);
        k.bits = (uint8_t*)(d + off_bits + bo);
        k.flags = (uint8_t*)(d + off_flags + bo);
        k.next = (int32_t*)(d + off_next + 8 * bo);
        k.groups = (uint16_t*)(d + off_groups) + go;
        k.st = b->d_states + stations[c];
        k.pr = (ArgusRdsProgress*)(d + off_prog) + c;
        k.out = d_outs + c;
        bo += align256((size_t)k.bit_cap);
        go += (size_t)k.group_cap * 4;
    }
    std::vector<char> head(off_outs, 0);
    std::memcpy(head.data(), calls.data(), (size_t)n_calls * sizeof(ArgusRdsCall));
    std::vector<ArgusRdsOut> outs((size_t)n_calls);
    std::vector<uint16_t> groups(group_total);
    int err = (int)hipMemcpyAsync(d, head.data(), off_outs, hipMemcpyHostToDevice, st);
    if (!err) err = argus_launch_rds((const ArgusRdsCall*)d, n_calls, n_calls, max_cap, st);
    if (!err)
        err = (int)hipMemcpyAsync(outs.data(), d_outs, (size_t)n_calls * sizeof(ArgusRdsOut), hipMemcpyDeviceToHost,
                                  st);
    if (!err && group_total)
        err = (int)hipMemcpyAsync(groups.data(), d + off_groups, group_total * sizeof(uint16_t),
                                  hipMemcpyDeviceToHost, st);
    hipFreeAsync(d, st);
    const hipError_t se = hipStreamSynchronize(st);
    if (!err) err = (int)se;
    if (err) return err;
    go = 0;
    for (int c = 0; c < n_calls; c++) {
        const ArgusRdsOut& o = outs[(size_t)c];
        found[c] = o.found;
        b->last[(size_t)stations[c]] = o;
        auto& lg = b->log[(size_t)stations[c]];
        lg.insert(lg.end(), groups.begin() + (ptrdiff_t)go, groups.begin() + (ptrdiff_t)(go + 4 * (size_t)o.written));
        go += (size_t)calls[(size_t)c].group_cap * 4;
    }
    return 0;
}

GPUD_API int gpud_rds_batch_state(gpud_rds_batch* b, int station, uint16_t* pi, char* ps9, char* rt65)
{
    if (!b || station < 0 || station >= b->n_stations) return -1;
    const ArgusRdsOut& s = b->last[(size_t)station];
    if (pi) *pi = (uint16_t)s.pi;
    if (ps9) argus_rds_text(s.ps, 8, ps9);
    if (rt65) argus_rds_text(s.rt, 64, rt65);
    return 0;
}

GPUD_API int gpud_rds_batch_groups(gpud_rds_batch* b, int station, uint16_t* groups4, int cap)
{
    if (!b || station < 0 || station >= b->n_stations || cap < 0 || (cap > 0 && !groups4)) return -1;
    const auto& lg = b->log[(size_t)station];
    const int total = (int)(lg.size() / 4);
    const int m = total < cap ? total : cap;
    if (m > 0) std::memcpy(groups4, lg.data(), (size_t)m * 4 * sizeof(uint16_t));
    return total;
}

}  // extern "C"
