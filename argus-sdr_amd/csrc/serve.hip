// serve.hip -- streaming service: channels are opened and closed while
// blocks flow, admitted against a fixed HBM budget, and every block runs one
// fused extraction launch plus one demodulation launch over all live
// channels.  Filter history, NCO phase, decimator phase, FM carry and BFO
// phase stay resident between blocks.
//
// Pipelining: each block occupies one of ARGUS_SERVE_SLOTS slots.  Its upload
// runs on the copy stream, the kernels on the compute stream, the download on
// the copy stream again, ordered by events, so block k+1's upload and block
// k-1's download overlap block k's kernels.  A channel opened while blocks are
// in flight joins at the next submitted block; a closed channel leaves the
// next block and its memory is released once the last block that used it has
// completed.
#include "extract.hpp"
#include "demod.hpp"
#include "serve.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

int argus_lowpass_fir(double cutoff_hz, int sample_rate, int taps, double* out, bool hamming);

namespace {

template <typename T>
hipError_t dev_alloc(T** p, size_t n)
{
    return hipMalloc((void**)p, (n ? n : 1) * sizeof(T));
}

template <typename T>
hipError_t host_alloc(T** p, size_t n)
{
    return hipHostMalloc((void**)p, (n ? n : 1) * sizeof(T), hipHostMallocDefault);
}

template <typename T>
int grow_dev(T** p, size_t* cap, size_t need)
{
    if (*p && *cap >= need) return 0;
    if (*p) hipFree(*p);
    *p = nullptr;
    *cap = 0;
    const size_t n = need < 64 ? 64 : need + need / 4;
    hipError_t e = dev_alloc(p, n);
    if (e != hipSuccess) return (int)e;
    *cap = n;
    return 0;
}

template <typename T>
int grow_host(T** p, size_t* cap, size_t need)
{
    if (*p && *cap >= need) return 0;
    if (*p) hipHostFree(*p);
    *p = nullptr;
    *cap = 0;
    const size_t n = need < 64 ? 64 : need + need / 4;
    hipError_t e = host_alloc(p, n);
    if (e != hipSuccess) return (int)e;
    *cap = n;
    return 0;
}

bool valid_mode(int m) { return m >= GPUD_DEMOD_NONE && m <= GPUD_DEMOD_WFM_STEREO; }

// Device block of a stereo channel: ArgusWfmState, then three ping-pong pairs
// of delay lines (mono, L-R, fallback), then the tap sets.
constexpr size_t kWfmState = 64;
constexpr size_t kWfmHistPair = 2 * ARGUS_WFM_HIST * sizeof(float);
constexpr size_t kWfmMonoHist = kWfmState;
constexpr size_t kWfmLrHist = kWfmMonoHist + kWfmHistPair;
constexpr size_t kWfmFbHist = kWfmLrHist + kWfmHistPair;
constexpr size_t kWfmTaps = kWfmFbHist + kWfmHistPair;
constexpr size_t kWfmBytes = kWfmTaps + ARGUS_WFM_NTAPSETS * ARGUS_WFM_TAPS * sizeof(float);
// audio-stage block of a channel (see ServeChannel::d_aud)
constexpr size_t kAudHist = 0;  // [2 halves][2 sides][ARGUS_AUDIO_TMAX] floats
constexpr size_t kAudDeemph = kAudHist + 4 * ARGUS_AUDIO_TMAX * sizeof(float);  // 2 doubles
constexpr size_t kAudBytes = kAudDeemph + 2 * sizeof(double);
static_assert(kAudDeemph % 8 == 0, "de-emphasis states are doubles");
static_assert(sizeof(ArgusWfmState) <= kWfmState, "state block");

// RDS mixer phase as a fixed-point turn: num / den of a turn, num < den.
uint64_t frac_to_fix(uint64_t num, uint64_t den)
{
    return (uint64_t)(((unsigned __int128)num << 64) / den);
}

}  // namespace

// Device block of a channel's pre-demod stage: FP64 taps, then two delay
// lines of T-1 inputs.
constexpr size_t kPreTaps = 0;
constexpr size_t kPreHist = ARGUS_PRE_TAPS * sizeof(double);
constexpr size_t kPreBytes = kPreHist + 2 * (ARGUS_PRE_TAPS - 1) * sizeof(float2);

// Device bytes a channel holds while open (taps, modulated taps, ping-pong
// history, FM carry) plus its share of the per-slot output buffers.
static size_t channel_footprint(const gpud_serve* s, int D, int T, bool stereo, int pre_factor)
{
    const size_t L = (size_t)(T + D - 1) / D;
    const size_t taps = (size_t)D * L * sizeof(float) + (size_t)T * sizeof(float2);
    const size_t hist = 2 * (size_t)std::max(1, T - 1) * sizeof(float2);
    const size_t carry = 2 * sizeof(float2);
    const size_t n_out = (size_t)s->block_samples / D + 2;
    size_t per_block = n_out * (sizeof(float2) + sizeof(float));
    size_t fixed = taps + hist + carry;
    if (argus_fast_eligible(D, T)) fixed += ARGUS_FAST_N * sizeof(float2);
    if (pre_factor > 1) {
        fixed += kPreBytes;
        per_block += (n_out / pre_factor + 1) * sizeof(float2);
    }
    if (stereo) {
        // mono + 4 scratch floats + a second audio float + RDS samples (<= 1 per 2 outputs)
        per_block += n_out * (6 * sizeof(float)) + (n_out / 2 + 1) * sizeof(float2) + sizeof(ArgusWfmBlock);
        fixed += kWfmBytes;
    }
    return fixed + per_block * ARGUS_SERVE_SLOTS;
}

static size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

char* ServeArena::alloc(size_t bytes)
{
    bytes = align256(bytes ? bytes : 1);
    for (int pass = 0; pass < 2; pass++) {
        for (auto it = free_list.begin(); it != free_list.end(); ++it) {
            if (it->second < bytes) continue;
            char* p = it->first;
            const size_t rest = it->second - bytes;
            free_list.erase(it);
            if (rest) free_list.emplace(p + bytes, rest);
            return p;
        }
        if (pass || total >= cap) break;
        // a new region: as large as everything so far (256 MiB first), at
        // least the request, within the cap
        size_t want = std::max<size_t>(total ? total : ((size_t)256 << 20), bytes);
        want = std::min(want, cap - total);
        if (want < bytes) break;
        char* base = nullptr;
        if (hipMalloc((void**)&base, want) != hipSuccess) break;
        regions.emplace(base, want);
        total += want;
        free_list.emplace(base, want);
    }
    return nullptr;
}

void ServeArena::release(char* p, size_t bytes)
{
    if (!p) return;
    bytes = align256(bytes ? bytes : 1);
    auto reg = std::prev(regions.upper_bound(p));  // the region holding p
    char* lo = reg->first;
    char* hi = reg->first + reg->second;
    auto next = free_list.lower_bound(p);
    if (next != free_list.end() && p + bytes == next->first && next->first < hi) {  // merge with the block after
        bytes += next->second;
        next = free_list.erase(next);
    }
    if (next != free_list.begin()) {  // and with the block before
        auto prev = std::prev(next);
        if (prev->first >= lo && prev->first + prev->second == p) {
            prev->second += bytes;
            return;
        }
    }
    free_list.emplace(p, bytes);
}

void ServeArena::destroy()
{
    for (auto& kv : regions) hipFree(kv.first);
    regions.clear();
    free_list.clear();
    total = 0;
}

// Copy `bytes` of host data to device memory `dst` through the pinned ring,
// queued on the compute stream (ahead of every later block).  Waits only when
// the ring wraps onto a region whose copy is still pending.
static int stage_upload(gpud_serve* s, void* dst, const void* src, size_t bytes)
{
    ServeStage& st = s->stage;
    if (!bytes) return 0;
    if (bytes > st.size) return -1;
    size_t at = align256(st.head);
    if (at + bytes > st.size) at = 0;
    // retire completed regions, then wait out any still overlapping [at, at + bytes)
    while (!st.busy.empty() && hipEventQuery(st.busy.front().second) == hipSuccess) {
        st.spare.push_back(st.busy.front().second);
        st.busy.pop_front();
    }
    for (size_t i = 0; i < st.busy.size(); i++) {
        const size_t o = st.busy[i].first.first, n = st.busy[i].first.second;
        if (o < at + bytes && at < o + n) {
            ARGUS_CHECK(hipEventSynchronize(st.busy[i].second));
            for (size_t j = 0; j <= i; j++) {  // ring order: everything before it is done too
                st.spare.push_back(st.busy.front().second);
                st.busy.pop_front();
            }
            i = (size_t)-1;
        }
    }
    std::memcpy(st.ring + at, src, bytes);
    ARGUS_CHECK(hipMemcpyAsync(dst, st.ring + at, bytes, hipMemcpyHostToDevice, s->compute));
    hipEvent_t ev = nullptr;
    if (!st.spare.empty()) {
        ev = st.spare.back();
        st.spare.pop_back();
    } else {
        ARGUS_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    }
    ARGUS_CHECK(hipEventRecord(ev, s->compute));
    st.busy.push_back({{at, bytes}, ev});
    st.head = at + bytes;
    return 0;
}

static void release_channel(gpud_serve* s, ServeChannel& c)
{
    s->arena.release(c.d_block, c.block_bytes);
    s->arena.release(c.d_aud, c.aud_bytes);
    s->arena.release(c.d_rdec, c.rdec_bytes);
    c.d_block = nullptr;
    c.block_bytes = c.aud_bytes = c.rdec_bytes = 0;
    c.d_taps = nullptr;
    c.d_ctaps = nullptr;
    c.d_hist[0] = c.d_hist[1] = nullptr;
    c.d_carry = nullptr;
    c.d_wfm = nullptr;
    c.d_fastG = nullptr;
    c.d_pre = nullptr;
    c.d_aud = nullptr;
    c.d_rdec = nullptr;
    c.d_fastG = nullptr;
    c.d_taps = nullptr;
    c.d_ctaps = nullptr;
    c.d_hist[0] = c.d_hist[1] = nullptr;
    c.d_carry = nullptr;
    c.d_wfm = nullptr;
}

// Release channels whose last block has completed.
static void reap(gpud_serve* s)
{
    for (size_t i = 0; i < s->closed.size();) {
        ServeChannel& c = s->closed[i];
        if (c.last_ticket <= s->completed_ticket) {
            s->bytes_in_use -= c.footprint;
            release_channel(s, c);
            s->closed[i] = s->closed.back();
            s->closed.pop_back();
        } else {
            i++;
        }
    }
}

// Initial device block of a new stereo channel (kWfmBytes at blob): PLL at
// 19 kHz (streamer.go:1532-1549), zero delay lines, the tap sets at the
// channel's demod rate.
static void wfm_channel_blob(ServeChannel& c, int rate, char* blob)
{
    ArgusWfmState st{};
    st.freq = 2.0 * M_PI * 19000.0 / (double)rate;
    std::memcpy(blob, &st, sizeof st);
    argus_wfm_design_taps(rate, (float*)(blob + kWfmTaps));
    c.rds_decim = std::max(1, rate / 19000);
    c.rds_count = 0;
    c.wfm_sel = 0;
}

// Free a slot's output buffers, tables and captured graph (not its input
// staging buffer or events).
static void slot_release(ServeSlot& sl)
{
    hipFree(sl.d_iq);
    hipFree(sl.d_audio);
    hipHostFree(sl.h_iq);
    hipHostFree(sl.h_audio);
    hipFree(sl.d_mono);
    hipFree(sl.d_wscratch);
    hipFree(sl.d_rds);
    hipFree(sl.d_wblk);
    hipHostFree(sl.h_rds);
    hipHostFree(sl.h_wblk);
    hipHostFree(sl.h_tab);
    hipFree(sl.d_tab);
    hipHostFree(sl.h_items);
    hipFree(sl.d_items);
    sl.h_items = sl.d_items = nullptr;
    sl.h_items_cap = sl.d_items_cap = 0;
    hipFree(sl.d_aout);
    hipHostFree(sl.h_aout);
    sl.d_aout = sl.h_aout = nullptr;
    hipFree(sl.d_pre);
    hipHostFree(sl.h_pre);
    sl.d_pre = sl.h_pre = nullptr;
    sl.pre_cap = sl.h_pre_cap = 0;
    sl.aout_cap = sl.h_aout_cap = 0;
    hipHostFree(sl.h_rtab);
    hipFree(sl.d_rtab);
    sl.h_rtab = sl.d_rtab = nullptr;
    sl.rtab_cap = sl.h_rtab_cap = 0;
    sl.n_rdec = 0;
    for (auto& ge : sl.gexec)
        if (ge) hipGraphExecDestroy(ge);
    sl.d_iq = nullptr;
    sl.d_audio = nullptr;
    sl.h_iq = nullptr;
    sl.h_audio = nullptr;
    sl.d_mono = nullptr;
    sl.d_wscratch = nullptr;
    sl.d_rds = nullptr;
    sl.d_wblk = nullptr;
    sl.h_rds = nullptr;
    sl.h_wblk = nullptr;
    sl.h_tab = nullptr;
    sl.d_tab = nullptr;
    for (auto& ge : sl.gexec) ge = nullptr;
    sl.iq_cap = sl.audio_cap = sl.h_iq_cap = sl.h_audio_cap = 0;
    sl.mono_cap = sl.wscratch_cap = sl.rds_cap = sl.h_rds_cap = sl.wblk_cap = sl.h_wblk_cap = 0;
    sl.tab_cap = 0;
}

// The RDS decodes of a completed block become their channels' latest state.
static void rds_collect(gpud_serve* s, const ServeSlot& sl)
{
    const ArgusRdsOut* outs = (const ArgusRdsOut*)(sl.h_rtab + sl.off_rout);
    for (int i = 0; i < sl.n_rdec; i++) {
        auto it = s->live.find(sl.rdec_channel[(size_t)i]);
        if (it == s->live.end() || it->second.r_last_ticket >= sl.ticket) continue;
        it->second.r_last = outs[i];
        it->second.r_last_ticket = sl.ticket;
    }
}

static int slot_wait(gpud_serve* s, ServeSlot& sl)
{
    if (!sl.busy) return 0;
    ARGUS_CHECK(hipEventSynchronize(sl.ev_done));
    sl.busy = false;
    rds_collect(s, sl);
    if (sl.ticket > s->completed_ticket) s->completed_ticket = sl.ticket;
    reap(s);
    return 0;
}

extern "C" {

GPUD_API int gpud_serve_create(int sample_rate, int block_samples, uint64_t hbm_budget_bytes, gpud_serve** out)
{
    if (!out || sample_rate <= 0 || block_samples <= 0 || hbm_budget_bytes == 0) return -1;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return (int)hipErrorNoDevice;
    gpud_serve* s = new gpud_serve();
    s->sample_rate = sample_rate;
    s->block_samples = block_samples;
    s->budget = hbm_budget_bytes;
    ARGUS_CHECK(hipStreamCreate(&s->compute));
    ARGUS_CHECK(hipStreamCreate(&s->copy));
    for (int i = 0; i < ARGUS_SERVE_SLOTS; i++) {
        ServeSlot& sl = s->slots[i];
        ARGUS_CHECK(hipEventCreateWithFlags(&sl.ev_in, hipEventDisableTiming));
        ARGUS_CHECK(hipEventCreate(&sl.ev_kern));  // timed: kernel_ms in gpud_serve_wait
        ARGUS_CHECK(hipEventCreate(&sl.ev_done));
        ARGUS_CHECK(hipEventCreate(&sl.ev_start));
        ARGUS_CHECK(hipEventCreate(&sl.ev_kstart));
        ARGUS_CHECK(hipEventCreate(&sl.ev_xend));  // timed: extract_ms (recorded inside the block's graph)
        for (auto& ev : sl.ev_part) ARGUS_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        ARGUS_CHECK(dev_alloc(&sl.d_in, (size_t)block_samples));
        ARGUS_CHECK(host_alloc(&sl.h_in, (size_t)block_samples));
    }
    // channel state arena (bounded by the budget) and the upload ring
    s->arena.cap = align256((size_t)std::min<uint64_t>(hbm_budget_bytes, (uint64_t)16 << 30));
    s->stage.size = (size_t)16 << 20;
    ARGUS_CHECK(hipHostMalloc((void**)&s->stage.ring, s->stage.size, hipHostMallocDefault));
    const int smax = argus_fast_segments(block_samples);
    ARGUS_CHECK(dev_alloc(&s->d_ftail, (size_t)ARGUS_FAST_TAIL));
    ARGUS_CHECK(hipMemset(s->d_ftail, 0, ARGUS_FAST_TAIL * sizeof(float2)));
    ARGUS_CHECK(dev_alloc(&s->d_fu, argus_fast_u_len(smax)));
    ARGUS_CHECK(dev_alloc(&s->d_fX, (size_t)smax * ARGUS_FAST_N));
    *out = s;
    return 0;
}

GPUD_API int gpud_serve_destroy(gpud_serve* s)
{
    if (!s) return 0;
    hipDeviceSynchronize();
    for (auto& kv : s->live) release_channel(s, kv.second);
    for (auto& c : s->closed) release_channel(s, c);
    s->arena.destroy();
    hipHostFree(s->stage.ring);
    for (auto& b : s->stage.busy) hipEventDestroy(b.second);
    for (hipEvent_t ev : s->stage.spare) hipEventDestroy(ev);
    for (int i = 0; i < ARGUS_SERVE_SLOTS; i++) {
        ServeSlot& sl = s->slots[i];
        hipEventDestroy(sl.ev_in);
        hipEventDestroy(sl.ev_kern);
        hipEventDestroy(sl.ev_done);
        hipEventDestroy(sl.ev_start);
        hipEventDestroy(sl.ev_kstart);
        hipEventDestroy(sl.ev_xend);
        for (hipEvent_t ev : sl.ev_part) hipEventDestroy(ev);
        hipFree(sl.d_in);
        hipHostFree(sl.h_in);
        slot_release(sl);
    }
    for (auto& kv : s->filters) hipFree(kv.second);
    hipFree(s->d_ftail);
    hipFree(s->d_fu);
    hipFree(s->d_fX);
    hipStreamDestroy(s->compute);
    hipStreamDestroy(s->copy);
    delete s;
    return 0;
}

// Admit a channel.  Returns 0, -1 on bad arguments or a non-integer
// decimation (cpu_oracle.go:155-163), GPUD_ERR_DUPLICATE if the id is live,
// GPUD_ERR_BUDGET if its device footprint does not fit the remaining budget.
GPUD_API int gpud_serve_open_ex(gpud_serve* s, const gpud_channel_spec* spec, const gpud_channel_opts* opts)
{
    if (!s || !spec || !valid_mode(spec->demod) || (spec->window != 0 && spec->window != 1)) return -1;
    if (spec->out_rate <= 0 || s->sample_rate % spec->out_rate != 0) return -1;
    if (opts && opts->demod_rate < 0) return -1;
    if (s->live.count(spec->channel_id)) return GPUD_ERR_DUPLICATE;
    const int D = s->sample_rate / spec->out_rate;
    const int T = spec->num_taps > 0 ? spec->num_taps : 101;
    if (T > ARGUS_MAX_TAPS) return -1;
    const bool stereo = spec->demod == GPUD_DEMOD_WFM_STEREO;
    // the pre-demod decimation (streamer.go:1098-1112): round(snipRate /
    // demodRate), at least 1; WFM forced to 2 where that keeps >= 200 kHz
    int pre = 1;
    if (opts && opts->demod_rate > 0) {
        pre = (int)std::lround((double)spec->out_rate / (double)opts->demod_rate);
        if (pre < 1) pre = 1;
        const bool wfm = spec->demod == GPUD_DEMOD_WFM || stereo;
        if (wfm && pre > 2 && spec->out_rate / 2 >= 200000) pre = 2;
    }
    const size_t fp = channel_footprint(s, D, T, stereo, pre);
    reap(s);
    if (s->bytes_in_use + fp > s->budget) return GPUD_ERR_BUDGET;

    ServeChannel c;
    c.spec = *spec;
    c.r_last_ticket = s->last_ticket;  // decodes of blocks before it are not its own
    c.decim = D;
    c.T = T;
    c.footprint = fp;
    c.stereo = stereo;
    c.pre_factor = pre;
    c.demod_rate = spec->out_rate / pre;
    double cutoff = spec->bandwidth / 2;
    if (cutoff < 200) cutoff = 200;
    std::vector<double> b64((size_t)T + 1);
    const int nb = argus_lowpass_fir(cutoff, s->sample_rate, T, b64.data(), spec->window == 1);
    const int L = (T + D - 1) / D;
    std::vector<float> poly((size_t)D * L, 0.0f);
    for (int p = 0; p < D; p++)
        for (int k = 0; k < L; k++) {
            const int src = p + k * D;
            if (src < std::min(nb, T)) poly[(size_t)p * L + k] = (float)b64[(size_t)src];
        }
    // the increment the block builder derives from the same offset
    const uint64_t inc = argus_rad_to_fix(-2.0 * M_PI * spec->offset_hz / (double)s->sample_rate);
    // one arena block: taps | modulated taps | history x2 | FM carry x2 |
    // fast-path spectrum | stereo block | pre-demod block
    const bool fast = argus_fast_eligible(D, T);
    const size_t H = (size_t)std::max(1, T - 1);
    const size_t o_ctaps = align256(poly.size() * sizeof(float));
    const size_t o_hist0 = o_ctaps + align256((size_t)T * sizeof(float2));
    const size_t o_hist1 = o_hist0 + align256(H * sizeof(float2));
    const size_t o_carry = o_hist1 + align256(H * sizeof(float2));
    const size_t o_fast = o_carry + align256(2 * sizeof(float2));
    const size_t o_wfm = o_fast + (fast ? align256((size_t)ARGUS_FAST_N * sizeof(float2)) : 0);
    const size_t o_pre = o_wfm + (stereo ? align256(kWfmBytes) : 0);
    const size_t bytes = o_pre + (pre > 1 ? align256(kPreBytes) : 0);
    std::vector<char> blob(bytes, 0);
    std::memcpy(blob.data(), poly.data(), poly.size() * sizeof(float));
    argus_modulated_taps(poly.data(), D, T, inc, (float2*)(blob.data() + o_ctaps));
    if (fast) {  // the fast path's transform of the modulated taps
        std::vector<float> h((size_t)T);
        for (int k = 0; k < T; k++) h[(size_t)k] = poly[(size_t)(k % D) * L + k / D];
        argus_fast_spectrum(h.data(), T, inc, (float2*)(blob.data() + o_fast));
    }
    if (stereo) wfm_channel_blob(c, c.demod_rate, blob.data() + o_wfm);
    if (pre > 1) {
        // LowpassFIR(cutoff, snipRate, 101) in FP64, cutoff 90 kHz for WFM,
        // else 0.8 of the demod Nyquist (streamer.go:1114-1124); zero delay lines
        const bool wfm = spec->demod == GPUD_DEMOD_WFM || stereo;
        const double pc = wfm ? 90000.0 : (double)c.demod_rate / 2.0 * 0.8;
        std::vector<double> h(ARGUS_PRE_TAPS + 1);
        argus_lowpass_fir(pc, spec->out_rate, ARGUS_PRE_TAPS, h.data(), false);
        std::memcpy(blob.data() + o_pre + kPreTaps, h.data(), ARGUS_PRE_TAPS * sizeof(double));
    }
    c.d_block = s->arena.alloc(bytes);
    if (!c.d_block) return GPUD_ERR_BUDGET;  // the arena cannot place it
    c.block_bytes = bytes;
    c.d_taps = (float*)c.d_block;
    c.d_ctaps = (float2*)(c.d_block + o_ctaps);
    c.d_hist[0] = (float2*)(c.d_block + o_hist0);
    c.d_hist[1] = (float2*)(c.d_block + o_hist1);
    c.d_carry = (float2*)(c.d_block + o_carry);
    if (fast) c.d_fastG = (float2*)(c.d_block + o_fast);
    if (stereo) c.d_wfm = c.d_block + o_wfm;
    if (pre > 1) c.d_pre = c.d_block + o_pre;
    const int e = stage_upload(s, c.d_block, blob.data(), bytes);
    if (e) {
        release_channel(s, c);
        return e;
    }
    s->bytes_in_use += fp;
    if (c.d_fastG) s->n_fast_capable++;
    s->live.emplace(spec->channel_id, c);
    s->epoch++;
    return 0;
}

GPUD_API int gpud_serve_open(gpud_serve* s, const gpud_channel_spec* spec)
{
    return gpud_serve_open_ex(s, spec, nullptr);
}

// Retire a channel: it is absent from the next submitted block; its memory is
// released when the blocks already in flight that use it have completed.
GPUD_API int gpud_serve_close(gpud_serve* s, int64_t channel_id)
{
    if (!s) return -1;
    auto it = s->live.find(channel_id);
    if (it == s->live.end()) return -1;
    if (it->second.rds_win > 0) s->n_rds_on--;
    if (it->second.d_fastG) s->n_fast_capable--;
    s->closed.push_back(it->second);
    s->live.erase(it);
    s->epoch++;
    reap(s);
    return 0;
}

GPUD_API int gpud_serve_info(gpud_serve* s, int* live_channels, uint64_t* bytes_in_use, uint64_t* budget)
{
    if (!s) return -1;
    if (live_channels) *live_channels = (int)s->live.size();
    if (bytes_in_use) *bytes_in_use = s->bytes_in_use;
    if (budget) *budget = s->budget;
    return 0;
}

}  // extern "C"

// Stereo-stage descriptor and work items of one channel for this block (the
// channel's carry advances in commit_block).
static void wfm_describe(const ServeChannel& c, const ServeOut& o, ServeSlot& sl, ArgusWfmDesc& wd,
                         std::vector<ArgusWfmItem>& items)
{
    const int rate = c.demod_rate;
    const int D = c.rds_decim;
    wd.mono = sl.d_mono + o.mono_off;
    wd.audio = sl.d_audio + o.audio_off;
    wd.scratch = sl.d_wscratch + 4 * o.mono_off;
    wd.rds = sl.d_rds + o.rds_off;
    float* mh = (float*)(c.d_wfm + kWfmMonoHist);
    float* lh = (float*)(c.d_wfm + kWfmLrHist);
    wd.mono_hist_rd = mh + c.wfm_sel * ARGUS_WFM_HIST;
    wd.mono_hist_wr = mh + (c.wfm_sel ^ 1) * ARGUS_WFM_HIST;
    wd.lr_hist_rd = lh + c.wfm_sel * ARGUS_WFM_HIST;
    wd.lr_hist_wr = lh + (c.wfm_sel ^ 1) * ARGUS_WFM_HIST;
    wd.fb_hist = (float*)(c.d_wfm + kWfmFbHist);
    wd.taps = (const float*)(c.d_wfm + kWfmTaps);
    wd.st = (ArgusWfmState*)c.d_wfm;
    wd.blk = sl.d_wblk + o.wfm_index;
    // stream sample c0 + i mixes with 2 pi ((c0 + i + 1) 57000 mod rate) / rate
    const uint64_t r = (uint64_t)rate;
    const uint64_t num0 = ((uint64_t)((c.rds_count + 1) % rate) * 57000u) % r;
    wd.mix0 = frac_to_fix(num0, r);
    wd.mix_inc = frac_to_fix(57000u % r, r);
    argus_wfm_loop_constants(rate, wd);
    wd.n = o.n_mono;
    wd.rds_first = (int)((D - c.rds_count % D) % D);
    wd.rds_decim = D;
    wd.n_rds = o.n_rds;
    wd.batch = ARGUS_WFM_STREAM;
    wd.pad = 0;
    for (int st = 0; st < o.n_mono; st += ARGUS_WFM_TILE) items.push_back({o.wfm_index, st, ARGUS_WFM_ITEM_TILE, 0});
    if (o.n_mono > 0) items.push_back({o.wfm_index, 0, ARGUS_WFM_ITEM_TAIL, 0});
}

// Audio-stage descriptor of one channel for this block (the resampler phase
// and the delay-line half advance in commit_block).
static void audio_describe(const ServeChannel& c, const ServeOut& o, ServeSlot& sl, ArgusAudioDesc& ad)
{
    const bool wfm = c.stereo || c.spec.demod == GPUD_DEMOD_WFM;
    float* hh = (float*)(c.d_aud + kAudHist);
    ad.in = sl.d_audio + o.audio_off;
    ad.out = sl.d_aout + o.aout_off;
    ad.arms = c.d_arms;
    ad.hist_rd = hh + c.a_sel * 2 * ARGUS_AUDIO_TMAX;
    ad.hist_wr = hh + (c.a_sel ^ 1) * 2 * ARGUS_AUDIO_TMAX;
    ad.deemph = (double*)(c.d_aud + kAudDeemph);
    ad.deemph_on = wfm && c.deemph_us > 0.0 ? 1 : 0;
    ad.alpha = ad.deemph_on ? std::exp(-1.0 / ((double)c.audio_rate * c.deemph_us * 1e-6)) : 0.0;
    ad.L = c.aL;
    ad.M = c.aM;
    ad.t0 = c.at0;
    ad.T = c.aT;
    ad.ident = c.a_ident ? 1 : 0;
    ad.n_in = c.stereo ? o.n_mono : o.n_audio;
    ad.n_out = o.n_aout;
    ad.channels = o.a_channels;
    ad.gain = wfm ? 0.35f : 1.0f;
}

static size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

// Grow a slot's table to at least `need` bytes (its contents are rewritten by the caller).
static int grow_table(ServeSlot& sl, size_t need)
{
    if (sl.h_tab && sl.tab_cap >= need) return 0;
    const size_t cap = std::max<size_t>(need + need / 4, 1 << 16);
    char* h = nullptr;
    char* d = nullptr;
    ARGUS_CHECK(hipHostMalloc((void**)&h, cap, hipHostMallocDefault));
    hipError_t e = hipMalloc((void**)&d, cap);
    if (e != hipSuccess) {
        hipHostFree(h);
        return (int)e;
    }
    if (sl.h_tab) hipHostFree(sl.h_tab);
    if (sl.d_tab) hipFree(sl.d_tab);
    sl.h_tab = h;
    sl.d_tab = d;
    sl.tab_cap = cap;
    return 0;
}

// Whether the channel takes the fast form this block (its history is the
// stream's own last T-1 samples).
static bool fast_now(const gpud_serve* s, const ServeChannel& c, int n)
{
    return s->fast && c.d_fastG && c.hist_len == c.T - 1 && n > 0 && n <= ARGUS_FAST_MAX_BLOCK;
}

// Whether its outputs are complete after the fast kernels and its own
// demodulation (no stage after the demodulator, none before it).
static bool early_now(const gpud_serve* s, const ServeChannel& c, int n)
{
    return fast_now(s, c, n) && c.pre_factor == 1 && !c.stereo && c.audio_rate == 0;
}

// Output layout of one block of n samples over the live channels (in the
// order of `live`, the early channels' IQ and demodulated outputs ahead of
// the others'): per-channel counts and offsets, and the totals.
static void layout_block(const gpud_serve* s, int n, BlockLayout& L)
{
    L.outs.clear();
    L.outs.reserve(s->live.size());
    L.iq_total = L.audio_total = L.mono_total = L.rds_total = L.aout_total = L.pre_total = 0;
    L.n_wfm = L.n_audio_chans = L.n_pre_chans = 0;
    // the early channels' IQ and demodulated outputs first (in id order),
    // then the others'
    int64_t iq_e = 0, au_e = 0;
    for (const auto& kv : s->live) {
        const ServeChannel& c = kv.second;
        if (!early_now(s, c, n)) continue;
        iq_e += (c.phase_count + n) / c.decim;
        au_e += argus_demod_count(c.spec.demod, (c.phase_count + n) / c.decim, c.have_carry);
    }
    L.iq_early = iq_e;
    L.audio_early = au_e;
    int64_t iq_at_e = 0, au_at_e = 0, iq_at = iq_e, au_at = au_e;
    for (const auto& kv : s->live) {
        const ServeChannel& c = kv.second;
        ServeOut o;
        o.channel_id = kv.first;
        o.early = early_now(s, c, n) ? 1 : 0;
        o.n_iq = (c.phase_count + n) / c.decim;
        o.iq_off = o.early ? iq_at_e : iq_at;
        // the demodulator's input: the pre-demod stage's outputs, or the baseband
        o.n_pre = o.n_iq;
        o.pre_off = -1;
        if (c.pre_factor > 1) {
            argus_pre_block(c.pre_phase, c.pre_factor, o.n_iq, &o.n_pre, nullptr);
            o.pre_off = L.pre_total;
            L.pre_total += o.n_pre;
            L.n_pre_chans++;
        }
        o.n_audio = argus_demod_count(c.spec.demod, o.n_pre, c.have_carry);
        o.audio_off = o.early ? au_at_e : au_at;
        o.rate = c.spec.out_rate;
        o.wfm_index = -1;
        o.mono_off = o.rds_off = 0;
        o.n_mono = o.n_rds = o.rds_rate = 0;
        o.a_index = -1;
        o.n_aout = o.a_channels = 0;
        o.a_tnext = c.at0;
        o.a_rate = c.audio_rate;
        o.aout_off = 0;
        o.r_on = c.rds_win > 0 ? 1 : 0;
        o.r_index = -1;
        if (c.stereo) {
            o.n_mono = argus_demod_count(ARGUS_DEMOD_WFM, o.n_pre, c.have_carry);
            o.n_audio = 2 * o.n_mono;
            o.wfm_index = L.n_wfm++;
            o.mono_off = L.mono_total;
            const int first = (int)((c.rds_decim - c.rds_count % c.rds_decim) % c.rds_decim);
            o.n_rds = o.n_mono > first ? (o.n_mono - 1 - first) / c.rds_decim + 1 : 0;
            o.rds_off = L.rds_total;
            o.rds_rate = c.demod_rate / c.rds_decim;
            L.mono_total += o.n_mono;
            L.rds_total += o.n_rds;
        }
        if (c.audio_rate > 0) {
            // the stage's input: the stereo stage's L/R frames, or the mono
            // demodulator output
            const int n_in = c.stereo ? o.n_mono : o.n_audio;
            o.a_index = L.n_audio_chans++;
            o.a_channels = c.stereo ? 2 : 1;
            argus_audio_block(c.aL, c.aM, c.at0, n_in, &o.n_aout, &o.a_tnext);
            o.aout_off = L.aout_total;
            L.aout_total += (int64_t)o.n_aout * o.a_channels;
        }
        (o.early ? iq_at_e : iq_at) += o.n_iq;
        (o.early ? au_at_e : au_at) += o.n_audio;
        L.outs.push_back(o);
    }
    L.iq_total = iq_at;
    L.audio_total = au_at;
}

// Whether the slot's output buffers already hold a block of this layout.
static bool slot_fits(const ServeSlot& sl, const BlockLayout& L)
{
    auto fits = [](size_t cap, int64_t need) { return cap >= (size_t)std::max<int64_t>(need, 1); };
    bool ok = fits(sl.iq_cap, L.iq_total) && fits(sl.audio_cap, L.audio_total) && fits(sl.h_iq_cap, L.iq_total) &&
              fits(sl.h_audio_cap, L.audio_total);
    if (L.n_wfm > 0)
        ok = ok && fits(sl.mono_cap, L.mono_total) && fits(sl.wscratch_cap, 4 * L.mono_total) &&
             fits(sl.rds_cap, L.rds_total) && fits(sl.h_rds_cap, L.rds_total) && fits(sl.wblk_cap, L.n_wfm) &&
             fits(sl.h_wblk_cap, L.n_wfm);
    if (L.n_audio_chans > 0)
        ok = ok && fits(sl.aout_cap, L.aout_total) && fits(sl.h_aout_cap, L.aout_total);
    if (L.n_pre_chans > 0) ok = ok && fits(sl.pre_cap, L.pre_total) && fits(sl.h_pre_cap, L.pre_total);
    return ok;
}

// Grow the slot's output buffers for a layout (invalidates the outputs of
// the block the slot held).
static int slot_grow(ServeSlot& sl, const BlockLayout& L)
{
    ARGUS_CHECK((hipError_t)grow_dev(&sl.d_iq, &sl.iq_cap, (size_t)std::max<int64_t>(L.iq_total, 1)));
    ARGUS_CHECK((hipError_t)grow_dev(&sl.d_audio, &sl.audio_cap, (size_t)std::max<int64_t>(L.audio_total, 1)));
    ARGUS_CHECK((hipError_t)grow_host(&sl.h_iq, &sl.h_iq_cap, (size_t)std::max<int64_t>(L.iq_total, 1)));
    ARGUS_CHECK((hipError_t)grow_host(&sl.h_audio, &sl.h_audio_cap, (size_t)std::max<int64_t>(L.audio_total, 1)));
    if (L.n_wfm > 0) {
        ARGUS_CHECK((hipError_t)grow_dev(&sl.d_mono, &sl.mono_cap, (size_t)std::max<int64_t>(L.mono_total, 1)));
        ARGUS_CHECK((hipError_t)grow_dev(&sl.d_wscratch, &sl.wscratch_cap, (size_t)std::max<int64_t>(4 * L.mono_total, 1)));
        ARGUS_CHECK((hipError_t)grow_dev(&sl.d_rds, &sl.rds_cap, (size_t)std::max<int64_t>(L.rds_total, 1)));
        ARGUS_CHECK((hipError_t)grow_host(&sl.h_rds, &sl.h_rds_cap, (size_t)std::max<int64_t>(L.rds_total, 1)));
        ARGUS_CHECK((hipError_t)grow_dev(&sl.d_wblk, &sl.wblk_cap, (size_t)L.n_wfm));
        ARGUS_CHECK((hipError_t)grow_host(&sl.h_wblk, &sl.h_wblk_cap, (size_t)L.n_wfm));
    }
    if (L.n_audio_chans > 0) {
        ARGUS_CHECK((hipError_t)grow_dev(&sl.d_aout, &sl.aout_cap, (size_t)std::max<int64_t>(L.aout_total, 1)));
        ARGUS_CHECK((hipError_t)grow_host(&sl.h_aout, &sl.h_aout_cap, (size_t)std::max<int64_t>(L.aout_total, 1)));
    }
    if (L.n_pre_chans > 0) {
        ARGUS_CHECK((hipError_t)grow_dev(&sl.d_pre, &sl.pre_cap, (size_t)std::max<int64_t>(L.pre_total, 1)));
        ARGUS_CHECK((hipError_t)grow_host(&sl.h_pre, &sl.h_pre_cap, (size_t)std::max<int64_t>(L.pre_total, 1)));
    }
    return 0;
}

// Make a layout the slot's current block (its totals and per-channel outputs).
static void slot_adopt(ServeSlot& sl, BlockLayout& L)
{
    sl.outs.swap(L.outs);
    sl.n_wfm = L.n_wfm;
    sl.rds_total = L.rds_total;
    sl.iq_total = L.iq_total;
    sl.audio_total = L.audio_total;
    sl.iq_early = L.iq_early;
    sl.audio_early = L.audio_early;
    sl.n_audio_chans = L.n_audio_chans;
    sl.aout_total = L.aout_total;
    sl.pre_total = L.pre_total;
    sl.n_pre_chans = L.n_pre_chans;
}

// The block's tables for the live channels' current state, written into the
// slot's host table for input `in` (a device pointer) of n samples: outs is
// the block's layout (the slot's buffers must hold it).  Channel state is
// not changed (commit_block advances it).  With cache_items the work items
// are kept for later blocks of the same channel set, length and tune.
static int describe_block(gpud_serve* s, ServeSlot& sl, const std::vector<ServeOut>& outs, int n_wfm, int n,
                          const float2* in, bool cache_items)
{
    const int nch = (int)s->live.size();
    // Work items: reused while the channel set, block length and tune are
    // unchanged (tiles are planned for the most outputs a block of n samples
    // gives; the kernels skip tiles past a block's last output).
    const bool reuse = sl.items_epoch == s->epoch && sl.items_n == n && sl.h_tab;
    if (!reuse) {
        s->items.clear();
    }
    size_t items_bytes = reuse ? (size_t)sl.n_items * sizeof(ArgusWorkItem) : 0;
    // descriptors are written straight into the table, after the items
    // region, once its size is known; stage them in the channel order first
    s->chans.resize((size_t)nch);
    s->ddescs.resize((size_t)nch);
    s->wdescs.resize((size_t)n_wfm);
    s->ditems.clear();
    s->ditems_e.clear();
    s->witems.clear();
    int n_audio = 0;
    for (const ServeOut& o : outs) n_audio += o.a_index >= 0 ? 1 : 0;
    s->adescs.resize((size_t)n_audio);
    s->fdescs.clear();
    s->f_ditems.clear();
    s->f_iq_end.clear();
    s->f_au_end.clear();
    s->pdescs.clear();
    s->pitems.clear();
    size_t lds = 0;
    int ci = 0;
    for (const auto& kv : s->live) {
        const ServeChannel& c = kv.second;
        const ServeOut& o = outs[(size_t)ci];
        const int keep = std::min(c.hist_len + n, c.T - 1);
        const double inc = -2.0 * M_PI * c.spec.offset_hz / (double)s->sample_rate;
        ArgusChanDesc& ch = s->chans[(size_t)ci];
        if (argus_fill_chan(&ch, c.d_hist[c.hist_sel], c.hist_len, c.d_hist[c.hist_sel ^ 1], keep, c.d_taps,
                            c.d_ctaps, sl.d_iq + o.iq_off, c.decim, c.T, c.phase_count, n, c.nco_phase, inc,
                            s->tune) != 0)
            return -1;
        // the fast path, once the channel's history is the stream's own
        // last T-1 samples; its direct items are then only the history
        const bool fast = fast_now(s, c, n);
        const int fi = fast ? (int)s->fdescs.size() : -1;
        if (fast) {
            s->fdescs.push_back(argus_fast_desc(ch, c.d_fastG));
        } else if ((c.decim - 1 + n) / c.decim > 0) {
            lds = std::max(lds, argus_item_lds_bytes(ch));
        }
        if (!reuse) {
            ArgusChanDesc full = ch;
            full.n_out = fast ? 0 : (c.decim - 1 + n) / c.decim;
            argus_append_items(full, ci, s->items);
        }

        if (c.pre_factor > 1) {  // pre-demod stage: baseband -> demodulator input
            ArgusPreDesc pd;
            double* taps = (double*)(c.d_pre + kPreTaps);
            float2* hist = (float2*)(c.d_pre + kPreHist);
            pd.in = sl.d_iq + o.iq_off;
            pd.out = sl.d_pre + o.pre_off;
            pd.taps = taps;
            pd.hist_rd = hist + (size_t)c.pre_sel * (ARGUS_PRE_TAPS - 1);
            pd.hist_wr = hist + (size_t)(c.pre_sel ^ 1) * (ARGUS_PRE_TAPS - 1);
            pd.n_in = o.n_iq;
            pd.n_out = o.n_pre;
            pd.first = c.pre_phase;
            pd.factor = c.pre_factor;
            pd.T = ARGUS_PRE_TAPS;
            pd.pad = 0;
            const int pi = (int)s->pdescs.size();
            s->pdescs.push_back(pd);
            for (int st = 0; st < o.n_pre; st += ARGUS_PRE_TILE)
                s->pitems.push_back({pi, st, std::min(ARGUS_PRE_TILE, o.n_pre - st), 0});
            if (o.n_iq > 0) s->pitems.push_back({pi, 0, 0, 1});
        }
        ArgusDemodDesc& dd = s->ddescs[(size_t)ci];
        dd.iq = c.pre_factor > 1 ? sl.d_pre + o.pre_off : sl.d_iq + o.iq_off;
        dd.audio = c.stereo ? sl.d_mono + o.mono_off : sl.d_audio + o.audio_off;
        dd.carry_rd = c.d_carry + c.carry_sel;
        dd.carry_wr = c.d_carry + (c.carry_sel ^ 1);
        dd.n = o.n_pre;
        dd.mode = c.stereo ? ARGUS_DEMOD_WFM : c.spec.demod;
        dd.have_carry = c.have_carry ? 1 : 0;
        dd.skip = (dd.mode == ARGUS_DEMOD_NFM || dd.mode == ARGUS_DEMOD_WFM) && !c.have_carry ? 1 : 0;
        const double bfo_hz = dd.mode == ARGUS_DEMOD_LSB ? -700.0 : 700.0;
        const double bfo_inc = 2.0 * M_PI * bfo_hz / (double)c.demod_rate;
        dd.bfo0 = argus_rad_to_fix(c.bfo_phase + bfo_inc);
        dd.bfo_inc = argus_rad_to_fix(bfo_inc);
        // the fast kernel demodulates its channel's baseband in its epilogue
        // when the demodulator reads that baseband directly
        const bool fm = dd.mode == ARGUS_DEMOD_NFM || dd.mode == ARGUS_DEMOD_WFM;
        const bool fused = fi >= 0 && s->fuse_demod && c.pre_factor == 1 && !c.stereo &&
                           (fm || dd.mode == ARGUS_DEMOD_USB || dd.mode == ARGUS_DEMOD_LSB ||
                            dd.mode == ARGUS_DEMOD_CW);
        if (fused) {
            ArgusFastDesc& fd = s->fdescs[(size_t)fi];
            fd.demod = dd.mode;
            fd.audio = dd.audio;
            fd.bfo0 = dd.bfo0;
            fd.bfo_inc = dd.bfo_inc;
            fd.skip = dd.skip;
        }
        // the direct form's narrow items (modulated taps) likewise
        const bool fused_direct = fi < 0 && s->fuse_demod && ch.narrow && ch.ctaps && c.pre_factor == 1 &&
                                  !c.stereo &&
                                  (fm || dd.mode == ARGUS_DEMOD_USB || dd.mode == ARGUS_DEMOD_LSB ||
                                   dd.mode == ARGUS_DEMOD_CW);
        if (fused_direct) {
            ch.demod = dd.mode;
            ch.audio = dd.audio;
            ch.bfo0 = dd.bfo0;
            ch.bfo_inc = dd.bfo_inc;
            ch.skip = dd.skip;
        }
        if (dd.mode != ARGUS_DEMOD_NONE && o.n_pre > 0) {
            std::vector<ArgusDemodItem>& di = o.early ? s->ditems_e : s->ditems;
            if (fused) {
                // FM: the segments' first outputs and the carry; SSB/CW: done
                if (fm) di.push_back({ci, ch.e0, ch.D, ARGUS_DEMOD_ITEM_SEAMS});
            } else if (fused_direct) {
                // FM: each narrow item's first output and the carry; SSB/CW: done
                if (fm) di.push_back({ci, 0, ch.M, ARGUS_DEMOD_ITEM_SEAMS_EVERY});
            } else if (dd.mode == ARGUS_DEMOD_AM) {
                di.push_back({ci, 0, o.n_pre, ARGUS_DEMOD_ITEM_OUT});
            } else {
                for (int st = 0; st < o.n_pre; st += ARGUS_DEMOD_TILE)
                    di.push_back({ci, st, std::min(ARGUS_DEMOD_TILE, o.n_pre - st), ARGUS_DEMOD_ITEM_OUT});
            }
        }
        if (fast) {  // early outputs are laid out in channel order from 0
            const int64_t pi = s->f_iq_end.empty() ? 0 : s->f_iq_end.back();
            const int64_t pa = s->f_au_end.empty() ? 0 : s->f_au_end.back();
            s->f_ditems.push_back((int)s->ditems_e.size());
            s->f_iq_end.push_back(o.early ? o.iq_off + o.n_iq : pi);
            s->f_au_end.push_back(o.early ? o.audio_off + o.n_audio : pa);
        }
        if (c.stereo) {
            if (o.wfm_index < 0 || o.wfm_index >= n_wfm) return -1;
            wfm_describe(c, o, sl, s->wdescs[(size_t)o.wfm_index], s->witems);
        }
        if (o.a_index >= 0) {
            if (o.a_index >= n_audio) return -1;  // a layout that is not this channel set's
            audio_describe(c, o, sl, s->adescs[(size_t)o.a_index]);
        }
        ci++;
    }
    if (!reuse) {
        argus_order_items(s->items);
        items_bytes = s->items.size() * sizeof(ArgusWorkItem);
    }
    const int n_items = reuse ? sl.n_items : (int)s->items.size();
    // the early channels' demod items first
    const int n_ditems_e = (int)s->ditems_e.size();
    s->ditems.insert(s->ditems.begin(), s->ditems_e.begin(), s->ditems_e.end());
    const int n_ditems = (int)s->ditems.size(), n_witems = (int)s->witems.size();
    const size_t off_hdr = 0;
    const size_t off_chans = align_up(off_hdr + sizeof(ArgusServeHdr));
    const size_t off_ddesc = align_up(off_chans + (size_t)nch * sizeof(ArgusChanDesc));
    const size_t off_wdesc = align_up(off_ddesc + (size_t)nch * sizeof(ArgusDemodDesc));
    const size_t off_ditems = align_up(off_wdesc + (size_t)n_wfm * sizeof(ArgusWfmDesc));
    const size_t off_witems = align_up(off_ditems + (size_t)n_ditems * sizeof(ArgusDemodItem));
    const size_t off_adesc = align_up(off_witems + (size_t)n_witems * sizeof(ArgusWfmItem));
    const int n_fast = (int)s->fdescs.size();
    const size_t off_fdesc = align_up(off_adesc + (size_t)n_audio * sizeof(ArgusAudioDesc));
    const int n_pre = (int)s->pdescs.size(), n_pitems = (int)s->pitems.size();
    const size_t off_pdesc = align_up(off_fdesc + (size_t)n_fast * sizeof(ArgusFastDesc));
    const size_t off_pitems = align_up(off_pdesc + (size_t)n_pre * sizeof(ArgusPreDesc));
    const size_t total = off_pitems + (size_t)n_pitems * sizeof(ArgusPreItem);
    ARGUS_CHECK((hipError_t)grow_table(sl, total));
    if (!reuse) {
        ARGUS_CHECK((hipError_t)grow_dev(&sl.d_items, &sl.d_items_cap, (size_t)std::max(n_items, 1)));
        ARGUS_CHECK((hipError_t)grow_host(&sl.h_items, &sl.h_items_cap, (size_t)std::max(n_items, 1)));
        if (n_items) std::memcpy(sl.h_items, s->items.data(), items_bytes);
        sl.items_epoch = cache_items ? s->epoch : 0;
        sl.items_n = n;
        sl.n_items = n_items;
        sl.items_dirty = true;
    }
    if (nch) {
        std::memcpy(sl.h_tab + off_chans, s->chans.data(), (size_t)nch * sizeof(ArgusChanDesc));
        std::memcpy(sl.h_tab + off_ddesc, s->ddescs.data(), (size_t)nch * sizeof(ArgusDemodDesc));
    }
    if (n_wfm) std::memcpy(sl.h_tab + off_wdesc, s->wdescs.data(), (size_t)n_wfm * sizeof(ArgusWfmDesc));
    if (n_ditems) std::memcpy(sl.h_tab + off_ditems, s->ditems.data(), (size_t)n_ditems * sizeof(ArgusDemodItem));
    if (n_witems) std::memcpy(sl.h_tab + off_witems, s->witems.data(), (size_t)n_witems * sizeof(ArgusWfmItem));
    // longest inputs first: workgroups are dispatched in order, so the
    // high-rate channels (many chunks each) start first
    std::stable_sort(s->adescs.begin(), s->adescs.end(), [](const ArgusAudioDesc& a, const ArgusAudioDesc& b) {
        return (int64_t)a.n_in * a.channels > (int64_t)b.n_in * b.channels;
    });
    if (n_audio) std::memcpy(sl.h_tab + off_adesc, s->adescs.data(), (size_t)n_audio * sizeof(ArgusAudioDesc));
    if (n_fast) std::memcpy(sl.h_tab + off_fdesc, s->fdescs.data(), (size_t)n_fast * sizeof(ArgusFastDesc));
    if (n_pre) std::memcpy(sl.h_tab + off_pdesc, s->pdescs.data(), (size_t)n_pre * sizeof(ArgusPreDesc));
    if (n_pitems) std::memcpy(sl.h_tab + off_pitems, s->pitems.data(), (size_t)n_pitems * sizeof(ArgusPreItem));
    ArgusServeHdr& hdr = *(ArgusServeHdr*)(sl.h_tab + off_hdr);
    const char* d = sl.d_tab;
    hdr.x = ArgusExtractHdr{(const ArgusChanDesc*)(d + off_chans), sl.d_items, in, n_items, n};
    // the fast channels in n_parts parts, one per part_min_bytes of early
    // outputs the block downloads (at most max_parts; parts after n_parts
    // are empty and not launched); part p's early demodulation items and
    // output ranges follow from the channel order
    {
        const int64_t iq_e = s->f_iq_end.empty() ? 0 : s->f_iq_end.back();
        const int64_t au_e = s->f_au_end.empty() ? 0 : s->f_au_end.back();
        const int64_t dl_bytes = ((s->dl_mask & GPUD_DL_IQ) ? (int64_t)sizeof(float2) * iq_e : 0) +
                                 ((s->dl_mask & GPUD_DL_DEMOD) ? (int64_t)sizeof(float) * au_e : 0);
        const int64_t by_bytes = s->part_min_bytes > 0 ? dl_bytes / s->part_min_bytes : s->max_parts;
        sl.n_parts = (int)std::max<int64_t>(1, std::min<int64_t>(s->max_parts, by_bytes));
    }
    // With several parts the first is half the size of the others, so the
    // first download starts sooner (part 1's kernels still end long before
    // part 0's outputs are down): boundaries at (2p + 1) / (2 n_parts - 1).
    int part_end[ARGUS_FAST_PARTS];
    int e_prev = 0;
    for (int p = 0; p < ARGUS_FAST_PARTS; p++) {
        const int k = p < sl.n_parts
                          ? (int)(((int64_t)n_fast * (2 * p + 1)) / (2 * sl.n_parts - 1))
                          : n_fast;
        part_end[p] = k;
        const int e_end = k ? s->f_ditems[(size_t)k - 1] : 0;
        hdr.de[p] = ArgusDemodHdr{(const ArgusDemodDesc*)(d + off_ddesc), (const ArgusDemodItem*)(d + off_ditems) + e_prev,
                                  e_end - e_prev, 0};
        e_prev = e_end;
        sl.iq_part_end[p] = k ? s->f_iq_end[(size_t)k - 1] : 0;
        sl.audio_part_end[p] = k ? s->f_au_end[(size_t)k - 1] : 0;
    }
    if (e_prev != n_ditems_e) return -1;  // every early item belongs to a fast channel
    hdr.d = ArgusDemodHdr{(const ArgusDemodDesc*)(d + off_ddesc), (const ArgusDemodItem*)(d + off_ditems) + n_ditems_e,
                          n_ditems - n_ditems_e, 0};
    hdr.w = ArgusWfmHdr{(const ArgusWfmDesc*)(d + off_wdesc), (const ArgusWfmItem*)(d + off_witems), n_wfm, n_witems};
    hdr.a = ArgusAudioHdr{(const ArgusAudioDesc*)(d + off_adesc), n_audio, 0};
    // the fast stage runs while any live channel can take the fast form
    // (channels come and go without changing what the captured graph holds)
    sl.fast_S = s->fast && s->n_fast_capable > 0 ? argus_fast_segments(n) : 0;
    hdr.f = ArgusFastHdr{(const ArgusFastDesc*)(d + off_fdesc), s->d_ftail, s->d_fu, s->d_fX, n_fast, sl.fast_S};
    std::memcpy(hdr.f.part_end, part_end, sizeof(part_end));
    sl.a_mode = n_audio > 0 ? 1 : 0;
    hdr.p = ArgusPreHdr{(const ArgusPreDesc*)(d + off_pdesc), (const ArgusPreItem*)(d + off_pitems), n_pitems, 0};
    sl.p_mode = n_pitems > 0 ? 1 : 0;
    sl.tab_bytes = total;
    sl.off_hdr = off_hdr;
    // dynamic LDS: the largest any block has needed since the tune was set,
    // in 8 KiB steps, so channels coming and going do not change what the
    // captured graphs launch (it only grows, a few times at most)
    const size_t lq = lds ? std::min<size_t>((lds + 8191) & ~(size_t)8191, 160 * 1024) : 0;
    if (cache_items) s->lds_hw = std::max(s->lds_hw, lq);
    sl.lds = cache_items ? s->lds_hw : lq;
    return 0;
}

// Advance every live channel's carried state past a block of n samples with
// the given layout (the values describe_block used become the previous
// block's): NCO and decimator phase, history and FM-carry halves, BFO phase,
// stereo delay-line half and RDS sample count.
static void commit_block(gpud_serve* s, const std::vector<ServeOut>& outs, int n, int64_t ticket)
{
    int ci = 0;
    bool warmed = false;
    for (auto& kv : s->live) {
        ServeChannel& c = kv.second;
        const ServeOut& o = outs[(size_t)ci++];
        const int keep = std::min(c.hist_len + n, c.T - 1);
        if (c.d_fastG && c.hist_len < c.T - 1 && keep == c.T - 1) warmed = true;
        const double inc = -2.0 * M_PI * c.spec.offset_hz / (double)s->sample_rate;
        const int mode = c.stereo ? ARGUS_DEMOD_WFM : c.spec.demod;
        const double bfo_inc = 2.0 * M_PI * (mode == ARGUS_DEMOD_LSB ? -700.0 : 700.0) / (double)c.demod_rate;
        c.phase_count = (c.phase_count + n) % c.decim;
        c.nco_phase = argus_reduce_phase(c.nco_phase + inc * (double)n);
        if (keep > 0) c.hist_sel ^= 1;
        c.hist_len = keep;
        if (c.pre_factor > 1) {
            argus_pre_block(c.pre_phase, c.pre_factor, o.n_iq, nullptr, &c.pre_phase);
            if (o.n_iq > 0) c.pre_sel ^= 1;
        }
        if ((mode == ARGUS_DEMOD_NFM || mode == ARGUS_DEMOD_WFM) && o.n_pre > 0) {
            c.carry_sel ^= 1;
            c.have_carry = true;
        }
        c.bfo_phase = argus_reduce_phase(c.bfo_phase + bfo_inc * (double)o.n_pre);
        c.last_ticket = ticket;
        if (c.stereo) {
            if (o.n_mono > 0) c.wfm_sel ^= 1;
            c.rds_count += o.n_mono;
        }
        if (o.a_index >= 0) {
            c.at0 = o.a_tnext;
            if ((c.stereo ? o.n_mono : o.n_audio) > 0) c.a_sel ^= 1;
        }
    }
    if (warmed && s->fast) s->epoch++;  // those channels' work items change (fast path from the next block)
}

// Layout, buffers and tables of a block in slot sl, built at submit time.
static int build_block(gpud_serve* s, ServeSlot& sl, int n, const float2* in, bool cache_items)
{
    BlockLayout L;
    layout_block(s, n, L);
    int err = slot_grow(sl, L);
    if (err) return err;
    slot_adopt(sl, L);
    return describe_block(s, sl, sl.outs, sl.n_wfm, n, in, cache_items);
}

// Upload the block's table: the work items only when they were rebuilt.
static int upload_table(gpud_serve* s, ServeSlot& sl)
{
    if (sl.items_dirty) {
        if (sl.n_items > 0)
            ARGUS_CHECK(hipMemcpyAsync(sl.d_items, sl.h_items, (size_t)sl.n_items * sizeof(ArgusWorkItem),
                                       hipMemcpyHostToDevice, s->compute));
        sl.items_dirty = false;
    }
    ARGUS_CHECK(hipMemcpyAsync(sl.d_tab, sl.h_tab, sl.tab_bytes, hipMemcpyHostToDevice, s->compute));
    return 0;
}

// The block's launches in ARGUS_FAST_PARTS + 2 stages, each reading its
// counts and pointers from the table header on the device: parts 0..P-1 of
// the fast form (part 0 also stages the input and transforms the segments),
// each with its early channels' demodulation; the direct form's extraction
// and the tail; then the other channels' pre-demod stage, demodulation,
// stereo and audio.
constexpr int kServeStages = ARGUS_FAST_PARTS + 2;
constexpr int kStageDirect = ARGUS_FAST_PARTS;

static int launch_stage(gpud_serve* s, ServeSlot& sl, int stage, bool wfm)
{
    const ArgusServeHdr* dh = (const ArgusServeHdr*)(sl.d_tab + sl.off_hdr);
    if (stage < ARGUS_FAST_PARTS) {
        int err = stage == 0 ? argus_launch_fast_input(&dh->f, &dh->x, sl.fast_S, s->d_fu, s->d_fX, s->compute) : 0;
        if (!err) err = argus_launch_fast_chans(&dh->f, sl.fast_S, s->tune.fast_grid, stage, s->compute);
        return err ? err : argus_launch_demod_hdr(&dh->de[stage], s->compute);
    }
    if (stage == kStageDirect) {
        int err = argus_launch_extract_hdr(&dh->x, sl.lds, s->tune.grid, s->compute);
        return err ? err : argus_launch_fast_tail(&dh->f, &dh->x, s->compute);
    }
    int err = sl.p_mode ? argus_launch_predemod_hdr(&dh->p, s->compute) : 0;
    if (err) return err;
    err = argus_launch_demod_hdr(&dh->d, s->compute);
    if (err) return err;
    if (wfm) {
        err = argus_launch_wfm_hdr(&dh->w, s->compute);
        if (err) return err;
    }
    return sl.a_mode ? argus_launch_audio_hdr(&dh->a, s->compute) : 0;
}

// Run the block's stages, recording ev_part[p] after fast part p (its early
// channels' outputs are complete: their download may start) and ev_xend
// after the direct form (the extraction's end, for extract_ms).  In graph
// mode each stage is the replay of a captured per-slot graph, captured again
// when what it depends on (table address, LDS size, grid, which stages run)
// has changed; the events stay outside the graphs.
static int launch_block(gpud_serve* s, ServeSlot& sl)
{
    const bool wfm = sl.n_wfm > 0;
    if (!s->graph) {
        for (int stage = 0; stage < kServeStages; stage++) {
            if (stage >= sl.n_parts && stage < ARGUS_FAST_PARTS) continue;  // an empty part
            int err = launch_stage(s, sl, stage, wfm);
            if (err) return err;
            if (stage < ARGUS_FAST_PARTS) ARGUS_CHECK(hipEventRecord(sl.ev_part[stage], s->compute));
            if (stage == kStageDirect) ARGUS_CHECK(hipEventRecord(sl.ev_xend, s->compute));
        }
        return 0;
    }
    const char* hdr = sl.d_tab + sl.off_hdr;
    if (!sl.gexec[0] || sl.g_hdr != hdr || sl.g_lds != sl.lds || sl.g_grid != s->tune.grid ||
        sl.g_fgrid != s->tune.fast_grid ||
        sl.g_wfm != (int)wfm || sl.g_audio != sl.a_mode || sl.g_fastS != sl.fast_S || sl.g_pre != sl.p_mode) {
        for (auto& ge : sl.gexec)
            if (ge) {
                hipGraphExecDestroy(ge);
                ge = nullptr;
            }
        int lerr = argus_extract_allow_lds(sl.lds);  // outside capture
        if (lerr) return lerr;
        for (int stage = 0; stage < kServeStages; stage++) {
            ARGUS_CHECK(hipStreamBeginCapture(s->compute, hipStreamCaptureModeThreadLocal));
            int err = launch_stage(s, sl, stage, wfm);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(s->compute, &g);
            if (err || e != hipSuccess) {
                if (g) hipGraphDestroy(g);
                return err ? err : (int)e;
            }
            e = hipGraphInstantiate(&sl.gexec[stage], g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) {
                sl.gexec[stage] = nullptr;
                return (int)e;
            }
        }
        s->graph_captures++;
        sl.g_hdr = hdr;
        sl.g_lds = sl.lds;
        sl.g_grid = s->tune.grid;
        sl.g_fgrid = s->tune.fast_grid;
        sl.g_wfm = (int)wfm;
        sl.g_audio = sl.a_mode;
        sl.g_fastS = sl.fast_S;
        sl.g_pre = sl.p_mode;
    }
    for (int stage = 0; stage < kServeStages; stage++) {
        if (stage >= sl.n_parts && stage < ARGUS_FAST_PARTS) continue;  // an empty part
        ARGUS_CHECK(hipGraphLaunch(sl.gexec[stage], s->compute));
        if (stage < ARGUS_FAST_PARTS) ARGUS_CHECK(hipEventRecord(sl.ev_part[stage], s->compute));
        if (stage == kStageDirect) ARGUS_CHECK(hipEventRecord(sl.ev_xend, s->compute));
    }
    return 0;
}

// A channel's RDS decoding block on the device (ServeChannel::d_rdec):
// decoder state | call progress | bits | match flags | next-hit table |
// two windows of W baseband samples.
struct RdsLayout {
    int bit_cap, group_cap;
    size_t off_pr, off_bits, off_flags, off_next, off_win, bytes;
};

static RdsLayout rds_layout(const ServeChannel& c, int W)
{
    RdsLayout l{};
    argus_rds_caps(W, c.demod_rate / c.rds_decim, &l.bit_cap, &l.group_cap);
    l.off_pr = align_up(sizeof(ArgusRdsState));
    l.off_bits = l.off_pr + align_up(sizeof(ArgusRdsProgress));
    l.off_flags = l.off_bits + align_up((size_t)l.bit_cap);
    l.off_next = l.off_flags + align_up((size_t)l.bit_cap);
    l.off_win = l.off_next + align_up(8 * (size_t)l.bit_cap);
    l.bytes = l.off_win + 2 * (size_t)W * sizeof(float2);
    return l;
}

// Device-side RDS decoding of a launched block, on the compute stream after
// its graphs.  The block's RDS baseband of every decoding channel goes into
// the channel's current window (one append launch); the symbol loop of each
// channel's Decode call advances over what its window holds (the calls run
// across the blocks that fill the window), and the calls whose window this
// block completes finish and run their group scan.  Advances the channels'
// window state, as commit_block does for the rest.
static int rds_stage(gpud_serve* s, ServeSlot& sl)
{
    sl.n_rdec = 0;
    sl.rdec_channel.clear();
    sl.rdec_groups.clear();
    if (s->n_rds_on == 0) return 0;
    auto& ap = s->rapp;
    auto& calls = s->rcalls;
    ap.clear();
    calls.clear();
    std::vector<ArgusRdsCall> streaming;
    size_t group_words = 0;
    int max_cap = 0;
    int ci = 0;
    for (auto& kv : s->live) {
        ServeChannel& c = kv.second;
        ServeOut& o = sl.outs[(size_t)ci++];
        o.r_index = -1;
        if (!o.r_on || c.rds_win <= 0 || o.n_rds <= 0) continue;
        const int W = c.rds_win;  // >= the most RDS samples of one block
        const RdsLayout l = rds_layout(c, W);
        float2* win0 = (float2*)(c.d_rdec + l.off_win);
        float2* cur = win0 + (size_t)c.rwin_sel * W;
        float2* nxt = win0 + (size_t)(c.rwin_sel ^ 1) * W;
        const int n0 = std::min(o.n_rds, W - c.rwin_fill);
        ap.push_back({sl.d_rds + o.rds_off, cur + c.rwin_fill, nxt, n0, o.n_rds - n0});
        c.rwin_fill += n0;
        ArgusRdsCall k{};
        k.x = cur;
        k.bits = (uint8_t*)(c.d_rdec + l.off_bits);
        k.flags = (uint8_t*)(c.d_rdec + l.off_flags);
        k.next = (int32_t*)(c.d_rdec + l.off_next);
        k.st = (ArgusRdsState*)c.d_rdec;
        k.pr = (ArgusRdsProgress*)(c.d_rdec + l.off_pr);
        k.n = W;
        k.avail = c.rwin_fill;
        k.rate = o.rds_rate;
        k.bit_cap = l.bit_cap;
        k.group_cap = l.group_cap;
        k.final = c.rwin_fill == W ? 1 : 0;
        if (!k.final) {
            streaming.push_back(k);
            continue;
        }
        o.r_index = (int)calls.size();
        calls.push_back(k);
        sl.rdec_channel.push_back(kv.first);
        sl.rdec_groups.push_back(group_words);
        group_words += 4 * (size_t)l.group_cap;
        max_cap = std::max(max_cap, l.bit_cap);
        c.rwin_sel ^= 1;
        c.rwin_fill = o.n_rds - n0;
    }
    if (ap.empty()) return 0;
    const int n_final = (int)calls.size();
    calls.insert(calls.end(), streaming.begin(), streaming.end());
    // table: appends | calls (final ones first) | outs | groups
    const size_t off_calls = align_up(ap.size() * sizeof(ArgusRdsAppend));
    const size_t off_outs = off_calls + align_up(calls.size() * sizeof(ArgusRdsCall));
    const size_t off_groups = off_outs + align_up(std::max(n_final, 1) * sizeof(ArgusRdsOut));
    const size_t tab = off_groups + std::max<size_t>(group_words, 1) * sizeof(uint16_t);
    ARGUS_CHECK((hipError_t)grow_dev(&sl.d_rtab, &sl.rtab_cap, tab));
    ARGUS_CHECK((hipError_t)grow_host(&sl.h_rtab, &sl.h_rtab_cap, tab));
    ArgusRdsOut* d_outs = (ArgusRdsOut*)(sl.d_rtab + off_outs);
    uint16_t* d_groups = (uint16_t*)(sl.d_rtab + off_groups);
    for (size_t i = 0; i < calls.size(); i++) {
        ArgusRdsCall& k = calls[i];
        const bool fin = (int)i < n_final;
        k.out = d_outs + (fin ? i : 0);  // a non-final call writes no result
        k.groups = fin ? d_groups + sl.rdec_groups[i] : nullptr;
    }
    std::memcpy(sl.h_rtab, ap.data(), ap.size() * sizeof(ArgusRdsAppend));
    std::memcpy(sl.h_rtab + off_calls, calls.data(), calls.size() * sizeof(ArgusRdsCall));
    ARGUS_CHECK(hipMemcpyAsync(sl.d_rtab, sl.h_rtab, off_outs, hipMemcpyHostToDevice, s->compute));
    int err = argus_launch_rds_append((const ArgusRdsAppend*)sl.d_rtab, (int)ap.size(), s->compute);
    if (err) return err;
    err = argus_launch_rds((const ArgusRdsCall*)(sl.d_rtab + off_calls), (int)calls.size(), n_final, max_cap,
                           s->compute);
    if (err) return err;
    sl.n_rdec = n_final;
    sl.off_rout = off_outs;
    sl.off_rgroups = off_groups;
    sl.rgroups_bytes = tab - off_outs;  // outs and groups come back in one copy
    return 0;
}

// Enqueue one block: upload (unless device-resident), extract, demodulate,
// download.  Returns immediately; *ticket identifies the block.
// Build and upload the next block's tables ahead, into the next slot, while
// the GPU works on the block just launched (ServePrep).  Skipped when that
// slot's previous block is still running or its buffers would have to grow
// (the outputs it holds stay valid until the next submit).
static void prepare_next(gpud_serve* s, int n, const float2* in_dev)
{
    const int idx = s->next_slot;
    ServeSlot& sl = s->slots[idx];
    if (sl.busy) {
        if (hipEventQuery(sl.ev_done) != hipSuccess) return;
        if (slot_wait(s, sl) != 0) return;  // already complete: only marks it so
    }
    ServePrep& pp = s->prep;
    layout_block(s, n, pp.layout);
    if (!slot_fits(sl, pp.layout)) return;
    const float2* in = in_dev ? in_dev : sl.d_in;
    if (describe_block(s, sl, pp.layout.outs, pp.layout.n_wfm, n, in, true) != 0) return;
    if (upload_table(s, sl) != 0) return;
    pp.slot = idx;
    pp.ticket = s->last_ticket + 1;
    pp.epoch = s->epoch;
    pp.n = n;
    pp.in = in;
    pp.valid = true;
}

static int serve_submit(gpud_serve* s, const gpud_float2* iq, int n, bool on_device, int64_t* ticket)
{
    if (!s || n < 0 || n > s->block_samples || (n > 0 && !iq)) return -1;
    const int idx = s->next_slot;
    ServeSlot& sl = s->slots[idx];
    int err = slot_wait(s, sl);  // the slot's previous block must be done
    if (err) return err;
    s->next_slot = (idx + 1) % ARGUS_SERVE_SLOTS;
    sl.ticket = ++s->last_ticket;
    sl.n = n;
    const float2* in = on_device ? (const float2*)iq : sl.d_in;
    // tables prepared while the previous block ran, if nothing they depend on
    // has changed since; else build them now.  Only the header holds the
    // input pointer, so a block from another device buffer (a ring of
    // ingest slots) keeps the prepared tables and re-sends the header.
    ServePrep& pp = s->prep;
    const bool prepared =
        pp.valid && pp.slot == idx && pp.ticket == sl.ticket && pp.epoch == s->epoch && pp.n == n;
    const bool repoint = prepared && pp.in != in;
    pp.valid = false;
    if (prepared) {
        slot_adopt(sl, pp.layout);
        if (repoint) ((ArgusServeHdr*)(sl.h_tab + sl.off_hdr))->x.in = in;
    } else {
        err = build_block(s, sl, n, in, true);
        if (err) return err;
    }

    // ---- copy stream: block upload -------------------------------------
    ARGUS_CHECK(hipEventRecord(sl.ev_start, s->copy));
    if (!on_device && n > 0) {
        std::memcpy(sl.h_in, iq, (size_t)n * sizeof(float2));
        ARGUS_CHECK(hipMemcpyAsync(sl.d_in, sl.h_in, (size_t)n * sizeof(float2), hipMemcpyHostToDevice, s->copy));
    }
    ARGUS_CHECK(hipEventRecord(sl.ev_in, s->copy));
    // ---- compute stream: one table copy, then the block's kernels ----------
    ARGUS_CHECK(hipStreamWaitEvent(s->compute, sl.ev_in, 0));
    if (!prepared) {  // a prepared table was uploaded when it was built
        err = upload_table(s, sl);
        if (err) return err;
    } else if (repoint) {
        ARGUS_CHECK(hipMemcpyAsync(sl.d_tab + sl.off_hdr, sl.h_tab + sl.off_hdr, sizeof(ArgusServeHdr),
                                   hipMemcpyHostToDevice, s->compute));
    }
    ARGUS_CHECK(hipEventRecord(sl.ev_kstart, s->compute));
    err = launch_block(s, sl);
    if (err) return err;
    err = rds_stage(s, sl);
    if (err) return err;
    ARGUS_CHECK(hipEventRecord(sl.ev_kern, s->compute));
    commit_block(s, sl.outs, n, sl.ticket);
    // ---- copy stream: download ------------------------------------------
    // each fast part's early outputs as soon as they are complete, while the
    // later parts and the direct form run; then the rest
    const int dl = s->dl_mask;
    int64_t iq_lo = 0, au_lo = 0;
    for (int p = 0; p < sl.n_parts; p++) {
        const int64_t iq_hi = sl.iq_part_end[p], au_hi = sl.audio_part_end[p];
        const bool iq_part = (dl & GPUD_DL_IQ) && iq_hi > iq_lo;
        const bool au_part = (dl & GPUD_DL_DEMOD) && au_hi > au_lo;
        if (iq_part || au_part) ARGUS_CHECK(hipStreamWaitEvent(s->copy, sl.ev_part[p], 0));
        if (iq_part)
            ARGUS_CHECK(hipMemcpyAsync(sl.h_iq + iq_lo, sl.d_iq + iq_lo, (size_t)(iq_hi - iq_lo) * sizeof(float2),
                                       hipMemcpyDeviceToHost, s->copy));
        if (au_part)
            ARGUS_CHECK(hipMemcpyAsync(sl.h_audio + au_lo, sl.d_audio + au_lo, (size_t)(au_hi - au_lo) * sizeof(float),
                                       hipMemcpyDeviceToHost, s->copy));
        iq_lo = std::max(iq_lo, iq_hi);
        au_lo = std::max(au_lo, au_hi);
    }
    ARGUS_CHECK(hipStreamWaitEvent(s->copy, sl.ev_kern, 0));
    if ((dl & GPUD_DL_IQ) && sl.iq_total > sl.iq_early)
        ARGUS_CHECK(hipMemcpyAsync(sl.h_iq + sl.iq_early, sl.d_iq + sl.iq_early,
                                   (size_t)(sl.iq_total - sl.iq_early) * sizeof(float2), hipMemcpyDeviceToHost,
                                   s->copy));
    if ((dl & GPUD_DL_DEMOD) && sl.audio_total > sl.audio_early)
        ARGUS_CHECK(hipMemcpyAsync(sl.h_audio + sl.audio_early, sl.d_audio + sl.audio_early,
                                   (size_t)(sl.audio_total - sl.audio_early) * sizeof(float), hipMemcpyDeviceToHost,
                                   s->copy));
    if ((dl & GPUD_DL_WFM) && sl.rds_total > 0)
        ARGUS_CHECK(hipMemcpyAsync(sl.h_rds, sl.d_rds, (size_t)sl.rds_total * sizeof(float2), hipMemcpyDeviceToHost,
                                   s->copy));
    if ((dl & GPUD_DL_WFM) && sl.n_wfm > 0)
        ARGUS_CHECK(hipMemcpyAsync(sl.h_wblk, sl.d_wblk, (size_t)sl.n_wfm * sizeof(ArgusWfmBlock),
                                   hipMemcpyDeviceToHost, s->copy));
    if ((dl & GPUD_DL_AUDIO) && sl.aout_total > 0)
        ARGUS_CHECK(hipMemcpyAsync(sl.h_aout, sl.d_aout, (size_t)sl.aout_total * sizeof(float),
                                   hipMemcpyDeviceToHost, s->copy));
    if ((dl & GPUD_DL_PREDEMOD) && sl.pre_total > 0)
        ARGUS_CHECK(hipMemcpyAsync(sl.h_pre, sl.d_pre, (size_t)sl.pre_total * sizeof(float2), hipMemcpyDeviceToHost,
                                   s->copy));
    if (sl.n_rdec > 0)  // decoded RDS groups come back whether or not outputs do
        ARGUS_CHECK(hipMemcpyAsync(sl.h_rtab + sl.off_rout, sl.d_rtab + sl.off_rout, sl.rgroups_bytes,
                                   hipMemcpyDeviceToHost, s->copy));
    ARGUS_CHECK(hipEventRecord(sl.ev_done, s->copy));
    sl.busy = true;
    if (ticket) *ticket = sl.ticket;
    prepare_next(s, n, on_device ? in : nullptr);
    return 0;
}

static ServeSlot* find_slot(gpud_serve* s, int64_t ticket)
{
    for (int i = 0; i < ARGUS_SERVE_SLOTS; i++)
        if (s->slots[i].ticket == ticket) return &s->slots[i];
    return nullptr;
}

extern "C" {

GPUD_API int gpud_serve_submit(gpud_serve* s, const gpud_float2* iq_host, int n, int64_t* ticket)
{
    return serve_submit(s, iq_host, n, false, ticket);
}

GPUD_API int gpud_serve_submit_device(gpud_serve* s, const gpud_float2* iq_dev, int n, int64_t* ticket)
{
    return serve_submit(s, iq_dev, n, true, ticket);
}

GPUD_API int gpud_serve_set_download(gpud_serve* s, int enable)
{
    if (!s) return -1;
    s->dl_mask = enable ? GPUD_DL_ALL : 0;
    return 0;
}

GPUD_API int gpud_serve_set_download_parts(gpud_serve* s, int mask)
{
    if (!s || (mask & ~GPUD_DL_ALL)) return -1;
    s->dl_mask = mask;
    return 0;
}

GPUD_API int gpud_serve_set_download_split(gpud_serve* s, int max_parts, int64_t min_part_bytes)
{
    if (!s || max_parts < 1 || max_parts > ARGUS_FAST_PARTS || min_part_bytes < 0) return -1;
    s->max_parts = max_parts;
    s->part_min_bytes = min_part_bytes;
    return 0;
}

GPUD_API int gpud_serve_poll(gpud_serve* s, int64_t ticket)
{
    if (!s) return -1;
    ServeSlot* sl = find_slot(s, ticket);
    if (!sl) return -2;
    if (!sl->busy) return 1;
    const hipError_t e = hipEventQuery(sl->ev_done);
    if (e == hipSuccess) return 1;
    return e == hipErrorNotReady ? 0 : (int)e;
}

GPUD_API int gpud_serve_counters(gpud_serve* s, int64_t* graph_captures, uint64_t* arena_free)
{
    if (!s) return -1;
    if (graph_captures) *graph_captures = s->graph_captures;
    if (arena_free) {
        uint64_t f = 0;
        for (const auto& kv : s->arena.free_list) f += kv.second;
        *arena_free = f;
    }
    return 0;
}

// Wait for a block; fills per-block totals and timing (event-timed GPU
// milliseconds from upload start to download end, and kernels only).
GPUD_API int gpud_serve_wait(gpud_serve* s, int64_t ticket, gpud_serve_block_info* info)
{
    if (!s) return -1;
    ServeSlot* sl = find_slot(s, ticket);
    if (!sl) return -2;  // never submitted, or its slot was reused
    int err = slot_wait(s, *sl);
    if (err) return err;
    if (info) {
        info->ticket = sl->ticket;
        info->n_channels = (int)sl->outs.size();
        info->n_samples = sl->n;
        info->iq_total = sl->iq_total;
        info->audio_total = sl->audio_total;
        float ms = 0.0f;
        ARGUS_CHECK(hipEventElapsedTime(&ms, sl->ev_start, sl->ev_done));
        info->gpu_ms = ms;
        ARGUS_CHECK(hipEventElapsedTime(&ms, sl->ev_kstart, sl->ev_kern));
        info->kernel_ms = ms;
        ARGUS_CHECK(hipEventElapsedTime(&ms, sl->ev_kstart, sl->ev_xend));
        info->extract_ms = ms;
    }
    return 0;
}

// Outputs of one channel in a completed block (pinned host memory, valid
// until the slot is reused by a later submit).  Returns -1 if the channel
// was not part of the block.
GPUD_API int gpud_serve_channel_output(gpud_serve* s, int64_t ticket, int64_t channel_id, const gpud_float2** iq,
                                       int* n_iq, const float** audio, int* n_audio, int* rate)
{
    if (!s) return -1;
    ServeSlot* sl = find_slot(s, ticket);
    if (!sl || sl->busy) return -2;
    for (const ServeOut& o : sl->outs)
        if (o.channel_id == channel_id) {
            if (iq) *iq = (const gpud_float2*)(sl->h_iq + o.iq_off);
            if (n_iq) *n_iq = o.n_iq;
            if (audio) *audio = sl->h_audio + o.audio_off;
            if (n_audio) *n_audio = o.n_audio;
            if (rate) *rate = o.rate;
            return 0;
        }
    return -1;
}

GPUD_API int gpud_serve_channel_predemod(gpud_serve* s, int64_t ticket, int64_t channel_id, const gpud_float2** iq,
                                         int* n, int* rate)
{
    if (!s) return -1;
    ServeSlot* sl = find_slot(s, ticket);
    if (!sl || sl->busy) return -2;
    for (const ServeOut& o : sl->outs)
        if (o.channel_id == channel_id) {
            if (o.pre_off < 0) return -3;
            auto it = s->live.find(channel_id);
            if (iq) *iq = (const gpud_float2*)(sl->h_pre + o.pre_off);
            if (n) *n = o.n_pre;
            if (rate) *rate = it != s->live.end() ? it->second.demod_rate : 0;
            return 0;
        }
    return -1;
}

GPUD_API int gpud_serve_channel_audio(gpud_serve* s, int64_t ticket, int64_t channel_id, const float** audio,
                                      int* n_frames, int* channels, int* rate)
{
    if (!s) return -1;
    ServeSlot* sl = find_slot(s, ticket);
    if (!sl || sl->busy) return -2;
    for (const ServeOut& o : sl->outs)
        if (o.channel_id == channel_id) {
            if (o.a_index < 0) return -3;
            if (audio) *audio = sl->h_aout + o.aout_off;
            if (n_frames) *n_frames = o.n_aout;
            if (channels) *channels = o.a_channels;
            if (rate) *rate = o.a_rate;
            return 0;
        }
    return -1;
}

// Device bytes of a channel's audio stage: its state block plus its share of
// the per-slot audio outputs.
static size_t audio_footprint(const gpud_serve* s, const ServeChannel& c, int L, int M)
{
    const int64_t n_in = (int64_t)s->block_samples / c.decim + 2;
    const int64_t n_out = n_in * L / M + 4;
    const int ch = c.stereo ? 2 : 1;
    return kAudBytes + (size_t)(n_out * ch) * sizeof(float) * ARGUS_SERVE_SLOTS;
}

GPUD_API int gpud_serve_set_audio(gpud_serve* s, int64_t channel_id, int audio_rate, double deemphasis_us)
{
    if (!s || audio_rate < 0 || !(deemphasis_us >= 0.0)) return -1;
    auto it = s->live.find(channel_id);
    if (it == s->live.end()) return -1;
    ServeChannel& c = it->second;
    for (auto& sl : s->slots) {  // blocks in flight read the stage's current state
        int err = slot_wait(s, sl);
        if (err) return err;
    }
    const size_t old_fp = c.audio_rate > 0 ? audio_footprint(s, c, c.aL, c.aM) : 0;
    if (audio_rate == 0) {
        s->arena.release(c.d_aud, c.aud_bytes);
        c.d_aud = nullptr;
        c.aud_bytes = 0;
        c.audio_rate = 0;
        c.footprint -= old_fp;
        s->bytes_in_use -= old_fp;
        s->epoch++;
        return 0;
    }
    // plain WFM is low-passed at 15 kHz first (folded into the filter); the
    // stereo stage's frames already are
    const bool lowpass = c.spec.demod == GPUD_DEMOD_WFM;
    int L = 1, M = 1, T = 1;
    if (argus_audio_filter(c.demod_rate, audio_rate, lowpass, &L, &M, &T, nullptr) != 0) return -1;
    const size_t fp = audio_footprint(s, c, L, M);
    if (s->bytes_in_use - old_fp + fp > s->budget) return GPUD_ERR_BUDGET;
    const auto key = std::make_tuple(c.demod_rate, audio_rate, lowpass);
    auto fi = s->filters.find(key);
    if (fi == s->filters.end()) {
        std::vector<float> h((size_t)L * T);
        if (argus_audio_filter(c.demod_rate, audio_rate, lowpass, nullptr, nullptr, nullptr, h.data()) != 0)
            return -1;
        float* d = nullptr;
        ARGUS_CHECK(dev_alloc(&d, h.size()));
        hipError_t e = hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            hipFree(d);
            return (int)e;
        }
        fi = s->filters.emplace(key, d).first;
    }
    // fresh state: zero delay lines and de-emphasis
    if (!c.d_aud) {
        c.d_aud = s->arena.alloc(kAudBytes);
        if (!c.d_aud) return GPUD_ERR_BUDGET;
        c.aud_bytes = kAudBytes;
    }
    const std::vector<char> zero(kAudBytes, 0);
    const int ue = stage_upload(s, c.d_aud, zero.data(), kAudBytes);
    if (ue) return ue;
    c.audio_rate = audio_rate;
    c.deemph_us = deemphasis_us;
    c.aL = L;
    c.aM = M;
    c.aT = T;
    c.a_ident = L == M;
    c.at0 = 0;
    c.a_sel = 0;
    c.d_arms = fi->second;
    c.footprint += fp - old_fp;
    s->bytes_in_use += fp - old_fp;
    s->epoch++;
    return 0;
}

// Device bytes of a channel's RDS decoding: its block and its share of the
// per-slot decode tables.
static size_t rds_footprint(const ServeChannel& c, int W)
{
    const RdsLayout l = rds_layout(c, W);
    const size_t per_slot =
        sizeof(ArgusRdsAppend) + sizeof(ArgusRdsCall) + sizeof(ArgusRdsOut) + 8 * (size_t)l.group_cap;
    return l.bytes + per_slot * ARGUS_SERVE_SLOTS;
}

GPUD_API int gpud_serve_set_rds(gpud_serve* s, int64_t channel_id, int window_samples)
{
    if (!s || window_samples < 0) return -1;
    auto it = s->live.find(channel_id);
    if (it == s->live.end() || !it->second.stereo) return -1;
    ServeChannel& c = it->second;
    // the most RDS samples one block can carry, and a window that holds a group
    const int per_block = (s->block_samples / c.decim / c.pre_factor + 2) / c.rds_decim + 2;
    if (window_samples > 0 && window_samples < std::max(per_block, 104)) return -1;
    if ((int64_t)window_samples * 2 * (int64_t)sizeof(float2) > ((int64_t)1 << 31)) return -1;
    for (auto& sl : s->slots) {  // blocks in flight read and write the decoder's state
        int err = slot_wait(s, sl);
        if (err) return err;
    }
    const size_t old_fp = c.rds_win > 0 ? rds_footprint(c, c.rds_win) : 0;
    const size_t fp = window_samples > 0 ? rds_footprint(c, window_samples) : 0;
    if (s->bytes_in_use - old_fp + fp > s->budget) return GPUD_ERR_BUDGET;
    if (c.d_rdec && window_samples != c.rds_win) {
        s->arena.release(c.d_rdec, c.rdec_bytes);
        c.d_rdec = nullptr;
        c.rdec_bytes = 0;
    }
    if (window_samples > 0) {  // a fresh decoder: zero state and progress
        const RdsLayout l = rds_layout(c, window_samples);
        if (!c.d_rdec) {
            c.d_rdec = s->arena.alloc(l.bytes);
            if (!c.d_rdec) return GPUD_ERR_BUDGET;
            c.rdec_bytes = l.bytes;
        }
        const std::vector<char> zero(l.off_bits, 0);
        const int ue = stage_upload(s, c.d_rdec, zero.data(), zero.size());
        if (ue) return ue;
    }
    s->n_rds_on += (window_samples > 0 ? 1 : 0) - (c.rds_win > 0 ? 1 : 0);
    c.rds_win = window_samples;
    c.rwin_sel = 0;
    c.rwin_fill = 0;
    c.r_last = ArgusRdsOut{};
    c.r_last_ticket = s->last_ticket;
    c.footprint += fp - old_fp;
    s->bytes_in_use += fp - old_fp;
    s->epoch++;
    return 0;
}

GPUD_API int gpud_serve_channel_rds(gpud_serve* s, int64_t ticket, int64_t channel_id, int* decoded, int* found,
                                    uint16_t* pi, char* ps9, char* rt65, const uint16_t** groups, int* n_groups)
{
    if (!s) return -1;
    ServeSlot* sl = find_slot(s, ticket);
    if (!sl || sl->busy) return -2;
    for (const ServeOut& o : sl->outs)
        if (o.channel_id == channel_id) {
            if (!o.r_on) return -3;
            ArgusRdsOut r{};
            const bool dec = o.r_index >= 0;
            if (dec) {
                r = ((const ArgusRdsOut*)(sl->h_rtab + sl->off_rout))[o.r_index];
            } else {
                auto it = s->live.find(channel_id);
                if (it != s->live.end()) r = it->second.r_last;
            }
            if (decoded) *decoded = dec ? 1 : 0;
            if (found) *found = dec ? r.found : 0;
            if (pi) *pi = (uint16_t)r.pi;
            if (ps9) argus_rds_text(r.ps, 8, ps9);
            if (rt65) argus_rds_text(r.rt, 64, rt65);
            if (groups)
                *groups = dec ? (const uint16_t*)(sl->h_rtab + sl->off_rgroups) + sl->rdec_groups[(size_t)o.r_index]
                              : nullptr;
            if (n_groups) *n_groups = dec ? r.written : 0;
            return 0;
        }
    return -1;
}

// The fast-convolution path for eligible channels (decimation a multiple of
// 8, D < taps <= 512, full history): on by default.
GPUD_API int gpud_serve_set_fast(gpud_serve* s, int enable)
{
    if (!s) return -1;
    s->fast = enable != 0;
    s->epoch++;
    return 0;
}

// Demodulation in the epilogues of the fast kernel and of the direct form's
// narrow items for the channels whose demodulator reads their baseband (FM,
// SSB, CW; no pre-demod or stereo stage): on by default.  Off, those
// channels' baseband is demodulated by the separate pass as every other
// channel's (same outputs).
GPUD_API int gpud_serve_set_demod_fusion(gpud_serve* s, int enable)
{
    if (!s) return -1;
    s->fuse_demod = enable != 0;
    s->epoch++;
    return 0;
}

GPUD_API int gpud_serve_set_graph(gpud_serve* s, int enable)
{
    if (!s) return -1;
    for (auto& sl : s->slots) slot_wait(s, sl);
    s->graph = enable != 0;
    return 0;
}

GPUD_API int gpud_serve_get_tune(gpud_serve* s, gpud_tune* t)
{
    if (!s || !t) return -1;
    *t = gpud_tune{s->tune.lds_entries, s->tune.narrow_outputs, s->tune.grid, s->tune.fast_grid, 0.0f, 0.0f};
    return 0;
}

static bool valid_tune(const gpud_tune& t)
{
    return t.lds_entries >= 1024 && t.lds_entries <= ARGUS_MAX_LDS_ENTRIES && t.narrow_outputs >= 1 &&
           t.narrow_outputs <= 1024 && t.grid >= ARGUS_XCDS && t.grid <= 65536 && t.grid % ARGUS_XCDS == 0 &&
           t.fast_grid >= 0 && t.fast_grid <= 65536 && t.fast_grid % ARGUS_XCDS == 0;
}

GPUD_API int gpud_serve_set_tune(gpud_serve* s, const gpud_tune* t)
{
    if (!s || !t || !valid_tune(*t)) return -1;
    s->tune = ArgusTune{t->lds_entries, t->narrow_outputs, t->grid, t->fast_grid ? t->fast_grid : s->tune.fast_grid};
    s->lds_hw = 0;
    s->epoch++;
    return 0;
}

// Time the fused extraction launch of the live channel set on the block at
// iq_dev (n samples, device memory) for tile and launch shapes around the
// current one, one parameter at a time (grid, then LDS budget per dense tile,
// then outputs per narrow item; two rounds), and keep the fastest.  Channel state does
// not advance: the launches write a scratch slot's outputs and the inactive
// history halves only.  Blocks in flight are completed first.
// Event-timed fused extraction launches of the live channel set at the
// current tune on the block at iq_dev, using a scratch slot (outputs and the
// inactive history halves only; channel state does not advance): one
// warm-up launch, then reps timed ones.
static int time_extract(gpud_serve* s, ServeSlot& scratch, const float2* iq_dev, int n, int reps, hipEvent_t e0,
                        hipEvent_t e1, float* min_ms, float* avg_ms)
{
    s->epoch++;  // the scratch slot's table is built for this call only
    int rc = build_block(s, scratch, n, iq_dev, false);
    if (rc) return rc;
    rc = argus_extract_allow_lds(scratch.lds);
    if (rc) return rc;
    rc = upload_table(s, scratch);
    if (rc) return rc;
    const ArgusServeHdr* dh = (const ArgusServeHdr*)(scratch.d_tab + scratch.off_hdr);
    // the extraction stage as a block runs it (the fast path's transforms and
    // per-channel pass, then the direct kernel), without the tail update that
    // would advance the stream's state
    auto stage = [&]() -> int {
        int e = argus_launch_fast(&dh->f, &dh->x, scratch.fast_S, s->d_fu, s->d_fX, s->tune.fast_grid, s->compute);
        return e ? e : argus_launch_extract_hdr(&dh->x, scratch.lds, s->tune.grid, s->compute);
    };
    rc = stage();  // warm-up
    if (rc) return rc;
    float bestm = 1e30f, sum = 0.0f;
    for (int r = 0; r < reps; r++) {
        ARGUS_CHECK(hipEventRecord(e0, s->compute));
        rc = stage();
        if (rc) return rc;
        ARGUS_CHECK(hipEventRecord(e1, s->compute));
        ARGUS_CHECK(hipEventSynchronize(e1));
        float m = 0.0f;
        ARGUS_CHECK(hipEventElapsedTime(&m, e0, e1));
        bestm = std::min(bestm, m);
        sum += m;
    }
    if (min_ms) *min_ms = bestm;
    if (avg_ms) *avg_ms = sum / (float)reps;
    return 0;
}

GPUD_API int gpud_serve_extract_time(gpud_serve* s, const gpud_float2* iq_dev, int n, int reps, float* avg_ms,
                                     float* min_ms)
{
    if (!s || !iq_dev || n <= 0 || n > s->block_samples || reps <= 0) return -1;
    for (auto& sl : s->slots) {
        int err = slot_wait(s, sl);
        if (err) return err;
    }
    ServeSlot scratch;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int err = (int)hipEventCreate(&e0);
    if (!err) err = (int)hipEventCreate(&e1);
    if (!err) err = time_extract(s, scratch, (const float2*)iq_dev, n, reps, e0, e1, min_ms, avg_ms);
    hipStreamSynchronize(s->compute);
    slot_release(scratch);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    return err;
}

GPUD_API int gpud_serve_autotune(gpud_serve* s, const gpud_float2* iq_dev, int n, int reps, gpud_tune* best)
{
    if (!s || !iq_dev || n <= 0 || n > s->block_samples || reps <= 0) return -1;
    for (auto& sl : s->slots) {
        int err = slot_wait(s, sl);
        if (err) return err;
    }
    ServeSlot scratch;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int err = (int)hipEventCreate(&e0);
    if (!err) err = (int)hipEventCreate(&e1);
    auto eval = [&](const ArgusTune& t, float* ms) -> int {
        s->tune = t;
        return time_extract(s, scratch, (const float2*)iq_dev, n, reps, e0, e1, ms, nullptr);
    };
    ArgusTune cur = s->tune;
    float cur_ms = 0.0f;
    if (!err) err = eval(cur, &cur_ms);
    // coordinate descent, twice over (grid and LDS budget interact: smaller
    // tiles mean more workgroups resident per CU); the fast form's grid last
    const int grids[] = {1024, 2048, 4096, 6144, 8192, 12288};
    const int ldss[] = {2048, 3072, 4096, 6144, 8192, 12288};
    const int narrows[] = {16, 32, 64};
    const int fgrids[] = {512, 1024, 2048, 4096};
    for (int pass = 0; pass < 8 && !err; pass++) {
        const int axis = pass % 4;
        const int* vals = axis == 0 ? grids : axis == 1 ? ldss : axis == 2 ? narrows : fgrids;
        const int nv = axis == 0 || axis == 1 ? 6 : axis == 2 ? 3 : 4;
        for (int i = 0; i < nv && !err; i++) {
            ArgusTune t = cur;
            int& field = axis == 0 ? t.grid : axis == 1 ? t.lds_entries : axis == 2 ? t.narrow_outputs : t.fast_grid;
            if (field == vals[i]) continue;
            field = vals[i];
            float ms = 0.0f;
            err = eval(t, &ms);
            if (!err && ms < cur_ms) {
                cur = t;
                cur_ms = ms;
            }
        }
    }
    s->tune = cur;
    s->lds_hw = 0;
    s->epoch++;
    hipStreamSynchronize(s->compute);
    slot_release(scratch);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (err) return err;
    if (best) *best = gpud_tune{cur.lds_entries, cur.narrow_outputs, cur.grid, cur.fast_grid, cur_ms, 0.0f};
    return 0;
}

// RDS baseband and stereo lock of a WFM stereo channel in a completed block.
// -1 if the channel was not part of the block, -3 if it is not a stereo
// channel.
GPUD_API int gpud_serve_channel_wfm(gpud_serve* s, int64_t ticket, int64_t channel_id, const gpud_float2** rds,
                                    int* n_rds, int* rds_rate, int* locked, int* stereo)
{
    if (!s) return -1;
    ServeSlot* sl = find_slot(s, ticket);
    if (!sl || sl->busy) return -2;
    for (const ServeOut& o : sl->outs)
        if (o.channel_id == channel_id) {
            if (o.wfm_index < 0) return -3;
            if (rds) *rds = (const gpud_float2*)(sl->h_rds + o.rds_off);
            if (n_rds) *n_rds = o.n_rds;
            if (rds_rate) *rds_rate = o.rds_rate;
            const ArgusWfmBlock& b = sl->h_wblk[o.wfm_index];
            if (locked) *locked = o.n_mono > 0 ? b.locked : 0;
            if (stereo) *stereo = b.stereo;
            return 0;
        }
    return -1;
}

}  // extern "C"
