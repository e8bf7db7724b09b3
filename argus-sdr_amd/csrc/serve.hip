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

int argus_launch_extract_batch(const ArgusChanDesc* d_chans, const ArgusWorkItem* d_items, int n_items,
                               size_t lds_bytes, const float2* in, int n_new, hipStream_t stream);
int argus_fill_chan(ArgusChanDesc* ch, const float2* hist_rd, int H, float2* hist_wr, int keep, const float* taps,
                    float2* out, int D, int T, int phase_count, int n_new, double phase_start, double phase_inc);
size_t argus_item_lds_bytes(const ArgusChanDesc& ch);
void argus_append_items(const ArgusChanDesc& ch, int chan_index, int n_new, std::vector<ArgusWorkItem>& items,
                        std::vector<int>& pos);
void argus_order_items(std::vector<ArgusWorkItem>& items, const std::vector<int>& pos, int n_new);
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

bool valid_mode(int m) { return m >= GPUD_DEMOD_NONE && m <= GPUD_DEMOD_CW; }

}  // namespace

// Device bytes a channel holds while open (taps, ping-pong history, FM carry)
// plus its share of the per-slot output buffers.
static size_t channel_footprint(const gpud_serve* s, int D, int T)
{
    const size_t L = (size_t)(T + D - 1) / D;
    const size_t taps = (size_t)D * L * sizeof(float);
    const size_t hist = 2 * (size_t)std::max(1, T - 1) * sizeof(float2);
    const size_t carry = 2 * sizeof(float2);
    const size_t per_block = ((size_t)s->block_samples / D + 2) * (sizeof(float2) + sizeof(float));
    return taps + hist + carry + per_block * ARGUS_SERVE_SLOTS;
}

static void release_channel(ServeChannel& c)
{
    if (c.d_taps) hipFree(c.d_taps);
    if (c.d_hist[0]) hipFree(c.d_hist[0]);
    if (c.d_hist[1]) hipFree(c.d_hist[1]);
    if (c.d_carry) hipFree(c.d_carry);
    c.d_taps = nullptr;
    c.d_hist[0] = c.d_hist[1] = nullptr;
    c.d_carry = nullptr;
}

// Release channels whose last block has completed.
static void reap(gpud_serve* s)
{
    for (size_t i = 0; i < s->closed.size();) {
        ServeChannel& c = s->closed[i];
        if (c.last_ticket <= s->completed_ticket) {
            s->bytes_in_use -= c.footprint;
            release_channel(c);
            s->closed[i] = s->closed.back();
            s->closed.pop_back();
        } else {
            i++;
        }
    }
}

static int slot_wait(gpud_serve* s, ServeSlot& sl)
{
    if (!sl.busy) return 0;
    ARGUS_CHECK(hipEventSynchronize(sl.ev_done));
    sl.busy = false;
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
        ARGUS_CHECK(dev_alloc(&sl.d_in, (size_t)block_samples));
        ARGUS_CHECK(host_alloc(&sl.h_in, (size_t)block_samples));
    }
    *out = s;
    return 0;
}

GPUD_API int gpud_serve_destroy(gpud_serve* s)
{
    if (!s) return 0;
    hipDeviceSynchronize();
    for (auto& kv : s->live) release_channel(kv.second);
    for (auto& c : s->closed) release_channel(c);
    for (int i = 0; i < ARGUS_SERVE_SLOTS; i++) {
        ServeSlot& sl = s->slots[i];
        hipEventDestroy(sl.ev_in);
        hipEventDestroy(sl.ev_kern);
        hipEventDestroy(sl.ev_done);
        hipEventDestroy(sl.ev_start);
        hipEventDestroy(sl.ev_kstart);
        hipFree(sl.d_in);
        hipHostFree(sl.h_in);
        hipFree(sl.d_iq);
        hipFree(sl.d_audio);
        hipHostFree(sl.h_iq);
        hipHostFree(sl.h_audio);
        hipFree(sl.d_chans);
        hipFree(sl.d_items);
        hipFree(sl.d_ddesc);
        hipFree(sl.d_ditems);
        hipHostFree(sl.h_chans);
        hipHostFree(sl.h_items);
        hipHostFree(sl.h_ddesc);
        hipHostFree(sl.h_ditems);
    }
    hipStreamDestroy(s->compute);
    hipStreamDestroy(s->copy);
    delete s;
    return 0;
}

// Admit a channel.  Returns 0, -1 on bad arguments or a non-integer
// decimation (cpu_oracle.go:155-163), GPUD_ERR_DUPLICATE if the id is live,
// GPUD_ERR_BUDGET if its device footprint does not fit the remaining budget.
GPUD_API int gpud_serve_open(gpud_serve* s, const gpud_channel_spec* spec)
{
    if (!s || !spec || !valid_mode(spec->demod) || (spec->window != 0 && spec->window != 1)) return -1;
    if (spec->out_rate <= 0 || s->sample_rate % spec->out_rate != 0) return -1;
    if (s->live.count(spec->channel_id)) return GPUD_ERR_DUPLICATE;
    const int D = s->sample_rate / spec->out_rate;
    const int T = spec->num_taps > 0 ? spec->num_taps : 101;
    if (T > ARGUS_MAX_TAPS) return -1;
    const size_t fp = channel_footprint(s, D, T);
    reap(s);
    if (s->bytes_in_use + fp > s->budget) return GPUD_ERR_BUDGET;

    ServeChannel c;
    c.spec = *spec;
    c.decim = D;
    c.T = T;
    c.footprint = fp;
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
    if (dev_alloc(&c.d_taps, poly.size()) != hipSuccess || dev_alloc(&c.d_hist[0], (size_t)std::max(1, T - 1)) ||
        dev_alloc(&c.d_hist[1], (size_t)std::max(1, T - 1)) || dev_alloc(&c.d_carry, 2)) {
        release_channel(c);
        return (int)hipErrorOutOfMemory;
    }
    hipError_t e = hipMemcpy(c.d_taps, poly.data(), poly.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        release_channel(c);
        return (int)e;
    }
    s->bytes_in_use += fp;
    s->live.emplace(spec->channel_id, c);
    return 0;
}

// Retire a channel: it is absent from the next submitted block; its memory is
// released when the blocks already in flight that use it have completed.
GPUD_API int gpud_serve_close(gpud_serve* s, int64_t channel_id)
{
    if (!s) return -1;
    auto it = s->live.find(channel_id);
    if (it == s->live.end()) return -1;
    s->closed.push_back(it->second);
    s->live.erase(it);
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

// Enqueue one block: upload (unless device-resident), extract, demodulate,
// download.  Returns immediately; *ticket identifies the block.
static int serve_submit(gpud_serve* s, const gpud_float2* iq, int n, bool on_device, int64_t* ticket)
{
    if (!s || n < 0 || n > s->block_samples || (n > 0 && !iq)) return -1;
    ServeSlot& sl = s->slots[s->next_slot];
    int err = slot_wait(s, sl);  // the slot's previous block must be done
    if (err) return err;
    s->next_slot = (s->next_slot + 1) % ARGUS_SERVE_SLOTS;
    sl.ticket = ++s->last_ticket;
    sl.n = n;

    // ---- per-channel descriptors (host) --------------------------------
    const int nch = (int)s->live.size();
    sl.outs.clear();
    sl.outs.reserve((size_t)nch);
    int64_t iq_total = 0, audio_total = 0;
    for (auto& kv : s->live) {
        ServeChannel& c = kv.second;
        ServeOut o;
        o.channel_id = kv.first;
        o.n_iq = (c.phase_count + n) / c.decim;
        o.iq_off = iq_total;
        o.n_audio = argus_demod_count(c.spec.demod, o.n_iq, c.have_carry);
        o.audio_off = audio_total;
        o.rate = c.spec.out_rate;
        iq_total += o.n_iq;
        audio_total += o.n_audio;
        sl.outs.push_back(o);
    }
    ARGUS_CHECK((hipError_t)grow_dev(&sl.d_iq, &sl.iq_cap, (size_t)std::max<int64_t>(iq_total, 1)));
    ARGUS_CHECK((hipError_t)grow_dev(&sl.d_audio, &sl.audio_cap, (size_t)std::max<int64_t>(audio_total, 1)));
    ARGUS_CHECK((hipError_t)grow_host(&sl.h_iq, &sl.h_iq_cap, (size_t)std::max<int64_t>(iq_total, 1)));
    ARGUS_CHECK((hipError_t)grow_host(&sl.h_audio, &sl.h_audio_cap, (size_t)std::max<int64_t>(audio_total, 1)));
    ARGUS_CHECK((hipError_t)grow_host(&sl.h_chans, &sl.h_chans_cap, (size_t)std::max(nch, 1)));
    ARGUS_CHECK((hipError_t)grow_dev(&sl.d_chans, &sl.d_chans_cap, (size_t)std::max(nch, 1)));
    ARGUS_CHECK((hipError_t)grow_host(&sl.h_ddesc, &sl.h_ddesc_cap, (size_t)std::max(nch, 1)));
    ARGUS_CHECK((hipError_t)grow_dev(&sl.d_ddesc, &sl.d_ddesc_cap, (size_t)std::max(nch, 1)));

    s->items.clear();
    s->item_pos.clear();
    s->ditems.clear();
    size_t lds = 0;
    int ci = 0;
    const float2* in = on_device ? (const float2*)iq : sl.d_in;
    for (auto& kv : s->live) {
        ServeChannel& c = kv.second;
        const ServeOut& o = sl.outs[(size_t)ci];
        const int keep = std::min(c.hist_len + n, c.T - 1);
        const double inc = -2.0 * M_PI * c.spec.offset_hz / (double)s->sample_rate;
        ArgusChanDesc& ch = sl.h_chans[ci];
        if (argus_fill_chan(&ch, c.d_hist[c.hist_sel], c.hist_len, c.d_hist[c.hist_sel ^ 1], keep, c.d_taps,
                            sl.d_iq + o.iq_off, c.decim, c.T, c.phase_count, n, c.nco_phase, inc) != 0)
            return -1;
        if (ch.n_out > 0) lds = std::max(lds, argus_item_lds_bytes(ch));
        argus_append_items(ch, ci, n, s->items, s->item_pos);

        ArgusDemodDesc& dd = sl.h_ddesc[ci];
        dd.iq = sl.d_iq + o.iq_off;
        dd.audio = sl.d_audio + o.audio_off;
        dd.carry_rd = c.d_carry + c.carry_sel;
        dd.carry_wr = c.d_carry + (c.carry_sel ^ 1);
        dd.n = o.n_iq;
        dd.mode = c.spec.demod;
        dd.have_carry = c.have_carry ? 1 : 0;
        dd.skip = (dd.mode == ARGUS_DEMOD_NFM || dd.mode == ARGUS_DEMOD_WFM) && !c.have_carry ? 1 : 0;
        const double bfo_hz = dd.mode == ARGUS_DEMOD_LSB ? -700.0 : 700.0;
        const double bfo_inc = 2.0 * M_PI * bfo_hz / (double)c.spec.out_rate;
        dd.bfo0 = argus_rad_to_fix(c.bfo_phase + bfo_inc);
        dd.bfo_inc = argus_rad_to_fix(bfo_inc);
        if (dd.mode != ARGUS_DEMOD_NONE && o.n_iq > 0) {
            if (dd.mode == ARGUS_DEMOD_AM) {
                s->ditems.push_back({ci, 0, o.n_iq, 0});
            } else {
                for (int st = 0; st < o.n_iq; st += ARGUS_DEMOD_TILE)
                    s->ditems.push_back({ci, st, std::min(ARGUS_DEMOD_TILE, o.n_iq - st), 0});
            }
        }
        // carry state forward (the kernels compute with the values above)
        c.phase_count = (c.phase_count + n) % c.decim;
        c.nco_phase = argus_reduce_phase(c.nco_phase + inc * (double)n);
        if (keep > 0) c.hist_sel ^= 1;
        c.hist_len = keep;
        if ((dd.mode == ARGUS_DEMOD_NFM || dd.mode == ARGUS_DEMOD_WFM) && o.n_iq > 0) {
            c.carry_sel ^= 1;
            c.have_carry = true;
        }
        c.bfo_phase = argus_reduce_phase(c.bfo_phase + bfo_inc * (double)o.n_iq);
        c.last_ticket = sl.ticket;
        ci++;
    }
    argus_order_items(s->items, s->item_pos, n);
    const int n_items = (int)s->items.size(), n_ditems = (int)s->ditems.size();
    ARGUS_CHECK((hipError_t)grow_host(&sl.h_items, &sl.h_items_cap, (size_t)std::max(n_items, 1)));
    ARGUS_CHECK((hipError_t)grow_dev(&sl.d_items, &sl.d_items_cap, (size_t)std::max(n_items, 1)));
    ARGUS_CHECK((hipError_t)grow_host(&sl.h_ditems, &sl.h_ditems_cap, (size_t)std::max(n_ditems, 1)));
    ARGUS_CHECK((hipError_t)grow_dev(&sl.d_ditems, &sl.d_ditems_cap, (size_t)std::max(n_ditems, 1)));
    if (n_items) std::memcpy(sl.h_items, s->items.data(), (size_t)n_items * sizeof(ArgusWorkItem));
    if (n_ditems) std::memcpy(sl.h_ditems, s->ditems.data(), (size_t)n_ditems * sizeof(ArgusDemodItem));
    sl.iq_total = iq_total;
    sl.audio_total = audio_total;

    // ---- copy stream: block upload -------------------------------------
    ARGUS_CHECK(hipEventRecord(sl.ev_start, s->copy));
    if (!on_device && n > 0) {
        std::memcpy(sl.h_in, iq, (size_t)n * sizeof(float2));
        ARGUS_CHECK(hipMemcpyAsync(sl.d_in, sl.h_in, (size_t)n * sizeof(float2), hipMemcpyHostToDevice, s->copy));
    }
    ARGUS_CHECK(hipEventRecord(sl.ev_in, s->copy));
    // ---- compute stream: tables, extraction, demodulation ----------------
    ARGUS_CHECK(hipStreamWaitEvent(s->compute, sl.ev_in, 0));
    ARGUS_CHECK(hipMemcpyAsync(sl.d_chans, sl.h_chans, (size_t)nch * sizeof(ArgusChanDesc), hipMemcpyHostToDevice,
                               s->compute));
    ARGUS_CHECK(hipMemcpyAsync(sl.d_items, sl.h_items, (size_t)n_items * sizeof(ArgusWorkItem),
                               hipMemcpyHostToDevice, s->compute));
    ARGUS_CHECK(hipMemcpyAsync(sl.d_ddesc, sl.h_ddesc, (size_t)nch * sizeof(ArgusDemodDesc), hipMemcpyHostToDevice,
                               s->compute));
    ARGUS_CHECK(hipMemcpyAsync(sl.d_ditems, sl.h_ditems, (size_t)n_ditems * sizeof(ArgusDemodItem),
                               hipMemcpyHostToDevice, s->compute));
    ARGUS_CHECK(hipEventRecord(sl.ev_kstart, s->compute));
    err = argus_launch_extract_batch(sl.d_chans, sl.d_items, n_items, lds, in, n, s->compute);
    if (err) return err;
    err = argus_launch_demod(sl.d_ddesc, sl.d_ditems, n_ditems, s->compute);
    if (err) return err;
    ARGUS_CHECK(hipEventRecord(sl.ev_kern, s->compute));
    // ---- copy stream: download ------------------------------------------
    ARGUS_CHECK(hipStreamWaitEvent(s->copy, sl.ev_kern, 0));
    if (s->download) {
        if (iq_total > 0)
            ARGUS_CHECK(hipMemcpyAsync(sl.h_iq, sl.d_iq, (size_t)iq_total * sizeof(float2), hipMemcpyDeviceToHost,
                                       s->copy));
        if (audio_total > 0)
            ARGUS_CHECK(hipMemcpyAsync(sl.h_audio, sl.d_audio, (size_t)audio_total * sizeof(float),
                                       hipMemcpyDeviceToHost, s->copy));
    }
    ARGUS_CHECK(hipEventRecord(sl.ev_done, s->copy));
    sl.busy = true;
    if (ticket) *ticket = sl.ticket;
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
    s->download = enable != 0;
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

}  // extern "C"
