// rds_runtime.hip -- the daemon's RDS caller flow on the device: updateRDS
// (cmd/sdrd/pipeline_runtime.go:1044-1156) with extractRDS
// (cmd/sdrd/helpers.go:117-160) and the fallback chain it takes when that
// extraction fails (pipeline_runtime.go:1104-1124), the prune of
// maintenance (1158-1176), StereoPilotPresent (internal/classifier/
// pll.go:82-131), RDSBasebandComplexInto (internal/demod/fm.go:234-271) and
// rds.Decoder.Decode (internal/rds/rds.go:78-264).
//
// Bookkeeping per signal, in call order (pipeline_runtime.go:1048-1080):
// signals with ID 0 are skipped; the state of the ID is created on first
// sight (a fresh decoder) and stamped last_seen = now; states not seen for
// more than 30 s are dropped; a state whose last decode is 4 s or more ago
// takes its decode turn (last_decode = now).  States live until pruned or
// until maintain() is called without their ID.
//
// Where the reference starts one goroutine per due station, each running
// extraction, pilot test, RDS baseband and decoder in turn on the last 4 s
// of the recorder's ring, one update call here takes every station due
// this tick through batched device passes over the one window:
//   1. extraction, every due station a job of one fused extraction launch:
//      route A (extractRDS -> ShiftFilterDecimateBatch): D = fs / 250000
//      (at least 1), 101-tap LowpassFIR at 100 kHz (bw 200 kHz), shift from
//      phase 0, keep inputs 0, D, 2D, ... (n / D outputs), rate fs / D --
//      taken when fs / D divides fs (ExactIntegerDecimation); route B
//      otherwise (the fallback chain): D1 = fs / 1 MHz, shift with the
//      pre-incremented phase of FreqShift, 51-tap low-pass at 0.4 (fs/D1),
//      keep every D1-th (ceil(n / D1) outputs), then D2 = (fs/D1) / 250000,
//      101-tap low-pass at 0.4 (fs/D1/D2) and every D2-th;
//   2. the stereo pilot test on each station's decimated baseband
//      (pilot.hip, one batched call);
//   3. the RDS baseband of each station (argus_rds_bb_kernel below): FM
//      discriminator (n - 1 outputs, FP64 atan2), mix by +57 kHz with the
//      pre-incremented phase, 101-tap complex low-pass at 7.5 kHz from zero
//      history, keep every floor(rate / 19000)-th, rate / that;
//   4. one call of the device RDS decoders (rds_batch.hip), one decoder per
//      live state.
// The call returns when the results are on the host.  The reference's
// goroutine publishes its result on a later frame; here it is in place when
// update returns.  A window shorter than one second (or rate <= 0) skips
// the decodes of the tick after their turn was taken, as the reference's
// goroutine returns early.
#include "common.hpp"
#include "runner.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <vector>

int argus_lowpass_fir(double cutoff_hz, int sample_rate, int taps, double* out, bool hamming);

namespace {

constexpr double kDecodeEvery = 4.0;  // pipeline_runtime.go:1069
constexpr double kForgetAfter = 30.0; // pipeline_runtime.go:1064
constexpr int kRdsTaps = 101;
constexpr int kBbOut = 128;           // RDS outputs per work item
constexpr int kBbThreads = 256;
constexpr int kBbMaxSpan = 8192;      // LDS samples of one item: (kBbOut - 1) * D + kRdsTaps

struct RdsBbDesc {
    const float2* x;   // decimated baseband, n_disc + 1 samples
    float2* out;       // n_out RDS baseband samples
    uint64_t mix_inc;  // +57 kHz per sample (turns, 2^64 == 1)
    int n_disc;        // discriminator outputs
    int D;             // RDS decimation
    int n_out;         // ceil(n_disc / D)
    int pad;
};

struct RdsBbItem {
    int st;  // descriptor
    int r0;  // first output of the item
};

// One item: the discriminator and mixer over the inputs its outputs read
// (staged once in LDS), then the low-pass at the kept positions.
__global__ __launch_bounds__(kBbThreads) void argus_rds_bb_kernel(const RdsBbDesc* __restrict__ descs,
                                                                  const RdsBbItem* __restrict__ items, int n_items,
                                                                  const float* __restrict__ taps)
{
    extern __shared__ float2 ms[];
    __shared__ float h[kRdsTaps];
    for (int k = threadIdx.x; k < kRdsTaps; k += kBbThreads) h[k] = taps[k];
    for (int w = blockIdx.x; w < n_items; w += gridDim.x) {
        const RdsBbItem it = items[w];
        const RdsBbDesc d = descs[it.st];
        const int cnt = min(kBbOut, d.n_out - it.r0);
        const int lo = it.r0 * d.D - (kRdsTaps - 1);
        const int span = (cnt - 1) * d.D + kRdsTaps;
        for (int j = threadIdx.x; j < span; j += kBbThreads) {
            const int idx = lo + j;  // < n_disc: the last kept position is (n_out - 1) * D <= n_disc - 1
            float2 v = make_float2(0.0f, 0.0f);
            if (idx >= 0) {
                const float2 p = d.x[idx], c = d.x[idx + 1];
                const double num = (double)p.x * (double)c.y - (double)p.y * (double)c.x;
                const double den = (double)p.x * (double)c.x + (double)p.y * (double)c.y;
                const float b = (float)atan2(num, den);
                float s, co;
                argus_nco(d.mix_inc * (uint64_t)(idx + 1), s, co);  // FreqShift pre-increments its phase
                v = make_float2(b * co, b * s);
            }
            ms[j] = v;
        }
        __syncthreads();
        for (int r = threadIdx.x; r < cnt; r += kBbThreads) {
            const float2* mp = ms + r * d.D + (kRdsTaps - 1);
            float ar = 0.0f, ai = 0.0f;
            for (int k = 0; k < kRdsTaps; k++) {
                const float2 v = mp[-k];
                ar = fmaf(h[k], v.x, ar);
                ai = fmaf(h[k], v.y, ai);
            }
            d.out[it.r0 + r] = make_float2(ar, ai);
        }
        __syncthreads();
    }
}

template <typename T>
int grow(T** p, size_t* cap, size_t need)
{
    if (*p && *cap >= need) return 0;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    const size_t n = std::max<size_t>(need + need / 4, 256);
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return (int)e;
    *cap = n;
    return 0;
}

void trim_space(const char* in, char* out9)
{
    int a = 0, b = (int)std::strlen(in);
    // the decoder's PS holds printable ASCII only (rds.go:600-605), so
    // strings.TrimSpace reduces to spaces
    while (a < b && in[a] == ' ') a++;
    while (b > a && in[b - 1] == ' ') b--;
    const int m = std::min(b - a, 8);
    std::memcpy(out9, in + a, (size_t)m);
    out9[m] = 0;
}

}  // namespace

struct RdsRtState {
    double last_seen = 0.0;
    double last_decode = -std::numeric_limits<double>::infinity();  // never: the first sight is due
    int slot = -1;
    bool fresh = true;       // the slot's decoder still needs its reset
    int stereo = 0;
    bool has_result = false; // a decode has reported a PS
    uint16_t pi = 0;
    char ps[9] = {0};
    char rt[65] = {0};
    // the last decode's device outputs (valid until the next decoding update)
    int64_t base_off = -1, base_n = 0, rds_off = -1, rds_n = 0;
    int base_rate = 0, rds_rate = 0;
    int64_t tick = -1;
};

struct gpud_rds_runtime {
    int max_stations = 0;
    std::unordered_map<int64_t, RdsRtState> states;
    std::vector<int> free_slots;
    int64_t tick = 0;  // decoding updates so far
    // device side, created on the first decode
    hipStream_t stream = nullptr;
    gpud_rds_batch* dec = nullptr;
    float2* d_win = nullptr;
    size_t win_cap = 0;
    float2* d_mid = nullptr;
    size_t mid_cap = 0;
    float2* d_base = nullptr;
    size_t base_cap = 0;
    float2* d_rds = nullptr;
    size_t rds_cap = 0;
    float* d_taps = nullptr;
    size_t taps_cap = 0;
    RdsBbDesc* d_desc = nullptr;
    size_t desc_cap = 0;
    RdsBbItem* d_items = nullptr;
    size_t items_cap = 0;
    std::vector<ArgusLegacyScratch*> scratch;  // one per launch of a call
    ArgusTune tune = argus_default_tune();
};

namespace {

void release_state(gpud_rds_runtime* rt, RdsRtState& s)
{
    if (s.slot >= 0) rt->free_slots.push_back(s.slot);
    s.slot = -1;
}

int device_init(gpud_rds_runtime* rt)
{
    if (rt->stream) return 0;
    ARGUS_CHECK(hipStreamCreateWithFlags(&rt->stream, hipStreamNonBlocking));
    int e = gpud_rds_batch_create(&rt->dec, rt->max_stations);
    return e;
}

ArgusLegacyScratch*& scratch_at(gpud_rds_runtime* rt, size_t i)
{
    if (rt->scratch.size() <= i) rt->scratch.resize(i + 1, nullptr);
    return rt->scratch[i];
}

struct Due {
    int64_t id;
    double offset_hz;
};

// Steps 1-4 for the due stations over the window.
int decode_tick(gpud_rds_runtime* rt, const std::vector<Due>& due, const gpud_float2* window, int64_t n_window,
                int window_on_device, int rate, double window_center_hz)
{
    int e = device_init(rt);
    if (e) return e;
    if (n_window > (int64_t)0x7fffffff) return -2;
    const int n = (int)n_window;
    const int nd = (int)due.size();
    hipStream_t st = rt->stream;
    const float2* d_in = (const float2*)window;
    if (!window_on_device) {
        e = grow(&rt->d_win, &rt->win_cap, (size_t)n);
        if (e) return e;
        ARGUS_CHECK(hipMemcpyAsync(rt->d_win, window, (size_t)n * sizeof(float2), hipMemcpyHostToDevice, st));
        d_in = rt->d_win;
    }
    // 1. extraction
    const int dA = std::max(1, rate / 250000);
    const int out_rate = rate / dA;
    const bool route_a = out_rate > 0 && rate % out_rate == 0;
    int base_rate, n_base;
    std::vector<double> b64(kRdsTaps + 2);
    std::vector<std::vector<float>> taps;
    auto design = [&](double cutoff, int fs, int T) {
        const int m = argus_lowpass_fir(cutoff, fs, T, b64.data(), false);
        taps.emplace_back(b64.begin(), b64.begin() + m);
        return (int)taps.size() - 1;
    };
    if (route_a) {
        base_rate = rate / dA;
        n_base = n / dA;
        e = grow(&rt->d_base, &rt->base_cap, (size_t)nd * (size_t)std::max(n_base, 1));
        if (e) return e;
        const int ti = design(100000.0, rate, 101);  // max(bw / 2, 200) with bw 200 kHz
        std::vector<ArgusLegacyJob> jobs((size_t)nd);
        for (int i = 0; i < nd; i++) {
            const double inc = -2.0 * M_PI * due[(size_t)i].offset_hz / (double)rate;
            jobs[(size_t)i] = ArgusLegacyJob{taps[(size_t)ti].data(), (int)taps[(size_t)ti].size(), dA, 0.0, inc,
                                             n_base, rt->d_base + (size_t)i * n_base};
        }
        e = argus_legacy_launch(scratch_at(rt, 0), st, rt->tune, d_in, n, jobs);
        if (e) return e;
    } else {
        const int d1 = std::max(1, rate / 1000000);
        const int rate1 = rate / d1;
        const int d2 = std::max(1, rate1 / 250000);
        const int n1 = (n + d1 - 1) / d1;
        n_base = (n1 + d2 - 1) / d2;
        base_rate = rate1 / d2;
        e = grow(&rt->d_mid, &rt->mid_cap, (size_t)nd * (size_t)std::max(n1, 1));
        if (!e) e = grow(&rt->d_base, &rt->base_cap, (size_t)nd * (size_t)std::max(n_base, 1));
        if (e) return e;
        const int t1 = design((double)(rate / d1) / 2.0 * 0.8, rate, 51);
        const int t2 = design((double)(rate1 / d2) / 2.0 * 0.8, rate1, 101);
        std::vector<ArgusLegacyJob> jobs((size_t)nd);
        for (int i = 0; i < nd; i++) {
            const double inc = -2.0 * M_PI * due[(size_t)i].offset_hz / (double)rate;
            jobs[(size_t)i] = ArgusLegacyJob{taps[(size_t)t1].data(), (int)taps[(size_t)t1].size(), d1, inc, inc, n1,
                                             rt->d_mid + (size_t)i * n1};
        }
        e = argus_legacy_launch(scratch_at(rt, 0), st, rt->tune, d_in, n, jobs);
        if (e) return e;
        // stage 2: each station filters its own stage-1 output (one launch each)
        for (int i = 0; i < nd; i++) {
            std::vector<ArgusLegacyJob> j2{ArgusLegacyJob{taps[(size_t)t2].data(), (int)taps[(size_t)t2].size(), d2,
                                                          0.0, 0.0, n_base, rt->d_base + (size_t)i * n_base}};
            e = argus_legacy_launch(scratch_at(rt, (size_t)i + 1), st, rt->tune, rt->d_mid + (size_t)i * n1, n1, j2);
            if (e) return e;
        }
    }
    // 2. stereo pilot on each decimated baseband
    std::vector<int64_t> offs((size_t)nd + 1);
    std::vector<int32_t> rates((size_t)nd, base_rate), present((size_t)nd, 0);
    for (int i = 0; i <= nd; i++) offs[(size_t)i] = (int64_t)i * n_base;
    e = gpud_stereo_pilot_present((const gpud_float2*)rt->d_base, offs.data(), rates.data(), nd, present.data(),
                                  nullptr, nullptr, nullptr, (gpud_stream_handle)st);
    if (e) return e;
    // 3. RDS baseband
    const int n_disc = std::max(0, n_base - 1);
    const int rD = std::max(1, base_rate / 19000);
    const int n_rds = (n_disc + rD - 1) / rD;
    const int rds_rate = base_rate / rD;
    if ((kBbOut - 1) * rD + kRdsTaps > kBbMaxSpan) return -5;
    if (n_rds > 0) {
        const int tr = design(7500.0, base_rate, kRdsTaps);
        e = grow(&rt->d_rds, &rt->rds_cap, (size_t)nd * n_rds);
        if (!e) e = grow(&rt->d_taps, &rt->taps_cap, (size_t)kRdsTaps);
        if (!e) e = grow(&rt->d_desc, &rt->desc_cap, (size_t)nd);
        const int per = (n_rds + kBbOut - 1) / kBbOut;
        if (!e) e = grow(&rt->d_items, &rt->items_cap, (size_t)nd * per);
        if (e) return e;
        std::vector<RdsBbDesc> desc((size_t)nd);
        std::vector<RdsBbItem> items;
        items.reserve((size_t)nd * per);
        const uint64_t mix = argus_rad_to_fix(2.0 * M_PI * 57000.0 / (double)base_rate);
        for (int i = 0; i < nd; i++) {
            desc[(size_t)i] = RdsBbDesc{rt->d_base + (size_t)i * n_base, rt->d_rds + (size_t)i * n_rds, mix, n_disc, rD,
                                        n_rds, 0};
            for (int r = 0; r < n_rds; r += kBbOut) items.push_back(RdsBbItem{i, r});
        }
        std::vector<float> hf(kRdsTaps, 0.0f);
        for (int k = 0; k < (int)taps[(size_t)tr].size() && k < kRdsTaps; k++) hf[(size_t)k] = taps[(size_t)tr][(size_t)k];
        // the pilot call above has drained the stream, so these synchronous copies do not stall queued work
        ARGUS_CHECK(hipMemcpy(rt->d_taps, hf.data(), kRdsTaps * sizeof(float), hipMemcpyHostToDevice));
        ARGUS_CHECK(hipMemcpy(rt->d_desc, desc.data(), desc.size() * sizeof(RdsBbDesc), hipMemcpyHostToDevice));
        ARGUS_CHECK(hipMemcpy(rt->d_items, items.data(), items.size() * sizeof(RdsBbItem), hipMemcpyHostToDevice));
        const size_t lds = (size_t)((kBbOut - 1) * rD + kRdsTaps) * sizeof(float2);
        const int grid = (int)std::min<size_t>(items.size(), 4096);
        argus_clear_error();
        hipLaunchKernelGGL(argus_rds_bb_kernel, dim3(grid), dim3(kBbThreads), lds, st, (const RdsBbDesc*)rt->d_desc,
                           (const RdsBbItem*)rt->d_items, (int)items.size(), (const float*)rt->d_taps);
        ARGUS_CHECK(hipGetLastError());
    }
    // 4. the decoders
    std::vector<int32_t> slots, rds_rates, found;
    std::vector<int64_t> rofs{0};
    for (int i = 0; i < nd; i++) {
        RdsRtState& s = rt->states[due[(size_t)i].id];
        s.stereo = present[(size_t)i] ? 1 : 0;
        s.base_off = (int64_t)i * n_base;
        s.base_n = n_base;
        s.base_rate = base_rate;
        s.rds_off = (int64_t)i * n_rds;
        s.rds_n = n_rds;
        s.rds_rate = rds_rate;
        s.tick = rt->tick;
        if (n_rds == 0) continue;  // empty RDS baseband: no Decode call
        if (s.fresh) {
            e = gpud_rds_batch_reset(rt->dec, s.slot);
            if (e) return e;
            s.fresh = false;
        }
        slots.push_back(s.slot);
        rds_rates.push_back(rds_rate);
        rofs.push_back(rofs.back() + n_rds);
    }
    if (!slots.empty()) {
        found.assign(slots.size(), 0);
        // the stations' RDS baseband is contiguous in due order (offsets i * n_rds)
        e = gpud_rds_batch_decode_device(rt->dec, (int)slots.size(), slots.data(), (const gpud_float2*)rt->d_rds,
                                         rofs.data(), rds_rates.data(), found.data(), (gpud_stream_handle)st);
        if (e) return e;
        for (int i = 0; i < nd; i++) {
            RdsRtState& s = rt->states[due[(size_t)i].id];
            uint16_t pi = 0;
            char ps[9] = {0}, rtx[65] = {0};
            e = gpud_rds_batch_state(rt->dec, s.slot, &pi, ps, rtx);
            if (e) return e;
            if (ps[0]) {  // pipeline_runtime.go:1136-1138: a result is kept once it carries a PS
                s.has_result = true;
                s.pi = pi;
                std::memcpy(s.ps, ps, sizeof(ps));
                std::memcpy(s.rt, rtx, sizeof(rtx));
            }
        }
    }
    ARGUS_CHECK(hipStreamSynchronize(st));
    rt->tick++;
    return 0;
}

}  // namespace

extern "C" {

GPUD_API int gpud_rds_runtime_create(int max_stations, gpud_rds_runtime** out)
{
    if (!out || max_stations <= 0 || max_stations > 65535) return -1;
    auto* rt = new gpud_rds_runtime();
    rt->max_stations = max_stations;
    for (int s = max_stations - 1; s >= 0; s--) rt->free_slots.push_back(s);
    *out = rt;
    return 0;
}

GPUD_API int gpud_rds_runtime_destroy(gpud_rds_runtime* rt)
{
    if (!rt) return 0;
    if (rt->stream) (void)hipStreamSynchronize(rt->stream);
    if (rt->dec) gpud_rds_batch_destroy(rt->dec);
    for (ArgusLegacyScratch* sc : rt->scratch) argus_legacy_release(sc);
    (void)hipFree(rt->d_win);
    (void)hipFree(rt->d_mid);
    (void)hipFree(rt->d_base);
    (void)hipFree(rt->d_rds);
    (void)hipFree(rt->d_taps);
    (void)hipFree(rt->d_desc);
    (void)hipFree(rt->d_items);
    if (rt->stream) (void)hipStreamDestroy(rt->stream);
    delete rt;
    return 0;
}

GPUD_API int gpud_rds_runtime_update(gpud_rds_runtime* rt, double now_s, const int64_t* ids, const double* center_hz,
                                     int n_signals, const gpud_float2* window, int64_t n_window, int window_on_device,
                                     int window_rate, double window_center_hz, int32_t* stereo, char* station9,
                                     int32_t* turn)
{
    if (!rt || n_signals < 0 || (n_signals > 0 && (!ids || !center_hz)) || n_window < 0 ||
        (n_window > 0 && !window))
        return -1;
    std::vector<Due> due;
    bool pruned = false;
    for (int i = 0; i < n_signals; i++) {
        if (turn) turn[i] = 0;
        const int64_t id = ids[i];
        if (id == 0) {  // pipeline_runtime.go:1054-1056
            if (turn) turn[i] = -1;
            continue;
        }
        auto it = rt->states.find(id);
        if (it == rt->states.end()) {
            if (rt->free_slots.empty()) return -8;  // more live states than decoder slots
            RdsRtState s;
            s.slot = rt->free_slots.back();
            rt->free_slots.pop_back();
            it = rt->states.emplace(id, s).first;
        }
        it->second.last_seen = now_s;
        if (!pruned) {
            // every state not seen this tick has the same age at each signal's
            // call, so one sweep after the first touch drops what each call would
            for (auto p = rt->states.begin(); p != rt->states.end();) {
                if (now_s - p->second.last_seen > kForgetAfter) {
                    release_state(rt, p->second);
                    p = rt->states.erase(p);
                } else {
                    ++p;
                }
            }
            pruned = true;
        }
        RdsRtState& s = rt->states[id];
        if (now_s - s.last_decode >= kDecodeEvery) {
            s.last_decode = now_s;
            due.push_back(Due{id, center_hz[i] - window_center_hz});
            if (turn) turn[i] = 1;
        }
    }
    if (!due.empty() && window_rate > 0 && n_window >= window_rate) {
        int e = decode_tick(rt, due, window, n_window, window_on_device, window_rate, window_center_hz);
        if (e) return e;
    }
    // pipeline_runtime.go:1142-1155: what the signal reports this frame
    for (int i = 0; i < n_signals; i++) {
        auto it = ids[i] ? rt->states.find(ids[i]) : rt->states.end();
        if (stereo) stereo[i] = it != rt->states.end() ? it->second.stereo : 0;
        if (station9) {
            char* o = station9 + (size_t)i * 9;
            o[0] = 0;
            if (it != rt->states.end() && it->second.has_result) trim_space(it->second.ps, o);
        }
    }
    return 0;
}

GPUD_API int gpud_rds_runtime_maintain(gpud_rds_runtime* rt, const int64_t* active_ids, int n_active)
{
    if (!rt || n_active < 0 || (n_active > 0 && !active_ids)) return -1;
    std::vector<int64_t> keep(active_ids, active_ids + n_active);
    std::sort(keep.begin(), keep.end());
    for (auto p = rt->states.begin(); p != rt->states.end();) {
        if (!std::binary_search(keep.begin(), keep.end(), p->first)) {
            release_state(rt, p->second);
            p = rt->states.erase(p);
        } else {
            ++p;
        }
    }
    return 0;
}

GPUD_API int gpud_rds_runtime_num_states(const gpud_rds_runtime* rt) { return rt ? (int)rt->states.size() : -1; }

GPUD_API int gpud_rds_runtime_state(const gpud_rds_runtime* rt, int64_t id, double* last_seen, double* last_decode,
                                    int32_t* stereo, uint16_t* pi, char* ps9, char* rt65)
{
    if (!rt) return -1;
    auto it = rt->states.find(id);
    if (it == rt->states.end()) return -1;
    const RdsRtState& s = it->second;
    if (last_seen) *last_seen = s.last_seen;
    if (last_decode) *last_decode = s.last_decode;
    if (stereo) *stereo = s.stereo;
    if (pi) *pi = s.has_result ? s.pi : 0;
    if (ps9) std::memcpy(ps9, s.has_result ? s.ps : "", s.has_result ? 9 : 1);
    if (rt65) std::memcpy(rt65, s.has_result ? s.rt : "", s.has_result ? 65 : 1);
    return s.has_result ? 1 : 0;
}

GPUD_API int gpud_rds_runtime_last(gpud_rds_runtime* rt, int64_t id, gpud_float2* base, int64_t base_cap,
                                   int64_t* n_base, int32_t* base_rate, gpud_float2* rds, int64_t rds_cap,
                                   int64_t* n_rds, int32_t* rds_rate)
{
    if (!rt) return -1;
    auto it = rt->states.find(id);
    if (it == rt->states.end()) return -1;
    const RdsRtState& s = it->second;
    if (s.tick < 0 || s.tick != rt->tick - 1) return -2;  // not decoded by the last decoding update
    if (n_base) *n_base = s.base_n;
    if (base_rate) *base_rate = s.base_rate;
    if (n_rds) *n_rds = s.rds_n;
    if (rds_rate) *rds_rate = s.rds_rate;
    if (base && base_cap > 0 && s.base_n > 0)
        ARGUS_CHECK(hipMemcpy(base, rt->d_base + s.base_off, (size_t)std::min(base_cap, s.base_n) * sizeof(float2),
                              hipMemcpyDeviceToHost));
    if (rds && rds_cap > 0 && s.rds_n > 0)
        ARGUS_CHECK(hipMemcpy(rds, rt->d_rds + s.rds_off, (size_t)std::min(rds_cap, s.rds_n) * sizeof(float2),
                              hipMemcpyDeviceToHost));
    return 0;
}

GPUD_API int gpud_rds_runtime_groups(gpud_rds_runtime* rt, int64_t id, uint16_t* groups4, int cap)
{
    if (!rt) return -1;
    auto it = rt->states.find(id);
    if (it == rt->states.end()) return -1;
    if (!rt->dec || it->second.fresh) return 0;
    return gpud_rds_batch_groups(rt->dec, it->second.slot, groups4, cap);
}

}  // extern "C"
