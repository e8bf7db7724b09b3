// wfm_batch.hip -- gpud_wfm_stereo_batch: the one-shot stereo decoder of
// whole recordings (internal/recorder/wfm_batch_stereo.go:76-157) on the
// stereo-stage kernels of wfm.hip in batch mode, one descriptor per station.
//
// Workspace, one stream-ordered allocation:
//   [descs | tile items | state per station | lock block per station |
//    tap sets per station | zero delay line | scratch 4 x total samples]
// The PLL of every station runs on its own wavefront, so a batch of
// recordings decodes in the time of the longest one.
#include "wfm.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace {

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

GPUD_API int gpud_wfm_stereo_batch(const float* mono_dev, const int64_t* offsets, const int32_t* rates,
                                   int n_stations, int mono_fallback, float* out_dev, int32_t* locked,
                                   gpud_stream_handle stream)
{
    if (n_stations < 0 || (n_stations > 0 && (!offsets || !rates || !locked))) return -1;
    if (mono_fallback != 0 && mono_fallback != 1) return -1;
    if (n_stations == 0) return 0;
    if (offsets[0] < 0) return -1;
    std::vector<int> live;  // stations with samples and a rate
    int64_t total = 0;
    for (int i = 0; i < n_stations; i++) {
        const int64_t n = offsets[i + 1] - offsets[i];
        if (n < 0 || n > (int64_t)0x7fffffff - ARGUS_WFM_TILE) return -1;
        locked[i] = 0;
        if (n > 0 && rates[i] > 0) {
            live.push_back(i);
            total += n;
        }
    }
    if (live.empty()) return 0;
    if (!mono_dev || !out_dev) return -1;
    hipStream_t st = (hipStream_t)stream;
    const int nl = (int)live.size();

    std::vector<ArgusWfmItem> items;
    for (int k = 0; k < nl; k++) {
        const int n = (int)(offsets[live[k] + 1] - offsets[live[k]]);
        for (int s0 = 0; s0 < n; s0 += ARGUS_WFM_TILE) items.push_back({k, s0, ARGUS_WFM_ITEM_TILE, 0});
    }
    const size_t tap_floats = (size_t)ARGUS_WFM_NTAPSETS * ARGUS_WFM_TAPS;
    const size_t off_items = align256((size_t)nl * sizeof(ArgusWfmDesc));
    const size_t off_state = off_items + align256(items.size() * sizeof(ArgusWfmItem));
    const size_t off_blk = off_state + align256((size_t)nl * sizeof(ArgusWfmState));
    const size_t off_taps = off_blk + align256((size_t)nl * sizeof(ArgusWfmBlock));
    const size_t off_zero = off_taps + align256((size_t)nl * tap_floats * sizeof(float));
    const size_t off_scratch = off_zero + align256(ARGUS_WFM_HIST * sizeof(float));
    const size_t bytes = off_scratch + (size_t)total * 4 * sizeof(float);
    const size_t head = off_scratch;  // the part uploaded from the host

    char* d = nullptr;
    ARGUS_CHECK(hipMallocAsync((void**)&d, bytes, st));
    std::vector<char> h(head, 0);
    ArgusWfmDesc* descs = (ArgusWfmDesc*)h.data();
    ArgusWfmState* states = (ArgusWfmState*)(h.data() + off_state);
    float* taps = (float*)(h.data() + off_taps);
    int last_rate = -1;
    int64_t sc = 0;
    for (int k = 0; k < nl; k++) {
        const int i = live[k];
        const int rate = rates[i];
        const int64_t o = offsets[i];
        const int n = (int)(offsets[i + 1] - o);
        float* tk = taps + (size_t)k * tap_floats;
        if (rate == last_rate)
            std::memcpy(tk, tk - tap_floats, tap_floats * sizeof(float));
        else
            argus_wfm_design_taps(rate, tk);
        last_rate = rate;
        states[k].freq = 2.0 * M_PI * 19000.0 / (double)rate;  // wfm_batch_stereo.go:100
        ArgusWfmDesc& wd = descs[k];
        wd.mono = mono_dev + o;
        wd.audio = out_dev + 2 * o;
        wd.scratch = (float*)(d + off_scratch) + 4 * sc;
        wd.rds = nullptr;
        wd.mono_hist_rd = (const float*)(d + off_zero);
        wd.mono_hist_wr = nullptr;  // no tail items in batch mode
        wd.lr_hist_rd = (const float*)(d + off_zero);
        wd.lr_hist_wr = nullptr;
        wd.fb_hist = nullptr;
        wd.taps = (const float*)(d + off_taps) + (size_t)k * tap_floats;
        wd.st = (ArgusWfmState*)(d + off_state) + k;
        wd.blk = (ArgusWfmBlock*)(d + off_blk) + k;
        wd.mix0 = 0;
        wd.mix_inc = 0;
        argus_wfm_loop_constants(rate, wd);
        wd.n = n;
        wd.rds_first = 0;
        wd.rds_decim = 1;
        wd.n_rds = 0;
        wd.batch = mono_fallback ? ARGUS_WFM_BATCH_FALLBACK : ARGUS_WFM_BATCH_MATRIX;
        wd.pad = 0;
        sc += n;
    }
    std::memcpy(h.data() + off_items, items.data(), items.size() * sizeof(ArgusWfmItem));
    std::vector<ArgusWfmBlock> blk((size_t)nl);
    hipError_t e = hipMemcpyAsync(d, h.data(), head, hipMemcpyHostToDevice, st);
    int err = (int)e;
    if (e == hipSuccess)
        err = argus_launch_wfm((const ArgusWfmDesc*)d, nl, (const ArgusWfmItem*)(d + off_items), (int)items.size(),
                               st);
    if (!err)
        err = (int)hipMemcpyAsync(blk.data(), d + off_blk, (size_t)nl * sizeof(ArgusWfmBlock), hipMemcpyDeviceToHost,
                                  st);
    hipFreeAsync(d, st);
    const hipError_t se = hipStreamSynchronize(st);
    if (!err) err = (int)se;
    if (err) return err;
    for (int k = 0; k < nl; k++) locked[live[k]] = blk[(size_t)k].locked;
    return 0;
}

}  // extern "C"
