// serve.hpp -- state of the streaming service.
#pragma once
#include "extract.hpp"
#include "demod.hpp"
#include "wfm.hpp"
#include "audio.hpp"
#include "rds.hpp"
#include "fastconv.hpp"
#include "predemod.hpp"
#include <deque>
#include <map>
#include <tuple>
#include <utility>
#include <vector>

#define ARGUS_SERVE_SLOTS 3

struct ServeChannel {
    gpud_channel_spec spec{};
    int decim = 0, T = 0;
    size_t footprint = 0;
    // the channel's state block in the service's device arena (taps,
    // modulated taps, histories, carries and the optional stage blocks point
    // into it), and the separate arena blocks of the audio and RDS stages
    char* d_block = nullptr;
    size_t block_bytes = 0;
    size_t aud_bytes = 0, rdec_bytes = 0;
    float* d_taps = nullptr;
    float2* d_ctaps = nullptr;  // modulated taps for the channel's NCO increment (T)
    float2* d_hist[2] = {nullptr, nullptr};
    int hist_sel = 0, hist_len = 0;
    float2* d_carry = nullptr;  // 2 slots, ping-pong
    int carry_sel = 0;
    bool have_carry = false;
    double nco_phase = 0.0;
    int phase_count = 0;
    double bfo_phase = 0.0;
    int64_t last_ticket = 0;  // last block that used the channel
    // WFM stereo + RDS (demod GPUD_DEMOD_WFM_STEREO): one device block of
    // ArgusWfmState | mono delay lines [2] | L-R delay lines [2] |
    // fallback delay lines [2] | tap sets
    bool stereo = false;
    char* d_wfm = nullptr;
    int wfm_sel = 0;         // ping-pong half of the mono and L-R delay lines
    int64_t rds_count = 0;   // discriminator samples seen so far
    int rds_decim = 1;
    // post-demodulation audio stage (gpud_serve_set_audio): one device block of
    // delay lines [2 halves][2 sides][ARGUS_AUDIO_TMAX] | de-emphasis states [2]
    int audio_rate = 0;      // 0: stage off
    double deemph_us = 0.0;
    int aL = 1, aM = 1, aT = 1, at0 = 0;
    bool a_ident = false;    // equal rates
    int a_sel = 0;           // delay-line half read by the next block
    const float* d_arms = nullptr;  // shared per filter (gpud_serve::filters)
    char* d_aud = nullptr;
    // device-side RDS group decoding (gpud_serve_set_rds): one device block of
    // ArgusRdsState | two decode windows of rds_win samples; the block's RDS
    // baseband fills window rwin_sel, a full window is decoded and the
    // other one fills next
    // fast-convolution path (fastconv.hip): the transform of the channel's
    // modulated taps, for channels whose decimation and length allow it
    float2* d_fastG = nullptr;
    int rds_win = 0;         // 0: off
    char* d_rdec = nullptr;
    int rwin_sel = 0, rwin_fill = 0;
    ArgusRdsOut r_last{};    // state after the latest completed decode
    int64_t r_last_ticket = 0;
    // pre-demod decimating FIR (gpud_serve_open_ex with a demod rate): one
    // device block of FP64 taps [ARGUS_PRE_TAPS] | delay lines [2][T-1]
    // (ping-pong); pre_phase counts the inputs until its next output
    int pre_factor = 1;      // 1: no stage
    int pre_phase = 0;
    int pre_sel = 0;
    char* d_pre = nullptr;
    int demod_rate = 0;      // rate the demodulator (and stereo, RDS, audio) runs at
};

struct ServeOut {
    int64_t channel_id;
    int64_t iq_off, audio_off;
    int n_iq, n_audio, rate;
    int wfm_index;           // index among the block's stereo channels, or -1
    int64_t mono_off, rds_off;
    int n_mono, n_rds, rds_rate;
    int a_index;             // index among the block's audio-stage channels, or -1
    int n_aout, a_channels, a_tnext;  // audio frames, channels per frame, next resampler phase
    int a_rate;                       // audio rate (0: no audio stage)
    int64_t aout_off;                 // audio output offset (floats)
    int r_on;                         // RDS decoding on for the channel in this block
    int r_index;                      // index of its decode among the block's, or -1
    int n_pre;                        // demodulator input samples (the pre-demod stage's outputs, or n_iq)
    int64_t pre_off;                  // offset of the pre-demod outputs, -1 without the stage
    int early;                        // outputs complete after the fast kernels (see BlockLayout)
};

struct ServeSlot {
    int64_t ticket = 0;
    bool busy = false;
    int n = 0;
    hipEvent_t ev_start = nullptr, ev_in = nullptr, ev_kstart = nullptr, ev_kern = nullptr, ev_done = nullptr;
    hipEvent_t ev_xend = nullptr;  // after the extraction launch
    // after part p of the fast kernels and its early channels' demodulation
    hipEvent_t ev_part[ARGUS_FAST_PARTS] = {};
    float2* d_in = nullptr;
    float2* h_in = nullptr;
    float2* d_iq = nullptr;
    float* d_audio = nullptr;
    size_t iq_cap = 0, audio_cap = 0;
    float2* h_iq = nullptr;
    float* h_audio = nullptr;
    size_t h_iq_cap = 0, h_audio_cap = 0;
    // The block's tables, one pinned host buffer mirrored by one device
    // buffer: extraction work items | ArgusServeHdr | channel descriptors |
    // demod descriptors | stereo descriptors | demod items | stereo items |
    // audio descriptors.
    // The work items depend only on the channel set, the block length and
    // the tune, so they are rebuilt and uploaded only when those change
    // (items_epoch, items_n); everything from the header on is uploaded with
    // every block.
    // the extraction work items, in buffers of their own (the header points
    // at them, so rebuilding or moving them does not change what a captured
    // graph launches; the header sits at the start of the table)
    ArgusWorkItem* h_items = nullptr;
    ArgusWorkItem* d_items = nullptr;
    size_t h_items_cap = 0, d_items_cap = 0;
    char* h_tab = nullptr;
    char* d_tab = nullptr;
    size_t tab_cap = 0;
    size_t tab_bytes = 0;
    size_t off_hdr = 0;
    size_t lds = 0;
    uint64_t items_epoch = 0;
    int items_n = -1;
    int n_items = 0;
    bool items_dirty = true;
    // captured launches of this slot (extraction; demodulation + stereo
    // stage) and what they were captured for
    hipGraphExec_t gexec[ARGUS_FAST_PARTS + 2] = {};
    const char* g_hdr = nullptr;
    size_t g_lds = 0;
    int g_grid = 0;
    int g_fgrid = 0;
    int g_wfm = -1;
    // WFM stereo stage
    float* d_mono = nullptr;
    size_t mono_cap = 0;
    float* d_wscratch = nullptr;
    size_t wscratch_cap = 0;
    float2* d_rds = nullptr;
    size_t rds_cap = 0;
    float2* h_rds = nullptr;
    size_t h_rds_cap = 0;
    ArgusWfmBlock* d_wblk = nullptr;
    size_t wblk_cap = 0;
    ArgusWfmBlock* h_wblk = nullptr;
    size_t h_wblk_cap = 0;
    int n_wfm = 0;
    int64_t rds_total = 0;
    int64_t iq_total = 0, audio_total = 0;
    int64_t iq_early = 0, audio_early = 0;  // the early channels' share (BlockLayout)
    // end of the early outputs of fast-kernel parts 0..p (describe_block)
    int64_t iq_part_end[ARGUS_FAST_PARTS] = {}, audio_part_end[ARGUS_FAST_PARTS] = {};
    int n_parts = 1;  // fast-kernel parts launched for the described block (the rest are empty)
    // post-demodulation audio stage
    float* d_aout = nullptr;
    size_t aout_cap = 0;
    float* h_aout = nullptr;
    size_t h_aout_cap = 0;
    int n_audio_chans = 0;
    int64_t aout_total = 0;
    // pre-demod stage outputs
    float2* d_pre = nullptr;
    size_t pre_cap = 0;
    float2* h_pre = nullptr;
    size_t h_pre_cap = 0;
    int64_t pre_total = 0;
    int n_pre_chans = 0;
    int p_mode = 0;          // the described block runs the pre-demod stage (1) or not (0)
    int g_pre = -1;          // p_mode the stage graph was captured for
    int fast_S = 0;          // segments of the fast path in the described block (0: none)
    int g_fastS = -1;        // fast_S the extraction graph was captured for
    int a_mode = 0;          // the described block runs the audio stage (1) or not (0)
    int g_audio = -1;        // a_mode the stage graph was captured for
    // device-side RDS decoding of the block (after its graphs): a pinned
    // table mirrored on the device, [appends | calls | outs | groups], the
    // outs and groups of the block's completed decodes coming back
    char* h_rtab = nullptr;
    char* d_rtab = nullptr;
    size_t rtab_cap = 0, h_rtab_cap = 0;
    int n_rdec = 0;                      // decodes completed in the block
    size_t off_rout = 0;                 // outs in the table
    size_t off_rgroups = 0;          // groups in the table
    size_t rgroups_bytes = 0;            // outs + groups
    std::vector<int64_t> rdec_channel;   // channel id per decode
    std::vector<size_t> rdec_groups;     // first group word per decode
    std::vector<ServeOut> outs;
};

// A block's output layout.  Channels whose outputs are complete once the
// fast-convolution kernels and their demodulation have run (fast form this
// block, plain demodulation: no pre-demod, stereo or audio stage) come first
// in the IQ and demodulated buffers, [0, iq_early) and [0, audio_early), in
// channel order, so each part of the fast kernels' outputs starts down while
// the later parts and the direct form still run.
struct BlockLayout {
    std::vector<ServeOut> outs;
    int n_wfm = 0;
    int n_audio_chans = 0;
    int n_pre_chans = 0;
    int64_t iq_total = 0, audio_total = 0, mono_total = 0, rds_total = 0, aout_total = 0, pre_total = 0;
    int64_t iq_early = 0, audio_early = 0;
};

// A block prepared ahead: after a block is launched, the next block's
// tables are built and uploaded into the next slot while the GPU works, on
// the assumption that it comes with the same channel set, length and input
// buffer.  Its layout waits here (the slot still holds the outputs of its
// previous block until the next submit).
struct ServePrep {
    bool valid = false;
    int slot = -1;
    int64_t ticket = 0;
    uint64_t epoch = 0;
    int n = 0;
    const float2* in = nullptr;
    BlockLayout layout;
};

// What the graph-mode kernels of one block read (first bytes of the table).
struct ArgusServeHdr {
    ArgusExtractHdr x;
    ArgusDemodHdr d;   // demodulation of the channels that are not early
    ArgusDemodHdr de[ARGUS_FAST_PARTS];  // demodulation of the early channels, per fast-kernel part
    ArgusWfmHdr w;
    ArgusAudioHdr a;
    ArgusFastHdr f;
    ArgusPreHdr p;
};

// Device memory for channel state, taken from the budget in regions: the
// first of 256 MiB (or the cap), each further one as large as everything
// before it, up to `cap` in all.  Free space is a first-fit list of
// (address, bytes) that coalesces within a region; a region is kept until
// the service is destroyed.  Growing by regions keeps a service's footprint
// near what its channels use: allocating (and later freeing) the whole cap
// at once costs the driver a full clear of that memory.
struct ServeArena {
    size_t cap = 0;                              // bytes the regions may total
    size_t total = 0;                            // bytes allocated in regions
    std::map<char*, size_t> regions;             // base -> bytes
    std::map<char*, size_t> free_list;           // address -> bytes
    char* alloc(size_t bytes);
    void release(char* p, size_t bytes);
    void destroy();
};

// Pinned staging of channel-state uploads: regions taken in ring order, each
// copied to the device on the compute stream (ahead of the channel's first
// block) and reusable once the event recorded after its copy has completed.
struct ServeStage {
    char* ring = nullptr;
    size_t size = 0, head = 0;
    std::deque<std::pair<std::pair<size_t, size_t>, hipEvent_t>> busy;  // (offset, bytes), event
    std::vector<hipEvent_t> spare;
};

struct gpud_serve {
    ArgusTune tune = argus_default_tune();
    ServeArena arena;
    ServeStage stage;
    int64_t graph_captures = 0;  // instantiations of the per-slot graphs
    size_t lds_hw = 0;           // largest dynamic LDS of an extraction launch since the tune was set
    int n_fast_capable = 0;      // live channels with a fast-form spectrum
    uint64_t epoch = 1;  // bumped when the channel set or the tune changes
    bool graph = true;  // replay a captured per-slot graph of the block's launches
    int sample_rate = 0;
    int block_samples = 0;
    uint64_t budget = 0;
    uint64_t bytes_in_use = 0;
    int dl_mask = GPUD_DL_ALL;  // parts of each block's outputs downloaded (GPUD_DL_*)
    // the fast form runs in up to max_parts launches, one per part_min_bytes
    // of early outputs downloaded (gpud_serve_set_download_split)
    int max_parts = ARGUS_FAST_PARTS;
    int64_t part_min_bytes = (int64_t)16 << 20;
    hipStream_t compute = nullptr, copy = nullptr;
    std::map<int64_t, ServeChannel> live;  // ordered by id: a deterministic channel order per block
    std::vector<ServeChannel> closed;     // awaiting their last block
    ServeSlot slots[ARGUS_SERVE_SLOTS];
    int next_slot = 0;
    int64_t last_ticket = 0;
    int64_t completed_ticket = 0;
    ServePrep prep;
    std::map<std::tuple<int, int, bool>, float*> filters;  // audio filters per (in rate, out rate, low-pass)
    std::vector<ArgusAudioDesc> adescs;
    std::vector<ArgusWorkItem> items;
    std::vector<ArgusChanDesc> chans;
    std::vector<ArgusDemodDesc> ddescs;
    std::vector<ArgusWfmDesc> wdescs;
    std::vector<ArgusDemodItem> ditems;
    std::vector<ArgusDemodItem> ditems_e;  // the early channels' (serve.hip: layout_block)
    // per fast channel of the block being described: early demod items and
    // early output ends up to and including it
    std::vector<int> f_ditems;
    std::vector<int64_t> f_iq_end, f_au_end;
    std::vector<ArgusWfmItem> witems;
    std::vector<ArgusPreDesc> pdescs;
    std::vector<ArgusPreItem> pitems;
    int n_rds_on = 0;  // live channels with RDS decoding on
    // fast-convolution path: on/off, the stream's last input samples, the
    // staged segments and their transforms (sized for the largest block)
    bool fast = true;
    bool fuse_demod = true;  // demodulation in the fast kernel's epilogue (gpud_serve_set_demod_fusion)
    float2* d_ftail = nullptr;
    float2* d_fu = nullptr;
    float2* d_fX = nullptr;
    std::vector<ArgusFastDesc> fdescs;
    std::vector<ArgusRdsAppend> rapp;
    std::vector<ArgusRdsCall> rcalls;
};
