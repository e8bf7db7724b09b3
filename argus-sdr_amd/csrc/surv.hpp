// surv.hpp -- surveillance engine state.
#pragma once
#include "common.hpp"

// Most CFAR runs per frame the detection list takes (device finish).
#define ARGUS_DETECT_MAX_RUNS 2048

struct gpud_surv {
    int n = 0;             // FFT size
    int sample_rate = 0;
    hipStream_t stream = nullptr;
    float* d_win = nullptr;  // symmetric Hann, n floats
    float2* d_scratch = nullptr;  // complex spectra of gpud_surv_spectrum_device
    size_t scratch_cap = 0;       // complex samples
    // CFAR: per-frame decision masks and largest dB values (bytes)
    void* d_cfar = nullptr;
    size_t cfar_cap = 0;
    // decimated level (gpud_surv_level_spectrum_device): zero-stuffed frames,
    // FP64 low-pass taps and the level's Hann window for level_decim
    float2* d_zstuff = nullptr;
    size_t zstuff_cap = 0;
    double* d_ltaps = nullptr;
    float* d_lwin = nullptr;
    int level_decim = 0, level_taps = 0;
    // per-frame EMA alphas of gpud_surv_smooth_device (device and pinned staging)
    double* d_alpha = nullptr;
    double* h_alpha = nullptr;
    size_t alpha_cap = 0;
};
