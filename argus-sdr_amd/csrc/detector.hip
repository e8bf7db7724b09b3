// detector.hip -- the detector's settings and its spectrum smoothing.
//
//  * gpud_detector_settings: detector.New's mapping of the configuration to
//    what the detection path reads (detector.go:101-261): CFAR guard and
//    training cells from Hz as ceil(Hz / binWidth) (train >= 1, guard >= 0),
//    the OS rank default round(0.75 * 2 * train), and the defaults of the
//    hysteresis, scale, edge margin, maximum bandwidth, merge gap and EMA
//    alpha.  Host code.
//  * gpud_detector_adaptive_alpha: Process's frame-rate-adaptive alpha
//    (detector.go:263-278).  Host code.
//  * gpud_surv_smooth_device: smoothSpectrum (detector.go:806-817) over a
//    run of frames on the device: ema = alpha_f * v + (1 - alpha_f) * ema per
//    bin, FP64 state, the first frame of a fresh detector copied.  One thread
//    per bin walks the frames in order (the recurrence is serial in time,
//    independent across bins): coalesced float reads and writes, the FP64
//    state in a register.  HBM-bound: 8 bytes per bin and frame.
#include "surv.hpp"
#include <cmath>

extern "C" {

GPUD_API int gpud_detector_settings(const gpud_detector_config* cfg, int sample_rate, int fft_size,
                                    gpud_cfar_params* cfar, gpud_detector_params* det, double* ema_alpha,
                                    int* fft_bins)
{
    if (!cfg || sample_rate <= 0) return -1;
    // fftSize bounded to (0, 4 Mi], else 65536 (detector.go:102-111)
    const int n = fft_size > 0 && fft_size <= 4194304 ? fft_size : 65536;
    const double bin_width = (double)sample_rate / (double)n;
    int guard = (int)std::ceil(cfg->cfar_guard_hz / bin_width);
    if (guard < 0) guard = 0;
    int train = (int)std::ceil(cfg->cfar_train_hz / bin_width);
    if (train < 1) train = 1;
    double alpha = cfg->ema_alpha;
    if (alpha <= 0 || alpha > 1) alpha = 0.2;
    double hyst = cfg->hysteresis_db;
    if (hyst <= 0) hyst = 3;
    double scale = cfg->cfar_scale_db;
    if (scale <= 0) scale = 6;
    double edge = cfg->edge_margin_db;
    if (edge <= 0) edge = 3.0;
    double max_bw = cfg->max_signal_bw_hz;
    if (max_bw <= 0) max_bw = 150000;
    double gap = cfg->merge_gap_hz;
    if (gap <= 0) gap = 5000;
    int rank = cfg->cfar_rank;
    if (rank <= 0 || rank > 2 * train) {
        rank = (int)std::round(0.75 * (double)(2 * train));
        if (rank <= 0) rank = 1;
    }
    if (cfar) {
        cfar->mode = cfg->cfar_mode;
        cfar->guard = guard;
        cfar->train = train;
        cfar->wrap = cfg->cfar_wrap ? 1 : 0;
        cfar->scale_db = scale;
        cfar->hysteresis_db = hyst;
        cfar->rank = rank;
        cfar->pad = 0;
    }
    if (det) {
        det->threshold_db = cfg->threshold_db;
        det->hysteresis_db = hyst;
        det->cfar_scale_db = scale;
        det->edge_margin_db = edge;
        det->max_signal_bw_hz = max_bw;
        det->merge_gap_hz = gap;
    }
    if (ema_alpha) *ema_alpha = alpha;
    if (fft_bins) *fft_bins = n;
    return 0;
}

GPUD_API double gpud_detector_adaptive_alpha(double ema_alpha, double dt_seconds)
{
    double dt = dt_seconds;
    if (!(dt > 0) || dt > 1.0) dt = 1.0 / 15.0;  // first frame, a clock step back or a stall
    const double ratio = dt / (1.0 / 15.0);
    double a = 1.0 - std::pow(1.0 - ema_alpha, ratio);
    if (a < 0.01) a = 0.01;
    if (a > 0.99) a = 0.99;
    return a;
}

}  // extern "C"

namespace {

constexpr int kSmoothBlock = 256;

__global__ __launch_bounds__(kSmoothBlock) void argus_ema_kernel(const float* __restrict__ psd, int n_bins,
                                                                   int n_frames, const double* __restrict__ alphas,
                                                                   int first, double* __restrict__ ema,
                                                                   float* __restrict__ out)
{
    // separate FP64 roundings in the reference's order: no fused multiply-add
#pragma clang fp contract(off)
    const int b = blockIdx.x * kSmoothBlock + threadIdx.x;
    if (b >= n_bins) return;
    double e = first ? 0.0 : ema[b];
    for (int f = 0; f < n_frames; f++) {
        const double v = (double)psd[(size_t)f * n_bins + b];
        if (first && f == 0) {
            e = v;
        } else {
            const double a = alphas[f];
            // alpha*v + (1-alpha)*ema, each operation rounded (contraction is off here)
            e = a * v + (1.0 - a) * e;
        }
        out[(size_t)f * n_bins + b] = (float)e;
    }
    ema[b] = e;
}

}  // namespace

extern "C" GPUD_API int gpud_surv_smooth_device(gpud_surv* s, const float* psd_dev, int n_frames,
                                                const double* alphas, int first, double* ema_dev, float* out_dev)
{
    if (!s || n_frames < 0 || (n_frames > 0 && (!psd_dev || !alphas || !ema_dev || !out_dev))) return -1;
    if (n_frames == 0) return 0;
    // the alphas travel through the engine's pinned staging, after the
    // copies already queued from it have been read
    ARGUS_CHECK(hipStreamSynchronize(s->stream));
    if (s->alpha_cap < (size_t)n_frames) {
        if (s->d_alpha) ARGUS_CHECK(hipFree(s->d_alpha));
        if (s->h_alpha) ARGUS_CHECK(hipHostFree(s->h_alpha));
        s->d_alpha = nullptr;
        s->h_alpha = nullptr;
        s->alpha_cap = 0;
        const size_t cap = (size_t)n_frames + 64;
        ARGUS_CHECK(hipMalloc((void**)&s->d_alpha, cap * sizeof(double)));
        ARGUS_CHECK(hipHostMalloc((void**)&s->h_alpha, cap * sizeof(double), hipHostMallocDefault));
        s->alpha_cap = cap;
    }
    for (int f = 0; f < n_frames; f++) s->h_alpha[f] = alphas[f];
    ARGUS_CHECK(hipMemcpyAsync(s->d_alpha, s->h_alpha, (size_t)n_frames * sizeof(double), hipMemcpyHostToDevice,
                               s->stream));
    argus_clear_error();
    hipLaunchKernelGGL(argus_ema_kernel, dim3((s->n + kSmoothBlock - 1) / kSmoothBlock), dim3(kSmoothBlock), 0,
                       s->stream, psd_dev, s->n, n_frames, (const double*)s->d_alpha, first ? 1 : 0, ema_dev,
                       out_dev);
    return (int)hipGetLastError();
}
