// detect.hip -- the detection list after the CFAR runs: edge expansion,
// merging and the peak-bounded power centroid (detector.go:483-639,
// 698-713), on the host for one frame (gpud_surv_finish_frame) and on the
// device for many frames at once (gpud_surv_finish_device), plus the whole
// surveillance pass (Welch -> CFAR runs -> detection list) over
// device-resident IQ.
//
// argus_detect_finish_kernel: one workgroup per frame.  Threads build the
// signals from the runs and expand their edges one signal per thread (the
// outward walks are short and independent); the global median the
// expansion falls back to near the band edges is an exact order statistic
// found by a four-pass 8-bit radix select over the frame's float keys; the
// stable sort by first bin is a rank computation (rank = earlier-or-equal
// keys before it); the merge is a serial pass of one thread over the sorted
// signals; the centroids are again one signal per thread.  Every step
// follows the host code's arithmetic in the same order, so bins are
// identical and centres differ only by the last bits of pow().
#include "common.hpp"
#include "surv.hpp"
#include <algorithm>
#include <cmath>
#include <vector>

namespace {

// detector.go:685-696 (sorted copy, mean of the middle pair when even)
double median_of(const float* s, int n)
{
    if (n <= 0) return 0;
    std::vector<double> c(s, s + n);
    const int mid = n / 2;
    std::nth_element(c.begin(), c.begin() + mid, c.end());
    const double hi = c[(size_t)mid];
    if (n % 2) return hi;
    const double lo = *std::max_element(c.begin(), c.begin() + mid);
    return (lo + hi) / 2;
}

double min_in_range(const float* s, int n, int from, int to)
{
    if (n == 0) return 0;
    if (from < 0) from = 0;
    if (to >= n) to = n - 1;
    if (from > to) return s[std::min(std::max(from, 0), n - 1)];
    double m = s[from];
    for (int i = from + 1; i <= to; i++) m = std::min(m, (double)s[i]);
    return m;
}

struct Det {
    double bin_width;
    int nbins;
    const gpud_detector_params* p;
    double center_for_bin(double bin) const { return p->center_hz + (bin - (double)nbins / 2.0) * bin_width; }
};

// detector.go:602-639
double power_weighted_center(const Det& d, const float* s, int n, int first, int last)
{
    if (first > last || first < 0 || last >= n) return d.center_for_bin((double)(first + last) / 2.0);
    int pk = first;
    double pv = s[first];
    for (int i = first; i <= last; i++)
        if (s[i] > pv) {
            pv = s[i];
            pk = i;
        }
    int win = d.bin_width > 0 ? (int)(30000.0 / d.bin_width) : 1;
    if (win < 1) win = 1;
    const int lo = std::max(pk - win, first), hi = std::min(pk + win, last);
    double sp = 0, sw = 0;
    for (int i = lo; i <= hi; i++) {
        const double pw = std::pow(10.0, (double)s[i] / 10.0);
        sp += pw;
        sw += pw * (double)i;
    }
    if (sp <= 0) return d.center_for_bin((double)pk);
    return d.center_for_bin(sw / sp);
}

}  // namespace

extern "C" GPUD_API int gpud_surv_finish_frame(const float* s, int n, int sample_rate, const gpud_surv_run* runs,
                                               int n_runs, const gpud_detector_params* params, gpud_surv_signal* out,
                                               int cap)
{
    if (!s || n <= 0 || sample_rate <= 0 || !params || n_runs < 0 || (n_runs > 0 && !runs)) return -1;
    Det d{(double)sample_rate / (double)n, n, params};
    std::vector<gpud_surv_signal> sig((size_t)n_runs);
    // detector.go:698-713 makeSignal
    for (int i = 0; i < n_runs; i++) {
        const gpud_surv_run& r = runs[i];
        gpud_surv_signal& g = sig[(size_t)i];
        g.first_bin = r.first_bin;
        g.last_bin = r.last_bin;
        g.peak_bin = r.peak_bin;
        g.pad = 0;
        g.center_hz = params->center_hz + ((double)r.peak_bin - (double)n / 2.0) * d.bin_width;
        g.bw_hz = (double)(r.last_bin - r.first_bin + 1) * d.bin_width;
        g.peak_db = r.peak_db;
        g.snr_db = r.peak_db - r.noise_db;
        g.noise_db = r.noise_db;
    }
    // detector.go:483-549 expandSignalEdges (median only when a probe leaves the band)
    const double margin = params->edge_margin_db > 0 ? params->edge_margin_db : 3.0;
    int max_exp = (int)(params->max_signal_bw_hz / d.bin_width);
    if (max_exp < 10) max_exp = 10;
    bool have_median = false;
    double noise_global = 0;
    for (auto& g : sig) {
        double local;
        const int lp = g.first_bin - 50, rp = g.last_bin + 50;
        if (lp >= 0 && rp < n) {
            const double ln = min_in_range(s, n, std::max(0, lp), std::max(0, g.first_bin - 5));
            const double rn = min_in_range(s, n, std::min(n - 1, g.last_bin + 5), std::min(n - 1, rp));
            local = std::min(ln, rn);
        } else {
            if (!have_median) {
                noise_global = median_of(s, n);
                have_median = true;
            }
            local = noise_global;
        }
        const double edge = local + margin;
        int nf = g.first_bin;
        double prev = s[nf];
        for (int j = 0; j < max_exp; j++) {
            const int nx = nf - 1;
            if (nx < 0) break;
            const double val = s[nx];
            if (val <= edge) break;
            if (val > prev + 1.0 && val < g.peak_db - 6.0) break;
            prev = val;
            nf = nx;
        }
        int nl = g.last_bin;
        prev = s[nl];
        for (int j = 0; j < max_exp; j++) {
            const int nx = nl + 1;
            if (nx >= n) break;
            const double val = s[nx];
            if (val <= edge) break;
            if (val > prev + 1.0 && val < g.peak_db - 6.0) break;
            prev = val;
            nl = nx;
        }
        g.first_bin = nf;
        g.last_bin = nl;
        g.bw_hz = (double)(nl - nf + 1) * d.bin_width;
    }
    // detector.go:551-588 mergeOverlapping (stable order of equal first bins;
    // the merge is symmetric in the operands, so the order does not change it)
    int ns = n_runs;
    if (ns > 1) {
        int gap_bins = 0;
        if (params->merge_gap_hz > 0 && d.bin_width > 0) gap_bins = (int)std::ceil(params->merge_gap_hz / d.bin_width);
        std::stable_sort(sig.begin(), sig.end(),
                         [](const gpud_surv_signal& a, const gpud_surv_signal& b) { return a.first_bin < b.first_bin; });
        int m = 1;
        for (int i = 1; i < ns; i++) {
            gpud_surv_signal& last = sig[(size_t)m - 1];
            const gpud_surv_signal cur = sig[(size_t)i];
            const int gap = cur.first_bin - last.last_bin - 1;
            if (gap <= gap_bins) {
                if (cur.last_bin > last.last_bin) last.last_bin = cur.last_bin;
                if (cur.peak_db > last.peak_db) last.peak_db = cur.peak_db;
                if (cur.snr_db > last.snr_db) last.snr_db = cur.snr_db;
                last.bw_hz = (double)(last.last_bin - last.first_bin + 1) * d.bin_width;
                last.center_hz = d.center_for_bin((double)(last.first_bin + last.last_bin) / 2.0);
                if (cur.noise_db < last.noise_db || last.noise_db == 0) last.noise_db = cur.noise_db;
            } else {
                sig[(size_t)m++] = cur;
            }
        }
        ns = m;
    }
    // detector.go:475-479 centroid and final bandwidth
    const int nout = std::min(ns, cap);
    for (int i = 0; i < nout; i++) {
        gpud_surv_signal g = sig[(size_t)i];
        g.center_hz = power_weighted_center(d, s, n, g.first_bin, g.last_bin);
        g.bw_hz = (double)(g.last_bin - g.first_bin + 1) * d.bin_width;
        out[i] = g;
    }
    return ns;
}

namespace {

constexpr int kFinishThreads = 256;

__device__ __forceinline__ uint32_t order_key(float v)
{
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float key_value(uint32_t k)
{
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// The k-th smallest (0-based) of s[0..n), for every thread of the block.
__device__ float block_select(const float* s, int n, int k, uint32_t* hist, int* st)
{
    uint32_t prefix = 0, mask = 0;
    for (int pass = 0; pass < 4; pass++) {
        const int shift = 24 - 8 * pass;
        for (int i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0;
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const uint32_t u = order_key(s[i]);
            if ((u & mask) == prefix) atomicAdd(&hist[(u >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int d = 0; d < 256; d++) {
                if (acc + (int)hist[d] > k) {
                    st[0] = d;
                    st[1] = k - acc;
                    break;
                }
                acc += (int)hist[d];
            }
        }
        __syncthreads();
        prefix |= (uint32_t)st[0] << shift;
        mask |= 255u << shift;
        k = st[1];
        __syncthreads();
    }
    return key_value(prefix);
}

__device__ double dev_min_in_range(const float* s, int n, int from, int to)
{
    if (n == 0) return 0;
    if (from < 0) from = 0;
    if (to >= n) to = n - 1;
    if (from > to) return s[min(max(from, 0), n - 1)];
    double m = s[from];
    for (int i = from + 1; i <= to; i++) m = fmin(m, (double)s[i]);
    return m;
}

struct FinishArgs {
    const float* psd;
    const gpud_surv_run* runs;
    const int* run_counts;
    gpud_surv_signal* work;  // 2 x max_runs per frame
    gpud_surv_signal* out;
    int* out_counts;
    int n, max_runs, cap;
    double bin_width;
    gpud_detector_params p;
};

__global__ __launch_bounds__(kFinishThreads) void argus_detect_finish_kernel(FinishArgs a)
{
    __shared__ uint32_t hist[256];
    __shared__ int sel[2];
    __shared__ int need_median;
    __shared__ double noise_global;
    __shared__ int keys[ARGUS_DETECT_MAX_RUNS];
    __shared__ int merged;
    const int f = blockIdx.x, tid = threadIdx.x, n = a.n;
    const float* s = a.psd + (size_t)f * n;
    const int nr = min(a.run_counts[f], a.max_runs);
    gpud_surv_signal* A = a.work + (size_t)f * 2 * a.max_runs;
    gpud_surv_signal* B = A + a.max_runs;
    const gpud_detector_params& p = a.p;
    const double bw = a.bin_width;
    if (tid == 0) need_median = 0;
    __syncthreads();
    // detector.go:698-713 makeSignal
    for (int i = tid; i < nr; i += kFinishThreads) {
        const gpud_surv_run r = a.runs[(size_t)f * a.max_runs + i];
        gpud_surv_signal g;
        g.first_bin = r.first_bin;
        g.last_bin = r.last_bin;
        g.peak_bin = r.peak_bin;
        g.pad = 0;
        g.center_hz = p.center_hz + ((double)r.peak_bin - (double)n / 2.0) * bw;
        g.bw_hz = (double)(r.last_bin - r.first_bin + 1) * bw;
        g.peak_db = r.peak_db;
        g.snr_db = r.peak_db - r.noise_db;
        g.noise_db = r.noise_db;
        A[i] = g;
        if (!(r.first_bin - 50 >= 0 && r.last_bin + 50 < n)) need_median = 1;
    }
    __syncthreads();
    if (need_median) {  // detector.go:685-696
        const int mid = n / 2;
        const double hi = block_select(s, n, mid, hist, sel);
        double med = hi;
        if (n % 2 == 0) med = ((double)block_select(s, n, mid - 1, hist, sel) + hi) / 2;
        if (tid == 0) noise_global = med;
    }
    __syncthreads();
    // detector.go:483-549 expandSignalEdges
    const double margin = p.edge_margin_db > 0 ? p.edge_margin_db : 3.0;
    int max_exp = (int)(p.max_signal_bw_hz / bw);
    if (max_exp < 10) max_exp = 10;
    for (int i = tid; i < nr; i += kFinishThreads) {
        gpud_surv_signal g = A[i];
        double local;
        const int lp = g.first_bin - 50, rp = g.last_bin + 50;
        if (lp >= 0 && rp < n) {
            const double ln = dev_min_in_range(s, n, max(0, lp), max(0, g.first_bin - 5));
            const double rn = dev_min_in_range(s, n, min(n - 1, g.last_bin + 5), min(n - 1, rp));
            local = fmin(ln, rn);
        } else {
            local = noise_global;
        }
        const double edge = local + margin;
        int nf = g.first_bin;
        double prev = s[nf];
        for (int j = 0; j < max_exp; j++) {
            const int nx = nf - 1;
            if (nx < 0) break;
            const double val = s[nx];
            if (val <= edge) break;
            if (val > prev + 1.0 && val < g.peak_db - 6.0) break;
            prev = val;
            nf = nx;
        }
        int nl = g.last_bin;
        prev = s[nl];
        for (int j = 0; j < max_exp; j++) {
            const int nx = nl + 1;
            if (nx >= n) break;
            const double val = s[nx];
            if (val <= edge) break;
            if (val > prev + 1.0 && val < g.peak_db - 6.0) break;
            prev = val;
            nl = nx;
        }
        g.first_bin = nf;
        g.last_bin = nl;
        g.bw_hz = (double)(nl - nf + 1) * bw;
        A[i] = g;
        keys[i] = nf;
    }
    __syncthreads();
    // detector.go:551-588 mergeOverlapping: stable sort by first bin (rank =
    // smaller keys + equal keys earlier), then one serial pass
    for (int i = tid; i < nr; i += kFinishThreads) {
        const int ki = keys[i];
        int rank = 0;
        for (int j = 0; j < nr; j++) rank += (keys[j] < ki || (keys[j] == ki && j < i)) ? 1 : 0;
        B[rank] = A[i];
    }
    __syncthreads();
    if (tid == 0) {
        int m = nr;
        if (nr > 1) {
            int gap_bins = 0;
            if (p.merge_gap_hz > 0 && bw > 0) gap_bins = (int)ceil(p.merge_gap_hz / bw);
            m = 1;
            for (int i = 1; i < nr; i++) {
                gpud_surv_signal last = B[m - 1];
                const gpud_surv_signal cur = B[i];
                const int gap = cur.first_bin - last.last_bin - 1;
                if (gap <= gap_bins) {
                    if (cur.last_bin > last.last_bin) last.last_bin = cur.last_bin;
                    if (cur.peak_db > last.peak_db) last.peak_db = cur.peak_db;
                    if (cur.snr_db > last.snr_db) last.snr_db = cur.snr_db;
                    last.bw_hz = (double)(last.last_bin - last.first_bin + 1) * bw;
                    last.center_hz = p.center_hz + ((double)(last.first_bin + last.last_bin) / 2.0 - (double)n / 2.0) * bw;
                    if (cur.noise_db < last.noise_db || last.noise_db == 0) last.noise_db = cur.noise_db;
                    B[m - 1] = last;
                } else {
                    B[m++] = cur;
                }
            }
        }
        merged = m;
        a.out_counts[f] = m;
    }
    __syncthreads();
    // detector.go:602-639 powerWeightedCenter and the final bandwidth
    const int nout = min(merged, a.cap);
    for (int i = tid; i < nout; i += kFinishThreads) {
        gpud_surv_signal g = B[i];
        const int first = g.first_bin, last = g.last_bin;
        double c;
        if (first > last || first < 0 || last >= n) {
            c = p.center_hz + ((double)(first + last) / 2.0 - (double)n / 2.0) * bw;
        } else {
            int pk = first;
            double pv = s[first];
            for (int k = first; k <= last; k++)
                if (s[k] > pv) {
                    pv = s[k];
                    pk = k;
                }
            int win = bw > 0 ? (int)(30000.0 / bw) : 1;
            if (win < 1) win = 1;
            const int lo = max(pk - win, first), hi = min(pk + win, last);
            double sp = 0, sw = 0;
            for (int k = lo; k <= hi; k++) {
                const double pw = pow(10.0, (double)s[k] / 10.0);
                sp += pw;
                sw += pw * (double)k;
            }
            c = sp <= 0 ? p.center_hz + ((double)pk - (double)n / 2.0) * bw
                        : p.center_hz + (sw / sp - (double)n / 2.0) * bw;
        }
        g.center_hz = c;
        g.bw_hz = (double)(last - first + 1) * bw;
        a.out[(size_t)f * a.cap + i] = g;
    }
}

}  // namespace

extern "C" GPUD_API int gpud_surv_finish_device(gpud_surv* s, const float* psd_dev, int n_frames,
                                                const gpud_surv_run* runs_dev, const int* run_counts_dev,
                                                int max_runs, const gpud_detector_params* params,
                                                gpud_surv_signal* signals_dev, int cap_per_frame, int* counts_dev)
{
    if (!s || !params || n_frames < 0 || max_runs <= 0 || max_runs > ARGUS_DETECT_MAX_RUNS || cap_per_frame <= 0)
        return -1;
    if (n_frames == 0) return 0;
    if (!psd_dev || !runs_dev || !run_counts_dev || !signals_dev || !counts_dev) return -1;
    gpud_surv_signal* work = nullptr;
    ARGUS_CHECK(hipMallocAsync((void**)&work, sizeof(gpud_surv_signal) * 2 * (size_t)max_runs * n_frames, s->stream));
    FinishArgs a;
    a.psd = psd_dev;
    a.runs = runs_dev;
    a.run_counts = run_counts_dev;
    a.work = work;
    a.out = signals_dev;
    a.out_counts = counts_dev;
    a.n = s->n;
    a.max_runs = max_runs;
    a.cap = cap_per_frame;
    a.bin_width = (double)s->sample_rate / (double)s->n;
    a.p = *params;
    argus_clear_error();
    argus_detect_finish_kernel<<<n_frames, kFinishThreads, 0, s->stream>>>(a);
    const int err = (int)hipGetLastError();
    hipFreeAsync(work, s->stream);
    return err;
}

extern "C" GPUD_API int gpud_surv_detect(gpud_surv* s, const gpud_float2* iq_dev, int64_t frame_step, int seg_hop,
                                         int n_segs, int n_frames, const gpud_cfar_params* cfar,
                                         const gpud_detector_params* det, gpud_surv_signal* signals,
                                         int cap_per_frame, int* counts, float* psd_host)
{
    if (!s || !iq_dev || !cfar || !det || !signals || !counts || cap_per_frame <= 0) return -1;
    if (n_frames <= 0) return 0;
    const int n = s->n;
    const int max_runs = ARGUS_DETECT_MAX_RUNS;
    float* d_psd = nullptr;
    gpud_surv_run* d_runs = nullptr;
    int* d_counts = nullptr;
    gpud_surv_signal* d_sig = nullptr;
    ARGUS_CHECK(hipMallocAsync((void**)&d_psd, sizeof(float) * (size_t)n * n_frames, s->stream));
    ARGUS_CHECK(hipMallocAsync((void**)&d_runs, sizeof(gpud_surv_run) * (size_t)max_runs * n_frames, s->stream));
    ARGUS_CHECK(hipMallocAsync((void**)&d_counts, sizeof(int) * 2 * (size_t)n_frames, s->stream));
    ARGUS_CHECK(hipMallocAsync((void**)&d_sig, sizeof(gpud_surv_signal) * (size_t)cap_per_frame * n_frames,
                               s->stream));
    int* d_out_counts = d_counts + n_frames;
    int err = gpud_surv_welch_device(s, iq_dev, frame_step, seg_hop, n_segs, n_frames, d_psd);
    if (!err) err = gpud_surv_cfar_runs_device(s, d_psd, n_frames, cfar, d_runs, d_counts, max_runs);
    if (!err)
        err = gpud_surv_finish_device(s, d_psd, n_frames, d_runs, d_counts, max_runs, det, d_sig, cap_per_frame,
                                      d_out_counts);
    if (!err)
        err = (int)hipMemcpyAsync(counts, d_out_counts, sizeof(int) * (size_t)n_frames, hipMemcpyDeviceToHost,
                                  s->stream);
    if (!err) err = (int)hipStreamSynchronize(s->stream);
    if (!err) {
        // the signals of every frame in one strided copy, as many rows as the
        // fullest frame holds
        int rows = 0;
        for (int f = 0; f < n_frames; f++) rows = std::max(rows, std::min(counts[f], cap_per_frame));
        const size_t pitch = sizeof(gpud_surv_signal) * (size_t)cap_per_frame;
        if (rows > 0)
            err = (int)hipMemcpy2DAsync(signals, pitch, d_sig, pitch, sizeof(gpud_surv_signal) * (size_t)rows,
                                        (size_t)n_frames, hipMemcpyDeviceToHost, s->stream);
        if (!err && psd_host)
            err = (int)hipMemcpyAsync(psd_host, d_psd, sizeof(float) * (size_t)n * n_frames, hipMemcpyDeviceToHost,
                                      s->stream);
    }
    hipFreeAsync(d_psd, s->stream);
    hipFreeAsync(d_runs, s->stream);
    hipFreeAsync(d_counts, s->stream);
    hipFreeAsync(d_sig, s->stream);
    const hipError_t se = hipStreamSynchronize(s->stream);
    if (!err) err = (int)se;
    return err;
}
