// detect.hip -- host half of the detection list: edge expansion, merging
// and the peak-bounded power centroid over the runs the CFAR kernel found
// (detector.go:483-639, 698-713), plus the whole surveillance pass
// (Welch -> CFAR runs -> detection list) over device-resident IQ.
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

extern "C" GPUD_API int gpud_surv_detect(gpud_surv* s, const gpud_float2* iq_dev, int64_t frame_step, int seg_hop,
                                         int n_segs, int n_frames, const gpud_cfar_params* cfar,
                                         const gpud_detector_params* det, gpud_surv_signal* signals,
                                         int cap_per_frame, int* counts, float* psd_host)
{
    if (!s || !iq_dev || !cfar || !det || !signals || !counts || cap_per_frame <= 0) return -1;
    if (n_frames <= 0) return 0;
    const int n = s->n;
    const int max_runs = 2048;
    float* d_psd = nullptr;
    gpud_surv_run* d_runs = nullptr;
    int* d_counts = nullptr;
    ARGUS_CHECK(hipMallocAsync((void**)&d_psd, sizeof(float) * (size_t)n * n_frames, s->stream));
    ARGUS_CHECK(hipMallocAsync((void**)&d_runs, sizeof(gpud_surv_run) * (size_t)max_runs * n_frames, s->stream));
    ARGUS_CHECK(hipMallocAsync((void**)&d_counts, sizeof(int) * (size_t)n_frames, s->stream));
    int err = gpud_surv_welch_device(s, iq_dev, frame_step, seg_hop, n_segs, n_frames, d_psd);
    if (!err) err = gpud_surv_cfar_runs_device(s, d_psd, n_frames, cfar, d_runs, d_counts, max_runs);
    std::vector<float> psd((size_t)n * n_frames);
    std::vector<gpud_surv_run> runs((size_t)max_runs * n_frames);
    std::vector<int> rc((size_t)n_frames);
    if (!err) {
        ARGUS_CHECK(hipMemcpyAsync(psd.data(), d_psd, psd.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream));
        ARGUS_CHECK(hipMemcpyAsync(runs.data(), d_runs, runs.size() * sizeof(gpud_surv_run), hipMemcpyDeviceToHost,
                                   s->stream));
        ARGUS_CHECK(hipMemcpyAsync(rc.data(), d_counts, rc.size() * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    }
    hipFreeAsync(d_psd, s->stream);
    hipFreeAsync(d_runs, s->stream);
    hipFreeAsync(d_counts, s->stream);
    ARGUS_CHECK(hipStreamSynchronize(s->stream));
    if (err) return err;
    for (int f = 0; f < n_frames; f++) {
        const int nr = std::min(rc[(size_t)f], max_runs);
        const int m = gpud_surv_finish_frame(psd.data() + (size_t)f * n, n, s->sample_rate,
                                             runs.data() + (size_t)f * max_runs, nr, det,
                                             signals + (size_t)f * cap_per_frame, cap_per_frame);
        if (m < 0) return m;
        counts[f] = m;
    }
    if (psd_host) std::copy(psd.begin(), psd.end(), psd_host);
    return 0;
}
