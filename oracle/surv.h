/* surv.h -- oracle surveillance declarations (test infrastructure, see oracle.h). */
#ifndef ARGUS_ORACLE_SURV_H
#define ARGUS_ORACLE_SURV_H
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_CFAR_CA = 0, ORC_CFAR_GOSCA = 1, ORC_CFAR_CASO = 2, ORC_CFAR_OS = 3 };

/* The Detector fields the threshold path reads (detector.go:101-261). */
typedef struct orc_detector_cfg {
    double threshold_db;
    double hysteresis_db;
    double cfar_scale_db;
    double edge_margin_db;
    double max_signal_bw_hz;
    double merge_gap_hz;
    double bin_width;
    double center_hz;
    int nbins;
    int pad;
} orc_detector_cfg;

/* detector.go:82-94 Signal (the fields the detection list carries) */
typedef struct orc_signal {
    int first_bin;
    int last_bin;
    int peak_bin;
    int pad;
    double center_hz;
    double bw_hz;
    double peak_db;
    double snr_db;
    double noise_db;
} orc_signal;

void orc_hann(int n, double* w);
void orc_fft(double* x, int n);
void orc_spectrum(const float* iq, int n, const double* window, int wlen, double* out);
void orc_spectrum_from_fft(const float* X, int n, double* out);
int orc_welch_psd(const float* iq, int n, int seg, double overlap, const double* window, int wlen, double* out);
int orc_cfar_thresholds(int mode, int guard, int train, int rank, double scale_db, int wrap, const double* s, int n,
                        double* out);
double orc_median(const double* v, int n);
int orc_detect_signals(const orc_detector_cfg* d, const double* s, const double* thr, int n, orc_signal* out, int cap,
                       double* noise_global_out);

#ifdef __cplusplus
}
#endif
#endif
