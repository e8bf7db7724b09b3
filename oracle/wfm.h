/* wfm.h -- oracle WFM stereo / RDS declarations (test infrastructure, see oracle.h). */
#ifndef ARGUS_ORACLE_WFM_H
#define ARGUS_ORACLE_WFM_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_fir_state {
    double taps[103];
    double delay[103];
    int nt;
    int pad;
} orc_fir_state;

/* streamer.go session fields used by stereoDecodeStateful */
typedef struct orc_stereo_state {
    orc_fir_state lpf, bp_hi, bp_lo, lr_lpf, pilot_hi, pilot_lo;
    double phase, freq, alpha, beta, err_avg, i_state, q_state, lp_alpha;
    int rate;
    int on_count, off_count, enabled;
} orc_stereo_state;

#define ORC_RDS_MAX_GROUPS 4096

/* rds.go:10-31 Decoder state (+ a log of the groups applied) */
typedef struct orc_rds_decoder {
    char ps[8];
    char rt[64];
    uint16_t pi;
    uint16_t pad0;
    double costas_phase, costas_freq;
    double mu;
    float prev[2];
    float prev_dec;
    int has_prev;
    int last_rate;
    int last_hard, has_last_hard;
    int total_decodes, groups_found;
    int n_groups;
    uint16_t groups[ORC_RDS_MAX_GROUPS][4];
} orc_rds_decoder;

/* streaming RDS baseband state (fm.go:234-271 carried across blocks).  The
 * +57 kHz mixer phase of input sample c is 2 pi ((c + 1) * 57000 mod fs) / fs,
 * carried as the exact integer numerator, so it does not drift. */
typedef struct orc_rds_stream {
    double taps[103];
    double hr[103], hi[103];
    long long mix_num;  /* (count * 57000) mod rate */
    long long count;
    int nt, decim, rate, pad;
} orc_rds_stream;

void orc_rds_stream_init(orc_rds_stream* s, int sample_rate);
int orc_rds_stream_process(orc_rds_stream* s, const float* disc, int n, float* out);
void orc_pll_coefficients(double loop_bw, double damping, int sample_rate, double* alpha, double* beta);
void orc_stereo_init(orc_stereo_state* st, int sample_rate);
int orc_stereo_decode(orc_stereo_state* st, const float* mono, int n, float* out);
int orc_stereo_decode_batch(const float* mono, int n, int sample_rate, float* out);
int orc_stereo_pilot_present(const float* iq, int n, int sample_rate, double* ratio_out);
int orc_stereo_pilot_estimate(const float* iq, int n, int sample_rate, double* ratio_out, double* offset_out,
                              double* mags_out);
uint16_t orc_rds_syndrome(uint32_t raw);
void orc_rds_init(orc_rds_decoder* d);
int orc_rds_decode(orc_rds_decoder* d, const float* samples, int n, int sample_rate);
uint32_t orc_rds_block(uint16_t data, uint16_t offset);
int orc_rds_encode_groups(uint16_t pi, const char* ps8, const char* rt64, int n_groups, uint8_t* bits);

#ifdef __cplusplus
}
#endif
#endif
