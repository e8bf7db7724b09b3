// rds.hip -- host-side RDS group decoder behind gpud_rds_* (no device code).
//
// Follows internal/rds/rds.go Decoder.Decode (78-264): Mueller & Muller
// symbol timing with linear interpolation (gain 0.01), a Costas loop
// (alpha 0.132, beta alpha^2/4), sign(I) hard bits, differential decoding
// carried across calls, then block sync on the 26-bit CRC syndromes
// (generator 0x1B9, offset words A/B/C/C'/D of IEC 62106) with a unique
// 1-bit flip FEC, tried on the direct and then the inverted bit stream, and a
// +-3-bit flywheel around the next expected group boundary (365-430).
// Arithmetic follows the reference's float32/float64 mix.  The diagnostic
// scans of the reference (analyzeStream) change no decoder state and are not
// reproduced.
#include "common.hpp"
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr uint16_t kOffA = 0x0FC, kOffB = 0x198, kOffC = 0x168, kOffCp = 0x350, kOffD = 0x1B4;
constexpr int kBlockBits = 26;
constexpr int kGroupBits = 4 * kBlockBits;
constexpr int kFlywheel = 3;

// Remainder of the 26-bit word modulo the RDS generator x^10+x^8+x^7+x^5+x^4+x^3+1.
uint16_t syndrome(uint32_t word)
{
    for (int bit = 25; bit >= 10; --bit)
        if (word & (1u << bit)) word ^= 0x1B9u << (bit - 10);
    return (uint16_t)(word & 0x3FF);
}

struct Block {
    uint16_t data, offset;
};

bool is_offset(uint16_t s) { return s == kOffA || s == kOffB || s == kOffC || s == kOffCp || s == kOffD; }

// Block at bits[0..26) whose offset word is one of `allowed` (a bitmask over
// the order A, B, C, C', D), with the unique-single-flip correction.
bool block_at(const uint8_t* bits, unsigned allowed, uint16_t* data)
{
    auto ok = [allowed](uint16_t off) {
        const unsigned bit = off == kOffA ? 1u : off == kOffB ? 2u : off == kOffC ? 4u : off == kOffCp ? 8u : 16u;
        return (allowed & bit) != 0;
    };
    uint32_t word = 0;
    for (int i = 0; i < kBlockBits; i++) word = (word << 1) | (bits[i] & 1u);
    uint16_t s = syndrome(word);
    if (is_offset(s) && ok(s)) {
        *data = (uint16_t)(word >> 10);
        return true;
    }
    int found = 0;
    uint16_t cand = 0;
    for (int i = 0; i < kBlockBits; i++) {
        const uint32_t w = word ^ (1u << (25 - i));
        s = syndrome(w);
        if (!is_offset(s) || !ok(s)) continue;
        cand = (uint16_t)(w >> 10);
        if (++found > 1) return false;  // ambiguous correction
    }
    if (found == 1) {
        *data = cand;
        return true;
    }
    return false;
}

bool group_at(const std::vector<uint8_t>& bits, int start, uint16_t g[4])
{
    if (start < 0 || start + kGroupBits > (int)bits.size()) return false;
    const uint8_t* b = bits.data() + start;
    return block_at(b, 1u, &g[0]) && block_at(b + 26, 2u, &g[1]) && block_at(b + 52, 4u | 8u, &g[2]) &&
           block_at(b + 78, 16u, &g[3]);
}

char printable(unsigned v) { return (v < 32 || v > 126) ? ' ' : (char)v; }

}  // namespace

struct gpud_rds_decoder {
    char ps[8];
    char rt[64];
    uint16_t pi = 0;
    double costas_phase = 0, costas_freq = 0;
    double mu = 0;
    float prev_re = 0, prev_im = 0, prev_dec = 0;
    bool have_prev = false;
    int last_rate = 0;
    int last_hard = 0;
    bool have_last_hard = false;
    std::vector<uint16_t> log;  // applied groups, 4 words each

    gpud_rds_decoder()
    {
        std::memset(ps, 0, sizeof ps);
        std::memset(rt, 0, sizeof rt);
    }

    // rds.go:475-502 applyGroup
    void apply(const uint16_t g[4])
    {
        if (g[0] != 0) pi = g[0];
        const int type = (g[1] >> 12) & 0xF;
        const bool version_a = ((g[1] >> 11) & 1) == 0;
        if (type == 0 && version_a) {
            const int at = (g[1] & 3) * 2;
            ps[at] = printable(g[3] >> 8);
            ps[at + 1] = printable(g[3] & 0xFF);
        }
        if (type == 2 && version_a) {
            const int at = (g[1] & 0xF) * 4;
            const unsigned ch[4] = {(unsigned)(g[2] >> 8), (unsigned)(g[2] & 0xFF), (unsigned)(g[3] >> 8),
                                    (unsigned)(g[3] & 0xFF)};
            for (int j = 0; j < 4 && at + j < 64; j++) rt[at + j] = printable(ch[j]);
        }
        log.insert(log.end(), g, g + 4);
    }

    // rds.go:365-430 tryDecode
    int scan(const std::vector<uint8_t>& bits)
    {
        int groups = 0;
        const int nb = (int)bits.size();
        uint16_t g[4];
        for (int i = 0; i + kGroupBits <= nb;) {
            if (!group_at(bits, i, g)) {
                i++;
                continue;
            }
            groups++;
            apply(g);
            int next = i + kGroupBits;
            bool locked = true;
            while (locked && next + kGroupBits <= nb) {
                if (group_at(bits, next, g)) {
                    apply(g);
                    groups++;
                    next += kGroupBits;
                    continue;
                }
                bool hit = false;
                for (int dl = 1; dl <= kFlywheel && !hit; dl++) {
                    if (next - dl >= 0 && group_at(bits, next - dl, g)) {
                        apply(g);
                        groups++;
                        next = next - dl + kGroupBits;
                        hit = true;
                    } else if (next + dl + kGroupBits <= nb && group_at(bits, next + dl, g)) {
                        apply(g);
                        groups++;
                        next = next + dl + kGroupBits;
                        hit = true;
                    }
                }
                locked = hit;
            }
            i = next - kFlywheel > i ? next - kFlywheel : i + 1;
        }
        return groups;
    }

    int decode(const gpud_float2* x, int n, int rate)
    {
        if (n < kGroupBits || rate <= 0) return 0;
        if (last_rate != 0 && last_rate != rate) {
            costas_phase = costas_freq = 0;
            mu = 0;
            prev_re = prev_im = prev_dec = 0;
            have_prev = false;
            last_hard = 0;
            have_last_hard = false;
        }
        last_rate = rate;
        const double sps = (double)rate / 1187.5;
        double m = mu;
        if (m < 0 || m >= sps || std::isnan(m) || std::isinf(m)) m = sps / 2;
        // symbol timing (rds.go:103-150)
        std::vector<float> sr, si;
        sr.reserve((size_t)(n / (int)sps) + 2);
        si.reserve(sr.capacity());
        while (m < (double)(n - 1)) {
            const int k = (int)m;
            const double f = m - (double)k;
            if (k + 1 >= n) break;
            const float re = (float)((double)x[k].x * (1 - f) + (double)x[k + 1].x * f);
            const float im = (float)((double)x[k].y * (1 - f) + (double)x[k + 1].y * f);
            const float dec = re >= 0 ? 1.0f : -1.0f;
            if (have_prev) {
                const double e = (double)(dec - prev_dec) * (double)prev_re - (double)(re - prev_re) * (double)prev_dec;
                m += sps + 0.01 * e;
            } else {
                m += sps;
                have_prev = true;
            }
            prev_dec = dec;
            prev_re = re;
            prev_im = im;
            sr.push_back(re);
            si.push_back(im);
        }
        double res = m - (double)n;
        while (res < 0) res += sps;
        while (res >= sps) res -= sps;
        mu = res;
        const int ns = (int)sr.size();
        if (ns < kGroupBits) return 0;
        // Costas loop and hard decisions (rds.go:157-205)
        const double alpha = 0.132, beta = alpha * alpha / 4.0;
        std::vector<uint8_t> hard((size_t)ns);
        for (int i = 0; i < ns; i++) {
            const float c = (float)std::cos(costas_phase), s = (float)std::sin(costas_phase);
            const float re = sr[(size_t)i] * c + si[(size_t)i] * s;
            const float im = si[(size_t)i] * c - sr[(size_t)i] * s;
            const double err = re >= 0 ? (double)im : -(double)im;
            costas_freq += beta * err;
            costas_phase += costas_freq + alpha * err;
            while (costas_phase > M_PI) costas_phase -= 2 * M_PI;
            while (costas_phase < -M_PI) costas_phase += 2 * M_PI;
            hard[(size_t)i] = re >= 0 ? 1 : 0;
        }
        // differential decoding carried across calls (rds.go:214-233)
        std::vector<uint8_t> bits;
        bits.reserve((size_t)ns);
        if (have_last_hard) bits.push_back(hard[0] ^ (uint8_t)last_hard);
        for (int i = 1; i < ns; i++) bits.push_back(hard[(size_t)i] ^ hard[(size_t)i - 1]);
        last_hard = hard[(size_t)ns - 1];
        have_last_hard = true;
        int found = scan(bits);
        if (found == 0) {
            for (auto& b : bits) b ^= 1;
            found = scan(bits);
        }
        return found;
    }
};

extern "C" {

GPUD_API int gpud_rds_decoder_create(gpud_rds_decoder** out)
{
    if (!out) return -1;
    *out = new gpud_rds_decoder();
    return 0;
}

GPUD_API int gpud_rds_decoder_destroy(gpud_rds_decoder* d)
{
    delete d;
    return 0;
}

GPUD_API int gpud_rds_decode(gpud_rds_decoder* d, const gpud_float2* samples, int n, int sample_rate)
{
    if (!d || n < 0 || (n > 0 && !samples)) return -1;
    return d->decode(samples, n, sample_rate);
}

GPUD_API int gpud_rds_decoder_state(gpud_rds_decoder* d, uint16_t* pi, char* ps9, char* rt65)
{
    if (!d) return -1;
    if (pi) *pi = d->pi;
    auto text = [](const char* src, int len, char* dst) {
        int end = len;
        while (end > 0 && (src[end - 1] == ' ' || src[end - 1] == 0)) end--;
        for (int i = 0; i < end; i++) dst[i] = src[i] ? src[i] : ' ';
        dst[end] = 0;
    };
    if (ps9) text(d->ps, 8, ps9);
    if (rt65) text(d->rt, 64, rt65);
    return 0;
}

GPUD_API int gpud_rds_decoder_groups(gpud_rds_decoder* d, uint16_t* groups4, int cap)
{
    if (!d || cap < 0 || (cap > 0 && !groups4)) return -1;
    const int total = (int)(d->log.size() / 4);
    const int m = total < cap ? total : cap;
    if (m > 0) std::memcpy(groups4, d->log.data(), (size_t)m * 4 * sizeof(uint16_t));
    return total;
}

}  // extern "C"
