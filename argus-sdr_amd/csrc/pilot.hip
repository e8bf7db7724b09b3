// pilot.hip -- gpud_stereo_pilot_present: StereoPilotPresent /
// estimateWFMPilot (internal/classifier/pll.go:82-131) for a batch of
// signals.
//
// Per signal: the last rate/2 IQ samples, their FP64 FM discriminator
// (fmDemod, pll.go:160-176), and the magnitude of its spectrum at 19 kHz
// +- 1 kHz in 10 Hz steps (goertzelMagnitude, pll.go:178-193, which is
// |sum_i d[i] e^{-j 2 pi f i / fs}|).  Present when max / median > 4.
//
// Two launches:
//   dft    one workgroup per ARGUS_PILOT_CHUNK discriminator samples of a
//          signal: the chunk's discriminator goes to LDS (FP64), then thread
//          b accumulates bin b's partial DFT with a rotator anchored exactly
//          at the chunk start ((f * g) mod fs in integers);
//   reduce one workgroup per signal: partial sums in chunk order, the 201
//          magnitudes, the first maximum and the median (rank 100).
// FP64 throughout, as the reference; the sum is the direct DFT rather than
// the Goertzel recursion, which agrees with it to rounding.
#include "common.hpp"

#include <cstring>
#include <mutex>
#include <vector>

#define ARGUS_PILOT_BINS 201
#define ARGUS_PILOT_CHUNK 1024
#define ARGUS_PILOT_MIN_RATE 40000

namespace {

struct PilotItem {
    const float2* iq;  // first analysed IQ sample of the signal
    int start;         // first discriminator sample of this chunk
    int cnt;
    int rate;
};

struct PilotSig {
    int first_item;
    int n_items;
};

struct PilotOut {
    double ratio;
    double offset_hz;
    int present;
    int pad;
};

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace

__global__ __launch_bounds__(256) void argus_pilot_dft_kernel(const PilotItem* __restrict__ items,
                                                              double2* __restrict__ partial)
{
    __shared__ double ds[ARGUS_PILOT_CHUNK];
    const PilotItem it = items[blockIdx.x];
    for (int j = threadIdx.x; j < it.cnt; j += blockDim.x) {
        const float2 p = it.iq[it.start + j];
        const float2 c = it.iq[it.start + j + 1];
        const double pr = p.x, pi = p.y, cr = c.x, ci = c.y;
        ds[j] = atan2(pr * ci - pi * cr, pr * cr + pi * ci);
    }
    __syncthreads();
    const int b = threadIdx.x;
    if (b >= ARGUS_PILOT_BINS) return;
    const int64_t fs = it.rate;
    const int64_t f = 18000 + 10 * b;  // 19 kHz - 1 kHz + 10 Hz * b
    double s, c, s1, c1;
    sincospi(2.0 * (double)((f * (int64_t)it.start) % fs) / (double)fs, &s, &c);
    sincospi(2.0 * (double)(f % fs) / (double)fs, &s1, &c1);
    double ar = 0.0, ai = 0.0;
    for (int j = 0; j < it.cnt; j++) {
        const double x = ds[j];
        ar = fma(x, c, ar);
        ai = fma(x, s, ai);
        const double cn = fma(c, c1, -s * s1);
        s = fma(s, c1, c * s1);
        c = cn;
    }
    partial[(size_t)blockIdx.x * ARGUS_PILOT_BINS + b] = make_double2(ar, ai);
}

__global__ __launch_bounds__(256) void argus_pilot_reduce_kernel(const PilotSig* __restrict__ sigs,
                                                                 const double2* __restrict__ partial,
                                                                 double* __restrict__ mags, PilotOut* __restrict__ out)
{
    __shared__ double mg[ARGUS_PILOT_BINS];
    __shared__ double med;
    const PilotSig sg = sigs[blockIdx.x];
    const int b = threadIdx.x;
    if (b == 0) med = 0.0;
    if (b < ARGUS_PILOT_BINS) {
        double ar = 0.0, ai = 0.0;
        for (int k = 0; k < sg.n_items; k++) {
            const double2 v = partial[(size_t)(sg.first_item + k) * ARGUS_PILOT_BINS + b];
            ar += v.x;
            ai += v.y;
        }
        const double m = sqrt(ar * ar + ai * ai);
        mg[b] = m;
        mags[(size_t)blockIdx.x * ARGUS_PILOT_BINS + b] = m;
    }
    __syncthreads();
    // medianFloats (pll.go:35-42): element len/2 of the ascending order; ties
    // broken by index so exactly one bin holds each rank
    if (b < ARGUS_PILOT_BINS) {
        const double v = mg[b];
        int rank = 0;
        for (int j = 0; j < ARGUS_PILOT_BINS; j++) {
            const double w = mg[j];
            rank += (w < v) || (w == v && j < b);
        }
        if (rank == ARGUS_PILOT_BINS / 2) med = v;
    }
    __syncthreads();
    if (b == 0) {
        double best = 0.0;
        int bi = -1;
        for (int j = 0; j < ARGUS_PILOT_BINS; j++)
            if (mg[j] > best) {  // first maximum above 0, as the reference's scan
                best = mg[j];
                bi = j;
            }
        PilotOut o;
        o.ratio = med > 0.0 ? best / med : 0.0;
        o.offset_hz = bi < 0 ? 0.0 : (double)(10 * bi - 1000);
        o.present = med > 0.0 && o.ratio > 4.0;
        o.pad = 0;
        out[blockIdx.x] = o;
    }
}

// Workspace of gpud_stereo_pilot_present_noalloc: a file-scope device
// array, one per GPU (each device holds its own copy until the process
// ends), shared by every call on that GPU.  Host code never passes the
// array's name to a kernel or a copy: it uses the device address
// hipGetSymbolAddress gives for the current device, taken once per device.
#define ARGUS_PILOT_SCRATCH_BYTES (16u << 20)
#define ARGUS_PILOT_MAX_DEVICES 64
static __device__ __align__(256) char scratch[ARGUS_PILOT_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static char* g_scratch_ptr[ARGUS_PILOT_MAX_DEVICES] = {};

static int scratch_address(char** out)
{
    int dev = 0;
    ARGUS_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= ARGUS_PILOT_MAX_DEVICES) return -1;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        g_scratch_ptr[dev] = (char*)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// The batch pilot test with its workspace either taken from the
// stream-ordered pool (ws == nullptr) or the caller's ws of ws_cap bytes.
static int pilot_present(const gpud_float2* iq_dev, const int64_t* offsets, const int32_t* rates, int n_signals,
                         int32_t* present, double* ratio, double* offset_hz, double* mags201,
                         gpud_stream_handle stream, char* ws, size_t ws_cap)
{
    if (n_signals < 0 || (n_signals > 0 && (!offsets || !rates || !present))) return -1;
    if (n_signals == 0) return 0;
    if (offsets[0] < 0) return -1;
    std::vector<int> live;
    std::vector<PilotItem> items;
    std::vector<PilotSig> sigs;
    for (int i = 0; i < n_signals; i++) {
        const int64_t n_all = offsets[i + 1] - offsets[i];
        if (n_all < 0) return -1;
        present[i] = 0;
        if (ratio) ratio[i] = 0.0;
        if (offset_hz) offset_hz[i] = 0.0;
        if (mags201)
            for (int b = 0; b < ARGUS_PILOT_BINS; b++) mags201[(size_t)i * ARGUS_PILOT_BINS + b] = 0.0;
        const int rate = rates[i];
        if (rate < ARGUS_PILOT_MIN_RATE) continue;  // pll.go:90-92
        const int64_t n = n_all > rate / 2 ? rate / 2 : n_all;  // pll.go:83-85
        if (n < 2) continue;                                      // fmDemod returns nothing
        const int m = (int)(n - 1);
        if (!iq_dev) return -1;
        const float2* base = (const float2*)iq_dev + offsets[i] + (n_all - n);
        PilotSig sg;
        sg.first_item = (int)items.size();
        for (int s0 = 0; s0 < m; s0 += ARGUS_PILOT_CHUNK)
            items.push_back({base, s0, m - s0 < ARGUS_PILOT_CHUNK ? m - s0 : ARGUS_PILOT_CHUNK, rate});
        sg.n_items = (int)items.size() - sg.first_item;
        sigs.push_back(sg);
        live.push_back(i);
    }
    if (live.empty()) return 0;
    hipStream_t st = (hipStream_t)stream;
    const int nl = (int)live.size();
    const size_t off_sigs = align256(items.size() * sizeof(PilotItem));
    const size_t off_part = off_sigs + align256((size_t)nl * sizeof(PilotSig));
    const size_t off_mags = off_part + align256(items.size() * ARGUS_PILOT_BINS * sizeof(double2));
    const size_t off_out = off_mags + align256((size_t)nl * ARGUS_PILOT_BINS * sizeof(double));
    const size_t bytes = off_out + (size_t)nl * sizeof(PilotOut);
    std::vector<char> h(off_part, 0);
    std::memcpy(h.data(), items.data(), items.size() * sizeof(PilotItem));
    std::memcpy(h.data() + off_sigs, sigs.data(), sigs.size() * sizeof(PilotSig));
    char* d = ws;
    if (ws) {
        if (bytes > ws_cap) return -7;  // the workspace this call needs exceeds the scratch array
    } else {
        ARGUS_CHECK(hipMallocAsync((void**)&d, bytes, st));
    }
    std::vector<double> mg((size_t)nl * ARGUS_PILOT_BINS);
    std::vector<PilotOut> out((size_t)nl);
    int err = (int)hipMemcpyAsync(d, h.data(), off_part, hipMemcpyHostToDevice, st);
    if (!err) {
        argus_clear_error();
        hipLaunchKernelGGL(argus_pilot_dft_kernel, dim3((unsigned)items.size()), dim3(256), 0, st,
                           (const PilotItem*)d, (double2*)(d + off_part));
        err = (int)hipGetLastError();
    }
    if (!err) {
        hipLaunchKernelGGL(argus_pilot_reduce_kernel, dim3((unsigned)nl), dim3(256), 0, st,
                           (const PilotSig*)(d + off_sigs), (const double2*)(d + off_part), (double*)(d + off_mags),
                           (PilotOut*)(d + off_out));
        err = (int)hipGetLastError();
    }
    if (!err)
        err = (int)hipMemcpyAsync(mg.data(), d + off_mags, mg.size() * sizeof(double), hipMemcpyDeviceToHost, st);
    if (!err)
        err = (int)hipMemcpyAsync(out.data(), d + off_out, out.size() * sizeof(PilotOut), hipMemcpyDeviceToHost, st);
    if (!ws) hipFreeAsync(d, st);
    const hipError_t se = hipStreamSynchronize(st);
    if (!err) err = (int)se;
    if (err) return err;
    for (int k = 0; k < nl; k++) {
        const int i = live[k];
        present[i] = out[(size_t)k].present;
        if (ratio) ratio[i] = out[(size_t)k].ratio;
        if (offset_hz) offset_hz[i] = out[(size_t)k].offset_hz;
        if (mags201)
            std::memcpy(mags201 + (size_t)i * ARGUS_PILOT_BINS, mg.data() + (size_t)k * ARGUS_PILOT_BINS,
                        ARGUS_PILOT_BINS * sizeof(double));
    }
    return 0;
}

extern "C" {

GPUD_API int gpud_stereo_pilot_present(const gpud_float2* iq_dev, const int64_t* offsets, const int32_t* rates,
                                       int n_signals, int32_t* present, double* ratio, double* offset_hz,
                                       double* mags201, gpud_stream_handle stream)
{
    return pilot_present(iq_dev, offsets, rates, n_signals, present, ratio, offset_hz, mags201, stream, nullptr, 0);
}

GPUD_API int gpud_stereo_pilot_present_noalloc(const gpud_float2* iq_dev, const int64_t* offsets,
                                               const int32_t* rates, int n_signals, int32_t* present, double* ratio,
                                               double* offset_hz, double* mags201, gpud_stream_handle stream)
{
    char* ws = nullptr;
    const int e = scratch_address(&ws);
    if (e) return e;
    return pilot_present(iq_dev, offsets, rates, n_signals, present, ratio, offset_hz, mags201, stream, ws,
                         ARGUS_PILOT_SCRATCH_BYTES);
}

}  // extern "C"
