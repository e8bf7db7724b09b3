// primitives.hip -- device/memory/stream helpers and the primitive launchers
// that succeed the reference's single-purpose kernels (exports.cu:18-326).
// Every elementwise kernel is a grid-stride loop over 256-thread blocks with
// coalesced complex64 access; the NCO uses the fixed-point turn accumulator.
#include "common.hpp"

namespace {

constexpr int kBlock = 256;

inline int grid_for(int n)
{
    int g = (n + kBlock - 1) / kBlock;
    return g < 8192 ? g : 8192;
}

__global__ void k_freq_shift(const float2* __restrict__ in, float2* __restrict__ out, int n, uint64_t inc,
                             uint64_t phase0)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float s, c;
        argus_nco(phase0 + inc * (uint64_t)(uint32_t)i, s, c);
        out[i] = argus_rot(in[i], s, c);
    }
}

__global__ void k_fm_discrim(const float2* __restrict__ in, float* __restrict__ out, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n - 1; i += gridDim.x * blockDim.x) {
        const float2 p = in[i], c = in[i + 1];
        out[i] = atan2f(p.x * c.y - p.y * c.x, p.x * c.x + p.y * c.y);
    }
}

__global__ void k_decimate(const float2* __restrict__ in, float2* __restrict__ out, int n_out, int factor)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_out; i += gridDim.x * blockDim.x)
        out[i] = in[(size_t)i * factor];
}

__constant__ float c_fir_taps[256];

// Causal FIR over an LDS tile with a (T-1)-sample halo; zero before sample 0.
template <bool kConstTaps>
__global__ __launch_bounds__(kBlock) void k_fir(const float2* __restrict__ in, float2* __restrict__ out,
                                                const float* __restrict__ taps, int n, int T)
{
    extern __shared__ float2 tile[];
    const int base = blockIdx.x * kBlock;
    const int halo = T - 1;
    for (int i = threadIdx.x; i < kBlock + halo; i += kBlock) {
        const int src = base - halo + i;
        tile[i] = (src >= 0 && src < n) ? in[src] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    const int g = base + threadIdx.x;
    if (g >= n) return;
    float ar = 0.0f, ai = 0.0f;
    const float2* s = tile + threadIdx.x + halo;
    for (int k = 0; k < T; ++k) {
        const float t = kConstTaps ? c_fir_taps[k] : taps[k];
        const float2 v = s[-k];
        ar = fmaf(t, v.x, ar);
        ai = fmaf(t, v.y, ai);
    }
    out[g] = make_float2(ar, ai);
}

__global__ void k_am_envelope(const float2* __restrict__ in, float* __restrict__ out, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float2 v = in[i];
        out[i] = sqrtf(v.x * v.x + v.y * v.y);
    }
}

__global__ void k_ssb_product(const float2* __restrict__ in, float* __restrict__ out, int n, uint64_t inc,
                              uint64_t phase0)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float s, c;
        argus_nco(phase0 + inc * (uint64_t)(uint32_t)i, s, c);
        const float2 v = in[i];
        out[i] = v.x * c - v.y * s;
    }
}

int launch_fir(bool const_taps, const float2* in, float2* out, const float* taps, int n, int T, hipStream_t st)
{
    if (n <= 0 || T <= 0 || T > 256) return 0;
    const size_t lds = (size_t)(kBlock + T - 1) * sizeof(float2);
    const int grid = (n + kBlock - 1) / kBlock;
    argus_clear_error();
    if (const_taps)
        hipLaunchKernelGGL(k_fir<true>, dim3(grid), dim3(kBlock), lds, st, in, out, taps, n, T);
    else
        hipLaunchKernelGGL(k_fir<false>, dim3(grid), dim3(kBlock), lds, st, in, out, taps, n, T);
    return (int)hipGetLastError();
}

}  // namespace

// Stream-ordered launchers of the demodulation primitives, shared by the C
// entry points below (default stream) and the legacy engine (its own stream).
int argus_launch_fm_discrim(const float2* in, float* out, int n, hipStream_t st)
{
    if (n <= 1) return 0;
    argus_clear_error();
    hipLaunchKernelGGL(k_fm_discrim, dim3(grid_for(n)), dim3(kBlock), 0, st, in, out, n);
    return (int)hipGetLastError();
}

int argus_launch_am_envelope(const float2* in, float* out, int n, hipStream_t st)
{
    if (n <= 0) return 0;
    argus_clear_error();
    hipLaunchKernelGGL(k_am_envelope, dim3(grid_for(n)), dim3(kBlock), 0, st, in, out, n);
    return (int)hipGetLastError();
}

int argus_launch_ssb_product(const float2* in, float* out, int n, double phase_inc, double phase_start,
                             hipStream_t st)
{
    if (n <= 0) return 0;
    argus_clear_error();
    hipLaunchKernelGGL(k_ssb_product, dim3(grid_for(n)), dim3(kBlock), 0, st, in, out, n,
                       argus_rad_to_fix(phase_inc), argus_rad_to_fix(phase_start));
    return (int)hipGetLastError();
}

extern "C" {

GPUD_API const char* gpud_version(void) { return "argus-hip 0.1 (gfx950)"; }

GPUD_API int gpud_event_create(gpud_event_handle* ev)
{
    if (!ev) return -1;
    hipEvent_t e;
    hipError_t err = hipEventCreate(&e);
    if (err != hipSuccess) return (int)err;
    *ev = (gpud_event_handle)e;
    return 0;
}
GPUD_API int gpud_event_destroy(gpud_event_handle ev) { return ev ? (int)hipEventDestroy((hipEvent_t)ev) : 0; }
GPUD_API int gpud_event_record(gpud_event_handle ev, gpud_stream_handle stream)
{
    if (!ev) return -1;
    return (int)hipEventRecord((hipEvent_t)ev, (hipStream_t)stream);
}
GPUD_API int gpud_event_elapsed_ms(gpud_event_handle start, gpud_event_handle stop, float* ms)
{
    if (!start || !stop || !ms) return -1;
    hipError_t e = hipEventSynchronize((hipEvent_t)stop);
    if (e != hipSuccess) return (int)e;
    return (int)hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop);
}

GPUD_API int gpud_device_count(int* count)
{
    if (!count) return -1;
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        *count = 0;
        return (int)e;
    }
    *count = c;
    return 0;
}

GPUD_API int gpud_malloc(void** ptr, size_t bytes)
{
    if (!ptr) return -1;
    return (int)hipMalloc(ptr, bytes ? bytes : 1);
}
GPUD_API int gpud_free(void* ptr) { return ptr ? (int)hipFree(ptr) : 0; }
GPUD_API int gpud_host_alloc(void** ptr, size_t bytes)
{
    if (!ptr) return -1;
    return (int)hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault);
}
GPUD_API int gpud_host_free(void* ptr) { return ptr ? (int)hipHostFree(ptr) : 0; }
GPUD_API int gpud_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
    return bytes ? (int)hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) : 0;
}
GPUD_API int gpud_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
    return bytes ? (int)hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) : 0;
}
GPUD_API int gpud_memcpy_d2d(void* dst, const void* src, size_t bytes)
{
    return bytes ? (int)hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) : 0;
}
GPUD_API int gpud_memset_zero(void* dst, size_t bytes) { return bytes ? (int)hipMemset(dst, 0, bytes) : 0; }
GPUD_API int gpud_device_sync(void) { return (int)hipDeviceSynchronize(); }

GPUD_API int gpud_stream_create(gpud_stream_handle* out)
{
    if (!out) return -1;
    hipStream_t s;
    hipError_t e = hipStreamCreate(&s);
    if (e != hipSuccess) return (int)e;
    *out = (gpud_stream_handle)s;
    return 0;
}
GPUD_API int gpud_stream_destroy(gpud_stream_handle stream)
{
    return stream ? (int)hipStreamDestroy((hipStream_t)stream) : 0;
}
GPUD_API int gpud_stream_sync(gpud_stream_handle stream)
{
    return stream ? (int)hipStreamSynchronize((hipStream_t)stream) : (int)hipDeviceSynchronize();
}

GPUD_API int gpud_launch_freq_shift_stream_hip(const gpud_float2* in, gpud_float2* out, int n, double phase_inc,
                                               double phase_start, gpud_stream_handle stream)
{
    if (n <= 0) return 0;
    argus_clear_error();
    hipLaunchKernelGGL(k_freq_shift, dim3(grid_for(n)), dim3(kBlock), 0, (hipStream_t)stream, (const float2*)in,
                       (float2*)out, n, argus_rad_to_fix(phase_inc), argus_rad_to_fix(phase_start));
    return (int)hipGetLastError();
}
GPUD_API int gpud_launch_freq_shift_hip(const gpud_float2* in, gpud_float2* out, int n, double phase_inc,
                                        double phase_start)
{
    return gpud_launch_freq_shift_stream_hip(in, out, n, phase_inc, phase_start, 0);
}
GPUD_API int gpud_launch_fm_discrim_hip(const gpud_float2* in, float* out, int n)
{
    return argus_launch_fm_discrim((const float2*)in, out, n, 0);
}
GPUD_API int gpud_upload_fir_taps_hip(const float* taps, int n)
{
    if (!taps || n <= 0 || n > 256) return -1;
    return (int)hipMemcpyToSymbol(HIP_SYMBOL(c_fir_taps), taps, (size_t)n * sizeof(float));
}
GPUD_API int gpud_launch_fir_hip(const gpud_float2* in, gpud_float2* out, int n, int num_taps)
{
    return launch_fir(true, (const float2*)in, (float2*)out, nullptr, n, num_taps, 0);
}
GPUD_API int gpud_launch_fir_stream_hip(const gpud_float2* in, gpud_float2* out, int n, int num_taps,
                                        gpud_stream_handle stream)
{
    return launch_fir(true, (const float2*)in, (float2*)out, nullptr, n, num_taps, (hipStream_t)stream);
}
GPUD_API int gpud_launch_fir_v2_stream_hip(const gpud_float2* in, gpud_float2* out, const float* taps, int n,
                                           int num_taps, gpud_stream_handle stream)
{
    if (!taps) return n <= 0 ? 0 : -1;
    return launch_fir(false, (const float2*)in, (float2*)out, taps, n, num_taps, (hipStream_t)stream);
}
GPUD_API int gpud_launch_decimate_stream_hip(const gpud_float2* in, gpud_float2* out, int n_out, int factor,
                                             gpud_stream_handle stream)
{
    if (n_out <= 0 || factor <= 0) return 0;
    argus_clear_error();
    hipLaunchKernelGGL(k_decimate, dim3(grid_for(n_out)), dim3(kBlock), 0, (hipStream_t)stream,
                       (const float2*)in, (float2*)out, n_out, factor);
    return (int)hipGetLastError();
}
GPUD_API int gpud_launch_decimate_hip(const gpud_float2* in, gpud_float2* out, int n_out, int factor)
{
    return gpud_launch_decimate_stream_hip(in, out, n_out, factor, 0);
}
GPUD_API int gpud_launch_am_envelope_hip(const gpud_float2* in, float* out, int n)
{
    return argus_launch_am_envelope((const float2*)in, out, n, 0);
}
GPUD_API int gpud_launch_ssb_product_hip(const gpud_float2* in, float* out, int n, double phase_inc,
                                         double phase_start)
{
    return argus_launch_ssb_product((const float2*)in, out, n, phase_inc, phase_start, 0);
}

}  // extern "C"
