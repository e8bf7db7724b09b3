// demod.hip -- demodulation epilogue over every channel of a block in one
// launch.  Semantics follow the reference CPU demodulators:
//   FM  : internal/demod/fm.go:179-196 atan2 of conj(prev)*curr, bridged
//         across blocks by the previous block's last sample
//         (internal/recorder/streamer.go:1287-1302);
//   AM  : internal/demod/am.go:11-27 |x| minus the block mean;
//   USB/LSB: internal/demod/ssb.go:19-39 Re(x e^{j phi}), phi pre-incremented
//         by 2 pi (+-700 Hz)/fs, carried across blocks;
//   CW  : internal/demod/cw.go:11-26 Re(x) cos(phi), +700 Hz.
//
// A channel whose baseband the fast-convolution kernel also demodulated
// (FM, SSB, CW) needs here only its FM seams: the first output of each
// 3584-sample segment, whose predecessor belongs to another wavefront's
// segment (or, at output 0, is the carried sample), and the new carry.
#include "demod.hpp"
#include "fastconv.hpp"

namespace {

// The FM outputs at the segment seams of a channel the fast kernel
// demodulated: output m_s = ceil((s hop - e0) / D), the first of segment s.
__device__ void fm_seams(const ArgusDemodDesc& d, int e0, int D)
{
    for (int sg = threadIdx.x;; sg += blockDim.x) {
        const int m = max(0, (sg * ARGUS_FAST_HOP - e0 + D - 1) / D);
        if (m >= d.n) break;
        if (m > 0 || d.have_carry) {
            const float2 prev = m > 0 ? d.iq[m - 1] : *d.carry_rd;
            d.audio[m - d.skip] = argus_fm_step(prev, d.iq[m]);
        }
    }
    if (threadIdx.x == 0 && d.n > 0) *d.carry_wr = d.iq[d.n - 1];
}

// The FM outputs m = k * M (the first of each narrow item of the direct
// form, which demodulated the rest) and the carry.
__device__ void fm_seams_every(const ArgusDemodDesc& d, int M)
{
    for (int k = threadIdx.x;; k += blockDim.x) {
        const int64_t m = (int64_t)k * M;
        if (m >= d.n) break;
        if (m > 0 || d.have_carry) {
            const float2 prev = m > 0 ? d.iq[m - 1] : *d.carry_rd;
            d.audio[m - d.skip] = argus_fm_step(prev, d.iq[m]);
        }
    }
    if (threadIdx.x == 0 && d.n > 0) *d.carry_wr = d.iq[d.n - 1];
}

__device__ void am_block(const ArgusDemodDesc& d)
{
    __shared__ double part[ARGUS_DEMOD_TILE / 64];
    double acc = 0.0;
    for (int i = threadIdx.x; i < d.n; i += blockDim.x) {
        const float2 v = d.iq[i];
        acc += (double)sqrtf(v.x * v.x + v.y * v.y);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) part[wave] = acc;
    __syncthreads();
    double mean = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) mean += part[w];
    mean /= (double)(d.n > 0 ? d.n : 1);
    const float fmean = (float)mean;
    for (int i = threadIdx.x; i < d.n; i += blockDim.x) {
        const float2 v = d.iq[i];
        d.audio[i] = sqrtf(v.x * v.x + v.y * v.y) - fmean;
    }
    __syncthreads();
}

}  // namespace

__device__ __forceinline__ void argus_demod_items(const ArgusDemodDesc* __restrict__ descs,
                                                  const ArgusDemodItem* __restrict__ items, int n_items)
{
    for (int w = blockIdx.x; w < n_items; w += gridDim.x) {
        const ArgusDemodItem it = items[w];
        const ArgusDemodDesc d = descs[it.chan];
        if (it.kind == ARGUS_DEMOD_ITEM_SEAMS) {
            fm_seams(d, it.start, it.count);
            continue;
        }
        if (it.kind == ARGUS_DEMOD_ITEM_SEAMS_EVERY) {
            fm_seams_every(d, it.count);
            continue;
        }
        const int end = it.start + it.count;
        switch (d.mode) {
        case ARGUS_DEMOD_NFM:
        case ARGUS_DEMOD_WFM:
            for (int i = it.start + threadIdx.x; i < end; i += blockDim.x) {
                if (i == 0 && !d.have_carry) continue;
                const float2 prev = i > 0 ? d.iq[i - 1] : *d.carry_rd;
                d.audio[i - d.skip] = argus_fm_step(prev, d.iq[i]);
                if (i == d.n - 1) *d.carry_wr = d.iq[i];
            }
            break;
        case ARGUS_DEMOD_AM:
            am_block(d);
            break;
        case ARGUS_DEMOD_USB:
        case ARGUS_DEMOD_LSB:
            for (int i = it.start + threadIdx.x; i < end; i += blockDim.x) {
                float s, c;
                argus_nco(d.bfo0 + d.bfo_inc * (uint64_t)(uint32_t)i, s, c);
                d.audio[i] = argus_ssb_out(d.iq[i], s, c);
            }
            break;
        case ARGUS_DEMOD_CW:
            for (int i = it.start + threadIdx.x; i < end; i += blockDim.x) {
                float s, c;
                argus_nco(d.bfo0 + d.bfo_inc * (uint64_t)(uint32_t)i, s, c);
                d.audio[i] = d.iq[i].x * c;
            }
            break;
        default:
            break;
        }
    }
}

__global__ __launch_bounds__(256) void argus_demod_kernel(const ArgusDemodDesc* __restrict__ descs,
                                                          const ArgusDemodItem* __restrict__ items, int n_items)
{
    argus_demod_items(descs, items, n_items);
}

__global__ __launch_bounds__(256) void argus_demod_hdr_kernel(const ArgusDemodHdr* __restrict__ hdr)
{
    const ArgusDemodHdr h = *hdr;
    argus_demod_items(h.descs, h.items, h.n_items);
}

int argus_launch_demod(const ArgusDemodDesc* d_desc, const ArgusDemodItem* d_items, int n_items,
                       hipStream_t stream)
{
    if (n_items <= 0) return 0;
    const int grid = n_items < 2048 ? n_items : 2048;
    argus_clear_error();
    hipLaunchKernelGGL(argus_demod_kernel, dim3(grid), dim3(256), 0, stream, d_desc, d_items, n_items);
    return (int)hipGetLastError();
}

int argus_launch_demod_hdr(const ArgusDemodHdr* d_hdr, hipStream_t stream)
{
    argus_clear_error();
    hipLaunchKernelGGL(argus_demod_hdr_kernel, dim3(ARGUS_DEMOD_GRID), dim3(256), 0, stream, d_hdr);
    return (int)hipGetLastError();
}
