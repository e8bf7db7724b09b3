// predemod.hip -- the pre-demod decimating FIR of every channel that has one,
// in one launch per block (see predemod.hpp).  An item is ARGUS_PRE_TILE
// outputs of one channel, one thread per output: the output's T inputs
// (block samples, or the carried history before the block) against the
// channel's FP64 taps, accumulated in FP64 in tap order and rounded once to
// complex64, as the reference's delay-line loop does.  A history item writes
// the channel's next T-1 inputs into the other half of a ping-pong pair.
#include "predemod.hpp"

namespace {

__global__ __launch_bounds__(ARGUS_PRE_TILE) void argus_predemod_hdr_kernel(const ArgusPreHdr* __restrict__ hdr)
{
    // separate FP64 roundings in the reference's order: no fused multiply-add
#pragma clang fp contract(off)
    const ArgusPreHdr h = *hdr;
    for (int w = blockIdx.x; w < h.n_items; w += gridDim.x) {
        const ArgusPreItem it = h.items[w];
        const ArgusPreDesc d = h.descs[it.chan];
        const int t = threadIdx.x;
        if (it.kind == 1) {
            // new history: the last T-1 samples of (history ++ block)
            for (int q = t; q < d.T - 1; q += ARGUS_PRE_TILE) {
                const int e = d.n_in + q - (d.T - 1);
                d.hist_wr[q] = e >= 0 ? d.in[e] : d.hist_rd[d.T - 1 + e];
            }
            continue;
        }
        if (t >= it.count) continue;
        const int i = d.first + (it.start + t) * d.factor;
        double ar = 0.0, ai = 0.0;
        for (int k = 0; k < d.T; k++) {
            const int e = i - k;
            const float2 v = e >= 0 ? d.in[e] : d.hist_rd[d.T - 1 + e];
            const double w8 = d.taps[k];
            ar += (double)v.x * w8;
            ai += (double)v.y * w8;
        }
        d.out[it.start + t] = make_float2((float)ar, (float)ai);
    }
}

}  // namespace

int argus_launch_predemod_hdr(const ArgusPreHdr* d_hdr, hipStream_t stream)
{
    argus_clear_error();
    hipLaunchKernelGGL(argus_predemod_hdr_kernel, dim3(ARGUS_PRE_GRID), dim3(ARGUS_PRE_TILE), 0, stream, d_hdr);
    return (int)hipGetLastError();
}
