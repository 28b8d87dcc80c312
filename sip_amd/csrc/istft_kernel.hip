// Overlap-add stage of the ISTFT: windowed sum of the inverse-transformed frames, the
// squared-window envelope, the 1/nfft scale and the centre crop in one pass.
//
// Thread (b, i) produces output sample i of channel b.  Its frames t are those whose window
// covers m = start + i; for each it reads frames[b][t][m - t*hop] and window[m - t*hop].
// Consecutive lanes take consecutive m, so for one t their frame reads are consecutive complex
// values of one row: every frame element is read once, in coalesced runs, and each output
// sample is written once.  HBM-bound: 8 bytes read per frame element, 8 written per sample.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_kernel_args.h"
#include "kernels.h"

namespace asdfft {

namespace {

constexpr int kOlaBlock = 256;

__global__ void __launch_bounds__(kOlaBlock) istft_ola_kernel(const OlaArgs a)
{
    // workgroup: samples [c * kOlaBlock, (c + 1) * kOlaBlock) of channel b, c = blockIdx.x % chunks
    const int64_t chunks = (a.out_len + kOlaBlock - 1) / kOlaBlock;
    const int64_t b = (int64_t)blockIdx.x / chunks;
    const int64_t i = ((int64_t)blockIdx.x - b * chunks) * kOlaBlock + threadIdx.x;
    if (i >= a.out_len) return;
    const int64_t m = a.start + i;  // position in the full overlap-added signal
    const int nfft = a.nfft, hop = a.hop;
    int64_t t_hi = m / hop;
    if (t_hi > a.nframes - 1) t_hi = a.nframes - 1;
    const int64_t t_lo = (m >= nfft) ? (m - nfft) / hop + 1 : 0;  // first frame with m - t*hop < nfft
    const float2* f = reinterpret_cast<const float2*>(a.frames) + b * (int64_t)a.nframes * nfft;
    float yr = 0.0f, yi = 0.0f, env = 0.0f;
    for (int64_t t = t_lo; t <= t_hi; ++t) {
        const int n = (int)(m - t * hop);
        const float w = a.window[n];
        const float2 v = f[t * nfft + n];
        yr += v.x * w;
        yi += v.y * w;
        env += w * w;
    }
    const float s = a.scale / env;
    reinterpret_cast<float2*>(a.out)[b * a.out_len + i] = make_float2(yr * s, yi * s);
}

}  // namespace

hipError_t launch_istft_ola(const OlaArgs& args, hipStream_t stream)
{
    if (args.channel <= 0 || args.out_len <= 0 || args.nfft <= 0 || args.hop <= 0 || args.nframes <= 0)
        return hipErrorInvalidValue;
    const int64_t chunks = (args.out_len + kOlaBlock - 1) / kOlaBlock;
    const int64_t grid = args.channel * chunks;
    if (grid > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(istft_ola_kernel, dim3((unsigned)grid), dim3(kOlaBlock), 0, stream, args);
    return hipGetLastError();
}

}  // namespace asdfft
