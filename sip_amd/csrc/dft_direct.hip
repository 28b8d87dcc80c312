// Direct O(N^2) DFT kernel and the launch dispatcher for the Stockham variants.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_kernel_args.h"
#include "kernels.h"

namespace asdfft {

// ------------------------------------------------------------------ direct DFT
// O(N^2) transform for lengths with prime factors outside the codelet set and
// for N == 1.  One workgroup holds tpb transforms in LDS (interleaved float2).
__device__ __forceinline__ float2 dft_load(const DftArgs& a, int64_t base, int i)
{
    const float2* in2 = reinterpret_cast<const float2*>(a.in);
    const int64_t s = a.in_stride;
    float2 x;
    if (a.prologue == kProReal) {
        x = make_float2(reinterpret_cast<const float*>(a.in)[base + i * s], 0.0f);
    } else if (a.prologue == kProPlanar) {
        x = make_float2(reinterpret_cast<const float*>(a.in)[base + i * s], a.in_im[base + i * s]);
    } else if (a.prologue == kProHerm) {
        const int h = a.n >> 1;
        if (i <= h) {
            x = in2[base + i * s];
            if (i == 0 || 2 * i == a.n) x.y = 0.0f;
        } else {
            x = in2[base + (a.n - i) * s];
            x.y = -x.y;
        }
    } else {
        x = in2[base + i * s];
    }
    if (a.conj_in) x.y = -x.y;
    return x;
}

__global__ void __launch_bounds__(256) dft_direct_kernel(const DftArgs a)
{
    extern __shared__ float smem[];
    float2* buf = reinterpret_cast<float2*>(smem);
    const int n = a.n;
    const int nblk = a.tpb * n;
    for (int e = threadIdx.x; e < nblk; e += blockDim.x) {
        const int tt = e / n;
        const int i = e - tt * n;
        const int64_t t = (int64_t)blockIdx.x * a.tpb + tt;
        buf[e] = (t < a.count) ? dft_load(a, (t % a.in_cols) + (t / a.in_cols) * a.in_dist, i) : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    const int nout = (a.epilogue == kEpiHalf) ? n / 2 + 1 : n;
    const float2* w = reinterpret_cast<const float2*>(a.roots);
    for (int e = threadIdx.x; e < a.tpb * nout; e += blockDim.x) {
        const int tt = e / nout;
        const int k = e - tt * nout;
        const int64_t t = (int64_t)blockIdx.x * a.tpb + tt;
        if (t >= a.count) continue;
        float accx = 0.0f, accy = 0.0f;
        int idx = 0;
        for (int i = 0; i < n; ++i) {
            const float2 x = buf[tt * n + i];
            const float2 wi = w[idx];
            accx += x.x * wi.x - x.y * wi.y;
            accy += x.x * wi.y + x.y * wi.x;
            idx += k;
            if (idx >= n) idx -= n;
        }
        if (a.epilogue == kEpiReal) {
            reinterpret_cast<float*>(a.out)[t * a.out_dist + k] = accx;
        } else if (a.epilogue == kEpiPlanar) {
            if (a.conj_out) accy = -accy;
            reinterpret_cast<float*>(a.out)[t * a.out_dist + k] = accx;
            a.out_im[t * a.out_dist + k] = accy;
        } else {
            if (a.conj_out) accy = -accy;
            reinterpret_cast<float2*>(a.out)[t * a.out_dist + k] = make_float2(accx, accy);
        }
    }
}


hipError_t launch_dft_direct(const LaunchShape& s, const DftArgs& args, hipStream_t stream)
{
    if (s.grid <= 0 || s.block <= 0 || s.block > 256) return hipErrorInvalidConfiguration;
    if (s.lds_bytes > 65536) {
        static std::atomic<uint64_t> raised{0};
        hipError_t e = raise_lds_limit_once(reinterpret_cast<const void*>(dft_direct_kernel), raised);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(dft_direct_kernel, dim3((unsigned)s.grid), dim3((unsigned)s.block), s.lds_bytes, stream, args);
    return hipGetLastError();
}

#define ASDFFT_DECL(V, E) hipError_t launch_stockham_v##V##_e##E(const LaunchShape&, const FftArgs&, hipStream_t);
ASDFFT_FOR_EACH_INSTANCE(ASDFFT_DECL)
#undef ASDFFT_DECL

hipError_t launch_stockham(const LaunchShape& s, const FftArgs& args, hipStream_t stream)
{
    if (s.grid <= 0 || s.block <= 0 || s.block > 1024) return hipErrorInvalidConfiguration;
    if (s.static_id >= 0) {
        if (s.static_id >= kNumStaticInstances) return hipErrorInvalidValue;
        return kStaticInstances[s.static_id].launch(s, args, stream);
    }
#define ASDFFT_CASE(V, E) \
    if (s.variant == V && s.emax == E) return launch_stockham_v##V##_e##E(s, args, stream);
    ASDFFT_FOR_EACH_INSTANCE(ASDFFT_CASE)
#undef ASDFFT_CASE
    return hipErrorInvalidValue;
}

int find_static(int variant, int emax, int n, int threads, int tpb, const int* radices, int nradices)
{
    for (int i = 0; i < kNumStaticInstances; i++) {
        const StaticEntry& e = kStaticInstances[i];
        if (e.variant != variant || e.emax != emax || e.n != n || e.threads != threads || e.nradices != nradices) continue;
        if (e.tpb != 0 && e.tpb != tpb) continue;
        bool same = true;
        for (int p = 0; p < nradices; p++) same = same && (e.radices[p] == radices[p]);
        if (same) return i;
    }
    return -1;
}

}  // namespace asdfft
