// Elementwise passes around a long complex transform that make it a real one (R2C / C2R of
// lengths above one workgroup's reach: even N = 2M with M > 16384, odd N > 16384).
//
//   kAuxR2CPost     even R2C: Z = FFT_M(x[2n] + i x[2n+1]) in natural order ->
//                   X[k] = 1/2 [(Z_k + conj Z_{M-k}) - i W_N^k (Z_k - conj Z_{M-k})], k in [0, M]
//                   (one thread per pair k, M - k: each Z element is read once)
//   kAuxC2RPre      even C2R: X[0..M] -> the length-M sequence whose forward FFT, conjugated,
//                   is the real output read as M complex values (the pre-twiddle of the one-
//                   workgroup C2R kernel, written out)
//   kAuxPromote     odd R2C: real x -> complex (x, 0)
//   kAuxHermExtend  odd C2R: Hermitian half X[0..N/2] -> full spectrum of length N
//   kAuxTruncate    odd R2C: first N/2 + 1 bins of the full spectrum
//   kAuxRealPart    odd C2R: real parts
// Each is HBM-bound: one read and one write of every element.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_kernel_args.h"
#include "kernels.h"

namespace asdfft {

namespace {

constexpr int kAuxBlock = 256;

__device__ __forceinline__ float2 cmul2(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ float2 conj_if2(float2 x, int c) { return c ? make_float2(x.x, -x.y) : x; }

// items per transform for each op
__device__ __forceinline__ int64_t aux_items(const RealAuxArgs& a)
{
    const int64_t m = a.n / 2;
    switch (a.op) {
        case kAuxR2CPost: return m / 2 + 1;
        case kAuxC2RPre: return m;
        case kAuxTruncate: return a.n / 2 + 1;
        default: return a.n;
    }
}

__global__ void __launch_bounds__(kAuxBlock) real_aux_kernel(const RealAuxArgs a)
{
    const int64_t items = aux_items(a);
    const int64_t chunks = (items + kAuxBlock - 1) / kAuxBlock;
    const int64_t t = (int64_t)blockIdx.x / chunks;
    const int64_t i = ((int64_t)blockIdx.x - t * chunks) * kAuxBlock + threadIdx.x;
    if (t >= a.count || i >= items) return;
    const float2* in2 = reinterpret_cast<const float2*>(a.in);
    float2* out2 = reinterpret_cast<float2*>(a.out);
    const float2* w = reinterpret_cast<const float2*>(a.tw);
    switch (a.op) {
        case kAuxR2CPost: {
            const int64_t m = a.n / 2;
            const float2* z = in2 + t * a.in_dist;
            float2* x = out2 + t * a.out_dist;
            const int64_t k = i;
            const int64_t km = (k == 0) ? 0 : m - k;
            const float2 zk = z[k], zm = z[km];
            const float2 wk = w[k], wm = w[m - k];
            const float2 s0 = make_float2(zk.x + zm.x, zk.y - zm.y);
            const float2 d0 = make_float2(zk.x - zm.x, zk.y + zm.y);
            const float2 q0 = cmul2(wk, d0);
            x[k] = conj_if2(make_float2(0.5f * (s0.x + q0.y), 0.5f * (s0.y - q0.x)), a.conj);
            if (m - k != k) {
                const float2 s1 = make_float2(zm.x + zk.x, zm.y - zk.y);
                const float2 d1 = make_float2(zm.x - zk.x, zm.y + zk.y);
                const float2 q1 = cmul2(wm, d1);
                x[m - k] = conj_if2(make_float2(0.5f * (s1.x + q1.y), 0.5f * (s1.y - q1.x)), a.conj);
            }
            break;
        }
        case kAuxC2RPre: {
            const int64_t m = a.n / 2;
            const float2* x = in2 + t * a.in_dist;
            float2 xa = conj_if2(x[i], a.conj);
            float2 xb = conj_if2(x[m - i], a.conj);
            if (i == 0) {
                xa.y = 0.0f;
                xb.y = 0.0f;
            }
            const float2 sm = make_float2(xa.x + xb.x, xa.y - xb.y);  // xa + conj(xb)
            const float2 d = make_float2(xa.x - xb.x, xa.y + xb.y);   // xa - conj(xb)
            const float2 wi = w[i];
            const float2 q = cmul2(make_float2(wi.x, -wi.y), d);      // W_N^{-i} d
            // Z = sm + i q; the forward FFT consumes conj(Z)
            out2[t * a.out_dist + i] = make_float2(sm.x - q.y, -(sm.y + q.x));
            break;
        }
        case kAuxPromote: {
            const float* x = reinterpret_cast<const float*>(a.in) + t * a.in_dist;
            out2[t * a.out_dist + i] = make_float2(x[i], 0.0f);
            break;
        }
        case kAuxHermExtend: {
            const float2* x = in2 + t * a.in_dist;
            const int64_t h = a.n >> 1;
            float2 v;
            if (i <= h) {
                v = x[i];
                if (i == 0 || 2 * i == a.n) v.y = 0.0f;
            } else {
                v = x[a.n - i];
                v.y = -v.y;
            }
            out2[t * a.out_dist + i] = conj_if2(v, a.conj);
            break;
        }
        case kAuxTruncate:
            out2[t * a.out_dist + i] = conj_if2(in2[t * a.in_dist + i], a.conj);
            break;
        default:  // kAuxRealPart
            reinterpret_cast<float*>(a.out)[t * a.out_dist + i] = in2[t * a.in_dist + i].x;
            break;
    }
}

}  // namespace

hipError_t launch_real_aux(const RealAuxArgs& args, hipStream_t stream)
{
    if (args.count <= 0 || args.n <= 1) return hipErrorInvalidValue;
    const int64_t m = args.n / 2;
    int64_t items;
    switch (args.op) {
        case kAuxR2CPost: items = m / 2 + 1; break;
        case kAuxC2RPre: items = m; break;
        case kAuxTruncate: items = args.n / 2 + 1; break;
        default: items = args.n; break;
    }
    const int64_t grid = args.count * ((items + kAuxBlock - 1) / kAuxBlock);
    if (grid > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(real_aux_kernel, dim3((unsigned)grid), dim3(kAuxBlock), 0, stream, args);
    return hipGetLastError();
}

}  // namespace asdfft
