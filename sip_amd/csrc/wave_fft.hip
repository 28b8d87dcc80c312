// Lane-butterfly form of short C2C rows (n = e * lanes, lanes a power of two up to 64).
//
// A transform lives in `lanes` consecutive lanes of one wavefront, e points per lane, and never
// touches LDS memory: lane l holds x[e l + j], j < e (a contiguous run of the row).  With
// n = E L, k = kl + L ke:
//   X[kl + L ke] = sum_j W_E^(j ke) W_n^(j kl) sum_l x[E l + j] W_L^(l kl)
//   1. the inner L-point DFTs run across the lanes as log2(L) radix-2 decimation-in-frequency
//      stages.  At the stage of half-size h a lane exchanges every point with lane l ^ h through
//      DPP (h = 1, 2: quad_perm; h = 8: row_ror:8), ds_swizzle (h = 4, 16: xor within 32 lanes)
//      or ds_bpermute (h = 32), then keeps a + b (bit h of l clear) or (a - b) W_2h^(l mod h)
//      (bit h set), a being the lower lane's point.  Lane l ends with bin kl = bitrev(l).
//   2. twiddle: point j of the lane times W_n^(j kl).
//   3. the E-point DFT in registers (the Stockham codelets); point ke is X[kl + L ke].
// Stores are coalesced: for each ke the lanes of a transform cover L consecutive outputs.
// Inverse transforms run as conj(DFT(conj x)).  No workgroup barrier and no LDS allocation,
// so residency is bounded by registers alone.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "stockham_impl.h"
#include "wave_fft.h"

namespace asdfft {

namespace {

constexpr int kWaveBlock = 256;  // four wavefronts per workgroup

// Value of x held by lane (lane ^ H), H a power of two below 64.
template <int H>
__device__ __forceinline__ float lane_xor(float x)
{
    const int i = __float_as_int(x);
    int r;
    if constexpr (H == 1) {
        r = __builtin_amdgcn_mov_dpp(i, 0xB1, 0xF, 0xF, false);  // quad_perm [1, 0, 3, 2]
    } else if constexpr (H == 2) {
        r = __builtin_amdgcn_mov_dpp(i, 0x4E, 0xF, 0xF, false);  // quad_perm [2, 3, 0, 1]
    } else if constexpr (H == 8) {
        r = __builtin_amdgcn_mov_dpp(i, 0x128, 0xF, 0xF, false);  // row_ror:8 within 16 lanes
    } else if constexpr (H == 4 || H == 16) {
        r = __builtin_amdgcn_ds_swizzle(i, 0x1F | (H << 10));  // bitmask mode: and 31, xor H
    } else {
        static_assert(H == 32, "lane distance");
        r = __builtin_amdgcn_ds_bpermute((int)(((threadIdx.x & 63) ^ 32) << 2), i);
    }
    return __int_as_float(r);
}

// One decimation-in-frequency stage of the cross-lane DFT (half-size H).
template <int E, int H>
__device__ __forceinline__ void lane_stage(float2 (&v)[E], int l, const float2* tw, int n)
{
    const bool up = (l & H) != 0;
    const float s = up ? -1.0f : 1.0f;
    // W_2H^(l mod H) for the upper lane of each pair, W^0 = 1 for the lower one
    const float2 w = tw[up ? (l & (H - 1)) * (n / (2 * H)) : 0];
#pragma unroll
    for (int j = 0; j < E; ++j) {
        const float2 b = make_float2(lane_xor<H>(v[j].x), lane_xor<H>(v[j].y));
        const float2 u = make_float2(fmaf(s, v[j].x, b.x), fmaf(s, v[j].y, b.y));
        v[j] = (H > 1) ? cmul(u, w) : u;
    }
}

template <int E, int H>
__device__ __forceinline__ void lane_stages(float2 (&v)[E], int l, const float2* tw, int n)
{
    if constexpr (H >= 1) {
        lane_stage<E, H>(v, l, tw, n);
        lane_stages<E, H / 2>(v, l, tw, n);
    }
}

template <int E, int L>
__global__ void __launch_bounds__(kWaveBlock) wave_fft_kernel(const WaveFftArgs a)
{
    constexpr int kBits = __builtin_ctz(L);
    const int l = threadIdx.x & (L - 1);
    const int64_t t = ((int64_t)blockIdx.x * kWaveBlock + threadIdx.x) / L;  // this lane's transform
    const bool live = t < a.count;
    float2 v[E];
    if (live) {
        const float2* src = a.in + t * a.in_dist + E * l;
        bool loaded = false;
        if constexpr (E % 2 == 0) {
            if (a.vec4) {  // uniform: every lane's run starts on a 16-byte boundary
                const float4* s4 = reinterpret_cast<const float4*>(src);
#pragma unroll
                for (int j = 0; j < E / 2; ++j) {
                    const float4 q = s4[j];
                    v[2 * j] = make_float2(q.x, q.y);
                    v[2 * j + 1] = make_float2(q.z, q.w);
                }
                loaded = true;
            }
        }
        if (!loaded) {
#pragma unroll
            for (int j = 0; j < E; ++j) v[j] = src[j];
        }
        if (a.conj_in) {
#pragma unroll
            for (int j = 0; j < E; ++j) v[j].y = -v[j].y;
        }
    } else {
        // lanes past the last transform still take part in the exchanges
#pragma unroll
        for (int j = 0; j < E; ++j) v[j] = make_float2(0.0f, 0.0f);
    }
    lane_stages<E, L / 2>(v, l, a.tw, a.n);
    const int kl = (int)(__brev((unsigned)l) >> (32 - kBits));
#pragma unroll
    for (int j = 1; j < E; ++j) v[j] = cmul(v[j], a.tw[j * kl]);  // j kl < n: no reduction
    Dft<E>::run(v);
    if (live) {
        if (a.conj_out) {
#pragma unroll
            for (int j = 0; j < E; ++j) v[j].y = -v[j].y;
        }
        float2* dst = a.out + t * a.out_dist + kl;
#pragma unroll
        for (int k = 0; k < E; ++k) dst[L * k] = v[k];
    }
}

// (points per lane, lanes per transform) of every instance
#define ASDFFT_WAVE_INSTANCES(X)                                                                       \
    X(16, 4) X(16, 8) X(16, 16) X(16, 32) X(16, 64) X(12, 4) X(12, 8) X(12, 16) X(12, 32) X(12, 64) \
    X(10, 4) X(10, 8) X(10, 16) X(10, 32) X(10, 64) X(9, 4) X(9, 8) X(9, 16) X(9, 32) X(9, 64)

constexpr int kWaveE[] = {16, 12, 10, 9};  // preference order: larger e, fewer cross-lane stages

}  // namespace

bool wave_fft_shape(int64_t n, int* e, int* lanes)
{
    for (int E : kWaveE) {
        if (n % E != 0) continue;
        const int64_t L = n / E;
        if (L >= 4 && L <= 64 && (L & (L - 1)) == 0) {
            *e = E;
            *lanes = (int)L;
            return true;
        }
    }
    return false;
}

hipError_t launch_wave_fft(const WaveFftArgs& args, hipStream_t stream)
{
    WaveFftArgs a = args;
    int e = 0, lanes = 0;
    if (!wave_fft_shape(a.n, &e, &lanes) || e != a.e || lanes != a.lanes) return hipErrorInvalidValue;
    if (a.count <= 0 || a.in == nullptr || a.out == nullptr || a.tw == nullptr) return hipErrorInvalidValue;
    if (a.in_dist < a.n || a.out_dist < a.n) return hipErrorInvalidValue;
    a.vec4 = (e % 2 == 0 && reinterpret_cast<uintptr_t>(a.in) % 16 == 0 && a.in_dist % 2 == 0) ? 1 : 0;
    const int64_t threads = a.count * lanes;
    const int64_t grid = (threads + kWaveBlock - 1) / kWaveBlock;
    if (grid > INT32_MAX) return hipErrorInvalidValue;
#define ASDFFT_WAVE_CASE(E, L)                                                                                  \
    if (e == E && lanes == L) {                                                                                 \
        hipLaunchKernelGGL((wave_fft_kernel<E, L>), dim3((unsigned)grid), dim3(kWaveBlock), 0, stream, a); \
        return hipGetLastError();                                                                               \
    }
    ASDFFT_WAVE_INSTANCES(ASDFFT_WAVE_CASE)
#undef ASDFFT_WAVE_CASE
    return hipErrorInvalidValue;
}

}  // namespace asdfft
