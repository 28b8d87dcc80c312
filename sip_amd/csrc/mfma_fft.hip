// Split-FP16 MFMA form of the 1024-point C2C transform (1024 = 32 x 32).
//
// Each wave transforms one row at a time:
//   stage 1  Y[k1][n2]  = sum_n1 F[k1][n1] x[32 n1 + n2]            (32x32x32 complex matmul)
//   twiddle  Y'[k1][n2] = Y[k1][n2] W_1024^(k1 n2)                  (FP32)
//   stage 2  X[k1 + 32 k2] = sum_n2 F[k2][n2] Y'[k1][n2]             (32x32x32 complex matmul)
// with F[k][n] = W_32^(k n).  Each real matmul runs on v_mfma_f32_32x32x16_f16 with both
// operands split into FP16 hi + lo parts and FP32 accumulation:
//   a b ~= a_hi b_hi + a_hi b_lo + a_lo b_hi      (the lo*lo term, ~2^-22 relative, is dropped)
// A power-of-two scale per row (exact) puts the largest input component in [2^9, 2^10), so
// the lo parts stay FP16-normal for components down to 2^-12 of the largest and no FP16
// operand can overflow (stage-1 sums stay below 46 * 2^10 and are scaled by 2^-5 before
// stage 2).  Inverse transforms run as conj(DFT(conj x)).
//
// MFMA lane layout (32x32x16, wave64): operand A holds A[i][k] at lane l for i = l % 32 and
// k = 8 (l / 32) + j, j = 0..7; operand B holds B[k][n] for n = l % 32, same k; the 16
// accumulators hold D[i][n] for n = l % 32, i = 8 (v / 4) + 4 (l / 32) + v % 4.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mfma_fft.h"

namespace asdfft {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f16acc __attribute__((ext_vector_type(16)));

constexpr int kWaves = 4;              // waves per workgroup, one row in flight each
constexpr int kPitch = 33;             // LDS row pitch (floats) of the stage-2 transpose
constexpr int kPlane = 32 * kPitch;    // one 32x32 FP32 plane

struct Split {
    half8 hi, lo;
};

// FP32 -> FP16 hi part (round to nearest even) and the FP16 rounding of the remainder.
__device__ __forceinline__ _Float16 hi_of(float v) { return (_Float16)v; }
__device__ __forceinline__ _Float16 lo_of(float v, _Float16 hi) { return (_Float16)(v - (float)hi); }

__device__ __forceinline__ f16acc mfma(const half8& a, const half8& b, f16acc c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ half8 neg8(half8 v) { return -v; }

// Complex product accumulation: (re, im) += (Fr + i Fi) (Xr + i Xi), split three-term.
__device__ __forceinline__ void cmatmul(const Split& fr, const Split& fi, const Split& xr, const Split& xi,
                                        f16acc& re, f16acc& im)
{
    const half8 xih_n = neg8(xi.hi), xil_n = neg8(xi.lo);
    re = mfma(fr.hi, xr.hi, re);
    re = mfma(fr.hi, xr.lo, re);
    re = mfma(fr.lo, xr.hi, re);
    re = mfma(fi.hi, xih_n, re);
    re = mfma(fi.hi, xil_n, re);
    re = mfma(fi.lo, xih_n, re);
    im = mfma(fr.hi, xi.hi, im);
    im = mfma(fr.hi, xi.lo, im);
    im = mfma(fr.lo, xi.hi, im);
    im = mfma(fi.hi, xr.hi, im);
    im = mfma(fi.hi, xr.lo, im);
    im = mfma(fi.lo, xr.hi, im);
}

__device__ __forceinline__ float wave_max(float v)
{
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

__global__ void __launch_bounds__(64 * kWaves) mfma_fft1024_kernel(const MfmaFftArgs a)
{
    __shared__ float2 tw[1024];              // W_1024^m
    __shared__ float tr[kWaves][2 * kPlane];  // per-wave transpose buffer (re plane, im plane)
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    for (int m = threadIdx.x; m < 1024; m += blockDim.x) tw[m] = a.tw1024[m];

    // A operands: F[i][k] for i = lane % 32, k = 16 c + 8 (lane / 32) + j
    Split fr[2], fi[2];
    for (int c = 0; c < 2; c++) {
        for (int j = 0; j < 8; j++) {
            const int k = 16 * c + 8 * (lane >> 5) + j;
            const float2 w = a.w32[((lane & 31) * k) & 31];
            const _Float16 rh = hi_of(w.x), ih = hi_of(w.y);
            fr[c].hi[j] = rh;
            fr[c].lo[j] = lo_of(w.x, rh);
            fi[c].hi[j] = ih;
            fi[c].lo[j] = lo_of(w.y, ih);
        }
    }
    __syncthreads();
    const float csign_in = a.conj_in ? -1.0f : 1.0f;
    const float csign_out = a.conj_out ? -1.0f : 1.0f;
    float* tre = tr[wave];
    float* tim = tr[wave] + kPlane;
    const int64_t rows_per_step = (int64_t)gridDim.x * kWaves;
    for (int64_t row = (int64_t)blockIdx.x * kWaves + wave; row < a.count; row += rows_per_step) {
        const float2* x = a.in + row * 1024;
        // ---- load: B operand of stage 1, B[n1][n2] = x[32 n1 + n2], n2 = lane % 32
        float2 v[2][8];
        for (int c = 0; c < 2; c++) {
            for (int j = 0; j < 8; j++) {
                const int n1 = 16 * c + 8 * (lane >> 5) + j;
                float2 t = x[32 * n1 + (lane & 31)];
                t.y *= csign_in;
                v[c][j] = t;
            }
        }
        float m = 0.0f;
        for (int c = 0; c < 2; c++) {
            for (int j = 0; j < 8; j++) m = fmaxf(m, fmaxf(fabsf(v[c][j].x), fabsf(v[c][j].y)));
        }
        m = wave_max(m);
        // scale 2^(9 - e) with m in [2^e, 2^(e+1)): the largest component lands in [2^9, 2^10)
        const int e = (m > 0.0f) ? ilogbf(m) : 0;
        const float s = ldexpf(1.0f, 9 - e);
        Split xr[2], xi[2];
        for (int c = 0; c < 2; c++) {
            for (int j = 0; j < 8; j++) {
                const float re = v[c][j].x * s, im = v[c][j].y * s;
                const _Float16 rh = hi_of(re), ih = hi_of(im);
                xr[c].hi[j] = rh;
                xr[c].lo[j] = lo_of(re, rh);
                xi[c].hi[j] = ih;
                xi[c].lo[j] = lo_of(im, ih);
            }
        }
        // ---- stage 1
        f16acc yr = {}, yi = {};
        for (int c = 0; c < 2; c++) cmatmul(fr[c], fi[c], xr[c], xi[c], yr, yi);
        // ---- twiddle (and 2^-5), then transpose through LDS: tre/tim[k1][n2]
        for (int q = 0; q < 16; q++) {
            const int k1 = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
            const int n2 = lane & 31;
            const float2 w = tw[(k1 * n2) & 1023];
            const float re = yr[q] * w.x - yi[q] * w.y;
            const float im = yr[q] * w.y + yi[q] * w.x;
            tre[k1 * kPitch + n2] = re * 0.03125f;
            tim[k1 * kPitch + n2] = im * 0.03125f;
        }
        // this wave's transpose writes complete before its lanes read other lanes' values
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        // ---- stage 2 B operand: B[n2][k1] = Y'[k1][n2], k1 = lane % 32, n2 = 16 c + 8 (lane / 32) + j
        for (int c = 0; c < 2; c++) {
            for (int j = 0; j < 8; j++) {
                const int n2 = 16 * c + 8 * (lane >> 5) + j;
                const int at = (lane & 31) * kPitch + n2;
                const float re = tre[at], im = tim[at];
                const _Float16 rh = hi_of(re), ih = hi_of(im);
                xr[c].hi[j] = rh;
                xr[c].lo[j] = lo_of(re, rh);
                xi[c].hi[j] = ih;
                xi[c].lo[j] = lo_of(im, ih);
            }
        }
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // reads done before the next row's writes
        f16acc zr = {}, zi = {};
        for (int c = 0; c < 2; c++) cmatmul(fr[c], fi[c], xr[c], xi[c], zr, zi);
        // ---- store: Z^T[k2][k1] -> X[k1 + 32 k2], k1 = lane % 32
        float2* y = a.out + row * 1024;
        const float back = 32.0f / s;
        for (int q = 0; q < 16; q++) {
            const int k2 = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
            y[(lane & 31) + 32 * k2] = make_float2(zr[q] * back, zi[q] * back * csign_out);
        }
    }
}

}  // namespace

hipError_t launch_mfma_fft1024(const MfmaFftArgs& args, int64_t grid, hipStream_t stream)
{
    if (grid <= 0 || args.count <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mfma_fft1024_kernel, dim3((unsigned)grid), dim3(64 * kWaves), 0, stream, args);
    return hipGetLastError();
}

int mfma_fft1024_waves() { return kWaves; }

}  // namespace asdfft
