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
//
// Two kernels share the row transform:
//  * mfma_fft1024_kernel: contiguous rows in and out (a 1D plan of length 1024).  Each wave
//    loops over rows with the next two rows' loads in flight while it transforms the current
//    one; the A operands F (the same for every wave) are read from LDS per stage.
//  * mfma_fft1024_tr_kernel: the second pass of a four-step transform whose output is stored
//    transposed: row t = g * R + k (R rows per group) writes element m to
//    out[g * R * 1024 + k + R * m].  A workgroup transforms a tile of 16 consecutive rows of
//    one group in LDS (8 waves, two rows each) and then stores the tile as 128-byte runs.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>

#include "kernels.h"
#include "mfma_fft.h"

namespace asdfft {

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f16acc __attribute__((ext_vector_type(16)));

constexpr int kWaves = 4;              // plain kernel: waves per workgroup, one row in flight each
constexpr int kPitch = 33;             // plain kernel: LDS row pitch (floats) of the stage-2 transpose
constexpr int kPlane = 32 * kPitch;    // one 32x32 FP32 plane
constexpr int kTrWaves = 8;            // transposed kernel: waves per workgroup
constexpr int kTrRows = 16;            // transposed kernel: rows per tile (two per wave)
constexpr int kTrPitch = kTrRows + 1;  // float2 slots per element index m in the tile

// Tile slot of element m of tile row r.  The pitch of 17 and the skew m >> 5 keep both
// stage-2 transpose patterns (fixed k1 with n2 = lane, fixed n2 with k1 = lane; m = 32 k1 + n2)
// on distinct banks, and leave the 16 rows of one m contiguous for the final store.
__device__ __forceinline__ int tr_slot(int m, int r) { return m * kTrPitch + (m >> 5) + r; }
constexpr int kTrSlots = 1023 * kTrPitch + 31 + kTrRows;  // one past the largest slot
constexpr size_t kTrLdsBytes = (size_t)kTrSlots * 8 + 1024 * 8;  // tile + W_1024 table

struct Split {
    half8 hi, lo;
};

// FP32 -> FP16 hi part (round to nearest even) and the FP16 rounding of the remainder.
__device__ __forceinline__ _Float16 hi_of(float v) { return (_Float16)v; }
// v - hi is exact in FP32; written as an FMA so that it maps onto the mixed-precision FMA
// (f16 operand, f16 result) instead of a convert, a subtract and a convert.
__device__ __forceinline__ _Float16 lo_of(float v, _Float16 hi) { return (_Float16)fmaf((float)hi, -1.0f, v); }

__device__ __forceinline__ f16acc mfma(const half8& a, const half8& b, f16acc c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ half8 neg8(half8 v) { return -v; }

// Complex product accumulation: (re, im) += (Fr + i Fi) (Xr + i Xi), split three-term;
// fin = -Fi (negating an FP16 operand is exact).
__device__ __forceinline__ void cmatmul(const Split& fr, const Split& fi, const Split& fin, const Split& xr,
                                        const Split& xi, f16acc& re, f16acc& im)
{
    re = mfma(fr.hi, xr.hi, re);
    re = mfma(fr.hi, xr.lo, re);
    re = mfma(fr.lo, xr.hi, re);
    re = mfma(fin.hi, xi.hi, re);
    re = mfma(fin.hi, xi.lo, re);
    re = mfma(fin.lo, xi.hi, re);
    im = mfma(fr.hi, xi.hi, im);
    im = mfma(fr.hi, xi.lo, im);
    im = mfma(fr.lo, xi.hi, im);
    im = mfma(fi.hi, xr.hi, im);
    im = mfma(fi.hi, xr.lo, im);
    im = mfma(fi.lo, xr.hi, im);
}

// The same product with the negation on the B operand (for callers that hold F, not -Fi, in
// registers).
__device__ __forceinline__ void cmatmul_negx(const Split& fr, const Split& fi, const Split& xr, const Split& xi,
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

typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// hi and lo FP16 parts of the pair (a, b), each packed into one dword (a in the low half).
// hi: one packed round-to-nearest convert.  lo = a - hi(a), b - hi(b): exact in FP32 and rounded
// once to FP16 by the mixed-precision FMAs (f16 operand selected by op_sel, f32 operands),
// two instructions for the pair.
__device__ __forceinline__ void split_pair(float a, float b, uint32_t& hi, uint32_t& lo)
{
    const half2v h = {(_Float16)a, (_Float16)b};
    hi = __builtin_bit_cast(uint32_t, h);
    asm("v_fma_mixlo_f16 %0, %1, -1.0, %2 op_sel:[0,0,0] op_sel_hi:[1,0,0]\n\t"
        "v_fma_mixhi_f16 %0, %1, -1.0, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
        : "=&v"(lo)
        : "v"(hi), "v"(a), "v"(b));
}

// Points j and j + 1 (j even) of the B operands xr / xi.
__device__ __forceinline__ void split_into2(float re0, float im0, float re1, float im1, Split& xr, Split& xi, int j)
{
    u32x4 rh = __builtin_bit_cast(u32x4, xr.hi), rl = __builtin_bit_cast(u32x4, xr.lo);
    u32x4 ih = __builtin_bit_cast(u32x4, xi.hi), il = __builtin_bit_cast(u32x4, xi.lo);
    uint32_t h, l;
    split_pair(re0, re1, h, l);
    rh[j >> 1] = h;
    rl[j >> 1] = l;
    split_pair(im0, im1, h, l);
    ih[j >> 1] = h;
    il[j >> 1] = l;
    xr.hi = __builtin_bit_cast(half8, rh);
    xr.lo = __builtin_bit_cast(half8, rl);
    xi.hi = __builtin_bit_cast(half8, ih);
    xi.lo = __builtin_bit_cast(half8, il);
}

// A operands: F[i][k] = W_32^(i k) for i = lane % 32, k = 16 c + 8 (lane / 32) + j
__device__ __forceinline__ void load_dft_matrix(const float2* w32, int lane, Split (&fr)[2], Split (&fi)[2])
{
    for (int c = 0; c < 2; c++) {
        for (int j = 0; j < 8; j++) {
            const int k = 16 * c + 8 * (lane >> 5) + j;
            const float2 w = w32[((lane & 31) * k) & 31];
            const _Float16 rh = hi_of(w.x), ih = hi_of(w.y);
            fr[c].hi[j] = rh;
            fr[c].lo[j] = lo_of(w.x, rh);
            fi[c].hi[j] = ih;
            fi[c].lo[j] = lo_of(w.y, ih);
        }
    }
}

// Stage-1 B operand of one row: B[n1][n2] = x[32 n1 + n2], n2 = lane % 32 (256-byte runs).
// The lane's base pointer plus compile-time offsets below 4 KiB: one address per 16 loads
// of each half instead of one per load.
__device__ __forceinline__ void load_row(const float2* x, int lane, float2 (&v)[2][8])
{
    const float2* p = x + 256 * (lane >> 5) + (lane & 31);
    for (int c = 0; c < 2; c++) {
        const float2* pc = p + 512 * c;
        for (int j = 0; j < 8; j++) v[c][j] = pc[32 * j];
    }
}

// The 1024-point transform of one row held in v (stage-1 B layout).  xpose_put(k1, n2, re, im)
// and xpose_get(k1, n2, &re, &im) store and fetch the twiddled stage-1 result through LDS
// (wave-private locations); put(mq, re, im) receives output element m = mlane + mq with
// mlane = (lane % 32) + 128 (lane / 32) and mq a compile-time constant per call.  tw holds
// W_1024^m / 32 (the 2^-5 between the stages folded in).  mm(c, xr, xi, re, im) accumulates
// (re, im) += F[c] (xr + i xi) for the A operand block c of the DFT matrix.
template <typename MM, typename XPut, typename XGet, typename Put>
__device__ __forceinline__ void row_fft(float2 (&v)[2][8], MM mm, const float2* tw, int lane, float csign_in,
                                        float csign_out, XPut xpose_put, XGet xpose_get, Put put)
{
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
        for (int j = 0; j < 8; j += 2) {
            split_into2(v[c][j].x * s, v[c][j].y * s * csign_in, v[c][j + 1].x * s, v[c][j + 1].y * s * csign_in,
                        xr[c], xi[c], j);
        }
    }
    // ---- stage 1
    f16acc yr = {}, yi = {};
    for (int c = 0; c < 2; c++) mm(c, xr[c], xi[c], yr, yi);
    // ---- twiddle (tw carries the 2^-5), then transpose through LDS
    for (int q = 0; q < 16; q++) {
        const int k1 = 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
        const int n2 = lane & 31;
        const float2 w = tw[(k1 * n2) & 1023];
        const float re = yr[q] * w.x - yi[q] * w.y;
        const float im = yr[q] * w.y + yi[q] * w.x;
        xpose_put(k1, n2, re, im);
    }
    // this wave's transpose writes complete before its lanes read other lanes' values
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    // ---- stage 2 B operand: B[n2][k1] = Y'[k1][n2], k1 = lane % 32, n2 = 16 c + 8 (lane / 32) + j
    for (int c = 0; c < 2; c++) {
        for (int j = 0; j < 8; j += 2) {
            float re0, im0, re1, im1;
            xpose_get(lane & 31, 16 * c + 8 * (lane >> 5) + j, &re0, &im0);
            xpose_get(lane & 31, 16 * c + 8 * (lane >> 5) + j + 1, &re1, &im1);
            split_into2(re0, im0, re1, im1, xr[c], xi[c], j);
        }
    }
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // reads done before the locations are reused
    f16acc zr = {}, zi = {};
    for (int c = 0; c < 2; c++) mm(c, xr[c], xi[c], zr, zi);
    // ---- output: Z^T[k2][k1] -> X[k1 + 32 k2], k1 = lane % 32
    const float back = 32.0f / s;
    for (int q = 0; q < 16; q++) {
        // k2 = 8 (q / 4) + 4 (lane / 32) + q % 4; m = (lane % 32) + 32 k2
        put(32 * (8 * (q >> 2) + (q & 3)), zr[q] * back, zi[q] * back * csign_out);
    }
}

__global__ void __launch_bounds__(64 * kWaves) mfma_fft1024_kernel(const MfmaFftArgs a)
{
    __shared__ float2 tw[1024];              // W_1024^m
    __shared__ float tr[kWaves][2 * kPlane];  // per-wave transpose buffer (re plane, im plane)
    __shared__ Split sf[2][3][64];            // A operands F[c] (real, imaginary, -imaginary) per lane
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    for (int m = threadIdx.x; m < 1024; m += blockDim.x) {
        const float2 w = a.tw1024[m];
        tw[m] = make_float2(w.x * 0.03125f, w.y * 0.03125f);  // exact: a power of two
    }
    if (wave == 0) {
        // the operands are the same for every wave: kept in LDS and read per stage, so they
        // do not take registers for the whole kernel
        Split fr[2], fi[2];
        load_dft_matrix(a.w32, lane, fr, fi);
        for (int c = 0; c < 2; c++) {
            sf[c][0][lane] = fr[c];
            sf[c][1][lane] = fi[c];
            sf[c][2][lane] = Split{neg8(fi[c].hi), neg8(fi[c].lo)};
        }
    }
    __syncthreads();
    auto mm = [&](int c, const Split& xr, const Split& xi, f16acc& re, f16acc& im) {
        // read at every use: the barrier keeps the compiler from hoisting the loop-invariant
        // reads out of the row loop and holding all of F in registers again
        asm volatile("" ::: "memory");
        const Split fr = sf[c][0][lane], fi = sf[c][1][lane], fin = sf[c][2][lane];
        cmatmul(fr, fi, fin, xr, xi, re, im);
    };
    const float csign_in = a.conj_in ? -1.0f : 1.0f;
    const float csign_out = a.conj_out ? -1.0f : 1.0f;
    float* tre = tr[wave];
    float* tim = tr[wave] + kPlane;
    auto xput = [&](int k1, int n2, float re, float im) {
        tre[k1 * kPitch + n2] = re;
        tim[k1 * kPitch + n2] = im;
    };
    auto xget = [&](int k1, int n2, float* re, float* im) {
        *re = tre[k1 * kPitch + n2];
        *im = tim[k1 * kPitch + n2];
    };
    const int64_t rows_per_step = (int64_t)gridDim.x * kWaves;
    int64_t row = (int64_t)blockIdx.x * kWaves + wave;
    // two rows ahead are in flight while a row is transformed
    float2 next[2][8], next2[2][8];
    if (row < a.count) load_row(a.in + row * 1024, lane, next);
    if (row + rows_per_step < a.count) load_row(a.in + (row + rows_per_step) * 1024, lane, next2);
    for (; row < a.count; row += rows_per_step) {
        float2 v[2][8];
        for (int c = 0; c < 2; c++) {
            for (int j = 0; j < 8; j++) {
                v[c][j] = next[c][j];
                next[c][j] = next2[c][j];
            }
        }
        if (row + 2 * rows_per_step < a.count) load_row(a.in + (row + 2 * rows_per_step) * 1024, lane, next2);
        float2* y = a.out + row * 1024 + (lane & 31) + 128 * (lane >> 5);
        row_fft(v, mm, tw, lane, csign_in, csign_out, xput, xget,
                [&](int mq, float re, float im) { y[mq] = make_float2(re, im); });
    }
}

__global__ void __launch_bounds__(64 * kTrWaves) mfma_fft1024_tr_kernel(const MfmaFftArgs a)
{
    extern __shared__ float2 lds[];
    float2* tile = lds;               // kTrSlots slots: tile rows through the stage-2 transpose and output
    float2* tw = lds + kTrSlots;      // W_1024^m
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    for (int m = threadIdx.x; m < 1024; m += blockDim.x) {
        const float2 w = a.tw1024[m];
        tw[m] = make_float2(w.x * 0.03125f, w.y * 0.03125f);  // exact: a power of two
    }
    Split frr[2], fir[2];
    load_dft_matrix(a.w32, lane, frr, fir);
    __syncthreads();
    auto mm = [&](int c, const Split& xr, const Split& xi, f16acc& re, f16acc& im) {
        cmatmul_negx(frr[c], fir[c], xr, xi, re, im);
    };
    const float csign_in = a.conj_in ? -1.0f : 1.0f;
    const float csign_out = a.conj_out ? -1.0f : 1.0f;
    // tile: rows t0 .. t0 + 15 of group g (R = out_rows rows per group, R % 16 == 0)
    const int64_t t0 = (int64_t)blockIdx.x * kTrRows;
    const int64_t g = t0 / a.out_rows;
    const int64_t k0 = t0 - g * a.out_rows;
    const int r0 = 2 * wave;
    float2 v0[2][8], v1[2][8];
    if (t0 + r0 < a.count) load_row(a.in + (t0 + r0) * 1024, lane, v0);
    if (t0 + r0 + 1 < a.count) load_row(a.in + (t0 + r0 + 1) * 1024, lane, v1);
    for (int h = 0; h < 2; h++) {
        const int r = r0 + h;
        if (t0 + r >= a.count) break;  // uniform per wave
        auto xput = [&](int k1, int n2, float re, float im) { tile[tr_slot(32 * k1 + n2, r)] = make_float2(re, im); };
        auto xget = [&](int k1, int n2, float* re, float* im) {
            const float2 x = tile[tr_slot(32 * k1 + n2, r)];
            *re = x.x;
            *im = x.y;
        };
        auto put = [&](int mq, float re, float im) {
            tile[tr_slot((lane & 31) + 128 * (lane >> 5) + mq, r)] = make_float2(re, im);
        };
        if (h == 0) {
            row_fft(v0, mm, tw, lane, csign_in, csign_out, xput, xget, put);
        } else {
            row_fft(v1, mm, tw, lane, csign_in, csign_out, xput, xget, put);
        }
    }
    __syncthreads();
    // store: out[g * R * 1024 + k0 + r + R * m], 16 consecutive rows r per element m
    const int live = (int)((a.count - t0 < kTrRows) ? a.count - t0 : kTrRows);
    float2* out = a.out + g * a.out_rows * 1024 + k0;
    for (int e = threadIdx.x; e < 1024 * kTrRows; e += 64 * kTrWaves) {
        const int m = e / kTrRows;
        const int r = e - m * kTrRows;
        if (r < live) out[r + a.out_rows * (int64_t)m] = tile[tr_slot(m, r)];
    }
}

}  // namespace

hipError_t launch_mfma_fft1024(const MfmaFftArgs& args, int64_t grid, hipStream_t stream)
{
    if (grid <= 0 || args.count <= 0) return hipErrorInvalidValue;
    if (!args.transposed) {
        hipLaunchKernelGGL(mfma_fft1024_kernel, dim3((unsigned)grid), dim3(64 * kWaves), 0, stream, args);
        return hipGetLastError();
    }
    // one tile of 16 rows per workgroup; tiles never straddle two groups of out_rows rows
    if (args.out_rows <= 0 || args.out_rows % kTrRows != 0 || args.count % args.out_rows != 0 ||
        grid != (args.count + kTrRows - 1) / kTrRows)
        return hipErrorInvalidValue;
    static std::atomic<uint64_t> raised{0};
    hipError_t e = raise_lds_limit_once(reinterpret_cast<const void*>(mfma_fft1024_tr_kernel), raised);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mfma_fft1024_tr_kernel, dim3((unsigned)grid), dim3(64 * kTrWaves), kTrLdsBytes, stream,
                       args);
    return hipGetLastError();
}

int mfma_fft1024_waves() { return kWaves; }
int mfma_fft1024_tr_rows() { return kTrRows; }

}  // namespace asdfft
