// Launch parameters shared by the host planner and the HIP kernels.
//
// One launch performs `count` independent transforms of length `n` (complex
// points).  Transform t lives at
//     base(t) = (t % cols) * inner + (t / cols) * outer      (element units)
// for input and output separately, element i at base(t) + i * elem_stride.
// Planar (separated real/imaginary) inputs and outputs use the same element
// offsets in each of their two float planes.
// Row batches use cols = 1; column passes of 2D/3D plans use cols = number of
// adjacent columns, inner = 1.
#pragma once
#include <stdint.h>

namespace asdfft {

constexpr int kMaxPasses = 10;

// Radix codelets compiled into the Stockham kernels.
constexpr int kRadixSet[] = {16, 12, 10, 9, 8, 6, 4, 2, 19, 17, 13, 11, 7, 5, 3};

// Thread mapping inside a workgroup.
enum Mapping : int {
    kMapRows = 0,  // lanes run along a transform (tid fastest): contiguous rows
    kMapCols = 1,  // lanes run across transforms (transform fastest): strided columns
};

// How the first pass obtains its inputs.
enum Prologue : int {
    kProComplex = 0,   // complex load (conjugated when conj_in)
    kProReal = 1,      // real load, imaginary part 0 (odd-length R2C)
    kProHerm = 2,      // Hermitian half -> full spectrum (odd-length C2R)
    kProC2REven = 3,   // even-length C2R pre-twiddle: Z[k] from X[k], X[M-k]
    kProPlanar = 4,    // real and imaginary parts from two float planes (in, in_im)
    kProShift = 5,     // complex load through the input shift (roll_in, in_o_mod / in_o_roll)
};

// How the last pass disposes of its outputs.
enum Epilogue : int {
    kEpiComplex = 0,   // complex store (conjugated when conj_out)
    kEpiHalf = 1,      // store bins 0..n/2 only (odd-length R2C)
    kEpiReal = 2,      // store the real part (odd-length C2R)
    kEpiR2CEven = 3,   // even-length R2C post-twiddle through LDS
    kEpiTwiddle = 4,   // multiply by W_big^(column * i), then store (four-step pass 1)
    kEpiTranspose = 5, // store transposed through LDS (four-step pass 2)
    kEpiPlanar = 6,    // real and imaginary parts to two float planes (out, out_im)
    kEpiTransposePlanar = 7,  // kEpiTranspose into two float planes
    kEpiRowsOut = 8,   // column mapping: each transform stored contiguously through LDS
    kEpiShift = 9,     // complex store through the output shift (roll_out, out_col_roll), times out_scale
    kEpiPairRadix2 = 10,  // rows n and n + o_split of a pair, through LDS: u + v and (u - v) W^n (tw_aux)
    kEpiQuadRadix4 = 11,  // rows n + q o_split, q < 4, through LDS: their 4-point DFT, bin k times W^(n k) (tw_aux)
};

struct PassDesc {
    int radix;        // R
    int span;         // L: product of the radices of the earlier passes
    uint32_t magic;   // floor(2^32 / L) + 1 (L > 1): j / L == umulhi(j, magic)
    int tw_offset;    // first entry of this pass in the twiddle table (layout [k][r-1])
};

struct FftArgs {
    const void* in;
    void* out;
    const float* in_im;     // imaginary plane of a planar input (kProPlanar), else unused
    float* out_im;          // imaginary plane of a planar output (kEpiPlanar, kEpiTransposePlanar)
    const float* tw;        // per-pass twiddles, interleaved (re, im)
    const float* tw_aux;    // R2C/C2R half-length twiddles W_N^k, k in [0, n]
    const float* tw_big_lo; // four-step twiddles: W_big^m = hi[m >> big_shift] * lo[m & mask]
    const float* tw_big_hi;
    int n;                  // complex transform length
    int threads;            // T: threads per transform
    int tpb;                // transforms per workgroup
    int passes;
    int lds_tstride;        // row mapping: padded LDS words per transform
    int lds_plane;          // LDS words per plane (re plane, then im plane)
    int conj_in, conj_out;
    int big_shift;          // four-step twiddle split
    int xcd_remap;          // 1: give each XCD a contiguous range of workgroups (shared L2 lines)
    int stream_in;          // 1: the first pass reads its complex input with non-temporal loads
    int64_t count;          // transforms in this launch
    int64_t cols;
    int64_t in_inner, in_outer, in_stride;
    int64_t out_inner, out_outer, out_stride;
    int64_t out_cols;       // transposed store: transforms per output group (kEpiTranspose)
    // shifts folded into the first / last pass (kProShift, kEpiShift): element i of the
    // transform is loaded from element (i - roll_in) mod n, and stored to (i + roll_out) mod n;
    // the outer index o (t / cols) is read as (o - (o mod in_o_mod) + ((o mod in_o_mod) -
    // in_o_roll) mod in_o_mod) (in_o_mod 1: none); the column index (t mod cols) is stored as
    // (col + out_col_roll) mod cols; stored values are multiplied by out_scale
    int roll_in, roll_out;
    int64_t in_o_mod, in_o_roll, out_col_roll;
    float out_scale;
    // split outer index (o_split > 0): the outer index o = t / cols addresses
    // (o % o_split) * outer + (o / o_split) * outer2 instead of o * outer
    int64_t o_split, in_outer2, out_outer2;
    PassDesc pass[kMaxPasses];
};

// ---- ragged batches: transforms of different lengths and kinds in one launch ----------
// A class is every transform of one (kind, direction, length) in the batch; its FftArgs
// carry the geometry, pass list, twiddle tables and conjugation (in/out/count unused).
struct RaggedClass {
    FftArgs a;
    int variant;   // kVarRowsC2C, kVarR2CEven, kVarC2REven, kVarR2COdd or kVarC2ROdd
    int reserved;
};

// A workgroup's unit of work: `count` (<= the class's tpb) transforms of one class whose
// byte offsets are entries first .. first + count - 1 of the offset table.
struct RaggedTile {
    int cls;
    int count;
    int64_t first;
};

// Kernel arguments besides the class and tile tables (those are separate restrict
// pointers so the uniform descriptor reads can use scalar loads).
struct RaggedArgs {
    const void* in_base;
    void* out_base;
    const int64_t* offsets;  // per transform: {input byte offset, output byte offset}
    int ntiles;
};

// Direct O(N^2) DFT kernel parameters (lengths with factors outside the radix set, N == 1).
struct DftArgs {
    const void* in;
    void* out;
    const float* in_im;     // planar input / output planes (kProPlanar, kEpiPlanar)
    float* out_im;
    const float* roots;     // W_N^m, m in [0, N)
    int n;
    int tpb;
    int conj_in, conj_out;
    int prologue, epilogue; // kProComplex/kProReal/kProHerm/kProPlanar, kEpiComplex/kEpiHalf/kEpiReal/kEpiPlanar
    int64_t count;
    int64_t in_dist, out_dist;
    // input of transform t: (t % in_cols) + (t / in_cols) * in_dist, element i at + i * in_stride
    // (1 and 1 for contiguous rows)
    int64_t in_cols, in_stride;
};

// Elementwise passes that turn a long complex transform into a real one (real_aux.hip).
enum RealAuxOp : int {
    kAuxR2CPost = 0,     // even R2C post-twiddle: Z[0..M) -> X[0..M]
    kAuxC2RPre = 1,      // even C2R pre-twiddle: X[0..M] -> sequence of length M for a forward FFT
    kAuxPromote = 2,     // odd R2C: real -> complex
    kAuxHermExtend = 3,  // odd C2R: X[0..N/2] -> full Hermitian spectrum of length N
    kAuxTruncate = 4,    // odd R2C: bins 0..N/2 of the full spectrum
    kAuxRealPart = 5,    // odd C2R: real parts
};

struct RealAuxArgs {
    const void* in;
    void* out;
    const float* tw;  // W_N^k, k in [0, N/2] (even ops)
    int op;
    int conj;         // conjugate the output (R2C post, truncate) or the input (C2R pre, extension)
    int64_t n;        // real length N
    int64_t count;    // transforms
    int64_t in_dist, out_dist;  // per-transform distance in the op's input / output elements
};

// Overlap-add of inverse-transformed STFT frames (ISTFT): for output sample i of channel b,
//   m = start + i,  y = sum_t frames[b][t][m - t*hop] * window[m - t*hop],
//   env = sum_t window[m - t*hop]^2,  out[b][i] = y * scale / env
// over the frames t with 0 <= m - t*hop < nfft.
struct OlaArgs {
    const void* frames;  // complex [channel][nframes][nfft]
    const float* window; // [nfft]
    void* out;           // complex [channel][out_len]
    int nfft, hop, nframes, start;
    int64_t channel, out_len;
    float scale;         // 1 / nfft
};

}  // namespace asdfft
