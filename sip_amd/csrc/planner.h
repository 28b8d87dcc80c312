// gfx950 FFT planner: turns a plan description (type, direction, sizes, batch,
// layout) into a list of kernel launches.
//
// Replaces the reference's plan framework (core/fft/fft_api.cpp addFFTSteps
// :662-766, init2DSteps :796-897, init3DSteps :938-992, route selectors
// core/fft/fft1d_optimized_selection.h, core/fft/fft_mixed_selection.h and
// ops/include/params/fft_*_dispatch.h).  Instead of Ascend route tables it
// factors each axis length into radix-{16,12,10,9,8,6,4,2} and odd-prime
// {3,5,7,11,13,17,19} Stockham passes fused into one workgroup, and composes
// multi-dimensional plans from row and strided column passes with no transposes.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "fft_kernel_args.h"
#include "kernels.h"
#include "mfma_fft.h"
#include "twiddles.h"
#include "wave_fft.h"

namespace asdfft {

enum class Kind { C2C, R2C, C2R };

// Largest complex length one workgroup transforms (all of it resident in LDS).
constexpr int64_t kMaxSingleKernel = 16384;
// Largest length the O(N^2) direct kernel accepts (lengths with prime factors > 19).
constexpr int64_t kMaxDirect = 8192;

struct Schedule {
    std::vector<int> radices;
    int emax = 0;     // complex points held per thread
    int threads = 0;  // threads per transform
};

// Radix sequence with the fewest passes (then the cheapest butterflies); empty
// when n has a prime factor outside {2,3,5,7,11,13,17,19} or n == 1.
std::vector<int> choose_radices(int64_t n);

// Radix multisets (listed in descending order) whose product is n, with at most
// `extra` passes more than the fewest possible; sorted by passes, then by whether
// every radix is <= 16, then by butterfly cost.  At most `limit` of them.
std::vector<std::vector<int>> radix_multisets(int64_t n, int extra, int limit);

// choose_radices(n), or when that sequence needs more threads per transform than a
// workgroup holds, the fewest-pass alternative multiset that fits.
std::vector<int> fitting_radices(int64_t n);

// Threads per transform needed to cover every pass with EMAX/R butterflies per
// thread, or 0 when this EMAX cannot hold a radix of the sequence or the
// transform does not fit the EMAX's workgroup limit.
int threads_for(int64_t n, const std::vector<int>& radices, int emax);

int max_block_for(int emax);

enum Buffer : int { kBufIn = 0, kBufOut = 1, kBufWork = 2 };

// Arithmetic of a plan's kernels.
enum class Precision {
    kFp32 = 0,     // FP32 Stockham passes (default)
    kSplit16 = 1,  // split-FP16 MFMA large-radix stages where a kernel exists (C2C N = 1024 rows, and the
                   // 1024-point pass of four-step lengths N = 1024 * n2, n2 a multiple of 16)
};

struct Step {
    bool direct = false;
    bool mfma = false;       // split-FP16 MFMA kernel (margs, mfma_grid)
    bool aux = false;        // elementwise real pass (raux, real_aux.hip)
    bool wave = false;       // lane-butterfly kernel for short rows (wargs, wave_fft.hip)
    WaveFftArgs wargs{};
    RealAuxArgs raux{};
    MfmaFftArgs margs{};
    int64_t mfma_grid = 0;
    LaunchShape shape{};
    FftArgs args{};
    DftArgs dargs{};
    int src = kBufIn;
    int dst = kBufOut;
    bool planar_in = false;   // src is a pair of float planes (real, imaginary)
    bool planar_out = false;  // dst is a pair of float planes
    size_t src_offset = 0;  // bytes into the buffer
    size_t dst_offset = 0;
    std::vector<TablePtr> tables;
    bool tables_ready = true;  // false when no device was present at plan time
    std::string describe;
    // algorithmic bytes moved per launch (read + write of the data it transforms)
    double bytes = 0.0;
};

// Shift folded into a plan (asdFftExtSetShift): the permutation of asdFftShift / asdIfftShift
// on every FFT axis.
enum ShiftMode : int { kShiftNone = 0, kShiftFft = 1, kShiftIfft = 2 };

struct PlanSpec {
    Kind kind = Kind::C2C;
    bool inverse = false;
    std::vector<int64_t> dims;  // FFT sizes; the last one is the contiguous axis
    int64_t batch = 1;
    bool vertical = false;      // 1D VERTICAL layout: tensor (N, batch)
    bool planar = false;        // C2C on separated real / imaginary float planes (C2C_SEP)
    int in_shift = kShiftNone;  // applied to the input before the transform
    int out_shift = kShiftNone; // applied to the transform's output
    float out_scale = 1.0f;     // output multiplier (float32)
    Precision precision = Precision::kFp32;
};

struct BuiltPlan {
    std::vector<Step> steps;
    size_t workspace = 0;  // bytes of caller workspace needed
};

// Returns "" on success, otherwise the reason the plan is unsupported.
std::string build_plan(const PlanSpec& spec, BuiltPlan* out);

// ISTFT frame transforms: unnormalised inverse C2C of length nfft along the strided frequency
// axis of a (channel, nfft, nframes) complex spectrum, frame t of channel b stored as row
// b * nframes + t of a (channel, nframes, nfft) complex workspace (one launch, src kBufIn,
// dst kBufWork).  Returns "" on success, otherwise the reason.
std::string build_istft_frames(int64_t nfft, int64_t channel, int64_t nframes, BuiltPlan* out);

// Launch of one step on `stream`: bufs = {input, output, workspace}; planar steps take the
// imaginary planes from in_im / out_im.
hipError_t launch_step(const Step& s, char* const bufs[3], const float* in_im, float* out_im, hipStream_t stream);

// Human-readable description of the steps (ASDFFT_PLAN_TRACE=1 prints it).
std::string describe_plan(const BuiltPlan& plan);

}  // namespace asdfft
