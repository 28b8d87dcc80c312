// Host-side entry points of the HIP kernels (compiled in stockham.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include <atomic>
#include <stdint.h>

#include "fft_kernel_args.h"

namespace asdfft {

// Kernel variants (mapping / prologue / epilogue combinations that are compiled).
enum Variant : int {
    kVarRowsC2C = 0,
    kVarColsC2C = 1,
    kVarR2CEven = 2,
    kVarC2REven = 3,
    kVarR2COdd = 4,
    kVarC2ROdd = 5,
    kVarColsTwiddle = 6,
    kVarRowsTranspose = 7,
    kVarDirect = 8,  // dft_direct_kernel
    // separated real/imaginary planes (ASCEND_FFT_C2C_SEP plans)
    kVarRowsSep = 9,            // rows, planar in and out
    kVarRowsSepIn = 10,         // rows, planar in, complex out
    kVarColsSepOut = 11,        // columns, complex in, planar out
    kVarColsTwiddleSep = 12,    // four-step pass 1 with a planar input
    kVarRowsTransposeSep = 13,  // four-step pass 2 with a planar output
    kVarColsToRows = 14,        // strided columns in, each transform stored as a contiguous row
    kVarColsC2CSplit = 15,      // columns C2C exchanging through one LDS plane (specialised kernels only)
    // shifts folded into the transform (asdFftExtSetShift)
    kVarRowsShiftIn = 16,       // rows, input shift
    kVarRowsShiftBoth = 17,     // rows, input and output shift
    kVarColsShiftOut = 18,      // columns, output shift
    kVarColsShiftOutSplit = 19, // kVarColsShiftOut exchanging through one LDS plane (specialised only)
    kVarRowsC2CSplit = 20,      // rows C2C exchanging through one LDS plane (specialised only)
    kVarRowsPair2 = 21,         // 2D rows in pairs n, n + X/2 with the columns' first radix-2 stage (specialised only)
    kVarRowsC2CTwPre = 22,      // rows C2C with every pass's twiddles loaded at kernel start (specialised only)
    kVarRowsQuad4 = 23,         // 2D rows in fours n + q X/4 with the columns' first radix-4 stage (specialised only)
    kVarRowsQuad4Split = 24,    // kVarRowsQuad4 exchanging through one LDS plane (specialised only)
};

// Dynamic LDS limit set on every kernel that needs more than 64 KiB: the whole LDS of a
// workgroup.  The limit only admits launches, so it is never lowered to a launch's own size
// (launches of different sizes from several threads could otherwise find it lowered).
constexpr int kMaxDynamicLds = 160 * 1024;

// Raises fn's dynamic-LDS limit to kMaxDynamicLds the first time it is launched on the
// current device; `raised` (one per kernel) keeps one bit per device, so later launches make
// no attribute call.
inline hipError_t raise_lds_limit_once(const void* fn, std::atomic<uint64_t>& raised)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const uint64_t bit = uint64_t(1) << (dev & 63);
    if (raised.load(std::memory_order_acquire) & bit) return hipSuccess;
    e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxDynamicLds);
    if (e == hipSuccess) raised.fetch_or(bit, std::memory_order_release);
    return e;
}

struct LaunchShape {
    int static_id = -1;  // index into kStaticInstances, or -1 for the run-time kernel
    int variant;
    int emax;          // registers per thread (complex points): 16, 32 or 48
    int block;         // threads per workgroup
    size_t lds_bytes;  // dynamic LDS
    int64_t grid;      // workgroups
};

hipError_t launch_stockham(const LaunchShape& s, const FftArgs& args, hipStream_t stream);

// Specialised kernels (generated from static_instances.txt by the build).
struct StaticEntry {
    int variant, emax, n, threads, tpb, nradices;
    int radices[kMaxPasses];
    hipError_t (*launch)(const LaunchShape&, const FftArgs&, hipStream_t);
};
extern const StaticEntry kStaticInstances[];
extern const int kNumStaticInstances;

// Index of the specialised kernel for this launch, or -1.  tpb only has to match
// for the column mapping (rows keep transforms-per-workgroup at run time).
int find_static(int variant, int emax, int n, int threads, int tpb, const int* radices, int nradices);
hipError_t launch_dft_direct(const LaunchShape& s, const DftArgs& args, hipStream_t stream);
hipError_t launch_istft_ola(const OlaArgs& args, hipStream_t stream);
hipError_t launch_real_aux(const RealAuxArgs& args, hipStream_t stream);

}  // namespace asdfft

// Compiled (variant, EMAX) instances; each lives in its own translation unit
// (stockham_inst.hip built with -DASDFFT_VARIANT=v -DASDFFT_EMAX=e).
#define ASDFFT_FOR_EACH_INSTANCE(X) \
    X(0, 16) X(1, 16) X(2, 16) X(3, 16) X(4, 16) X(5, 16) X(6, 16) X(7, 16) \
    X(0, 32) X(1, 32) X(2, 32) X(3, 32) X(4, 32) X(5, 32) X(6, 32) X(7, 32) \
    X(0, 48) X(1, 48) X(2, 48) X(3, 48) X(4, 48) X(5, 48) \
    X(9, 16) X(10, 16) X(11, 16) X(12, 16) X(13, 16) \
    X(9, 32) X(10, 32) X(11, 32) X(12, 32) X(13, 32) \
    X(9, 48) X(10, 48) X(11, 48) \
    X(14, 16) X(14, 32) \
    X(16, 16) X(17, 16) X(18, 16) X(16, 32) X(17, 32) X(18, 32) X(16, 48) X(17, 48)

namespace asdfft {
// True when the run-time kernel of (variant, emax) is compiled in.
inline bool has_instance(int variant, int emax)
{
#define ASDFFT_HAS(V, E) \
    if (variant == V && emax == E) return true;
    ASDFFT_FOR_EACH_INSTANCE(ASDFFT_HAS)
#undef ASDFFT_HAS
    return false;
}
}  // namespace asdfft
