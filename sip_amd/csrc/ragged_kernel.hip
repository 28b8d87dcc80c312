// Persistent ragged-batch kernel: one launch transforms tiles of different lengths and
// kinds (C2C, even/odd R2C and C2R rows) taken from a device-side work table.
//
// Work table: tiles[k] = {class, transforms in the tile, first offset-table entry};
// classes[c] holds the Stockham pass list and geometry of one (kind, direction, length).
// Tiles are sorted by decreasing cost on the host, so the largest start first and the
// small ones fill the tail.  The kernel keeps no state between launches: graph replays and
// concurrent launches of one plan need no reset.
#include "ragged.h"
#include "stockham_impl.h"

namespace asdfft {

namespace {

// Lanes past tpb * T of a fixed-size workgroup own no LDS block; this thread index lies
// past every pass's butterfly range, so they issue no loads, LDS accesses or stores but
// still reach every barrier.
constexpr int kIdleThread = 1 << 20;

template <int EMAX, int PRO, int EPI>
__device__ __forceinline__ void ragged_tile(const RaggedArgs& r, const FftArgs& a, int count, int64_t first,
                                            float* smem)
{
    const Geo g{__builtin_amdgcn_readfirstlane(a.n), __builtin_amdgcn_readfirstlane(a.threads),
                __builtin_amdgcn_readfirstlane(a.tpb), __builtin_amdgcn_readfirstlane(a.lds_tstride)};
    float* lre = smem;
    float* lim = smem + __builtin_amdgcn_readfirstlane(a.lds_plane);
    int tt = (int)threadIdx.x / g.T;  // rows mapping: lanes along a transform
    int tid = (int)threadIdx.x - tt * g.T;
    const bool lane_ok = tt < g.tpb;
    const bool live = lane_ok && tt < count;
    if (!lane_ok) {
        tt = 0;
        tid = kIdleThread;
    }
    Io io;
    io.in = r.in_base;
    io.out = r.out_base;
    io.in_im = nullptr;  // ragged classes are interleaved complex or real
    io.out_im = nullptr;
    io.col = 0;
    io.cin = __builtin_amdgcn_readfirstlane(a.conj_in) ? 0x80000000u : 0u;
    io.cout = __builtin_amdgcn_readfirstlane(a.conj_out) ? 0x80000000u : 0u;
    io.in_stride = 1;
    io.out_stride = 1;
    if (live) {
        const int64_t* o = r.offsets + 2 * (first + tt);
        io.in = static_cast<const char*>(r.in_base) + o[0];
        io.out = static_cast<char*>(r.out_base) + o[1];
    }
    const int passes = __builtin_amdgcn_readfirstlane(a.passes);
    for (int p = 0; p < passes; ++p) {
        switch (__builtin_amdgcn_readfirstlane(a.pass[p].radix)) {
#define ASDFFT_CASE(RR) \
    case RR: maybe_pass<EMAX, RR, kMapRows, PRO, EPI, true>(a, g, p, lre, lim, tt, tid, live, io); break;
            ASDFFT_CASE(16)
            ASDFFT_CASE(12)
            ASDFFT_CASE(10)
            ASDFFT_CASE(9)
            ASDFFT_CASE(8)
            ASDFFT_CASE(6)
            ASDFFT_CASE(4)
            ASDFFT_CASE(2)
            ASDFFT_CASE(19)
            ASDFFT_CASE(17)
            ASDFFT_CASE(13)
            ASDFFT_CASE(11)
            ASDFFT_CASE(7)
            ASDFFT_CASE(5)
            ASDFFT_CASE(3)
#undef ASDFFT_CASE
            default:
                break;
        }
    }
    lds_epilogue<kMapRows, EPI>(a, g, lre, lim, tt, tid, 0, live, io);
}

template <int EMAX>
__global__ void __launch_bounds__(ragged_max_block(EMAX)) __attribute__((amdgpu_waves_per_eu(min_waves<EMAX>())))
ragged_kernel(const RaggedArgs r, const RaggedClass* __restrict__ classes, const RaggedTile* __restrict__ tiles)
{
    extern __shared__ float smem[];
    // Workgroup w runs tiles w, w + G, w + 2G, ... of the cost-sorted table (G = gridDim.x).
    // With G = tiles (the default) every workgroup runs one tile and the hardware
    // dispatcher, which starts workgroups in index order as CUs free up, balances the load;
    // a smaller G makes the launch persistent.  No atomics: many workgroups incrementing one
    // queue word serialise in L2.
    for (int k = blockIdx.x; k < r.ntiles; k += gridDim.x) {
        const int cls = __builtin_amdgcn_readfirstlane(tiles[k].cls);
        const int count = __builtin_amdgcn_readfirstlane(tiles[k].count);
        const int64_t first = tiles[k].first;
        const RaggedClass& c = classes[cls];
        switch (__builtin_amdgcn_readfirstlane(c.variant)) {
            case kVarRowsC2C:
                ragged_tile<EMAX, kProComplex, kEpiComplex>(r, c.a, count, first, smem);
                break;
            case kVarR2CEven:
                ragged_tile<EMAX, kProComplex, kEpiR2CEven>(r, c.a, count, first, smem);
                break;
            case kVarC2REven:
                ragged_tile<EMAX, kProC2REven, kEpiComplex>(r, c.a, count, first, smem);
                break;
            case kVarR2COdd:
                ragged_tile<EMAX, kProReal, kEpiHalf>(r, c.a, count, first, smem);
                break;
            case kVarC2ROdd:
                ragged_tile<EMAX, kProHerm, kEpiReal>(r, c.a, count, first, smem);
                break;
            default:
                break;
        }
        // LDS reuse across tiles: the next tile's first LDS write sits behind its first
        // exchange barrier, which every wave reaches only after this tile's LDS reads;
        // a barrier here covers the epilogue reads of a tile followed by a single-pass one
        lds_barrier();
    }
}

}  // namespace

// Entry points of this translation unit's register budget (the build compiles this file
// once per EMAX in {16, 32, 48}; ragged.cpp dispatches on the budget).
#ifndef ASDFFT_RAGGED_EMAX
#error "ASDFFT_RAGGED_EMAX must be defined"
#endif
#define ASDFFT_RNAME2(B, E) B##_e##E
#define ASDFFT_RNAME(B, E) ASDFFT_RNAME2(B, E)

const void* ASDFFT_RNAME(ragged_kernel_fn, ASDFFT_RAGGED_EMAX)()
{
    return reinterpret_cast<const void*>(&ragged_kernel<ASDFFT_RAGGED_EMAX>);
}

hipError_t ASDFFT_RNAME(ragged_prepare, ASDFFT_RAGGED_EMAX)(size_t lds_bytes, int block, int* per_cu)
{
    const void* fn = reinterpret_cast<const void*>(&ragged_kernel<ASDFFT_RAGGED_EMAX>);
    if (lds_bytes > 65536) {
        // the kernel's static LDS (the tile index) comes out of the same 160 KiB
        hipFuncAttributes attr{};
        hipError_t e = hipFuncGetAttributes(&attr, fn);
        if (e != hipSuccess) return e;
        const size_t limit = (size_t)kMaxDynamicLds - attr.sharedSizeBytes;
        if (lds_bytes > limit) return hipErrorInvalidValue;
        e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)limit);
        if (e != hipSuccess) return e;
    }
    int n = 0;
    const hipError_t e =
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, ragged_kernel<ASDFFT_RAGGED_EMAX>, block, lds_bytes);
    if (e != hipSuccess) return e;
    *per_cu = n;
    return hipSuccess;
}

hipError_t ASDFFT_RNAME(launch_ragged, ASDFFT_RAGGED_EMAX)(int64_t grid, int block, size_t lds_bytes,
                                                            const RaggedArgs& args, const RaggedClass* classes,
                                                            const RaggedTile* tiles, hipStream_t stream)
{
    if (block <= 0 || block > ragged_max_block(ASDFFT_RAGGED_EMAX) || block % 64 != 0 || grid <= 0)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(ragged_kernel<ASDFFT_RAGGED_EMAX>, dim3((unsigned)grid), dim3((unsigned)block), lds_bytes,
                       stream, args, classes, tiles);
    return hipGetLastError();
}

}  // namespace asdfft
