// Batched 2D circular shift of complex64 slices (the asdFftShift / asdIfftShift kernels).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace asdfft {

struct ShiftArgs {
    const float2* in;
    float2* out;
    int64_t row, col, batch;
    int64_t dist;            // elements between input slices (>= row * col)
    int64_t shift_r, shift_c;
    float scale;
    int apply_scale;         // 0: pure permutation (bit-exact copy)
    int rows_per_block;      // output rows per workgroup (rows_per_block * col < 2^31)
};

hipError_t launch_shift2d(const ShiftArgs& args, hipStream_t stream);

}  // namespace asdfft
