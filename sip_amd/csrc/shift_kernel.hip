// Batched 2D circular shift of complex64 slices (asdFftShift / asdIfftShift):
//   out[s][(r + shift_r) % row][(c + shift_c) % col] = scale * in[s * dist + r * col + c]
// written as a gather over the contiguous output, so stores are fully coalesced and each
// output row reads at most two contiguous runs of its source row.  One read and one write
// per element: the kernel is HBM-bound (16 bytes per element).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "shift.h"

namespace asdfft {

namespace {

__global__ void __launch_bounds__(256) shift2d_kernel(const ShiftArgs a)
{
    // workgroup: output rows q0 .. q0 + rows_per_block - 1 (global row q = s * row + r);
    // the host keeps row, col and rows_per_block * col below 2^31, so the per-element index
    // arithmetic is 32-bit and the only 64-bit division is the workgroup's first row
    const int64_t q0 = (int64_t)blockIdx.x * a.rows_per_block;
    const int64_t rows_total = a.batch * a.row;
    const int nrows = (int)((q0 + a.rows_per_block <= rows_total) ? a.rows_per_block : rows_total - q0);
    const int row = (int)a.row;
    const int col = (int)a.col;
    const int64_t s0 = q0 / a.row;
    const int r0 = (int)(q0 - s0 * a.row);
    const int sr = (int)a.shift_r, sc = (int)a.shift_c;
    const int elems = nrows * col;
    for (int e = threadIdx.x; e < elems; e += blockDim.x) {
        const int lr = e / col;
        const int cd = e - lr * col;    // destination column
        const int rr = r0 + lr;
        const int sadd = rr / row;      // slices crossed inside this workgroup
        const int rd = rr - sadd * row; // destination row within its slice
        int rs = rd - sr;
        if (rs < 0) rs += row;
        int cs = cd - sc;
        if (cs < 0) cs += col;
        const int64_t s = s0 + sadd;
        float2 v = a.in[s * a.dist + (int64_t)rs * col + cs];
        if (a.apply_scale) {
            v.x *= a.scale;
            v.y *= a.scale;
        }
        a.out[(q0 + lr) * (int64_t)col + cd] = v;
    }
}

}  // namespace

hipError_t launch_shift2d(const ShiftArgs& args, hipStream_t stream)
{
    if (args.row <= 0 || args.col <= 0 || args.batch <= 0 || args.rows_per_block <= 0) return hipErrorInvalidValue;
    const int64_t rows_total = args.batch * args.row;
    const int64_t grid = (rows_total + args.rows_per_block - 1) / args.rows_per_block;
    if (grid > 0x7fffffff) return hipErrorInvalidValue;
    hipLaunchKernelGGL(shift2d_kernel, dim3((unsigned)grid), dim3(256), 0, stream, args);
    return hipGetLastError();
}

}  // namespace asdfft
