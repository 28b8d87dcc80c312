// Split-FP16 MFMA transforms (precision form): host-side entry points.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace asdfft {

struct MfmaFftArgs {
    const float2* in;      // count rows of 1024 complex points
    float2* out;
    const float2* w32;     // W_32^m, m < 32
    const float2* tw1024;  // W_1024^m, m < 1024
    int64_t count;
    int conj_in, conj_out;
    // transposed output (second pass of a four-step transform): row t = g * out_rows + k writes
    // element m to out[g * out_rows * 1024 + k + out_rows * m]; out_rows is a multiple of
    // mfma_fft1024_tr_rows() and divides count, and the grid is count / mfma_fft1024_tr_rows()
    int transposed;
    int64_t out_rows;
};

hipError_t launch_mfma_fft1024(const MfmaFftArgs& args, int64_t grid, hipStream_t stream);
int mfma_fft1024_waves();    // rows in flight per workgroup (contiguous rows)
int mfma_fft1024_tr_rows();  // rows per workgroup tile (transposed output)

}  // namespace asdfft
