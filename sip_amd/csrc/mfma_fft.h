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
};

hipError_t launch_mfma_fft1024(const MfmaFftArgs& args, int64_t grid, hipStream_t stream);
int mfma_fft1024_waves();  // rows in flight per workgroup

}  // namespace asdfft
