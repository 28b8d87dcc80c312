// Lane-butterfly form of short C2C rows: host-side entry points.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace asdfft {

struct WaveFftArgs {
    const float2* in;   // count transforms of n points, in_dist points apart
    float2* out;        // count transforms, out_dist points apart (may equal in)
    const float2* tw;   // W_n^m, m < n
    int64_t count;
    int64_t in_dist, out_dist;
    int n;
    int e, lanes;       // n = e * lanes: points per lane and lanes per transform
    int conj_in, conj_out;
    int vec4;           // set by the launch: 16-byte loads (even e, aligned runs)
};

// Whether the kernel has an instance for length n; sets the points per lane and lanes per
// transform (n = e * lanes, lanes a power of two from 4 to 64).
bool wave_fft_shape(int64_t n, int* e, int* lanes);

// Validates the arguments against the instance table and launches one thread per lane.
hipError_t launch_wave_fft(const WaveFftArgs& a, hipStream_t stream);

}  // namespace asdfft
