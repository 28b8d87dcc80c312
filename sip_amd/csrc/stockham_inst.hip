// One (variant, EMAX) instance of the Stockham kernel; the build compiles this
// file once per entry of ASDFFT_FOR_EACH_INSTANCE so the instances build in parallel.
#include "stockham_impl.h"

#ifndef ASDFFT_VARIANT
#error "ASDFFT_VARIANT must be defined"
#endif
#ifndef ASDFFT_EMAX
#error "ASDFFT_EMAX must be defined"
#endif

namespace asdfft {
#define ASDFFT_NAME2(V, E) launch_stockham_v##V##_e##E
#define ASDFFT_NAME(V, E) ASDFFT_NAME2(V, E)
hipError_t ASDFFT_NAME(ASDFFT_VARIANT, ASDFFT_EMAX)(const LaunchShape& s, const FftArgs& args, hipStream_t stream)
{
    return launch_variant<ASDFFT_EMAX, ASDFFT_VARIANT>(s, args, stream);
}
}  // namespace asdfft
