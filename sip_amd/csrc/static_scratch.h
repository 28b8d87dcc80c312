// Library-resident device scratch for the exec entry point that allocates nothing
// (asdFftExtExecNoAlloc).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace asdfft {

// Capacity of the per-device scratch array: the largest plan workspace that
// asdFftExtExecNoAlloc accepts (64 MiB; e.g. a 65536-point C2C batch of 128 needs 64 MiB).
constexpr size_t kStaticScratchBytes = size_t(64) << 20;

// Device address of the current device's copy of the scratch array.  The address is looked
// up once per device (with that device current) and kept; nothing is allocated.
hipError_t static_scratch(void** ptr);

}  // namespace asdfft
