// Twiddle tables: generated in extended precision on the host, rounded once to
// float32 and kept resident on the device.  Tables are shared between plans
// through a process-wide cache keyed by their content recipe (the reference keeps a
// process-wide coefficient cache too: core/fft/utils/fftensor_cache.cpp:82).
#pragma once
#include <stdint.h>

#include <memory>
#include <string>
#include <vector>

namespace asdfft {

struct DeviceTable {
    float* dev = nullptr;  // interleaved (re, im)
    size_t entries = 0;
    int device = -1;
    ~DeviceTable();
};

using TablePtr = std::shared_ptr<DeviceTable>;

// e^{-2 pi i m / n} for m in [0, count), computed in long double.
std::vector<float> roots_table(int64_t n, int64_t count, int64_t step = 1);

// Per-pass Stockham twiddles for a radix sequence: pass p with span L and radix R
// contributes entries [k][r-1] = W_{L R}^{r k}, k < L, r in 1..R-1.  Returns the
// table and writes each pass's first entry to offsets.
std::vector<float> stockham_table(int n, const std::vector<int>& radices, std::vector<int>* offsets);

// Uploads (or reuses) a table on the current device.  Returns null when no device
// is available (plans can still be validated on a host without a GPU).
TablePtr upload_table(const std::string& key, const std::vector<float>& host);

// True when a HIP device is present and usable.
bool device_available();

}  // namespace asdfft
