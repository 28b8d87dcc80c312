#include "twiddles.h"

#include <hip/hip_runtime.h>
#include <math.h>

#include <map>
#include <mutex>

namespace asdfft {

DeviceTable::~DeviceTable()
{
    if (dev != nullptr) {
        int prev = -1;
        if (hipGetDevice(&prev) == hipSuccess && prev != device) (void)hipSetDevice(device);
        (void)hipFree(dev);
        if (prev >= 0 && prev != device) (void)hipSetDevice(prev);
    }
}

std::vector<float> roots_table(int64_t n, int64_t count, int64_t step)
{
    static const long double kTwoPi = 6.283185307179586476925286766559005768L;
    std::vector<float> out(2 * (size_t)count);
    for (int64_t m = 0; m < count; m++) {
        int64_t e = (m * step) % n;
        long double a = kTwoPi * (long double)e / (long double)n;
        out[2 * m] = (float)cosl(a);
        out[2 * m + 1] = (float)(-sinl(a));
    }
    return out;
}

std::vector<float> stockham_table(int n, const std::vector<int>& radices, std::vector<int>* offsets)
{
    static const long double kTwoPi = 6.283185307179586476925286766559005768L;
    std::vector<float> tab;
    offsets->clear();
    int span = 1;
    for (int r : radices) {
        offsets->push_back((int)(tab.size() / 2));
        if (span > 1) {
            const int64_t base = (int64_t)span * r;  // W_{L R}
            for (int k = 0; k < span; k++) {
                for (int q = 1; q < r; q++) {
                    int64_t e = ((int64_t)q * k) % base;
                    long double a = kTwoPi * (long double)e / (long double)base;
                    tab.push_back((float)cosl(a));
                    tab.push_back((float)(-sinl(a)));
                }
            }
        }
        span *= r;
    }
    (void)n;
    if (tab.empty()) tab.assign(2, 0.0f);  // keep a valid device pointer
    return tab;
}

bool device_available()
{
    static int state = -1;  // -1 unknown, 0 no, 1 yes
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    if (state < 0) {
        int count = 0;
        state = (hipGetDeviceCount(&count) == hipSuccess && count > 0) ? 1 : 0;
        (void)hipGetLastError();
    }
    return state == 1;
}

TablePtr upload_table(const std::string& key, const std::vector<float>& host)
{
    if (!device_available()) return nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    static std::mutex mu;
    static std::map<std::pair<int, std::string>, std::weak_ptr<DeviceTable>> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({dev, key});
    if (it != cache.end()) {
        if (TablePtr p = it->second.lock()) return p;
    }
    auto t = std::make_shared<DeviceTable>();
    t->device = dev;
    t->entries = host.size() / 2;
    if (hipMalloc(reinterpret_cast<void**>(&t->dev), host.size() * sizeof(float)) != hipSuccess) {
        t->dev = nullptr;
        return nullptr;
    }
    if (hipMemcpy(t->dev, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        return nullptr;
    }
    cache[{dev, key}] = t;
    return t;
}

}  // namespace asdfft
