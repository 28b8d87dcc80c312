// C++ caller of include/fft_api.h written the way the reference's own callers are
// (example/A2/FFT/1d/example_acl_fft_c2c_1d.cpp:121-141,
//  example/Ascend950/FFT/fftshift/example_acl_fft_shift.cpp:117,
//  example/Ascend950/FFT/ifftshift/example_acl_ifft_shift.cpp:120,
//  tests/ut/unittest/fft/istft/test_fft_istft.cpp:125), plus a multi-threaded section that
// drives the handle table and the exec-time checks from several threads at once.
//
// Everything here stays on the host: plans are made, queried and destroyed, and every
// exec is one that the library rejects before it would touch a device (wrong dtype, wrong
// shape, destroyed handle).  It runs on a machine without a GPU, and under AddressSanitizer,
// UndefinedBehaviorSanitizer or ThreadSanitizer builds of the host library.
//
// Exit status 0 when every check passes; the failing check is printed otherwise.
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "acl_meta.h"
#include "fft_api.h"

using namespace AsdSip;

namespace {

std::atomic<int> g_failures{0};

#define CHECK(cond)                                                             \
    do {                                                                        \
        if (!(cond)) {                                                          \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            g_failures.fetch_add(1);                                            \
        }                                                                       \
    } while (0)

// The reference example's ASD_STATUS_CHECK, counting instead of returning.
#define ASD_STATUS_CHECK(err)                              \
    do {                                                   \
        AsdSip::AspbStatus err_ = (err);                   \
        if (err_ != AsdSip::ErrorType::ACL_SUCCESS) {      \
            std::fprintf(stderr, "%s:%d: status %d\n", __FILE__, __LINE__, (int)err_); \
            g_failures.fetch_add(1);                       \
        }                                                  \
    } while (0)

// Contiguous ND tensor over a dummy device address that is never dereferenced.
aclTensor* make_tensor(const std::vector<int64_t>& shape, aclDataType dt, uintptr_t addr)
{
    std::vector<int64_t> strides(shape.size(), 1);
    for (int64_t i = (int64_t)shape.size() - 2; i >= 0; i--) strides[i] = shape[i + 1] * strides[i + 1];
    return aclCreateTensor(shape.data(), shape.size(), dt, strides.data(), 0, aclFormat::ACL_FORMAT_ND, shape.data(),
                           shape.size(), reinterpret_cast<void*>(addr));
}

void reference_call_shapes()
{
    // example_acl_fft_c2c_1d.cpp:121-141: MakePlan1D with the default dimType argument,
    // workspace size through a size_t reference, stream as void*.
    int64_t Nfft = 4096, batch = 512;
    asdFftHandle handle;
    ASD_STATUS_CHECK(asdFftCreate(handle));
    ASD_STATUS_CHECK(
        asdFftMakePlan1D(handle, Nfft, asdFftType::ASCEND_FFT_C2C, asdFftDirection::ASCEND_FFT_FORWARD, batch));
    size_t work_size = 1;
    ASD_STATUS_CHECK(asdFftGetWorkspaceSize(handle, work_size));
    CHECK(work_size == 0);  // a 4096-point row transform needs no workspace
    void* workspaceAddr = nullptr;
    ASD_STATUS_CHECK(asdFftSetWorkspace(handle, (uint8_t*)workspaceAddr));
    void* stream = reinterpret_cast<void*>(0x10);
    ASD_STATUS_CHECK(asdFftSetStream(handle, stream));
    asdFftType type;
    ASD_STATUS_CHECK(asdFftGetType(handle, type));
    CHECK(type == ASCEND_FFT_C2C);
    // a FLOAT tensor for a C2C plan is rejected before any launch
    aclTensor* wrong = make_tensor({batch, Nfft}, aclDataType::ACL_FLOAT, 0x100000);
    aclTensor* out = make_tensor({batch, Nfft}, aclDataType::ACL_COMPLEX64, 0x40000000);
    CHECK(asdFftExecC2C(handle, wrong, out) == ACL_ERROR_INVALID_PARAM);
    CHECK(asdFftMakePlan1D(handle, Nfft, ASCEND_FFT_C2C, ASCEND_FFT_FORWARD, batch) ==
          ACL_ERROR_INVALID_PARAM);  // a handle is initialised once
    ASD_STATUS_CHECK(asdFftDestroy(handle));
    aclDestroyTensor(wrong);
    aclDestroyTensor(out);

    // example_acl_fft_shift.cpp:117 / example_acl_ifft_shift.cpp:120: seven and eight
    // arguments, the workspace defaulted.  FLOAT tensors are refused before any launch.
    int64_t row = 4, col = 6, sb = 2, dist = 0;
    double scale = 1.0 / (row * col);
    aclTensor* fin = make_tensor({sb, row, col}, aclDataType::ACL_FLOAT, 0x100000);
    aclTensor* fout = make_tensor({sb, row, col}, aclDataType::ACL_FLOAT, 0x200000);
    CHECK(asdFftShift(row, col, sb, dist, fin, fout, stream) == ACL_ERROR_UNSUPPORTED_DATA_TYPE);
    CHECK(asdIfftShift(row, col, sb, dist, scale, fin, fout, stream) == ACL_ERROR_UNSUPPORTED_DATA_TYPE);
    CHECK(asdFftShift(0, col, sb, dist, fin, fout, stream) == ACL_ERROR_INVALID_PARAM);
    aclDestroyTensor(fin);
    aclDestroyTensor(fout);

    // test_fft_istft.cpp:125 spells every argument; the short form takes the reference's
    // defaults (center, not normalized, two-sided, no length, complex output).
    int64_t nFft = 64, hop = 16, frames = 8, channel = 2;
    aclTensor* spec = make_tensor({channel, nFft, frames}, aclDataType::ACL_COMPLEX64, 0x100000);
    asdFftHandle h1, h2;
    ASD_STATUS_CHECK(asdFftCreate(h1));
    ASD_STATUS_CHECK(asdFftCreate(h2));
    ASD_STATUS_CHECK(asdFftIstftMakePlan(h1, spec, nFft, hop, nFft, true, false, false, 0, true));
    ASD_STATUS_CHECK(asdFftIstftMakePlan(h2, spec, nFft, hop, nFft));
    size_t ws1 = 0, ws2 = 0;
    ASD_STATUS_CHECK(asdFftGetWorkspaceSize(h1, ws1));
    ASD_STATUS_CHECK(asdFftGetWorkspaceSize(h2, ws2));
    CHECK(ws1 == ws2 && ws1 == (size_t)(channel * frames * nFft * 8));
    ASD_STATUS_CHECK(asdFftDestroy(h1));
    ASD_STATUS_CHECK(asdFftDestroy(h2));
    aclDestroyTensor(spec);

    // tensor accessors as the reference reads them (view shape released with delete[])
    aclTensor* t3 = make_tensor({3, 5, 7}, aclDataType::ACL_COMPLEX64, 0x100000);
    int64_t* dims = nullptr;
    uint64_t nd = 0;
    CHECK(aclGetViewShape(t3, &dims, &nd) == 0 && nd == 3 && dims[0] == 3 && dims[2] == 7);
    delete[] dims;
    aclDataType dt;
    CHECK(aclGetDataType(t3, &dt) == 0 && dt == ACL_COMPLEX64);
    aclDestroyTensor(t3);
}

// Several threads create, plan, query, mis-execute and destroy handles of their own, while
// others hit one shared handle with concurrent queries and rejected execs, and one thread
// destroys a third handle that others keep calling (once it is gone, exec returns
// ACL_ERROR_INVALID_PARAM and GetType ACL_ERROR_INTERNAL_ERROR, the reference's codes for an
// unknown handle, never a crash).
void concurrent_handles(int threads, int iters)
{
    asdFftHandle shared;
    ASD_STATUS_CHECK(asdFftCreate(shared));
    ASD_STATUS_CHECK(asdFftMakePlan1D(shared, 1000, ASCEND_FFT_C2C, ASCEND_FFT_INVERSE, 64));
    asdFftHandle doomed;
    ASD_STATUS_CHECK(asdFftCreate(doomed));
    ASD_STATUS_CHECK(asdFftMakePlan2D(doomed, 64, 48, ASCEND_FFT_C2C, ASCEND_FFT_FORWARD, 3));
    aclTensor* wrong = make_tensor({64, 1000}, aclDataType::ACL_FLOAT, 0x100000);
    aclTensor* right = make_tensor({64, 1000}, aclDataType::ACL_COMPLEX64, 0x40000000);
    std::atomic<bool> go{false};
    std::atomic<int> distinct_fail{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) {
        pool.emplace_back([&, t]() {
            while (!go.load()) std::this_thread::yield();
            const int64_t sizes[] = {16, 64, 100, 243, 1000, 4096, 6000, 65536};
            for (int i = 0; i < iters; i++) {
                if (t == 0 && i == iters / 2) CHECK(asdFftDestroy(doomed) == ACL_SUCCESS);
                asdFftHandle h = nullptr;
                if (asdFftCreate(h) != ACL_SUCCESS || h == nullptr || h == shared) distinct_fail.fetch_add(1);
                const int64_t n = sizes[(t * 7 + i) % 8];
                const asdFftType ty = (i % 3 == 0) ? ASCEND_FFT_R2C : (i % 3 == 1) ? ASCEND_FFT_C2R : ASCEND_FFT_C2C;
                CHECK(asdFftMakePlan1D(h, n, ty, ASCEND_FFT_FORWARD, 1 + (i % 5)) == ACL_SUCCESS);
                size_t ws = 0;
                CHECK(asdFftGetWorkspaceSize(h, ws) == ACL_SUCCESS);
                asdFftType got;
                CHECK(asdFftGetType(h, got) == ACL_SUCCESS && got == ty);
                // a C2C exec on a real plan is refused before any launch
                CHECK(asdFftExecC2C(h, right, right) == ACL_ERROR_INVALID_PARAM);
                CHECK(asdFftSetStream(h, reinterpret_cast<void*>(0x20 + t)) == ACL_SUCCESS);
                CHECK(asdFftDestroy(h) == ACL_SUCCESS);
                CHECK(asdFftDestroy(h) == ACL_ERROR_INVALID_PARAM);  // already gone
                // the shared handle: concurrent queries and rejected execs
                asdFftType st;
                CHECK(asdFftGetType(shared, st) == ACL_SUCCESS && st == ASCEND_FFT_C2C);
                size_t sws = 7;
                CHECK(asdFftGetWorkspaceSize(shared, sws) == ACL_SUCCESS && sws == 0);
                CHECK(asdFftExecC2C(shared, wrong, right) == ACL_ERROR_INVALID_PARAM);
                CHECK(asdFftExecR2C(shared, wrong, right) == ACL_ERROR_INVALID_PARAM);
                CHECK(asdFftSetStream(shared, reinterpret_cast<void*>(0x40 + t)) == ACL_SUCCESS);
                // the handle another thread destroys half-way through
                asdFftType dty;
                const AspbStatus ds = asdFftGetType(doomed, dty);
                CHECK(ds == ACL_SUCCESS || ds == ACL_ERROR_INTERNAL_ERROR);  // as fft_api.cpp:1395-1397
                const AspbStatus de = asdFftExecC2C(doomed, wrong, right);
                CHECK(de == ACL_ERROR_INVALID_PARAM);
            }
        });
    }
    go.store(true);
    for (auto& th : pool) th.join();
    CHECK(distinct_fail.load() == 0);
    asdFftType after;
    CHECK(asdFftGetType(doomed, after) == ACL_ERROR_INTERNAL_ERROR);
    ASD_STATUS_CHECK(asdFftDestroy(shared));
    aclDestroyTensor(wrong);
    aclDestroyTensor(right);
}

}  // namespace

int main(int argc, char** argv)
{
    const int threads = argc > 1 ? std::atoi(argv[1]) : 8;
    const int iters = argc > 2 ? std::atoi(argv[2]) : 200;
    reference_call_shapes();
    concurrent_handles(threads, iters);
    const int f = g_failures.load();
    std::printf("contract_driver: %d failed checks (%d threads x %d iterations)\n", f, threads, iters);
    return f == 0 ? 0 : 1;
}
