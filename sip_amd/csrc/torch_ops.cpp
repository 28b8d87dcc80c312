// Torch operators of the FFT plugin (library `torch_sip`; reference sip_pta/csrc/fft/*.cpp) over the
// C ABI of libasdfft.so, for HIP tensors (dispatch key CUDA in ROCm builds of torch).
//
// Each operator checks its arguments as the reference operator does (TORCH_CHECK, so a bad call
// raises RuntimeError), takes the plan from a 10-entry LRU keyed by the plan parameters (reference
// sip_pta/csrc/FFTMixCache.cpp:18-88), takes the workspace from torch's caching allocator and
// launches on the current stream of the input's device.  The same functions are registered for CPU
// tensors only to run the argument checks; they then refuse: there is no host path.
#include <ATen/ATen.h>
#include <ATen/hip/impl/HIPGuardImplMasqueradingAsCUDA.h>
#include <c10/hip/HIPStream.h>
#include <torch/library.h>

#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "acl_meta.h"
#include "fft_api.h"

namespace {

using namespace AsdSip;

// ---- plan cache -------------------------------------------------------------------------
struct PlanKey {
    int type = 0;
    int direction = 0;
    std::vector<int64_t> dims;  // FFT sizes (1 to 3)
    int64_t batch = 0;
    int device = 0;
    bool operator==(const PlanKey& o) const
    {
        return type == o.type && direction == o.direction && dims == o.dims && batch == o.batch &&
               device == o.device;
    }
};

// One cached plan.  Binding the workspace and stream and launching happen under `mu`, so two
// threads that share a plan do not exchange workspaces; the handle is destroyed with the last
// reference (an evicted entry stays valid for a call that still holds it).
struct PlanEntry {
    asdFftHandle handle = nullptr;
    std::mutex mu;
    ~PlanEntry()
    {
        if (handle != nullptr) asdFftDestroy(handle);
    }
};

class PlanLru {
public:
    static constexpr int64_t kDefaultCapacity = 10;  // FFTMixCache::DEFAULT_CACHE_SIZE

    std::shared_ptr<PlanEntry> get(const PlanKey& key, AspbStatus* status)
    {
        std::lock_guard<std::mutex> lock(mu_);
        for (auto it = list_.begin(); it != list_.end(); ++it) {
            if (it->first == key) {
                list_.splice(list_.end(), list_, it);
                return list_.back().second;
            }
        }
        auto entry = std::make_shared<PlanEntry>();
        *status = make_plan(key, &entry->handle);
        if (*status != ACL_SUCCESS) return nullptr;  // a failed plan is not cached
        while ((int64_t)list_.size() >= capacity_ && !list_.empty()) list_.pop_front();
        if (capacity_ > 0) list_.emplace_back(key, entry);
        return entry;
    }

    void set_capacity(int64_t cap)
    {
        std::lock_guard<std::mutex> lock(mu_);
        capacity_ = cap;
        while ((int64_t)list_.size() > capacity_ && !list_.empty()) list_.pop_front();
    }

    int64_t size()
    {
        std::lock_guard<std::mutex> lock(mu_);
        return (int64_t)list_.size();
    }

    void clear()
    {
        std::lock_guard<std::mutex> lock(mu_);
        list_.clear();
    }

private:
    static AspbStatus make_plan(const PlanKey& k, asdFftHandle* out)
    {
        asdFftHandle h = nullptr;
        AspbStatus st = asdFftCreate(h);
        if (st != ACL_SUCCESS) return st;
        const auto type = static_cast<asdFftType>(k.type);
        const auto dir = static_cast<asdFftDirection>(k.direction);
        if (k.dims.size() == 1) {
            st = asdFftMakePlan1D(h, k.dims[0], type, dir, k.batch, ASCEND_FFT_HORIZONTAL);
        } else if (k.dims.size() == 2) {
            st = asdFftMakePlan2D(h, k.dims[0], k.dims[1], type, dir, (int32_t)k.batch);
        } else {
            st = asdFftMakePlan3D(h, k.dims[0], k.dims[1], k.dims[2], type, dir, (int32_t)k.batch);
        }
        if (st != ACL_SUCCESS) {
            asdFftDestroy(h);
            return st;
        }
        *out = h;
        return ACL_SUCCESS;
    }

    std::mutex mu_;
    int64_t capacity_ = kDefaultCapacity;
    std::list<std::pair<PlanKey, std::shared_ptr<PlanEntry>>> list_;
};

// Process lifetime on purpose: plans are not destroyed from a static destructor that may run
// after the HIP runtime has shut down.
PlanLru& plan_cache()
{
    static PlanLru* cache = new PlanLru();
    return *cache;
}

// ---- helpers ----------------------------------------------------------------------------
class AclView {
public:
    AclView(const at::Tensor& t, aclDataType dtype)
    {
        std::vector<int64_t> dims(t.sizes().begin(), t.sizes().end());
        std::vector<int64_t> strides(dims.size(), 1);
        for (int64_t i = (int64_t)dims.size() - 2; i >= 0; i--) strides[i] = strides[i + 1] * dims[i + 1];
        ptr_ = aclCreateTensor(dims.data(), dims.size(), dtype, strides.data(), 0, ACL_FORMAT_ND, dims.data(),
                               dims.size(), t.data_ptr());
        TORCH_CHECK(ptr_ != nullptr, "aclCreateTensor failed");
    }
    ~AclView() { aclDestroyTensor(ptr_); }
    AclView(const AclView&) = delete;
    AclView& operator=(const AclView&) = delete;
    aclTensor* get() const { return ptr_; }

private:
    aclTensor* ptr_ = nullptr;
};

void require_device(const at::Tensor& t, const char* op)
{
    TORCH_CHECK(t.is_cuda(), op, ": tensors must be on a HIP device (this library has no host path), got ",
                t.device());
}

// torch's default stream is the null stream, which asdFftSetStream refuses (as the reference
// does); hipStreamLegacy names that same legacy default stream.
hipStream_t current_stream(const at::Tensor& t)
{
    hipStream_t s = c10::hip::getCurrentHIPStream(t.device().index()).stream();
    return s != nullptr ? s : hipStreamLegacy;
}

// Cached plan for (type, direction, dims, batch) with its workspace and the current stream bound;
// `run` launches under the entry's lock.  EXEC_FFT_FUNC (pytorch_npu_helper.hpp:448-467).
template <typename Run>
void exec_fft(const char* op, asdFftType type, bool forward, const std::vector<int64_t>& dims, int64_t batch,
              const at::Tensor& like, Run run)
{
    PlanKey key;
    key.type = type;
    key.direction = forward ? ASCEND_FFT_FORWARD : ASCEND_FFT_INVERSE;
    key.dims = dims;
    key.batch = batch;
    key.device = like.device().index();
    AspbStatus st = ACL_SUCCESS;
    std::shared_ptr<PlanEntry> e = plan_cache().get(key, &st);
    TORCH_CHECK(e != nullptr, op, ": plan creation failed, status: ", st);
    std::lock_guard<std::mutex> lock(e->mu);
    size_t ws = 0;
    st = asdFftGetWorkspaceSize(e->handle, ws);
    TORCH_CHECK(st == ACL_SUCCESS, "asdFftGetWorkspaceSize failed, status: ", st);
    at::Tensor workspace;
    void* wp = nullptr;
    if (ws > 0) {
        workspace = at::empty({(int64_t)ws}, like.options().dtype(at::kByte));
        wp = workspace.data_ptr();
    }
    st = asdFftSetWorkspace(e->handle, wp);
    TORCH_CHECK(st == ACL_SUCCESS, "asdFftSetWorkspace failed, status: ", st);
    st = asdFftSetStream(e->handle, current_stream(like));
    TORCH_CHECK(st == ACL_SUCCESS, "asdFftSetStream failed, status: ", st);
    st = run(e->handle);
    TORCH_CHECK(st == ACL_SUCCESS, op, " failed, status: ", st);
}

std::vector<int64_t> fft_dims(at::IntArrayRef shape) { return std::vector<int64_t>(shape.begin() + 1, shape.end()); }

// ---- operators --------------------------------------------------------------------------
// asd_fft_c2c (sip_pta/csrc/fft/asd_fft_c2c.cpp:20-56)
at::Tensor asd_fft_c2c(const at::Tensor& in, bool is_forward)
{
    const char* op = "asd_fft_c2c";
    TORCH_CHECK(in.scalar_type() == at::kComplexFloat, op, ": input must be complex64, got ", in.scalar_type());
    at::Tensor input = in.contiguous();
    const int64_t rank = input.dim();
    TORCH_CHECK(rank >= 2 && rank <= 4, "Input tensor must have 2, 3 or 4 dimensions, but got ", rank);
    require_device(input, op);
    c10::hip::HIPGuardMasqueradingAsCUDA guard(input.device());
    at::Tensor output = at::empty_like(input);
    AclView vi(input, ACL_COMPLEX64), vo(output, ACL_COMPLEX64);
    exec_fft(op, ASCEND_FFT_C2C, is_forward, fft_dims(input.sizes()), input.size(0), input,
             [&](asdFftHandle h) { return asdFftExecC2C(h, vi.get(), vo.get()); });
    return output;
}

// asd_fft_r2c (sip_pta/csrc/fft/asd_fft_r2c.cpp:21-68): last axis N -> N/2+1 complex
at::Tensor asd_fft_r2c(const at::Tensor& in, bool is_forward)
{
    const char* op = "asd_fft_r2c";
    at::Tensor input = in.contiguous();
    TORCH_CHECK(input.scalar_type() == at::kFloat, "asd_fft_r2c: Input tensor must be Float32.");
    const int64_t rank = input.dim();
    TORCH_CHECK(rank >= 2 && rank <= 4, "asd_fft_r2c: Input tensor must have 2, 3 or 4 dimensions, but got ", rank);
    require_device(input, op);
    c10::hip::HIPGuardMasqueradingAsCUDA guard(input.device());
    std::vector<int64_t> out_shape(input.sizes().begin(), input.sizes().end());
    out_shape.back() = out_shape.back() / 2 + 1;
    at::Tensor output = at::empty(out_shape, input.options().dtype(at::kComplexFloat));
    AclView vi(input, ACL_FLOAT), vo(output, ACL_COMPLEX64);
    exec_fft(op, ASCEND_FFT_R2C, is_forward, fft_dims(input.sizes()), input.size(0), input,
             [&](asdFftHandle h) { return asdFftExecR2C(h, vi.get(), vo.get()); });
    return output;
}

// asd_fft_c2r (sip_pta/csrc/fft/asd_fft_c2r.cpp:22-70): last axis M -> 2(M-1) real
at::Tensor asd_fft_c2r(const at::Tensor& in, bool is_forward)
{
    const char* op = "asd_fft_c2r";
    at::Tensor input = in.contiguous();
    TORCH_CHECK(input.scalar_type() == at::kComplexFloat, "asd_fft_c2r: Input tensor must be ComplexFloat (COMPLEX64).");
    const int64_t rank = input.dim();
    TORCH_CHECK(rank >= 2 && rank <= 4, "asd_fft_c2r: Input tensor must have 2, 3 or 4 dimensions, but got ", rank);
    require_device(input, op);
    c10::hip::HIPGuardMasqueradingAsCUDA guard(input.device());
    std::vector<int64_t> out_shape(input.sizes().begin(), input.sizes().end());
    out_shape.back() = (out_shape.back() - 1) * 2;
    at::Tensor output = at::empty(out_shape, input.options().dtype(at::kFloat));
    AclView vi(input, ACL_COMPLEX64), vo(output, ACL_FLOAT);
    exec_fft(op, ASCEND_FFT_C2R, is_forward, fft_dims(output.sizes()), input.size(0), input,
             [&](asdFftHandle h) { return asdFftExecC2R(h, vi.get(), vo.get()); });
    return output;
}

// asd_fft_c2c_sep (sip_pta/csrc/fft/asd_fft_c2c_sep.cpp:22-65): (batch, N) or (batch, X, Y, Z) planes
std::tuple<at::Tensor, at::Tensor> asd_fft_c2c_sep(const at::Tensor& in_real, const at::Tensor& in_imag,
                                                   bool is_forward)
{
    const char* op = "asd_fft_c2c_sep";
    TORCH_CHECK(in_real.scalar_type() == at::kFloat && in_imag.scalar_type() == at::kFloat,
                "asd_fft_c2c_sep: input real/imag must be float32, got ", in_real.scalar_type(), "/",
                in_imag.scalar_type());
    at::Tensor re = in_real.contiguous();
    at::Tensor im = in_imag.contiguous();
    const int64_t rank = re.dim();
    TORCH_CHECK(rank == 2 || rank == 4, "Separated FFT support 2 or 4 dimensions (including Batch), but got ", rank);
    TORCH_CHECK(re.sizes() == im.sizes(), "Real and Imaginary tensors must have the same shape");
    require_device(re, op);
    require_device(im, op);
    c10::hip::HIPGuardMasqueradingAsCUDA guard(re.device());
    at::Tensor out_re = at::empty_like(re);
    at::Tensor out_im = at::empty_like(im);
    AclView a(re, ACL_FLOAT), b(im, ACL_FLOAT), c(out_re, ACL_FLOAT), d(out_im, ACL_FLOAT);
    exec_fft(op, ASCEND_FFT_C2C_SEP, is_forward, fft_dims(re.sizes()), re.size(0), re,
             [&](asdFftHandle h) { return asdFftExecC2CSeparated(h, a.get(), b.get(), c.get(), d.get()); });
    return {out_re, out_im};
}

// asd_fft_shift / asd_ifft_shift (sip_pta/csrc/fft/asd_fft_shift.cpp:29-67, asd_ifft_shift.cpp:29-70)
at::Tensor shift_op(const char* op, const at::Tensor& in, int64_t row, int64_t col, int64_t batch, int64_t dist,
                    bool inverse, double scale)
{
    TORCH_CHECK(in.scalar_type() == at::kComplexFloat, op, ": input must be complex64, got ", in.scalar_type());
    TORCH_CHECK(row > 0 && col > 0 && batch > 0, op, ": row/col/batch must be positive, got row=", row,
                " col=", col, " batch=", batch);
    const int64_t dist_val = (dist > 0) ? dist : row * col;
    TORCH_CHECK(dist_val >= row * col, op, ": dist (", dist_val, ") must be >= row * col (", row * col, ")");
    const int64_t required = (batch - 1) * dist_val + row * col;
    TORCH_CHECK(in.numel() >= required, op, ": input needs at least ", required, " elements, got ", in.numel());
    if (in.dim() == 3) {
        TORCH_CHECK(batch == in.size(0) && row == in.size(1) && col == in.size(2), op, ": 3D input shape [",
                    in.size(0), ", ", in.size(1), ", ", in.size(2), "] must match explicit batch=", batch,
                    " row=", row, " col=", col);
    }
    at::Tensor input = in.contiguous().reshape({-1});
    require_device(input, op);
    c10::hip::HIPGuardMasqueradingAsCUDA guard(input.device());
    at::Tensor output = at::empty({batch, row, col}, input.options());
    AclView vi(input, ACL_COMPLEX64), vo(output, ACL_COMPLEX64);
    void* stream = current_stream(input);
    const AspbStatus st = inverse ? asdIfftShift(row, col, batch, dist_val, scale, vi.get(), vo.get(), stream, nullptr)
                                  : asdFftShift(row, col, batch, dist_val, vi.get(), vo.get(), stream, nullptr);
    TORCH_CHECK(st == ACL_SUCCESS, inverse ? "asdIfftShift" : "asdFftShift", " failed, status: ", st);
    return output;
}

at::Tensor asd_fft_shift(const at::Tensor& in, int64_t row, int64_t col, int64_t batch, int64_t dist)
{
    return shift_op("asd_fft_shift", in, row, col, batch, dist, false, 1.0);
}

at::Tensor asd_fft_shift_auto(const at::Tensor& in)
{
    TORCH_CHECK(in.dim() == 3, "asd_fft_shift: auto mode expects a 3D [batch, row, col] tensor, got ", in.dim(),
                " dimensions");
    return asd_fft_shift(in, in.size(1), in.size(2), in.size(0), 0);
}

at::Tensor asd_ifft_shift(const at::Tensor& in, int64_t row, int64_t col, int64_t batch, int64_t dist, double scale)
{
    return shift_op("asd_ifft_shift", in, row, col, batch, dist, true, scale);
}

at::Tensor asd_ifft_shift_auto(const at::Tensor& in)
{
    TORCH_CHECK(in.dim() == 3, "asd_ifft_shift: auto mode expects a 3D [batch, row, col] tensor, got ", in.dim(),
                " dimensions");
    return asd_ifft_shift(in, in.size(1), in.size(2), in.size(0), 0, 1.0);
}

// plan cache control (FFTMixCache::setCapacity / getSize / clear; no operator in the reference)
void asd_fft_plan_cache_set_capacity(int64_t capacity) { plan_cache().set_capacity(capacity); }
int64_t asd_fft_plan_cache_size() { return plan_cache().size(); }
void asd_fft_plan_cache_clear() { plan_cache().clear(); }

}  // namespace

TORCH_LIBRARY(torch_sip, m)
{
    m.def("asd_fft_c2c(Tensor input, bool is_forward) -> Tensor");
    m.def("asd_fft_c2c_sep(Tensor in_real, Tensor in_imag, bool is_forward) -> (Tensor, Tensor)");
    m.def("asd_fft_r2c(Tensor input, bool is_forward) -> Tensor");
    m.def("asd_fft_c2r(Tensor input, bool is_forward) -> Tensor");
    m.def("asd_fft_shift(Tensor input, int row, int col, int batch, int dist=0) -> Tensor");
    m.def("asd_fft_shift_auto(Tensor input) -> Tensor");
    m.def("asd_ifft_shift(Tensor input, int row, int col, int batch, int dist=0, float scale=1.0) -> Tensor");
    m.def("asd_ifft_shift_auto(Tensor input) -> Tensor");
    m.def("asd_fft_plan_cache_set_capacity(int capacity) -> ()", &asd_fft_plan_cache_set_capacity);
    m.def("asd_fft_plan_cache_size() -> int", &asd_fft_plan_cache_size);
    m.def("asd_fft_plan_cache_clear() -> ()", &asd_fft_plan_cache_clear);
}

#define ASDFFT_TORCH_IMPLS(m)                               \
    m.impl("asd_fft_c2c", &asd_fft_c2c);                    \
    m.impl("asd_fft_c2c_sep", &asd_fft_c2c_sep);            \
    m.impl("asd_fft_r2c", &asd_fft_r2c);                    \
    m.impl("asd_fft_c2r", &asd_fft_c2r);                    \
    m.impl("asd_fft_shift", &asd_fft_shift);                \
    m.impl("asd_fft_shift_auto", &asd_fft_shift_auto);      \
    m.impl("asd_ifft_shift", &asd_ifft_shift);              \
    m.impl("asd_ifft_shift_auto", &asd_ifft_shift_auto);

TORCH_LIBRARY_IMPL(torch_sip, CUDA, m) { ASDFFT_TORCH_IMPLS(m) }
TORCH_LIBRARY_IMPL(torch_sip, CPU, m) { ASDFFT_TORCH_IMPLS(m) }
