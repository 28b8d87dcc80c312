// Implementation of include/fft_api.h on HIP.
//
// Validation order, status codes and plan-state rules follow the reference
// implementation core/fft/fft_api.cpp (cited per function); the plans
// themselves come from the gfx950 planner (planner.cpp) and run as HIP kernels
// on the bound stream.  There is no host fallback: without a HIP device, plans
// can be built and validated but every exec returns ACL_ERROR_INTERNAL_ERROR.
#include "fft_api.h"

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "acl_meta.h"
#include "asdfft_ext.h"
#include "planner.h"
#include "ragged.h"
#include "shift.h"
#include "static_scratch.h"
#include "tensor_desc.h"

namespace {

using namespace AsdSip;

constexpr int64_t kMaxFftSize = int64_t(1) << 27;  // fft_api.cpp:82

enum class PlanState { kCreated, kInitialized, kFailed };

// ISTFT plan parameters (IstftDesc, core/fft/istft.cpp): the spectrum shape recorded at plan
// time and the overlap-add geometry.
struct IstftDesc {
    int64_t channel = 0, nfft = 0, nframes = 0, hop = 0, win_length = 0, out_len = 0;
};

struct FftPlan {
    std::mutex mu;
    PlanState state = PlanState::kCreated;
    asdFftType type = ASCEND_FFT_C2C;
    asdFftDirection direction = ASCEND_FFT_FORWARD;
    std::vector<int64_t> sizes;
    int64_t batch = 1;          // transforms per exec (horizontal / multi-dim)
    int64_t vertical_batch = 1; // VERTICAL 1D: tensor (N, vertical_batch)
    bool vertical = false;
    void* stream = nullptr;
    void* workspace = nullptr;
    size_t workspace_recorded = 0;  // set by GetWorkspaceSize (fft_api.cpp:1339-1360)
    bool workspace_queried = false;
    asdfft::BuiltPlan built;
    std::unique_ptr<asdfft::RaggedPlan> ragged;  // set by asdFftExtMakePlanRagged
    std::unique_ptr<IstftDesc> istft;            // set by asdFftIstftMakePlan
    asdfft::Precision precision = asdfft::Precision::kFp32;  // asdFftExtSetPrecision
    int in_shift = asdfft::kShiftNone;                       // asdFftExtSetShift
    int out_shift = asdfft::kShiftNone;
    float out_scale = 1.0f;
    bool shifted() const { return in_shift != asdfft::kShiftNone || out_shift != asdfft::kShiftNone || out_scale != 1.0f; }
};

// Handles are keys of this table, drawn from a counter that never repeats a value, not plan
// addresses: a stale handle (used after asdFftDestroy) then never names a plan created later,
// so it is rejected instead of acting on another caller's plan.
std::mutex g_table_mu;
uintptr_t g_next_key = 0x1000;  // guarded by g_table_mu
std::unordered_map<void*, std::shared_ptr<FftPlan>>& table()
{
    static auto* t = new std::unordered_map<void*, std::shared_ptr<FftPlan>>();
    return *t;
}

std::shared_ptr<FftPlan> lookup(asdFftHandle h)
{
    std::lock_guard<std::mutex> lock(g_table_mu);
    auto it = table().find(h);
    if (it == table().end()) return nullptr;
    return it->second;
}

void log_error(const char* fn, const std::string& msg)
{
    static const bool quiet = [] {
        const char* e = getenv("ASDFFT_QUIET");
        return e != nullptr && e[0] == '1';
    }();
    if (!quiet) fprintf(stderr, "[asdfft] %s: %s\n", fn, msg.c_str());
}

bool trace_enabled()
{
    const char* e = getenv("ASDFFT_PLAN_TRACE");
    return e != nullptr && strcmp(e, "1") == 0;
}

bool valid_direction(asdFftDirection d)  // CheckFftDirection, fft_api.cpp:1088-1096
{
    return d == ASCEND_FFT_FORWARD || d == ASCEND_FFT_INVERSE;
}

asdfft::Kind kind_of(asdFftType t)
{
    if (t == ASCEND_FFT_R2C) return asdfft::Kind::R2C;
    if (t == ASCEND_FFT_C2R) return asdfft::Kind::C2R;
    return asdfft::Kind::C2C;
}

AspbStatus finish_plan(FftPlan& p, const char* fn)
{
    asdfft::PlanSpec spec;
    spec.kind = kind_of(p.type);
    spec.inverse = (p.direction == ASCEND_FFT_INVERSE);
    spec.dims = p.sizes;
    spec.batch = p.vertical ? p.vertical_batch : p.batch;
    spec.vertical = p.vertical;
    spec.planar = (p.type == ASCEND_FFT_C2C_SEP);
    spec.precision = p.precision;
    spec.in_shift = p.in_shift;
    spec.out_shift = p.out_shift;
    spec.out_scale = p.out_scale;
    asdfft::BuiltPlan built;
    std::string why;
    try {
        why = asdfft::build_plan(spec, &built);
    } catch (const std::exception& e) {
        why = e.what();
    }
    if (!why.empty()) {
        log_error(fn, why);
        p.built = asdfft::BuiltPlan{};
        return p.shifted() ? ACL_ERROR_API_NOT_SUPPORT : ACL_ERROR_INVALID_PARAM;
    }
    p.built = std::move(built);
    p.state = PlanState::kInitialized;
    if (trace_enabled()) fprintf(stderr, "[asdfft] %s", asdfft::describe_plan(p.built).c_str());
    return ACL_SUCCESS;
}

// ---- tensor / plan matching (fft_api.cpp:1529-1722) -------------------------------------
bool leading_extent_matches(const std::vector<int64_t>& dims, int64_t lead_count, int64_t batch)
{
    if (batch <= 0) return false;
    int64_t ext = 1;
    for (int64_t j = 0; j < lead_count; j++) {
        if (dims[j] <= 0 || ext > batch / dims[j]) return false;
        ext *= dims[j];
    }
    return ext == batch;
}

int64_t plan_batch_extent(const FftPlan& p) { return p.vertical ? p.vertical_batch : p.batch; }

// matchC2C1D_ (fft_api.cpp:1549-1566)
bool match_c2c_1d(const FftPlan& p, const std::vector<int64_t>& d)
{
    if (d.size() != 2) return false;
    const int64_t n = p.sizes[0];
    if (p.vertical && p.vertical_batch > 1) return d[0] == n && d[1] == p.vertical_batch;
    if (plan_batch_extent(p) == 1) return (d[0] == 1 && d[1] == n) || (d[0] == n && d[1] == 1);
    return d[0] == p.batch && d[1] == n;
}

// matchC2C_ (fft_api.cpp:1568-1618); multi-dimensional plans additionally require the
// leading dimensions to cover exactly the plan's batch, as the real variants do.
bool match_c2c(const FftPlan& p, const aclTensor* t)
{
    if (t == nullptr || t->dtype != ACL_COMPLEX64) return false;
    const auto& d = t->view;
    if (p.sizes.size() == 1) return match_c2c_1d(p, d);
    if (d.size() < p.sizes.size()) return false;
    size_t i = d.size() - 1;
    for (size_t k = p.sizes.size(); k-- > 0; i--) {
        if (d[i] != p.sizes[k]) return false;
    }
    return leading_extent_matches(d, (int64_t)(d.size() - p.sizes.size()), p.batch);
}

// matchC2R_ (fft_api.cpp:1620-1674): complex half spectrum
bool match_half_complex(const FftPlan& p, const aclTensor* t)
{
    if (t == nullptr || t->dtype != ACL_COMPLEX64) return false;
    const auto& d = t->view;
    if (d.size() < p.sizes.size()) return false;
    int64_t i = (int64_t)d.size() - 1;
    for (size_t k = p.sizes.size(); k-- > 0; i--) {
        const int64_t want = (i == (int64_t)d.size() - 1) ? p.sizes[k] / 2 + 1 : p.sizes[k];
        if (d[i] != want) return false;
    }
    return leading_extent_matches(d, i + 1, plan_batch_extent(p));
}

// matchR2C_ (fft_api.cpp:1676-1722): real signal
bool match_real(const FftPlan& p, const aclTensor* t)
{
    if (t == nullptr || t->dtype != ACL_FLOAT) return false;
    const auto& d = t->view;
    if (d.size() < p.sizes.size()) return false;
    int64_t i = (int64_t)d.size() - 1;
    for (size_t k = p.sizes.size(); k-- > 0; i--) {
        if (d[i] != p.sizes[k]) return false;
    }
    return leading_extent_matches(d, i + 1, plan_batch_extent(p));
}

// matchSeparatedPlan_ (fft_api.cpp:1728-1750): trailing extents are the plan's sizes and
// the leading ones cover exactly its batch
bool match_separated(const FftPlan& p, const std::vector<int64_t>& d)
{
    if (p.sizes.empty() || d.size() < p.sizes.size()) return false;
    int64_t i = (int64_t)d.size() - 1;
    for (size_t k = p.sizes.size(); k-- > 0; i--) {
        if (d[i] != p.sizes[k]) return false;
    }
    return leading_extent_matches(d, i + 1, plan_batch_extent(p));
}

bool ranges_overlap(const void* a, size_t na, const void* b, size_t nb)
{
    const char* x = static_cast<const char*>(a);
    const char* y = static_cast<const char*>(b);
    return x < y + nb && y < x + na;
}

// The launch list of a built plan.  bufs = {input, output, workspace}; a step whose
// source (destination) is a pair of float planes takes the real plane from bufs and the
// imaginary plane from in_im (out_im).  only_step >= 0 launches that one step.
AspbStatus launch_steps(const asdfft::BuiltPlan& bp, char* const bufs[3], const float* in_im, float* out_im,
                        hipStream_t stream, const char* fn, int only_step)
{
    for (size_t si = 0; si < bp.steps.size(); si++) {
        if (only_step >= 0 && (size_t)only_step != si) continue;
        const asdfft::Step& s = bp.steps[si];
        if ((s.planar_in && in_im == nullptr) || (s.planar_out && out_im == nullptr)) {
            log_error(fn, "separated-plane step without its imaginary plane");
            return ACL_ERROR_INTERNAL_ERROR;
        }
        const hipError_t e = asdfft::launch_step(s, bufs, in_im, out_im, stream);
        if (e != hipSuccess) {
            log_error(fn, std::string("kernel launch failed: ") + hipGetErrorString(e));
            return ACL_ERROR_INTERNAL_ERROR;
        }
    }
    return ACL_SUCCESS;
}

// CheckWorkspaceBinding (fft_api.cpp:1015-1032): a plan that needs workspace must have
// had its size queried and a large enough buffer bound.
AspbStatus check_workspace(const FftPlan& p, const char* fn)
{
    const asdfft::BuiltPlan& bp = p.built;
    if (bp.workspace > 0 && p.workspace == nullptr) {
        log_error(fn, "workspace has not been set");
        return ACL_ERROR_INTERNAL_ERROR;
    }
    if (bp.workspace > 0 && !p.workspace_queried) {
        log_error(fn, "workspace size has not been queried, call asdFftGetWorkspaceSize before exec");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (bp.workspace > p.workspace_recorded) {
        log_error(fn, "workspace smaller than the plan needs");
        return ACL_ERROR_INVALID_PARAM;
    }
    return ACL_SUCCESS;
}

// Every step's device tables exist (a plan made without a device has none).
AspbStatus check_device(const asdfft::BuiltPlan& bp, const char* fn)
{
    if (!asdfft::device_available()) {
        log_error(fn, "no HIP device: this library has no host execution path");
        return ACL_ERROR_INTERNAL_ERROR;
    }
    for (const asdfft::Step& s : bp.steps) {
        if (!s.tables_ready) {
            log_error(fn, "plan was built without a HIP device; create it again on a device");
            return ACL_ERROR_INTERNAL_ERROR;
        }
    }
    return ACL_SUCCESS;
}

// asdFftExecV2 (fft_api.cpp:1446-1511): workspace guard, then the launch list
// (only_step >= 0 launches that one step, for per-kernel timing).  no_alloc: the plan's
// workspace is the library's static device scratch (asdFftExtExecNoAlloc) instead of the
// caller's bound buffer; plans needing more than its capacity are refused.
AspbStatus run_plan(FftPlan& p, const aclTensor* input, const aclTensor* output, const char* fn, int only_step = -1,
                    bool no_alloc = false)
{
    const asdfft::BuiltPlan& bp = p.built;
    void* workspace = p.workspace;
    if (no_alloc) {
        if (bp.workspace > asdfft::kStaticScratchBytes) {
            log_error(fn, "the plan needs " + std::to_string(bp.workspace) + " bytes of workspace, more than the " +
                              std::to_string(asdfft::kStaticScratchBytes) +
                              "-byte static scratch; use asdFftExec* with a bound workspace");
            return ACL_ERROR_INVALID_PARAM;
        }
    } else {
        const AspbStatus ws = check_workspace(p, fn);
        if (ws != ACL_SUCCESS) return ws;
    }
    void* in = asdfft::data_address(input);
    void* out = asdfft::data_address(output);
    if (in == nullptr || out == nullptr) {
        log_error(fn, "tensor data is nullptr");
        return ACL_ERROR_INTERNAL_ERROR;
    }
    if (!asdfft::is_dense(input) || !asdfft::is_dense(output)) {
        log_error(fn, "only dense row-major tensors are supported");
        return ACL_ERROR_INVALID_PARAM;
    }
    const size_t in_bytes = (size_t)asdfft::numel(input) * (size_t)asdfft::element_bytes(input->dtype);
    const size_t out_bytes = (size_t)asdfft::numel(output) * (size_t)asdfft::element_bytes(output->dtype);
    // exact in-place C2C is allowed, except with folded shifts (they move data between workgroups)
    const bool same = (in == out && in_bytes == out_bytes && p.type == ASCEND_FFT_C2C && !p.shifted());
    if (!same && ranges_overlap(in, in_bytes, out, out_bytes)) {
        log_error(fn, "input and output overlap (only exact in-place C2C is supported)");
        return ACL_ERROR_INVALID_PARAM;
    }
    const AspbStatus dev = check_device(bp, fn);
    if (dev != ACL_SUCCESS) return dev;
    if (no_alloc && bp.workspace > 0) {
        const hipError_t e = asdfft::static_scratch(&workspace);
        if (e != hipSuccess) {
            log_error(fn, std::string("static scratch lookup failed: ") + hipGetErrorString(e));
            return ACL_ERROR_INTERNAL_ERROR;
        }
    }
    char* bufs[3] = {static_cast<char*>(in), static_cast<char*>(out), static_cast<char*>(workspace)};
    return launch_steps(bp, bufs, nullptr, nullptr, static_cast<hipStream_t>(p.stream), fn, only_step);
}

template <typename MatchIn, typename MatchOut>
AspbStatus exec_common(asdFftHandle handle, const aclTensor* input, const aclTensor* output, asdFftType want,
                       MatchIn match_in, MatchOut match_out, const char* fn, int only_step = -1,
                       bool no_alloc = false)
{
    auto sp = lookup(handle);
    if (!sp) {
        log_error(fn, "invalid handle");
        return ACL_ERROR_INVALID_PARAM;
    }
    FftPlan& p = *sp;
    std::lock_guard<std::mutex> lock(p.mu);
    if (p.state != PlanState::kInitialized) {
        log_error(fn, "plan is not initialized");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (p.ragged) {
        log_error(fn, "ragged-batch plans run through asdFftExtExecRagged");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (p.istft) {
        log_error(fn, "ISTFT plans run through asdFftExecIstft");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (p.type != want) {  // fft_api.cpp:1764-1767
        log_error(fn, "plan fftType does not match this exec entry point");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (!match_in(p, input)) {
        log_error(fn, "input does not match plan");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (!match_out(p, output)) {
        log_error(fn, "output does not match plan");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (only_step >= (int)p.built.steps.size()) {
        log_error(fn, "step index out of range");
        return ACL_ERROR_INVALID_PARAM;
    }
    return run_plan(p, input, output, fn, only_step, no_alloc);
}

// asdFftShift / asdIfftShift (core/fft/fftshift.cpp:101-138 and ifftshift.cpp): the
// reference's parameter checks in its order, then one HBM-bound gather kernel.
//   asdFftShift : out[s][(r + ceil(row/2)) % row][(c + ceil(col/2)) % col] = in[s*dist + r*col + c]
//   asdIfftShift: out[s][(r + row/2) % row][(c + col/2) % col] = float(scale) * in[...]
bool storage_numel(const aclTensor* t, int64_t* n)
{
    int64_t v = 1;
    for (int64_t d : t->storage) {
        if (d < 0 || (d > 0 && v > std::numeric_limits<int64_t>::max() / d)) return false;
        v *= d;
    }
    *n = v;
    return true;
}

AspbStatus run_shift(const char* fn, int64_t row, int64_t col, int64_t batch, int64_t dist, double scale,
                     bool inverse, const aclTensor* input, aclTensor* output, void* stream)
{
    constexpr int64_t kMax = std::numeric_limits<int64_t>::max();
    if (row <= 0 || col <= 0 || batch <= 0) {
        log_error(fn, "row, col and batch must be positive");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (row > kMax / col) {
        log_error(fn, "row * col overflows int64");
        return ACL_ERROR_INVALID_PARAM;
    }
    const int64_t slice = row * col;
    if (dist <= 0) {
        dist = slice;
    } else if (dist < slice) {
        log_error(fn, "dist must be at least row * col when positive");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (input == nullptr || output == nullptr) {
        log_error(fn, "tensor is nullptr");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (input->dtype != ACL_COMPLEX64 || output->dtype != ACL_COMPLEX64) {
        log_error(fn, "input and output must be complex64");
        return ACL_ERROR_UNSUPPORTED_DATA_TYPE;
    }
    if (batch - 1 > (kMax - slice) / dist) {
        log_error(fn, "(batch - 1) * dist + row * col overflows int64");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (batch > kMax / slice) {
        log_error(fn, "batch * row * col overflows int64");
        return ACL_ERROR_INVALID_PARAM;
    }
    const int64_t need_in = (batch - 1) * dist + slice;
    const int64_t need_out = batch * slice;
    int64_t have_in = 0, have_out = 0;
    if (!storage_numel(input, &have_in) || have_in < need_in) {
        log_error(fn, "input holds fewer elements than the slices need");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (!storage_numel(output, &have_out) || have_out < need_out) {
        log_error(fn, "output holds fewer elements than the slices need");
        return ACL_ERROR_INVALID_PARAM;
    }
    void* in = asdfft::data_address(input);
    void* out = asdfft::data_address(output);
    if (in == nullptr || out == nullptr) {
        log_error(fn, "tensor data is nullptr");
        return ACL_ERROR_INTERNAL_ERROR;
    }
    if (ranges_overlap(in, (size_t)need_in * 8, out, (size_t)need_out * 8)) {
        log_error(fn, "input and output overlap (the shift is a gather into a separate buffer)");
        return ACL_ERROR_INVALID_PARAM;
    }
    // the kernel indexes rows and columns in 32-bit arithmetic (row + rows per workgroup < 2^31)
    if (row >= ((int64_t)1 << 30) || col >= ((int64_t)1 << 30) || batch * row >= ((int64_t)1 << 40)) {
        log_error(fn, "slice extents of 2^30 rows or columns and more are not supported");
        return ACL_ERROR_API_NOT_SUPPORT;
    }
    if (!asdfft::device_available()) {
        log_error(fn, "no HIP device: this library has no host execution path");
        return ACL_ERROR_INTERNAL_ERROR;
    }
    asdfft::ShiftArgs a{};
    a.in = static_cast<const float2*>(in);
    a.out = static_cast<float2*>(out);
    a.row = row;
    a.col = col;
    a.batch = batch;
    a.dist = dist;
    a.shift_r = inverse ? row / 2 : (row + 1) / 2;
    a.shift_c = inverse ? col / 2 : (col + 1) / 2;
    a.scale = (float)scale;
    a.apply_scale = inverse && scale != 1.0;
    a.rows_per_block = (int)std::max<int64_t>(1, std::min<int64_t>(4096 / col, (int64_t)1 << 20));
    const hipError_t e = asdfft::launch_shift2d(a, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        log_error(fn, std::string("kernel launch failed: ") + hipGetErrorString(e));
        return ACL_ERROR_INTERNAL_ERROR;
    }
    return ACL_SUCCESS;
}

}  // namespace

namespace AsdSip {

extern "C" {

AspbStatus asdFftCreate(asdFftHandle& handle)  // fft_api.cpp:1047-1065
{
    try {
        auto p = std::make_shared<FftPlan>();
        std::lock_guard<std::mutex> lock(g_table_mu);
        g_next_key += 0x10;
        void* key = reinterpret_cast<void*>(g_next_key);
        table()[key] = std::move(p);
        handle = key;
    } catch (const std::exception& e) {
        handle = nullptr;
        log_error("asdFftCreate", e.what());
        return ACL_ERROR_INTERNAL_ERROR;
    }
    return ACL_SUCCESS;
}

AspbStatus asdFftSetStream(asdFftHandle handle, void* stream)  // fft_api.cpp:1067-1082
{
    if (stream == nullptr) {
        log_error("asdFftSetStream", "stream is nullptr");
        return ACL_ERROR_INVALID_PARAM;
    }
    auto sp = lookup(handle);
    if (!sp) {
        log_error("asdFftSetStream", "invalid handle");
        return ACL_ERROR_INVALID_PARAM;
    }
    std::lock_guard<std::mutex> lock(sp->mu);
    sp->stream = stream;
    return ACL_SUCCESS;
}

AspbStatus asdFftMakePlan1D(asdFftHandle handle, int64_t fftSize, asdFftType fftType, asdFftDirection direction,
                            int64_t batchSize, asdFft1dDimType dimType)  // fft_api.cpp:1124-1190
{
    const char* fn = "asdFftMakePlan1D";
    auto sp = lookup(handle);
    if (!sp) {
        log_error(fn, "invalid handle");
        return ACL_ERROR_INVALID_PARAM;
    }
    FftPlan& p = *sp;
    std::lock_guard<std::mutex> lock(p.mu);
    if (fftSize <= 0 || fftSize > kMaxFftSize) {
        log_error(fn, "invalid fft_size");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (batchSize <= 0 || batchSize > std::numeric_limits<int32_t>::max()) {
        log_error(fn, "invalid batch_size");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (fftType != ASCEND_FFT_C2C && fftType != ASCEND_FFT_C2R && fftType != ASCEND_FFT_R2C &&
        fftType != ASCEND_FFT_C2C_SEP) {
        log_error(fn, "unsupported fftType for a 1D plan");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (!valid_direction(direction)) {
        log_error(fn, "invalid direction");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (p.state == PlanState::kInitialized) {
        log_error(fn, "plan is already initialized");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (dimType != ASCEND_FFT_HORIZONTAL && dimType != ASCEND_FFT_VERTICAL) {  // Setup1dBatchLayout
        log_error(fn, "invalid dimType");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (dimType == ASCEND_FFT_VERTICAL && batchSize > 1 && fftType != ASCEND_FFT_C2C) {
        log_error(fn, "VERTICAL layout with batch > 1 is supported for C2C only");
        return ACL_ERROR_API_NOT_SUPPORT;
    }
    p.type = fftType;
    p.direction = direction;
    p.sizes = {fftSize};
    p.vertical = (dimType == ASCEND_FFT_VERTICAL);
    p.batch = p.vertical ? 1 : batchSize;
    p.vertical_batch = p.vertical ? batchSize : 1;
    if (p.vertical && batchSize == 1) p.vertical = false;  // (1, N) and (N, 1) both accepted
    return finish_plan(p, fn);
}

AspbStatus asdFftMakePlan2D(asdFftHandle handle, int64_t fftSizeX, int64_t fftSizeY, asdFftType fftType,
                            asdFftDirection direction, int32_t batchSize)  // fft_api.cpp:1245-1289
{
    const char* fn = "asdFftMakePlan2D";
    auto sp = lookup(handle);
    if (!sp) {
        log_error(fn, "invalid handle");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (fftSizeX <= 0 || fftSizeX > kMaxFftSize || fftSizeY <= 0 || fftSizeY > kMaxFftSize || batchSize <= 0) {
        log_error(fn, "invalid size or batch");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (fftType != ASCEND_FFT_C2C && fftType != ASCEND_FFT_C2R && fftType != ASCEND_FFT_R2C) {
        log_error(fn, "unsupported fftType for a 2D plan");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (!valid_direction(direction)) {
        log_error(fn, "invalid direction");
        return ACL_ERROR_INVALID_PARAM;
    }
    FftPlan& p = *sp;
    std::lock_guard<std::mutex> lock(p.mu);
    if (p.state == PlanState::kInitialized) {
        log_error(fn, "plan is already initialized");
        return ACL_ERROR_INVALID_PARAM;
    }
    const int64_t imax = std::numeric_limits<int32_t>::max();  // derived batch guard, fft_api.cpp:666-673
    if ((int64_t)batchSize * fftSizeX > imax || (int64_t)batchSize * fftSizeY > imax) {
        log_error(fn, "derived batch size exceeds INT32_MAX");
        return ACL_ERROR_INVALID_PARAM;
    }
    p.type = fftType;
    p.direction = direction;
    p.sizes = {fftSizeX, fftSizeY};
    p.batch = batchSize;
    p.vertical = false;
    p.vertical_batch = 1;
    return finish_plan(p, fn);
}

AspbStatus asdFftMakePlan3D(asdFftHandle handle, int64_t fftSizeX, int64_t fftSizeY, int64_t fftSizeZ,
                            asdFftType fftType, asdFftDirection direction, int32_t batchSize)  // fft_api.cpp:1291-1337
{
    const char* fn = "asdFftMakePlan3D";
    auto sp = lookup(handle);
    if (!sp) {
        log_error(fn, "invalid handle");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (fftSizeX <= 0 || fftSizeX > kMaxFftSize || fftSizeY <= 0 || fftSizeY > kMaxFftSize || batchSize <= 0 ||
        fftSizeZ <= 0 || fftSizeZ > kMaxFftSize) {
        log_error(fn, "invalid size or batch");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (fftType != ASCEND_FFT_C2C && fftType != ASCEND_FFT_C2R && fftType != ASCEND_FFT_R2C &&
        fftType != ASCEND_FFT_C2C_SEP) {
        log_error(fn, "unsupported fftType for a 3D plan");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (!valid_direction(direction)) {
        log_error(fn, "invalid direction");
        return ACL_ERROR_INVALID_PARAM;
    }
    FftPlan& p = *sp;
    std::lock_guard<std::mutex> lock(p.mu);
    if (p.state == PlanState::kInitialized) {
        log_error(fn, "plan is already initialized");
        return ACL_ERROR_INVALID_PARAM;
    }
    const int64_t imax = std::numeric_limits<int32_t>::max();
    const int64_t b = batchSize;
    if (fftSizeX > imax / b / fftSizeY || fftSizeX > imax / b / fftSizeZ || fftSizeY > imax / b / fftSizeZ) {
        log_error(fn, "derived batch size exceeds INT32_MAX");
        return ACL_ERROR_INVALID_PARAM;
    }
    p.type = fftType;
    p.direction = direction;
    p.sizes = {fftSizeX, fftSizeY, fftSizeZ};
    p.batch = batchSize;
    p.vertical = false;
    p.vertical_batch = 1;
    return finish_plan(p, fn);
}

AspbStatus asdFftGetWorkspaceSize(asdFftHandle handle, size_t& workspaceSize)  // fft_api.cpp:1339-1360
{
    auto sp = lookup(handle);
    if (!sp) {
        log_error("asdFftGetWorkspaceSize", "invalid handle");
        return ACL_ERROR_INVALID_PARAM;
    }
    std::lock_guard<std::mutex> lock(sp->mu);
    workspaceSize = sp->built.workspace;
    sp->workspace_recorded = workspaceSize;
    sp->workspace_queried = true;
    return ACL_SUCCESS;
}

AspbStatus asdFftSetWorkspace(asdFftHandle handle, void* workspace)  // fft_api.cpp:1362-1374
{
    auto sp = lookup(handle);
    if (!sp) {
        log_error("asdFftSetWorkspace", "invalid handle");
        return ACL_ERROR_INVALID_PARAM;
    }
    std::lock_guard<std::mutex> lock(sp->mu);
    sp->workspace = workspace;
    return ACL_SUCCESS;
}

AspbStatus asdFftSynchronize(asdFftHandle handle)  // fft_api.cpp:1376-1390
{
    auto sp = lookup(handle);
    if (!sp) {
        log_error("asdFftSynchronize", "invalid handle");
        return ACL_ERROR_INTERNAL_ERROR;
    }
    void* s;
    {
        std::lock_guard<std::mutex> lock(sp->mu);
        s = sp->stream;
    }
    if (!asdfft::device_available()) {
        log_error("asdFftSynchronize", "no HIP device");
        return ACL_ERROR_INTERNAL_ERROR;
    }
    hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(s));
    if (e != hipSuccess) {
        log_error("asdFftSynchronize", hipGetErrorString(e));
        return ACL_ERROR_INTERNAL_ERROR;
    }
    return ACL_SUCCESS;
}

AspbStatus asdFftGetType(asdFftHandle handle, asdFftType& fftType)  // fft_api.cpp:1392-1404
{
    auto sp = lookup(handle);
    if (!sp) {
        log_error("asdFftGetType", "invalid handle");
        return ACL_ERROR_INTERNAL_ERROR;
    }
    std::lock_guard<std::mutex> lock(sp->mu);
    fftType = sp->type;
    return ACL_SUCCESS;
}

AspbStatus asdFftDestroy(asdFftHandle handle)  // fft_api.cpp:1234-1243
{
    std::shared_ptr<FftPlan> victim;
    {
        std::lock_guard<std::mutex> lock(g_table_mu);
        auto it = table().find(handle);
        if (it == table().end()) {
            log_error("asdFftDestroy", "invalid handle");
            return ACL_ERROR_INVALID_PARAM;
        }
        victim = std::move(it->second);
        table().erase(it);
    }
    // wait for any in-flight exec on another thread before the tables go away
    std::lock_guard<std::mutex> lock(victim->mu);
    return ACL_SUCCESS;
}

AspbStatus asdFftExecC2C(asdFftHandle handle, const aclTensor* input, const aclTensor* output)  // :1752-1788
{
    return exec_common(handle, input, output, ASCEND_FFT_C2C, match_c2c, match_c2c, "asdFftExecC2C");
}

AspbStatus asdFftExecC2R(asdFftHandle handle, const aclTensor* input, const aclTensor* output)  // :1790-1816
{
    return exec_common(handle, input, output, ASCEND_FFT_C2R, match_half_complex, match_real, "asdFftExecC2R");
}

AspbStatus asdFftExecR2C(asdFftHandle handle, const aclTensor* input, const aclTensor* output)  // :1818-1844
{
    return exec_common(handle, input, output, ASCEND_FFT_R2C, match_real, match_half_complex, "asdFftExecR2C");
}

AspbStatus asdFftIstftMakePlan(asdFftHandle handle, const aclTensor* input, const int64_t nFft,
                               const int64_t hopLengthOpt, const int64_t winLengthOpt, const bool center,
                               const bool normalized, const bool onesidedOpt, int64_t lengthOpt,
                               const bool returnComplex)  // core/fft/istft.cpp:525-570
{
    const char* fn = "asdFftIstftMakePlan";
    auto sp = lookup(handle);
    if (!sp) {
        log_error(fn, "invalid handle");
        return ACL_ERROR_INVALID_PARAM;
    }
    // IstftParamsCheck (istft.cpp:343-433): the only supported form is center, not normalised,
    // two-sided, full length, complex output; other flag values are replaced with a warning
    if (input == nullptr || input->dtype != ACL_COMPLEX64) {
        log_error(fn, "Invalid input or invalid input dtype! Only Support complex64!");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (input->view.size() != 3) {
        log_error(fn, "Invalid input! Only 3-dimensional tensors are supported!");
        return ACL_ERROR_INVALID_PARAM;
    }
    IstftDesc d;
    d.channel = input->view[0];
    const int64_t fft_size = input->view[1];
    d.nframes = input->view[2];
    if (!center) log_error(fn, "warning: Currently, the supported center is only true.");
    if (normalized) log_error(fn, "warning: Currently, the supported normalized is only false.");
    if (onesidedOpt) log_error(fn, "warning: Currently, the supported onesidedOpt is only false.");
    if (!returnComplex) log_error(fn, "warning: Currently, the supported returnComplex is only true.");
    if (d.nframes <= 1) {
        log_error(fn, "Currently, the nFrames dim value of input should be more than 1");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (lengthOpt != 0) log_error(fn, "warning: Currently, the supported lengthOpt is only 0.");
    if (nFft != fft_size) {
        log_error(fn, "expected the frequency dimension of the input tensor to match n_fft when onesided=False");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (!(0 < hopLengthOpt && hopLengthOpt <= winLengthOpt)) {
        log_error(fn, "expected 0 < hop_length <= win_length");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (!(winLengthOpt > 0 && winLengthOpt == nFft)) {
        log_error(fn, "expected 0 < win_length == n_fft");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (nFft >= 1500) {
        log_error(fn, "Currently, the nfft size should be less than 1500");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (hopLengthOpt >= 1500) {
        log_error(fn, "Currently, the hoplen size should be less than 1500.");
        return ACL_ERROR_INVALID_PARAM;
    }
    d.nfft = nFft;
    d.hop = hopLengthOpt;
    d.win_length = winLengthOpt;
    d.out_len = nFft + hopLengthOpt * (d.nframes - 1) - 2 * (nFft / 2);
    // MakePlan1DFft (istft.cpp:495-518): the frame transforms are a batch of channel * nframes
    // inverse C2C; where that plan cannot be made the status is INTERNAL_ERROR
    FftPlan& p = *sp;
    std::lock_guard<std::mutex> lock(p.mu);
    if (p.state == PlanState::kInitialized) {
        log_error(fn, "plan is already initialized");
        return ACL_ERROR_INTERNAL_ERROR;
    }
    const int64_t imax = std::numeric_limits<int32_t>::max();
    if (d.channel <= 0 || d.nframes > imax / d.channel) {
        log_error(fn, "channel * nFrames must be in [1, INT32_MAX]");
        return ACL_ERROR_INTERNAL_ERROR;
    }
    asdfft::BuiltPlan built;
    std::string why;
    try {
        why = asdfft::build_istft_frames(d.nfft, d.channel, d.nframes, &built);
    } catch (const std::exception& e) {
        why = e.what();
    }
    if (!why.empty()) {
        log_error(fn, why);
        return ACL_ERROR_INTERNAL_ERROR;
    }
    p.type = ASCEND_FFT_C2C;
    p.direction = ASCEND_FFT_INVERSE;
    p.sizes = {nFft};
    p.batch = d.channel * d.nframes;
    p.vertical = false;
    p.vertical_batch = 1;
    p.built = std::move(built);
    p.istft.reset(new IstftDesc(d));
    p.state = PlanState::kInitialized;
    if (trace_enabled()) fprintf(stderr, "[asdfft] %s", asdfft::describe_plan(p.built).c_str());
    return ACL_SUCCESS;
}

AspbStatus asdFftExecIstft(asdFftHandle handle, const aclTensor* input, const aclTensor* windowOpt,
                           const aclTensor* output)  // core/fft/istft.cpp:572-595, 248-316
{
    const char* fn = "asdFftExecIstft";
    auto sp = lookup(handle);
    if (!sp) {
        log_error(fn, "invalid handle");
        return ACL_ERROR_INVALID_PARAM;
    }
    FftPlan& p = *sp;
    std::lock_guard<std::mutex> lock(p.mu);
    // MatchIstftC2C (istft.cpp:189-222): the spectrum shape recorded by the plan
    const IstftDesc* d = p.istft.get();
    if (d == nullptr || input == nullptr || input->dtype != ACL_COMPLEX64 || input->view.size() != 3 ||
        input->view[0] != d->channel || input->view[1] != d->nfft || input->view[2] != d->nframes) {
        log_error(fn, "input does not match c2c plan");
        return ACL_ERROR_INVALID_PARAM;
    }
    // IstftOutTensorCheck, then IstftWindowTensorCheck (istft.cpp:437-523)
    if (output == nullptr || output->dtype != ACL_COMPLEX64 || output->view.size() != 2 ||
        output->view[0] != d->channel || output->view[1] != d->out_len) {
        log_error(fn, "Invalid out tensor: expected complex64 (channel, " + std::to_string(d->out_len) + ")");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (windowOpt == nullptr || windowOpt->dtype != ACL_FLOAT || windowOpt->view.size() != 1 ||
        windowOpt->view[0] != d->win_length || windowOpt->view[0] != d->nfft) {
        log_error(fn, "Invalid window tensor: expected float (n_fft)");
        return ACL_ERROR_INVALID_PARAM;
    }
    // asdFftExecIstftV2 (istft.cpp:224-286)
    if (p.state != PlanState::kInitialized) {
        log_error(fn, "plan is not initialized");
        return ACL_ERROR_INVALID_PARAM;
    }
    const AspbStatus ws = check_workspace(p, fn);
    if (ws != ACL_SUCCESS) return ws;
    char* in = static_cast<char*>(asdfft::data_address(input));
    char* win = static_cast<char*>(asdfft::data_address(windowOpt));
    char* out = static_cast<char*>(asdfft::data_address(output));
    if (in == nullptr || win == nullptr || out == nullptr) {
        log_error(fn, "input, window or output data is nullptr");
        return ACL_ERROR_INTERNAL_ERROR;
    }
    if (!asdfft::is_dense(input) || !asdfft::is_dense(windowOpt) || !asdfft::is_dense(output)) {
        log_error(fn, "only dense row-major tensors are supported");
        return ACL_ERROR_INVALID_PARAM;
    }
    // no in-place update (the reference's documented constraint); the output and the frames in
    // the workspace must not overlap the operands either
    const size_t in_bytes = (size_t)asdfft::numel(input) * 8;
    const size_t out_bytes = (size_t)asdfft::numel(output) * 8;
    const size_t win_bytes = (size_t)d->nfft * 4;
    const char* work = static_cast<const char*>(p.workspace);
    if (ranges_overlap(out, out_bytes, in, in_bytes) || ranges_overlap(out, out_bytes, win, win_bytes) ||
        ranges_overlap(out, out_bytes, work, p.built.workspace) || ranges_overlap(in, in_bytes, work, p.built.workspace) ||
        ranges_overlap(win, win_bytes, work, p.built.workspace)) {
        log_error(fn, "output, input, window and workspace must not overlap");
        return ACL_ERROR_INVALID_PARAM;
    }
    const AspbStatus dev = check_device(p.built, fn);
    if (dev != ACL_SUCCESS) return dev;
    hipStream_t stream = static_cast<hipStream_t>(p.stream);
    char* bufs[3] = {in, out, static_cast<char*>(p.workspace)};
    const AspbStatus st = launch_steps(p.built, bufs, nullptr, nullptr, stream, fn, -1);
    if (st != ACL_SUCCESS) return st;
    asdfft::OlaArgs a{};
    a.frames = p.workspace;
    a.window = reinterpret_cast<const float*>(win);
    a.out = out;
    a.nfft = (int)d->nfft;
    a.hop = (int)d->hop;
    a.nframes = (int)d->nframes;
    a.start = (int)(d->nfft / 2);
    a.channel = d->channel;
    a.out_len = d->out_len;
    a.scale = 1.0f / (float)d->nfft;
    const hipError_t e = asdfft::launch_istft_ola(a, stream);
    if (e != hipSuccess) {
        log_error(fn, std::string("kernel launch failed: ") + hipGetErrorString(e));
        return ACL_ERROR_INTERNAL_ERROR;
    }
    return ACL_SUCCESS;
}

AspbStatus asdFftExecC2CSeparated(asdFftHandle handle, const aclTensor* inputReal, const aclTensor* inputImag,
                                  const aclTensor* outputReal, const aclTensor* outputImag)  // :1912-1985
{
    const char* fn = "asdFftExecC2CSeparated";
    auto sp = lookup(handle);
    if (!sp) {
        log_error(fn, "invalid handle");
        return ACL_ERROR_INVALID_PARAM;
    }
    const aclTensor* t[4] = {inputReal, inputImag, outputReal, outputImag};
    const char* names[4] = {"inputReal", "inputImag", "outputReal", "outputImag"};
    for (int i = 0; i < 4; i++) {
        if (t[i] == nullptr) {  // aclGetViewShape fails
            log_error(fn, std::string("no view shape for ") + names[i]);
            return ACL_ERROR_INTERNAL_ERROR;
        }
    }
    for (int i = 1; i < 4; i++) {  // four equal shapes (rank, then extents)
        if (t[i]->view != t[0]->view) {
            log_error(fn, "inputReal, inputImag, outputReal and outputImag must have one shape");
            return ACL_ERROR_INTERNAL_ERROR;
        }
    }
    for (int i = 0; i < 4; i++) {
        if (t[i]->dtype != ACL_FLOAT) {
            log_error(fn, std::string("requires a FLOAT ") + names[i] + " tensor");
            return ACL_ERROR_OP_INPUT_NOT_MATCH;
        }
    }
    FftPlan& p = *sp;
    std::lock_guard<std::mutex> lock(p.mu);
    if (p.state != PlanState::kInitialized) {
        log_error(fn, "plan is not initialized");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (p.ragged) {
        log_error(fn, "ragged-batch plans run through asdFftExtExecRagged");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (!match_separated(p, t[0]->view)) {
        log_error(fn, "input/output shape does not match plan");
        return ACL_ERROR_INVALID_PARAM;
    }
    // asdFftExecV2Separated (:1846-1910)
    const AspbStatus ws = check_workspace(p, fn);
    if (ws != ACL_SUCCESS) return ws;
    char* data[4];
    for (int i = 0; i < 4; i++) {
        data[i] = static_cast<char*>(asdfft::data_address(t[i]));
        if (data[i] == nullptr) {
            log_error(fn, std::string(names[i]) + " data is nullptr");
            return ACL_ERROR_INTERNAL_ERROR;
        }
    }
    if (p.type != ASCEND_FFT_C2C_SEP) {  // only separated plans have plane kernels (:1889-1898)
        log_error(fn, "the plan does not implement separated execution (make it with ASCEND_FFT_C2C_SEP)");
        return ACL_ERROR_INTERNAL_ERROR;
    }
    for (int i = 0; i < 4; i++) {
        if (!asdfft::is_dense(t[i])) {
            log_error(fn, "only dense row-major tensors are supported");
            return ACL_ERROR_INVALID_PARAM;
        }
    }
    // the planes may be updated in place pairwise (outputReal == inputReal and outputImag ==
    // inputImag); any other overlap between them is rejected
    const size_t bytes = (size_t)asdfft::numel(t[0]) * sizeof(float);
    const int pairs[6][2] = {{0, 1}, {2, 3}, {0, 3}, {1, 2}, {0, 2}, {1, 3}};
    for (int k = 0; k < 6; k++) {
        const char* x = data[pairs[k][0]];
        const char* y = data[pairs[k][1]];
        const bool in_place = (k >= 4 && x == y);
        if (!in_place && ranges_overlap(x, bytes, y, bytes)) {
            log_error(fn, "planes overlap (only outputReal == inputReal and outputImag == inputImag in place)");
            return ACL_ERROR_INVALID_PARAM;
        }
    }
    const AspbStatus dev = check_device(p.built, fn);
    if (dev != ACL_SUCCESS) return dev;
    char* bufs[3] = {data[0], data[2], static_cast<char*>(p.workspace)};
    return launch_steps(p.built, bufs, reinterpret_cast<const float*>(data[1]), reinterpret_cast<float*>(data[3]),
                        static_cast<hipStream_t>(p.stream), fn, -1);
}

AspbStatus asdFftShift(int64_t row, int64_t col, int64_t batch, int64_t dist, const aclTensor* input,
                       aclTensor* output, void* stream, void* workspace)  // core/fft/fftshift.cpp:101-138
{
    (void)workspace;
    return run_shift("asdFftShift", row, col, batch, dist, 1.0, false, input, output, stream);
}

AspbStatus asdIfftShift(int64_t row, int64_t col, int64_t batch, int64_t dist, double scale, const aclTensor* input,
                        aclTensor* output, void* stream, void* workspace)  // core/fft/ifftshift.cpp
{
    (void)workspace;
    return run_shift("asdIfftShift", row, col, batch, dist, scale, true, input, output, stream);
}

}  // extern "C"

}  // namespace AsdSip

// ---- extension entry points (asdfft_ext.h) -------------------------------------------------
extern "C" int32_t asdFftExtDescribePlan(void* handle, char* buffer, size_t length, size_t* needed)
{
    auto sp = lookup(handle);
    if (!sp) return AsdSip::ACL_ERROR_INVALID_PARAM;
    std::lock_guard<std::mutex> lock(sp->mu);
    const std::string s = sp->ragged ? sp->ragged->describe : asdfft::describe_plan(sp->built);
    if (needed != nullptr) *needed = s.size() + 1;
    if (buffer != nullptr && length > 0) {
        const size_t n = std::min(length - 1, s.size());
        memcpy(buffer, s.data(), n);
        buffer[n] = '\0';
    }
    return AsdSip::ACL_SUCCESS;
}

extern "C" int32_t asdFftExtPlanSteps(void* handle, int32_t* steps, double* algorithmic_bytes)
{
    auto sp = lookup(handle);
    if (!sp) return AsdSip::ACL_ERROR_INVALID_PARAM;
    std::lock_guard<std::mutex> lock(sp->mu);
    if (sp->ragged) {
        const asdfft::RaggedPlan& r = *sp->ragged;
        if (steps != nullptr) *steps = (int32_t)r.launches.size();
        if (algorithmic_bytes != nullptr) {
            for (size_t i = 0; i < r.launches.size(); i++) {
                double b = 0;
                for (int ci : r.launches[i].class_ids) b += r.classes[ci].bytes;
                algorithmic_bytes[i] = b;
            }
        }
        return AsdSip::ACL_SUCCESS;
    }
    if (steps != nullptr) *steps = (int32_t)sp->built.steps.size();
    if (algorithmic_bytes != nullptr) {
        for (size_t i = 0; i < sp->built.steps.size(); i++) algorithmic_bytes[i] = sp->built.steps[i].bytes;
    }
    return AsdSip::ACL_SUCCESS;
}

extern "C" int32_t asdFftExtExecStep(void* handle, const aclTensor* input, const aclTensor* output, int32_t step)
{
    auto sp = lookup(handle);
    if (!sp) return AsdSip::ACL_ERROR_INVALID_PARAM;
    asdFftType t;
    {
        std::lock_guard<std::mutex> lock(sp->mu);
        t = sp->type;
    }
    if (step < 0) return AsdSip::ACL_ERROR_INVALID_PARAM;
    if (t == AsdSip::ASCEND_FFT_R2C)
        return exec_common(handle, input, output, t, match_real, match_half_complex, "asdFftExtExecStep", step);
    if (t == AsdSip::ASCEND_FFT_C2R)
        return exec_common(handle, input, output, t, match_half_complex, match_real, "asdFftExtExecStep", step);
    return exec_common(handle, input, output, t, match_c2c, match_c2c, "asdFftExtExecStep", step);
}

static_assert(ASDFFT_NOALLOC_SCRATCH_BYTES == asdfft::kStaticScratchBytes, "documented scratch size");

extern "C" int32_t asdFftExtExecNoAlloc(void* handle, const aclTensor* input, const aclTensor* output)
{
    const char* fn = "asdFftExtExecNoAlloc";
    auto sp = lookup(handle);
    if (!sp) {
        log_error(fn, "invalid handle");
        return AsdSip::ACL_ERROR_INVALID_PARAM;
    }
    asdFftType t;
    {
        std::lock_guard<std::mutex> lock(sp->mu);
        t = sp->type;
    }
    if (t == AsdSip::ASCEND_FFT_R2C) return exec_common(handle, input, output, t, match_real, match_half_complex, fn, -1, true);
    if (t == AsdSip::ASCEND_FFT_C2R) return exec_common(handle, input, output, t, match_half_complex, match_real, fn, -1, true);
    if (t == AsdSip::ASCEND_FFT_C2C) return exec_common(handle, input, output, t, match_c2c, match_c2c, fn, -1, true);
    log_error(fn, "only C2C, R2C and C2R plans run through this entry point");
    return AsdSip::ACL_ERROR_INVALID_PARAM;
}

extern "C" int32_t asdFftExtMakePlanRagged(void* handle, const asdFftExtRaggedEntry* entries, int64_t num_entries,
                                           uint32_t flags)
{
    const char* fn = "asdFftExtMakePlanRagged";
    auto sp = lookup(handle);
    if (!sp) {
        log_error(fn, "invalid handle");
        return ACL_ERROR_INVALID_PARAM;
    }
    FftPlan& p = *sp;
    std::lock_guard<std::mutex> lock(p.mu);
    if (p.state == PlanState::kInitialized) {
        log_error(fn, "plan is already initialized");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (entries == nullptr || num_entries <= 0 || num_entries > std::numeric_limits<int32_t>::max()) {
        log_error(fn, "entries must be a non-empty array");
        return ACL_ERROR_INVALID_PARAM;
    }
    if ((flags & ~(ASDFFT_RAGGED_GRAPH | ASDFFT_RAGGED_AUTOTUNE | ASDFFT_RAGGED_PER_BUDGET)) != 0) {
        log_error(fn, "unknown flags");
        return ACL_ERROR_INVALID_PARAM;
    }
    std::vector<asdfft::RaggedEntry> list;
    list.reserve((size_t)num_entries);
    for (int64_t i = 0; i < num_entries; i++) {
        const asdFftExtRaggedEntry& e = entries[i];
        asdfft::RaggedEntry r;
        if (e.type == ASCEND_FFT_C2C) {
            r.kind = asdfft::RaggedKind::C2C;
        } else if (e.type == ASCEND_FFT_R2C) {
            r.kind = asdfft::RaggedKind::R2C;
        } else if (e.type == ASCEND_FFT_C2R) {
            r.kind = asdfft::RaggedKind::C2R;
        } else {
            log_error(fn, "entry " + std::to_string(i) + ": unsupported type");
            return ACL_ERROR_INVALID_PARAM;
        }
        if (!valid_direction(static_cast<asdFftDirection>(e.direction))) {
            log_error(fn, "entry " + std::to_string(i) + ": invalid direction");
            return ACL_ERROR_INVALID_PARAM;
        }
        if (e.n <= 0 || e.n > kMaxFftSize || e.count <= 0 || e.count > std::numeric_limits<int32_t>::max()) {
            log_error(fn, "entry " + std::to_string(i) + ": invalid length or count");
            return ACL_ERROR_INVALID_PARAM;
        }
        r.inverse = (e.direction == ASCEND_FFT_INVERSE);
        r.n = e.n;
        r.count = e.count;
        r.in_offset = e.in_offset;
        r.out_offset = e.out_offset;
        list.push_back(r);
    }
    asdfft::RaggedOptions opt;
    opt.graph = (flags & ASDFFT_RAGGED_GRAPH) != 0;
    opt.autotune = (flags & ASDFFT_RAGGED_AUTOTUNE) != 0;
    opt.one_launch = (flags & ASDFFT_RAGGED_PER_BUDGET) == 0;
    // ASDFFT_RAGGED_WIDE16=1: the EMAX-16 classes on wide workgroups (up to 1024 threads per
    // transform) in a launch of their own beside the single launch of the others (measured
    // within 3% of the single launch: 100.6 vs 103.3 us on the 4000-transform workload)
    opt.wide16 = false;
    if (const char* w = getenv("ASDFFT_RAGGED_WIDE16")) opt.wide16 = opt.one_launch && w[0] == '1';
    auto plan = std::make_unique<asdfft::RaggedPlan>();
    bool unsupported = false;
    std::string why;
    try {
        why = asdfft::build_ragged(list, opt, plan.get(), &unsupported);
    } catch (const std::exception& e) {
        why = e.what();
    }
    if (!why.empty()) {
        log_error(fn, why);
        return unsupported ? ACL_ERROR_API_NOT_SUPPORT : ACL_ERROR_INVALID_PARAM;
    }
    p.type = static_cast<asdFftType>(entries[0].type);
    p.direction = static_cast<asdFftDirection>(entries[0].direction);
    if (trace_enabled()) fprintf(stderr, "[asdfft] %s", plan->describe.c_str());
    p.ragged = std::move(plan);
    p.state = PlanState::kInitialized;
    return ACL_SUCCESS;
}

extern "C" int32_t asdFftExtExecRagged(void* handle, const void* in_base, void* out_base)
{
    const char* fn = "asdFftExtExecRagged";
    auto sp = lookup(handle);
    if (!sp) {
        log_error(fn, "invalid handle");
        return ACL_ERROR_INVALID_PARAM;
    }
    FftPlan& p = *sp;
    std::lock_guard<std::mutex> lock(p.mu);
    if (p.state != PlanState::kInitialized || !p.ragged) {
        log_error(fn, "handle holds no ragged-batch plan");
        return ACL_ERROR_INVALID_PARAM;
    }
    asdfft::RaggedPlan& r = *p.ragged;
    if (in_base == nullptr || out_base == nullptr) {
        log_error(fn, "data pointer is nullptr");
        return ACL_ERROR_INTERNAL_ERROR;
    }
    const bool in_place = (in_base == out_base) && r.all_inplace_c2c;
    if (!in_place && ranges_overlap(in_base, r.in_extent, out_base, r.out_extent)) {
        log_error(fn, "input and output ranges overlap");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (!asdfft::device_available()) {
        log_error(fn, "no HIP device: this library has no host execution path");
        return ACL_ERROR_INTERNAL_ERROR;
    }
    if (!r.ready) {
        log_error(fn, "plan was built without a HIP device; create it again on a device");
        return ACL_ERROR_INTERNAL_ERROR;
    }
    const hipError_t e = asdfft::exec_ragged(r, in_base, out_base, static_cast<hipStream_t>(p.stream));
    if (e != hipSuccess) {
        log_error(fn, std::string("launch failed: ") + hipGetErrorString(e));
        return ACL_ERROR_INTERNAL_ERROR;
    }
    return ACL_SUCCESS;
}

extern "C" int32_t asdFftExtRaggedInfo(void* handle, int64_t* in_extent, int64_t* out_extent, double* bytes,
                                       double* flops)
{
    auto sp = lookup(handle);
    if (!sp) return ACL_ERROR_INVALID_PARAM;
    std::lock_guard<std::mutex> lock(sp->mu);
    if (!sp->ragged) return ACL_ERROR_INVALID_PARAM;
    const asdfft::RaggedPlan& r = *sp->ragged;
    if (in_extent != nullptr) *in_extent = (int64_t)r.in_extent;
    if (out_extent != nullptr) *out_extent = (int64_t)r.out_extent;
    if (bytes != nullptr) *bytes = r.bytes;
    if (flops != nullptr) *flops = r.flops;
    return ACL_SUCCESS;
}

extern "C" int32_t asdFftExtSetShift(void* handle, int32_t in_shift, int32_t out_shift, double scale)
{
    const char* fn = "asdFftExtSetShift";
    auto sp = lookup(handle);
    if (!sp) {
        log_error(fn, "invalid handle");
        return ACL_ERROR_INVALID_PARAM;
    }
    std::lock_guard<std::mutex> lock(sp->mu);
    if (sp->state == PlanState::kInitialized) {
        log_error(fn, "shifts must be chosen before the plan is made");
        return ACL_ERROR_INVALID_PARAM;
    }
    auto valid = [](int32_t m) { return m == ASDFFT_SHIFT_NONE || m == ASDFFT_SHIFT_FFT || m == ASDFFT_SHIFT_IFFT; };
    if (!valid(in_shift) || !valid(out_shift)) {
        log_error(fn, "unknown shift mode");
        return ACL_ERROR_INVALID_PARAM;
    }
    const double fmax = (double)std::numeric_limits<float>::max();
    if (!(scale >= -fmax && scale <= fmax)) {  // NaN fails both comparisons
        log_error(fn, "scale must be a finite float32 value");
        return ACL_ERROR_INVALID_PARAM;
    }
    sp->in_shift = in_shift == ASDFFT_SHIFT_FFT ? asdfft::kShiftFft
                                                 : in_shift == ASDFFT_SHIFT_IFFT ? asdfft::kShiftIfft : asdfft::kShiftNone;
    sp->out_shift = out_shift == ASDFFT_SHIFT_FFT ? asdfft::kShiftFft
                                                   : out_shift == ASDFFT_SHIFT_IFFT ? asdfft::kShiftIfft
                                                                                    : asdfft::kShiftNone;
    sp->out_scale = (float)scale;
    return ACL_SUCCESS;
}

extern "C" int32_t asdFftExtSetPrecision(void* handle, int32_t mode)
{
    const char* fn = "asdFftExtSetPrecision";
    auto sp = lookup(handle);
    if (!sp) {
        log_error(fn, "invalid handle");
        return ACL_ERROR_INVALID_PARAM;
    }
    std::lock_guard<std::mutex> lock(sp->mu);
    if (sp->state == PlanState::kInitialized) {
        log_error(fn, "precision must be chosen before the plan is made");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (mode != ASDFFT_PRECISION_FP32 && mode != ASDFFT_PRECISION_SPLIT16) {
        log_error(fn, "unknown precision mode");
        return ACL_ERROR_INVALID_PARAM;
    }
    sp->precision = mode == ASDFFT_PRECISION_SPLIT16 ? asdfft::Precision::kSplit16 : asdfft::Precision::kFp32;
    return ACL_SUCCESS;
}
