// Ragged batches: many transforms of different lengths and kinds (C2C, R2C, C2R) run
// as one persistent launch per register budget, driven by a device-side work table,
// optionally replayed from a hipGraph.  The reference has no counterpart; its nearest
// equivalent is one plan and one exec per distinct (type, length)
// (core/fft/fft_api.cpp:1124-1190 per size, sip_pta/csrc/FFTMixCache.cpp:18-40 caching).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "fft_kernel_args.h"
#include "kernels.h"
#include "twiddles.h"

namespace asdfft {

// Largest workgroup of the ragged kernel (its __launch_bounds__): 1024 threads for the EMAX-16
// kernel (at most 128 VGPRs, which that budget meets), 512 for the EMAX-32 and EMAX-48 ones.
constexpr int kRaggedMaxBlock = 512;
constexpr int kRaggedMaxBlock16 = 1024;
constexpr int ragged_max_block(int emax) { return emax <= 16 ? kRaggedMaxBlock16 : kRaggedMaxBlock; }
// Dynamic LDS a ragged launch may take: the workgroup's 160 KiB less room for the
// kernel's static tile index.
constexpr size_t kRaggedLdsLimit = 160 * 1024 - 256;

// ---- kernel side (ragged_kernel.hip) --------------------------------------------------
const void* ragged_kernel_fn(int emax);
// Raises the dynamic LDS limit when needed and reports resident workgroups per CU.
hipError_t ragged_prepare(int emax, size_t lds_bytes, int block, int* per_cu);
hipError_t launch_ragged(int emax, int64_t grid, int block, size_t lds_bytes, const RaggedArgs& args,
                         const RaggedClass* classes, const RaggedTile* tiles, hipStream_t stream);

// ---- host side (ragged.cpp) -------------------------------------------------------------
enum class RaggedKind { C2C, R2C, C2R };

struct RaggedEntry {
    RaggedKind kind = RaggedKind::C2C;
    bool inverse = false;
    int64_t n = 0;           // transform length (real length for R2C / C2R)
    int64_t count = 0;       // transforms stored back to back
    int64_t in_offset = 0;   // bytes from the input base
    int64_t out_offset = 0;  // bytes from the output base
};

// One class's schedule (what the autotuner varies).
struct ClassSchedule {
    std::vector<int> radices;  // pass order
    int threads = 0;           // threads per transform (>= the minimum for EMAX)
    int tpb = 0;               // transforms per tile (0: as many as the workgroup holds)
};

struct RaggedOptions {
    bool graph = false;     // replay the launches from an instantiated hipGraph
    bool autotune = false;  // time candidate schedules on the device at plan time
    bool one_launch = true;  // every class in one launch at the largest register budget (false:
                             // one concurrent launch per register budget)
    bool wide16 = false;     // EMAX-16 schedules of up to 1024 threads per transform before larger budgets
                             // (the EMAX-16 classes then take a launch of their own: two launches)
    int block16 = 0;        // workgroup size of the EMAX-16 launch (0: automatic)
    int block32 = 0;        // workgroup size of the EMAX-32 launch (0: automatic)
    int64_t max_grid = 0;   // workgroups per launch at most (0: one per tile)
    // schedules fixed by the caller or a previous tuning run, keyed by class key
    std::map<std::string, ClassSchedule> fixed;
};

struct RaggedClassInfo {
    std::string key;    // "c2c:f:4096", "r2c:i:1000", ...
    RaggedKind kind = RaggedKind::C2C;
    bool inverse = false;
    int variant = 0;
    int emax = 16;
    int64_t n = 0;      // complex length the Stockham passes run on
    int64_t nreal = 0;  // length given by the caller
    int64_t transforms = 0;
    ClassSchedule sched;
    size_t lds = 0;
    double bytes = 0;   // algorithmic bytes of all its transforms
};

struct RaggedLaunchPlan {
    int emax = 16;
    int block = 256;
    size_t lds_bytes = 0;
    int64_t grid = 0;
    int64_t ntiles = 0;
    std::vector<int> class_ids;  // indices into RaggedPlan::classes
    std::vector<RaggedClass> dev_classes;
    std::vector<RaggedTile> tiles;
    RaggedClass* d_classes = nullptr;
    RaggedTile* d_tiles = nullptr;
};

struct RaggedPlan {
    std::vector<RaggedEntry> entries;
    double tune_ms = 0;            // plan-time autotuning wall time (0 when not tuned)
    int tune_timed = 0;            // candidate schedules timed
    std::vector<RaggedClassInfo> classes;
    std::vector<RaggedLaunchPlan> launches;
    std::vector<int64_t> offsets;  // host copy of the offset table
    int64_t* d_offsets = nullptr;
    std::vector<TablePtr> tables;
    int device = -1;
    bool ready = false;            // device tables uploaded
    size_t in_extent = 0, out_extent = 0;  // bytes touched past each base
    bool all_inplace_c2c = false;
    double bytes = 0;
    double flops = 0;              // 5 N log2 N per complex-equivalent transform
    RaggedOptions opt;
    // stream exec: helper streams and events joining the register budgets' launches
    std::vector<hipStream_t> aux;
    std::vector<hipEvent_t> joins;
    hipEvent_t fork = nullptr;
    // graph replay state
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<hipGraphNode_t> nodes;
    const void* graph_in = nullptr;
    void* graph_out = nullptr;
    std::string describe;
    ~RaggedPlan();
};

// ---- class helpers (shared with the serving loop) ------------------------------------
// Default workgroup of a register budget: 256 threads hold EMAX-16 transforms of up to
// 4096 points (four workgroups per CU at 34 KiB of LDS each); larger budgets use 512.
int ragged_block_for(int emax);

// Kernel variant, complex pass length, conjugation and load alignment of a class (the
// same choices the batched planner makes for a plain 1D plan).
struct ClassShape {
    int variant;
    int64_t n;  // complex length of the passes
    int conj_in, conj_out;
    bool aux;   // needs the half-length W_N^k table
    int in_align, out_align;
};
bool ragged_shape(const RaggedEntry& e, ClassShape* s);
std::string ragged_class_key(const RaggedEntry& e);  // "c2c:f:4096", ...
int64_t ragged_in_bytes(const RaggedEntry& e);       // bytes of one transform's input
int64_t ragged_out_bytes(const RaggedEntry& e);
// Fills c->sched (the given schedule, or the planner's default) and c->emax; "" on success.
// max_t16: most threads per transform at the EMAX-16 budget (ragged_block_for(16) unless the
// caller runs wide EMAX-16 workgroups).
std::string ragged_schedule(RaggedClassInfo* c, const ClassSchedule* fixed, bool* unsupported, int max_t16 = 256);
size_t ragged_class_lds(const RaggedClassInfo& c, int tpb);
// Device descriptor of a class at tpb transforms per tile; uploads its twiddle tables.
bool ragged_device_class(const RaggedClassInfo& c, const ClassShape& s, int tpb, RaggedClass* rc,
                         std::vector<TablePtr>* tables);

// Validates the entries and builds the plan ("" on success, otherwise the reason).
// `unsupported` is set when the entries are valid but outside what the kernels handle.
std::string build_ragged(const std::vector<RaggedEntry>& entries, const RaggedOptions& opt, RaggedPlan* plan,
                         bool* unsupported);

// Enqueues the batch on `stream` (graph replay when the plan was built with graph = true).
hipError_t exec_ragged(RaggedPlan& plan, const void* in_base, void* out_base, hipStream_t stream);

// One line per class: key, radix order, threads per transform, transforms per tile.
std::string format_schedules(const RaggedPlan& plan);

// Candidate schedules the autotuner times for one class (radix multisets within one
// pass of the fewest, in descending and ascending order, at the minimum and at twice
// the minimum threads per transform).
std::vector<ClassSchedule> class_candidates(int64_t n, int limit);

// Drops the process-wide cache of tuned schedules (measurement and tests).
void clear_tuning_cache();

}  // namespace asdfft
