// Host planner of ragged batches (see ragged.h).
//
// Entries with the same (kind, direction, length) form a class.  Each class gets a
// Stockham schedule (radix order, threads per transform, transforms per tile) and an
// EMAX register budget; classes with one budget share one persistent launch.  Every
// class's transforms are listed in the offset table, cut into tiles of tpb transforms,
// and the tiles of a launch are sorted by decreasing cost (longest-processing-time
// first), so the workgroups the hardware starts last carry the cheapest tiles.
#include "ragged.h"

#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <sstream>

#include "planner.h"

namespace asdfft {

namespace {

constexpr size_t kLdsLimit = kRaggedLdsLimit;
constexpr size_t kLdsFloor = 64 * 1024;
constexpr int64_t kMaxComplexLength = 16384;  // one class's transform fits one workgroup's LDS

int64_t pad16(int64_t x) { return x + (x >> 4); }


uint32_t magic_for(int span)
{
    if (span <= 1) return 0;
    return (uint32_t)((((uint64_t)1) << 32) / (uint64_t)span + 1);
}

std::string table_key(int64_t n, const std::vector<int>& r)
{
    std::ostringstream os;
    os << "stk:" << n;
    for (int v : r) os << ":" << v;
    return os.str();
}

const char* kind_name(RaggedKind k)
{
    switch (k) {
        case RaggedKind::C2C: return "c2c";
        case RaggedKind::R2C: return "r2c";
        default: return "c2r";
    }
}

// Threads per transform so that every pass has EMAX / R butterflies per thread at most.
int min_threads(int64_t n, const std::vector<int>& radices, int emax)
{
    int64_t t = 0;
    for (int r : radices) {
        if (r > emax) return 0;
        const int64_t b = emax / r;
        t = std::max(t, (n / r + b - 1) / b);
    }
    return (int)t;
}

bool need_lds(const RaggedClassInfo& c)
{
    return c.sched.radices.size() > 1 || c.variant == kVarR2CEven;
}

size_t class_lds(const RaggedClassInfo& c, int tpb)
{
    return need_lds(c) ? (size_t)8 * (size_t)tpb * (size_t)(pad16(c.n) + 1) : 0;
}

int device_cus()
{
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) ==
                                                 hipSuccess && n > 0) {
        return n;
    }
    return 256;
}

void free_device(void* p)
{
    if (p != nullptr) (void)hipFree(p);
}

}  // namespace

int ragged_block_for(int emax) { return emax <= 16 ? 256 : 512; }

// ---- dispatch over the per-budget kernel translation units
#define ASDFFT_RAGGED_DECL(E)                                                                        \
    const void* ragged_kernel_fn_e##E();                                                             \
    hipError_t ragged_prepare_e##E(size_t lds_bytes, int block, int* per_cu);                         \
    hipError_t launch_ragged_e##E(int64_t grid, int block, size_t lds_bytes, const RaggedArgs& args, \
                                  const RaggedClass* classes, const RaggedTile* tiles, hipStream_t stream);
ASDFFT_RAGGED_DECL(16)
ASDFFT_RAGGED_DECL(32)
ASDFFT_RAGGED_DECL(48)
#undef ASDFFT_RAGGED_DECL

const void* ragged_kernel_fn(int emax)
{
    if (emax <= 16) return ragged_kernel_fn_e16();
    return emax <= 32 ? ragged_kernel_fn_e32() : ragged_kernel_fn_e48();
}

hipError_t ragged_prepare(int emax, size_t lds_bytes, int block, int* per_cu)
{
    if (emax <= 16) return ragged_prepare_e16(lds_bytes, block, per_cu);
    return emax <= 32 ? ragged_prepare_e32(lds_bytes, block, per_cu) : ragged_prepare_e48(lds_bytes, block, per_cu);
}

hipError_t launch_ragged(int emax, int64_t grid, int block, size_t lds_bytes, const RaggedArgs& args,
                         const RaggedClass* classes, const RaggedTile* tiles, hipStream_t stream)
{
    if (emax <= 16) return launch_ragged_e16(grid, block, lds_bytes, args, classes, tiles, stream);
    if (emax <= 32) return launch_ragged_e32(grid, block, lds_bytes, args, classes, tiles, stream);
    return launch_ragged_e48(grid, block, lds_bytes, args, classes, tiles, stream);
}

int64_t ragged_in_bytes(const RaggedEntry& e)
{
    switch (e.kind) {
        case RaggedKind::C2C: return 8 * e.n;
        case RaggedKind::R2C: return 4 * e.n;
        default: return 8 * (e.n / 2 + 1);
    }
}

int64_t ragged_out_bytes(const RaggedEntry& e)
{
    switch (e.kind) {
        case RaggedKind::C2C: return 8 * e.n;
        case RaggedKind::R2C: return 8 * (e.n / 2 + 1);
        default: return 4 * e.n;
    }
}

bool ragged_shape(const RaggedEntry& e, ClassShape* s)
{
    const bool inv = e.inverse;
    if (e.kind == RaggedKind::C2C) {
        *s = {kVarRowsC2C, e.n, inv, inv, false, 8, 8};
        return true;
    }
    const bool even = (e.n % 2) == 0 && e.n >= 4 && !choose_radices(e.n / 2).empty();
    if (e.kind == RaggedKind::R2C) {
        if (even) {
            *s = {kVarR2CEven, e.n / 2, 0, inv, true, 8, 8};
        } else {
            *s = {kVarR2COdd, e.n, 0, inv, false, 4, 8};
        }
    } else {
        if (even) {
            *s = {kVarC2REven, e.n / 2, inv ? 0 : 1, 1, true, 8, 8};
        } else {
            *s = {kVarC2ROdd, e.n, inv, 0, false, 8, 4};
        }
    }
    return true;
}

std::string ragged_class_key(const RaggedEntry& e)
{
    std::ostringstream os;
    os << kind_name(e.kind) << ":" << (e.inverse ? "i" : "f") << ":" << e.n;
    return os.str();
}

std::string ragged_schedule(RaggedClassInfo* cp, const ClassSchedule* fixed, bool* unsupported, int max_t16)
{
    RaggedClassInfo& c = *cp;
    ClassSchedule sch;
    if (fixed != nullptr) {
        sch = *fixed;
    } else {
        sch.radices = choose_radices(c.n);
    }
    int64_t prod = 1;
    for (int r : sch.radices) prod *= r;
    if (prod != c.n || sch.radices.empty() || (int)sch.radices.size() > kMaxPasses) {
        return "schedule for " + c.key + " does not factor its length";
    }
    auto limit = [&](int emax) { return emax <= 16 ? std::max(max_t16, ragged_block_for(16)) : ragged_block_for(emax); };
    auto fit = [&](const std::vector<int>& r, int* emax_out, int* t_out) {
        for (int emax : {16, 32, 48}) {
            const int t = min_threads(c.n, r, emax);
            if (t > 0 && t <= limit(emax)) {
                *emax_out = emax;
                *t_out = t;
                return true;
            }
        }
        return false;
    };
    c.emax = 0;
    int emax = 0, t = 0;
    if (!fit(sch.radices, &emax, &t) && fixed == nullptr) {
        // the fewest passes need too many threads: take the first multiset that fits
        for (const std::vector<int>& alt : radix_multisets(c.n, 3, 64)) {
            if (fit(alt, &emax, &t)) {
                sch.radices = alt;
                break;
            }
        }
    }
    if (fixed == nullptr && max_t16 > ragged_block_for(16) && emax > 16) {
        // wide EMAX-16 workgroups: a multiset with up to three more passes that fits 16 points
        // per thread in at most max_t16 threads beats a larger register budget
        for (const std::vector<int>& alt : radix_multisets(c.n, 3, 64)) {
            const int t16 = min_threads(c.n, alt, 16);
            if (t16 > 0 && t16 <= max_t16) {
                sch.radices = alt;
                emax = 16;
                t = t16;
                break;
            }
        }
    }
    if (emax > 0) {
        c.emax = emax;
        sch.threads = std::max(sch.threads, t);
    }
    if (c.emax == 0 || sch.threads > ragged_max_block(c.emax)) {
        *unsupported = true;
        return "no register budget holds " + c.key;
    }
    c.sched = sch;
    return "";
}

size_t ragged_class_lds(const RaggedClassInfo& c, int tpb) { return class_lds(c, tpb); }

bool ragged_device_class(const RaggedClassInfo& c, const ClassShape& s, int tpb, RaggedClass* rc,
                         std::vector<TablePtr>* tables)
{
    *rc = RaggedClass{};
    FftArgs& a = rc->a;
    std::vector<int> offs;
    TablePtr tw = upload_table(table_key(c.n, c.sched.radices), stockham_table((int)c.n, c.sched.radices, &offs));
    if (!tw) return false;
    tables->push_back(tw);
    a.tw = tw->dev;
    if (s.aux) {
        TablePtr aux = upload_table("aux:" + std::to_string(c.nreal), roots_table(c.nreal, c.nreal / 2 + 1));
        if (!aux) return false;
        tables->push_back(aux);
        a.tw_aux = aux->dev;
    }
    a.n = (int)c.n;
    a.threads = c.sched.threads;
    a.tpb = tpb;
    a.lds_tstride = (int)(pad16(c.n) + 1);
    a.lds_plane = tpb * a.lds_tstride;
    a.conj_in = s.conj_in;
    a.conj_out = s.conj_out;
    a.count = c.transforms;
    a.cols = 1;
    a.in_stride = 1;
    a.out_stride = 1;
    a.passes = (int)c.sched.radices.size();
    int span = 1;
    for (size_t p = 0; p < c.sched.radices.size(); p++) {
        a.pass[p].radix = c.sched.radices[p];
        a.pass[p].span = span;
        a.pass[p].magic = magic_for(span);
        a.pass[p].tw_offset = offs[p];
        span *= c.sched.radices[p];
    }
    rc->variant = c.variant;
    return true;
}

RaggedPlan::~RaggedPlan()
{
    int prev = -1;
    const bool have = device >= 0 && hipGetDevice(&prev) == hipSuccess;
    if (have && prev != device) (void)hipSetDevice(device);
    if (graph_exec != nullptr) (void)hipGraphExecDestroy(graph_exec);
    if (graph != nullptr) (void)hipGraphDestroy(graph);
    for (hipStream_t st : aux) (void)hipStreamDestroy(st);
    for (hipEvent_t ev : joins) (void)hipEventDestroy(ev);
    if (fork != nullptr) (void)hipEventDestroy(fork);
    for (RaggedLaunchPlan& l : launches) {
        free_device(l.d_classes);
        free_device(l.d_tiles);
    }
    free_device(d_offsets);
    if (have && prev != device) (void)hipSetDevice(prev);
}

namespace {

std::string build_once(const std::vector<RaggedEntry>& entries, const RaggedOptions& opt, RaggedPlan* plan,
                       bool* unsupported)
{
    *unsupported = false;
    plan->entries = entries;
    plan->opt = opt;
    if (entries.empty()) return "a ragged batch needs at least one entry";

    // ---- validation, extents, classes
    std::map<std::string, int> class_of;
    std::vector<std::vector<int>> class_entries;
    std::vector<ClassShape> shapes;
    std::vector<std::pair<int64_t, int64_t>> out_ranges;
    bool inplace = true;
    for (size_t i = 0; i < entries.size(); i++) {
        const RaggedEntry& e = entries[i];
        const std::string where = "entry " + std::to_string(i) + ": ";
        if (e.n < 1 || e.count < 1) return where + "length and count must be positive";
        if (e.in_offset < 0 || e.out_offset < 0) return where + "offsets must be non-negative";
        ClassShape s;
        ragged_shape(e, &s);
        if (e.in_offset % s.in_align != 0 || e.out_offset % s.out_align != 0) {
            return where + "offsets must be multiples of the element size (" + std::to_string(s.in_align) + " / " +
                   std::to_string(s.out_align) + " bytes)";
        }
        if (s.n < 2 || choose_radices(s.n).empty() || s.n > kMaxComplexLength) {
            *unsupported = true;
            return where + "length " + std::to_string(e.n) +
                   " is outside the ragged kernels (radix-{2..19}-smooth, at most 16384 complex points per pass)";
        }
        const int64_t ib = ragged_in_bytes(e), ob = ragged_out_bytes(e);
        plan->in_extent = std::max<size_t>(plan->in_extent, (size_t)(e.in_offset + e.count * ib));
        plan->out_extent = std::max<size_t>(plan->out_extent, (size_t)(e.out_offset + e.count * ob));
        out_ranges.push_back({e.out_offset, e.out_offset + e.count * ob});
        inplace = inplace && e.kind == RaggedKind::C2C && e.in_offset == e.out_offset;
        const std::string key = ragged_class_key(e);
        auto it = class_of.find(key);
        if (it == class_of.end()) {
            class_of[key] = (int)plan->classes.size();
            RaggedClassInfo c;
            c.key = key;
            c.kind = e.kind;
            c.inverse = e.inverse;
            c.variant = s.variant;
            c.n = s.n;
            c.nreal = e.n;
            plan->classes.push_back(c);
            class_entries.push_back({});
            shapes.push_back(s);
            it = class_of.find(key);
        }
        RaggedClassInfo& c = plan->classes[it->second];
        c.transforms += e.count;
        c.bytes += (double)(ib + ob) * (double)e.count;
        class_entries[it->second].push_back((int)i);
        plan->bytes += (double)(ib + ob) * (double)e.count;
        plan->flops += 5.0 * (double)e.n * log2((double)e.n) * (double)e.count;
    }
    std::sort(out_ranges.begin(), out_ranges.end());
    for (size_t i = 1; i < out_ranges.size(); i++) {
        if (out_ranges[i].first < out_ranges[i - 1].second) return "output ranges of two entries overlap";
    }
    plan->all_inplace_c2c = inplace;

    // ---- per-class schedules
    for (size_t ci = 0; ci < plan->classes.size(); ci++) {
        RaggedClassInfo& c = plan->classes[ci];
        auto f = opt.fixed.find(c.key);
        const std::string why = ragged_schedule(&c, f != opt.fixed.end() ? &f->second : nullptr, unsupported,
                                                opt.wide16 ? kRaggedMaxBlock16 : ragged_block_for(16));
        if (!why.empty()) return why;
    }

    // one launch for the whole batch: every class takes the largest register budget (a
    // class keeps its thread count; the extra butterfly slots of a larger budget stay idle).
    // With wide EMAX-16 schedules (up to 1024 threads per transform, more than the larger
    // budgets' 512-thread workgroups hold) the EMAX-16 classes keep a launch of their own and
    // only the others share the largest budget: two launches at most.
    if (opt.one_launch) {
        int top = 0, top_rest = 0;
        for (const RaggedClassInfo& c : plan->classes) {
            top = std::max(top, c.emax);
            if (c.emax > 16) top_rest = std::max(top_rest, c.emax);
        }
        for (RaggedClassInfo& c : plan->classes) {
            if (!opt.wide16) {
                c.emax = top;
            } else if (c.emax > 16) {
                c.emax = top_rest;
            }
        }
    }

    // ---- launches: one per register budget
    for (int emax : {16, 32, 48}) {
        RaggedLaunchPlan L;
        L.emax = emax;
        int maxT = 0;
        size_t one = 0;
        for (size_t ci = 0; ci < plan->classes.size(); ci++) {
            if (plan->classes[ci].emax != emax) continue;
            L.class_ids.push_back((int)ci);
            maxT = std::max(maxT, plan->classes[ci].sched.threads);
            one = std::max(one, class_lds(plan->classes[ci], 1));
        }
        if (L.class_ids.empty()) continue;
        const int want = (emax == 16) ? opt.block16 : (emax == 32 ? opt.block32 : 0);
        int block = want > 0 ? want : ragged_block_for(emax);
        block = std::max(block, (maxT + 63) / 64 * 64);
        block = std::min(block, ragged_max_block(emax));
        L.block = block;
        // whole workgroups of transforms: tiles fill every lane that LDS allows
        const size_t budget = std::min(kLdsLimit, std::max(kLdsFloor, one * (size_t)std::max(1, block / maxT)));
        for (int ci : L.class_ids) {
            RaggedClassInfo& c = plan->classes[ci];
            if (c.sched.threads > block) {  // every transform of a tile needs its threads in the workgroup
                *unsupported = true;
                return "schedule of " + c.key + " needs " + std::to_string(c.sched.threads) +
                       " threads, more than its launch's workgroup of " + std::to_string(block);
            }
            int tpb = block / c.sched.threads;
            if (c.sched.tpb > 0) tpb = std::min(tpb, c.sched.tpb);
            tpb = (int)std::max<int64_t>(1, std::min<int64_t>(tpb, c.transforms));
            while (tpb > 1 && class_lds(c, tpb) > budget) tpb--;
            c.sched.tpb = tpb;
            c.lds = class_lds(c, tpb);
            L.lds_bytes = std::max(L.lds_bytes, c.lds);
        }
        plan->launches.push_back(L);
    }

    // ---- offset table and tiles
    for (RaggedLaunchPlan& L : plan->launches) {
        struct Costed {
            double cost;
            RaggedTile t;
        };
        std::vector<Costed> tiles;
        for (size_t k = 0; k < L.class_ids.size(); k++) {
            const int ci = L.class_ids[k];
            const RaggedClassInfo& c = plan->classes[ci];
            const int64_t first = (int64_t)plan->offsets.size() / 2;
            for (int ei : class_entries[ci]) {
                const RaggedEntry& e = entries[ei];
                const int64_t ib = ragged_in_bytes(e), ob = ragged_out_bytes(e);
                for (int64_t t = 0; t < e.count; t++) {
                    plan->offsets.push_back(e.in_offset + t * ib);
                    plan->offsets.push_back(e.out_offset + t * ob);
                }
            }
            const double per = (double)c.n * (double)(c.sched.radices.size() + 1);
            for (int64_t f = 0; f < c.transforms; f += c.sched.tpb) {
                RaggedTile t;
                t.cls = (int)k;  // index into this launch's class table
                t.count = (int)std::min<int64_t>(c.sched.tpb, c.transforms - f);
                t.first = first + f;
                tiles.push_back({per * t.count, t});
            }
        }
        std::stable_sort(tiles.begin(), tiles.end(), [](const Costed& a, const Costed& b) { return a.cost > b.cost; });
        for (const Costed& t : tiles) L.tiles.push_back(t.t);
        L.ntiles = (int64_t)L.tiles.size();
    }

    // ---- description
    std::ostringstream os;
    os << "ragged plan: " << entries.size() << " entries, " << plan->classes.size() << " classes, "
       << plan->launches.size() << " launch(es), " << (plan->opt.graph ? "graph" : "stream") << " exec\n";
    for (const RaggedLaunchPlan& L : plan->launches) {
        os << "  launch emax=" << L.emax << " block=" << L.block << " lds=" << L.lds_bytes << " tiles=" << L.ntiles
           << "\n";
        for (int ci : L.class_ids) {
            const RaggedClassInfo& c = plan->classes[ci];
            os << "    " << c.key << " variant=" << c.variant << " n=" << c.n << " radices=";
            for (size_t i = 0; i < c.sched.radices.size(); i++) os << (i ? "x" : "") << c.sched.radices[i];
            os << " T=" << c.sched.threads << " tpb=" << c.sched.tpb << " transforms=" << c.transforms << "\n";
        }
    }
    plan->describe = os.str();

    // ---- device tables (skipped without a device: the plan still validates)
    if (!device_available()) return "";
    if (hipGetDevice(&plan->device) != hipSuccess) return "";
    for (RaggedLaunchPlan& L : plan->launches) {
        for (int ci : L.class_ids) {
            const RaggedClassInfo& c = plan->classes[ci];
            const ClassShape& s = shapes[ci];
            RaggedClass rc{};
            if (!ragged_device_class(c, s, c.sched.tpb, &rc, &plan->tables)) return "";
            L.dev_classes.push_back(rc);
        }
        if (hipMalloc(reinterpret_cast<void**>(&L.d_classes), L.dev_classes.size() * sizeof(RaggedClass)) !=
                hipSuccess ||
            hipMemcpy(L.d_classes, L.dev_classes.data(), L.dev_classes.size() * sizeof(RaggedClass),
                      hipMemcpyHostToDevice) != hipSuccess) {
            return "device allocation failed";
        }
        if (hipMalloc(reinterpret_cast<void**>(&L.d_tiles), L.tiles.size() * sizeof(RaggedTile)) != hipSuccess ||
            hipMemcpy(L.d_tiles, L.tiles.data(), L.tiles.size() * sizeof(RaggedTile), hipMemcpyHostToDevice) !=
                hipSuccess) {
            return "device allocation failed";
        }
        int per_cu = 0;
        if (ragged_prepare(L.emax, L.lds_bytes, L.block, &per_cu) != hipSuccess || per_cu < 1) {
            return "the ragged kernel cannot run with " + std::to_string(L.lds_bytes) + " bytes of LDS";
        }
        // one workgroup per tile (opt.max_grid > 0 caps it: the workgroups then loop)
        L.grid = opt.max_grid > 0 ? std::min<int64_t>(L.ntiles, opt.max_grid) : L.ntiles;
        (void)device_cus;
    }
    const size_t obytes = plan->offsets.size() * sizeof(int64_t);
    if (hipMalloc(reinterpret_cast<void**>(&plan->d_offsets), obytes) != hipSuccess ||
        hipMemcpy(plan->d_offsets, plan->offsets.data(), obytes, hipMemcpyHostToDevice) != hipSuccess) {
        return "device allocation failed";
    }
    if (hipDeviceSynchronize() != hipSuccess) return "device error while uploading the plan";
    plan->ready = true;
    return "";
}

RaggedArgs args_for(const RaggedPlan& p, size_t li, const void* in, void* out)
{
    RaggedArgs a{};
    a.in_base = in;
    a.out_base = out;
    a.offsets = p.d_offsets;
    a.ntiles = (int)p.launches[li].ntiles;
    return a;
}

hipError_t node_params(const RaggedPlan& p, size_t li, RaggedArgs* args, void** slots, hipKernelNodeParams* kp)
{
    const RaggedLaunchPlan& L = p.launches[li];
    slots[0] = args;
    slots[1] = const_cast<RaggedClass**>(&L.d_classes);
    slots[2] = const_cast<RaggedTile**>(&L.d_tiles);
    *kp = hipKernelNodeParams{};
    kp->func = const_cast<void*>(ragged_kernel_fn(L.emax));
    kp->gridDim = dim3((unsigned)L.grid);
    kp->blockDim = dim3((unsigned)L.block);
    kp->sharedMemBytes = (unsigned)L.lds_bytes;
    kp->kernelParams = slots;
    kp->extra = nullptr;
    return hipSuccess;
}

// Helper streams and events of the side-by-side launches, created all together or not at
// all: a failure part-way destroys what was made, so the next exec starts from nothing.
hipError_t ensure_helpers(RaggedPlan& p, size_t nl)
{
    if (nl <= 1 || !p.aux.empty()) return hipSuccess;
    std::vector<hipStream_t> aux;
    std::vector<hipEvent_t> joins;
    hipEvent_t fork = nullptr;
    hipError_t e = hipSuccess;
    for (size_t i = 0; i + 1 < nl && e == hipSuccess; i++) {
        hipStream_t st;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) break;
        aux.push_back(st);
        hipEvent_t ev;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) break;
        joins.push_back(ev);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
    if (e != hipSuccess) {
        for (hipStream_t st : aux) (void)hipStreamDestroy(st);
        for (hipEvent_t ev : joins) (void)hipEventDestroy(ev);
        return e;
    }
    p.aux = std::move(aux);
    p.joins = std::move(joins);
    p.fork = fork;
    return hipSuccess;
}

// Builds and instantiates the graph of the launches, or leaves no graph, nodes or exec
// behind when any step fails.
hipError_t build_graph(RaggedPlan& p, const void* in, void* out)
{
    hipGraph_t graph = nullptr;
    hipError_t e = hipGraphCreate(&graph, 0);
    if (e != hipSuccess) return e;
    std::vector<hipGraphNode_t> nodes;
    // the launches touch disjoint outputs: independent nodes
    for (size_t li = 0; li < p.launches.size() && e == hipSuccess; li++) {
        RaggedArgs a = args_for(p, li, in, out);
        void* slots[3];
        hipKernelNodeParams kp;
        e = node_params(p, li, &a, slots, &kp);
        hipGraphNode_t node;
        if (e == hipSuccess) e = hipGraphAddKernelNode(&node, graph, nullptr, 0, &kp);
        if (e == hipSuccess) nodes.push_back(node);
    }
    hipGraphExec_t exec = nullptr;
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return e;
    }
    p.graph = graph;
    p.nodes = std::move(nodes);
    p.graph_exec = exec;
    p.graph_in = in;
    p.graph_out = out;
    return hipSuccess;
}

}  // namespace

hipError_t exec_ragged(RaggedPlan& p, const void* in, void* out, hipStream_t stream)
{
    if (!p.ready) return hipErrorNotReady;
    if (!p.opt.graph) {
        const size_t nl = p.launches.size();
        // the register budgets' launches run side by side on helper streams: a budget
        // with few tiles (long transforms) would otherwise leave most CUs idle
        hipError_t e = ensure_helpers(p, nl);
        if (e != hipSuccess) return e;
        if (nl > 1) {
            e = hipEventRecord(p.fork, stream);
            if (e != hipSuccess) return e;
        }
        // launch order: the budgets with the longest transforms (fewest, slowest tiles) first
        for (size_t k = nl; k-- > 0;) {
            const RaggedLaunchPlan& L = p.launches[k];
            hipStream_t st = (k == 0) ? stream : p.aux[k - 1];
            if (k > 0) {
                e = hipStreamWaitEvent(st, p.fork, 0);
                if (e != hipSuccess) return e;
            }
            const RaggedArgs a = args_for(p, k, in, out);
            e = launch_ragged(L.emax, L.grid, L.block, L.lds_bytes, a, L.d_classes, L.d_tiles, st);
            if (e != hipSuccess) return e;
            if (k > 0) {
                e = hipEventRecord(p.joins[k - 1], st);
                if (e != hipSuccess) return e;
            }
        }
        for (size_t k = 1; k < nl; k++) {
            e = hipStreamWaitEvent(stream, p.joins[k - 1], 0);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    hipError_t e;
    if (p.graph_exec == nullptr) {
        e = build_graph(p, in, out);
        if (e != hipSuccess) return e;
    } else if (in != p.graph_in || out != p.graph_out) {
        for (size_t li = 0; li < p.launches.size(); li++) {
            RaggedArgs a = args_for(p, li, in, out);
            void* slots[3];
            hipKernelNodeParams kp;
            e = node_params(p, li, &a, slots, &kp);
            if (e != hipSuccess) return e;
            e = hipGraphExecKernelNodeSetParams(p.graph_exec, p.nodes[li], &kp);
            if (e != hipSuccess) return e;
        }
        p.graph_in = in;
        p.graph_out = out;
    }
    return hipGraphLaunch(p.graph_exec, stream);
}

std::string format_schedules(const RaggedPlan& plan)
{
    std::ostringstream os;
    for (const RaggedClassInfo& c : plan.classes) {
        os << c.key << " ";
        for (size_t i = 0; i < c.sched.radices.size(); i++) os << (i ? "x" : "") << c.sched.radices[i];
        os << " " << c.sched.threads << " " << c.sched.tpb << "\n";
    }
    return os.str();
}


// ---- autotuner -----------------------------------------------------------------------
namespace {

constexpr double kTuneBytes = 8.0 * 1024 * 1024;  // data per timed single-class launch
constexpr int kTuneReps = 4;
constexpr int kTuneCandidates = 8;

std::mutex g_tune_mu;
std::map<std::string, ClassSchedule>& tune_cache()
{
    static auto* m = new std::map<std::string, ClassSchedule>();
    return *m;
}

struct Scratch {
    void* in = nullptr;
    void* out = nullptr;
    size_t bytes = 0;
    bool ensure(size_t need)
    {
        if (need <= bytes) return true;
        release();
        if (hipMalloc(&in, need) != hipSuccess || hipMalloc(&out, need) != hipSuccess) {
            release();
            return false;
        }
        if (hipMemset(in, 0, need) != hipSuccess) return false;
        bytes = need;
        return true;
    }
    void release()
    {
        free_device(in);
        free_device(out);
        in = out = nullptr;
        bytes = 0;
    }
    ~Scratch() { release(); }
};

// Average event-timed duration of one exec (after one untimed exec), +inf on failure.
double time_exec(RaggedPlan& p, void* in, void* out, hipStream_t s, int reps)
{
    const double inf = 1e30;
    if (exec_ragged(p, in, out, s) != hipSuccess) return inf;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return inf;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return inf;
    }
    double ms = inf;
    bool ok = hipEventRecord(e0, s) == hipSuccess;
    for (int r = 0; ok && r < reps; r++) ok = exec_ragged(p, in, out, s) == hipSuccess;
    ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
    float f = 0;
    if (ok && hipEventElapsedTime(&f, e0, e1) == hipSuccess) ms = (double)f / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

std::string cache_key(int device, const std::string& key, const RaggedOptions& opt)
{
    return std::to_string(device) + "|" + key + "|" + std::to_string(opt.block16) + "|" + std::to_string(opt.block32);
}

// Times every candidate schedule of class c (alone, replicated to about kTuneBytes) and
// returns the fastest.  Falls back to the default schedule when nothing could be timed.
ClassSchedule tune_class(const RaggedClassInfo& c, const RaggedOptions& base, Scratch& scratch, hipStream_t s,
                         int* timed)
{
    RaggedEntry e;
    e.kind = c.kind;
    e.inverse = c.inverse;
    e.n = c.nreal;
    const int64_t ib = ragged_in_bytes(e), ob = ragged_out_bytes(e);
    const int64_t want = (int64_t)(kTuneBytes / (double)(ib + ob)) + 1;
    e.count = std::max<int64_t>(c.transforms, std::min<int64_t>(want, 1 << 16));
    const size_t need = (size_t)std::max(ib, ob) * (size_t)e.count;
    ClassSchedule best = c.sched;
    best.tpb = 0;
    if (!scratch.ensure(need)) return best;
    double best_ms = 1e30;
    for (const ClassSchedule& cand : class_candidates(c.n, kTuneCandidates)) {
        RaggedOptions o = base;
        o.graph = false;
        o.autotune = false;
        o.fixed.clear();
        o.fixed[c.key] = cand;
        RaggedPlan p;
        bool unsup = false;
        if (!build_once({e}, o, &p, &unsup).empty() || !p.ready) continue;
        const double ms = time_exec(p, scratch.in, scratch.out, s, kTuneReps);
        (*timed)++;
        if (ms < best_ms) {
            best_ms = ms;
            best = cand;
        }
    }
    return best;
}

}  // namespace

std::vector<ClassSchedule> class_candidates(int64_t n, int limit)
{
    std::vector<ClassSchedule> out;
    auto add = [&](const std::vector<int>& r, int threads) {
        for (const ClassSchedule& c : out) {
            if (c.radices == r && c.threads == threads) return;
        }
        ClassSchedule c;
        c.radices = r;
        c.threads = threads;
        out.push_back(c);
    };
    for (const std::vector<int>& ms : radix_multisets(n, 1, 4)) {
        std::vector<int> asc(ms.rbegin(), ms.rend());
        for (const std::vector<int>& order : {ms, asc}) {
            int t0 = 0;
            for (int emax : {16, 32, 48}) {
                const int t = min_threads(n, order, emax);
                if (t > 0 && t <= kRaggedMaxBlock) {
                    t0 = t;
                    break;
                }
            }
            if (t0 == 0) continue;
            add(order, t0);
            if (2 * t0 <= kRaggedMaxBlock && 2 * t0 <= n) add(order, 2 * t0);
        }
    }
    // the default schedule first, then the rest in enumeration order
    if ((int)out.size() > limit) out.resize(limit);
    return out;
}

void clear_tuning_cache()
{
    std::lock_guard<std::mutex> lock(g_tune_mu);
    tune_cache().clear();
}

std::string build_ragged(const std::vector<RaggedEntry>& entries, const RaggedOptions& opt, RaggedPlan* plan,
                         bool* unsupported)
{
    if (!opt.autotune || !device_available()) return build_once(entries, opt, plan, unsupported);
    const auto t0 = std::chrono::steady_clock::now();
    // validate and learn the classes with the default schedules first
    RaggedOptions plain = opt;
    plain.autotune = false;
    plain.graph = false;
    RaggedPlan probe;
    std::string why = build_once(entries, plain, &probe, unsupported);
    if (!why.empty() || !probe.ready) return why.empty() ? build_once(entries, opt, plan, unsupported) : why;
    int device = 0;
    (void)hipGetDevice(&device);
    hipStream_t s = nullptr;
    if (hipStreamCreate(&s) != hipSuccess) return build_once(entries, plain, plan, unsupported);
    int timed = 0;
    RaggedOptions tuned = plain;
    {
        Scratch scratch;
        std::lock_guard<std::mutex> lock(g_tune_mu);
        for (const RaggedClassInfo& c : probe.classes) {
            if (opt.fixed.count(c.key)) continue;
            const std::string ck = cache_key(device, c.key, plain);
            auto it = tune_cache().find(ck);
            if (it == tune_cache().end()) it = tune_cache().emplace(ck, tune_class(c, plain, scratch, s, &timed)).first;
            tuned.fixed[c.key] = it->second;
        }
        // workgroup size of each launch, timed on the whole batch with the tuned schedules
        if (opt.block16 == 0 || opt.block32 == 0) {
            size_t need = std::max(probe.in_extent, probe.out_extent);
            if (scratch.ensure(need)) {
                double best = 1e30;
                int b16 = 512, b32 = 512;
                for (int cand : {256, 512}) {
                    RaggedOptions o = tuned;
                    if (opt.block16 == 0) o.block16 = cand;
                    if (opt.block32 == 0) o.block32 = 512;
                    RaggedPlan p;
                    bool unsup = false;
                    if (!build_once(entries, o, &p, &unsup).empty() || !p.ready) continue;
                    const double ms = time_exec(p, scratch.in, scratch.out, s, kTuneReps);
                    timed++;
                    if (ms < best) {
                        best = ms;
                        b16 = o.block16;
                        b32 = o.block32;
                    }
                }
                if (opt.block16 == 0) tuned.block16 = b16;
                if (opt.block32 == 0) tuned.block32 = b32;
            }
        }
    }
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    tuned.graph = opt.graph;
    tuned.autotune = false;
    why = build_once(entries, tuned, plan, unsupported);
    plan->opt.autotune = true;
    plan->tune_timed = timed;
    plan->tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::ostringstream os;
    os << plan->describe << "  autotuned: " << timed << " timings in " << (int)plan->tune_ms << " ms, block16="
       << tuned.block16 << "\n";
    plan->describe = os.str();
    return why;
}

}  // namespace asdfft
