// Serving loop (see server.h).
#include "server.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "planner.h"
#include "ragged.h"

namespace asdfft {

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kPoolClasses = 4096;           // resident ragged class descriptors
constexpr int64_t kSlotTransforms = 1 << 16;  // ragged transforms per batch
constexpr int kSlotTiles = 1 << 16;
// 1D requests above this many bytes run as their own launch group (a batch's ragged
// launch completes as a whole, so one large request would hold its small neighbours)
constexpr size_t kRaggedMaxRequestBytes = (size_t)8 << 20;
constexpr size_t kStageBytes = sizeof(int64_t) * 2 * kSlotTransforms + sizeof(RaggedTile) * kSlotTiles;
constexpr int kGroups = 3;                   // register budgets: EMAX 16, 32, 48
constexpr size_t kServeLdsCap = 80 * 1024;   // per-class tile LDS (two workgroups per CU)
constexpr int kTypeC2C = 0x10, kTypeC2R = 0x11, kTypeR2C = 0x12;

struct Req {
    uint64_t ticket = 0;
    ServeRequest r;
    bool ragged = false;
    RaggedEntry entry;      // ragged: one entry of `batch` transforms
    std::string plain_key;  // plain: plan cache key
    size_t bytes = 0;
    Clock::time_point submitted, admitted, launched, completed;
    int status = -1;        // -1 in flight, 0 done, 3 device error
};

struct PoolClass {
    RaggedClassInfo info;
    ClassShape shape{};
    int tpb = 1;
    size_t lds = 0;
};

struct PlainPlan {
    BuiltPlan built;
};

struct Slot {
    hipStream_t stream = nullptr;
    // completion of a batch's launch groups, in launch order: the ragged launches, then
    // each plain request; every group's requests complete as soon as its event fires
    std::vector<hipEvent_t> events;
    std::vector<std::vector<uint64_t>> groups;
    std::vector<size_t> group_bytes;
    std::vector<uint64_t> failed;  // tickets of a batch whose launches could not be enqueued
    // a batch's work table: the offset table, then the tiles, staged in pinned memory and
    // moved to the device with one copy
    char* d_stage = nullptr;
    char* h_stage = nullptr;
    // plain plans of a batch run one after another on the slot's stream: one workspace,
    // grown to the largest workspace of any plan submitted so far.  A workspace that is
    // outgrown is kept (and counted as resident) until the server closes: hipFree waits for
    // the whole device, which stalled every stream for about a millisecond per growth
    void* workspace = nullptr;
    size_t workspace_bytes = 0;
    std::vector<void*> retired;
    size_t bytes = 0;
    bool busy = false;
    bool launched = false;
    int error = 0;
    std::thread waiter;
};

int64_t kind_in_bytes(int type, int64_t inner_points, int64_t last)
{
    // inner_points: points of the leading FFT axes; last: the contiguous axis length
    if (type == kTypeC2C) return 8 * inner_points * last;
    if (type == kTypeR2C) return 4 * inner_points * last;
    return 8 * inner_points * (last / 2 + 1);
}

int64_t kind_out_bytes(int type, int64_t inner_points, int64_t last)
{
    if (type == kTypeC2C) return 8 * inner_points * last;
    if (type == kTypeR2C) return 8 * inner_points * (last / 2 + 1);
    return 4 * inner_points * last;
}

Kind plan_kind(int type)
{
    if (type == kTypeR2C) return Kind::R2C;
    if (type == kTypeC2R) return Kind::C2R;
    return Kind::C2C;
}

RaggedKind ragged_kind(int type)
{
    if (type == kTypeR2C) return RaggedKind::R2C;
    if (type == kTypeC2R) return RaggedKind::C2R;
    return RaggedKind::C2C;
}

}  // namespace

class Server {
public:
    ServeConfig cfg;
    int device = 0;
    int cus = 256;
    std::mutex mu;
    std::condition_variable cv;  // scheduler, waiters and serve_wait callers
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Req> reqs;
    uint64_t next_ticket = 1;
    size_t inflight = 0;
    size_t resident = 0;
    bool stop = false;
    ServeStats stats;
    std::vector<std::unique_ptr<Slot>> slots;
    std::thread scheduler;
    hipStream_t setup = nullptr;  // uploads of new resident state
    // resident ragged classes, created at submission (pool_mu) and read by the scheduler
    std::mutex pool_mu;
    RaggedClass* d_pool = nullptr;
    std::vector<PoolClass> pool;
    std::map<std::string, int> pool_index;
    std::unordered_map<uint64_t, int> pool_fast;  // (kind, direction, n) -> pool index, no key string
    std::vector<TablePtr> tables;
    std::mutex plan_mu;  // plain plans are created by submitters and read by the scheduler
    std::map<std::string, std::unique_ptr<PlainPlan>> plans;
    std::atomic<size_t> workspace_max{0};  // largest workspace of any plan built so far
    // small requests in a batch from which they share one ragged launch (a ragged launch pays a
    // work-table copy)
    int min_ragged = 2;
    // pending requests from which batches form (0: one more than the slots; measured at 4, 8
    // and 16 with three slots, the light-load p99s stayed within the run-to-run spread and 16
    // doubled the 50k req/s p50)
    int min_queue = 0;

    ~Server();
    bool init(std::string* why);
    void scheduler_loop();
    void waiter_loop(Slot* s);
    int launch_batch(Slot& s, int slot_index, const std::vector<Req*>& batch);
    int pool_class(const RaggedEntry& e);
    PlainPlan* plain_plan(const Req& r, std::string* why);
};

bool Server::init(std::string* why)
{
    if (!device_available() || hipGetDevice(&device) != hipSuccess) {
        *why = "no HIP device";
        return false;
    }
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && n > 0) cus = n;
    // ASDFFT_SERVE_MIN_RAGGED=<k>: batches share a ragged launch from k small requests (measurement)
    if (const char* e = getenv("ASDFFT_SERVE_MIN_RAGGED")) min_ragged = std::max(2, atoi(e));
    if (const char* e = getenv("ASDFFT_SERVE_MIN_QUEUE")) min_queue = std::max(0, atoi(e));
    if (hipStreamCreateWithFlags(&setup, hipStreamNonBlocking) != hipSuccess) {
        *why = "stream creation failed";
        return false;
    }
    if (hipMalloc(reinterpret_cast<void**>(&d_pool), sizeof(RaggedClass) * kPoolClasses) != hipSuccess) {
        *why = "device allocation failed";
        return false;
    }
    resident += sizeof(RaggedClass) * kPoolClasses;
    for (int emax : {16, 32, 48}) {
        int pc = 0;
        if (ragged_prepare(emax, kRaggedLdsLimit, ragged_block_for(emax), &pc) != hipSuccess) {
            *why = "ragged kernel setup failed";
            return false;
        }
    }
    for (int i = 0; i < cfg.slots; i++) {
        auto s = std::make_unique<Slot>();
        // non-blocking streams: uploads of new classes on other streams never wait for batches
        bool ok = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) == hipSuccess &&

                  hipMalloc(reinterpret_cast<void**>(&s->d_stage), kStageBytes) == hipSuccess &&
                  hipHostMalloc(reinterpret_cast<void**>(&s->h_stage), kStageBytes, 0) == hipSuccess;
        resident += kStageBytes;
        s->events.assign((size_t)cfg.max_batch + 1, nullptr);
        for (auto& ev : s->events) ok = ok && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
        slots.push_back(std::move(s));
        if (!ok) {
            *why = "slot allocation failed";
            return false;
        }
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        *why = "device error during setup";
        return false;
    }
    if (resident >= cfg.hbm_budget) {
        *why = "HBM budget smaller than the server's resident state";
        return false;
    }
    for (auto& s : slots) {
        Slot* sp = s.get();
        sp->waiter = std::thread([this, sp] { waiter_loop(sp); });
    }
    scheduler = std::thread([this] { scheduler_loop(); });
    return true;
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> lock(mu);
        stop = true;
    }
    cv.notify_all();
    if (scheduler.joinable()) scheduler.join();
    for (auto& s : slots) {
        if (s->waiter.joinable()) s->waiter.join();
    }
    int prev = -1;
    const bool have = hipGetDevice(&prev) == hipSuccess;
    if (have && prev != device) (void)hipSetDevice(device);
    (void)hipDeviceSynchronize();
    for (auto& s : slots) {
        if (s->stream) (void)hipStreamDestroy(s->stream);
        for (hipEvent_t ev : s->events) {
            if (ev) (void)hipEventDestroy(ev);
        }
        if (s->d_stage) (void)hipFree(s->d_stage);
        if (s->h_stage) (void)hipHostFree(s->h_stage);
        if (s->workspace) (void)hipFree(s->workspace);
        for (void* w : s->retired) (void)hipFree(w);
    }
    if (d_pool) (void)hipFree(d_pool);
    if (setup) (void)hipStreamDestroy(setup);
    if (have && prev != device) (void)hipSetDevice(prev);
}

// Index of the resident class for this entry (created and uploaded on first use), or -1.
int Server::pool_class(const RaggedEntry& e)
{
    const uint64_t fast = ((uint64_t)e.n << 3) | ((uint64_t)e.inverse << 2) | (uint64_t)e.kind;
    std::lock_guard<std::mutex> lock(pool_mu);
    auto f = pool_fast.find(fast);
    if (f != pool_fast.end()) return f->second;
    const std::string key = ragged_class_key(e);
    auto it = pool_index.find(key);
    if (it != pool_index.end()) return it->second;
    if ((int)pool.size() >= kPoolClasses) return -1;
    PoolClass pc;
    ragged_shape(e, &pc.shape);
    pc.info.key = key;
    pc.info.kind = e.kind;
    pc.info.inverse = e.inverse;
    pc.info.variant = pc.shape.variant;
    pc.info.n = pc.shape.n;
    pc.info.nreal = e.n;
    pc.info.transforms = 1;
    bool unsup = false;
    if (!ragged_schedule(&pc.info, nullptr, &unsup).empty()) return -1;
    const size_t one = ragged_class_lds(pc.info, 1);
    const size_t cap = std::max(kServeLdsCap, one);
    int tpb = std::max(1, ragged_block_for(pc.info.emax) / pc.info.sched.threads);
    while (tpb > 1 && ragged_class_lds(pc.info, tpb) > cap) tpb--;
    pc.tpb = tpb;
    pc.lds = ragged_class_lds(pc.info, tpb);
    RaggedClass rc;
    if (!ragged_device_class(pc.info, pc.shape, tpb, &rc, &tables)) return -1;
    const int idx = (int)pool.size();
    if (hipMemcpyAsync(d_pool + idx, &rc, sizeof(rc), hipMemcpyHostToDevice, setup) != hipSuccess ||
        hipStreamSynchronize(setup) != hipSuccess) {
        return -1;
    }
    pool.push_back(pc);
    pool_index[key] = idx;
    pool_fast[fast] = idx;
    return idx;
}

PlainPlan* Server::plain_plan(const Req& q, std::string* why)
{
    std::lock_guard<std::mutex> lock(plan_mu);
    auto it = plans.find(q.plain_key);
    if (it != plans.end()) return it->second.get();
    PlanSpec spec;
    spec.kind = plan_kind(q.r.type);
    spec.inverse = q.r.inverse;
    spec.dims.assign(q.r.dims, q.r.dims + q.r.rank);
    spec.batch = q.r.batch;
    auto p = std::make_unique<PlainPlan>();
    *why = build_plan(spec, &p->built);
    if (!why->empty()) return nullptr;
    for (const Step& st : p->built.steps) {
        if (!st.tables_ready) {
            *why = "plan tables could not be uploaded";
            return nullptr;
        }
    }
    PlainPlan* raw = p.get();
    size_t seen = workspace_max.load();
    while (raw->built.workspace > seen && !workspace_max.compare_exchange_weak(seen, raw->built.workspace)) {
    }
    plans[q.plain_key] = std::move(p);
    return raw;
}

int Server::launch_batch(Slot& s, int slot_index, const std::vector<Req*>& batch)
{
    // ---- ragged part: transforms grouped per resident class
    std::map<int, std::vector<int64_t>> per_class;  // pool index -> (in, out) address pairs
    std::vector<Req*> plain;
    std::vector<uint64_t> ragged_tickets;
    size_t ragged_bytes = 0;
    s.groups.clear();
    s.group_bytes.clear();
    // the ragged launch pays a work-table copy: it runs when at least min_ragged small 1D
    // requests share the batch; fewer run back to back through their plain plans
    int eligible = 0;
    for (Req* q : batch) eligible += (q->ragged && q->bytes <= kRaggedMaxRequestBytes) ? 1 : 0;
    for (Req* q : batch) {
        const bool take = eligible >= min_ragged && q->ragged && q->bytes <= kRaggedMaxRequestBytes;
        int ci = take ? pool_class(q->entry) : -1;
        if (ci < 0) {
            plain.push_back(q);
            continue;
        }
        ragged_tickets.push_back(q->ticket);
        ragged_bytes += q->bytes;
        const int64_t ib = ragged_in_bytes(q->entry), ob = ragged_out_bytes(q->entry);
        auto& v = per_class[ci];
        const int64_t in0 = (int64_t)reinterpret_cast<uintptr_t>(q->r.in);
        const int64_t out0 = (int64_t)reinterpret_cast<uintptr_t>(q->r.out);
        for (int64_t t = 0; t < q->entry.count; t++) {
            v.push_back(in0 + t * ib);
            v.push_back(out0 + t * ob);
        }
    }
    int64_t nt = 0;
    struct Costed {
        double cost;
        RaggedTile t;
        int group;
    };
    auto group_of = [](int emax) { return emax <= 16 ? 0 : (emax <= 32 ? 1 : 2); };
    std::vector<Costed> tiles;
    size_t lds[kGroups] = {0, 0, 0};
    std::vector<PoolClass> used;  // copies: the pool vector may grow under other submitters
    {
        std::lock_guard<std::mutex> lock(pool_mu);
        for (auto& kv : per_class) used.push_back(pool[kv.first]);
    }
    size_t ui = 0;
    for (auto& kv : per_class) {
        const PoolClass& pc = used[ui++];
        const int64_t count = (int64_t)kv.second.size() / 2;
        memcpy(reinterpret_cast<int64_t*>(s.h_stage) + 2 * nt, kv.second.data(), kv.second.size() * sizeof(int64_t));
        const double per = (double)pc.info.n * (double)(pc.info.sched.radices.size() + 1);
        const int g = group_of(pc.info.emax);
        for (int64_t f = 0; f < count; f += pc.tpb) {
            RaggedTile t;
            t.cls = kv.first;
            t.count = (int)std::min<int64_t>(pc.tpb, count - f);
            t.first = nt + f;
            tiles.push_back({per * t.count, t, g});
        }
        lds[g] = std::max(lds[g], pc.lds);
        nt += count;
    }
    std::stable_sort(tiles.begin(), tiles.end(), [](const Costed& a, const Costed& b) {
        if (a.group != b.group) return a.group < b.group;
        return a.cost > b.cost;
    });
    int64_t counts[kGroups] = {0, 0, 0};
    const size_t tiles_at = (size_t)nt * 2 * sizeof(int64_t);  // 16-byte aligned
    RaggedTile* h_tiles = reinterpret_cast<RaggedTile*>(s.h_stage + tiles_at);
    for (size_t i = 0; i < tiles.size(); i++) {
        h_tiles[i] = tiles[i].t;
        counts[tiles[i].group]++;
    }
    const int64_t* d_offsets = reinterpret_cast<const int64_t*>(s.d_stage);
    const RaggedTile* d_tiles = reinterpret_cast<const RaggedTile*>(s.d_stage + tiles_at);
    if (!tiles.empty()) {
        // one copy of the whole work table from pinned memory (a kernel reading the table
        // from host memory instead measured slower: 50k req/s p50 2.7 ms against 0.14 ms)
        if (hipMemcpyAsync(s.d_stage, s.h_stage, tiles_at + sizeof(RaggedTile) * tiles.size(), hipMemcpyHostToDevice,
                           s.stream) != hipSuccess) {
            return 3;
        }
        int64_t first = 0;
        for (int g = 0; g < kGroups; g++) {
            const int64_t count = counts[g];
            if (count == 0) continue;
            const int emax = 16 * (g + 1);
            RaggedArgs a{};
            a.in_base = nullptr;  // offsets hold absolute device addresses
            a.out_base = nullptr;
            a.offsets = d_offsets;
            a.ntiles = (int)count;
            const int64_t grid = count;  // one workgroup per tile
            if (launch_ragged(emax, grid, ragged_block_for(emax), lds[g], a, d_pool, d_tiles + first, s.stream) !=
                hipSuccess) {
                return 3;
            }
            first += count;
        }
        if (hipEventRecord(s.events[s.groups.size()], s.stream) != hipSuccess) return 3;
        s.groups.push_back(ragged_tickets);
        s.group_bytes.push_back(ragged_bytes);
    }
    // ---- plain part: cached plans, launched on this slot's stream
    std::vector<PlainPlan*> pplans;
    size_t ws_need = 0;
    for (Req* q : plain) {
        std::string why;
        PlainPlan* p = plain_plan(*q, &why);
        if (p == nullptr) return 3;
        pplans.push_back(p);
        ws_need = std::max(ws_need, p->built.workspace);
    }
    if (ws_need > s.workspace_bytes) {
        // grow straight to the largest workspace submitted so far (one growth per slot while
        // the request mix stays the same); the old one stays allocated, see Slot::retired
        const size_t want = std::max(ws_need, workspace_max.load());
        void* fresh = nullptr;
        if (hipMalloc(&fresh, want) != hipSuccess) return 3;
        if (s.workspace) s.retired.push_back(s.workspace);
        s.workspace = fresh;
        s.workspace_bytes = want;
        std::lock_guard<std::mutex> lock(mu);
        resident += want;
    }
    (void)slot_index;
    for (size_t k = 0; k < plain.size(); k++) {
        Req* q = plain[k];
        PlainPlan* p = pplans[k];
        char* bufs[3] = {static_cast<char*>(const_cast<void*>(q->r.in)), static_cast<char*>(q->r.out),
                         static_cast<char*>(s.workspace)};
        for (const Step& st : p->built.steps) {
            if (launch_step(st, bufs, nullptr, nullptr, s.stream) != hipSuccess) return 3;
        }
        if (hipEventRecord(s.events[s.groups.size()], s.stream) != hipSuccess) return 3;
        s.groups.push_back({q->ticket});
        s.group_bytes.push_back(q->bytes);
    }
    return 0;
}

void Server::scheduler_loop()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        int si = -1;
        auto ready = [&] {
            if (stop) return true;
            if (pending.empty()) return false;
            for (size_t i = 0; i < slots.size(); i++) {
                if (!slots[i]->busy) {
                    si = (int)i;
                    return true;
                }
            }
            return false;
        };
        cv.wait(lk, ready);
        if (stop) break;
        // admission: arrival order while the budget and the slot's tables allow.  The pending
        // requests are shared among the idle slots (each takes its share now, the others take
        // theirs as the scheduler comes round), so under light load every request gets a slot
        // and a stream of its own and the batches grow only when requests queue up
        std::vector<Req*> batch;
        std::vector<uint64_t> tickets;
        size_t add = 0;
        int64_t transforms = 0;
        const size_t budget = cfg.hbm_budget - std::min(cfg.hbm_budget, resident);
        size_t idle_slots = 0;
        for (const auto& sl : slots) idle_slots += sl->busy ? 0 : 1;
        const size_t share = (pending.size() + std::max<size_t>(idle_slots, 1) - 1) / std::max<size_t>(idle_slots, 1);
        size_t limit = std::min<size_t>((size_t)cfg.max_batch, std::max<size_t>(share, 1));
        if (cfg.batching && limit > 1) {
            // several requests share a slot only when more requests wait than there are slots
            // (the device is saturated) and they share a ragged launch: otherwise fewer than
            // min_ragged small requests would run back to back on this slot's stream, or a short
            // queue would wait for one ragged launch, while a slot that frees up meanwhile could
            // run the next request beside it
            size_t small_head = 0;
            for (auto it = pending.begin(); it != pending.end() && small_head < limit; ++it) {
                const Req& h = reqs[*it];
                if (!(h.ragged && h.bytes <= kRaggedMaxRequestBytes)) break;
                small_head++;
            }
            const size_t queue = min_queue > 0 ? (size_t)min_queue : slots.size() + 1;
            if (small_head < (size_t)min_ragged || pending.size() < queue) limit = 1;
        }
        while (!pending.empty() && batch.size() < limit) {
            Req& q = reqs[pending.front()];
            // with nothing in flight the head request is admitted even if resident state
            // (grown workspaces) has eaten into the budget since it was submitted
            const bool idle = inflight == 0 && batch.empty();
            if (inflight + add + q.bytes > budget && !idle) break;
            if (!cfg.batching && !batch.empty()) break;
            // only small 1D requests share a batch (one ragged launch); any other request runs
            // through its plain plan in a batch of its own, so large requests never queue behind
            // one another on one slot's stream while other slots free up
            const bool small = q.ragged && q.bytes <= kRaggedMaxRequestBytes;
            if (!small && !batch.empty()) break;
            if (q.ragged && transforms + q.entry.count > kSlotTransforms) break;
            if (q.ragged) transforms += q.entry.count;
            batch.push_back(&q);
            tickets.push_back(pending.front());
            add += q.bytes;
            pending.pop_front();
            if (!small) break;
        }
        if (batch.empty()) {
            // the head request waits for in-flight batches to release budget
            cv.wait(lk);
            continue;
        }
        Slot& s = *slots[si];
        s.busy = true;
        s.bytes = add;
        inflight += add;
        stats.peak_inflight_bytes = std::max<uint64_t>(stats.peak_inflight_bytes, inflight);
        stats.batches++;
        lk.unlock();
        const Clock::time_point admitted = Clock::now();
        const int st = launch_batch(s, si, batch);
        const Clock::time_point launched = Clock::now();
        lk.lock();
        for (Req* q : batch) {
            q->admitted = admitted;
            q->launched = launched;
        }
        s.error = st;
        s.failed.clear();
        if (st != 0) {
            // requests whose launch group was not enqueued complete with the error
            for (uint64_t t : tickets) {
                bool grouped = false;
                for (const auto& g : s.groups) grouped = grouped || std::find(g.begin(), g.end(), t) != g.end();
                if (!grouped) s.failed.push_back(t);
            }
        }
        s.launched = true;
        cv.notify_all();
    }
}

void Server::waiter_loop(Slot* s)
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv.wait(lk, [&] { return stop || s->launched; });
        if (!s->launched) break;  // stopping with nothing in flight
        auto finish = [&](const std::vector<uint64_t>& tickets, int st) {
            const Clock::time_point now = Clock::now();
            for (uint64_t t : tickets) {
                auto it = reqs.find(t);
                if (it == reqs.end()) continue;
                it->second.completed = now;
                it->second.status = st;
                stats.completed++;
            }
        };
        // groups complete in launch order (one stream); each releases its requests at once
        for (size_t i = 0; i < s->groups.size(); i++) {
            lk.unlock();
            const int st = hipEventSynchronize(s->events[i]) == hipSuccess ? 0 : 3;
            lk.lock();
            finish(s->groups[i], st);
            inflight -= s->group_bytes[i];
            s->bytes -= s->group_bytes[i];
            cv.notify_all();
        }
        finish(s->failed, 3);
        inflight -= s->bytes;  // the failed requests' bytes
        s->bytes = 0;
        s->groups.clear();
        s->group_bytes.clear();
        s->failed.clear();
        s->launched = false;
        s->busy = false;
        cv.notify_all();
    }
}

Server* serve_create(const ServeConfig& cfg, std::string* why)
{
    if (cfg.slots < 1 || cfg.slots > 8 || cfg.max_batch < 1 || cfg.hbm_budget == 0) {
        *why = "invalid server configuration";
        return nullptr;
    }
    auto* s = new Server();
    s->cfg = cfg;
    if (!s->init(why)) {
        delete s;
        return nullptr;
    }
    return s;
}

void serve_destroy(Server* s)
{
    if (s == nullptr) return;
    // let in-flight batches finish before the threads stop
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv.wait(lk, [&] {
            if (!s->pending.empty()) return false;
            for (auto& sl : s->slots) {
                if (sl->busy) return false;
            }
            return true;
        });
    }
    delete s;
}

int serve_submit(Server* s, const ServeRequest& r, uint64_t* ticket, std::string* why)
{
    if (r.type != kTypeC2C && r.type != kTypeC2R && r.type != kTypeR2C) {
        *why = "unsupported type";
        return 1;
    }
    if (r.rank < 1 || r.rank > 2 || r.batch < 1 || r.in == nullptr || r.out == nullptr) {
        *why = "invalid rank, batch or buffer";
        return 1;
    }
    int64_t inner = 1;
    for (int d = 0; d < r.rank; d++) {
        if (r.dims[d] < 1 || r.dims[d] > ((int64_t)1 << 27)) {
            *why = "invalid size";
            return 1;
        }
        if (d + 1 < r.rank) inner *= r.dims[d];
    }
    const int64_t last = r.dims[r.rank - 1];
    Req q;
    q.r = r;
    q.bytes = (size_t)((kind_in_bytes(r.type, inner, last) + kind_out_bytes(r.type, inner, last)) * r.batch);
    q.submitted = Clock::now();
    if (r.rank == 1) {
        q.entry.kind = ragged_kind(r.type);
        q.entry.inverse = r.inverse;
        q.entry.n = last;
        q.entry.count = r.batch;
        ClassShape cs;
        ragged_shape(q.entry, &cs);
        q.ragged = s->cfg.batching && cs.n >= 2 && cs.n <= 16384 && !choose_radices(cs.n).empty() &&
                   r.batch <= kSlotTransforms && reinterpret_cast<uintptr_t>(r.in) % cs.in_align == 0 &&
                   reinterpret_cast<uintptr_t>(r.out) % cs.out_align == 0;
    }
    char key[160];
    snprintf(key, sizeof(key), "%d:%d:%d:%lld:%lld:%lld", r.type, (int)r.inverse, r.rank, (long long)r.dims[0],
             (long long)(r.rank > 1 ? r.dims[1] : 0), (long long)r.batch);
    q.plain_key = key;
    if (q.ragged && s->pool_class(q.entry) < 0) q.ragged = false;  // class resident before any batch
    {
        // every request's plain plan is built (tables resident) at submission: a plan
        // failure is reported here, and a ragged-eligible request that ends up alone in a
        // batch runs through it without plan work on the scheduler thread
        std::string w;
        PlainPlan* p = s->plain_plan(q, &w);
        if (p == nullptr) {
            *why = w;
            return 2;
        }
        if (p->built.workspace > s->cfg.hbm_budget / 4) {
            *why = "plan workspace larger than a quarter of the HBM budget";
            return 1;
        }
    }
    std::lock_guard<std::mutex> lock(s->mu);
    const size_t budget = s->cfg.hbm_budget - std::min(s->cfg.hbm_budget, s->resident);
    if (q.bytes > budget) {
        *why = "request larger than the HBM budget";
        return 1;
    }
    const uint64_t t = s->next_ticket++;
    q.ticket = t;
    s->reqs[t] = q;
    s->pending.push_back(t);
    s->stats.submitted++;
    if (q.ragged) {
        s->stats.ragged_requests++;
    } else {
        s->stats.plain_requests++;
    }
    *ticket = t;
    s->cv.notify_all();
    return 0;
}

int serve_wait(Server* s, uint64_t ticket, int64_t timeout_us, double* latency_us)
{
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->reqs.find(ticket) == s->reqs.end()) return 1;
    auto ready = [&] {
        auto f = s->reqs.find(ticket);
        return f == s->reqs.end() || f->second.status >= 0;
    };
    if (timeout_us < 0) {
        s->cv.wait(lk, ready);
    } else if (!s->cv.wait_for(lk, std::chrono::microseconds(timeout_us), ready)) {
        return 4;
    }
    auto f = s->reqs.find(ticket);
    if (f == s->reqs.end()) return 1;  // retired by another waiter
    Req& q = f->second;
    const int st = q.status;
    if (latency_us != nullptr) {
        *latency_us = std::chrono::duration<double, std::micro>(q.completed - q.submitted).count();
    }
    s->reqs.erase(ticket);
    return st;
}

int serve_replay(Server* s, const ServeRequest* reqs, const double* arrival_us, int64_t n, double* latency_us)
{
    std::vector<uint64_t> tickets((size_t)n, 0);
    std::vector<Clock::time_point> due((size_t)n);
    const char* tr = getenv("ASDFFT_SERVE_TRACE");
    const bool trace = tr != nullptr && tr[0] == '1';
    std::vector<double> parts[4];
    const Clock::time_point t0 = Clock::now() + std::chrono::milliseconds(1);
    int worst = 0;
    for (int64_t i = 0; i < n; i++) {
        due[i] = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double, std::micro>(
                          arrival_us[i]));
        while (Clock::now() < due[i]) {
        }
        std::string why;
        const int st = serve_submit(s, reqs[i], &tickets[i], &why);
        if (st != 0) {
            tickets[i] = 0;
            if (worst == 0) {
                const char* q = getenv("ASDFFT_QUIET");
                if (q == nullptr || q[0] != '1') {
                    fprintf(stderr, "[asdfft] asdFftExtServeReplay: request %lld refused: %s\n", (long long)i,
                            why.c_str());
                }
            }
            worst = std::max(worst, st);
        }
    }
    for (int64_t i = 0; i < n; i++) {
        if (tickets[i] == 0) {
            latency_us[i] = -1.0;
            continue;
        }
        std::unique_lock<std::mutex> lk(s->mu);
        // a ticket that another thread's asdFftExtServeWait already retired is gone from the
        // table: it counts as done (its latency went to that caller), never as a new entry
        s->cv.wait(lk, [&] {
            auto it = s->reqs.find(tickets[i]);
            return it == s->reqs.end() || it->second.status >= 0;
        });
        auto it = s->reqs.find(tickets[i]);
        if (it == s->reqs.end()) {
            latency_us[i] = -1.0;
            continue;
        }
        Req& q = it->second;
        latency_us[i] = std::chrono::duration<double, std::micro>(q.completed - due[i]).count();
        if (trace) {
            auto us = [](Clock::time_point a, Clock::time_point b) {
                return std::chrono::duration<double, std::micro>(b - a).count();
            };
            parts[0].push_back(us(due[i], q.submitted));
            parts[1].push_back(us(q.submitted, q.admitted));
            parts[2].push_back(us(q.admitted, q.launched));
            parts[3].push_back(us(q.launched, q.completed));
        }
        if (q.status != 0) worst = std::max(worst, q.status);
        s->reqs.erase(it);
    }
    if (trace && !parts[0].empty()) {
        // ASDFFT_SERVE_TRACE=1: where a replay's latency goes, p50 / p99 of each part
        static const char* names[4] = {"arrival->submitted", "submitted->admitted", "admitted->launched",
                                       "launched->completed"};
        fprintf(stderr, "[asdfft] replay of %lld requests (batching %d):", (long long)n, s->cfg.batching ? 1 : 0);
        for (int k = 0; k < 4; k++) {
            std::vector<double>& v = parts[k];
            std::sort(v.begin(), v.end());
            fprintf(stderr, " %s %.1f/%.1f", names[k], v[v.size() / 2], v[std::min(v.size() - 1, v.size() * 99 / 100)]);
        }
        fprintf(stderr, " us\n");
    }
    return worst;
}

void serve_stats(Server* s, ServeStats* out)
{
    std::lock_guard<std::mutex> lock(s->mu);
    *out = s->stats;
    out->resident_bytes = s->resident;
}

}  // namespace asdfft
