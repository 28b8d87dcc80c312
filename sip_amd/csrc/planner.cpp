#include "planner.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <functional>
#include <map>
#include <mutex>
#include <sstream>
#include <unordered_map>

namespace asdfft {

namespace {

constexpr int kPrimes[] = {2, 3, 5, 7, 11, 13, 17, 19};
constexpr int kNumPrimes = 8;
constexpr size_t kLdsLimit = 160 * 1024;      // gfx950 LDS per workgroup
constexpr size_t kLdsPreferred = 64 * 1024;   // keep >= 2 workgroups per CU when batching rows
constexpr int64_t kNumCUs = 256;               // MI355X compute units (8 XCDs x 32)

struct RadixInfo {
    int radix;
    int exps[kNumPrimes];
    double cost;  // relative per-point butterfly cost, tie-breaker only
};

const RadixInfo kRadices[] = {
    {16, {4, 0, 0, 0, 0, 0, 0, 0}, 1.25}, {12, {2, 1, 0, 0, 0, 0, 0, 0}, 1.6}, {10, {1, 0, 1, 0, 0, 0, 0, 0}, 1.7},
    {9, {0, 2, 0, 0, 0, 0, 0, 0}, 1.6},   {8, {3, 0, 0, 0, 0, 0, 0, 0}, 1.1},  {6, {1, 1, 0, 0, 0, 0, 0, 0}, 1.4},
    {4, {2, 0, 0, 0, 0, 0, 0, 0}, 1.0},   {2, {1, 0, 0, 0, 0, 0, 0, 0}, 1.0},  {3, {0, 1, 0, 0, 0, 0, 0, 0}, 1.3},
    {5, {0, 0, 1, 0, 0, 0, 0, 0}, 1.6},   {7, {0, 0, 0, 1, 0, 0, 0, 0}, 2.2},  {11, {0, 0, 0, 0, 1, 0, 0, 0}, 3.0},
    {13, {0, 0, 0, 0, 0, 1, 0, 0}, 3.4},  {17, {0, 0, 0, 0, 0, 0, 1, 0}, 4.3}, {19, {0, 0, 0, 0, 0, 0, 0, 1}, 4.8},
};

struct Best {
    int passes = 1 << 30;
    int maxr = 0;
    double cost = 0;
    std::vector<int> radices;
};

bool better(const Best& a, const Best& b)
{
    if (a.passes != b.passes) return a.passes < b.passes;
    const bool a16 = a.maxr <= 16, b16 = b.maxr <= 16;
    if (a16 != b16) return a16;
    return a.cost < b.cost - 1e-9;
}

Best search(std::vector<int>& e, std::map<std::vector<int>, Best>& memo)
{
    bool done = true;
    for (int v : e) done = done && (v == 0);
    if (done) {
        Best b;
        b.passes = 0;
        return b;
    }
    auto it = memo.find(e);
    if (it != memo.end()) return it->second;
    Best best;
    for (const RadixInfo& r : kRadices) {
        bool ok = true;
        for (int i = 0; i < kNumPrimes; i++) ok = ok && (r.exps[i] <= e[i]);
        if (!ok) continue;
        for (int i = 0; i < kNumPrimes; i++) e[i] -= r.exps[i];
        Best sub = search(e, memo);
        for (int i = 0; i < kNumPrimes; i++) e[i] += r.exps[i];
        if (sub.passes >= (1 << 30)) continue;
        Best cand = sub;
        cand.passes += 1;
        cand.maxr = std::max(cand.maxr, r.radix);
        cand.cost += r.cost;
        cand.radices.push_back(r.radix);
        if (better(cand, best)) best = cand;
    }
    memo[e] = best;
    return best;
}

int64_t pad16(int64_t x) { return x + (x >> 4); }

uint32_t magic_for(int span)
{
    if (span <= 1) return 0;
    return (uint32_t)((((uint64_t)1) << 32) / (uint64_t)span + 1);
}

std::string radix_key(int64_t n, const std::vector<int>& r)
{
    std::ostringstream os;
    os << "stk:" << n;
    for (int v : r) os << ":" << v;
    return os.str();
}

struct Config {
    std::vector<int> radices;
    int emax = 0;
    int threads = 0;
    int tpb = 1;
    int64_t tstride = 0;  // rows mapping
    int64_t plane = 0;
    size_t lds = 0;
    bool split = false;  // one-plane LDS exchange (the variant's specialised split kernel)
};

// Rows mapping: lanes along a transform; several short transforms per workgroup.
// want_tpb > 0 asks for that many transforms per workgroup (transposed stores need
// >= 16 adjacent transforms for 128-byte runs) and may use the whole LDS.
// The fewest-pass radix sequence, or, when its largest per-thread share does not fit
// the register budgets, the first radix multiset (more passes) that does.
std::vector<int> fitting_radices(int64_t n)
{
    std::vector<int> r = choose_radices(n);
    if (r.empty()) return r;
    for (int emax : {16, 32, 48}) {
        if (threads_for(n, r, emax) > 0) return r;
    }
    for (const std::vector<int>& alt : radix_multisets(n, 3, 64)) {
        for (int emax : {16, 32, 48}) {
            if (threads_for(n, alt, emax) > 0) return alt;
        }
    }
    return r;
}

bool rows_config(int64_t n, int64_t count, bool epi_lds, int want_tpb, int variant, Config* c)
{
    std::vector<int> radices = fitting_radices(n);
    if (radices.empty()) return false;
    // 4096-point C2C rows with at most one transform per CU: four radix-8 passes on 512 threads
    // x 8 points instead of three radix-16 passes on 256 x 16, so each CU has twice the waves to
    // hide one transform's latency (4096 x 256: 10.24 against 10.52-10.56 us per sample; at 512
    // transforms, two per CU, level: 13.84 against 13.76-13.80).  ASDFFT_WIDE8=0 turns it off.
    struct WideRows {
        int n, threads;
        int radices[4];
    };
    // (1000 x 256 on 5.5.5.8 with 200 threads measured slower: 7.94 against 7.40-7.44 us)
    static const WideRows kWide[] = {{4096, 512, {8, 8, 8, 8}}};
    const char* w8 = getenv("ASDFFT_WIDE8");
    const char* nst8 = getenv("ASDFFT_NO_STATIC");  // the geometry exists only as an instance
    const char* rs8 = getenv("ASDFFT_ROWS_SPLIT");  // the opt-in one-plane rows keep their radices
    const WideRows* wide = nullptr;
    if (!(w8 != nullptr && w8[0] == '0') && !(nst8 != nullptr && nst8[0] == '1') && !(rs8 != nullptr && rs8[0] == '1') &&
        variant == kVarRowsC2C && want_tpb <= 0 && count <= kNumCUs) {
        for (const WideRows& w : kWide) {
            if (w.n == n && find_static(kVarRowsC2C, 16, w.n, w.threads, 0, w.radices, 4) >= 0) wide = &w;
        }
    }
    const bool wide8 = wide != nullptr;
    if (wide8) radices.assign(wide->radices, wide->radices + 4);
    // Small latency-bound C2C row launches (<= 16 MiB of input) whose passes are all below radix
    // 12: the smallest radix first.  Measured at batch 256: 2187 (3.9.9.9) 7.75 vs 8.21 us, 6000
    // (6.10.10.10) 16.46 vs 17.10, 3000 12.89 vs 13.39, 1215 6.63 vs 6.97; at batch 4096 and for
    // sequences with radix 12 or 16 the reorder is neutral or slower.
    if (variant == kVarRowsC2C && want_tpb <= 0 && radices.size() >= 3 && count * n * 8 <= (16 << 20) &&
        *std::max_element(radices.begin(), radices.end()) < 12) {
        std::rotate(radices.begin(), std::min_element(radices.begin(), radices.end()), radices.end());
    }
    bool found = false;
    Config best;
    for (int emax : {16, 32, 48}) {
        if (!has_instance(variant, emax)) continue;
        const int T = threads_for(n, radices, emax);
        if (T == 0) continue;
        const int maxb = max_block_for(emax);
        const int target = want_tpb > 0 ? want_tpb : std::max(1, 256 / T);
        int tpb = std::min(target, maxb / T);
        if (want_tpb <= 0) {
            // small batches: fewer transforms per workgroup until the launch has two workgroups
            // per CU (1000 x 256: 4.77 us at one transform per workgroup, 5.44 us at two, whose
            // 128 workgroups leave half the CUs idle), keeping at least one full wavefront
            const int64_t spread = std::max<int64_t>((64 + T - 1) / T, count / (2 * kNumCUs));
            tpb = (int)std::min<int64_t>(tpb, spread);
        }
        tpb = (int)std::max<int64_t>(1, std::min<int64_t>(tpb, count));
        const bool need_lds = radices.size() > 1 || epi_lds;
        const int64_t tstride = pad16(n) + 1;
        const size_t cap = (want_tpb > 0) ? kLdsLimit : kLdsPreferred;
        while (tpb > 1 && need_lds && (size_t)(8 * tpb * tstride) > cap) tpb--;
        const size_t lds = need_lds ? (size_t)(8 * tpb * tstride) : 0;
        if (lds > kLdsLimit) continue;
        Config cand;
        cand.radices = radices;
        cand.emax = emax;
        cand.threads = T;
        cand.tpb = tpb;
        cand.tstride = tstride;
        cand.plane = tpb * tstride;
        cand.lds = lds;
        if (!found || (want_tpb > 0 && cand.tpb > best.tpb)) best = cand;
        found = true;
        if (want_tpb <= 0) break;  // first feasible register budget (EMAX 16 preferred)
    }
    // Launches of one transform per workgroup and at most two per CU: where a specialised
    // instance exists, the EMAX-16 geometry with more threads per transform (e.g. 6000 on
    // 600 threads x 10 points instead of 200 x 30), so each CU has more waves and shorter
    // per-thread chains to hide the latency of a single transform.  ASDFFT_WIDE_ROWS=0 turns
    // it off.
    const char* wr = getenv("ASDFFT_WIDE_ROWS");
    const char* nst = getenv("ASDFFT_NO_STATIC");  // the wide geometry exists only as an instance
    if (found && variant == kVarRowsC2C && want_tpb <= 0 && best.tpb == 1 && best.emax > 16 &&
        count <= 2 * kNumCUs && !(wr != nullptr && wr[0] == '0') && !(nst != nullptr && nst[0] == '1')) {
        int64_t t = 0;
        for (int r : radices) {
            t = (r > 16) ? 0 : std::max<int64_t>(t, (n / r + 16 / r - 1) / (16 / r));
            if (t == 0) break;
        }
        if (t > 0 && t <= 1024 &&
            find_static(kVarRowsC2C, 16, (int)n, (int)t, 0, radices.data(), (int)radices.size()) >= 0) {
            best.emax = 16;
            best.threads = (int)t;
        }
    }
    if (found && wide8 && best.emax == 16) {
        best.threads = wide->threads;
        best.tpb = 1;
        best.plane = best.tstride;
        best.lds = (size_t)(8 * best.tstride);
    }
    // ASDFFT_ROWS_SPLIT=1: rows C2C on a specialised one-plane instance for the chosen geometry
    // (half the LDS, twice the resident workgroups).  Opt-in: measured no faster than the
    // two-plane kernel (1024 x 8192: 28.6 vs 28.7 us; 1000 / 2187 / 6000 x 256: 9-10% slower),
    // the points carried in registers across passes cost more than the occupancy gains.
    const char* rs = getenv("ASDFFT_ROWS_SPLIT");
    const char* ns = getenv("ASDFFT_NO_STATIC");
    if (found && variant == kVarRowsC2C && best.lds > 0 && rs != nullptr && rs[0] == '1' &&
        !(ns != nullptr && ns[0] == '1') &&
        find_static(kVarRowsC2CSplit, best.emax, (int)n, best.threads, best.tpb, radices.data(),
                    (int)radices.size()) >= 0) {
        best.split = true;
        best.lds = (size_t)(4 * best.tpb * best.tstride);
    }
    if (found) *c = best;
    return found;
}

// The one-plane specialised variant of a two-plane variant (-1: none).
int split_variant_of(int variant)
{
    switch (variant) {
        case kVarColsC2C: return kVarColsC2CSplit;
        case kVarColsShiftOut: return kVarColsShiftOutSplit;
        case kVarRowsC2C: return kVarRowsC2CSplit;
        case kVarRowsQuad4: return kVarRowsQuad4Split;
        default: return -1;
    }
}

// Column mapping: lanes across adjacent transforms (columns) for coalesced strided access.
bool cols_config(int64_t n, int64_t count, int variant, Config* c, bool prefer_split = false)
{
    std::vector<int> radices = fitting_radices(n);
    if (radices.empty()) return false;
    // ASDFFT_COLS="n:emax:T:tpb" forces a column geometry that has a specialised kernel
    // (measurement of alternatives)
    if (const char* e = getenv("ASDFFT_COLS")) {
        long fn = 0, fe = 0, ft = 0, fp = 0;
        if (sscanf(e, "%ld:%ld:%ld:%ld", &fn, &fe, &ft, &fp) == 4 && fn == n &&
            find_static(kVarColsC2C, (int)fe, (int)n, (int)ft, (int)fp, radices.data(), (int)radices.size()) >= 0) {
            c->radices = radices;
            c->emax = (int)fe;
            c->threads = (int)ft;
            c->tpb = (int)fp;
            c->tstride = 0;
            c->plane = n * (fp + 1);
            c->lds = (size_t)(8 * c->plane);
            return true;
        }
    }
    // 16 adjacent complex columns are 128-byte runs; a float plane needs 32 for the same
    // (a half-line read by two workgroups on different XCDs is fetched twice)
    // ISTFT frames (kVarColsToRows) read columns nframes * 8 bytes apart, a stride that is in
    // general not a multiple of a line: runs of 4 columns and more workgroups per CU measured
    // fastest (1024 x 16000 columns: 119.7 us at 4 per workgroup, 134.3 at 8, 150.9 at 16;
    // 400 x 32000: 93.4 / 117.6 / 127.7 us)
    const int first_tpb = (variant == kVarColsTwiddleSep) ? 32 : (variant == kVarColsToRows) ? 4 : 16;
    // ASDFFT_COLS_LDS_CAP=<bytes> caps a column workgroup's LDS (measurement)
    const char* capenv = getenv("ASDFFT_COLS_LDS_CAP");
    const size_t lds_cap = capenv ? (size_t)atol(capenv) : kLdsLimit;
    bool found = false;
    for (int tpb : {32, 16, 8, 4, 2, 1}) {
        if (tpb > first_tpb) continue;
        for (int emax : {16, 32, 48}) {
            if (!has_instance(variant, emax)) continue;
            const int T = threads_for(n, radices, emax);
            if (T == 0 || tpb * T > max_block_for(emax)) continue;
            const int64_t plane = n * (tpb + 1);  // column layout: row pitch tpb + 1 words
            const bool need_lds = radices.size() > 1 || variant == kVarColsToRows;  // (epilogue via LDS)
            const size_t lds = need_lds ? (size_t)(8 * plane) : 0;
            if (lds > kLdsLimit || (tpb > 1 && lds > lds_cap)) continue;
            c->radices = radices;
            c->emax = emax;
            c->threads = T;
            c->tpb = (int)std::max<int64_t>(1, std::min<int64_t>(tpb, count));
            c->tstride = 0;
            c->plane = plane;
            c->lds = lds;
            found = true;
            break;
        }
        if (found) break;
    }
    // Long columns whose two LDS planes leave one workgroup per CU: a specialised kernel
    // exchanging through one plane takes the most columns (16, else 8) whose plane fits the
    // LDS.  Full 128-byte runs matter more than a second workgroup per CU once the data comes
    // from HBM rather than the device cache: the 2048 x 2048 x 4 column pass, timed over a ring
    // of outputs larger than that cache, takes 74.5 us with 16 columns on one plane (139 KiB,
    // one workgroup per CU), 83.3 us with 8 (72 KiB, two per CU) and 95.3 us on the two-plane
    // kernel (8 columns).  ASDFFT_SPLIT_EMAX=0 disables it; ASDFFT_SPLIT_EMAX / ASDFFT_SPLIT_TPB
    // force the register budget / columns per workgroup of a specialised instance (measurement).
    const char* ns = getenv("ASDFFT_NO_STATIC");
    // ASDFFT_COLS_SPLIT=1: the one-plane kernel for every column pass that has an instance
    // (measurement)
    const char* cs = getenv("ASDFFT_COLS_SPLIT");
    if (cs != nullptr && cs[0] == '1') prefer_split = true;
    const char* se = getenv("ASDFFT_SPLIT_EMAX");
    const int split_only = se ? atoi(se) : -1;
    const char* st = getenv("ASDFFT_SPLIT_TPB");
    const int split_tpb_forced = st ? atoi(st) : 0;
    const int split_variant = split_variant_of(variant);
    // the one-plane kernel wherever the two planes leave fewer than four workgroups per CU and an
    // instance exists: 512-point columns (512^2 x 64: 100.4 against 122.7 us, the column pass 65.4
    // against 81.2 us), 1024 and longer.  At 256 points (four per CU either way) the two-plane
    // kernel stays (60.4 against 62.4 us).
    const size_t four_per_cu = kLdsLimit / 4;
    if (split_variant >= 0 && found && (c->lds > four_per_cu || prefer_split) && split_only != 0 &&
        !(ns != nullptr && ns[0] == '1')) {
        bool done = false;
        for (int tpb : {16, 8}) {
            if (split_tpb_forced > 0 && tpb != split_tpb_forced) continue;
            const size_t lds = (size_t)n * (tpb + 1) * 4;
            if (count < tpb || lds > kLdsLimit) continue;
            for (int emax : {16, 32, 64}) {
                if (split_only > 0 && emax != split_only) continue;
                const int T = threads_for(n, radices, emax);
                if (T == 0 || tpb * T > 1024) continue;
                if (find_static(split_variant, emax, (int)n, T, tpb, radices.data(), (int)radices.size()) < 0) continue;
                c->emax = emax;
                c->threads = T;
                c->tpb = tpb;
                c->plane = n * (tpb + 1);
                c->lds = lds;
                c->split = true;
                done = true;
                break;
            }
            if (done) break;
        }
    }
    return found;
}

void fill_passes(const Config& c, FftArgs* a, std::vector<int>* offsets)
{
    a->passes = (int)c.radices.size();
    int span = 1;
    for (size_t p = 0; p < c.radices.size(); p++) {
        a->pass[p].radix = c.radices[p];
        a->pass[p].span = span;
        a->pass[p].magic = magic_for(span);
        a->pass[p].tw_offset = (*offsets)[p];
        span *= c.radices[p];
    }
}

// Column passes with runs shorter than a 128-byte line share lines between neighbouring
// workgroups; keeping neighbours on one XCD lets its L2 serve the second half.
// ASDFFT_XCD_REMAP=0/1 overrides (for measurement).
int xcd_remap_for(bool cols_mapping, int tpb)
{
    const char* e = getenv("ASDFFT_XCD_REMAP");
    if (e != nullptr && (e[0] == '0' || e[0] == '1')) return e[0] - '0';
    return (cols_mapping && tpb < 16) ? 1 : 0;
}

// Row launches of at most this many workgroups take a twiddle-preloading instance where one
// exists (latency-bound launches: a few workgroups per CU, so the extra registers cost no
// occupancy that matters); ASDFFT_TW_PRELOAD=0 / 1 turns them off / on for every row launch.
bool tw_preload_wanted(int64_t grid)
{
    const char* e = getenv("ASDFFT_TW_PRELOAD");
    if (e != nullptr && (e[0] == '0' || e[0] == '1')) return e[0] == '1';
    return grid <= 1024;
}

// A Stockham launch.  Addressing: transform t at (t % cols)*inner + (t / cols)*outer.
struct StepIo {
    int64_t count = 0;
    int64_t cols = 1;
    int64_t in_inner = 0, in_outer = 0, in_stride = 1;
    int64_t out_inner = 0, out_outer = 0, out_stride = 1;
    int64_t o_split = 0, in_outer2 = 0, out_outer2 = 0;  // split outer index (FftArgs::o_split)
    bool prefer_split = false;  // columns: the one-plane kernel even where two planes fit twice per CU
};

bool stockham_step(int64_t n, int variant, bool cols_mapping, const StepIo& io, int conj_in, int conj_out,
                   int want_tpb, Step* st, std::string* why)
{
    Config c;
    const bool pair = (variant == kVarRowsPair2);
    const bool quad = (variant == kVarRowsQuad4);
    const bool epi_lds = (variant == kVarR2CEven || variant == kVarRowsTranspose || variant == kVarRowsTransposeSep ||
                          pair || quad);
    // the pair and quad variants have the row C2C kernel's geometry with two / four transforms
    // per workgroup (a quad workgroup may exceed the C2C kernel's block: its instances fix it)
    const bool ok = cols_mapping ? cols_config(n, io.count, variant, &c, io.prefer_split)
                                 : rows_config(n, io.count, epi_lds, pair ? 2 : quad ? 1 : want_tpb,
                                               (pair || quad) ? (int)kVarRowsC2C : variant, &c);
    if (!ok || (pair && (c.tpb != 2 || c.radices.size() < 2)) || (quad && (c.emax != 16 || c.radices.size() < 2))) {
        *why = "no single-workgroup schedule for length " + std::to_string(n);
        return false;
    }
    if (quad) {
        c.tpb = 4;
        c.plane = 4 * c.tstride;
    }
    if (pair || quad) {
        c.split = false;
        c.lds = (size_t)(8 * c.tpb * c.tstride);
    }
    // rows in fours through one LDS plane where an instance exists (ASDFFT_QUAD_SPLIT=0: two planes)
    const char* qs = getenv("ASDFFT_QUAD_SPLIT");
    const char* nst = getenv("ASDFFT_NO_STATIC");
    if (quad && !(qs != nullptr && qs[0] == '0') && !(nst != nullptr && nst[0] == '1') &&
        find_static(kVarRowsQuad4Split, c.emax, (int)n, c.threads, c.tpb, c.radices.data(), (int)c.radices.size()) >=
            0) {
        c.split = true;
        c.lds = (size_t)(4 * c.tpb * c.tstride);
    }
    std::vector<int> offsets;
    std::vector<float> tab = stockham_table((int)n, c.radices, &offsets);
    TablePtr t = upload_table(radix_key(n, c.radices), tab);
    st->tables_ready = (t != nullptr);
    if (t) st->tables.push_back(t);
    FftArgs& a = st->args;
    a = FftArgs{};
    a.tw = t ? t->dev : nullptr;
    a.n = (int)n;
    a.threads = c.threads;
    a.tpb = c.tpb;
    a.lds_tstride = (int)c.tstride;
    a.lds_plane = (int)c.plane;
    a.conj_in = conj_in;
    a.conj_out = conj_out;
    a.count = io.count;
    a.cols = io.cols;
    a.in_inner = io.in_inner;
    a.in_outer = io.in_outer;
    a.in_stride = io.in_stride;
    a.out_inner = io.out_inner;
    a.out_outer = io.out_outer;
    a.out_stride = io.out_stride;
    a.xcd_remap = xcd_remap_for(cols_mapping, c.tpb);
    a.in_o_mod = 1;
    a.out_scale = 1.0f;
    a.o_split = io.o_split;
    a.in_outer2 = io.in_outer2;
    a.out_outer2 = io.out_outer2;
    fill_passes(c, &a, &offsets);
    st->direct = false;
    st->shape.variant = !c.split ? variant : split_variant_of(variant);
    st->shape.emax = c.emax;
    st->shape.block = c.tpb * c.threads;
    st->shape.lds_bytes = c.lds;
    st->shape.grid = (io.count + c.tpb - 1) / c.tpb;
    st->shape.static_id = -1;
    const char* ns = getenv("ASDFFT_NO_STATIC");
    if (ns == nullptr || ns[0] != '1') {
        st->shape.static_id = find_static(st->shape.variant, c.emax, (int)n, c.threads, c.tpb, c.radices.data(),
                                          (int)c.radices.size());
        // small row launches: the instance with every pass's twiddles loaded at kernel start
        if (st->shape.variant == kVarRowsC2C && tw_preload_wanted(st->shape.grid)) {
            const int pre = find_static(kVarRowsC2CTwPre, c.emax, (int)n, c.threads, c.tpb, c.radices.data(),
                                        (int)c.radices.size());
            if (pre >= 0) {
                st->shape.variant = kVarRowsC2CTwPre;
                st->shape.static_id = pre;
            }
        }
    }
    std::ostringstream os;
    os << "stockham variant=" << st->shape.variant << " n=" << n << " count=" << io.count << " radices=";
    for (size_t i = 0; i < c.radices.size(); i++) os << (i ? "x" : "") << c.radices[i];
    os << " emax=" << c.emax << " T=" << c.threads << " tpb=" << c.tpb << " block=" << st->shape.block
       << " lds=" << c.lds << " grid=" << st->shape.grid;
    if (st->shape.static_id >= 0) os << " static=" << st->shape.static_id;
    // element strides and the largest element offset within one transform; the kernels form
    // offsets in 64 bits, so these may exceed 2^31
    if (io.in_stride != 1 || io.out_stride != 1) {
        os << " stride=" << io.in_stride << "/" << io.out_stride << " span=" << (n - 1) * io.in_stride << "/"
           << (n - 1) * io.out_stride;
    }
    st->describe = os.str();
    return true;
}

bool attach_aux(Step* st, int64_t nreal)
{
    const int64_t m = nreal / 2;
    TablePtr t = upload_table("aux:" + std::to_string(nreal), roots_table(nreal, m + 1));
    if (t) st->tables.push_back(t);
    st->tables_ready = st->tables_ready && (t != nullptr);
    st->args.tw_aux = t ? t->dev : nullptr;
    return true;
}

bool direct_step(int64_t n, int prologue, int epilogue, int64_t count, int64_t in_dist, int64_t out_dist, int conj_in,
                 int conj_out, Step* st, std::string* why)
{
    if (n > kMaxDirect) {
        *why = "length " + std::to_string(n) + " has prime factors above 19 and exceeds the direct-DFT limit";
        return false;
    }
    TablePtr t = upload_table("roots:" + std::to_string(n), roots_table(n, n));
    st->tables_ready = (t != nullptr);
    if (t) st->tables.push_back(t);
    DftArgs& d = st->dargs;
    d = DftArgs{};
    d.roots = t ? t->dev : nullptr;
    d.n = (int)n;
    int tpb = (int)std::max<int64_t>(1, std::min<int64_t>(count, std::max<int64_t>(1, 4096 / n)));
    d.tpb = tpb;
    d.conj_in = conj_in;
    d.conj_out = conj_out;
    d.prologue = prologue;
    d.epilogue = epilogue;
    d.count = count;
    d.in_dist = in_dist;
    d.out_dist = out_dist;
    d.in_cols = 1;
    d.in_stride = 1;
    st->direct = true;
    st->shape.variant = kVarDirect;
    st->shape.emax = 0;
    st->shape.block = 256;
    st->shape.lds_bytes = (size_t)tpb * (size_t)n * 8;
    st->shape.grid = (count + tpb - 1) / tpb;
    std::ostringstream os;
    os << "direct n=" << n << " count=" << count << " pro=" << prologue << " epi=" << epilogue << " tpb=" << tpb;
    st->describe = os.str();
    return true;
}

bool smooth_single(int64_t n) { return n > 1 && n <= kMaxSingleKernel && !choose_radices(n).empty(); }

// ---- 1D building blocks -------------------------------------------------------------
// Contiguous complex rows of length n: count rows, in/out row distance n.  pin / pout:
// the input / output is a pair of float planes (real, imaginary) instead of complex.
// Lane-butterfly kernel for short C2C rows: ASDFFT_WAVE=1 wherever it has an instance, 0 never;
// unset: never (the Stockham kernels measured faster at every length it covers).
int wave_mode()
{
    const char* e = getenv("ASDFFT_WAVE");
    if (e != nullptr && (e[0] == '0' || e[0] == '1')) return e[0] - '0';
    return -1;
}

bool wave_wanted(int64_t n)
{
    int e = 0, lanes = 0;
    return wave_mode() == 1 && wave_fft_shape(n, &e, &lanes);
}

// Contiguous C2C rows on the lane-butterfly kernel (wave_fft.hip).
void wave_rows(int64_t n, int64_t count, bool inverse, Step* st)
{
    int e = 0, lanes = 0;
    wave_fft_shape(n, &e, &lanes);
    st->wave = true;
    TablePtr w = upload_table("roots:" + std::to_string(n), roots_table(n, n));
    st->tables_ready = (bool)w;
    if (w) st->tables.push_back(w);
    WaveFftArgs& a = st->wargs;
    a.tw = w ? reinterpret_cast<const float2*>(w->dev) : nullptr;
    a.count = count;
    a.in_dist = n;
    a.out_dist = n;
    a.n = (int)n;
    a.e = e;
    a.lanes = lanes;
    a.conj_in = inverse;
    a.conj_out = inverse;
    std::ostringstream os;
    os << "wave n=" << n << " count=" << count << " points/lane=" << e << " lanes=" << lanes
       << " (cross-lane radix-2 stages, no LDS)";
    st->describe = os.str();
}

bool c2c_rows(int64_t n, int64_t count, bool inverse, int src, int dst, BuiltPlan* bp, std::string* why,
              bool pin = false, bool pout = false)
{
    Step st;
    st.src = src;
    st.dst = dst;
    st.planar_in = pin;
    st.planar_out = pout;
    st.bytes = 16.0 * (double)n * (double)count;
    if (n == 1 || !smooth_single(n)) {
        if (n > kMaxSingleKernel && !choose_radices(n).empty()) {
            *why = "large length handled by four-step";
            return false;
        }
        if (!direct_step(n, pin ? kProPlanar : kProComplex, pout ? kEpiPlanar : kEpiComplex, count, n, n, inverse,
                         inverse, &st, why))
            return false;
    } else if (!pin && !pout && wave_wanted(n)) {
        wave_rows(n, count, inverse, &st);
    } else {
        if (pout && !pin) {
            *why = "no row kernel with a complex input and a planar output";
            return false;
        }
        StepIo io;
        io.count = count;
        io.cols = 1;
        io.in_outer = n;
        io.out_outer = n;
        const int variant = pin ? (pout ? kVarRowsSep : kVarRowsSepIn) : kVarRowsC2C;
        if (!stockham_step(n, variant, false, io, inverse, inverse, 0, &st, why)) return false;
    }
    bp->steps.push_back(std::move(st));
    return true;
}

// Strided complex transforms of length n along an axis with element stride `stride`;
// `cols` adjacent transforms per outer group, groups `outer` elements apart.
// pout: the output is a pair of float planes (real, imaginary).
bool c2c_cols(int64_t n, int64_t count, int64_t cols, int64_t outer, int64_t stride, bool inverse, int src, int dst,
              BuiltPlan* bp, std::string* why, bool pout = false)
{
    if (n == 1) {
        if (src == dst) return true;  // identity along this axis
        // length-1 axis: the transforms are the elements themselves, contiguous
        Step st;
        st.src = src;
        st.dst = dst;
        st.planar_out = pout;
        st.bytes = 16.0 * (double)count;
        if (!direct_step(1, kProComplex, pout ? kEpiPlanar : kEpiComplex, count, 1, 1, inverse, inverse, &st, why))
            return false;
        bp->steps.push_back(std::move(st));
        return true;
    }
    if (!smooth_single(n)) {
        *why = "axis length " + std::to_string(n) + " unsupported for a strided pass";
        return false;
    }
    Step st;
    st.src = src;
    st.dst = dst;
    st.planar_out = pout;
    st.bytes = 16.0 * (double)n * (double)count;
    StepIo io;
    io.count = count;
    io.cols = cols;
    io.in_inner = 1;
    io.in_outer = outer;
    io.in_stride = stride;
    io.out_inner = 1;
    io.out_outer = outer;
    io.out_stride = stride;
    if (!stockham_step(n, pout ? kVarColsSepOut : kVarColsC2C, true, io, inverse, inverse, 0, &st, why)) return false;
    bp->steps.push_back(std::move(st));
    return true;
}

// First pass of a four-step transform of length n = n1 * n2: n2-point columns at stride n1
// (count * n1 of them) times W_n^(column * k), from src into the workspace in the same layout.
bool fourstep_pass_a(int64_t n, int64_t n1, int64_t count, int src, size_t src_offset, size_t work_offset, bool planar,
                     int conj_in, Step* out, std::string* why)
{
    const int64_t n2 = n / n1;
    int shift = 0;
    while (((int64_t)1 << shift) * ((int64_t)1 << shift) < n) shift++;
    TablePtr lo = upload_table("biglo:" + std::to_string(n) + ":" + std::to_string(shift),
                               roots_table(n, (int64_t)1 << shift));
    const int64_t nhi = (n + ((int64_t)1 << shift) - 1) >> shift;
    TablePtr hi = upload_table("bighi:" + std::to_string(n) + ":" + std::to_string(shift),
                               roots_table(n, nhi, (int64_t)1 << shift));
    Step& a = *out;
    a.src = src;
    a.src_offset = src_offset;
    a.dst = kBufWork;
    a.dst_offset = work_offset;
    a.planar_in = planar;
    a.bytes = 16.0 * (double)n * (double)count;
    StepIo ia;
    ia.count = count * n1;
    ia.cols = n1;
    ia.in_inner = 1;
    ia.in_outer = n;
    ia.in_stride = n1;
    ia.out_inner = 1;
    ia.out_outer = n;
    ia.out_stride = n1;
    if (!stockham_step(n2, planar ? kVarColsTwiddleSep : kVarColsTwiddle, true, ia, conj_in, 0, 0, &a, why)) return false;
    a.args.tw_big_lo = lo ? lo->dev : nullptr;
    a.args.tw_big_hi = hi ? hi->dev : nullptr;
    a.args.big_shift = shift;
    if (lo) a.tables.push_back(lo);
    if (hi) a.tables.push_back(hi);
    a.tables_ready = a.tables_ready && lo && hi;
    a.describe += " (four-step A n=" + std::to_string(n) + ")";
    return true;
}

// Large complex length: n = n1 * n2, pass A = length-n2 columns (stride n1) with the
// W_n^{n1 k2} twiddle folded into its store, pass B = length-n1 rows stored transposed.
// planar: pass A reads and pass B writes a pair of float planes (real, imaginary).
// conj_in / conj_out (-1: `inverse`) conjugate pass A's input and pass B's output; src_offset /
// dst_offset are byte offsets into the source / destination buffers.
bool c2c_fourstep(int64_t n, int64_t count, bool inverse, int src, int dst, size_t work_offset, BuiltPlan* bp,
                  std::string* why, bool planar = false, int conj_in = -1, int conj_out = -1, size_t src_offset = 0,
                  size_t dst_offset = 0)
{
    if (conj_in < 0) conj_in = inverse ? 1 : 0;
    if (conj_out < 0) conj_out = inverse ? 1 : 0;
    int64_t best1 = 0;
    double best_score = -1e30;
    for (int64_t n1 = 2; n1 <= 1024 && n1 < n; n1++) {
        if (n % n1) continue;
        const int64_t n2 = n / n1;
        if (n2 > 4096 || !smooth_single(n1) || !smooth_single(n2)) continue;
        Config ca, cb;
        if (!cols_config(n2, count * n1, planar ? kVarColsTwiddleSep : kVarColsTwiddle, &ca)) continue;
        if (!rows_config(n1, count * n2, true, planar ? 32 : 16, planar ? kVarRowsTransposeSep : kVarRowsTranspose,
                         &cb))
            continue;
        const double score = 4.0 * std::min(ca.tpb, cb.tpb) - fabs(log2((double)n1) - log2((double)n2));
        if (score > best_score) {
            best_score = score;
            best1 = n1;
        }
    }
    if (best1 == 0) {
        *why = "length " + std::to_string(n) + " has no supported two-pass split";
        return false;
    }
    const int64_t n1 = best1, n2 = n / n1;
    Step a;
    if (!fourstep_pass_a(n, n1, count, src, src_offset, work_offset, planar, conj_in, &a, why)) return false;

    Step b;
    b.src = kBufWork;
    b.src_offset = work_offset;
    b.dst = dst;
    b.dst_offset = dst_offset;
    b.planar_out = planar;
    b.bytes = 16.0 * (double)n * (double)count;
    StepIo ib;
    ib.count = count * n2;
    ib.cols = n2;
    ib.in_inner = n1;
    ib.in_outer = n;
    ib.in_stride = 1;
    ib.out_inner = 1;
    ib.out_outer = n;
    ib.out_stride = n2;
    // transposed stores: 16 transforms per workgroup write 128-byte runs of complex
    // values, 32 those of a float plane
    if (!stockham_step(n1, planar ? kVarRowsTransposeSep : kVarRowsTranspose, false, ib, 0, conj_out,
                       planar ? 32 : 16, &b, why))
        return false;
    b.describe += " (four-step B)";
    bp->steps.push_back(std::move(a));
    bp->steps.push_back(std::move(b));
    return true;
}

// Real rows: R2C (in real [count][n] -> out [count][n/2+1]) or C2R (reverse).
bool real_rows(Kind kind, int64_t n, int64_t count, bool inverse, int src, int dst, BuiltPlan* bp, std::string* why)
{
    const int64_t h = n / 2 + 1;
    Step st;
    st.src = src;
    st.dst = dst;
    st.bytes = (4.0 * (double)n + 8.0 * (double)h) * (double)count;
    const bool even = (n % 2) == 0;
    if (even && n >= 4 && smooth_single(n / 2)) {
        const int64_t m = n / 2;
        StepIo io;
        io.count = count;
        io.cols = 1;
        if (kind == Kind::R2C) {
            io.in_outer = m;       // float2 units
            io.out_outer = m + 1;
            if (!stockham_step(m, kVarR2CEven, false, io, 0, inverse, 0, &st, why)) return false;
        } else {
            io.in_outer = m + 1;
            io.out_outer = m;
            if (!stockham_step(m, kVarC2REven, false, io, inverse ? 0 : 1, 1, 0, &st, why)) return false;
        }
        attach_aux(&st, n);
    } else if (smooth_single(n)) {
        StepIo io;
        io.count = count;
        io.cols = 1;
        if (kind == Kind::R2C) {
            io.in_outer = n;  // floats
            io.out_outer = h;
            if (!stockham_step(n, kVarR2COdd, false, io, 0, inverse, 0, &st, why)) return false;
        } else {
            io.in_outer = h;
            io.out_outer = n;  // floats
            if (!stockham_step(n, kVarC2ROdd, false, io, inverse, 0, 0, &st, why)) return false;
        }
    } else {
        if (n > kMaxDirect && !choose_radices(n).empty()) {
            *why = "odd real length " + std::to_string(n) + " above " + std::to_string(kMaxSingleKernel) +
                   " is not supported (one workgroup transforms at most that many complex points)";
            return false;
        }
        if (kind == Kind::R2C) {
            if (!direct_step(n, kProReal, kEpiHalf, count, n, h, 0, inverse, &st, why)) return false;
        } else {
            if (!direct_step(n, kProHerm, kEpiReal, count, h, n, inverse, 0, &st, why)) return false;
        }
    }
    bp->steps.push_back(std::move(st));
    return true;
}

// True when a real length is beyond one workgroup but its complex transform has a four-step
// schedule: even N with N/2 > 16384, odd N > 16384 (prime factors <= 19 either way).
bool real_needs_large(int64_t n)
{
    const int64_t m = (n % 2 == 0) ? n / 2 : n;
    return m > kMaxSingleKernel && !choose_radices(m).empty();
}

Step aux_step(int op, int64_t n, int64_t count, int conj, int src, size_t src_offset, int dst, size_t dst_offset,
              int64_t in_dist, int64_t out_dist, double bytes)
{
    Step st;
    st.aux = true;
    st.src = src;
    st.src_offset = src_offset;
    st.dst = dst;
    st.dst_offset = dst_offset;
    st.bytes = bytes;
    RealAuxArgs& a = st.raux;
    a = RealAuxArgs{};
    a.op = op;
    a.conj = conj;
    a.n = n;
    a.count = count;
    a.in_dist = in_dist;
    a.out_dist = out_dist;
    if (op == kAuxR2CPost || op == kAuxC2RPre) {
        const int64_t m = n / 2;
        TablePtr t = upload_table("aux:" + std::to_string(n), roots_table(n, m + 1));
        if (t) st.tables.push_back(t);
        st.tables_ready = (t != nullptr);
        a.tw = t ? t->dev : nullptr;
    }
    static const char* names[] = {"r2c-post", "c2r-pre", "promote", "hermitian-extend", "truncate", "real-part"};
    std::ostringstream os;
    os << "real " << names[op] << " n=" << n << " count=" << count;
    st.describe = os.str();
    return st;
}

// Real transforms beyond one workgroup (1D): the complex four-step around elementwise passes.
//   even R2C: four-step FFT_M of the input read as M complex -> post-twiddle into the output
//   even C2R: pre-twiddle -> four-step FFT_M (conjugated output) written as the real output
//   odd R2C:  promote to complex -> four-step FFT_N -> first N/2 + 1 bins
//   odd C2R:  Hermitian extension -> four-step FFT_N -> real parts
// Workspace: two complex buffers of the four-step's length per transform.
bool real_large(Kind kind, int64_t n, int64_t count, bool inverse, BuiltPlan* bp, std::string* why)
{
    const bool even = (n % 2) == 0;
    const int64_t m = even ? n / 2 : n;  // complex length of the four-step
    const int64_t h = n / 2 + 1;
    const size_t w = (size_t)(m * count * 8);
    const double real_bytes = 4.0 * (double)n * (double)count, spec_bytes = 8.0 * (double)h * (double)count;
    const double cplx_bytes = 8.0 * (double)m * (double)count;
    bp->workspace = 2 * w;
    if (kind == Kind::R2C) {
        if (even) {
            if (!c2c_fourstep(m, count, false, kBufIn, kBufWork, 0, bp, why, false, 0, 0, 0, w)) return false;
            bp->steps.push_back(aux_step(kAuxR2CPost, n, count, inverse ? 1 : 0, kBufWork, w, kBufOut, 0, m, m + 1,
                                         cplx_bytes + spec_bytes));
        } else {
            bp->steps.push_back(aux_step(kAuxPromote, n, count, 0, kBufIn, 0, kBufWork, 0, n, n, real_bytes + cplx_bytes));
            if (!c2c_fourstep(m, count, false, kBufWork, kBufWork, w, bp, why, false, 0, 0, 0, 0)) return false;
            bp->steps.push_back(aux_step(kAuxTruncate, n, count, inverse ? 1 : 0, kBufWork, 0, kBufOut, 0, n, h,
                                         2.0 * spec_bytes));
        }
    } else {
        if (even) {
            bp->steps.push_back(aux_step(kAuxC2RPre, n, count, inverse ? 0 : 1, kBufIn, 0, kBufWork, 0, m + 1, m,
                                         spec_bytes + cplx_bytes));
            if (!c2c_fourstep(m, count, false, kBufWork, kBufOut, w, bp, why, false, 0, 1, 0, 0)) return false;
        } else {
            bp->steps.push_back(aux_step(kAuxHermExtend, n, count, inverse ? 1 : 0, kBufIn, 0, kBufWork, 0, h, n,
                                         spec_bytes + cplx_bytes));
            if (!c2c_fourstep(m, count, false, kBufWork, kBufWork, w, bp, why, false, 0, 0, 0, 0)) return false;
            bp->steps.push_back(aux_step(kAuxRealPart, n, count, 0, kBufWork, 0, kBufOut, 0, n, n,
                                         cplx_bytes + real_bytes));
        }
    }
    return true;
}

// 1024-point C2C rows on the split-FP16 MFMA kernel (two radix-32 matmul stages).
void mfma_rows_1024(int64_t count, bool inverse, BuiltPlan* bp)
{
    Step st;
    st.mfma = true;
    st.bytes = 16.0 * 1024.0 * (double)count;
    TablePtr w32 = upload_table("roots:32", roots_table(32, 32));
    TablePtr w1024 = upload_table("roots:1024", roots_table(1024, 1024));
    st.tables_ready = w32 && w1024;
    if (w32) st.tables.push_back(w32);
    if (w1024) st.tables.push_back(w1024);
    st.margs.w32 = w32 ? reinterpret_cast<const float2*>(w32->dev) : nullptr;
    st.margs.tw1024 = w1024 ? reinterpret_cast<const float2*>(w1024->dev) : nullptr;
    st.margs.count = count;
    st.margs.conj_in = inverse;
    st.margs.conj_out = inverse;
    const int waves = mfma_fft1024_waves();
    // two resident workgroups per CU (LDS and registers), all rows covered
    st.mfma_grid = std::min<int64_t>((count + waves - 1) / waves, 2 * 256);
    std::ostringstream os;
    os << "mfma split-fp16 n=1024 count=" << count << " stages=32x32 grid=" << st.mfma_grid;
    st.describe = os.str();
    bp->steps.push_back(std::move(st));
}

// Whether a long C2C length has the split-FP16 four-step: n = 1024 * n2 with n2 a whole number
// of the transposed MFMA kernel's row tiles and an FP32 single-kernel column length.
bool mfma_fourstep_fits(int64_t n)
{
    if (n <= kMaxSingleKernel || n % 1024 != 0) return false;
    const int64_t n2 = n / 1024;
    return n2 % mfma_fft1024_tr_rows() == 0 && n2 <= 4096 && smooth_single(n2);
}

// Long C2C, n = 1024 * n2, with the four-step's 1024-point rows on the split-FP16 MFMA kernel:
// pass A is the FP32 n2-point column pass with the W_n twiddle (into the workspace), pass B
// transforms the count * n2 rows of 1024 points (two radix-32 matmul stages) and stores
// element j of row k of each transform to k + n2 j.
bool mfma_fourstep(int64_t n, int64_t count, bool inverse, BuiltPlan* bp, std::string* why)
{
    const int64_t n1 = 1024, n2 = n / n1;
    Step a;
    if (!fourstep_pass_a(n, n1, count, kBufIn, 0, 0, false, inverse ? 1 : 0, &a, why)) return false;
    Step b;
    b.mfma = true;
    b.src = kBufWork;
    b.dst = kBufOut;
    b.bytes = 16.0 * (double)n * (double)count;
    TablePtr w32 = upload_table("roots:32", roots_table(32, 32));
    TablePtr w1024 = upload_table("roots:1024", roots_table(1024, 1024));
    b.tables_ready = w32 && w1024;
    if (w32) b.tables.push_back(w32);
    if (w1024) b.tables.push_back(w1024);
    b.margs.w32 = w32 ? reinterpret_cast<const float2*>(w32->dev) : nullptr;
    b.margs.tw1024 = w1024 ? reinterpret_cast<const float2*>(w1024->dev) : nullptr;
    b.margs.count = count * n2;
    b.margs.conj_in = 0;
    b.margs.conj_out = inverse;
    b.margs.transposed = 1;
    b.margs.out_rows = n2;
    b.mfma_grid = count * n2 / mfma_fft1024_tr_rows();
    std::ostringstream os;
    os << "mfma split-fp16 n=1024 count=" << count * n2 << " stages=32x32 transposed rows=" << n2
       << " grid=" << b.mfma_grid << " (four-step B)";
    b.describe = os.str();
    bp->steps.push_back(std::move(a));
    bp->steps.push_back(std::move(b));
    return true;
}

}  // namespace

int max_block_for(int emax) { return emax <= 16 ? 256 : 512; }

namespace {

// 2D C2C with the first radix-2 stage of the columns (length X) folded into the row pass
// (ASDFFT_PAIR2D=1): rows n and n + X/2 of a slice share a workgroup, which stores u + v to row
// n and (u - v) W_X^n to row n + X/2 of the workspace; one column launch then transforms the
// two halves as length-X/2 columns, storing point k of half h to row 2k + h of the output.
// Paired rows for 2D C2C: on by default where both launches have specialised instances
// (ASDFFT_PAIR2D=0 turns it off, =1 forces it where the row instance exists).
int pair2d_mode()
{
    const char* e = getenv("ASDFFT_PAIR2D");
    if (e != nullptr && (e[0] == '0' || e[0] == '1')) return e[0] - '0';
    return -1;
}

bool pair2d_wanted(int64_t X, int64_t Y)
{
    return pair2d_mode() != 0 && X % 2 == 0 && X / 2 >= 2 && Y >= 2;
}

bool pair2d(int64_t X, int64_t Y, int64_t B, bool inv, BuiltPlan* out)
{
    const int64_t h = X / 2;
    std::string why;
    BuiltPlan bp;
    // W_X^n, n < X/2, forward sign, from long double
    std::vector<float> tab;
    static const long double kTwoPi = 6.283185307179586476925286766559005768L;
    for (int64_t m = 0; m < h; m++) {
        const long double ang = kTwoPi * (long double)m / (long double)X;
        tab.push_back((float)cosl(ang));
        tab.push_back((float)(-sinl(ang)));
    }
    TablePtr pt = upload_table("pair2d:" + std::to_string(X), tab);
    // The row pass reads rows n and n + X/2 of the input and writes their sum and twiddled
    // difference to rows 2n and 2n + 1 of the output.  Each column transform (parity p, column c)
    // then reads and rewrites exactly the elements (2i + p, c), i < X/2, so the column pass is in
    // place on the output with identical input and output addresses: no workgroup can overwrite
    // what another has still to read, and the plan needs no workspace.  (Writing the halves to
    // rows n and n + X/2 and interleaving them in the column pass raced: the parity-0 workgroup
    // of a column writes rows that the parity-1 workgroup of the same column reads.)
    Step rows;
    rows.src = kBufIn;
    rows.dst = kBufOut;
    rows.bytes = 16.0 * (double)(X * Y * B);
    StepIo ri;
    ri.count = B * X;
    ri.cols = 2;
    ri.in_inner = h * Y;
    ri.in_outer = Y;
    ri.in_outer2 = X * Y;
    ri.out_inner = Y;
    ri.out_outer = 2 * Y;
    ri.out_outer2 = X * Y;
    ri.o_split = h;
    if (!stockham_step(Y, kVarRowsPair2, false, ri, inv, 0, 2, &rows, &why) || rows.shape.static_id < 0) return false;
    rows.args.tw_aux = pt ? pt->dev : nullptr;
    rows.tables_ready = rows.tables_ready && pt != nullptr;
    if (pt) rows.tables.push_back(pt);
    Step cols;
    cols.src = kBufOut;
    cols.dst = kBufOut;
    cols.bytes = 16.0 * (double)(X * Y * B);
    StepIo ci;
    ci.count = B * 2 * Y;
    ci.cols = Y;
    ci.in_inner = 1;
    ci.in_outer = Y;
    ci.in_outer2 = X * Y;
    ci.in_stride = 2 * Y;
    ci.out_inner = 1;
    ci.out_outer = Y;
    ci.out_outer2 = X * Y;
    ci.out_stride = 2 * Y;
    ci.o_split = 2;
    if (!stockham_step(h, kVarColsC2C, true, ci, 0, inv, 0, &cols, &why)) return false;
    // by default only where the half-length columns have a specialised kernel too (512-point
    // columns on the run-time kernel made 1024 x 1024 slower: 121-123 against 105 us)
    if (pair2d_mode() != 1 && cols.shape.static_id < 0) return false;
    out->steps.push_back(std::move(rows));
    out->steps.push_back(std::move(cols));
    return true;
}

// 2D C2C in two launches with the column transforms' first radix-4 stage in the row pass:
// ASDFFT_QUAD2D=0 never, 1 wherever the geometry allows; unset, where both launches have
// specialised kernels (the half-plane exchange column kernel included).
int quad2d_mode()
{
    const char* e = getenv("ASDFFT_QUAD2D");
    if (e != nullptr && (e[0] == '0' || e[0] == '1')) return e[0] - '0';
    return -1;
}

bool quad2d(int64_t X, int64_t Y, int64_t B, bool inv, BuiltPlan* out)
{
    if (quad2d_mode() == 0 || X % 4 != 0 || X / 4 < 2 || Y < 2) return false;
    const int64_t q = X / 4;
    std::string why;
    // W_X^m, m < 3X/4 (the row pass multiplies bin k of row n by W_X^(n k)), forward sign
    std::vector<float> tab;
    static const long double kTwoPi = 6.283185307179586476925286766559005768L;
    for (int64_t m = 0; m < 3 * q; m++) {
        const long double ang = kTwoPi * (long double)m / (long double)X;
        tab.push_back((float)cosl(ang));
        tab.push_back((float)(-sinl(ang)));
    }
    TablePtr pt = upload_table("quad2d:" + std::to_string(X), tab);
    // The row pass reads rows n + q X/4 (q < 4) of the input and writes bin k of their 4-point
    // DFT, twiddled, to row 4n + k of the output.  Each column transform (k, column c) then
    // reads and rewrites exactly the elements (4i + k, c), i < X/4: in place on the output with
    // identical input and output addresses, no workspace.
    Step rows;
    rows.src = kBufIn;
    rows.dst = kBufOut;
    rows.bytes = 16.0 * (double)(X * Y * B);
    StepIo ri;
    ri.count = B * X;
    ri.cols = 4;
    ri.in_inner = q * Y;
    ri.in_outer = Y;
    ri.in_outer2 = X * Y;
    ri.out_inner = Y;
    ri.out_outer = 4 * Y;
    ri.out_outer2 = X * Y;
    ri.o_split = q;
    // the four-row workgroup (more threads than the run-time kernel's block) exists only as an instance
    if (!stockham_step(Y, kVarRowsQuad4, false, ri, inv, 0, 4, &rows, &why) || rows.shape.static_id < 0) return false;
    rows.args.tw_aux = pt ? pt->dev : nullptr;
    rows.tables_ready = rows.tables_ready && pt != nullptr;
    if (pt) rows.tables.push_back(pt);
    Step cols;
    cols.src = kBufOut;
    cols.dst = kBufOut;
    cols.bytes = 16.0 * (double)(X * Y * B);
    StepIo ci;
    ci.count = B * 4 * Y;
    ci.cols = Y;
    ci.in_inner = 1;
    ci.in_outer = Y;
    ci.in_outer2 = X * Y;
    ci.in_stride = 4 * Y;
    ci.out_inner = 1;
    ci.out_outer = Y;
    ci.out_outer2 = X * Y;
    ci.out_stride = 4 * Y;
    ci.o_split = 4;
    // the quarter-length columns through one LDS plane: 16 columns (128-byte runs) in half the
    // LDS of the two-plane kernel, so twice the workgroups fit a CU
    ci.prefer_split = true;
    if (!stockham_step(q, kVarColsC2C, true, ci, 0, inv, 0, &cols, &why)) return false;
    if (quad2d_mode() != 1 && cols.shape.variant != kVarColsC2CSplit) return false;
    out->steps.push_back(std::move(rows));
    out->steps.push_back(std::move(cols));
    return true;
}

}  // namespace

static std::vector<int> search_radices(int64_t n)
{
    if (n <= 1) return {};
    std::vector<int> e(kNumPrimes, 0);
    int64_t m = n;
    for (int i = 0; i < kNumPrimes; i++) {
        while (m % kPrimes[i] == 0) {
            m /= kPrimes[i];
            e[i]++;
        }
    }
    if (m != 1) return {};
    std::map<std::vector<int>, Best> memo;
    Best b = search(e, memo);
    if (b.passes >= (1 << 30)) return {};
    std::vector<int> r = b.radices;
    std::sort(r.begin(), r.end(), [](int x, int y) { return x > y; });
    if ((int)r.size() > kMaxPasses) return {};
    return r;
}

// The search above, once per length for the process (the serving loop asks at every
// submission; the search itself takes microseconds to tens of microseconds).
std::vector<int> choose_radices(int64_t n)
{
    static std::mutex mu;
    static auto* cache = new std::unordered_map<int64_t, std::vector<int>>();
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache->find(n);
        if (it != cache->end()) return it->second;
    }
    std::vector<int> r = search_radices(n);
    std::lock_guard<std::mutex> lock(mu);
    if (cache->size() < 65536) (*cache)[n] = r;
    return r;
}

std::vector<std::vector<int>> radix_multisets(int64_t n, int extra, int limit)
{
    const std::vector<int> best = choose_radices(n);
    if (best.empty()) return {};
    const int max_passes = std::min<int>(kMaxPasses, (int)best.size() + extra);
    std::vector<const RadixInfo*> order;
    for (const RadixInfo& r : kRadices) order.push_back(&r);
    std::sort(order.begin(), order.end(), [](const RadixInfo* a, const RadixInfo* b) { return a->radix > b->radix; });
    struct Cand {
        std::vector<int> radices;
        double cost;
        int maxr;
    };
    std::vector<Cand> found;
    std::vector<int> cur;
    // non-increasing radix sequences: each factor is taken from position `start` on
    std::function<void(int64_t, size_t, double)> dfs = [&](int64_t m, size_t start, double cost) {
        if (m == 1) {
            found.push_back({cur, cost, cur.empty() ? 0 : cur.front()});
            return;
        }
        if ((int)cur.size() >= max_passes) return;
        for (size_t i = start; i < order.size(); i++) {
            const int r = order[i]->radix;
            if (m % r != 0) continue;
            cur.push_back(r);
            dfs(m / r, i, cost + order[i]->cost);
            cur.pop_back();
        }
    };
    dfs(n, 0, 0.0);
    std::sort(found.begin(), found.end(), [](const Cand& a, const Cand& b) {
        if (a.radices.size() != b.radices.size()) return a.radices.size() < b.radices.size();
        const bool a16 = a.maxr <= 16, b16 = b.maxr <= 16;
        if (a16 != b16) return a16;
        return a.cost < b.cost;
    });
    std::vector<std::vector<int>> out;
    for (const Cand& c : found) {
        if ((int)out.size() >= limit) break;
        out.push_back(c.radices);
    }
    return out;
}

int threads_for(int64_t n, const std::vector<int>& radices, int emax)
{
    int64_t t = 0;
    for (int r : radices) {
        if (r > emax) return 0;
        const int64_t b = emax / r;
        const int64_t need = (n / r + b - 1) / b;
        t = std::max(t, need);
    }
    if (t <= 0 || t > max_block_for(emax)) return 0;
    return (int)t;
}

namespace {

// Roll of an axis of length n for a shift mode: the forward shift (asdFftShift) moves element
// k to (k + ceil(n/2)) mod n, the inverse shift (asdIfftShift) to (k + floor(n/2)) mod n.
int shift_roll(int mode, int64_t n)
{
    if (mode == kShiftFft) return (int)((n + 1) / 2);
    if (mode == kShiftIfft) return (int)(n / 2);
    return 0;
}

// C2C with the shifts folded into the first and last pass (1D rows; 2D rows into the workspace,
// then columns into the output: the output column roll moves data between workgroups, so the
// column pass cannot run in place).
std::string build_shift_plan(const PlanSpec& spec, BuiltPlan* out)
{
    std::string why;
    if (spec.kind != Kind::C2C || spec.planar || spec.vertical) return "shifts are folded into C2C plans only";
    const size_t nd = spec.dims.size();
    if (nd != 1 && nd != 2) return "shifts are folded into 1D and 2D plans only";
    for (int64_t d : spec.dims) {
        if (d < 2 || !smooth_single(d)) return "shifts are folded only for axis lengths 2..16384 with prime factors <= 19";
    }
    const int64_t B = spec.batch;
    const bool inv = spec.inverse;
    if (nd == 1) {
        const int64_t n = spec.dims[0];
        Step st;
        st.src = kBufIn;
        st.dst = kBufOut;
        st.bytes = 16.0 * (double)n * (double)B;
        StepIo io;
        io.count = B;
        io.cols = 1;
        io.in_outer = n;
        io.out_outer = n;
        if (!stockham_step(n, kVarRowsShiftBoth, false, io, inv, inv, 0, &st, &why)) return why;
        st.args.roll_in = shift_roll(spec.in_shift, n);
        st.args.roll_out = shift_roll(spec.out_shift, n);
        st.args.out_scale = spec.out_scale;
        st.describe += " (shift in " + std::to_string(st.args.roll_in) + " out " + std::to_string(st.args.roll_out) + ")";
        out->steps.push_back(std::move(st));
        return "";
    }
    const int64_t X = spec.dims[0], Y = spec.dims[1];
    out->workspace = (size_t)(B * X * Y * 8);
    Step r;
    r.src = kBufIn;
    r.dst = kBufWork;
    r.bytes = 16.0 * (double)X * (double)Y * (double)B;
    StepIo ir;
    ir.count = B * X;
    ir.cols = 1;
    ir.in_outer = Y;
    ir.out_outer = Y;
    if (!stockham_step(Y, kVarRowsShiftIn, false, ir, inv, inv, 0, &r, &why)) return why;
    r.args.roll_in = shift_roll(spec.in_shift, Y);
    r.args.in_o_mod = X;
    r.args.in_o_roll = shift_roll(spec.in_shift, X);
    r.describe += " (shift in " + std::to_string(r.args.in_o_roll) + "," + std::to_string(r.args.roll_in) + ")";
    Step c;
    c.src = kBufWork;
    c.dst = kBufOut;
    c.bytes = r.bytes;
    StepIo ic;
    ic.count = B * Y;
    ic.cols = Y;
    ic.in_inner = 1;
    ic.in_outer = X * Y;
    ic.in_stride = Y;
    ic.out_inner = 1;
    ic.out_outer = X * Y;
    ic.out_stride = Y;
    if (!stockham_step(X, kVarColsShiftOut, true, ic, inv, inv, 0, &c, &why)) return why;
    c.args.roll_out = shift_roll(spec.out_shift, X);
    c.args.out_col_roll = shift_roll(spec.out_shift, Y);
    c.args.out_scale = spec.out_scale;
    c.describe += " (shift out " + std::to_string(c.args.roll_out) + "," + std::to_string(c.args.out_col_roll) + ")";
    out->steps.push_back(std::move(r));
    out->steps.push_back(std::move(c));
    return "";
}

}  // namespace

static std::string build_plan_steps(const PlanSpec& spec, BuiltPlan* out)
{
    out->steps.clear();
    out->workspace = 0;
    std::string why;
    const bool inv = spec.inverse;
    const int64_t B = spec.batch;
    const size_t nd = spec.dims.size();
    if (nd < 1 || nd > 3) return "unsupported rank";

    if (spec.planar && spec.kind != Kind::C2C) return "separated real/imaginary planes are C2C only";
    if (spec.in_shift != kShiftNone || spec.out_shift != kShiftNone || spec.out_scale != 1.0f) {
        return build_shift_plan(spec, out);
    }
    if (nd == 1) {
        const int64_t n = spec.dims[0];
        if (spec.planar) {
            // (batch, n) real plane and imaginary plane; a VERTICAL plan of batch 1 is the same
            if (spec.vertical && B > 1) return "separated planes with a VERTICAL batch are not supported";
            if (n > kMaxSingleKernel && !choose_radices(n).empty()) {
                out->workspace = (size_t)(B * n * 8);
                if (!c2c_fourstep(n, B, inv, kBufIn, kBufOut, 0, out, &why, true)) return why;
                return "";
            }
            if (!c2c_rows(n, B, inv, kBufIn, kBufOut, out, &why, true, true)) return why;
            return "";
        }
        if (spec.vertical) {
            // tensor (n, batch): transform along axis 0, `batch` adjacent columns
            if (spec.kind != Kind::C2C) return "VERTICAL layout is supported for C2C only";
            if (B == 1) {
                if (!c2c_rows(n, 1, inv, kBufIn, kBufOut, out, &why)) return why;
                return "";
            }
            if (!c2c_cols(n, B, B, 0, B, inv, kBufIn, kBufOut, out, &why)) return why;
            return "";
        }
        if (spec.kind == Kind::C2C && spec.precision == Precision::kSplit16 && n == 1024) {
            mfma_rows_1024(B, inv, out);
            return "";
        }
        if (spec.kind == Kind::C2C && spec.precision == Precision::kSplit16 && mfma_fourstep_fits(n)) {
            out->workspace = (size_t)(B * n * 8);
            if (!mfma_fourstep(n, B, inv, out, &why)) return why;
            return "";
        }
        if (spec.kind == Kind::C2C) {
            if (n > kMaxSingleKernel && !choose_radices(n).empty()) {
                out->workspace = (size_t)(B * n * 8);
                if (!c2c_fourstep(n, B, inv, kBufIn, kBufOut, 0, out, &why)) return why;
                return "";
            }
            if (!c2c_rows(n, B, inv, kBufIn, kBufOut, out, &why)) return why;
            return "";
        }
        if (real_needs_large(n)) {
            if (!real_large(spec.kind, n, B, inv, out, &why)) return why;
            return "";
        }
        if (!real_rows(spec.kind, n, B, inv, kBufIn, kBufOut, out, &why)) return why;
        return "";
    }

    // multi-dimensional: dims = (X, Y) or (X, Y, Z); last axis contiguous
    const int64_t L = spec.dims[nd - 1];
    const int64_t W = (spec.kind == Kind::C2C) ? L : L / 2 + 1;  // complex width of the last axis
    int64_t lead = 1;
    for (size_t d = 0; d + 1 < nd; d++) lead *= spec.dims[d];
    const int64_t rows = B * lead;

    // column passes over the complex array [B][dims[0..nd-2]][W], axis d (d < nd-1)
    auto col_pass = [&](size_t d, int src, int dst, bool pout = false) -> bool {
        const int64_t len = spec.dims[d];
        int64_t inner = W;  // elements between consecutive points along axis d
        for (size_t e = d + 1; e + 1 < nd; e++) inner *= spec.dims[e];
        const int64_t outer = inner * len;
        const int64_t count = B * lead / len * W;
        if (len == 1) {
            if (src == dst) return true;
        }
        return c2c_cols(len, count, inner, outer, inner, inv, src, dst, out, &why, pout);
    };

    if (spec.planar) {
        // rows read the planes into the complex workspace; the last non-trivial column
        // pass writes the planes (a plan whose leading axes all have length 1 is rows only)
        if (L > kMaxSingleKernel) return "last-axis length above " + std::to_string(kMaxSingleKernel);
        std::vector<size_t> axes;
        for (size_t d = nd - 1; d-- > 0;) {
            if (spec.dims[d] > 1) axes.push_back(d);
        }
        if (axes.empty()) {
            if (!c2c_rows(L, rows, inv, kBufIn, kBufOut, out, &why, true, true)) return why;
            return "";
        }
        out->workspace = (size_t)(rows * L * 8);
        if (!c2c_rows(L, rows, inv, kBufIn, kBufWork, out, &why, true, false)) return why;
        for (size_t i = 0; i < axes.size(); i++) {
            const bool last = (i + 1 == axes.size());
            if (!col_pass(axes[i], kBufWork, last ? kBufOut : kBufWork, last)) return why;
        }
        return "";
    }

    if (spec.kind == Kind::C2C) {
        if (L > kMaxSingleKernel) return "last-axis length above " + std::to_string(kMaxSingleKernel);
        if (nd == 2 && quad2d(spec.dims[0], L, B, inv, out)) return "";
        if (nd == 2 && pair2d_wanted(spec.dims[0], L) && pair2d(spec.dims[0], L, B, inv, out)) return "";
        if (!c2c_rows(L, rows, inv, kBufIn, kBufOut, out, &why)) return why;
        for (size_t d = nd - 1; d-- > 0;) {
            if (!col_pass(d, kBufOut, kBufOut)) return why;
        }
        return "";
    }
    if (L / 2 > kMaxSingleKernel) return "last-axis length above " + std::to_string(2 * kMaxSingleKernel);
    if (spec.kind == Kind::R2C) {
        if (!real_rows(Kind::R2C, L, rows, inv, kBufIn, kBufOut, out, &why)) return why;
        for (size_t d = nd - 1; d-- > 0;) {
            if (!col_pass(d, kBufOut, kBufOut)) return why;
        }
        return "";
    }
    // C2R: complex column passes into the workspace, then C2R rows into the output
    out->workspace = (size_t)(B * lead * W * 8);
    bool first = true;
    for (size_t d = nd - 1; d-- > 0;) {
        const int src = first ? kBufIn : kBufWork;
        if (spec.dims[d] == 1 && first) {
            // identity axis: still has to move the data into the workspace
            if (!c2c_rows(1, B * lead * W, inv, src, kBufWork, out, &why)) return why;
        } else if (!col_pass(d, src, kBufWork)) {
            return why;
        }
        first = false;
    }
    if (!real_rows(Kind::C2R, L, rows, inv, kBufWork, kBufOut, out, &why)) return why;
    return "";
}

// The first launch of a plan whose later launch reads back what that first launch wrote (2D and
// 3D transforms, the four-step) loads its input non-temporally when that input and the
// intermediate together are at least the 256 MiB device cache: the input is read once, and loads
// that allocate in the cache would displace the intermediate the next pass reads.  Measured on
// whole transforms over rings of buffers larger than the cache: 2048^2 x 4 C2C 105.7 against
// 111.6 us, 1024^2 x 16 93.0 against 100.0, 4096^2 119.8 against 129.8, 256^3 131.6 against
// 136.7.  Smaller plans keep ordinary loads: there a call repeated on the same buffers can find
// all of them in the cache, and streaming gave that up (128^3 x 4 C2C repeated on one buffer
// pair: 78.3 against 73.2 us), although on a ring the four-step 65536 x 64 gained (34.6 against
// 37.2 us).  ASDFFT_STREAM_IN=0 / =1 turns it off / on wherever the first kernel has the path.
static void mark_streamed_input(BuiltPlan* plan)
{
    const char* e = getenv("ASDFFT_STREAM_IN");
    if (e != nullptr && e[0] == '0') return;
    const bool forced = (e != nullptr && e[0] == '1');
    if (plan->steps.size() < 2) return;
    Step& first = plan->steps[0];
    if (first.direct || first.mfma || first.aux || first.wave || first.planar_in || first.src != kBufIn) return;
    // kernels with a streamed path: row passes that read interleaved complex points, and the
    // four-step's first (column) pass where each workgroup reads whole 128-byte lines (16
    // columns or more, rows 128-byte aligned): a non-temporal load is not kept in L2 for a
    // neighbouring workgroup that reads the rest of a shared line (ISTFT frames read from a
    // spectrum of 500 frames per row measured 97.7 against 85.3 us streamed)
    switch (first.shape.variant) {
        case kVarRowsC2C:
        case kVarR2CEven:
        case kVarRowsC2CSplit:
        case kVarRowsPair2:
        case kVarRowsQuad4:
        case kVarRowsQuad4Split:
            break;
        case kVarColsTwiddle:
            if (first.args.tpb < 16 || first.args.in_stride % 16 != 0) return;
            break;
        default:
            return;
    }
    bool read_back = false;
    for (size_t i = 1; i < plan->steps.size(); i++) read_back |= (plan->steps[i].src == first.dst);
    if (!read_back) return;
    constexpr double kDeviceCacheBytes = 256.0 * 1024 * 1024;
    if (!forced && first.bytes < kDeviceCacheBytes) return;
    first.args.stream_in = 1;
    first.describe += " streamed-input";
}

std::string build_plan(const PlanSpec& spec, BuiltPlan* out)
{
    const std::string why = build_plan_steps(spec, out);
    if (why.empty()) mark_streamed_input(out);
    return why;
}

std::string build_istft_frames(int64_t nfft, int64_t channel, int64_t nframes, BuiltPlan* out)
{
    out->steps.clear();
    out->workspace = 0;
    if (nfft <= 0 || channel <= 0 || nframes <= 0) return "empty ISTFT input";
    const int64_t count = channel * nframes;
    std::string why;
    Step st;
    st.src = kBufIn;
    st.dst = kBufWork;
    st.bytes = 16.0 * (double)nfft * (double)count;
    if (smooth_single(nfft)) {
        // transform t = (b, frame): column `frame` of channel b's (nfft, nframes) block
        StepIo io;
        io.count = count;
        io.cols = nframes;
        io.in_inner = 1;
        io.in_outer = nfft * nframes;
        io.in_stride = nframes;
        io.out_inner = nfft;
        io.out_outer = nfft * nframes;
        io.out_stride = 1;
        if (!stockham_step(nfft, kVarColsToRows, true, io, 1, 1, 0, &st, &why)) return why;
    } else {
        if (!direct_step(nfft, kProComplex, kEpiComplex, count, nfft * nframes, nfft, 1, 1, &st, &why)) return why;
        st.dargs.in_cols = nframes;
        st.dargs.in_stride = nframes;
    }
    st.describe += " (ISTFT frames)";
    out->steps.push_back(std::move(st));
    out->workspace = (size_t)(count * nfft * 8);
    return "";
}

hipError_t launch_step(const Step& s, char* const bufs[3], const float* in_im, float* out_im, hipStream_t stream)
{
    if (s.mfma) {
        MfmaFftArgs a = s.margs;
        a.in = reinterpret_cast<const float2*>(bufs[s.src] + s.src_offset);
        a.out = reinterpret_cast<float2*>(bufs[s.dst] + s.dst_offset);
        return launch_mfma_fft1024(a, s.mfma_grid, stream);
    }
    if (s.wave) {
        WaveFftArgs a = s.wargs;
        a.in = reinterpret_cast<const float2*>(bufs[s.src] + s.src_offset);
        a.out = reinterpret_cast<float2*>(bufs[s.dst] + s.dst_offset);
        return launch_wave_fft(a, stream);
    }
    if (s.aux) {
        RealAuxArgs a = s.raux;
        a.in = bufs[s.src] + s.src_offset;
        a.out = bufs[s.dst] + s.dst_offset;
        return launch_real_aux(a, stream);
    }
    if (s.direct) {
        DftArgs a = s.dargs;
        a.in = bufs[s.src] + s.src_offset;
        a.out = bufs[s.dst] + s.dst_offset;
        a.in_im = s.planar_in ? in_im : nullptr;
        a.out_im = s.planar_out ? out_im : nullptr;
        return launch_dft_direct(s.shape, a, stream);
    }
    FftArgs a = s.args;
    a.in = bufs[s.src] + s.src_offset;
    a.out = bufs[s.dst] + s.dst_offset;
    a.in_im = s.planar_in ? in_im : nullptr;
    a.out_im = s.planar_out ? out_im : nullptr;
    return launch_stockham(s.shape, a, stream);
}

std::string describe_plan(const BuiltPlan& plan)
{
    std::ostringstream os;
    os << "plan: " << plan.steps.size() << " step(s), workspace=" << plan.workspace << " B\n";
    const char* names[] = {"in", "out", "work"};
    for (const Step& s : plan.steps) {
        os << "  [" << names[s.src] << (s.planar_in ? "(re,im)" : "") << "->" << names[s.dst]
           << (s.planar_out ? "(re,im)" : "") << "] " << s.describe << "\n";
    }
    return os.str();
}

}  // namespace asdfft
