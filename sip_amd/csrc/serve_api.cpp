// C ABI of the serving loop (asdfft_ext.h, asdFftExtServe*).
#include <stdio.h>
#include <stdlib.h>

#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "asdfft_ext.h"
#include "aspb_status.h"
#include "fft_api.h"
#include "server.h"

namespace {

using namespace AsdSip;

std::mutex g_mu;
std::set<void*>& live()
{
    static auto* s = new std::set<void*>();
    return *s;
}

asdfft::Server* find(void* h)
{
    std::lock_guard<std::mutex> lock(g_mu);
    return live().count(h) ? static_cast<asdfft::Server*>(h) : nullptr;
}

void log_error(const char* fn, const std::string& msg)
{
    const char* e = getenv("ASDFFT_QUIET");
    if (e == nullptr || e[0] != '1') fprintf(stderr, "[asdfft] %s: %s\n", fn, msg.c_str());
}

int32_t status_of(int code)
{
    switch (code) {
        case 0: return ACL_SUCCESS;
        case 1: return ACL_ERROR_INVALID_PARAM;
        case 2: return ACL_ERROR_API_NOT_SUPPORT;
        default: return ACL_ERROR_INTERNAL_ERROR;
    }
}

asdfft::ServeRequest to_request(const asdFftExtServeRequest& r)
{
    asdfft::ServeRequest q;
    q.type = r.type;
    q.inverse = (r.direction == ASCEND_FFT_INVERSE);
    q.rank = r.rank;
    q.dims[0] = r.dims[0];
    q.dims[1] = r.dims[1];
    q.batch = r.batch;
    q.in = r.in;
    q.out = r.out;
    return q;
}

bool valid_direction(int32_t d) { return d == ASCEND_FFT_FORWARD || d == ASCEND_FFT_INVERSE; }

}  // namespace

extern "C" int32_t asdFftExtServeCreate(void** server, const asdFftExtServeConfig* config)
{
    const char* fn = "asdFftExtServeCreate";
    if (server == nullptr) return ACL_ERROR_INVALID_PARAM;
    *server = nullptr;
    asdfft::ServeConfig cfg;
    if (config != nullptr) {
        if (config->hbm_budget_bytes) cfg.hbm_budget = (size_t)config->hbm_budget_bytes;
        if (config->slots) cfg.slots = config->slots;
        if (config->max_batch_requests) cfg.max_batch = config->max_batch_requests;
        cfg.batching = config->batching != 0;
    }
    std::string why;
    asdfft::Server* s = asdfft::serve_create(cfg, &why);
    if (s == nullptr) {
        log_error(fn, why);
        return why == "invalid server configuration" ? ACL_ERROR_INVALID_PARAM : ACL_ERROR_INTERNAL_ERROR;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    live().insert(s);
    *server = s;
    return ACL_SUCCESS;
}

extern "C" int32_t asdFftExtServeSubmit(void* server, const asdFftExtServeRequest* request, uint64_t* ticket)
{
    const char* fn = "asdFftExtServeSubmit";
    asdfft::Server* s = find(server);
    if (s == nullptr || request == nullptr || ticket == nullptr) {
        log_error(fn, "invalid server, request or ticket pointer");
        return ACL_ERROR_INVALID_PARAM;
    }
    if (!valid_direction(request->direction)) {
        log_error(fn, "invalid direction");
        return ACL_ERROR_INVALID_PARAM;
    }
    std::string why;
    const int st = asdfft::serve_submit(s, to_request(*request), ticket, &why);
    if (st != 0) log_error(fn, why);
    return status_of(st);
}

extern "C" int32_t asdFftExtServeWait(void* server, uint64_t ticket, int64_t timeout_us, double* latency_us)
{
    asdfft::Server* s = find(server);
    if (s == nullptr) return ACL_ERROR_INVALID_PARAM;
    const int st = asdfft::serve_wait(s, ticket, timeout_us, latency_us);
    if (st == 4) return ACL_ERROR_INTERNAL_ERROR;
    return status_of(st);
}

extern "C" int32_t asdFftExtServeReplay(void* server, const asdFftExtServeRequest* requests, const double* arrival_us,
                                        int64_t count, double* latency_us)
{
    const char* fn = "asdFftExtServeReplay";
    asdfft::Server* s = find(server);
    if (s == nullptr || requests == nullptr || arrival_us == nullptr || latency_us == nullptr || count < 0) {
        log_error(fn, "invalid arguments");
        return ACL_ERROR_INVALID_PARAM;
    }
    std::vector<asdfft::ServeRequest> list;
    list.reserve((size_t)count);
    for (int64_t i = 0; i < count; i++) {
        if (!valid_direction(requests[i].direction)) {
            log_error(fn, "request " + std::to_string(i) + ": invalid direction");
            return ACL_ERROR_INVALID_PARAM;
        }
        list.push_back(to_request(requests[i]));
    }
    return status_of(asdfft::serve_replay(s, list.data(), arrival_us, count, latency_us));
}

extern "C" int32_t asdFftExtServeStats(void* server, asdFftExtServeCounters* stats)
{
    asdfft::Server* s = find(server);
    if (s == nullptr || stats == nullptr) return ACL_ERROR_INVALID_PARAM;
    asdfft::ServeStats st;
    asdfft::serve_stats(s, &st);
    stats->submitted = st.submitted;
    stats->completed = st.completed;
    stats->batches = st.batches;
    stats->ragged_requests = st.ragged_requests;
    stats->plain_requests = st.plain_requests;
    stats->peak_inflight_bytes = st.peak_inflight_bytes;
    stats->resident_bytes = st.resident_bytes;
    return ACL_SUCCESS;
}

extern "C" int32_t asdFftExtServeDestroy(void* server)
{
    asdfft::Server* s = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (!live().count(server)) return ACL_ERROR_INVALID_PARAM;
        live().erase(server);
        s = static_cast<asdfft::Server*>(server);
    }
    asdfft::serve_destroy(s);
    return ACL_SUCCESS;
}
