// Serving loop: continuous batching of FFT requests on one GPU.
//
// Requests (1D C2C / R2C / C2R of any batch, 2D C2C / R2C / C2R) are admitted in
// arrival order while their input + output bytes fit the HBM budget.  Whenever one of
// the `slots` in-flight batches completes (or a slot is idle), everything admitted so
// far forms the next batch: its 1D requests whose lengths the ragged kernel handles run
// as one persistent ragged launch per register budget (class descriptors and twiddle
// tables stay resident on the device), the rest run through cached plain plans on the
// same stream.  Each slot has its own stream, so batches overlap on the device.
// With batching off, every request is its own launch group through the plain plan path
// (the per-request exec the reference's callers use).
//
// The reference has no serving loop; its nearest equivalent is the torch plugin's LRU
// plan cache plus one exec per call (sip_pta/csrc/FFTMixCache.cpp:18-40).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace asdfft {

struct ServeConfig {
    size_t hbm_budget = (size_t)8 << 30;  // admitted request bytes (inputs + outputs) in flight
    int slots = 3;                        // batches in flight, one stream each
    int max_batch = 256;                  // requests per batch
    bool batching = true;
};

struct ServeRequest {
    int type = 0x10;       // asdFftType value
    bool inverse = false;
    int rank = 1;          // 1 or 2
    int64_t dims[2] = {0, 0};
    int64_t batch = 1;
    const void* in = nullptr;
    void* out = nullptr;
};

struct ServeStats {
    uint64_t submitted = 0, completed = 0, batches = 0;
    uint64_t ragged_requests = 0, plain_requests = 0;
    uint64_t peak_inflight_bytes = 0, resident_bytes = 0;
};

class Server;

// Status codes of the functions below: 0 success, 1 invalid argument, 2 unsupported,
// 3 device error, 4 timeout.
Server* serve_create(const ServeConfig& cfg, std::string* why);
void serve_destroy(Server* s);
int serve_submit(Server* s, const ServeRequest& r, uint64_t* ticket, std::string* why);
// Waits for a ticket; on completion the ticket is retired and latency_us receives the
// time from submission to the completion of its batch.
int serve_wait(Server* s, uint64_t ticket, int64_t timeout_us, double* latency_us);
// Submits reqs[i] at arrival_us[i] microseconds after the call (spinning in the calling
// thread), waits for all, and stores each latency measured from its scheduled arrival.
int serve_replay(Server* s, const ServeRequest* reqs, const double* arrival_us, int64_t n, double* latency_us);
void serve_stats(Server* s, ServeStats* out);

}  // namespace asdfft
