/*
 * asdfft_ext.h -- extension entry points of libasdfft.so.
 *
 * Status values are AspbStatus codes (aspb_status.h).
 * These have no counterpart in the reference's include/fft_api.h.  They serve
 * tooling that needs to see what a plan launches (traces, benchmarks, tests).
 */
#ifndef ASDFFT_EXT_H
#define ASDFFT_EXT_H

#include <stddef.h>
#include <stdint.h>

#include "aspb_status.h"

#ifdef __cplusplus
extern "C" {
#endif

/* No reference counterpart (introspection).  Writes a text description of the
 * plan's kernel launches into buffer (NUL-terminated, truncated to length) and
 * the full size including the NUL into *needed.  The same text is printed to
 * stderr at plan time when ASDFFT_PLAN_TRACE=1 (the reference's
 * ASDSIP_FFT_PLAN_TRACE plays that role, core/fft/fft_api.cpp:465-469). */
int32_t asdFftExtDescribePlan(void* handle, char* buffer, size_t length, size_t* needed);

/* No reference counterpart (introspection).  *steps = number of kernel launches
 * of one exec; if algorithmic_bytes is not NULL it receives, per launch, the
 * bytes of data that launch transforms (one read + one write). */
int32_t asdFftExtPlanSteps(void* handle, int32_t* steps, double* algorithmic_bytes);

/* No reference counterpart (profiling).  Runs only kernel launch `step` (0-based) of
 * the plan's exec with the same validation as asdFftExecC2C / C2R / R2C, so a
 * profiler can time each launch of a multi-launch plan on its own. */
typedef struct aclTensor aclTensor;
int32_t asdFftExtExecStep(void* handle, const aclTensor* input, const aclTensor* output, int32_t step);

/* No reference counterpart.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): runs a C2C, R2C or
 * C2R plan exactly as asdFftExecC2C / R2C / C2R do (same checks and status codes, the plan
 * type chooses which), except that the plan's workspace is a device array inside the library
 * (64 MiB, ASDFFT_NOALLOC_SCRATCH_BYTES, one per GPU, held until the process ends) instead of
 * the buffer bound with asdFftSetWorkspace, so neither GetWorkspaceSize nor SetWorkspace is
 * needed and the call allocates nothing.  A plan whose workspace exceeds the array is refused
 * with ACL_ERROR_INVALID_PARAM.  Every call on one GPU uses that GPU's one array: calls that
 * overlap on a GPU (other streams, other threads) race; serialise them.  Nothing in the array
 * needs to start at zero (each call's first pass writes the parts its second pass reads). */
#define ASDFFT_NOALLOC_SCRATCH_BYTES (64ull << 20)
int32_t asdFftExtExecNoAlloc(void* handle, const aclTensor* input, const aclTensor* output);

/* ---- ragged batches -------------------------------------------------------------------
 * One exec transforms entries of different lengths, types and directions.  Entry e holds
 * `count` transforms stored back to back from in_base + in_offset (inputs) and
 * out_base + out_offset (outputs), each laid out as a plain 1D transform of its type
 * (C2C: n complex64 in and out; R2C: n float32 in, n/2+1 complex64 out; C2R: the
 * reverse).  No normalisation, the same sign convention as asdFftExecC2C / R2C / C2R.
 * The reference has no counterpart: there, each distinct (type, length) needs its own
 * plan and exec (core/fft/fft_api.cpp:1124-1190, 1752-1844). */
typedef struct asdFftExtRaggedEntry {
    int32_t type;        /* asdFftType value: C2C 0x10, C2R 0x11, R2C 0x12 */
    int32_t direction;   /* asdFftDirection value: FORWARD 0x10, INVERSE 0x11 */
    int64_t n;           /* transform length (the real length for R2C / C2R) */
    int64_t count;       /* transforms in this entry */
    int64_t in_offset;   /* bytes from the exec's input base to the first input */
    int64_t out_offset;  /* bytes from the exec's output base to the first output */
} asdFftExtRaggedEntry;

#define ASDFFT_RAGGED_GRAPH 0x1u    /* exec replays an instantiated hipGraph of the launches */
#define ASDFFT_RAGGED_AUTOTUNE 0x2u /* time candidate schedules on the device while planning */
#define ASDFFT_RAGGED_PER_BUDGET 0x4u /* one concurrent launch per register budget (EMAX 16 / 32 /
                                       48) instead of the default single launch of every class at
                                       the largest budget (measured slower; kept for comparison) */

/* No reference counterpart (batching form).  Plans a ragged batch on a handle from
 * asdFftCreate.  Outputs of different entries must not overlap.  Returns
 * ACL_ERROR_INVALID_PARAM for invalid entries and ACL_ERROR_API_NOT_SUPPORT for lengths
 * outside the ragged kernels (complex pass length above 16384 or a prime factor above
 * 19).  With ASDFFT_RAGGED_AUTOTUNE the plan call runs short timing launches on the
 * current device (and allocates scratch for them). */
int32_t asdFftExtMakePlanRagged(void* handle, const asdFftExtRaggedEntry* entries, int64_t num_entries,
                                uint32_t flags);

/* No reference counterpart (batching form).  Enqueues the ragged batch on the handle's
 * stream (asdFftSetStream).  Inputs and outputs must not overlap, except that an
 * all-C2C batch whose entries have in_offset == out_offset may run in place with
 * in_base == out_base.  Execs of one plan share its tile queue: run them on one stream. */
int32_t asdFftExtExecRagged(void* handle, const void* in_base, void* out_base);

/* No reference counterpart (batching form).  Bytes past each base that an exec reads
 * and writes, the algorithmic bytes (inputs read once + outputs written once) and the
 * flop count 5 n log2 n summed over the transforms. */
int32_t asdFftExtRaggedInfo(void* handle, int64_t* in_extent, int64_t* out_extent, double* bytes, double* flops);

/* ---- precision form ------------------------------------------------------------------
 * ASDFFT_PRECISION_SPLIT16 runs large-radix stages as matrix products on FP16 MFMA with
 * split hi + lo operands (a b ~= a_hi b_hi + a_hi b_lo + a_lo b_hi) and FP32 accumulation,
 * after a per-row power-of-two scale.  Kernels exist for 1D C2C N = 1024 (two radix-32
 * stages) and for the 1024-point pass of long 1D C2C lengths N = 1024 * n2 (n2 a multiple of
 * 16, N > 16384, e.g. 65536: an FP32 n2-point column pass with the W_N twiddle, then the
 * 1024-point rows on MFMA with a transposed store); plans of other shapes keep the FP32
 * kernels.  Typical relative L2 error vs FP64 is a few 1e-7 (the FP32 kernels: about 1e-7). */
#define ASDFFT_PRECISION_FP32 0
#define ASDFFT_PRECISION_SPLIT16 1

/* No reference counterpart (precision form).  Chooses the arithmetic of the plan that the
 * next asdFftMakePlan1D / 2D / 3D call builds on this handle; ACL_ERROR_INVALID_PARAM for an
 * unknown mode or a handle whose plan already exists. */
int32_t asdFftExtSetPrecision(void* handle, int32_t mode);

/* ---- shifts folded into the transform ----------------------------------------------------
 * ASDFFT_SHIFT_FFT is the permutation of asdFftShift (every FFT axis of length n rolled by
 * ceil(n/2)), ASDFFT_SHIFT_IFFT that of asdIfftShift (rolled by floor(n/2)).  A plan made after
 * asdFftExtSetShift(h, in_shift, out_shift, scale) computes
 *     out = scale * out_shift( FFT( in_shift(in) ) )
 * with the input permutation applied while the first pass reads and the output permutation
 * and scale while the last pass writes: no extra pass over HBM (the separate asdIfftShift +
 * exec + asdFftShift sequence makes two).  For 1D and 2D C2C plans whose axis lengths are
 * 2..16384 with prime factors <= 19; MakePlan returns ACL_ERROR_API_NOT_SUPPORT for other
 * plans.  2D plans need a workspace of batch*X*Y*8 bytes; exec refuses in-place calls
 * (ACL_ERROR_INVALID_PARAM). */
#define ASDFFT_SHIFT_NONE 0
#define ASDFFT_SHIFT_FFT 1
#define ASDFFT_SHIFT_IFFT 2

/* No reference counterpart (fused shifts; callers that follow or precede a transform with
 * asdFftShift / asdIfftShift).  Chooses the shifts and output scale of the plan that the next
 * asdFftMakePlan1D / 2D call builds on this handle; ACL_ERROR_INVALID_PARAM for an unknown mode
 * or a handle whose plan already exists. */
int32_t asdFftExtSetShift(void* handle, int32_t in_shift, int32_t out_shift, double scale);

/* ---- serving loop ------------------------------------------------------------------------
 * A server admits FFT requests in arrival order while their input + output bytes (plus a
 * plan workspace) fit an HBM budget, and runs them in batches: whenever one of `slots`
 * in-flight batches completes, everything admitted meanwhile forms the next batch.  1D
 * requests of lengths the ragged kernel handles run as one persistent launch per batch
 * (class descriptors and twiddle tables stay resident); other requests (2D, longer 1D, odd
 * factors) run through cached plans on the batch's stream.  Each slot has its own stream,
 * so batches overlap on the device.  With batching = 0 every request is its own launch
 * group through the plain plan path.  Data pointers are device pointers the caller owns
 * until the request completes.  No reference counterpart: the reference's callers make
 * one plan and one exec per request (sip_pta/csrc/FFTMixCache.cpp:18-40). */
typedef struct asdFftExtServeConfig {
    uint64_t hbm_budget_bytes;   /* 0: 8 GiB */
    int32_t slots;               /* batches in flight, 1..8 (0: 3) */
    int32_t max_batch_requests;  /* requests per batch (0: 256) */
    int32_t batching;            /* 1: continuous batching, 0: one plain exec per request */
    int32_t reserved;
} asdFftExtServeConfig;

typedef struct asdFftExtServeRequest {
    int32_t type;       /* asdFftType value: C2C 0x10, C2R 0x11, R2C 0x12 */
    int32_t direction;  /* asdFftDirection value */
    int32_t rank;       /* 1 or 2 */
    int32_t reserved;
    int64_t dims[2];    /* n, or (X, Y); real length on the last axis for R2C / C2R */
    int64_t batch;
    const void* in;
    void* out;
} asdFftExtServeRequest;

typedef struct asdFftExtServeCounters {
    uint64_t submitted, completed, batches;
    uint64_t ragged_requests, plain_requests;
    uint64_t peak_inflight_bytes, resident_bytes;
} asdFftExtServeCounters;

/* No reference counterpart (serving form).  Creates a server on the current device. */
int32_t asdFftExtServeCreate(void** server, const asdFftExtServeConfig* config);
/* Queues a request; *ticket identifies it.  ACL_ERROR_INVALID_PARAM for malformed requests
 * or one larger than the budget, ACL_ERROR_API_NOT_SUPPORT for sizes no plan handles. */
int32_t asdFftExtServeSubmit(void* server, const asdFftExtServeRequest* request, uint64_t* ticket);
/* Waits for a ticket (timeout_us < 0: no limit) and retires it; *latency_us = submission to
 * completion.  ACL_ERROR_INVALID_PARAM for an unknown ticket, ACL_ERROR_INTERNAL_ERROR if
 * its batch failed or on timeout (the ticket then stays valid). */
int32_t asdFftExtServeWait(void* server, uint64_t ticket, int64_t timeout_us, double* latency_us);
/* Submits requests[i] at arrival_us[i] microseconds after the call from the calling thread,
 * waits for all of them and writes each latency measured from its scheduled arrival (-1
 * for a request that was refused). */
int32_t asdFftExtServeReplay(void* server, const asdFftExtServeRequest* requests, const double* arrival_us,
                             int64_t count, double* latency_us);
int32_t asdFftExtServeStats(void* server, asdFftExtServeCounters* stats);
/* Waits for every queued request, then releases the server. */
int32_t asdFftExtServeDestroy(void* server);

#ifdef __cplusplus
}
#endif

#endif
