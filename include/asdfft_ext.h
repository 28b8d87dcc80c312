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

#ifdef __cplusplus
}
#endif

#endif
