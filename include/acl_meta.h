/*
 * Minimal tensor descriptor used by the FFT entry points.
 *
 * The reference takes CANN's opaque aclTensor (aclnn/acl_meta.h) and reads its
 * data type, view shape and storage address (core/fft/fft_api.cpp:1574-1587,
 * :1463-1472); callers build it with aclCreateTensor(viewDims, n, dtype,
 * strides, offset, ACL_FORMAT_ND, storageDims, n, data)
 * (tests/standalone/fft1d_optimized/bench.cpp:70-80,
 *  tests/ut/framework/test_util/util.cpp:107-115).  This library ships its own
 * descriptor with the same function names and argument meaning so that such
 * callers compile and link unchanged.  Only dense row-major (ND) views are
 * accepted by the FFT entry points; `data` is a HIP device pointer.
 */
#ifndef ASDFFT_ACL_META_H
#define ASDFFT_ACL_META_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aclTensor aclTensor;

typedef enum {
    ACL_DT_UNDEFINED = -1,
    ACL_FLOAT = 0,
    ACL_FLOAT16 = 1,
    ACL_INT8 = 2,
    ACL_INT32 = 3,
    ACL_UINT8 = 4,
    ACL_INT16 = 6,
    ACL_UINT16 = 7,
    ACL_UINT32 = 8,
    ACL_INT64 = 9,
    ACL_UINT64 = 10,
    ACL_DOUBLE = 11,
    ACL_BOOL = 12,
    ACL_COMPLEX64 = 16,
    ACL_COMPLEX128 = 17
} aclDataType;

typedef enum {
    ACL_FORMAT_UNDEFINED = -1,
    ACL_FORMAT_NCHW = 0,
    ACL_FORMAT_NHWC = 1,
    ACL_FORMAT_ND = 2
} aclFormat;

/* Returns NULL on invalid arguments (negative dims, NULL dims with n > 0, n > 16). */
aclTensor* aclCreateTensor(const int64_t* viewDims, uint64_t viewDimsNum, aclDataType dataType, const int64_t* stride,
                           int64_t offset, aclFormat format, const int64_t* storageDims, uint64_t storageDimsNum,
                           void* tensorData);

/* 0 on success. */
int32_t aclDestroyTensor(const aclTensor* tensor);

/* *viewDims is allocated with new[]; the caller releases it with delete[]
 * (C callers: aclFreeViewShape). */
int32_t aclGetViewShape(const aclTensor* tensor, int64_t** viewDims, uint64_t* viewDimsNum);

void aclFreeViewShape(int64_t* viewDims);

int32_t aclGetDataType(const aclTensor* tensor, aclDataType* dataType);

#ifdef __cplusplus
}
#endif

#endif
