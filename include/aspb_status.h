/*
 * Status codes returned by every FFT entry point.
 * Same names and values as the reference's core/utils/include/utils/aspb_status.h:15-25
 * (AspbStatus = int32; 0 success, 100000 invalid parameter, 100021/100022 input/output
 * mismatch, 100026 unsupported data type, 100027 format mismatch, 200001 API not
 * supported, 500000 internal error).
 */
#ifndef ASPB_STATUS_H
#define ASPB_STATUS_H

#include <stdint.h>

#ifdef __cplusplus
namespace AsdSip {
using AspbStatus = int32_t;

enum ErrorType : int {
    ACL_SUCCESS = 0,
    ACL_ERROR_INVALID_PARAM = 100000,
    ACL_ERROR_OP_INPUT_NOT_MATCH = 100021,
    ACL_ERROR_OP_OUTPUT_NOT_MATCH = 100022,
    ACL_ERROR_UNSUPPORTED_DATA_TYPE = 100026,
    ACL_ERROR_FORMAT_NOT_MATCH = 100027,
    ACL_ERROR_API_NOT_SUPPORT = 200001,
    ACL_ERROR_INTERNAL_ERROR = 500000
};
}  // namespace AsdSip
#else
typedef int32_t AspbStatus;
enum {
    ACL_SUCCESS = 0,
    ACL_ERROR_INVALID_PARAM = 100000,
    ACL_ERROR_OP_INPUT_NOT_MATCH = 100021,
    ACL_ERROR_OP_OUTPUT_NOT_MATCH = 100022,
    ACL_ERROR_UNSUPPORTED_DATA_TYPE = 100026,
    ACL_ERROR_FORMAT_NOT_MATCH = 100027,
    ACL_ERROR_API_NOT_SUPPORT = 200001,
    ACL_ERROR_INTERNAL_ERROR = 500000
};
#endif

#endif
