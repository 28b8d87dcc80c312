"""ctypes binding of libasdfft.so (the C ABI declared in include/fft_api.h,
include/acl_meta.h and include/asdfft_ext.h).

The library must have been built (python -m sip_amd.build); there is no
Python or host fallback for any transform.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libasdfft.so")
TORCH_LIB_PATH = os.path.join(HERE, "lib", "libasdfft_torch.so")

# asdFftType / asdFftDirection / asdFft1dDimType (include/fft_api.h)
ASCEND_FFT_C2C = 0x10
ASCEND_FFT_C2R = 0x11
ASCEND_FFT_R2C = 0x12
ASCEND_STFT_C2C = 0x20
ASCEND_STFT_C2R = 0x21
ASCEND_STFT_R2C = 0x22
ASCEND_FFT_C2C_SEP = 0x30
ASCEND_FFT_FORWARD = 0x10
ASCEND_FFT_INVERSE = 0x11
ASCEND_FFT_HORIZONTAL = 0x10
ASCEND_FFT_VERTICAL = 0x11

# AspbStatus (include/aspb_status.h)
ACL_SUCCESS = 0
ACL_ERROR_INVALID_PARAM = 100000
ACL_ERROR_OP_INPUT_NOT_MATCH = 100021
ACL_ERROR_OP_OUTPUT_NOT_MATCH = 100022
ACL_ERROR_UNSUPPORTED_DATA_TYPE = 100026
ACL_ERROR_FORMAT_NOT_MATCH = 100027
ACL_ERROR_API_NOT_SUPPORT = 200001
ACL_ERROR_INTERNAL_ERROR = 500000

# aclDataType / aclFormat (include/acl_meta.h)
ACL_FLOAT = 0
ACL_COMPLEX64 = 16
ACL_FORMAT_ND = 2

STATUS_NAMES = {
    ACL_SUCCESS: "ACL_SUCCESS",
    ACL_ERROR_INVALID_PARAM: "ACL_ERROR_INVALID_PARAM",
    ACL_ERROR_OP_INPUT_NOT_MATCH: "ACL_ERROR_OP_INPUT_NOT_MATCH",
    ACL_ERROR_OP_OUTPUT_NOT_MATCH: "ACL_ERROR_OP_OUTPUT_NOT_MATCH",
    ACL_ERROR_UNSUPPORTED_DATA_TYPE: "ACL_ERROR_UNSUPPORTED_DATA_TYPE",
    ACL_ERROR_FORMAT_NOT_MATCH: "ACL_ERROR_FORMAT_NOT_MATCH",
    ACL_ERROR_API_NOT_SUPPORT: "ACL_ERROR_API_NOT_SUPPORT",
    ACL_ERROR_INTERNAL_ERROR: "ACL_ERROR_INTERNAL_ERROR",
}


class NativeLibraryMissing(RuntimeError):
    pass


class FftError(RuntimeError):
    def __init__(self, fn: str, status: int):
        super().__init__(f"{fn} returned {status} ({STATUS_NAMES.get(status, 'unknown')})")
        self.status = status


_lib = None


class ServeConfig(ctypes.Structure):
    """asdFftExtServeConfig (include/asdfft_ext.h)."""
    _fields_ = [
        ("hbm_budget_bytes", ctypes.c_uint64),
        ("slots", ctypes.c_int32),
        ("max_batch_requests", ctypes.c_int32),
        ("batching", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class ServeRequest(ctypes.Structure):
    """asdFftExtServeRequest (include/asdfft_ext.h)."""
    _fields_ = [
        ("type", ctypes.c_int32),
        ("direction", ctypes.c_int32),
        ("rank", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("dims", ctypes.c_int64 * 2),
        ("batch", ctypes.c_int64),
        ("input", ctypes.c_void_p),
        ("output", ctypes.c_void_p),
    ]


class ServeCounters(ctypes.Structure):
    """asdFftExtServeCounters (include/asdfft_ext.h)."""
    _fields_ = [(name, ctypes.c_uint64) for name in
                ("submitted", "completed", "batches", "ragged_requests", "plain_requests", "peak_inflight_bytes",
                 "resident_bytes")]

ASDFFT_PRECISION_FP32 = 0
ASDFFT_PRECISION_SPLIT16 = 1
ASDFFT_SHIFT_NONE = 0
ASDFFT_SHIFT_FFT = 1
ASDFFT_SHIFT_IFFT = 2
ASDFFT_RAGGED_GRAPH = 0x1
ASDFFT_RAGGED_AUTOTUNE = 0x2
ASDFFT_RAGGED_PER_BUDGET = 0x4


class RaggedEntry(ctypes.Structure):
    """asdFftExtRaggedEntry (include/asdfft_ext.h)."""
    _fields_ = [
        ("type", ctypes.c_int32),
        ("direction", ctypes.c_int32),
        ("n", ctypes.c_int64),
        ("count", ctypes.c_int64),
        ("in_offset", ctypes.c_int64),
        ("out_offset", ctypes.c_int64),
    ]

# (name, restype, argtypes) of every exported entry point
_VP = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
SIGNATURES = [
    ("asdFftCreate", _I32, [ctypes.POINTER(_VP)]),
    ("asdFftSetStream", _I32, [_VP, _VP]),
    ("asdFftMakePlan1D", _I32, [_VP, _I64, ctypes.c_int, ctypes.c_int, _I64, ctypes.c_int]),
    ("asdFftMakePlan2D", _I32, [_VP, _I64, _I64, ctypes.c_int, ctypes.c_int, _I32]),
    ("asdFftMakePlan3D", _I32, [_VP, _I64, _I64, _I64, ctypes.c_int, ctypes.c_int, _I32]),
    ("asdFftGetWorkspaceSize", _I32, [_VP, ctypes.POINTER(ctypes.c_size_t)]),
    ("asdFftSetWorkspace", _I32, [_VP, _VP]),
    ("asdFftExecC2C", _I32, [_VP, _VP, _VP]),
    ("asdFftExecC2R", _I32, [_VP, _VP, _VP]),
    ("asdFftExecR2C", _I32, [_VP, _VP, _VP]),
    ("asdFftSynchronize", _I32, [_VP]),
    ("asdFftDestroy", _I32, [_VP]),
    ("asdFftGetType", _I32, [_VP, ctypes.POINTER(ctypes.c_int)]),
    ("asdFftIstftMakePlan", _I32, [_VP, _VP, _I64, _I64, _I64, ctypes.c_bool, ctypes.c_bool, ctypes.c_bool, _I64,
                                   ctypes.c_bool]),
    ("asdFftExecIstft", _I32, [_VP, _VP, _VP, _VP]),
    ("asdFftExecC2CSeparated", _I32, [_VP, _VP, _VP, _VP, _VP]),
    ("asdFftShift", _I32, [_I64, _I64, _I64, _I64, _VP, _VP, _VP, _VP]),
    ("asdIfftShift", _I32, [_I64, _I64, _I64, _I64, ctypes.c_double, _VP, _VP, _VP, _VP]),
    ("aclCreateTensor", _VP, [ctypes.POINTER(_I64), ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(_I64), _I64,
                              ctypes.c_int, ctypes.POINTER(_I64), ctypes.c_uint64, _VP]),
    ("aclDestroyTensor", _I32, [_VP]),
    ("aclGetViewShape", _I32, [_VP, ctypes.POINTER(ctypes.POINTER(_I64)), ctypes.POINTER(ctypes.c_uint64)]),
    ("aclFreeViewShape", None, [ctypes.POINTER(_I64)]),
    ("aclGetDataType", _I32, [_VP, ctypes.POINTER(ctypes.c_int)]),
    ("asdFftExtDescribePlan", _I32, [_VP, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    ("asdFftExtPlanSteps", _I32, [_VP, ctypes.POINTER(_I32), ctypes.POINTER(ctypes.c_double)]),
    ("asdFftExtExecStep", _I32, [_VP, _VP, _VP, _I32]),
    ("asdFftExtExecNoAlloc", _I32, [_VP, _VP, _VP]),
    ("asdFftExtMakePlanRagged", _I32, [_VP, ctypes.POINTER(RaggedEntry), _I64, ctypes.c_uint32]),
    ("asdFftExtExecRagged", _I32, [_VP, _VP, _VP]),
    ("asdFftExtRaggedInfo", _I32, [_VP, ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_double)]),
    ("asdFftExtSetPrecision", _I32, [_VP, _I32]),
    ("asdFftExtSetShift", _I32, [_VP, _I32, _I32, ctypes.c_double]),
    ("asdFftExtServeCreate", _I32, [ctypes.POINTER(_VP), ctypes.POINTER(ServeConfig)]),
    ("asdFftExtServeSubmit", _I32, [_VP, ctypes.POINTER(ServeRequest), ctypes.POINTER(ctypes.c_uint64)]),
    ("asdFftExtServeWait", _I32, [_VP, ctypes.c_uint64, _I64, ctypes.POINTER(ctypes.c_double)]),
    ("asdFftExtServeReplay", _I32, [_VP, ctypes.POINTER(ServeRequest), ctypes.POINTER(ctypes.c_double), _I64,
                                    ctypes.POINTER(ctypes.c_double)]),
    ("asdFftExtServeStats", _I32, [_VP, ctypes.POINTER(ServeCounters)]),
    ("asdFftExtServeDestroy", _I32, [_VP]),
]


def lib():
    """Load libasdfft.so; raises NativeLibraryMissing if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(f"{LIB_PATH} not built; run `python -m sip_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


_torch_ops = None


def torch_ops():
    """Load libasdfft_torch.so (the torch_sip operators) once and return torch.ops.torch_sip;
    raises NativeLibraryMissing if it has not been built."""
    global _torch_ops
    if _torch_ops is None:
        import torch
        if not os.path.exists(TORCH_LIB_PATH):
            raise NativeLibraryMissing(f"{TORCH_LIB_PATH} not built; run `python -m sip_amd.build`")
        lib()  # the C ABI library first (the operators link against it)
        torch.ops.load_library(TORCH_LIB_PATH)
        _torch_ops = torch.ops.torch_sip
    return _torch_ops


def check(fn: str, status: int) -> None:
    if status != ACL_SUCCESS:
        raise FftError(fn, status)


class AclTensor:
    """Owns one aclTensor descriptor over a device (or any) address."""

    def __init__(self, shape, dtype: int, data_ptr: int):
        L = lib()
        shape = [int(s) for s in shape]
        n = len(shape)
        dims = (_I64 * max(n, 1))(*shape)
        strides = [1] * n
        for i in range(n - 2, -1, -1):
            strides[i] = strides[i + 1] * shape[i + 1]
        st = (_I64 * max(n, 1))(*strides)
        self.ptr = L.aclCreateTensor(dims, n, dtype, st, 0, ACL_FORMAT_ND, dims, n, ctypes.c_void_p(data_ptr))
        if not self.ptr:
            raise ValueError(f"aclCreateTensor rejected shape {shape}")

    def close(self):
        if self.ptr:
            lib().aclDestroyTensor(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
