"""C-ABI contract tests (CPU): exported symbols, plan validation, status codes,
plan-state rules, tensor matching and the workspace guard.

Negative cases mirror the reference's API-contract unit tests
(tests/ut/unittest/fft/1d/c2c/test_fft_c2c_1d.cpp:231-826,
 tests/ut/unittest/fft/fftworkspace/test_fft_workspace_negative.cpp:61-85,
 tests/ut/unittest/fft/2d/c2c/test_fft_c2c_2d.cpp:173-201).  None of them needs
a GPU: every rejection happens before a kernel would be launched.  Tensors
carry a dummy non-null address that is never dereferenced.
"""
import ctypes
import os
import re

import pytest

from sip_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FAKE_PTR = 0x1000

OK = N.ACL_SUCCESS
INVALID = N.ACL_ERROR_INVALID_PARAM
INTERNAL = N.ACL_ERROR_INTERNAL_ERROR
NOT_SUPPORTED = N.ACL_ERROR_API_NOT_SUPPORT
C2C, C2R, R2C = N.ASCEND_FFT_C2C, N.ASCEND_FFT_C2R, N.ASCEND_FFT_R2C
FWD, INV = N.ASCEND_FFT_FORWARD, N.ASCEND_FFT_INVERSE
H, V = N.ASCEND_FFT_HORIZONTAL, N.ASCEND_FFT_VERTICAL


@pytest.fixture(scope="module")
def L():
    return N.lib()


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


class Handle:
    def __init__(self, L):
        self.L = L
        self.h = ctypes.c_void_p()
        assert L.asdFftCreate(ctypes.byref(self.h)) == OK

    def __enter__(self):
        return self.h

    def __exit__(self, *exc):
        self.L.asdFftDestroy(self.h)


def tensor(shape, dtype, ptr=FAKE_PTR):
    return N.AclTensor(shape, dtype, ptr)


def declared_functions():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:AspbStatus|int32_t|aclTensor\*|void)\s+(\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(L):
    names = declared_functions()
    assert {"asdFftCreate", "asdFftExecC2C", "asdFftMakePlan3D", "aclCreateTensor"} <= names
    for n in sorted(names):
        assert hasattr(L, n), f"{n} declared in include/ but not exported"
    declared_in_binding = {s[0] for s in N.SIGNATURES}
    assert names <= declared_in_binding


def test_invalid_handle_codes(L):
    bogus = ctypes.c_void_p(0xdead)
    assert L.asdFftMakePlan1D(bogus, 64, C2C, FWD, 1, H) == INVALID
    assert L.asdFftMakePlan2D(bogus, 8, 8, C2C, FWD, 1) == INVALID
    assert L.asdFftSetStream(bogus, ctypes.c_void_p(1)) == INVALID
    assert L.asdFftSetWorkspace(bogus, None) == INVALID
    size = ctypes.c_size_t()
    assert L.asdFftGetWorkspaceSize(bogus, ctypes.byref(size)) == INVALID
    assert L.asdFftDestroy(bogus) == INVALID
    t = ctypes.c_int()
    assert L.asdFftSynchronize(bogus) == INTERNAL  # fft_api.cpp:1376-1390
    assert L.asdFftGetType(bogus, ctypes.byref(t)) == INTERNAL


def test_set_stream_null_rejected(L):
    with Handle(L) as h:
        assert L.asdFftSetStream(h, None) == INVALID


def test_make_plan_parameter_ranges(L):
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 0, C2C, FWD, 1, H) == INVALID
        assert L.asdFftMakePlan1D(h, (1 << 27) + 1, C2C, FWD, 1, H) == INVALID
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 0, H) == INVALID
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 2 ** 31, H) == INVALID
        assert L.asdFftMakePlan2D(h, 0, 8, C2C, FWD, 1) == INVALID
        assert L.asdFftMakePlan2D(h, 8, 8, C2C, FWD, 0) == INVALID
        assert L.asdFftMakePlan3D(h, 8, 8, 0, C2C, FWD, 1) == INVALID
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 1, H) == OK


def test_make_plan_2d_rejects_c2c_sep(L):
    with Handle(L) as h:
        assert L.asdFftMakePlan2D(h, 64, 64, N.ASCEND_FFT_C2C_SEP, FWD, 1) == INVALID


def test_repeated_init_rejected(L):
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 1, H) == OK
        assert L.asdFftMakePlan1D(h, 128, C2C, FWD, 1, H) == INVALID
        assert L.asdFftMakePlan2D(h, 64, 64, C2C, FWD, 1) == INVALID


def test_derived_batch_overflow_rejected(L):
    with Handle(L) as h:
        assert L.asdFftMakePlan2D(h, 64, 256, C2C, FWD, 67108864) == INVALID


def test_stft_and_unknown_types_rejected(L):
    with Handle(L) as h:
        for t in (N.ASCEND_STFT_C2C, N.ASCEND_STFT_C2R, N.ASCEND_STFT_R2C, 0x99):
            assert L.asdFftMakePlan1D(h, 64, t, FWD, 1, H) == INVALID
        # rejected calls leave the handle initialisable
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 1, H) == OK


def test_invalid_direction_rejected(L):
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, C2C, 0, 1, H) == INVALID
        assert L.asdFftMakePlan1D(h, 64, C2C, 0x99, 1, H) == INVALID
        assert L.asdFftMakePlan2D(h, 8, 8, C2C, 0, 1) == INVALID
        assert L.asdFftMakePlan3D(h, 8, 8, 8, C2C, 0, 1) == INVALID
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 1, H) == OK


def test_invalid_dim_type_rejected(L):
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 1, 0) == INVALID
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 1, 0x99) == INVALID
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 1, H) == OK


def test_get_type(L):
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 16, C2R, FWD, 1, H) == OK
        t = ctypes.c_int()
        assert L.asdFftGetType(h, ctypes.byref(t)) == OK
        assert t.value == C2R


def test_cross_variant_exec_rejected(L):
    a = tensor((1, 16), N.ACL_COMPLEX64)
    b = tensor((1, 16), N.ACL_COMPLEX64)
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 16, C2R, FWD, 1, H) == OK
        assert L.asdFftExecC2C(h, a.ptr, b.ptr) == INVALID
        assert L.asdFftExecR2C(h, a.ptr, b.ptr) == INVALID
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 16, C2C, FWD, 1, H) == OK
        assert L.asdFftExecC2R(h, a.ptr, b.ptr) == INVALID


def test_exec_before_plan_rejected(L):
    a = tensor((1, 16), N.ACL_COMPLEX64)
    with Handle(L) as h:
        assert L.asdFftExecC2C(h, a.ptr, a.ptr) == INVALID


@pytest.mark.parametrize("plan_type,in_shape,out_shape", [
    (C2R, (1, 33), (8, 64)), (C2R, (8, 33), (1, 64)), (R2C, (1, 64), (8, 33)), (R2C, (8, 64), (1, 33)),
    (C2R, (33,), (8, 64)), (R2C, (64,), (8, 33)),
])
def test_real_exec_batch_mismatch_rejected(L, plan_type, in_shape, out_shape):
    if plan_type == C2R:
        i, o = tensor(in_shape, N.ACL_COMPLEX64), tensor(out_shape, N.ACL_FLOAT)
    else:
        i, o = tensor(in_shape, N.ACL_FLOAT), tensor(out_shape, N.ACL_COMPLEX64)
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, plan_type, FWD, 8, H) == OK
        fn = L.asdFftExecC2R if plan_type == C2R else L.asdFftExecR2C
        assert fn(h, i.ptr, o.ptr) == INVALID


@pytest.mark.parametrize("in_shape,out_shape,dim_type,batch", [
    ((1, 256), (128, 256), H, 128),
    ((7, 128, 256), (7, 128, 256), H, 128),
    ((128, 256), (128, 256), V, 128),  # VERTICAL plans take (N, batch)
])
def test_c2c_shape_mismatch_rejected(L, in_shape, out_shape, dim_type, batch):
    i, o = tensor(in_shape, N.ACL_COMPLEX64), tensor(out_shape, N.ACL_COMPLEX64)
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 256, C2C, FWD, batch, dim_type) == OK
        assert L.asdFftExecC2C(h, i.ptr, o.ptr) == INVALID


def test_dtype_mismatch_rejected(L):
    i, o = tensor((4, 64), N.ACL_FLOAT), tensor((4, 64), N.ACL_COMPLEX64)
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 4, H) == OK
        assert L.asdFftExecC2C(h, i.ptr, o.ptr) == INVALID


def test_multidim_c2c_batch_must_match(L):
    with Handle(L) as h:
        assert L.asdFftMakePlan2D(h, 16, 16, C2C, FWD, 2) == OK
        good, bad = tensor((2, 16, 16), N.ACL_COMPLEX64), tensor((3, 16, 16), N.ACL_COMPLEX64)
        assert L.asdFftExecC2C(h, bad.ptr, good.ptr) == INVALID
        assert L.asdFftExecC2C(h, good.ptr, bad.ptr) == INVALID


def _four_step_plan(L, h):
    assert L.asdFftMakePlan1D(h, 65536, C2C, FWD, 2, H) == OK


def test_workspace_not_set_is_internal_error(L):
    i, o = tensor((2, 65536), N.ACL_COMPLEX64), tensor((2, 65536), N.ACL_COMPLEX64, FAKE_PTR * 4096)
    with Handle(L) as h:
        _four_step_plan(L, h)
        size = ctypes.c_size_t()
        assert L.asdFftGetWorkspaceSize(h, ctypes.byref(size)) == OK
        assert size.value == 2 * 65536 * 8
        assert L.asdFftExecC2C(h, i.ptr, o.ptr) == INTERNAL


def test_workspace_not_queried_is_invalid(L):
    i, o = tensor((2, 65536), N.ACL_COMPLEX64), tensor((2, 65536), N.ACL_COMPLEX64, FAKE_PTR * 4096)
    with Handle(L) as h:
        _four_step_plan(L, h)
        assert L.asdFftSetWorkspace(h, ctypes.c_void_p(FAKE_PTR * 65536)) == OK
        assert L.asdFftExecC2C(h, i.ptr, o.ptr) == INVALID


def test_null_data_is_internal_error(L):
    i = tensor((4, 64), N.ACL_COMPLEX64, 0)
    o = tensor((4, 64), N.ACL_COMPLEX64)
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 4, H) == OK
        assert L.asdFftExecC2C(h, i.ptr, o.ptr) == INTERNAL


def test_partial_overlap_rejected(L):
    i = tensor((4, 64), N.ACL_COMPLEX64, 0x100000)
    o = tensor((4, 64), N.ACL_COMPLEX64, 0x100000 + 64)
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 4, H) == OK
        assert L.asdFftExecC2C(h, i.ptr, o.ptr) == INVALID


@pytest.mark.skipif(has_gpu(), reason="checks the behaviour without a HIP device")
def test_exec_without_device_fails_loudly(L):
    """No host execution path: a valid exec on a machine without a GPU is an error."""
    i, o = tensor((4, 64), N.ACL_COMPLEX64), tensor((4, 64), N.ACL_COMPLEX64, 0x800000)
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 4, H) == OK
        assert L.asdFftExecC2C(h, i.ptr, o.ptr) == INTERNAL
        assert L.asdFftSynchronize(h) == INTERNAL


def test_unsupported_lengths_rejected(L):
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 10007, C2C, FWD, 1, H) == INVALID  # prime above the direct-DFT limit
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 1009, C2C, FWD, 1, H) == OK  # prime handled by the direct DFT


def test_next_entries_reject_null_tensors(L):
    with Handle(L) as h:
        # separated planes and ISTFT are implemented (tests/test_separated_contract.py,
        # tests/test_istft_contract.py): null tensors are invalid parameters
        assert L.asdFftExecIstft(h, None, None, None) == INVALID
    # fftshift / ifftshift are implemented (tests/test_shift_contract.py): null tensors are invalid
    assert L.asdFftShift(4, 4, 1, 0, None, None, None, None) == INVALID
    assert L.asdIfftShift(4, 4, 1, 0, 1.0, None, None, None, None) == INVALID


def test_acl_tensor_accessors(L):
    t = tensor((3, 5, 7), N.ACL_COMPLEX64)
    dims = ctypes.POINTER(ctypes.c_int64)()
    n = ctypes.c_uint64()
    assert L.aclGetViewShape(t.ptr, ctypes.byref(dims), ctypes.byref(n)) == 0
    assert [dims[i] for i in range(n.value)] == [3, 5, 7]
    L.aclFreeViewShape(dims)
    dt = ctypes.c_int()
    assert L.aclGetDataType(t.ptr, ctypes.byref(dt)) == 0 and dt.value == N.ACL_COMPLEX64
    neg = (ctypes.c_int64 * 2)(3, -1)
    assert not L.aclCreateTensor(neg, 2, N.ACL_FLOAT, None, 0, N.ACL_FORMAT_ND, neg, 2, None)


def test_set_precision_rules(L):
    with Handle(L) as h:
        assert L.asdFftExtSetPrecision(h, 7) == INVALID
        assert L.asdFftExtSetPrecision(h, N.ASDFFT_PRECISION_SPLIT16) == OK
        assert L.asdFftMakePlan1D(h, 1024, C2C, FWD, 8, H) == OK
        need = ctypes.c_size_t()
        L.asdFftExtDescribePlan(h, None, 0, ctypes.byref(need))
        buf = ctypes.create_string_buffer(need.value)
        L.asdFftExtDescribePlan(h, buf, len(buf), None)
        assert "mfma split-fp16 n=1024" in buf.value.decode()
        assert L.asdFftExtSetPrecision(h, N.ASDFFT_PRECISION_FP32) == INVALID  # plan already made
    with Handle(L) as h:  # other lengths keep the FP32 kernels
        assert L.asdFftExtSetPrecision(h, N.ASDFFT_PRECISION_SPLIT16) == OK
        assert L.asdFftMakePlan1D(h, 2048, C2C, FWD, 8, H) == OK
        need = ctypes.c_size_t()
        L.asdFftExtDescribePlan(h, None, 0, ctypes.byref(need))
        buf = ctypes.create_string_buffer(need.value)
        L.asdFftExtDescribePlan(h, buf, len(buf), None)
        assert "stockham" in buf.value.decode()
    assert L.asdFftExtSetPrecision(ctypes.c_void_p(0x1234), 0) == INVALID
