"""asdFftExecC2CSeparated / ASCEND_FFT_C2C_SEP contract (CPU).

Status codes and their order follow the reference's asdFftExecC2CSeparated and
asdFftExecV2Separated (core/fft/fft_api.cpp:1846-1985); the shape-mismatch cases
are the reference's own (tests/ut/unittest/fft/1d/c2c/test_fft_c2c_1d.cpp:230-240,
410-480).  Tensors carry dummy addresses that are never dereferenced: every case
here is rejected before a kernel would be launched.
"""
import ctypes

import pytest

from sip_amd import _native as N

OK = N.ACL_SUCCESS
INVALID = N.ACL_ERROR_INVALID_PARAM
INTERNAL = N.ACL_ERROR_INTERNAL_ERROR
NOT_MATCH = N.ACL_ERROR_OP_INPUT_NOT_MATCH
NOT_SUPPORTED = N.ACL_ERROR_API_NOT_SUPPORT
SEP, C2C = N.ASCEND_FFT_C2C_SEP, N.ASCEND_FFT_C2C
FWD, INV = N.ASCEND_FFT_FORWARD, N.ASCEND_FFT_INVERSE
H, V = N.ASCEND_FFT_HORIZONTAL, N.ASCEND_FFT_VERTICAL
F32, C64 = N.ACL_FLOAT, N.ACL_COMPLEX64


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


def planes(shape, dtypes=(F32, F32, F32, F32), base=0x100000, gap=1 << 24, ptrs=None):
    """Four tensors of one shape at distinct, non-overlapping dummy addresses."""
    if ptrs is None:
        ptrs = [base + k * gap for k in range(4)]
    return [N.AclTensor(shape, d, p) for d, p in zip(dtypes, ptrs)]


def sep(L, h, ts):
    return L.asdFftExecC2CSeparated(h, *[t.ptr if t is not None else None for t in ts])


def test_make_plan_rules(L):
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, SEP, FWD, 8, H) == OK
        t = ctypes.c_int(0)
        assert L.asdFftGetType(h, ctypes.byref(t)) == OK and t.value == SEP
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, SEP, INV, 1, V) == OK  # (N, 1) == (1, N)
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, SEP, FWD, 8, V) == NOT_SUPPORTED
    with Handle(L) as h:  # 2D plans reject C2C_SEP (test_fft_c2c_1d.cpp:230-240)
        assert L.asdFftMakePlan2D(h, 64, 64, SEP, FWD, 1) == INVALID
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 1, H) == OK  # the rejection left the handle usable
    with Handle(L) as h:
        assert L.asdFftMakePlan3D(h, 8, 8, 8, SEP, FWD, 2) == OK
    with Handle(L) as h:  # every supported length plans, including four-step and direct-DFT ones
        for n in (1, 97, 100, 1000, 16384, 65536, 1 << 20):
            with Handle(L) as h2:
                assert L.asdFftMakePlan1D(h2, n, SEP, FWD, 2, H) == OK, n


def test_invalid_handle(L):
    assert sep(L, ctypes.c_void_p(0xdead), planes((8, 64))) == INVALID


def test_null_tensor_is_internal_error(L):
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, SEP, FWD, 8, H) == OK
        for k in range(4):
            ts = planes((8, 64))
            ts[k] = None
            assert sep(L, h, ts) == INTERNAL, k


def test_shapes_must_agree(L):
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, SEP, FWD, 8, H) == OK
        for k in range(4):
            ts = planes((8, 64))
            ts[k] = N.AclTensor((8, 64, 1), F32, 0x900000000)  # rank differs
            assert sep(L, h, ts) == INTERNAL, k
            ts[k] = N.AclTensor((8, 32), F32, 0x900000000)  # extent differs
            assert sep(L, h, ts) == INTERNAL, k


def test_planes_must_be_float(L):
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, SEP, FWD, 8, H) == OK
        for k in range(4):
            d = [F32] * 4
            d[k] = C64
            assert sep(L, h, planes((8, 64), tuple(d))) == NOT_MATCH, k
    with Handle(L) as h:  # the dtype check comes before the plan-state check
        assert sep(L, h, planes((8, 64), (C64,) * 4)) == NOT_MATCH
        assert sep(L, h, planes((8, 64))) == INVALID  # not initialised


@pytest.mark.parametrize("shape", [(8, 32), (4, 64), (64,)])
def test_1d_shape_mismatch(L, shape):
    # test_fft_c2c_1d.cpp:412-452: plan batch 8, fftSize 64; one tensor passed four times
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, SEP, FWD, 8, H) == OK
        t = N.AclTensor(shape, F32, 0x100000)
        assert L.asdFftExecC2CSeparated(h, t.ptr, t.ptr, t.ptr, t.ptr) == INVALID


@pytest.mark.parametrize("shape", [(2, 8, 8, 4), (8, 8, 8)])
def test_3d_shape_mismatch(L, shape):
    # test_fft_c2c_1d.cpp:454-480: plan batch 2, (8, 8, 8)
    with Handle(L) as h:
        assert L.asdFftMakePlan3D(h, 8, 8, 8, SEP, FWD, 2) == OK
        t = N.AclTensor(shape, F32, 0x100000)
        assert L.asdFftExecC2CSeparated(h, t.ptr, t.ptr, t.ptr, t.ptr) == INVALID


@pytest.mark.skipif(has_gpu(), reason="the legal shape would launch")
def test_legal_shapes_reach_the_device_check(L):
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, SEP, FWD, 8, H) == OK
        assert sep(L, h, planes((8, 64))) == INTERNAL  # no HIP device
        assert sep(L, h, planes((2, 4, 64))) == INTERNAL  # leading extents 2 x 4 cover the batch
        # pairwise in place is allowed
        assert sep(L, h, planes((8, 64), ptrs=[0x100000, 0x2000000, 0x100000, 0x2000000])) == INTERNAL
    with Handle(L) as h:
        assert L.asdFftMakePlan3D(h, 8, 8, 8, SEP, FWD, 2) == OK
        size = ctypes.c_size_t(0)
        assert L.asdFftGetWorkspaceSize(h, ctypes.byref(size)) == OK
        assert size.value == 2 * 512 * 8
        assert L.asdFftSetWorkspace(h, ctypes.c_void_p(0x7000000000)) == OK
        assert sep(L, h, planes((2, 8, 8, 8))) == INTERNAL


def test_overlapping_planes_rejected(L):
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, SEP, FWD, 8, H) == OK
        n = 8 * 64 * 4
        bad = [
            [0x100000, 0x100000 + n // 2, 0x2000000, 0x3000000],  # input planes overlap
            [0x100000, 0x2000000, 0x3000000, 0x3000000],          # output planes coincide
            [0x100000, 0x2000000, 0x2000000, 0x3000000],          # outputReal on inputImag
            [0x100000, 0x2000000, 0x100004, 0x3000000],           # shifted in place
        ]
        for ptrs in bad:
            assert sep(L, h, planes((8, 64), ptrs=ptrs)) == INVALID, ptrs


def test_other_plan_types_rejected(L):
    # a plan without plane kernels fails as the reference's non-separated cores do
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 8, H) == OK
        assert sep(L, h, planes((8, 64))) == INTERNAL
    # and a separated plan is not a C2C plan
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, SEP, FWD, 8, H) == OK
        i = N.AclTensor((8, 64), C64, 0x100000)
        o = N.AclTensor((8, 64), C64, 0x2000000)
        assert L.asdFftExecC2C(h, i.ptr, o.ptr) == INVALID


def test_workspace_guard_on_four_step(L):
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 65536, SEP, FWD, 2, H) == OK
        assert sep(L, h, planes((2, 65536))) == INTERNAL  # workspace not set
        assert L.asdFftSetWorkspace(h, ctypes.c_void_p(0x7000000000)) == OK
        assert sep(L, h, planes((2, 65536))) == INVALID  # size never queried


def test_describe_marks_plane_steps(L):
    from sip_amd import Plan
    p = Plan("c2c_sep", (4096,), 4)
    d = p.describe()
    assert "in(re,im)->out(re,im)" in d, d
    p3 = Plan("c2c_sep", (4, 8, 16), 2)
    d3 = p3.describe()
    assert "in(re,im)->work" in d3 and "work->out(re,im)" in d3, d3
    p4 = Plan("c2c_sep", (65536,), 1)
    d4 = p4.describe()
    assert "in(re,im)->work" in d4 and "work->out(re,im)" in d4, d4
    p5 = Plan("c2c_sep", (1, 1, 64), 3)  # leading axes of length 1: rows only
    assert "in(re,im)->out(re,im)" in p5.describe()


def test_torch_binding_checks():
    # the operator's own checks (sip_pta/test/fft/c2c_sep.py:57-81 expects RuntimeError naming float32)
    torch = pytest.importorskip("torch")
    from sip_amd import asd_fft_c2c_sep
    with pytest.raises(RuntimeError, match="float32"):
        asd_fft_c2c_sep(torch.zeros(4, 8, dtype=torch.complex64), torch.zeros(4, 8, dtype=torch.complex64))
    with pytest.raises(RuntimeError, match="float32"):
        asd_fft_c2c_sep(torch.zeros(4, 8, dtype=torch.int32), torch.zeros(4, 8, dtype=torch.int32))
    with pytest.raises(RuntimeError, match="2 or 4"):
        asd_fft_c2c_sep(torch.zeros(4, 8, 8), torch.zeros(4, 8, 8))
    with pytest.raises(RuntimeError, match="same shape"):
        asd_fft_c2c_sep(torch.zeros(4, 8), torch.zeros(4, 16))
    with pytest.raises(RuntimeError, match="HIP device"):
        asd_fft_c2c_sep(torch.zeros(4, 8), torch.zeros(4, 8))
