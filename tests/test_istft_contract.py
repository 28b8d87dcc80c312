"""asdFftIstftMakePlan / asdFftExecIstft contract (CPU): parameter checks and status codes in
the reference's order (core/fft/istft.cpp:343-433 IstftParamsCheck, :437-523 the output and
window checks, :572-595 asdFftExecIstft, :224-286 asdFftExecIstftV2).  Tensors carry dummy
addresses that are never dereferenced."""
import ctypes

import pytest

from sip_amd import _native as N

OK = N.ACL_SUCCESS
INVALID = N.ACL_ERROR_INVALID_PARAM
INTERNAL = N.ACL_ERROR_INTERNAL_ERROR
C64, F32 = N.ACL_COMPLEX64, N.ACL_FLOAT


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


def spec(shape=(2, 64, 8), dtype=C64, ptr=0x100000):
    return N.AclTensor(shape, dtype, ptr)


def make(L, h, t, nfft=64, hop=16, win=64, center=True, normalized=False, onesided=False, length=0, cplx=True):
    return L.asdFftIstftMakePlan(h, t.ptr if t is not None else None, nfft, hop, win, center, normalized, onesided,
                                 length, cplx)


def test_make_plan_accepts_the_reference_form(L):
    with Handle(L) as h:
        assert make(L, h, spec()) == OK
        size = ctypes.c_size_t(0)
        assert L.asdFftGetWorkspaceSize(h, ctypes.byref(size)) == OK
        assert size.value == 2 * 8 * 64 * 8  # the frames: channel x nframes x nfft complex
        t = ctypes.c_int(0)
        assert L.asdFftGetType(h, ctypes.byref(t)) == OK and t.value == N.ASCEND_FFT_C2C  # as the reference


def test_unsupported_flags_are_replaced_not_refused(L):
    # istft.cpp:371-408: center / normalized / onesided / returnComplex / lengthOpt are forced
    with Handle(L) as h:
        assert make(L, h, spec(), center=False, normalized=True, onesided=True, length=5, cplx=False) == OK


@pytest.mark.parametrize("kw,shape,dtype", [
    ({}, (2, 64, 8), F32),                    # dtype
    ({}, (64, 8), C64),                       # rank
    ({}, (2, 64, 1), C64),                    # nFrames must be > 1
    ({"nfft": 32, "win": 32}, (2, 64, 8), C64),  # nFft != frequency dim
    ({"hop": 0}, (2, 64, 8), C64),            # 0 < hop
    ({"hop": 65}, (2, 64, 8), C64),           # hop <= win
    ({"win": 32, "hop": 16}, (2, 64, 8), C64),  # win == nFft
    ({"nfft": 1500, "win": 1500}, (2, 1500, 8), C64),  # nFft < 1500
])
def test_make_plan_rejections(L, kw, shape, dtype):
    with Handle(L) as h:
        assert make(L, h, spec(shape, dtype), **kw) == INVALID
        assert make(L, h, spec()) == OK  # the rejection left the handle usable


def test_make_plan_null_input_and_handle(L):
    with Handle(L) as h:
        assert make(L, h, None) == INVALID
    t = spec()
    assert L.asdFftIstftMakePlan(ctypes.c_void_p(0xdead), t.ptr, 64, 16, 64, True, False, False, 0, True) == INVALID


def test_make_plan_on_an_initialised_handle_is_internal(L):
    # MakePlan1DFft failing (asdFftMakePlan1D on an initialised plan) -> INTERNAL_ERROR, istft.cpp:545-546
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, N.ASCEND_FFT_C2C, N.ASCEND_FFT_FORWARD, 4, N.ASCEND_FFT_HORIZONTAL) == OK
        assert make(L, h, spec()) == INTERNAL


def exec_(L, h, i, w, o):
    return L.asdFftExecIstft(h, *(t.ptr if t is not None else None for t in (i, w, o)))


def ready(L, h):
    assert make(L, h, spec()) == OK
    size = ctypes.c_size_t(0)
    assert L.asdFftGetWorkspaceSize(h, ctypes.byref(size)) == OK
    assert L.asdFftSetWorkspace(h, ctypes.c_void_p(0x70000000)) == OK


def good():
    return spec(), N.AclTensor((64,), F32, 0x200000), N.AclTensor((2, 112), C64, 0x300000)


def test_exec_checks_in_order(L):
    assert exec_(L, ctypes.c_void_p(0xdead), *good()) == INVALID
    with Handle(L) as h:  # a handle without an ISTFT plan does not match
        assert exec_(L, h, *good()) == INVALID
    with Handle(L) as h:
        assert L.asdFftMakePlan1D(h, 64, N.ASCEND_FFT_C2C, N.ASCEND_FFT_INVERSE, 16, N.ASCEND_FFT_HORIZONTAL) == OK
        assert exec_(L, h, *good()) == INVALID
    with Handle(L) as h:
        ready(L, h)
        i, w, o = good()
        assert exec_(L, h, spec((2, 64, 9)), w, o) == INVALID       # frames differ from the plan
        assert exec_(L, h, spec((2, 64, 8), F32), w, o) == INVALID  # input dtype
        assert exec_(L, h, i, w, N.AclTensor((2, 111), C64, 0x300000)) == INVALID  # length
        assert exec_(L, h, i, w, N.AclTensor((3, 112), C64, 0x300000)) == INVALID  # channel
        assert exec_(L, h, i, w, N.AclTensor((2, 112), F32, 0x300000)) == INVALID  # out dtype
        assert exec_(L, h, i, w, N.AclTensor((224,), C64, 0x300000)) == INVALID    # out rank
        assert exec_(L, h, i, N.AclTensor((63,), F32, 0x200000), o) == INVALID     # window length
        assert exec_(L, h, i, N.AclTensor((64,), C64, 0x200000), o) == INVALID     # window dtype
        assert exec_(L, h, i, N.AclTensor((1, 64), F32, 0x200000), o) == INVALID   # window rank
        assert exec_(L, h, i, None, o) == INVALID
        assert exec_(L, h, i, w, None) == INVALID
        # null data (a descriptor over address 0) -> INTERNAL_ERROR, after the shape checks
        assert exec_(L, h, spec(ptr=0), w, o) == INTERNAL
        # overlap with the output is refused (no in-place update)
        assert exec_(L, h, i, w, N.AclTensor((2, 112), C64, 0x100000)) == INVALID


def test_exec_workspace_guard(L):
    with Handle(L) as h:
        assert make(L, h, spec()) == OK
        assert exec_(L, h, *good()) == INTERNAL  # workspace not set
        assert L.asdFftSetWorkspace(h, ctypes.c_void_p(0x70000000)) == OK
        assert exec_(L, h, *good()) == INVALID  # size never queried


@pytest.mark.skipif(has_gpu(), reason="the legal call would launch")
def test_legal_exec_reaches_the_device_check(L):
    with Handle(L) as h:
        ready(L, h)
        assert exec_(L, h, *good()) == INTERNAL  # no HIP device


def test_istft_plan_is_not_a_c2c_plan(L):
    with Handle(L) as h:
        ready(L, h)
        i = N.AclTensor((16, 64), C64, 0x100000)
        o = N.AclTensor((16, 64), C64, 0x400000)
        assert L.asdFftExecC2C(h, i.ptr, o.ptr) == INVALID
