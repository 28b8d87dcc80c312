"""asdFftExtSetShift (shifts folded into 1D / 2D C2C plans), CPU: argument rules, supported
plans, workspace and the in-place refusal.  Tensors carry dummy addresses."""
import ctypes

import pytest

from sip_amd import _native as N

OK, INVALID, INTERNAL = N.ACL_SUCCESS, N.ACL_ERROR_INVALID_PARAM, N.ACL_ERROR_INTERNAL_ERROR
NOT_SUPPORTED = N.ACL_ERROR_API_NOT_SUPPORT
C2C, R2C = N.ASCEND_FFT_C2C, N.ASCEND_FFT_R2C
FWD, H, V = N.ASCEND_FFT_FORWARD, N.ASCEND_FFT_HORIZONTAL, N.ASCEND_FFT_VERTICAL
SF, SI, S0 = N.ASDFFT_SHIFT_FFT, N.ASDFFT_SHIFT_IFFT, N.ASDFFT_SHIFT_NONE


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


def test_set_shift_rules(L):
    assert L.asdFftExtSetShift(ctypes.c_void_p(0xdead), SF, SF, 1.0) == INVALID
    with Handle(L) as h:
        assert L.asdFftExtSetShift(h, 3, S0, 1.0) == INVALID
        assert L.asdFftExtSetShift(h, S0, -1, 1.0) == INVALID
        assert L.asdFftExtSetShift(h, SI, SF, float("nan")) == INVALID
        assert L.asdFftExtSetShift(h, SI, SF, float("inf")) == INVALID
        assert L.asdFftExtSetShift(h, SI, SF, 1e300) == INVALID  # overflows float32
        assert L.asdFftExtSetShift(h, SI, SF, 0.25) == OK
        assert L.asdFftMakePlan2D(h, 64, 48, C2C, FWD, 2) == OK
        assert L.asdFftExtSetShift(h, SF, SF, 1.0) == INVALID  # the plan exists


@pytest.mark.parametrize("make", [
    lambda L, h: L.asdFftMakePlan1D(h, 64, R2C, FWD, 2, H),           # real plans
    lambda L, h: L.asdFftMakePlan3D(h, 8, 8, 8, C2C, FWD, 1),         # 3D
    lambda L, h: L.asdFftMakePlan1D(h, 65536, C2C, FWD, 2, H),        # four-step length
    lambda L, h: L.asdFftMakePlan1D(h, 97, C2C, FWD, 2, H),           # prime above 19
    lambda L, h: L.asdFftMakePlan1D(h, 64, C2C, FWD, 4, V),           # VERTICAL
    lambda L, h: L.asdFftMakePlan2D(h, 1, 64, C2C, FWD, 1),           # axis of length 1
    lambda L, h: L.asdFftMakePlan1D(h, 64, N.ASCEND_FFT_C2C_SEP, FWD, 2, H),  # separated planes
])
def test_unsupported_plans(L, make):
    with Handle(L) as h:
        assert L.asdFftExtSetShift(h, SI, SF, 1.0) == OK
        assert make(L, h) == NOT_SUPPORTED


def test_plans_and_workspace():
    from sip_amd.fft import Plan
    p = Plan("c2c", (1000,), 3, in_shift="ifft", out_shift="fft", scale=0.5)
    d = p.describe()
    assert p.workspace_size == 0 and "variant=17" in d and "shift in 500 out 500" in d, d
    p = Plan("c2c", (63, 40), 2, in_shift="fft")
    d = p.describe()
    assert p.workspace_size == 2 * 63 * 40 * 8, d
    assert "[in->work]" in d and "variant=16" in d and "shift in 32,20" in d, d
    assert "[work->out]" in d and "variant=18" in d and "shift out 0,0" in d, d
    p = Plan("c2c", (63, 40), 2, scale=2.0)  # a scale alone also folds
    assert "variant=18" in p.describe()


@pytest.mark.skipif(has_gpu(), reason="the legal call would launch")
def test_in_place_refused(L):
    with Handle(L) as h:
        assert L.asdFftExtSetShift(h, SI, SF, 1.0) == OK
        assert L.asdFftMakePlan1D(h, 64, C2C, FWD, 4, H) == OK
        t = N.AclTensor((4, 64), N.ACL_COMPLEX64, 0x100000)
        assert L.asdFftExecC2C(h, t.ptr, t.ptr) == INVALID
        o = N.AclTensor((4, 64), N.ACL_COMPLEX64, 0x900000)
        assert L.asdFftExecC2C(h, t.ptr, o.ptr) == INTERNAL  # out of place: reaches the device check
