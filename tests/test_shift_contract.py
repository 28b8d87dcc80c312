"""asdFftShift / asdIfftShift parameter checks (CPU), in the reference's order
(core/fft/fftshift.cpp:101-138, core/fft/ifftshift.cpp; negative cases of
tests/ut/unittest/fft/fftshift/test_fftshift_negative.cpp).  Tensors carry dummy
addresses that are never dereferenced."""
import ctypes

import pytest

from sip_amd import _native as N

OK, INVALID, INTERNAL = N.ACL_SUCCESS, N.ACL_ERROR_INVALID_PARAM, N.ACL_ERROR_INTERNAL_ERROR
DTYPE, NOT_SUPPORTED = N.ACL_ERROR_UNSUPPORTED_DATA_TYPE, N.ACL_ERROR_API_NOT_SUPPORT


@pytest.fixture(scope="module")
def L():
    return N.lib()


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def t(n, dtype=N.ACL_COMPLEX64, ptr=0x100000):
    return N.AclTensor((n,), dtype, ptr)


def both(L, row, col, batch, dist, tin, tout):
    a = L.asdFftShift(row, col, batch, dist, tin.ptr if tin else None, tout.ptr if tout else None, None, None)
    b = L.asdIfftShift(row, col, batch, dist, 1.0, tin.ptr if tin else None, tout.ptr if tout else None, None, None)
    return a, b


@pytest.mark.parametrize("row,col,batch,dist", [(0, 4, 1, 0), (4, 0, 1, 0), (4, 4, 0, 0), (-1, 4, 1, 0)])
def test_non_positive_extents(L, row, col, batch, dist):
    assert both(L, row, col, batch, dist, t(16), t(16, ptr=0x900000)) == (INVALID, INVALID)


def test_overflow_guards(L):
    big = 1 << 32
    assert both(L, big, big, 1, 0, t(16), t(16, ptr=0x900000)) == (INVALID, INVALID)        # row*col
    assert both(L, 1, 1, 1 << 32, 1 << 33, t(16), t(16, ptr=0x900000)) == (INVALID, INVALID)  # (batch-1)*dist


def test_overlapping_slices_rejected(L):
    assert both(L, 2, 2, 4, 3, t(16), t(16, ptr=0x900000)) == (INVALID, INVALID)  # 0 < dist < row*col


def test_dtype_checked(L):
    assert both(L, 2, 2, 1, 0, t(4, N.ACL_FLOAT), t(4, ptr=0x900000)) == (DTYPE, DTYPE)
    assert both(L, 2, 2, 1, 0, t(4), t(4, N.ACL_FLOAT, ptr=0x900000)) == (DTYPE, DTYPE)


def test_element_counts_checked(L):
    assert both(L, 1000, 1000, 1, 0, t(16), t(1000000, ptr=0x9000000)) == (INVALID, INVALID)  # input short
    assert both(L, 2, 2, 4, 0, t(16), t(8, ptr=0x900000)) == (INVALID, INVALID)               # output short
    assert both(L, 4, 6, 3, 32, t(24 * 3), t(72, ptr=0x900000)) == (INVALID, INVALID)          # strided input short


def test_null_tensor_and_overlap(L):
    assert both(L, 2, 2, 1, 0, None, t(4)) == (INVALID, INVALID)
    assert both(L, 2, 2, 1, 0, t(4, ptr=0x100000), t(4, ptr=0x100010)) == (INVALID, INVALID)


def test_valid_call_without_device(L):
    if has_gpu():
        pytest.skip("needs a host without a HIP device")
    assert both(L, 6, 8, 3, 0, t(144), t(144, ptr=0x900000)) == (INTERNAL, INTERNAL)
