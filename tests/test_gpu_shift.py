"""GPU parity of asdFftShift / asdIfftShift: bit-exact against the oracle (a permutation,
and with a scale the same float32 product), on the reference's unit-test shapes
(tests/ut/unittest/fft/{fftshift,ifftshift}/test_*.cpp; the largest batches reduced)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [(6, 8, 3, 0), (7, 9, 2, 0), (1, 16, 4, 0), (16, 1, 2, 0), (4, 6, 3, 32), (33, 67, 5, 0),
         (128, 50, 6, 0), (64, 48, 8, 0), (500, 50, 64, 0), (2, 10000, 2, 0)]
SCALED = [(7, 9, 2, 0, 1.0 / 63.0), (500, 50, 16, 0, 1.0 / 25000.0), (5, 8000, 2, 0, 0.25), (3, 10001, 2, 0, 0.5),
          (16, 1, 2, 0, -2.5)]


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from sip_amd import _native
    _native.lib()
    return torch.device("cuda:0")


def run(torch_dev, flat, row, col, batch, dist, inverse, scale=1.0):
    import torch

    from sip_amd import _native as N
    L = N.lib()
    x = torch.from_numpy(flat).to(torch_dev)
    y = torch.empty(batch * row * col, dtype=torch.complex64, device=torch_dev)
    tin = N.AclTensor(x.shape, N.ACL_COMPLEX64, x.data_ptr())
    tout = N.AclTensor(y.shape, N.ACL_COMPLEX64, y.data_ptr())
    s = torch.cuda.current_stream().cuda_stream
    if inverse:
        st = L.asdIfftShift(row, col, batch, dist, scale, tin.ptr, tout.ptr, s, None)
    else:
        st = L.asdFftShift(row, col, batch, dist, tin.ptr, tout.ptr, s, None)
    assert st == 0
    torch.cuda.synchronize()
    tin.close()
    tout.close()
    return y.cpu().numpy().reshape(batch, row, col)


def make(rng, row, col, batch, dist):
    d = row * col if dist <= 0 else dist
    n = (batch - 1) * d + row * col
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)


@pytest.mark.parametrize("row,col,batch,dist", CASES)
@pytest.mark.parametrize("inverse", [False, True])
def test_shift_bit_exact(torch_dev, oracle, row, col, batch, dist, inverse):
    rng = np.random.default_rng(row * 7 + col)
    flat = make(rng, row, col, batch, dist)
    got = run(torch_dev, flat, row, col, batch, dist, inverse)
    want = oracle.shift2d(flat, row, col, batch, dist, inverse=inverse)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("row,col,batch,dist,scale", SCALED)
def test_inverse_shift_scale_bit_exact(torch_dev, oracle, row, col, batch, dist, scale):
    rng = np.random.default_rng(col)
    flat = make(rng, row, col, batch, dist)
    got = run(torch_dev, flat, row, col, batch, dist, True, scale)
    want = oracle.shift2d(flat, row, col, batch, dist, inverse=True, scale=scale)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
