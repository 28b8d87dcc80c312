"""GPU parity of the torch_sip operators (libasdfft_torch.so over the C ABI) against the FP64
oracle: C2C / R2C / C2R of rank 2-4 tensors (1D / 2D / 3D plans), the shifts bit for bit, the
current-stream contract, non-contiguous inputs and the plan LRU (capacity, eviction, reuse).
Tolerance: relative L2 per transform <= 1e-6 * log2(N), N = points per transform."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def tol_for(dims):
    n = int(np.prod(dims))
    return 1e-6 * max(1.0, math.log2(max(n, 2)))


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from sip_amd import _native
    ops = _native.torch_ops()
    return torch, ops, torch.device("cuda:0")


def rel(got, want, ndim):
    from oracle.oracle import rel_l2_per_row
    return rel_l2_per_row(got, want, ndim).max()


def cplx(rng, shape):
    return (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(np.complex64)


@pytest.mark.parametrize("shape", [(16, 128), (32, 100), (4, 4096), (3, 48, 40), (2, 64, 64), (2, 16, 32, 64),
                                   (1, 8, 6, 10)], ids=str)
@pytest.mark.parametrize("forward", [True, False])
def test_c2c(env, oracle, shape, forward):
    torch, ops, dev = env
    x = cplx(np.random.default_rng(sum(shape)), shape)
    y = ops.asd_fft_c2c(torch.from_numpy(x).to(dev), forward)
    torch.cuda.synchronize()
    want = oracle.c2c_nd(x, len(shape) - 1, not forward)
    assert y.shape == x.shape and y.dtype == torch.complex64
    assert rel(y.cpu().numpy(), want, len(shape) - 1) <= tol_for(shape[1:])


@pytest.mark.parametrize("shape", [(8, 256), (4, 1000), (3, 97), (2, 24, 40), (2, 8, 16, 30)], ids=str)
@pytest.mark.parametrize("forward", [True, False])
def test_r2c_then_c2r(env, oracle, shape, forward):
    torch, ops, dev = env
    x = np.random.default_rng(shape[-1]).uniform(-1, 1, shape).astype(np.float32)
    nd = len(shape) - 1
    s = ops.asd_fft_r2c(torch.from_numpy(x).to(dev), forward)
    torch.cuda.synchronize()
    assert s.shape == shape[:-1] + (shape[-1] // 2 + 1,) and s.dtype == torch.complex64
    want = oracle.r2c_nd(x, nd, not forward)
    assert rel(s.cpu().numpy(), want, nd) <= tol_for(shape[1:])
    if shape[-1] % 2 == 0:  # the operator infers the real length 2(M-1)
        spec = s.cpu().numpy()
        back = ops.asd_fft_c2r(s, forward)
        torch.cuda.synchronize()
        assert back.shape == shape and back.dtype == torch.float32
        want_back = oracle.c2r_nd(spec, shape[1:], not forward)
        assert rel(back.cpu().numpy(), want_back, nd) <= tol_for(shape[1:])


@pytest.mark.parametrize("row,col,batch", [(6, 8, 3), (7, 9, 2), (33, 67, 5), (500, 50, 4)])
def test_shifts_bit_exact(env, oracle, row, col, batch):
    torch, ops, dev = env
    from sip_amd import fft_shift, ifft_shift
    x = cplx(np.random.default_rng(row + col), (batch, row, col))
    xt = torch.from_numpy(x).to(dev)
    a = fft_shift(xt)
    b = ifft_shift(xt, scale=0.25)
    c = ops.asd_fft_shift(xt.reshape(-1), row, col, batch, 0)
    torch.cuda.synchronize()
    flat = x.reshape(-1)
    want_a = oracle.shift2d(flat, row, col, batch, 0, inverse=False)
    want_b = oracle.shift2d(flat, row, col, batch, 0, inverse=True, scale=0.25)
    assert np.array_equal(a.cpu().numpy().view(np.uint64), want_a.view(np.uint64))
    assert np.array_equal(c.cpu().numpy().view(np.uint64), want_a.view(np.uint64))
    assert np.array_equal(b.cpu().numpy().view(np.uint64), want_b.view(np.uint64))


def test_padded_slices_through_dist(env, oracle):
    torch, ops, dev = env
    row, col, batch, dist = 5, 6, 3, 40
    flat = cplx(np.random.default_rng(5), ((batch - 1) * dist + row * col,))
    got = ops.asd_ifft_shift(torch.from_numpy(flat).to(dev), row, col, batch, dist, 1.0 / 30)
    torch.cuda.synchronize()
    want = oracle.shift2d(flat, row, col, batch, dist, inverse=True, scale=1.0 / 30)
    assert np.array_equal(got.cpu().numpy().view(np.uint64), want.view(np.uint64))


def test_runs_on_the_current_stream(env, oracle):
    torch, ops, dev = env
    x = cplx(np.random.default_rng(3), (64, 2048))
    xt = torch.from_numpy(x).to(dev)
    s = torch.cuda.Stream(device=dev)
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        y = ops.asd_fft_c2c(xt, True)
        z = ops.asd_fft_c2c(y, False)  # ordered after y on s
    s.synchronize()
    assert rel(y.cpu().numpy(), oracle.c2c_nd(x, 1, False), 1) <= tol_for((2048,))
    np.testing.assert_allclose(z.cpu().numpy() / 2048, x, atol=2e-6)


def test_non_contiguous_input(env, oracle):
    torch, ops, dev = env
    x = cplx(np.random.default_rng(4), (256, 12))
    xt = torch.from_numpy(x).to(dev).t()  # (12, 256) view, not contiguous
    y = ops.asd_fft_c2c(xt, True)
    torch.cuda.synchronize()
    want = oracle.c2c_nd(np.ascontiguousarray(x.T), 1, False)
    assert rel(y.cpu().numpy(), want, 1) <= tol_for((256,))


def test_plan_lru(env, oracle):
    torch, ops, dev = env
    ops.asd_fft_plan_cache_clear()
    ops.asd_fft_plan_cache_set_capacity(2)
    try:
        xs = [torch.from_numpy(cplx(np.random.default_rng(n), (4, n))).to(dev) for n in (64, 128, 256)]
        for x in xs:
            ops.asd_fft_c2c(x, True)
        assert ops.asd_fft_plan_cache_size() == 2  # the first plan was evicted
        ops.asd_fft_c2c(xs[2], True)  # hit
        ops.asd_fft_c2c(xs[2], False)  # another direction: another plan
        assert ops.asd_fft_plan_cache_size() == 2
        ops.asd_fft_plan_cache_set_capacity(0)
        assert ops.asd_fft_plan_cache_size() == 0
        y = ops.asd_fft_c2c(xs[0], True)  # capacity 0: a plan per call, still correct
        torch.cuda.synchronize()
        want = oracle.c2c_nd(xs[0].cpu().numpy(), 1, False)
        assert rel(y.cpu().numpy(), want, 1) <= tol_for((64,))
    finally:
        ops.asd_fft_plan_cache_set_capacity(10)


def test_unsupported_length_raises(env):
    torch, ops, dev = env
    x = torch.zeros(2, 10007, dtype=torch.complex64, device=dev)  # prime above the direct-DFT limit
    with pytest.raises(RuntimeError, match="plan creation failed"):
        ops.asd_fft_c2c(x, True)
