"""GPU parity of C2C on separated real / imaginary planes (ASCEND_FFT_C2C_SEP through
asdFftExecC2CSeparated) against the FP64 oracle's C2C of re + i*im.

Tolerance as for the interleaved C2C: relative L2 error per transform <= 1e-6 * log2(N),
N = points per transform.  The reference's torch test of this entry
(sip_pta/test/fft/c2c_sep.py:150-185: (16,128), (32,100), inverse (16,128), 3D
(2,16,32,64) both directions, unnormalised inverse) is run with its own comparator
(allclose rtol 1e-3, atol 2e-3) as well.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (fft dims, batch): direct DFT (1, 97), Stockham rows (100 .. 16384, 14400 on the
# 48-point register budget), four-step (65536, 98304)
CASES_1D = [((1,), 3), ((97,), 5), ((100,), 32), ((128,), 16), ((1000,), 7), ((4096,), 9), ((14400,), 2),
            ((16384,), 3), ((65536,), 2), ((98304,), 1)]
# 3D: column passes over both leading axes, one of them, or none (rows only)
CASES_3D = [((16, 32, 64), 2), ((8, 8, 8), 2), ((5, 1, 7), 3), ((1, 1, 64), 2), ((4, 1, 1), 2), ((3, 6, 10), 1)]


def tol_for(dims):
    n = int(np.prod(dims))
    return 1e-6 * max(1.0, math.log2(max(n, 2)))


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from sip_amd import _native
    _native.lib()
    return torch.device("cuda:0")


def planes(rng, shape):
    return (rng.standard_normal(shape).astype(np.float32), rng.standard_normal(shape).astype(np.float32))


def run_sep(torch_dev, dims, batch, inverse, re, im, in_place=False):
    import torch

    from sip_amd.fft import Plan
    p = Plan("c2c_sep", dims, batch, inverse)
    try:
        tr = torch.from_numpy(re).to(torch_dev)
        ti = torch.from_numpy(im).to(torch_dev)
        if in_place:
            orr, oi = p.execute_separated(tr, ti, tr, ti)
        else:
            orr, oi = p.execute_separated(tr, ti)
        torch.cuda.synchronize()
        return orr.cpu().numpy(), oi.cpu().numpy()
    finally:
        p.close()


def check(oracle, dims, re, im, got_re, got_im, inverse):
    from oracle.oracle import rel_l2_per_row
    x = re.astype(np.complex128) + 1j * im.astype(np.complex128)
    want = oracle.c2c_nd(x, len(dims), inverse)
    got = got_re.astype(np.complex128) + 1j * got_im.astype(np.complex128)
    err = rel_l2_per_row(got, want, len(dims)).max()
    assert err <= tol_for(dims), (dims, inverse, err, tol_for(dims))
    return want, got


@pytest.mark.parametrize("dims,batch", CASES_1D + CASES_3D, ids=lambda v: str(v))
@pytest.mark.parametrize("inverse", [False, True])
def test_separated_matches_oracle(torch_dev, oracle, dims, batch, inverse):
    rng = np.random.default_rng(sum(dims) * 31 + batch)
    re, im = planes(rng, (batch,) + dims)
    got_re, got_im = run_sep(torch_dev, dims, batch, inverse, re, im)
    check(oracle, dims, re, im, got_re, got_im, inverse)


@pytest.mark.parametrize("dims,batch", [((128,), 16), ((65536,), 2), ((16, 32, 64), 2), ((97,), 3)],
                         ids=lambda v: str(v))
def test_separated_in_place(torch_dev, oracle, dims, batch):
    rng = np.random.default_rng(7)
    re, im = planes(rng, (batch,) + dims)
    got_re, got_im = run_sep(torch_dev, dims, batch, False, re, im, in_place=True)
    check(oracle, dims, re, im, got_re, got_im, False)


def test_separated_agrees_with_interleaved(torch_dev, oracle):
    """The planes path and the interleaved C2C path give the same transform."""
    import torch

    from sip_amd.fft import Plan
    rng = np.random.default_rng(11)
    for dims, batch in [((4096,), 4), ((65536,), 1), ((16, 32, 64), 2)]:
        re, im = planes(rng, (batch,) + dims)
        got_re, got_im = run_sep(torch_dev, dims, batch, False, re, im)
        p = Plan("c2c", dims, batch, False)
        try:
            x = torch.from_numpy((re + 1j * im).astype(np.complex64)).to(torch_dev)
            y = p.execute(x)
            torch.cuda.synchronize()
            y = y.cpu().numpy()
        finally:
            p.close()
        scale = np.linalg.norm(y)
        diff = np.linalg.norm((got_re + 1j * got_im) - y) / scale
        assert diff <= 2 * tol_for(dims), (dims, diff)


@pytest.mark.parametrize("shape,forward", [((16, 128), True), ((32, 100), True), ((16, 128), False),
                                           ((2, 16, 32, 64), True), ((2, 16, 32, 64), False)])
def test_torch_binding_reference_cases(torch_dev, oracle, shape, forward):
    import torch

    from sip_amd import asd_fft_c2c_sep
    rng = np.random.default_rng(len(shape) * 100 + shape[-1])
    re, im = planes(rng, shape)
    out_re, out_im = asd_fft_c2c_sep(torch.from_numpy(re).to(torch_dev), torch.from_numpy(im).to(torch_dev), forward)
    torch.cuda.synchronize()
    dims = shape[1:]
    got_re, got_im = out_re.cpu().numpy(), out_im.cpu().numpy()
    want, got = check(oracle, dims, re, im, got_re, got_im, not forward)
    assert np.allclose(got, want, rtol=1e-3, atol=2e-3)
