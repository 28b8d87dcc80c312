"""GPU parity of the shifts folded into 1D / 2D C2C plans (asdFftExtSetShift) against the
oracle: out = scale * roll_out(FFT(roll_in(x))), with the rolls of asdFftShift (ceil(n/2)) and
asdIfftShift (floor(n/2)) on every FFT axis -- numpy.roll, on which the oracle's shift is pinned.
Tolerance: relative L2 per transform <= 1e-6 * log2(N) against FP64."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHIFTS = [(None, "fft"), ("ifft", None), ("ifft", "fft"), ("fft", "ifft"), (None, None)]
SHAPES = [((64,), 5), ((1000,), 3), ((63,), 4), ((7,), 9), ((16384,), 2), ((2048, 2048), 1), ((63, 40), 3),
          ((7, 9), 2), ((2, 1024), 4), ((256, 2), 3)]


def tol_for(dims):
    return 1e-6 * max(1.0, math.log2(max(int(np.prod(dims)), 2)))


def roll_amount(mode, n):
    return {None: 0, "fft": (n + 1) // 2, "ifft": n // 2}[mode]


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from sip_amd import _native
    _native.lib()
    return torch.device("cuda:0")


@pytest.mark.parametrize("dims,batch", SHAPES, ids=str)
@pytest.mark.parametrize("in_shift,out_shift", SHIFTS, ids=str)
def test_fused_shift_matches_oracle(torch_dev, oracle, dims, batch, in_shift, out_shift):
    import torch

    from oracle.oracle import rel_l2_per_row
    from sip_amd.fft import Plan
    inverse = (out_shift == "ifft")
    scale = 0.5 if in_shift == "ifft" else (1.0 if in_shift or out_shift else 0.25)
    rng = np.random.default_rng(sum(dims) + batch)
    shape = (batch,) + dims
    x = (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(np.complex64)
    axes = tuple(range(1, len(shape)))
    p = Plan("c2c", dims, batch, inverse, in_shift=in_shift, out_shift=out_shift, scale=scale)
    try:
        y = p.execute(torch.from_numpy(x).to(torch_dev))
        torch.cuda.synchronize()
        got = y.cpu().numpy()
    finally:
        p.close()
    xin = np.roll(x, tuple(roll_amount(in_shift, d) for d in dims), axis=axes)
    want = oracle.c2c_nd(xin, len(dims), inverse)
    want = scale * np.roll(want, tuple(roll_amount(out_shift, d) for d in dims), axis=axes)
    err = rel_l2_per_row(got, want, len(dims)).max()
    assert err <= tol_for(dims), (err, tol_for(dims))


def test_fused_equals_separate_sequence(torch_dev):
    """The fused plan against the three-call sequence asdIfftShift -> exec -> asdFftShift with
    the inverse shift's scale (the reference's way), on the GPU."""
    import torch

    from sip_amd import fft_shift, ifft_shift
    from sip_amd.fft import Plan
    X, Y, B = 96, 80, 3
    g = torch.Generator(device=torch_dev)
    g.manual_seed(5)
    x = torch.complex(torch.randn(B, X, Y, generator=g, device=torch_dev),
                      torch.randn(B, X, Y, generator=g, device=torch_dev))
    fused = Plan("c2c", (X, Y), B, False, in_shift="ifft", out_shift="fft", scale=1.0 / (X * Y))
    plain = Plan("c2c", (X, Y), B, False)
    try:
        a = fused.execute(x)
        b = fft_shift(plain.execute(ifft_shift(x, scale=1.0 / (X * Y))))
        torch.cuda.synchronize()
    finally:
        fused.close()
        plain.close()
    rel = (torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)).item()
    assert rel < 1e-6, rel
