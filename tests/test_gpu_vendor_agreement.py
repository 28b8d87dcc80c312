"""GPU: this library against rocFFT (through torch.fft) on the same device inputs.

A second, independent FP32 implementation beside the FP64 oracle of the parity tests: both
sides carry FP32 rounding (about 1.5e-7 relative L2 each against FP64), so their difference is
held to the same 1e-6 * log2(N) bound.  Semantics per kind and direction (numpy / torch names):
C2C forward fft, inverse ifft(norm="forward"); R2C forward rfft, inverse ihfft(norm="forward")
(the conjugate of rfft); C2R forward hfft, inverse irfft(norm="forward").
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    ("c2c", (4096,), 8), ("c2c", (1000,), 5), ("c2c", (2187,), 3), ("c2c", (6000,), 2), ("c2c", (17,), 4),
    ("c2c", (65536,), 2), ("c2c", (2048, 64), 2), ("c2c", (64, 96, 80), 1),
    ("r2c", (8192,), 4), ("r2c", (2187,), 3), ("r2c", (65536,), 2), ("r2c", (48, 40), 2),
    ("c2r", (8192,), 4), ("c2r", (1000,), 3), ("c2r", (59049,), 1), ("c2r", (48, 40), 2),
]


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from sip_amd import _native
    _native.lib()
    return torch.device("cuda:0")


def vendor(torch, kind, dims, inverse, x):
    axes = tuple(range(-len(dims), 0))
    if kind == "c2c":
        return torch.fft.ifftn(x, dim=axes, norm="forward") if inverse else torch.fft.fftn(x, dim=axes)
    if kind == "r2c":
        y = torch.fft.rfftn(x, dim=axes)
        return torch.conj(y).resolve_conj() if inverse else y
    if inverse:
        return torch.fft.irfftn(x, s=dims, dim=axes, norm="forward")
    # C2R forward: the e^- sign on a Hermitian half = N * irfft of the conjugate
    return torch.fft.irfftn(torch.conj(x).resolve_conj(), s=dims, dim=axes, norm="forward")


@pytest.mark.parametrize("kind,dims,batch", CASES, ids=str)
@pytest.mark.parametrize("inverse", [False, True])
def test_agrees_with_rocfft(dev, kind, dims, batch, inverse):
    import torch

    from sip_amd.fft import Plan
    g = torch.Generator(device=dev)
    g.manual_seed(math.prod(dims) + batch)
    shape = (batch,) + dims
    if kind == "c2c":
        x = torch.complex(torch.rand(shape, generator=g, device=dev) * 2 - 1,
                          torch.rand(shape, generator=g, device=dev) * 2 - 1)
    elif kind == "r2c":
        x = torch.rand(shape, generator=g, device=dev) * 2 - 1
    else:  # a Hermitian half: the spectrum of a real signal
        x = torch.fft.rfftn(torch.rand(shape, generator=g, device=dev) * 2 - 1, dim=tuple(range(-len(dims), 0)))
    x = x.contiguous()
    p = Plan(kind, dims, batch, inverse)
    try:
        got = p.execute(x)
    finally:
        torch.cuda.synchronize()
        p.close()
    want = vendor(torch, kind, dims, inverse, x)
    assert got.shape == want.shape, (got.shape, want.shape)
    flat = lambda t: t.reshape(batch, -1).to(torch.complex128 if t.is_complex() else torch.float64)
    err = (torch.linalg.vector_norm(flat(got) - flat(want), dim=1) / torch.linalg.vector_norm(flat(want), dim=1)).max()
    tol = 1e-6 * max(1.0, math.log2(math.prod(dims)))
    assert err.item() <= tol, (kind, dims, inverse, err.item(), tol)
