"""GPU: the streamed first-pass input of large multi-launch plans (non-temporal loads in the
first launch) changes where data is cached, never a value.  Each plan runs with streaming (the
default at these sizes) and with ASDFFT_STREAM_IN=0 on the same input; the outputs must be
bit-identical, and sampled transforms must match the FP64 oracle at relative L2 <= 1e-6 *
log2(N)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (kind, dims, batch): input and intermediate together fill at least the 256 MiB device cache
CASES = [("c2c", (1024, 1024), 32), ("c2c", (256, 256, 256), 1), ("c2c", (65536,), 256),
         ("r2c", (2048, 2048), 8)]


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from sip_amd import _native
    _native.lib()
    return torch.device("cuda:0")


def _input(torch, dev, kind, dims, batch):
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    shape = (batch,) + tuple(dims)
    if kind == "c2c":
        return torch.complex(torch.rand(shape, generator=g, device=dev) * 2 - 1,
                             torch.rand(shape, generator=g, device=dev) * 2 - 1)
    return torch.rand(shape, generator=g, device=dev) * 2 - 1


@pytest.mark.parametrize("kind,dims,batch", CASES, ids=lambda v: str(v))
def test_streamed_input_is_bit_identical(torch_dev, oracle, monkeypatch, kind, dims, batch):
    import torch

    from oracle.oracle import rel_l2_per_row
    from sip_amd.fft import Plan
    x = _input(torch, torch_dev, kind, dims, batch)
    outs = []
    for stream in ("1", "0"):
        monkeypatch.setenv("ASDFFT_STREAM_IN", stream)
        p = Plan(kind, dims, batch, False)
        monkeypatch.delenv("ASDFFT_STREAM_IN")
        try:
            first = [l for l in p.describe().splitlines() if l.strip().startswith("[")][0]
            assert ("streamed-input" in first) == (stream == "1"), first
            outs.append(p.execute(x).cpu().numpy())
        finally:
            p.close()
        torch.cuda.synchronize()
    assert np.array_equal(outs[0], outs[1]), (kind, dims, batch)
    rows = [0, batch - 1] if batch > 1 else [0]
    xs = x[rows].cpu().numpy()
    n = int(np.prod(dims))
    if kind == "c2c":
        want = oracle.c2c_nd(xs, len(dims), False)
    else:
        want = oracle.r2c_nd(xs, len(dims), False)
    err = rel_l2_per_row(outs[0][rows], want, len(dims)).max()
    assert err <= 1e-6 * max(1.0, math.log2(n)), (kind, dims, batch, err)
