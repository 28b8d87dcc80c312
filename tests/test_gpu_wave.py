"""GPU parity of the lane-butterfly kernel (wave_fft.hip): short C2C rows whose inner DFT runs
across the lanes of a wavefront through DPP, ds_swizzle and ds_bpermute exchanges.

ASDFFT_WAVE=1 makes the planner take it for every length it has an instance for
(n = e * lanes, e in {16, 12, 10, 9}, lanes 4..64).  Each case is checked against the FP64
oracle at relative L2 <= 1e-6 * log2(N) per transform, forward and inverse, with batches that
leave part of the last wavefront empty, in place, and from an input that is only 8-byte aligned
(the kernel's element-wise load path).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(e, lanes) for e in (16, 12, 10, 9) for lanes in (4, 8, 16, 32, 64)]


def tol(n):
    return 1e-6 * max(1.0, math.log2(n))


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from sip_amd import _native
    _native.lib()
    return torch.device("cuda:0")


def _rows(rng, batch, n):
    return (rng.uniform(-1, 1, (batch, n)) + 1j * rng.uniform(-1, 1, (batch, n))).astype(np.complex64)


def _plan(monkeypatch, n, batch, inverse, wave="1"):
    from sip_amd.fft import Plan
    monkeypatch.setenv("ASDFFT_WAVE", wave)
    p = Plan("c2c", (n,), batch, inverse)
    monkeypatch.delenv("ASDFFT_WAVE")
    return p


@pytest.mark.parametrize("e,lanes", SHAPES, ids=lambda v: str(v))
def test_wave_rows_match_oracle(torch_dev, oracle, monkeypatch, e, lanes):
    import torch

    from oracle.oracle import rel_l2_per_row
    n = e * lanes
    rng = np.random.default_rng(n)
    per_wave = 64 // lanes
    # 37 transforms; and three wavefronts plus one transform (the last wavefront mostly empty)
    for batch, inverse in ((37, False), (3 * per_wave + 1, True)):
        x = _rows(rng, batch, n)
        p = _plan(monkeypatch, n, batch, inverse)
        try:
            assert "wave n=%d" % n in p.describe() and "points/lane=%d lanes=%d" % (e, lanes) in p.describe(), \
                p.describe()
            got = p.execute(torch.from_numpy(x).to(torch_dev)).cpu().numpy()
        finally:
            p.close()
        err = rel_l2_per_row(got, oracle.c2c(x, inverse), 1).max()
        assert err <= tol(n), (n, batch, inverse, err)


@pytest.mark.parametrize("n", [1024, 256, 768, 640, 576, 36])
def test_wave_in_place_and_unaligned(torch_dev, oracle, monkeypatch, n):
    import torch

    from oracle.oracle import rel_l2_per_row
    rng = np.random.default_rng(n + 1)
    batch = 300
    x = _rows(rng, batch, n)
    want = oracle.c2c(x, False)
    p = _plan(monkeypatch, n, batch, False)
    try:
        # in place: each wavefront reads its transforms whole before it stores
        buf = torch.from_numpy(x).to(torch_dev)
        p.execute(buf, out=buf)
        torch.cuda.synchronize()
        assert rel_l2_per_row(buf.cpu().numpy(), want, 1).max() <= tol(n)
        # input one complex element past a 16-byte boundary: element-wise loads
        flat = torch.zeros(batch * n + 1, dtype=torch.complex64, device=torch_dev)
        flat[1:] = torch.from_numpy(x.reshape(-1)).to(torch_dev)
        xin = flat[1:].view(batch, n)
        assert xin.data_ptr() % 16 == 8
        got = p.execute(xin).cpu().numpy()
        assert rel_l2_per_row(got, want, 1).max() <= tol(n)
    finally:
        p.close()


def test_wave_large_batch_matches_stockham(torch_dev, oracle, monkeypatch):
    """1024 x 8192 (the FP32 companion of the precision form's N=1024 case): the two kernels
    agree to rounding and both match the oracle on sampled rows."""
    import torch

    from oracle.oracle import rel_l2_per_row
    rng = np.random.default_rng(7)
    n, batch = 1024, 8192
    x = _rows(rng, batch, n)
    xt = torch.from_numpy(x).to(torch_dev)
    outs = []
    for wave in ("1", "0"):
        p = _plan(monkeypatch, n, batch, False, wave)
        try:
            assert ("wave n=" in p.describe()) == (wave == "1"), p.describe()
            outs.append(p.execute(xt).cpu().numpy())
        finally:
            p.close()
    assert rel_l2_per_row(outs[0], outs[1].astype(np.complex128), 1).max() <= 2e-6
    rows = rng.choice(batch, 64, replace=False)
    want = oracle.c2c(x[rows], False)
    assert rel_l2_per_row(outs[0][rows], want, 1).max() <= tol(n)
