"""GPU tests of the split-FP16 MFMA precision form (asdFftExtSetPrecision).

Workload: 1D C2C N = 1024 and N = 65536 (the four-step whose 1024-point pass runs on MFMA;
also 32768 and 131072), inputs uniform[-1, 1] scaled by 10^k per transform with k uniform in
{-3..3}.  Gate: relative L2 per transform vs the FP64 oracle <= 1e-6 * log2 N, the same bound
as the FP32 kernels; both errors are printed side by side.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 1024
TOL = 1e-6 * math.log2(N)
MFMA_LENGTHS = (1024, 32768, 65536, 131072)  # lengths with a split-FP16 MFMA kernel


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from sip_amd import _native
    _native.lib()
    return torch.device("cuda:0")


def scaled_input(rng, batch, n=N):
    x = rng.uniform(-1, 1, (batch, n)) + 1j * rng.uniform(-1, 1, (batch, n))
    k = rng.integers(-3, 4, (batch, 1))
    return (x * 10.0 ** k).astype(np.complex64)


def run(torch_dev, x, inverse, precision, in_place=False):
    import torch

    from sip_amd.fft import Plan
    p = Plan("c2c", (x.shape[1],), x.shape[0], inverse, precision=precision)
    try:
        assert ("mfma split-fp16" in p.describe()) == (precision == "split16" and x.shape[1] in MFMA_LENGTHS)
        xt = torch.from_numpy(x).to(torch_dev)
        out = p.execute(xt, out=xt if in_place else None)
        torch.cuda.synchronize()
        return out.cpu().numpy()
    finally:
        p.close()


@pytest.mark.parametrize("inverse", [False, True])
def test_split16_matches_oracle(torch_dev, oracle, inverse):
    from oracle.oracle import rel_l2_per_row
    rng = np.random.default_rng(31 + inverse)
    x = scaled_input(rng, 2048)
    gold = oracle.c2c(x, inverse)
    e16 = rel_l2_per_row(run(torch_dev, x, inverse, "split16"), gold)
    e32 = rel_l2_per_row(run(torch_dev, x, inverse, "fp32"), gold)
    print(f"\nN=1024 inverse={inverse}: split16 max {e16.max():.3e} rms {np.sqrt((e16**2).mean()):.3e}; "
          f"fp32 max {e32.max():.3e} rms {np.sqrt((e32**2).mean()):.3e}")
    assert e16.max() <= TOL
    assert e32.max() <= TOL


@pytest.mark.parametrize("batch", [1, 3, 5, 7, 4097])
def test_split16_ragged_counts_and_in_place(torch_dev, oracle, batch):
    from oracle.oracle import rel_l2_per_row
    rng = np.random.default_rng(40 + batch)
    x = scaled_input(rng, batch)
    gold = oracle.c2c(x, False)
    got = run(torch_dev, x, False, "split16", in_place=True)
    assert rel_l2_per_row(got, gold).max() <= TOL


def test_split16_closed_forms(torch_dev):
    """Impulse -> constant, constant -> N delta, a zero row stays zero (scale guard), a
    single tone lands in its bin."""
    x = np.zeros((4, N), np.complex64)
    x[0, 0] = 1.0
    x[1, :] = 1.0 - 2.0j
    t = np.arange(N)
    x[3] = np.exp(2j * np.pi * 37 * t / N).astype(np.complex64)
    y = run(torch_dev, x, False, "split16")
    assert np.allclose(y[0], 1.0, atol=1e-6)
    want = np.zeros(N, np.complex128)
    want[0] = N * (1.0 - 2.0j)
    assert np.abs(y[1] - want).max() <= 1e-6 * N
    assert np.all(y[2] == 0)
    assert abs(y[3][37] - N) < 1e-3 * N and np.abs(np.delete(y[3], 37)).max() < 1e-2


def test_split16_unavailable_length_keeps_fp32(torch_dev, oracle):
    from oracle.oracle import rel_l2_per_row
    rng = np.random.default_rng(50)
    x = scaled_input(rng, 16, 2048)
    got = run(torch_dev, x, False, "split16")
    assert rel_l2_per_row(got, oracle.c2c(x, False)).max() <= 1e-6 * 11


@pytest.mark.parametrize("inverse", [False, True])
def test_split16_65536_matches_oracle(torch_dev, oracle, inverse):
    """The precision workload's long case at its full size: N = 65536, batch 64, 10^k scaling."""
    from oracle.oracle import rel_l2_per_row
    n = 65536
    rng = np.random.default_rng(8 + inverse)
    x = scaled_input(rng, 64, n)
    gold = oracle.c2c(x, inverse)
    e16 = rel_l2_per_row(run(torch_dev, x, inverse, "split16"), gold)
    e32 = rel_l2_per_row(run(torch_dev, x, inverse, "fp32"), gold)
    tol = 1e-6 * math.log2(n)
    print(f"\nN=65536 b64 inverse={inverse}: split16 max {e16.max():.3e} rms {np.sqrt((e16**2).mean()):.3e}; "
          f"fp32 max {e32.max():.3e} rms {np.sqrt((e32**2).mean()):.3e}")
    assert e16.max() <= tol
    assert e32.max() <= tol


@pytest.mark.parametrize("n,batch", [(32768, 3), (131072, 2), (65536, 1)])
def test_split16_fourstep_lengths_and_in_place(torch_dev, oracle, n, batch):
    from oracle.oracle import rel_l2_per_row
    rng = np.random.default_rng(n + batch)
    x = scaled_input(rng, batch, n)
    gold = oracle.c2c(x, False)
    got = run(torch_dev, x, False, "split16", in_place=True)
    assert rel_l2_per_row(got, gold).max() <= 1e-6 * math.log2(n)


def test_split16_65536_closed_forms(torch_dev):
    """Impulse at n0 -> the phase ramp W^(n0 k); a zero row stays zero; a single tone lands in
    its bin (checks the transposed store's index map)."""
    n = 65536
    x = np.zeros((3, n), np.complex64)
    x[0, 5] = 1.0
    t = np.arange(n)
    x[2] = np.exp(2j * np.pi * 4097 * t / n).astype(np.complex64)
    y = run(torch_dev, x, False, "split16")
    ramp = np.exp(-2j * np.pi * 5 * t / n)
    assert np.abs(y[0] - ramp).max() < 1e-5
    assert np.all(y[1] == 0)
    assert abs(y[2][4097] - n) < 1e-3 * n and np.abs(np.delete(y[2], 4097)).max() < 5e-2
