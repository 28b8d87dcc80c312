"""GPU parity of the ISTFT (asdFftIstftMakePlan / asdFftExecIstft through the C ABI) against
the FP64 oracle, which is itself pinned on torch.istft (tests/test_oracle.py).

Cases: the reference's unit tests (tests/ut/unittest/fft/istft/test_fft_istft.cpp:205-213,
including two executions of one plan with the window left unchanged, :149-173), odd and prime
lengths (the direct-DFT frame transform), radix-17/19 lengths (the 32-point register budget),
hop = nfft, hop = 1, the largest nfft (1499) and a long multi-channel signal.  Tolerance:
relative L2 per channel <= 1e-6 * log2(nfft) against FP64, and the reference UT comparator
(allclose rtol 1e-3, atol 1e-4)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [(2, 64, 8, 16), (1, 128, 6, 32), (3, 256, 5, 64), (1, 128, 4, 16), (2, 63, 9, 7), (2, 97, 6, 20),
         (1, 1499, 3, 700), (2, 272, 7, 68), (1, 1216, 5, 300), (2, 60, 5, 60), (1, 16, 40, 1), (4, 2, 9, 1),
         (16, 1024, 100, 256), (3, 1000, 12, 250)]


def tol_for(nfft):
    return 1e-6 * max(1.0, math.log2(max(nfft, 2)))


def window(n):
    return (0.6 + 0.4 * np.sin(2.0 * np.pi * np.arange(n) / n + 0.7)).astype(np.float32)


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from sip_amd import _native
    _native.lib()
    return torch.device("cuda:0")


@pytest.mark.parametrize("channel,nfft,nframes,hop", CASES, ids=str)
def test_istft_matches_oracle(torch_dev, oracle, channel, nfft, nframes, hop):
    import torch

    from oracle.oracle import rel_l2_per_row
    from sip_amd import Istft
    rng = np.random.default_rng(nfft * 7 + hop)
    spec = (rng.uniform(-1, 1, (channel, nfft, nframes)) +
            1j * rng.uniform(-1, 1, (channel, nfft, nframes))).astype(np.complex64)
    w = window(nfft)
    p = Istft(channel, nfft, nframes, hop)
    try:
        st = torch.from_numpy(spec).to(torch_dev)
        wt = torch.from_numpy(w).to(torch_dev)
        out = p.execute(st, wt)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
    finally:
        p.close()
    want = oracle.istft(spec, w, hop)
    assert got.shape == want.shape
    err = rel_l2_per_row(got, want, 1).max()
    assert err <= tol_for(nfft), (err, tol_for(nfft))
    assert np.allclose(got, want, rtol=1e-3, atol=1e-4)
    # the spectrum is left as it was
    assert np.array_equal(st.cpu().numpy(), spec)


def test_istft_repeat_exec_keeps_window(torch_dev, oracle):
    """Two executions of one plan with one window give the same output, and the window tensor is
    unchanged (test_fft_istft.cpp:211-213, the reference's regression for a window overwritten
    with its square)."""
    import torch

    from sip_amd import Istft
    channel, nfft, nframes, hop = 1, 128, 4, 16
    rng = np.random.default_rng(191)
    spec = (rng.uniform(-1, 1, (channel, nfft, nframes)) +
            1j * rng.uniform(-1, 1, (channel, nfft, nframes))).astype(np.complex64)
    w = window(nfft)
    p = Istft(channel, nfft, nframes, hop)
    try:
        st = torch.from_numpy(spec).to(torch_dev)
        wt = torch.from_numpy(w).to(torch_dev)
        first = p.execute(st, wt).clone()
        second = p.execute(st, wt)
        torch.cuda.synchronize()
        assert torch.equal(first, second)
        assert np.array_equal(wt.cpu().numpy(), w)
    finally:
        p.close()
    want = oracle.istft(spec, w, hop)
    assert np.allclose(second.cpu().numpy(), want, rtol=1e-3, atol=1e-4)


def test_istft_inverts_a_gpu_stft(torch_dev):
    """Round trip: a centred, zero-padded, windowed STFT made with torch on the GPU comes back as
    the signal."""
    import torch

    from sip_amd import Istft
    nfft, hop, channel = 512, 128, 3
    sig_len = nfft + hop * 40
    g = torch.Generator(device=torch_dev)
    g.manual_seed(7)
    sig = torch.complex(torch.randn(channel, sig_len, generator=g, device=torch_dev),
                        torch.randn(channel, sig_len, generator=g, device=torch_dev))
    wt = torch.from_numpy(window(nfft)).to(torch_dev)
    nframes = 1 + sig_len // hop
    padded = torch.nn.functional.pad(sig, (nfft // 2, nfft // 2))
    frames = padded.unfold(1, nfft, hop)[:, :nframes, :]          # (channel, nframes, nfft)
    spec = torch.fft.fft(frames * wt, dim=2).transpose(1, 2).contiguous()  # (channel, nfft, nframes)
    p = Istft(channel, nfft, nframes, hop)
    try:
        out = p.execute(spec, wt)
        torch.cuda.synchronize()
    finally:
        p.close()
    ref = sig[:, :out.shape[1]]
    err = (torch.linalg.vector_norm(out - ref) / torch.linalg.vector_norm(ref)).item()
    assert err < 1e-5, err
