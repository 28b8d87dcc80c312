"""Pins the CPU FP64 oracle (CPU only).

Each oracle function is checked three ways:
  1. against the committed numpy-complex128 goldens (tests/golden/), which
     restate the reference's own golden scripts for its own test shapes;
  2. the fast mixed-radix FFT against the long-double direct DFT (independent
     algorithms, brute force);
  3. against closed forms and invariants a plausible mistake would break
     (impulse, constant, single tone, Parseval, linearity, inverse∘forward = N·id,
     time shift, Hermitian symmetry, DC/Nyquist handling of C2R).
A planted mistake (sign flip, dropped conjugate) is shown to fail the same checks.
"""
import numpy as np
import pytest


def _ut_input(oracle, case):
    kind, dims, b = case["kind"], tuple(case["dims"]), case["batch"]
    if kind == "c2c":
        return oracle.ut_complex_input((b,) + dims)
    if kind == "r2c":
        return oracle.ut_real_input((b,) + dims)
    return oracle.ut_complex_input((b,) + dims[:-1] + (dims[-1] // 2 + 1,))


def _run(oracle, kind, dims, x, inverse):
    if kind == "c2c":
        return oracle.c2c_nd(x, len(dims), inverse)
    if kind == "r2c":
        return oracle.r2c_nd(x, len(dims), inverse)
    return oracle.c2r_nd(x, dims, inverse)


def _rel(a, b):
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)


def test_lcg_matches_independent_restatement(oracle, fixtures):
    manifest, _, _ = fixtures
    vals, _ = oracle.lcg_uniform(8)
    assert np.array_equal(vals, np.array(manifest["lcg2026_first8"], dtype=np.float32))
    # a longer stream against an integer restatement in numpy
    n = 4096
    s = np.empty(n, dtype=np.uint64)
    st = 2026
    for i in range(n):
        st = (st * 1664525 + 1013904223) & 0xFFFFFFFF
        s[i] = st
    ref = np.float32(-300.0) + (s.astype(np.float32) * np.float32(2.0 ** -32)) * np.float32(600.0)
    got, nxt = oracle.lcg_uniform(n)
    assert np.array_equal(got, ref)
    assert nxt == st


def test_oracle_matches_ut_goldens(oracle, fixtures):
    manifest, ut, _ = fixtures
    worst = 0.0
    for case in manifest["ut"]:
        x = _ut_input(oracle, case)
        got = _run(oracle, case["kind"], tuple(case["dims"]), x, case["inverse"])
        gold = ut[case["name"]]
        assert got.shape == gold.shape, case["name"]
        err = _rel(got, gold)
        worst = max(worst, err)
        assert err < 1e-13, (case["name"], err)
    assert worst > 0.0  # the comparison is not vacuous


def test_oracle_matches_workload_goldens(oracle, fixtures):
    manifest, _, wl = fixtures
    for case in manifest["workload"]:
        x = wl[case["name"] + "_in"]
        got = _run(oracle, case["kind"], tuple(case["dims"]), x, case["inverse"])
        err = _rel(got, wl[case["name"] + "_gold"])
        assert err < 1e-13, (case["name"], err)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 11, 13, 16, 17, 19, 23, 30, 49, 60, 64, 97, 125, 128, 210,
                               243, 289, 360, 1000, 1024, 2187])
@pytest.mark.parametrize("inverse", [False, True])
def test_fast_fft_matches_direct_dft(oracle, n, inverse):
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    a = oracle.fft_1d(x, inverse)
    b = oracle.dft_direct(x, inverse)
    assert _rel(a, b) < 1e-14


@pytest.mark.parametrize("n", [8, 15, 64, 81, 100, 4096])
def test_closed_forms(oracle, n):
    # impulse at 0 -> all ones; constant -> N at bin 0
    x = np.zeros((1, n), np.complex64)
    x[0, 0] = 1
    assert np.allclose(oracle.c2c(x), 1.0, atol=0, rtol=0)
    x = np.ones((1, n), np.complex64)
    y = oracle.c2c(x)
    assert abs(y[0, 0] - n) < 1e-9 and np.abs(y[0, 1:]).max() < 1e-9
    # single tone e^{+2 pi i k0 n/N} -> N at bin k0 (forward); sign of the exponent pinned
    k0 = 3 % n
    t = np.arange(n)
    x = np.exp(2j * np.pi * k0 * t / n).astype(np.complex128)
    y = oracle.fft_1d(x, False)
    expect = np.zeros(n, complex)
    expect[k0] = n
    assert np.abs(y - expect).max() < 1e-9 * n
    y = oracle.fft_1d(x, True)  # inverse: tone lands on bin -k0
    expect = np.zeros(n, complex)
    expect[(-k0) % n] = n
    assert np.abs(y - expect).max() < 1e-9 * n


@pytest.mark.parametrize("n", [12, 17, 256, 1000])
def test_invariants(oracle, n):
    rng = np.random.default_rng(7 + n)
    x = (rng.uniform(-1, 1, (2, n)) + 1j * rng.uniform(-1, 1, (2, n))).astype(np.complex64)
    y = oracle.c2c(x)
    # Parseval
    assert abs(np.sum(np.abs(y) ** 2) - n * np.sum(np.abs(x.astype(np.complex128)) ** 2)) < 1e-9 * n * n
    # inverse(forward(x)) = N x  (through float64 -> float32 rounding of the intermediate)
    z = oracle.c2c(y.astype(np.complex64), inverse=True)
    assert _rel(z, n * x) < 1e-6
    # linearity
    w = (rng.uniform(-1, 1, (2, n)) + 1j * rng.uniform(-1, 1, (2, n))).astype(np.complex64)
    lhs = oracle.c2c((x + 2 * w).astype(np.complex64))
    rhs = y + 2 * oracle.c2c(w)
    assert _rel(lhs, rhs) < 1e-6
    # time shift by 1 multiplies bin k by e^{-2 pi i k / N}
    ys = oracle.c2c(np.roll(x, 1, axis=1))
    k = np.arange(n)
    assert _rel(ys, y * np.exp(-2j * np.pi * k / n)) < 1e-12


@pytest.mark.parametrize("n", [16, 27, 30, 31, 1056])
def test_real_transforms(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, (3, n)).astype(np.float32)
    full = oracle.c2c(x.astype(np.complex64))
    h = n // 2 + 1
    assert _rel(oracle.r2c(x), full[:, :h]) < 1e-14
    assert _rel(oracle.r2c(x, inverse=True), np.conj(full[:, :h])) < 1e-14
    # C2R inverse of the R2C spectrum gives N x (Hermitian round trip)
    spec = oracle.r2c(x).astype(np.complex64)
    back = oracle.c2r(spec, n, inverse=True)
    assert _rel(back, n * x) < 1e-6
    # C2R forward (hfft) = C2R inverse of the conjugate
    assert _rel(oracle.c2r(spec, n, inverse=False), oracle.c2r(np.conj(spec), n, inverse=True)) < 1e-14
    # imaginary parts of DC (and Nyquist for even n) are ignored
    spec2 = spec.copy()
    spec2[:, 0] += 5j
    if n % 2 == 0:
        spec2[:, -1] -= 3j
    assert _rel(oracle.c2r(spec2, n, True), oracle.c2r(spec, n, True)) == 0.0


def test_nd_is_composition_of_1d(oracle):
    rng = np.random.default_rng(3)
    x = (rng.uniform(-1, 1, (2, 6, 10, 12)) + 1j * rng.uniform(-1, 1, (2, 6, 10, 12))).astype(np.complex64)
    y = oracle.c2c_nd(x, 3)
    ref = x.astype(np.complex128)
    for ax in (3, 2, 1):
        ref = np.apply_along_axis(lambda v: oracle.dft_direct(v), ax, ref)
    assert _rel(y, ref) < 1e-13


def test_planted_mistakes_are_caught(oracle, fixtures):
    """A sign flip or a dropped conjugate in the oracle fails the golden comparison."""
    manifest, ut, _ = fixtures
    case = next(c for c in manifest["ut"] if c["name"] == "c2c1d_n210_fwd")
    x = _ut_input(oracle, case)
    wrong_sign = oracle.c2c(x, inverse=True)
    assert _rel(wrong_sign, ut[case["name"]]) > 0.1
    case = next(c for c in manifest["ut"] if c["name"] == "r2c1d_n27_b1_inv")
    x = _ut_input(oracle, case)
    no_conj = oracle.r2c(x, inverse=False)
    assert _rel(no_conj, ut[case["name"]]) > 0.1
    case = next(c for c in manifest["ut"] if c["name"] == "c2r1d_n30_b1_fwd")
    x = _ut_input(oracle, case)
    wrong_dir = oracle.c2r(x, 30, inverse=True)
    assert _rel(wrong_dir, ut[case["name"]]) > 0.1


# ---- 2D circular shift (asdFftShift / asdIfftShift contract) ----------------------------------
@pytest.mark.parametrize("row,col,batch,dist", [(6, 8, 3, 0), (7, 9, 2, 0), (1, 16, 4, 0), (16, 1, 2, 0),
                                                (4, 6, 3, 32), (33, 67, 2, 0)])
def test_shift_matches_numpy_roll(oracle, row, col, batch, dist):
    """Pinned on numpy.roll, the definition the reference's gen_data.py goldens use:
    forward shift by ((row+1)//2, (col+1)//2), inverse by (row//2, col//2)."""
    rng = np.random.default_rng(row * 100 + col)
    d = row * col if dist <= 0 else dist
    flat = (rng.standard_normal(d * batch) + 1j * rng.standard_normal(d * batch)).astype(np.complex64)
    slices = np.stack([flat[s * d:s * d + row * col].reshape(row, col) for s in range(batch)])
    fwd = oracle.shift2d(flat, row, col, batch, dist, inverse=False)
    inv = oracle.shift2d(flat, row, col, batch, dist, inverse=True)
    assert np.array_equal(fwd, np.roll(slices, ((row + 1) // 2, (col + 1) // 2), axis=(1, 2)))
    assert np.array_equal(inv, np.roll(slices, (row // 2, col // 2), axis=(1, 2)))
    # numpy's own fftshift is the inverse shift for odd sizes (floor), the forward one agrees for even
    assert np.array_equal(inv, np.fft.fftshift(slices, axes=(1, 2)))
    # round trip: inverse undoes forward for every parity
    assert np.array_equal(oracle.shift2d(fwd, row, col, batch, 0, inverse=True), slices)


def test_shift_scale_is_float32_product(oracle):
    rng = np.random.default_rng(5)
    flat = (rng.standard_normal(63 * 2) + 1j * rng.standard_normal(63 * 2)).astype(np.complex64)
    scale = 1.0 / 63.0
    got = oracle.shift2d(flat, 7, 9, 2, 0, inverse=True, scale=scale)
    want = np.roll(flat.reshape(2, 7, 9), (3, 4), axis=(1, 2)) * np.float32(scale)
    assert want.dtype == np.complex64
    assert np.array_equal(got, want)


def test_shift_planted_mistake_is_caught(oracle):
    """A shift by floor instead of ceil on odd sizes (the easy mistake) differs."""
    flat = np.arange(7 * 9, dtype=np.float32).astype(np.complex64)
    fwd = oracle.shift2d(flat, 7, 9, 1, 0, inverse=False)
    assert not np.array_equal(fwd, np.roll(flat.reshape(1, 7, 9), (3, 4), axis=(1, 2)))


# ---- ISTFT (asdFftIstftMakePlan / asdFftExecIstft) ---------------------------------------------
def _istft_window(n):
    """The reference unit test's always-positive window (test_fft_istft.cpp:42-50)."""
    return (0.6 + 0.4 * np.sin(2.0 * np.pi * np.arange(n) / n + 0.7)).astype(np.float32)


def _torch_istft(spec, window, hop):
    torch = pytest.importorskip("torch")
    n = spec.shape[1]
    return torch.istft(torch.from_numpy(spec.astype(np.complex128)), n_fft=n, hop_length=hop, win_length=n,
                       window=torch.from_numpy(window.astype(np.float64)), center=True, normalized=False,
                       onesided=False, return_complex=True).numpy()


# (channel, nfft, nframes, hop): the reference's UT cases (test_fft_istft.cpp:205-213), odd and
# prime lengths, no overlap (hop = nfft) and hop 1
ISTFT_CASES = [(2, 64, 8, 16), (1, 128, 6, 32), (3, 256, 5, 64), (1, 128, 4, 16), (2, 63, 9, 7), (1, 97, 6, 20),
               (2, 60, 5, 60), (1, 16, 20, 1), (1, 2, 3, 1)]


@pytest.mark.parametrize("channel,nfft,nframes,hop", ISTFT_CASES)
def test_istft_matches_torch_istft(oracle, channel, nfft, nframes, hop):
    """Pinned on torch.istft (center, not normalised, two-sided, complex) in FP64: the library
    routine whose semantics the reference's golden (gen_data_istft.py) restates."""
    rng = np.random.default_rng(nfft * 10 + hop)
    spec = (rng.standard_normal((channel, nfft, nframes)) +
            1j * rng.standard_normal((channel, nfft, nframes))).astype(np.complex64)
    w = _istft_window(nfft)
    got = oracle.istft(spec, w, hop)
    want = _torch_istft(spec, w, hop)
    assert got.shape == (channel, oracle.istft_out_len(nfft, hop, nframes)) == want.shape
    assert _rel(got, want) < 1e-13


def test_istft_length_one_closed_form(oracle):
    """nfft = 1 (hop 1): every frame is one sample, so out[b][i] = spec[b][0][i] / w[0]."""
    spec = (np.arange(2 * 5) + 1j * np.arange(2 * 5)[::-1]).astype(np.complex64).reshape(2, 1, 5)
    w = np.array([0.8], dtype=np.float32)
    got = oracle.istft(spec, w, 1)
    assert got.shape == (2, 5)
    assert _rel(got, spec[:, 0, :].astype(np.complex128) / np.float64(w[0])) < 1e-15


def test_istft_inverts_the_centred_stft(oracle):
    """Round trip of the reference golden's own self-check (gen_data_istft.py:67-95): a signal's
    zero-padded, windowed STFT comes back as the signal."""
    nfft, hop = 128, 32
    w = _istft_window(nfft).astype(np.float64)
    rng = np.random.default_rng(42)
    sig_len = nfft + hop * 5
    t = np.arange(sig_len)
    signal = 0.5 * np.sin(2 * np.pi * 3.0 * t / nfft) + 0.3 * rng.standard_normal(sig_len)
    nframes = 1 + sig_len // hop
    padded = np.pad(signal, (nfft // 2, nfft // 2))
    frames = np.stack([padded[i * hop:i * hop + nfft] for i in range(nframes)], axis=1)
    spec = np.fft.fft(frames * w[:, None], axis=0)[None, :, :].astype(np.complex64)
    restored = oracle.istft(spec, w.astype(np.float32), hop)[0]
    ref = signal[:restored.shape[0]]
    assert np.max(np.abs(restored - ref)) / np.max(np.abs(ref)) < 1e-5


def test_istft_planted_mistakes_are_caught(oracle):
    """The plausible mistakes (forward sign, crop off by one sample, envelope left out) all fail
    the torch.istft comparison."""
    channel, nfft, nframes, hop = 2, 64, 8, 16
    rng = np.random.default_rng(3)
    spec = (rng.standard_normal((channel, nfft, nframes)) +
            1j * rng.standard_normal((channel, nfft, nframes))).astype(np.complex64)
    w = _istft_window(nfft)
    want = _torch_istft(spec, w, hop)
    good = oracle.istft(spec, w, hop)
    assert _rel(good, want) < 1e-13
    wrong_sign = np.conj(oracle.istft(np.conj(spec), w, hop))
    assert _rel(wrong_sign, want) > 0.1
    assert _rel(good[:, 1:], want[:, :-1]) > 0.1
    env = np.zeros(nfft + hop * (nframes - 1))
    for t in range(nframes):
        env[t * hop:t * hop + nfft] += w.astype(np.float64) ** 2
    no_env = good * env[nfft // 2:nfft // 2 + good.shape[1]]
    assert _rel(no_env, want) > 0.1
