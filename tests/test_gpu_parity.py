"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle and the
committed numpy goldens.

Tolerance (stated per config, from the engine's accuracy target): relative L2
error of every transform against the FP64 result <= 1e-6 * log2(N), N = the
transform's total point count (at least 1e-6).  The reference's own unit-test
comparator (numpy isclose, rtol 1e-3 / atol 1e-5, 2D atol 1e-4,
tests/ut/unittest/test_common.h:41-100) is applied to the UT cases as well.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def tol_for(dims):
    n = int(np.prod(dims))
    return 1e-6 * max(1.0, math.log2(max(n, 2)))


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from sip_amd import _native
    _native.lib()  # fail loudly if the native library is missing
    return torch.device("cuda:0")


def run_gpu(torch_dev, kind, dims, batch, inverse, x):
    import torch

    from sip_amd.fft import Plan
    p = Plan(kind, dims, batch, inverse)
    try:
        xt = torch.from_numpy(np.ascontiguousarray(x)).to(torch_dev)
        out = p.execute(xt)
        torch.cuda.synchronize()
        return out.cpu().numpy()
    finally:
        p.close()


def oracle_run(oracle, kind, dims, x, inverse):
    if kind == "c2c":
        return oracle.c2c_nd(x, len(dims), inverse)
    if kind == "r2c":
        return oracle.r2c_nd(x, len(dims), inverse)
    return oracle.c2r_nd(x, dims, inverse)


def per_row_rel(got, gold, ndim):
    from oracle.oracle import rel_l2_per_row
    return rel_l2_per_row(got, gold, ndim)


def ut_input(oracle, case):
    kind, dims, b = case["kind"], tuple(case["dims"]), case["batch"]
    if kind == "c2c":
        return oracle.ut_complex_input((b,) + dims)
    if kind == "r2c":
        return oracle.ut_real_input((b,) + dims)
    return oracle.ut_complex_input((b,) + dims[:-1] + (dims[-1] // 2 + 1,))


def test_ut_cases_match_numpy_golden(torch_dev, oracle, fixtures):
    manifest, ut, _ = fixtures
    worst = []
    for case in manifest["ut"]:
        kind, dims, b, inv = case["kind"], tuple(case["dims"]), case["batch"], case["inverse"]
        x = ut_input(oracle, case)
        got = run_gpu(torch_dev, kind, dims, b, inv, x)
        gold = ut[case["name"]]
        assert got.shape == gold.shape, case["name"]
        err = per_row_rel(got, gold, len(dims)).max()
        worst.append((err / tol_for(dims), case["name"], err))
        assert err <= tol_for(dims), (case["name"], err, tol_for(dims))
        if kind == "c2c" or len(dims) == 1:
            # the reference's elementwise comparator, on the shapes its own tests use it for
            atol = 1e-4 if len(dims) >= 2 else 1e-5
            g32 = gold.astype(np.complex64 if np.iscomplexobj(gold) else np.float32)
            assert np.isclose(got, g32, rtol=1e-3, atol=atol).all(), case["name"]
    worst.sort(reverse=True)
    print("worst UT cases (err/tol, name, rel-L2):", worst[:5])


def test_workload_slices_match_numpy_golden(torch_dev, fixtures):
    manifest, _, wl = fixtures
    for case in manifest["workload"]:
        kind, dims, b, inv = case["kind"], tuple(case["dims"]), case["batch"], case["inverse"]
        x = wl[case["name"] + "_in"]
        got = run_gpu(torch_dev, kind, dims, b, inv, x)
        err = per_row_rel(got, wl[case["name"] + "_gold"], len(dims)).max()
        print(f"{case['name']}: rel-L2 {err:.3e} (tol {tol_for(dims):.1e})")
        assert err <= tol_for(dims), (case["name"], err)


def _uniform_c(rng, shape):
    return (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(np.complex64)


# Full-size benchmark configurations; the oracle checks every transform of each.
FULL_CONFIGS = [
    ("c2c", (4096,), 512, 0, 512),
    ("c2c", (1000,), 256, 1, 256),
    ("c2c", (2187,), 256, 2, 256),
    ("c2c", (6000,), 256, 3, 256),
    ("c2c", (4096,), 256, 4, 256),
    ("c2c", (1024,), 8192, 9, 8192),
    ("c2c", (65536,), 64, 8, 64),
]


@pytest.mark.parametrize("kind,dims,batch,seed,check_rows", FULL_CONFIGS)
@pytest.mark.parametrize("inverse", [False, True])
def test_full_size_1d_vs_oracle(torch_dev, oracle, kind, dims, batch, seed, check_rows, inverse):
    rng = np.random.default_rng(seed)
    x = _uniform_c(rng, (batch,) + dims)
    if dims[0] in (65536, 1024):  # precision-form inputs: per-transform scale 10^k, k in -3..3
        x = (x * (10.0 ** rng.integers(-3, 4, (batch, 1)))).astype(np.complex64)
    got = run_gpu(torch_dev, kind, dims, batch, inverse, x)
    rows = np.sort(rng.choice(batch, size=min(check_rows, batch), replace=False))
    gold = oracle.c2c(x[rows], inverse)
    err = per_row_rel(got[rows], gold, 1)
    print(f"{dims} b{batch} inv={inverse}: max rel-L2 {err.max():.3e} rms {np.sqrt((err**2).mean()):.3e}")
    assert err.max() <= tol_for(dims)


@pytest.mark.parametrize("inverse", [False, True])
def test_full_size_real_8192(torch_dev, oracle, inverse):
    rng = np.random.default_rng(5)
    n, batch = 8192, 1024
    t = np.arange(n)
    x = rng.standard_normal((batch, n))
    k = rng.integers(1, n // 2, (batch, 3))
    amp = rng.uniform(0.5, 4.0, (batch, 3))
    ph = rng.uniform(0, 2 * np.pi, (batch, 3))
    for j in range(3):
        x += amp[:, j:j + 1] * np.cos(2 * np.pi * k[:, j:j + 1] * t / n + ph[:, j:j + 1])
    x = x.astype(np.float32)
    spec = run_gpu(torch_dev, "r2c", (n,), batch, inverse, x)
    rows = np.arange(batch)  # every transform
    err = per_row_rel(spec[rows], oracle.r2c(x[rows], inverse), 1)
    print(f"R2C 8192x1024 inv={inverse}: max rel-L2 {err.max():.3e}")
    assert err.max() <= tol_for((n,))
    # C2R of the GPU spectrum (the R2C-then-C2R pipeline of the benchmark)
    back = run_gpu(torch_dev, "c2r", (n,), batch, inverse, spec)
    gold = oracle.c2r(spec[rows], n, inverse)
    err = per_row_rel(back[rows], gold, 1)
    print(f"C2R 8192x1024 inv={inverse}: max rel-L2 {err.max():.3e}")
    assert err.max() <= tol_for((n,))
    if inverse:
        # conj(rfft) -> C2R inverse of the conjugate == N x
        pass
    else:
        rt = run_gpu(torch_dev, "c2r", (n,), batch, True, spec)
        assert per_row_rel(rt, n * x.astype(np.float64), 1).max() <= 2 * tol_for((n,))


@pytest.mark.parametrize("inverse", [False, True])
def test_full_size_2d_2048(torch_dev, oracle, inverse):
    rng = np.random.default_rng(6)
    x = _uniform_c(rng, (4, 2048, 2048))
    got = run_gpu(torch_dev, "c2c", (2048, 2048), 4, inverse, x)
    gold = oracle.c2c_nd(x, 2, inverse)  # all four slices
    err = per_row_rel(got, gold, 2)
    print(f"2D 2048^2 x4 inv={inverse}: per-slice rel-L2 {np.array2string(err, precision=3)}")
    assert err.max() <= tol_for((2048, 2048))


def test_full_size_3d_256(torch_dev, oracle):
    rng = np.random.default_rng(7)
    x = _uniform_c(rng, (1, 256, 256, 256))
    got = run_gpu(torch_dev, "c2c", (256, 256, 256), 1, False, x)
    gold = oracle.c2c_nd(x, 3, False)
    err = per_row_rel(got, gold, 3).max()
    print(f"3D 256^3: rel-L2 {err:.3e}")
    assert err <= tol_for((256, 256, 256))


@pytest.mark.parametrize("kind,dims,batch", [("c2c", (4096,), 512), ("c2c", (6000,), 64), ("c2c", (65536,), 8),
                                             ("c2c", (2048, 2048), 1), ("c2c", (64, 96, 80), 2)])
def test_round_trip_is_n_times_identity(torch_dev, kind, dims, batch):
    rng = np.random.default_rng(11)
    x = _uniform_c(rng, (batch,) + dims)
    y = run_gpu(torch_dev, kind, dims, batch, False, x)
    z = run_gpu(torch_dev, kind, dims, batch, True, y)
    n = int(np.prod(dims))
    err = per_row_rel(z, n * x.astype(np.complex128), len(dims)).max()
    assert err <= 2 * tol_for(dims)


def test_in_place_c2c(torch_dev, oracle):
    import torch

    from sip_amd.fft import Plan
    rng = np.random.default_rng(12)
    for dims, batch in (((4096,), 16), ((1000,), 5), ((128, 96), 2), ((65536,), 2)):
        x = _uniform_c(rng, (batch,) + dims)
        p = Plan("c2c", dims, batch, False)
        xt = torch.from_numpy(x).to(torch_dev)
        p.execute(xt, out=xt)
        torch.cuda.synchronize()
        gold = oracle.c2c_nd(x, len(dims))
        assert per_row_rel(xt.cpu().numpy(), gold, len(dims)).max() <= tol_for(dims)
        p.close()


def test_vertical_layout(torch_dev, oracle):
    rng = np.random.default_rng(13)
    n, batch = 512, 48
    x = _uniform_c(rng, (n, batch))
    from sip_amd.fft import Plan
    import torch
    p = Plan("c2c", (n,), batch, True, vertical=True)
    out = p.execute(torch.from_numpy(x).to(torch_dev))
    torch.cuda.synchronize()
    gold = oracle.c2c(np.ascontiguousarray(x.T), True).T
    assert per_row_rel(out.cpu().numpy().T, gold.T, 1).max() <= tol_for((n,))
    p.close()


def test_input_is_not_modified(torch_dev):
    import torch

    from sip_amd.fft import Plan
    rng = np.random.default_rng(14)
    for kind, dims, batch in (("c2c", (65536,), 2), ("c2r", (8, 12, 16), 2), ("c2c", (256, 256), 2)):
        shape = (batch,) + (dims[:-1] + (dims[-1] // 2 + 1,) if kind == "c2r" else dims)
        x = _uniform_c(rng, shape)
        xt = torch.from_numpy(x).to(torch_dev)
        p = Plan(kind, dims, batch, False)
        p.execute(xt)
        torch.cuda.synchronize()
        assert np.array_equal(xt.cpu().numpy(), x)
        p.close()


def _static_rows():
    import os
    rows = []
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sip_amd", "csrc",
                        "static_instances.txt")
    for line in open(path):
        line = line.split("#", 1)[0].strip()
        if line:
            rows.append([int(x) for x in line.split()])
    return rows


def test_specialised_kernels_match_runtime_kernel(torch_dev, oracle, monkeypatch):
    """Every compile-time specialised instance agrees with the run-time kernel and the oracle."""
    import torch

    from sip_amd.fft import Plan
    rng = np.random.default_rng(21)
    cases = {
        0: lambda n: [("c2c", (n,), 48, False), ("c2c", (n,), 17, True)],
        1: lambda n: [("c2c", (n, 64), 2, False), ("c2c", (n, 40), 1, True)],
        2: lambda n: [("r2c", (2 * n,), 24, False), ("r2c", (2 * n,), 9, True)],
        3: lambda n: [("c2r", (2 * n,), 24, True), ("c2r", (2 * n,), 9, False)],
        6: lambda n: [("c2c", (n * n,), 3, False)],
        7: lambda n: [("c2c", (n * n,), 2, True)],
    }
    checked = 0
    for row in _static_rows():
        variant, n = row[0], row[2]
        for kind, dims, batch, inv in cases.get(variant, lambda n: [])(n):
            shape = (batch,) + (dims[:-1] + (dims[-1] // 2 + 1,) if kind == "c2r" else dims)
            if kind == "r2c":
                x = rng.uniform(-1, 1, shape).astype(np.float32)
            else:
                x = _uniform_c(rng, shape)
            outs = []
            for no_static in ("0", "1"):
                monkeypatch.setenv("ASDFFT_NO_STATIC", no_static)
                p = Plan(kind, dims, batch, inv)
                desc = p.describe()
                if no_static == "1":
                    assert "static=" not in desc, desc
                outs.append(p.execute(torch.from_numpy(x).to(torch_dev)).cpu().numpy())
                p.close()
            monkeypatch.delenv("ASDFFT_NO_STATIC")
            gold = oracle_run(oracle, kind, dims, x, inv)
            for o in outs:
                assert per_row_rel(o, gold, len(dims)).max() <= tol_for(dims), (row, kind, dims)
            assert per_row_rel(outs[0], outs[1].astype(np.complex128) if np.iscomplexobj(outs[1])
                               else outs[1].astype(np.float64), len(dims)).max() <= 1e-6
            checked += 1
    assert checked >= 20


@pytest.mark.parametrize("n,batch", [(1000, 256), (2187, 256), (1000, 7), (2187, 3)])
def test_twiddle_preload_rows_match(torch_dev, oracle, monkeypatch, n, batch):
    """Small row launches load every pass's twiddles at kernel start (variant 22).  Both
    kernels multiply by the same table entries; the compiler may contract the products into
    FMAs differently, so they agree to rounding (1e-6 relative), and both match the oracle."""
    import torch

    from sip_amd.fft import Plan
    rng = np.random.default_rng(n + batch)
    x = _uniform_c(rng, (batch, n))
    outs = []
    for pre in ("1", "0"):
        monkeypatch.setenv("ASDFFT_TW_PRELOAD", pre)
        p = Plan("c2c", (n,), batch, False)
        assert ("variant=22" in p.describe()) == (pre == "1"), p.describe()
        outs.append(p.execute(torch.from_numpy(x).to(torch_dev)).cpu().numpy())
        p.close()
    monkeypatch.delenv("ASDFFT_TW_PRELOAD")
    assert per_row_rel(outs[0], outs[1].astype(np.complex128), 1).max() <= 1e-6
    gold = oracle.c2c(x, False)
    for o in outs:
        assert per_row_rel(o, gold, 1).max() <= tol_for((n,))


# long column passes (the one-plane LDS exchange kernel: 2048, 1024 and 1000 points with 16
# columns per workgroup, 4096 points with 8): full and partial workgroups, 2D, 3D and VERTICAL 1D
LONG_COLUMNS = [("c2c", (2048, 64), 1, False), ("c2c", (2048, 40), 2, True), ("c2c", (2048, 4, 8), 1, False),
                ("c2c", (2048, 24), 1, True), ("c2c", (1000, 40), 2, False), ("c2c", (1000, 1000), 1, True),
                ("c2c", (1024, 40), 1, False), ("c2c", (1024, 1024), 1, True), ("c2c", (4096, 40), 1, False),
                ("c2c", (4096, 12), 1, True), ("c2c", (512, 512), 3, False), ("c2c", (512, 96), 3, True),
                ("c2c", (512, 64, 32), 2, False), ("c2c", (512, 40), 1, True)]


@pytest.mark.parametrize("kind,dims,batch,inverse", LONG_COLUMNS, ids=str)
def test_long_column_passes(torch_dev, oracle, kind, dims, batch, inverse):
    rng = np.random.default_rng(sum(dims) + batch)
    shape = (batch,) + dims
    x = (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(np.complex64)
    got = run_gpu(torch_dev, kind, dims, batch, inverse, x)
    want = oracle.c2c_nd(x, len(dims), inverse)
    err = per_row_rel(got, want, len(dims)).max()
    assert err <= tol_for(dims), (err, tol_for(dims))


# one-plane geometries that plans reach only through the planner's overrides: rows C2C on one
# LDS plane (ASDFFT_ROWS_SPLIT=1) and the 2048-point column pass with 8 columns per workgroup
# or a 64-point register budget (ASDFFT_SPLIT_TPB / ASDFFT_SPLIT_EMAX)
OPT_IN_SPLIT = [
    ({"ASDFFT_ROWS_SPLIT": "1"}, (1024,), 37, "variant=20"),
    ({"ASDFFT_ROWS_SPLIT": "1"}, (4096,), 9, "variant=20"),
    ({"ASDFFT_SPLIT_TPB": "8"}, (2048, 40), 1, "variant=15 n=2048 count=40 radices=16x16x8 emax=32 T=64 tpb=8"),
    ({"ASDFFT_SPLIT_EMAX": "64", "ASDFFT_SPLIT_TPB": "8"}, (2048, 24), 1,
     "variant=15 n=2048 count=24 radices=16x16x8 emax=64 T=32 tpb=8"),
    ({"ASDFFT_SPLIT_EMAX": "64"}, (2048, 20), 2, "variant=15 n=2048 count=40 radices=16x16x8 emax=64 T=32 tpb=16"),
]


@pytest.mark.parametrize("env,dims,batch,expect", OPT_IN_SPLIT, ids=str)
def test_opt_in_one_plane_kernels(torch_dev, oracle, monkeypatch, env, dims, batch, expect):
    import torch

    from sip_amd.fft import Plan
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    rng = np.random.default_rng(sum(dims) + batch)
    x = _uniform_c(rng, (batch,) + dims)
    for inverse in (False, True):
        p = Plan("c2c", dims, batch, inverse)
        try:
            assert expect in p.describe() and "static=" in p.describe(), p.describe()
            got = p.execute(torch.from_numpy(x).to(torch_dev)).cpu().numpy()
        finally:
            p.close()
        want = oracle.c2c_nd(x, len(dims), inverse)
        err = per_row_rel(got, want, len(dims)).max()
        assert err <= tol_for(dims), (env, dims, inverse, err)


# 2D C2C with the columns' first radix-2 stage folded into the row pass (ASDFFT_PAIR2D=1, the
# radix-4 form off): rows n and n + X/2 in one workgroup written to rows 2n and 2n + 1, then the
# length-X/2 columns in place at stride 2Y
PAIR2D = [((2048, 2048), 1), ((2048, 2048), 2), ((2048, 1024), 2), ((1024, 1024), 3), ((1024, 2048), 1)]


@pytest.mark.parametrize("dims,batch", PAIR2D, ids=str)
@pytest.mark.parametrize("inverse", [False, True])
def test_pair2d_plans(torch_dev, oracle, monkeypatch, dims, batch, inverse):
    import torch

    from sip_amd.fft import Plan
    monkeypatch.setenv("ASDFFT_PAIR2D", "1")
    monkeypatch.setenv("ASDFFT_QUAD2D", "0")
    rng = np.random.default_rng(sum(dims) + batch + inverse)
    x = _uniform_c(rng, (batch,) + dims)
    p = Plan("c2c", dims, batch, inverse)
    try:
        assert "variant=21" in p.describe(), p.describe()
        xt = torch.from_numpy(x).to(torch_dev)
        got = p.execute(xt).cpu().numpy()
        assert np.array_equal(xt.cpu().numpy(), x)  # the input is left as it was
    finally:
        p.close()
    want = oracle.c2c_nd(x, 2, inverse)  # every slice
    err = per_row_rel(got, want, 2).max()
    assert err <= tol_for(dims), (dims, batch, inverse, err)


def test_vertical_2048_columns(torch_dev, oracle):
    import torch

    from sip_amd.fft import Plan
    n, batch = 2048, 20
    rng = np.random.default_rng(20)
    x = (rng.uniform(-1, 1, (n, batch)) + 1j * rng.uniform(-1, 1, (n, batch))).astype(np.complex64)
    p = Plan("c2c", (n,), batch, False, vertical=True)
    try:
        y = p.execute(torch.from_numpy(x).to(torch_dev))
        torch.cuda.synchronize()
        got = y.cpu().numpy()
    finally:
        p.close()
    want = oracle.c2c_nd(np.ascontiguousarray(x.T), 1, False).T
    err = per_row_rel(np.ascontiguousarray(got.T), np.ascontiguousarray(want.T), 1).max()
    assert err <= tol_for((n,)), err


# real lengths beyond one workgroup: even N with N/2 > 16384 (four-step on N/2 plus the
# post/pre-twiddle pass), odd N > 16384 (four-step on N between promote / Hermitian-extension
# and truncate / real-part passes)
LARGE_REAL = [(65536, 2), (131072, 1), (40000, 3), (59049, 2), (16807, 2)]


@pytest.mark.parametrize("n,batch", LARGE_REAL, ids=str)
@pytest.mark.parametrize("inverse", [False, True])
def test_large_real_r2c(torch_dev, oracle, n, batch, inverse):
    rng = np.random.default_rng(n + batch)
    x = rng.uniform(-1, 1, (batch, n)).astype(np.float32)
    got = run_gpu(torch_dev, "r2c", (n,), batch, inverse, x)
    want = oracle.r2c_nd(x, 1, inverse)
    assert got.shape == want.shape
    err = per_row_rel(got, want, 1).max()
    assert err <= tol_for((n,)), (err, tol_for((n,)))


@pytest.mark.parametrize("n,batch", LARGE_REAL, ids=str)
@pytest.mark.parametrize("inverse", [False, True])
def test_large_real_c2r(torch_dev, oracle, n, batch, inverse):
    rng = np.random.default_rng(n + 7 * batch)
    h = n // 2 + 1
    x = (rng.uniform(-1, 1, (batch, h)) + 1j * rng.uniform(-1, 1, (batch, h))).astype(np.complex64)
    got = run_gpu(torch_dev, "c2r", (n,), batch, inverse, x)
    want = oracle.c2r_nd(x, (n,), inverse)
    assert got.shape == want.shape
    err = per_row_rel(got, want, 1).max()
    assert err <= tol_for((n,)), (err, tol_for((n,)))


# 2D plans whose row pass takes the columns' first radix-4 stage (rows in fours, quarter-length
# columns in place on the output) where instances exist, else the paired or plain plans
QUAD_2D = [((2048, 2048), 1, True), ((1024, 1024), 3, False), ((2048, 1024), 2, True), ((1024, 2048), 2, False),
           ((1024, 512), 4, True), ((64, 256), 5, False)]


@pytest.mark.parametrize("dims,batch,inverse", QUAD_2D, ids=str)
def test_quad_2d_plans(torch_dev, oracle, monkeypatch, dims, batch, inverse):
    import torch

    from sip_amd.fft import Plan
    rng = np.random.default_rng(sum(dims) + batch)
    x = _uniform_c(rng, (batch,) + dims)
    want = oracle.c2c_nd(x, 2, inverse)
    outs = []
    # rows in fours through one LDS plane (variant 24, the default), through two (variant 23),
    # and the paired / plain plans
    for mode, split, expect in (("1", "1", "variant=24"), ("1", "0", "variant=23"), ("0", "1", None)):
        monkeypatch.setenv("ASDFFT_QUAD2D", mode)
        monkeypatch.setenv("ASDFFT_QUAD_SPLIT", split)
        p = Plan("c2c", dims, batch, inverse)
        monkeypatch.delenv("ASDFFT_QUAD2D")
        monkeypatch.delenv("ASDFFT_QUAD_SPLIT")
        try:
            quad = "variant=23" in p.describe() or "variant=24" in p.describe()
            if expect is None:
                assert not quad
            elif dims[0] in (1024, 2048) and dims[1] in (1024, 2048):
                assert expect in p.describe(), p.describe()
            got = p.execute(torch.from_numpy(x).to(torch_dev)).cpu().numpy()
        finally:
            p.close()
        err = per_row_rel(got, want, 2)
        assert err.max() <= tol_for(dims), (mode, split, dims, batch, inverse, err)
        outs.append(got)
    for other in outs[1:]:
        assert per_row_rel(outs[0], other.astype(np.complex128), 2).max() <= 1e-6
