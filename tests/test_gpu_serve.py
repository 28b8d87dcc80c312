"""GPU tests of the serving loop (asdFftExtServe*).

Every served result must meet the accuracy bound against the FP64 oracle (relative L2
<= 1e-6 * log2 N per transform), and the same request served alone and in a mix of
concurrent requests must agree to that bound.  (A request alone runs through the plain
kernels; in a mix, small 1D requests share one ragged launch.  Both evaluate the same
radix schedule, so the results typically differ by a few 1e-8.)
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def tol(n):
    return 1e-6 * max(1.0, math.log2(max(n, 2)))


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from sip_amd import _native
    _native.lib()
    return torch.device("cuda:0")


# (kind, dims, batch, inverse): ragged-path 1D, a four-step 1D, an odd-prime 1D, 2D
REQUESTS = [
    ("c2c", (4096,), 8, False), ("r2c", (1000,), 5, False), ("c2c", (65536,), 2, False),
    ("c2r", (8192,), 3, True), ("c2c", (256, 512), 2, False), ("c2c", (1024, 1024), 1, True),
    ("r2c", (243,), 7, True), ("c2c", (8191,), 1, False), ("c2c", (16384,), 4, True),
    ("r2c", (64, 100), 2, False), ("c2c", (6000,), 16, False), ("c2r", (75,), 9, False),
]


def host_input(rng, kind, dims, batch):
    shape = (batch,) + tuple(dims)
    if kind == "r2c":
        return rng.uniform(-1, 1, shape).astype(np.float32)
    if kind == "c2r":
        shape = (batch,) + tuple(dims[:-1]) + (dims[-1] // 2 + 1,)
    return (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(np.complex64)


def out_shape(kind, dims, batch):
    if kind == "r2c":
        return (batch,) + tuple(dims[:-1]) + (dims[-1] // 2 + 1,)
    return (batch,) + tuple(dims)


def oracle_of(oracle, kind, dims, x, inverse):
    if kind == "c2c":
        return oracle.c2c_nd(x, len(dims), inverse)
    if kind == "r2c":
        return oracle.r2c_nd(x, len(dims), inverse)
    return oracle.c2r_nd(x, dims, inverse)


def make_io(torch, dev, rng):
    ios = []
    for kind, dims, batch, inv in REQUESTS:
        x = host_input(rng, kind, dims, batch)
        dt = torch.float32 if kind == "c2r" else torch.complex64
        ios.append((torch.from_numpy(x).to(dev), torch.empty(out_shape(kind, dims, batch), dtype=dt, device=dev), x))
    return ios


def test_served_results_match_oracle_and_alone(torch_dev, oracle):
    import torch

    from oracle.oracle import rel_l2_per_row
    from sip_amd.fft import Server
    rng = np.random.default_rng(21)
    ios = make_io(torch, torch_dev, rng)
    srv = Server(hbm_budget=2 << 30, slots=3)
    try:
        # alone: one request in flight at a time
        alone = []
        for (kind, dims, batch, inv), (xd, yd, _) in zip(REQUESTS, ios):
            lat = srv.wait(srv.submit(kind, dims, batch, inv, xd, yd))
            assert lat > 0
            alone.append(yd.cpu().clone())
        # mixed: everything submitted at once, twice in a row
        for rep in range(2):
            for _, yd, _ in ios:
                yd.fill_(float("nan") if yd.dtype == torch.float32 else complex(float("nan"), 0))
            torch.cuda.synchronize()
            tickets = [srv.submit(k, d, b, i, xd, yd) for (k, d, b, i), (xd, yd, _) in zip(REQUESTS, ios)]
            for t in tickets:
                srv.wait(t)
            for (kind, dims, batch, inv), (xd, yd, x), a in zip(REQUESTS, ios, alone):
                got = yd.cpu()
                n = int(np.prod(dims))
                diff = rel_l2_per_row(got.numpy(), a.numpy().astype(np.complex128 if a.is_complex() else np.float64),
                                      len(dims))
                assert diff.max() <= tol(n), \
                    f"{kind} {dims} x{batch}: mixed result differs from the request served alone (pass {rep})"
        for (kind, dims, batch, inv), (_, _, x), a in zip(REQUESTS, ios, alone):
            gold = oracle_of(oracle, kind, dims, x, inv)
            err = rel_l2_per_row(a.numpy(), gold, len(dims)).max()
            n = int(np.prod(dims))
            assert err <= tol(n), f"{kind} {dims}: rel L2 {err:.3e} > {tol(n):.3e}"
        st = srv.stats()
        assert st["completed"] == 3 * len(REQUESTS)
        assert st["ragged_requests"] > 0 and st["plain_requests"] > 0
    finally:
        srv.close()


def test_budget_bounds_inflight_bytes(torch_dev):
    """A budget that holds only a few requests at once: everything still completes and the
    bytes in flight never exceed what the budget leaves after the resident state."""
    import torch

    from sip_amd.fft import Server
    n, batch = 4096, 64
    req_bytes = 2 * 8 * n * batch
    x = torch.randn(batch, n, dtype=torch.complex64, device=torch_dev)
    outs = [torch.empty_like(x) for _ in range(24)]
    probe = Server(hbm_budget=1 << 30, slots=2)
    resident = probe.stats()["resident_bytes"]
    probe.close()
    budget = resident + 3 * req_bytes + 1024
    srv = Server(hbm_budget=budget, slots=2)
    try:
        tickets = [srv.submit("c2c", (n,), batch, False, x, y) for y in outs]
        for t in tickets:
            srv.wait(t)
        st = srv.stats()
        assert st["completed"] == len(outs)
        assert st["peak_inflight_bytes"] <= budget - st["resident_bytes"]
        assert st["batches"] >= len(outs) // 3
        ref = outs[0].cpu()
        for y in outs[1:]:  # ragged (shared batch) and plain (alone) launches agree to the bound
            assert float((y.cpu() - ref).abs().max() / ref.abs().max()) <= tol(n)
    finally:
        srv.close()


def test_per_request_mode_matches_batched(torch_dev):
    import torch

    from sip_amd.fft import Server
    rng = np.random.default_rng(22)
    ios = make_io(torch, torch_dev, rng)
    results = []
    for batching in (True, False):
        srv = Server(batching=batching)
        try:
            tickets = [srv.submit(k, d, b, i, xd, yd) for (k, d, b, i), (xd, yd, _) in zip(REQUESTS, ios)]
            for t in tickets:
                srv.wait(t)
            results.append([yd.cpu().clone() for _, yd, _ in ios])
        finally:
            srv.close()
    for (kind, dims, batch, inv), a, b in zip(REQUESTS, *results):
        num = (a.double() - b.double()).norm() if a.dtype == torch.float32 else (a.cdouble() - b.cdouble()).norm()
        den = b.double().norm() if b.dtype == torch.float32 else b.cdouble().norm()
        n = int(np.prod(dims))
        assert float(num / den) <= tol(n), f"{kind} {dims}: batched and per-request results differ"


def test_replay_reports_latencies(torch_dev):
    import torch

    from sip_amd.fft import Server
    x = torch.randn(16, 2048, dtype=torch.complex64, device=torch_dev)
    ys = [torch.empty_like(x) for _ in range(64)]
    srv = Server()
    try:
        reqs = [Server.request("c2c", (2048,), 16, False, x.data_ptr(), y.data_ptr()) for y in ys]
        arrivals = [20.0 * i for i in range(len(reqs))]  # 50k requests per second
        lat = srv.replay(reqs, arrivals)
        assert all(v > 0 for v in lat)
        ref = ys[0].cpu()
        for y in ys[1:]:
            assert float((y.cpu() - ref).abs().max() / ref.abs().max()) <= tol(2048)
    finally:
        srv.close()
