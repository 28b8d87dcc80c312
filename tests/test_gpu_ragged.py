"""GPU parity of ragged batches (asdFftExtMakePlanRagged / asdFftExtExecRagged).

Every transform of a ragged batch must match the FP64 oracle to the same bound as
the plain path (relative L2 <= 1e-6 * log2 N per transform) and must agree with the
plain plan (asdFftMakePlan1D + exec) of its own length and kind.  Graph replays must
reproduce the stream exec bit for bit (a tile computes the same arithmetic whichever
workgroup takes it from the queue).
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


def make_input(rng, kind, n, count):
    h = n // 2 + 1
    if kind == "c2c":
        return (rng.uniform(-1, 1, (count, n)) + 1j * rng.uniform(-1, 1, (count, n))).astype(np.complex64)
    if kind == "r2c":
        return rng.uniform(-1, 1, (count, n)).astype(np.float32)
    return (rng.uniform(-1, 1, (count, h)) + 1j * rng.uniform(-1, 1, (count, h))).astype(np.complex64)


def oracle_of(oracle, kind, n, x, inverse):
    if kind == "c2c":
        return oracle.c2c(x, inverse)
    if kind == "r2c":
        return oracle.r2c(x, inverse)
    return oracle.c2r(x, n, inverse)


def pack(plan, inputs):
    buf = np.zeros(plan.in_extent, dtype=np.uint8)
    for (kind, n, count, inv, i0, o0), x in zip(plan.entries, inputs):
        raw = np.ascontiguousarray(x).view(np.uint8).reshape(-1)
        buf[i0:i0 + raw.size] = raw
    return buf


def run_ragged(torch_dev, plan, inputs, out_fill=0):
    import torch
    xin = torch.from_numpy(pack(plan, inputs)).to(torch_dev)
    out = torch.full((plan.out_extent,), out_fill, dtype=torch.uint8, device=torch_dev)
    plan.execute(xin, out)
    torch.cuda.synchronize()
    return [v.cpu().numpy() for v in plan.views(out, "out")]


def check_entries(oracle, entries, inputs, outs):
    worst = 0.0
    for (kind, n, count, inv, *_), x, got in zip(entries, inputs, outs):
        gold = oracle_of(oracle, kind, n, x, inv)
        from oracle.oracle import rel_l2_per_row
        err = rel_l2_per_row(got, gold).max()
        assert err <= tol(n), f"{kind} n={n} inverse={inv}: rel L2 {err:.3e} > {tol(n):.3e}"
        worst = max(worst, err / tol(n))
    return worst


MIXED = [
    ("c2c", 1000, 3, False), ("c2c", 4096, 5, True), ("r2c", 8192, 2, False), ("r2c", 81, 7, True),
    ("c2r", 4096, 3, True), ("c2r", 75, 4, False), ("c2c", 16384, 2, False), ("c2c", 12000, 1, True),
    ("c2c", 64, 40, False), ("r2c", 16384, 2, False), ("c2r", 20000, 1, True), ("c2c", 1001, 2, False),
    ("c2c", 323, 3, True), ("r2c", 2, 3, False), ("c2c", 2, 5, False), ("c2c", 3, 4, True),
    ("r2c", 6000, 3, False), ("c2r", 6000, 2, False), ("r2c", 2187, 2, True), ("c2r", 2187, 3, False),
    ("c2c", 1000, 2, False),  # second entry of an existing class
]


def test_mixed_batch_matches_oracle(torch_dev, oracle):
    from sip_amd.fft import RaggedPlan
    rng = np.random.default_rng(11)
    plan = RaggedPlan(MIXED)
    try:
        inputs = [make_input(rng, k, n, c) for k, n, c, _ in MIXED]
        outs = run_ragged(torch_dev, plan, inputs)
        check_entries(oracle, plan.entries, inputs, outs)
    finally:
        plan.close()


def test_mixed_batch_matches_plain_plans(torch_dev):
    import torch

    from sip_amd.fft import Plan, RaggedPlan
    rng = np.random.default_rng(12)
    plan = RaggedPlan(MIXED)
    try:
        inputs = [make_input(rng, k, n, c) for k, n, c, _ in MIXED]
        outs = run_ragged(torch_dev, plan, inputs)
        for (kind, n, count, inv), x, got in zip(MIXED, inputs, outs):
            p = Plan(kind, (n,), count, inv)
            try:
                ref = p.execute(torch.from_numpy(x).to(torch_dev))
                torch.cuda.synchronize()
                ref = ref.cpu().numpy()
            finally:
                p.close()
            num = np.linalg.norm((got - ref).astype(np.complex128).reshape(count, -1), axis=1)
            den = np.linalg.norm(ref.astype(np.complex128).reshape(count, -1), axis=1)
            assert (num <= tol(n) * np.maximum(den, 1e-30)).all(), f"{kind} n={n}: ragged differs from plain plan"
    finally:
        plan.close()


def test_scattered_entries_and_repeat_exec(torch_dev, oracle):
    """One entry per transform in shuffled buffer order; exec twice (the tile queue resets)."""
    from sip_amd.fft import RaggedPlan, _side_bytes
    rng = np.random.default_rng(13)
    shapes = [("c2c", n) for n in (64, 96, 1000, 5000, 16384, 729)] + [("r2c", n) for n in (100, 2048, 375, 16000)]
    order = rng.permutation(len(shapes))
    entries, ioff, ooff = [None] * len(shapes), 0, 0
    for j in order:
        kind, n = shapes[j]
        ib, ob = _side_bytes(kind, n)
        entries[j] = (kind, n, 1, False, ioff, ooff)
        ioff += (ib + 7) // 8 * 8 + 64
        ooff += (ob + 7) // 8 * 8 + 64
    plan = RaggedPlan(entries)
    try:
        inputs = [make_input(rng, k, n, 1) for k, n in shapes]
        first = run_ragged(torch_dev, plan, inputs)
        check_entries(oracle, plan.entries, inputs, first)
        second = run_ragged(torch_dev, plan, inputs, out_fill=255)
        for a, b in zip(first, second):
            assert np.array_equal(a, b)
    finally:
        plan.close()


def test_in_place_c2c(torch_dev, oracle):
    import torch

    from sip_amd.fft import RaggedPlan
    rng = np.random.default_rng(14)
    spec = [("c2c", 4096, 3, False), ("c2c", 750, 5, True), ("c2c", 16384, 1, False)]
    entries, off = [], 0
    for kind, n, count, inv in spec:
        entries.append((kind, n, count, inv, off, off))
        off += 8 * n * count
    plan = RaggedPlan(entries)
    try:
        inputs = [make_input(rng, k, n, c) for k, n, c, _ in spec]
        buf = torch.from_numpy(pack(plan, inputs)).to(torch_dev)
        plan.execute(buf, buf)
        torch.cuda.synchronize()
        outs = [v.cpu().numpy() for v in plan.views(buf, "out")]
        check_entries(oracle, plan.entries, inputs, outs)
    finally:
        plan.close()


def test_graph_replays_match_stream_exec(torch_dev):
    """Graph replays on two buffer pairs (node parameters updated between them), with
    eager work and synchronisations in between; every replay equals the stream exec."""
    import torch

    from sip_amd.fft import RaggedPlan
    rng = np.random.default_rng(15)
    eager = RaggedPlan(MIXED)
    graph = RaggedPlan(MIXED, graph=True)
    try:
        sets = []
        for _ in range(2):
            inputs = [make_input(rng, k, n, c) for k, n, c, _ in MIXED]
            sets.append(torch.from_numpy(pack(eager, inputs)).to(torch_dev))
        refs = []
        for xin in sets:
            out = torch.zeros(eager.out_extent, dtype=torch.uint8, device=torch_dev)
            eager.execute(xin, out)
            torch.cuda.synchronize()
            refs.append([v.cpu() for v in eager.views(out, "out")])
        outs = [torch.zeros(graph.out_extent, dtype=torch.uint8, device=torch_dev) for _ in sets]
        for rep in range(3):
            for i, xin in enumerate(sets):
                outs[i].fill_(rep + 1)  # alignment gaps between entries keep this fill
                graph.execute(xin, outs[i])
                junk = torch.randn(1 << 16, device=torch_dev).sum()  # eager work between replays
                torch.cuda.synchronize()
                assert torch.isfinite(junk)
                for e, (got, ref) in enumerate(zip(graph.views(outs[i], "out"), refs[i])):
                    assert torch.equal(got.cpu().view(torch.uint8), ref.view(torch.uint8)), \
                        f"replay {rep} on buffer set {i}: entry {e} differs"
    finally:
        eager.close()
        graph.close()


def test_ragged_workload_full_size(torch_dev, oracle):
    """The 4000-transform workload (N log-uniform over 5-smooth lengths in [64, 16384],
    70% C2C / 30% R2C), one entry per transform in generation order, graph exec: every
    transform against the oracle."""
    import bench
    from sip_amd.fft import RaggedPlan
    work = bench.ragged_workload()
    rng = np.random.default_rng(16)
    entries = [(k, n, 1, False) for k, n in work]
    plan = RaggedPlan(entries, graph=True)
    try:
        inputs = [make_input(rng, k, n, 1) for k, n in work]
        outs = run_ragged(torch_dev, plan, inputs)
        # group by class so the oracle runs batched
        groups = {}
        for i, (k, n) in enumerate(work):
            groups.setdefault((k, n), []).append(i)
        from oracle.oracle import rel_l2_per_row
        worst = 0.0
        for (k, n), idx in groups.items():
            x = np.concatenate([inputs[i] for i in idx])
            got = np.concatenate([outs[i] for i in idx])
            err = rel_l2_per_row(got, oracle_of(oracle, k, n, x, False)).max()
            assert err <= tol(n), f"{k} n={n}: rel L2 {err:.3e}"
            worst = max(worst, err / tol(n))
        print(f"ragged workload: {len(work)} transforms, {len(groups)} classes, worst err/tol {worst:.3f}")
    finally:
        plan.close()


def test_autotuned_plan_matches_oracle(torch_dev, oracle):
    """Schedules picked by the plan-time autotuner (other radix orders, more threads per
    transform, another workgroup size) still meet the accuracy bound."""
    from sip_amd.fft import RaggedPlan
    rng = np.random.default_rng(17)
    spec = [("c2c", 4096, 6, False), ("c2c", 6000, 3, True), ("r2c", 8192, 4, False), ("c2r", 1000, 5, True),
            ("c2c", 16384, 2, False), ("r2c", 243, 9, False), ("c2c", 96, 30, True)]
    plan = RaggedPlan(spec, autotune=True)
    try:
        assert "autotuned:" in plan.describe()
        inputs = [make_input(rng, k, n, c) for k, n, c, _ in spec]
        outs = run_ragged(torch_dev, plan, inputs)
        check_entries(oracle, plan.entries, inputs, outs)
    finally:
        plan.close()


def test_launch_forms_match_oracle(torch_dev, oracle):
    """Every launch form meets the bound, C2C / R2C / C2R, both directions, lengths from every
    budget: the default single launch at the largest budget (classes scheduled for a smaller
    budget run with idle butterfly slots), wide EMAX-16 workgroups of up to 1024 threads with the
    classes that need a larger budget in one more launch (ASDFFT_RAGGED_WIDE16=1), and the
    per-budget form (ASDFFT_RAGGED_PER_BUDGET)."""
    from sip_amd.fft import RaggedPlan
    rng = np.random.default_rng(18)
    spec = [("c2c", 64, 40, False), ("c2c", 4096, 3, True), ("r2c", 8192, 4, False), ("c2r", 1000, 5, True),
            ("c2c", 16384, 2, False), ("r2c", 243, 9, True), ("c2c", 14400, 2, True), ("c2r", 6000, 3, False)]
    import os
    inputs = [make_input(rng, k, n, c) for k, n, c, _ in spec + [("c2c", 15625, 2, False)]]
    spec = spec + [("c2c", 15625, 2, False)]  # 5^6: no EMAX-16 schedule within 1024 threads
    # default: one launch; wide EMAX-16 workgroups + one launch at EMAX 48 for 15625; per
    # budget: EMAX 16 / 32 / 48
    for wide, per_budget, launches in (("0", False, "1 launch(es)"), ("1", False, "2 launch(es)"),
                                       (None, True, "3 launch(es)")):
        if wide is not None:
            os.environ["ASDFFT_RAGGED_WIDE16"] = wide
        try:
            plan = RaggedPlan(spec, per_budget=per_budget)
        finally:
            os.environ.pop("ASDFFT_RAGGED_WIDE16", None)
        try:
            assert launches in plan.describe(), plan.describe()
            outs = run_ragged(torch_dev, plan, inputs)
            check_entries(oracle, plan.entries, inputs, outs)
        finally:
            plan.close()
