#!/usr/bin/env python3
"""Batched FFT throughput on one MI355X, through the asdFft C ABI.

One step = one pass over the plain-FP32 coverage workload, all forward except
where noted, inputs resident in HBM:
    1D C2C complex64 N in {1000, 2187, 6000, 4096}, batch 256 each
    1D R2C float32 N=8192 batch 1024, then C2R (inverse) of that spectrum
    2D C2C 2048 x 2048, batch 4
    3D C2C 256^3
Inputs are synthetic uniform[-1, 1] (the R2C input is Gaussian noise plus three
random tones per row), generated on the device from fixed seeds.

value = aggregate GFLOP/s over the step, FLOP = 5 N log2 N per transform
(N = total points; the same formula for R2C/C2R, whose conventional 2.5 N log2 N
figure is reported beside it).  The step is timed with HIP events on the stream
the plans launch on, bracketed by device synchronisation.  Each configuration
and each kernel launch is also timed per sample, as the reference's benchmark
does (start event, exec, end event; --samples of them, median), with the stream
held by a spin kernel while each sample is enqueued so that the interval is
device time.  The JSON line carries the roofline of the dominant kernel launch
(algorithmic bytes / its mean per-sample duration vs 8 TB/s HBM), each
configuration's relative L2 error against a float64 FFT on sampled transforms,
a device-copy ceiling of the same bytes, and the reference's CPU golden (numpy
FFT as its golden generators call it) timed on the host.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "batched FFT GFLOP/s + % of HBM roofline; served FFT p50/p99 latency"

# (name, kind, dims, batch, inverse, seed)
WORKLOAD = [
    ("c2c_n1000_b256", "c2c", (1000,), 256, False, 1),
    ("c2c_n2187_b256", "c2c", (2187,), 256, False, 2),
    ("c2c_n6000_b256", "c2c", (6000,), 256, False, 3),
    ("c2c_n4096_b256", "c2c", (4096,), 256, False, 4),
    ("r2c_n8192_b1024", "r2c", (8192,), 1024, False, 5),
    ("c2r_n8192_b1024", "c2r", (8192,), 1024, True, 5),
    ("c2c2d_2048x2048_b4", "c2c", (2048, 2048), 4, False, 6),
    ("c2c3d_256x256x256_b1", "c2c", (256, 256, 256), 1, False, 7),
]


def smooth_lengths(lo=64, hi=16384):
    """Sorted {2^a 3^b 5^c} in [lo, hi]."""
    out = []
    p2 = 1
    while p2 <= hi:
        p3 = p2
        while p3 <= hi:
            p5 = p3
            while p5 <= hi:
                if p5 >= lo:
                    out.append(p5)
                p5 *= 5
            p3 *= 3
        p2 *= 2
    return sorted(out)


def ragged_workload(seed=10, total=4000, c2c_frac=0.7):
    """The ragged-batch workload: `total` transforms, N log-uniform over {2^a 3^b 5^c} in
    [64, 16384] (a uniform draw of log N snapped to the nearest such length in log
    space), 70% forward C2C complex64 / 30% forward R2C float32.  Returns [(kind, n)]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    lens = np.array(smooth_lengths(), dtype=np.int64)
    logs = np.log(lens.astype(np.float64))
    u = rng.uniform(math.log(64), math.log(16384), total)
    idx = np.abs(logs[None, :] - u[:, None]).argmin(axis=1)
    kinds = np.where(rng.uniform(size=total) < c2c_frac, "c2c", "r2c")
    return [(str(k), int(n)) for k, n in zip(kinds, lens[idx])]


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "stockham_plain_fp32_coverage_hbm_counters.md")


def measured_traffic():
    """HBM bytes per launch from the committed counter summary (two rocprofv3 --pmc runs,
    FETCH_SIZE and WRITE_SIZE, scaled by a copy of known size): {"<workload> launch <i>": bytes}."""
    out = {}
    try:
        for line in open(TRAFFIC_PROFILE):
            if line.startswith("traffic_bytes "):
                key, val = line[len("traffic_bytes "):].rsplit(":", 1)
                out[key.strip()] = int(val)
    except OSError:
        pass
    return out


def flops(dims, batch):
    n = math.prod(dims)
    return 5.0 * n * math.log2(n) * batch


def alg_bytes(kind, dims, batch):
    n = math.prod(dims)
    if kind == "c2c":
        return 16.0 * n * batch
    half = math.prod(dims[:-1]) * (dims[-1] // 2 + 1)
    return (4.0 * n + 8.0 * half) * batch


def make_inputs(torch, dev, kind, dims, batch, seed):
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    shape = (batch,) + tuple(dims)
    if kind == "c2c":
        re = torch.rand(shape, generator=g, device=dev) * 2 - 1
        im = torch.rand(shape, generator=g, device=dev) * 2 - 1
        return torch.complex(re, im).contiguous()
    if kind == "r2c":
        n = dims[-1]
        x = torch.randn(shape, generator=g, device=dev, dtype=torch.float64)
        t = torch.arange(n, device=dev, dtype=torch.float64)
        for _ in range(3):
            k = torch.randint(1, n // 2, (batch, 1), generator=g, device=dev).double()
            a = torch.rand((batch, 1), generator=g, device=dev, dtype=torch.float64) * 3.5 + 0.5
            ph = torch.rand((batch, 1), generator=g, device=dev, dtype=torch.float64) * 2 * math.pi
            x = x + a * torch.cos(2 * math.pi * k * t / n + ph)
        return x.float().contiguous()
    raise ValueError(kind)


# per-launch timing rotates over copies of the (input, output) pair such that each of the two
# alone spans at least this many bytes, more than twice the MI355X's 256 MB device cache (MALL):
# an in-place step touches only the output, so sizing by the pair would leave its working set
# in the cache
RING_BYTES = 512 << 20
RING_MAX = 256


def ring_copies(x, y):
    smaller = min(x.numel() * x.element_size(), y.numel() * y.element_size())
    return max(1, min(RING_MAX, -(-RING_BYTES // smaller)))


def event_time_ms(torch, stream, fn, reps):
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    fn()  # untimed: first launches load their code objects
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(reps):
        fn()
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / reps


_SLEEP_CYCLES_PER_US = []


def _spin_cycles(torch, stream, us):
    """torch.cuda._sleep cycles that hold `stream` for about `us` microseconds (calibrated once)."""
    if not _SLEEP_CYCLES_PER_US:
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(stream):
            torch.cuda._sleep(1000)
            torch.cuda.synchronize()
            s.record(stream)
            torch.cuda._sleep(1 << 20)
            e.record(stream)
        torch.cuda.synchronize()
        _SLEEP_CYCLES_PER_US.append((1 << 20) / max(s.elapsed_time(e) * 1e3, 1e-3))
    return max(1, int(us * _SLEEP_CYCLES_PER_US[0]))


def device_times_us(torch, stream, fn, samples=300, hold_us=80.0):
    """Per-sample device time of fn() as the reference's benchmark takes it
    (tests/standalone/fft1d_optimized/bench.cpp:176-190: start event, exec, end event, one
    sample each; the median is reported).  Before each sample a spin kernel holds the stream
    for hold_us, longer than the host needs to enqueue the two events and the exec, so the
    events bracket queued work: the interval is device time, not host submission latency.
    Returns {median, mean, p10, p90, n} in microseconds."""
    cyc = _spin_cycles(torch, stream, hold_us)
    with torch.cuda.stream(stream):
        fn()  # untimed: first launches load their code objects
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(samples)]
        for i, (s, e) in enumerate(ev):
            torch.cuda._sleep(cyc)
            s.record(stream)
            fn()
            e.record(stream)
            if i % 64 == 63:
                torch.cuda.synchronize()
    torch.cuda.synchronize()
    t = sorted(s.elapsed_time(e) * 1e3 for s, e in ev)
    n = len(t)
    return {"median": t[n // 2] if n % 2 else 0.5 * (t[n // 2 - 1] + t[n // 2]), "mean": sum(t) / n,
            "p10": t[n // 10], "p90": t[(9 * n) // 10], "n": n}


def median_ms(torch, stream, fn, samples=300):
    """Median per-sample device time of fn() in milliseconds (device_times_us)."""
    return device_times_us(torch, stream, fn, samples)["median"] * 1e-3


def copy_ceiling_us(torch, stream, x, y, samples=300):
    """Median device time of a plain device-to-device copy moving the bytes of one exec of a
    transform from x to y (torch's copy kernel, (|x| + |y|) / 2 bytes read and as many written,
    so read + write equal the transform's algorithmic bytes), rotated over a ring larger than
    the device cache like the transform itself."""
    half = (x.numel() * x.element_size() + y.numel() * y.element_size()) // 2
    half -= half % 16
    copies = max(1, min(RING_MAX, -(-RING_BYTES // half)))
    srcs = [torch.empty(half, dtype=torch.uint8, device=x.device) for _ in range(copies)]
    dsts = [torch.empty(half, dtype=torch.uint8, device=x.device) for _ in range(copies)]
    for t in srcs:
        t.fill_(1)
    cur = [0]

    def run():
        i = cur[0] = (cur[0] + 1) % copies
        dsts[i].copy_(srcs[i])
    us = device_times_us(torch, stream, run, samples)["median"]
    del srcs, dsts
    return us


def sampled_error(torch, kind, dims, batch, inverse, x, y, rows=16):
    """Relative L2 error per transform of y = exec(x) against numpy's float64 FFT (the
    reference's golden arithmetic in double precision) on sampled transforms, outside any timed
    region: up to `rows` evenly spaced transforms for 1D (first and last included), the first and
    last slice for 2D, the whole volume for 3D.  Returns max and RMS over the sample and the
    bound 1e-6 log2 N (N = points per transform)."""
    import numpy as np
    n = math.prod(dims)
    if len(dims) == 1:
        idx = sorted(set(int(round(v)) for v in np.linspace(0, batch - 1, min(rows, batch))))
    elif len(dims) == 2:
        idx = sorted({0, batch - 1})
    else:
        idx = list(range(batch))
    xs = x[idx].cpu().numpy()
    ys = y[idx].cpu().numpy()
    axes = tuple(range(1, len(dims) + 1))
    if kind == "c2c":
        xd = xs.astype(np.complex128)
        ref = np.fft.ifftn(xd, axes=axes, norm="forward") if inverse else np.fft.fftn(xd, axes=axes)
    elif kind == "r2c":
        ref = np.fft.rfftn(xs.astype(np.float64), axes=axes)
        if inverse:
            ref = np.conj(ref)
    else:
        ref = np.fft.irfftn(xs.astype(np.complex128), s=dims, axes=axes, norm="forward")
        if not inverse:
            ref = np.fft.hfft(xs.astype(np.complex128), n=dims[-1])
    err = np.linalg.norm((ys.astype(ref.dtype) - ref).reshape(len(idx), -1), axis=1) / \
        np.linalg.norm(ref.reshape(len(idx), -1), axis=1)
    bound = 1e-6 * math.log2(n)
    return {"max": float(f"{err.max():.3g}"), "rms": float(f"{math.sqrt(float(np.mean(err ** 2))):.3g}"),
            "bound": float(f"{bound:.3g}"), "within_bound": bool(err.max() <= bound), "transforms": len(idx),
            "vs": "numpy float64 FFT"}


def cpu_oracle(seconds_budget=20.0):
    """FP64 oracle (single-threaded C) on a bounded sample of the same workload."""
    import numpy as np

    from oracle import oracle
    oracle.build()
    rng = np.random.default_rng(0)
    # sample: 8 transforms of each 1D shape, one 2D 2048^2 slice, one 256^3 volume
    sample = [("c2c", (1000,), 8), ("c2c", (2187,), 8), ("c2c", (6000,), 8), ("c2c", (4096,), 8),
              ("r2c", (8192,), 8), ("c2r", (8192,), 8), ("c2c", (2048, 2048), 1), ("c2c", (256, 256, 256), 1)]
    total_flop = 0.0
    t0 = time.perf_counter()
    for kind, dims, b in sample:
        shape = (b,) + dims
        if kind == "c2c":
            x = (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(np.complex64)
            oracle.c2c_nd(x, len(dims), False)
        elif kind == "r2c":
            oracle.r2c(rng.standard_normal(shape).astype(np.float32))
        else:
            h = (b, dims[-1] // 2 + 1)
            x = (rng.standard_normal(h) + 1j * rng.standard_normal(h)).astype(np.complex64)
            oracle.c2r(x, dims[-1], True)
        total_flop += flops(dims, b)
        if time.perf_counter() - t0 > seconds_budget:
            break
    dt = time.perf_counter() - t0
    return {"value": round(total_flop / dt / 1e9, 3), "unit": "GFLOP/s", "cores": 1, "kind": "port",
            "sample": "FP64 C oracle, 1 thread: 8 transforms each of C2C N=1000/2187/6000/4096, R2C and C2R "
                      "N=8192, plus one 2048x2048 2D and one 256^3 3D transform; "
                      f"{total_flop / 1e9:.2f} GFLOP in {dt:.1f} s"}


def cpu_numpy(max_threads=16, runs=5):
    """numpy.fft on the host cores over one full step of the coverage workload, two ways, each
    timed as the median of `runs` runs after one untimed warm-up:
    golden: as the reference's unit-test golden generators call it, on one thread: one
      np.fft.fft / rfft / irfft(norm="forward") per row, cast to complex64
      (tests/ut/unittest/fft/1d/c2c/c2c_data/gen_data.py:42-65, r2c_data/gen_data.py:43,
      c2r_data/gen_data.py:48), and one np.fft.fft2 / fftn over the whole 2D / 3D array
      (2d/c2c/c2c_2d_data/gen_data_2d.py:65, 3d/c2c/c2c_3d_data/gen_data_3d.py:44);
    threads: the same transforms as batched numpy calls over chunks on a thread pool (numpy's
      FFT releases the GIL), one axis at a time for 2D / 3D, each split along the largest other
      axis."""
    import concurrent.futures as cf

    import numpy as np
    rng = np.random.default_rng(0)
    threads = max(1, min(max_threads, len(os.sched_getaffinity(0))))
    data = []
    for name, kind, dims, batch, inv, _ in WORKLOAD:
        shape = (batch,) + dims
        if kind == "c2c":
            x = (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(np.complex64)
        elif kind == "r2c":
            x = rng.standard_normal(shape).astype(np.float32)
        else:
            x = np.fft.rfft(rng.standard_normal(shape)).astype(np.complex64)
        data.append((kind, dims, batch, x))

    def golden():
        for kind, dims, batch, x in data:
            n = dims[-1]
            if len(dims) == 2:
                np.fft.fft2(x, s=dims).astype(np.complex64)
            elif len(dims) == 3:
                np.fft.fftn(x, axes=(-3, -2, -1)).astype(np.complex64)
            elif kind == "c2c":
                for b in range(batch):
                    np.fft.fft(x[b], n=n).astype(np.complex64)
            elif kind == "r2c":
                for b in range(batch):
                    np.fft.rfft(x[b], n=n).astype(np.complex64)
            else:
                for b in range(batch):
                    np.fft.irfft(x[b], n=n, norm="forward").astype(np.float32)

    def pooled(ex):
        for kind, dims, batch, x in data:
            n = dims[-1]
            if kind != "c2c":
                f = (lambda a: np.fft.rfft(a, axis=-1).astype(np.complex64)) if kind == "r2c" else \
                    (lambda a: np.fft.irfft(a, n=n, axis=-1, norm="forward").astype(np.float32))
                list(ex.map(f, np.array_split(x, threads * 4)))
                continue
            y = x.copy()
            for ax in range(y.ndim - 1, 0, -1):
                split = max((d for d in range(y.ndim) if d != ax), key=lambda d: y.shape[d])
                bounds = np.linspace(0, y.shape[split], min(y.shape[split], 4 * threads) + 1).astype(int)

                def work(lo_hi, ax=ax, split=split):
                    idx = [slice(None)] * y.ndim
                    idx[split] = slice(*lo_hi)
                    y[tuple(idx)] = np.fft.fft(y[tuple(idx)], axis=ax)
                list(ex.map(work, zip(bounds[:-1], bounds[1:])))

    total_flop = sum(flops(dims, batch) for _, dims, batch, _ in data)

    def median_s(f):
        f()  # untimed warm-up
        ts = []
        for _ in range(runs):
            t0 = time.perf_counter()
            f()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[len(ts) // 2], ts

    t_golden, ts_golden = median_s(golden)
    with cf.ThreadPoolExecutor(threads) as ex:
        t_pool, ts_pool = median_s(lambda: pooled(ex))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    golden_gflops = total_flop / t_golden / 1e9
    return {"baseline": {"value": round(golden_gflops, 3), "unit": "GFLOP/s", "cores": 1, "kind": "reference",
                         "sample": "the reference's CPU golden: numpy.fft called as its golden generators call it "
                                   "(one np.fft.fft / rfft / irfft(norm='forward') per row cast to complex64, "
                                   "gen_data.py:42-65; fft2 / fftn for 2D / 3D) over one full step of the same "
                                   f"workload (same shapes and batches, {total_flop / 1e9:.2f} GFLOP, 5 N log2 N), "
                                   f"one thread; median of {runs} runs after one warm-up; numpy {np.__version__}; "
                                   f"{model}",
                         "s_per_step": round(t_golden, 3), "runs_s": [round(x, 3) for x in ts_golden],
                         "cpu_model": model, "numpy": np.__version__},
            "threads": {"gflops": round(total_flop / t_pool / 1e9, 3), "s_per_step": round(t_pool, 3),
                        "threads": threads, "runs_s": [round(x, 3) for x in ts_pool]},
            "cpu_model": model, "cpus_visible": os.cpu_count(), "numpy": np.__version__,
            "sample": "one full step of the coverage workload (same shapes and batches), seeded numpy inputs"}


def bench_ragged(torch, dev, stream, reps, warmup):
    """Ragged-batch form: the 4000-transform workload as one persistent launch per
    register budget (stream exec and hipGraph replay) against one plain batched plan per
    distinct (kind, N) (host loop of C-ABI execs, and the same loop captured in a torch
    CUDA graph).  All variants transform the same bytes; inputs resident in HBM."""
    import collections

    from sip_amd import _native as N
    from sip_amd.fft import Plan, RaggedPlan
    L = N.lib()
    work = ragged_workload()
    groups = collections.OrderedDict()
    for k, n in sorted(work, key=lambda kn: (kn[0], kn[1])):
        groups[(k, n)] = groups.get((k, n), 0) + 1
    grouped = [(k, n, c, False) for (k, n), c in groups.items()]
    scattered = [(k, n, 1, False) for k, n in work]
    res = {"transforms": len(work), "classes": len(groups)}
    with torch.cuda.stream(stream):
        p_stream = RaggedPlan(grouped)
        p_graph = RaggedPlan(grouped, graph=True)
        p_scat = RaggedPlan(scattered, graph=True)
        # one concurrent launch per register budget (EMAX 16 / 32 / 48) instead of the default
        # single launch of every class
        p_budget = RaggedPlan(grouped, per_budget=True)
        # the EMAX-16 classes on wide workgroups (up to 1024 threads per transform) in a launch of
        # their own beside the single launch of the others
        os.environ["ASDFFT_RAGGED_WIDE16"] = "1"
        try:
            p_wide = RaggedPlan(grouped)
        finally:
            os.environ.pop("ASDFFT_RAGGED_WIDE16", None)
        res["wide16_describe_head"] = [l for l in p_wide.describe().splitlines() if "launch" in l]
        res["plan_describe_head"] = [l for l in p_stream.describe().splitlines() if "launch" in l]
        res["per_budget_describe_head"] = [l for l in p_budget.describe().splitlines() if "launch" in l]
        t_tune = time.perf_counter()
        p_tuned = RaggedPlan(grouped, graph=True, autotune=True)
        res["autotune_plan_s"] = round(time.perf_counter() - t_tune, 2)
        g = torch.Generator(device=dev)
        g.manual_seed(10)
        xin = (torch.rand(p_stream.in_extent // 4, generator=g, device=dev) * 2 - 1).view(torch.uint8)
        yout = torch.empty(p_stream.out_extent, dtype=torch.uint8, device=dev)
        xs = (torch.rand(p_scat.in_extent // 4, generator=g, device=dev) * 2 - 1).view(torch.uint8)
        ys = torch.empty(p_scat.out_extent, dtype=torch.uint8, device=dev)
        # one plain plan per class over the same packed buffers (each class is contiguous)
        plain = []
        ins, outs = p_stream.views(xin, "in"), p_stream.views(yout, "out")
        for (kind, n, c, inv), xi, yo in zip(grouped, ins, outs):
            pl = Plan(kind, (n,), c, inv)
            pl.bind_stream(stream.cuda_stream)
            dt_in = N.ACL_FLOAT if kind == "r2c" else N.ACL_COMPLEX64
            dt_out = N.ACL_FLOAT if kind == "c2r" else N.ACL_COMPLEX64
            tin = N.AclTensor(xi.shape, dt_in, xi.data_ptr())
            tout = N.AclTensor(yo.shape, dt_out, yo.data_ptr())
            fn = {"c2c": L.asdFftExecC2C, "r2c": L.asdFftExecR2C, "c2r": L.asdFftExecC2R}[kind]
            plain.append((pl, tin, tout, fn))
        torch.cuda.synchronize()

        def run_plain():
            for pl, tin, tout, fn in plain:
                st = fn(pl.handle, tin.ptr, tout.ptr)
                if st != 0:
                    raise RuntimeError(f"plain exec returned {st}")

        variants = {
            "ragged_stream": lambda: p_stream.execute(xin, yout, stream.cuda_stream),
            "ragged_graph": lambda: p_graph.execute(xin, yout, stream.cuda_stream),
            "ragged_graph_scattered": lambda: p_scat.execute(xs, ys, stream.cuda_stream),
            "ragged_graph_autotuned": lambda: p_tuned.execute(xin, yout, stream.cuda_stream),
            "ragged_per_budget_launches": lambda: p_budget.execute(xin, yout, stream.cuda_stream),
            "ragged_wide16_two_launches": lambda: p_wide.execute(xin, yout, stream.cuda_stream),
            "plain_per_size_loop": run_plain,
        }
        # the per-size loop captured once in a CUDA graph: the plain kernels without host overhead
        graph_ok = True
        try:
            run_plain()
            torch.cuda.synchronize()
            cg = torch.cuda.CUDAGraph()
            with torch.cuda.graph(cg, stream=stream):
                run_plain()
            variants["plain_per_size_graph"] = cg.replay
        except Exception as e:  # capture of the plain path is a comparison aid, not the product
            graph_ok = False
            res["plain_graph_error"] = str(e)[:200]
        times = {}
        for name, fn in variants.items():
            for _ in range(warmup):
                fn()
            torch.cuda.synchronize()
            times[name] = median_ms(torch, stream, fn, reps)
        # ragged output equals the per-size plans' output
        run_plain()
        torch.cuda.synchronize()
        y_plain = yout.clone()
        p_graph.execute(xin, yout, stream.cuda_stream)
        torch.cuda.synchronize()
        a = yout.view(torch.float32).double()
        b = y_plain.view(torch.float32).double()
        res["ragged_vs_plain_rel_l2"] = float((a - b).norm() / b.norm())
    res["alg_bytes"] = p_stream.algorithmic_bytes
    res["gflop"] = round(p_stream.flops / 1e9, 4)
    for name, ms in times.items():
        res[name] = {"us": round(ms * 1e3, 2), "gflops": round(p_stream.flops / ms / 1e6, 1),
                     "hbm_frac": round(p_stream.algorithmic_bytes / ms / 1e6 / HBM_PEAK_GBS, 4)}
    best_plain = min(times["plain_per_size_loop"], times.get("plain_per_size_graph", float("inf")))
    best_ragged = min(times["ragged_graph"], times["ragged_graph_autotuned"])
    res["speedup_vs_best_plain"] = round(best_plain / best_ragged, 2)
    res["plain_graph_captured"] = graph_ok
    res["launches"] = p_stream.describe().count("launch emax")
    for pl, tin, tout, fn in plain:
        tin.close()
        tout.close()
        pl.close()
    for p in (p_stream, p_graph, p_scat, p_tuned):
        p.close()
    return res


def serve_workload(seed=11, count=3000):
    """Serving requests: 85% 1D (half C2C with N log-uniform over 5-smooth lengths in
    [256, 65536], half R2C with N in [256, 32768], odd real lengths at most 16384), batch
    uniform in 1..128; 15% 2D C2C with X, Y in {64, ..., 1024}, batch 1.
    Returns [(kind, dims, batch)]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    lens = {"c2c": np.array(smooth_lengths(256, 65536), dtype=np.int64),
            # odd real lengths above 16384 have no single-workgroup kernel
            "r2c": np.array([n for n in smooth_lengths(256, 32768) if n <= 16384 or n % 2 == 0], dtype=np.int64)}
    out = []
    for _ in range(count):
        if rng.uniform() < 0.85:
            kind = "c2c" if rng.uniform() < 0.5 else "r2c"
            cand = lens[kind]
            u = rng.uniform(math.log(256), math.log(cand[-1]))
            n = int(cand[np.abs(np.log(cand.astype(np.float64)) - u).argmin()])
            out.append((kind, (n,), int(rng.integers(1, 129))))
        else:
            x, y = (int(v) for v in rng.choice([64, 128, 256, 512, 1024], 2))
            out.append(("c2c", (x, y), 1))
    return out


def bench_serve(torch, dev, rates=(5000, 20000, 50000, 100000, 200000), count=3000, reps=3):
    """Serving form: a Poisson request stream replayed by the library's native submitter,
    once through the continuous-batching server and once as per-request plain execs (the
    same server with batching off), under an 8 GiB HBM budget.  Latency = completion of a
    request's launch group minus its scheduled arrival."""
    import numpy as np

    from sip_amd.fft import Server
    work = serve_workload(count=count)

    def nbytes(kind, dims, batch):
        pts = math.prod(dims[:-1]) * batch
        h = dims[-1] // 2 + 1
        return {"c2c": (8 * pts * dims[-1], 8 * pts * dims[-1]),
                "r2c": (4 * pts * dims[-1], 8 * pts * h)}[kind]

    max_in = max(nbytes(*w)[0] for w in work)
    src = (torch.rand(max_in // 4 + 64, device=dev) * 2 - 1)
    arena_bytes = 3 << 30
    arena = torch.empty(arena_bytes, dtype=torch.uint8, device=dev)
    reqs, off = [], 0
    for kind, dims, batch in work:
        ob = nbytes(kind, dims, batch)[1]
        if off + ob > arena_bytes:
            off = 0
        reqs.append(Server.request(kind, dims, batch, False, src.data_ptr(), arena.data_ptr() + off))
        off = (off + ob + 255) // 256 * 256
    torch.cuda.synchronize()
    # one arrival sequence per rate, replayed by both modes
    rng = np.random.default_rng(12)
    arrivals_at = {rate: np.cumsum(rng.exponential(1e6 / rate, count)) for rate in rates}
    total_flop = sum(5.0 * math.prod(d) * math.log2(math.prod(d)) * b for _, d, b in work)
    res = {"requests": count, "replays_per_mode": reps, "mix": "85% 1D C2C/R2C N 256..65536 b1..128, 15% 2D C2C <=1024^2",
           "hbm_budget_bytes": 8 << 30, "rates": {}}
    # both servers live side by side; at each rate the two modes replay the stream `reps` times
    # each, back to back, the first of them alternating, so neither mode always runs on a warmer
    # device; percentiles are over the pooled latencies of the replays (a p99 of 3000 requests
    # alone moved by +-30 us from run to run at 20k req/s)
    servers = {}
    try:
        for batching in (True, False):
            srv = Server(hbm_budget=8 << 30, slots=3, batching=batching)
            servers[batching] = srv
            srv.replay(reqs, [0.0] * len(reqs))  # plans and resident classes built before timing
        for i, rate in enumerate(rates):
            arrivals = arrivals_at[rate]
            lats = {True: [], False: []}
            rps = {True: [], False: []}
            for rep in range(reps):
                for batching in ((True, False) if (i + rep) % 2 == 0 else (False, True)):
                    lat = np.array(servers[batching].replay(reqs, arrivals.tolist()))
                    lats[batching].append(lat)
                    rps[batching].append(count / float((arrivals + lat).max()) * 1e6)
            for batching in (True, False):
                lat = np.concatenate(lats[batching])
                achieved = float(np.mean(rps[batching]))
                entry = {"p50_us": round(float(np.percentile(lat, 50)), 1),
                         "p99_us": round(float(np.percentile(lat, 99)), 1),
                         "mean_us": round(float(lat.mean()), 1),
                         "achieved_req_per_s": round(achieved, 1),
                         "gflops": round(total_flop * achieved / count / 1e9, 1)}
                res["rates"].setdefault(str(rate), {})["batched" if batching else "per_request"] = entry
        for batching, srv in servers.items():
            res["batches_" + ("batched" if batching else "per_request")] = srv.stats()["batches"]
    finally:
        for srv in servers.values():
            srv.close()
    del arena, src
    return res


def bench_precision(torch, dev, stream, reps):
    """Precision form: 1D C2C N=65536 batch 64 and N=1024 batch 8192, inputs uniform[-1,1]
    scaled by 10^k (k uniform in {-3..3} per transform).  Times the FP32 plan and the split-FP16
    plan (N=1024: both radix-32 stages on MFMA; N=65536: the FP32 64-point column pass, then the
    1024-point rows on MFMA with a transposed store) per sample over a ring of (input, output)
    pairs larger than the device cache; reports the relative L2 error of each against numpy's
    FP64 FFT on 64 sampled transforms."""
    import numpy as np

    from sip_amd.fft import Plan
    res = {}
    for n, batch, seed in ((65536, 64, 8), (1024, 8192, 9)):
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        x = torch.complex(torch.rand(batch, n, generator=g, device=dev) * 2 - 1,
                          torch.rand(batch, n, generator=g, device=dev) * 2 - 1)
        k = torch.randint(-3, 4, (batch, 1), generator=g, device=dev).to(torch.float32)
        x = (x * torch.pow(10.0, k)).contiguous()
        rows = np.random.default_rng(seed).choice(batch, 64, replace=False)
        xs = x[torch.as_tensor(rows, device=dev)].cpu().numpy().astype(np.complex128)
        gold = np.fft.fft(xs, axis=1)
        entry = {}
        for prec in ("fp32", "split16"):
            with torch.cuda.stream(stream):
                p = Plan("c2c", (n,), batch, False, precision=prec)
                used = "mfma split-fp16" in p.describe()
                if prec == "split16" and not used:
                    entry[prec] = {"available": False}
                    p.close()
                    continue
                y = torch.empty_like(x)
                p.bind_workspace(dev)
                p.bind_stream(stream.cuda_stream)
                copies = ring_copies(x, y)
                xs_ring = [x] + [x.clone() for _ in range(copies - 1)]
                ys_ring = [y] + [torch.empty_like(y) for _ in range(copies - 1)]
                cur = [0]

                def run_ring():
                    i = cur[0] = (cur[0] + 1) % copies
                    p.execute(xs_ring[i], ys_ring[i])
                ms = median_ms(torch, stream, run_ring, reps)
                p.execute(x, y)
                torch.cuda.synchronize()
                del xs_ring, ys_ring
                got = y[torch.as_tensor(rows, device=dev)].cpu().numpy().astype(np.complex128)
                err = np.linalg.norm(got - gold, axis=1) / np.linalg.norm(gold, axis=1)
                entry[prec] = {"us": round(ms * 1e3, 2), "gflops": round(flops((n,), batch) / ms / 1e6, 1),
                               "hbm_frac": round(16.0 * n * batch / ms / 1e6 / HBM_PEAK_GBS, 4),
                               "rel_l2_max": float(f"{err.max():.3e}"), "rel_l2_rms": float(f"{np.sqrt((err**2).mean()):.3e}")}
                p.close()
        res[f"c2c_n{n}_b{batch}"] = entry
    return res


def bench_separated(torch, dev, stream, reps):
    """Separated-planes form (ASCEND_FFT_C2C_SEP): the same transforms on a pair of float32
    planes and on interleaved complex64, event-timed on one stream.  Algorithmic bytes are
    the same 16 per point either way (8 read, 8 written)."""
    from sip_amd import _native as N
    from sip_amd.fft import Plan
    L = N.lib()
    res = {}
    for dims, batch in (((4096,), 4096), ((65536,), 256), ((128, 128, 128), 4)):
        pts = 1
        for d in dims:
            pts *= d
        g = torch.Generator(device=dev)
        g.manual_seed(pts + batch)
        re = torch.rand((batch,) + dims, generator=g, device=dev) * 2 - 1
        im = torch.rand((batch,) + dims, generator=g, device=dev) * 2 - 1
        entry = {}
        with torch.cuda.stream(stream):
            ps = Plan("c2c_sep", dims, batch, False)
            ore, oim = torch.empty_like(re), torch.empty_like(im)
            ps.bind_workspace(dev)
            ps.bind_stream(stream.cuda_stream)
            ts = [N.AclTensor(t.shape, N.ACL_FLOAT, t.data_ptr()) for t in (re, im, ore, oim)]

            def run_sep():
                st = L.asdFftExecC2CSeparated(ps.handle, *[t.ptr for t in ts])
                if st != 0:
                    raise RuntimeError(f"asdFftExecC2CSeparated returned {st}")
            ms = median_ms(torch, stream, run_sep, reps)
            for t in ts:
                t.close()
            entry["separated"] = {"us": round(ms * 1e3, 2), "gflops": round(flops(dims, batch) / ms / 1e6, 1),
                                  "hbm_frac": round(16.0 * pts * batch / ms / 1e6 / HBM_PEAK_GBS, 4),
                                  "launches": len(ps.step_bytes())}
            ps.close()
            x = torch.complex(re, im).contiguous()
            y = torch.empty_like(x)
            pc = Plan("c2c", dims, batch, False)
            pc.bind_workspace(dev)
            pc.bind_stream(stream.cuda_stream)
            ts = [N.AclTensor(t.shape, N.ACL_COMPLEX64, t.data_ptr()) for t in (x, y)]

            def run_c2c():
                st = L.asdFftExecC2C(pc.handle, ts[0].ptr, ts[1].ptr)
                if st != 0:
                    raise RuntimeError(f"asdFftExecC2C returned {st}")
            ms = median_ms(torch, stream, run_c2c, reps)
            for t in ts:
                t.close()
            entry["interleaved"] = {"us": round(ms * 1e3, 2), "gflops": round(flops(dims, batch) / ms / 1e6, 1),
                                    "hbm_frac": round(16.0 * pts * batch / ms / 1e6 / HBM_PEAK_GBS, 4),
                                    "launches": len(pc.step_bytes())}
            pc.close()
        res["c2c_sep_" + "x".join(str(d) for d in dims) + f"_b{batch}"] = entry
    return res


# (name, kind, dims, batch, inverse): every configuration of the plain FP32 drop-in, the
# C2C N=4096 batch 512 case and the two long/short C2C shapes of the precision form
VENDOR_CASES = [
    ("c2c_n4096_b512", "c2c", (4096,), 512, False),
    ("c2c_n1000_b256", "c2c", (1000,), 256, False),
    ("c2c_n2187_b256", "c2c", (2187,), 256, False),
    ("c2c_n6000_b256", "c2c", (6000,), 256, False),
    ("c2c_n4096_b256", "c2c", (4096,), 256, False),
    ("r2c_n8192_b1024", "r2c", (8192,), 1024, False),
    ("c2r_n8192_b1024", "c2r", (8192,), 1024, True),
    ("c2c2d_2048x2048_b4", "c2c", (2048, 2048), 4, False),
    ("c2c3d_256x256x256_b1", "c2c", (256, 256, 256), 1, False),
    ("c2c_n65536_b64", "c2c", (65536,), 64, False),
    ("c2c_n1024_b8192", "c2c", (1024,), 8192, False),
]


# hipFFT (include/hipfft/hipfft.h): transform types and the forward direction
HIPFFT_C2C, HIPFFT_R2C, HIPFFT_C2R, HIPFFT_FORWARD = 0x29, 0x2A, 0x2C, -1


def _hipfft(torch):
    """The hipFFT library torch.fft itself calls (torch's bundled copy), through ctypes."""
    path = os.path.join(os.path.dirname(torch.__file__), "lib", "libhipfft.so")
    if not os.path.exists(path):
        return None
    H = ctypes.CDLL(path)
    vp, i = ctypes.c_void_p, ctypes.c_int
    H.hipfftPlanMany.argtypes = [ctypes.POINTER(vp), i, ctypes.POINTER(i), vp, i, i, vp, i, i, i, i]
    H.hipfftSetStream.argtypes = [vp, vp]
    H.hipfftExecC2C.argtypes = [vp, vp, vp, i]
    H.hipfftExecR2C.argtypes = [vp, vp, vp]
    H.hipfftExecC2R.argtypes = [vp, vp, vp]
    H.hipfftDestroy.argtypes = [vp]
    for f in ("hipfftPlanMany", "hipfftSetStream", "hipfftExecC2C", "hipfftExecR2C", "hipfftExecC2R",
              "hipfftDestroy"):
        getattr(H, f).restype = i
    return H


def bench_vendor(torch, dev, stream, reps, with_torch=True):
    """Vendor comparator: the same transforms through rocFFT, called two ways, beside this
    library's C-ABI exec, all event-timed on one stream over the same ring of (input, output)
    pairs (ring_copies: each side spans at least RING_BYTES, so the device cache cannot serve a
    repetition):
      hipfft: hipFFT's plan-many / exec entry points called directly on the ring's buffers (the
        library torch.fft uses, without torch: no extra device work);
      torch: torch.fft, which copies its input before each rocFFT call and allocates each
        output (the previous output of the same ring slot is released just before, so its writes
        rotate over as many blocks).
    rocFFT's C2R overwrites its input, so the hipfft path runs last on the ring.  agree_* =
    relative L2 between this library's output and the other path's, on the same input."""
    from sip_amd import _native as N
    from sip_amd.fft import Plan
    L = N.lib()
    H = _hipfft(torch)
    res = {} if H is not None else {"hipfft": "unavailable (torch ships no libhipfft.so)"}
    for name, kind, dims, batch, inv in VENDOR_CASES:
        rank = len(dims)
        with torch.cuda.stream(stream):
            plan = Plan(kind, dims, batch, inv)
            plan.bind_workspace(dev)
            plan.bind_stream(stream.cuda_stream)
            if kind == "c2r":
                n = dims[-1]
                g = torch.Generator(device=dev)
                g.manual_seed(5)
                x0 = torch.fft.rfft(torch.randn((batch, n), generator=g, device=dev)).contiguous()
            else:
                x0 = make_inputs(torch, dev, kind, dims, batch, len(name))
            pristine = x0.clone()
            out_dtype = torch.float32 if kind == "c2r" else torch.complex64
            y0 = torch.empty(plan.out_shape(), dtype=out_dtype, device=dev)
            copies = ring_copies(x0, y0)
            xs = [x0] + [x0.clone() for _ in range(copies - 1)]
            ys = [y0] + [torch.empty_like(y0) for _ in range(copies - 1)]
            acl = lambda t: N.AclTensor(t.shape, N.ACL_FLOAT if t.dtype == torch.float32 else N.ACL_COMPLEX64,
                                        t.data_ptr())
            tins, touts = [acl(t) for t in xs], [acl(t) for t in ys]
            fn = {"c2c": L.asdFftExecC2C, "r2c": L.asdFftExecR2C, "c2r": L.asdFftExecC2R}[kind]
            if kind == "c2r":
                tf = lambda x: torch.fft.irfft(x, n=dims[-1], norm="forward")
            elif kind == "r2c":
                tf = torch.fft.rfft
            elif rank == 1:
                tf = torch.fft.fft
            else:
                tf = lambda x: torch.fft.fftn(x, dim=tuple(range(-rank, 0)))
            hp = ctypes.c_void_p()
            if H is not None:
                nn = (ctypes.c_int * rank)(*dims)
                htype = {"c2c": HIPFFT_C2C, "r2c": HIPFFT_R2C, "c2r": HIPFFT_C2R}[kind]
                if H.hipfftPlanMany(ctypes.byref(hp), rank, nn, None, 1, 0, None, 1, 0, htype, batch) != 0:
                    raise RuntimeError(f"{name}: hipfftPlanMany failed")
                if H.hipfftSetStream(hp, ctypes.c_void_p(stream.cuda_stream)) != 0:
                    raise RuntimeError(f"{name}: hipfftSetStream failed")

            def hip_exec(x, y):
                if kind == "c2c":
                    st = H.hipfftExecC2C(hp, x, y, HIPFFT_FORWARD)
                elif kind == "r2c":
                    st = H.hipfftExecR2C(hp, x, y)
                else:
                    st = H.hipfftExecC2R(hp, x, y)
                if st != 0:
                    raise RuntimeError(f"{name}: hipfftExec returned {st}")
            torch.cuda.synchronize()
            cur = [0]

            def ours():
                i = cur[0] = (cur[0] + 1) % copies
                st = fn(plan.handle, tins[i].ptr, touts[i].ptr)
                if st != 0:
                    raise RuntimeError(f"{name}: exec returned {st}")
            ms_ours = median_ms(torch, stream, ours, reps)

            def rel(a, b):
                return float(f"{(torch.linalg.vector_norm((a - b).flatten()) / torch.linalg.vector_norm(b.flatten())).item():.3g}")
            entry = {"ours_us": round(ms_ours * 1e3, 2),
                     "ours_hbm_frac": round(alg_bytes(kind, dims, batch) / ms_ours / 1e6 / HBM_PEAK_GBS, 4)}
            if with_torch:
                outs = [None] * copies

                def via_torch():
                    i = cur[0] = (cur[0] + 1) % copies
                    outs[i] = None
                    outs[i] = tf(xs[i])
                ms_torch = median_ms(torch, stream, via_torch, reps)
                ours()
                i = cur[0]
                entry.update({"torch_fft_us": round(ms_torch * 1e3, 2), "speedup_vs_torch_fft": round(ms_torch / ms_ours, 3),
                              "agree_torch_fft": rel(ys[i], tf(xs[i]))})
                del outs

            if H is None:
                for t in tins + touts:
                    t.close()
                plan.close()
                res[name] = entry
                continue

            def via_hipfft():
                i = cur[0] = (cur[0] + 1) % copies
                hip_exec(xs[i].data_ptr(), ys[i].data_ptr())
            ms_hip = median_ms(torch, stream, via_hipfft, reps)
            # agreement on fresh copies of the input (rocFFT's C2R overwrites its input)
            xs[0].copy_(pristine)
            xs[1 % copies].copy_(pristine)
            y_ours = torch.empty_like(y0)
            to = acl(y_ours)
            st = fn(plan.handle, tins[0].ptr, to.ptr)
            to.close()
            if st != 0:
                raise RuntimeError(f"{name}: exec returned {st}")
            hip_exec(xs[1 % copies].data_ptr(), ys[1 % copies].data_ptr())
            torch.cuda.synchronize()
            entry.update({"rocfft_us": round(ms_hip * 1e3, 2), "speedup_vs_rocfft": round(ms_hip / ms_ours, 3),
                          "rocfft_hbm_frac": round(alg_bytes(kind, dims, batch) / ms_hip / 1e6 / HBM_PEAK_GBS, 4),
                          "agree_rocfft": rel(y_ours, ys[1 % copies]), "ring_pairs": copies})
            H.hipfftDestroy(hp)
            for t in tins + touts:
                t.close()
            plan.close()
            del xs, ys, y_ours, pristine
        res[name] = entry
    torch.cuda.empty_cache()
    return res


def bench_shift(torch, dev, stream, reps):
    """Shift form: y = fftshift(FFT2(ifftshift(x)) / (X*Y)) on a 2D C2C batch, as one plan with
    the shifts folded in (asdFftExtSetShift) and as the reference's three calls (asdIfftShift
    with the scale, exec, asdFftShift), event-timed through the C ABI on one stream.
    Algorithmic bytes of the whole operation: x read once + y written once."""
    from sip_amd import _native as N
    from sip_amd.fft import Plan
    L = N.lib()
    res = {}
    for X, Y, B in ((2048, 2048, 4), (1000, 1000, 16)):
        g = torch.Generator(device=dev)
        g.manual_seed(X + B)
        x = torch.complex(torch.rand(B, X, Y, generator=g, device=dev) * 2 - 1,
                          torch.rand(B, X, Y, generator=g, device=dev) * 2 - 1).contiguous()
        y = torch.empty_like(x)
        t1 = torch.empty_like(x)
        t2 = torch.empty_like(x)
        scale = 1.0 / (X * Y)
        with torch.cuda.stream(stream):
            fused = Plan("c2c", (X, Y), B, False, in_shift="ifft", out_shift="fft", scale=scale)
            plain = Plan("c2c", (X, Y), B, False)
            for p in (fused, plain):
                p.bind_workspace(dev)
                p.bind_stream(stream.cuda_stream)
            tx, ty, tt1, tt2 = [N.AclTensor(t.shape, N.ACL_COMPLEX64, t.data_ptr()) for t in (x, y, t1, t2)]
            sh = stream.cuda_stream

            def run_fused():
                if L.asdFftExecC2C(fused.handle, tx.ptr, ty.ptr) != 0:
                    raise RuntimeError("fused exec failed")

            def run_three():
                if (L.asdIfftShift(X, Y, B, 0, scale, tx.ptr, tt1.ptr, ctypes.c_void_p(sh), None) != 0 or
                        L.asdFftExecC2C(plain.handle, tt1.ptr, tt2.ptr) != 0 or
                        L.asdFftShift(X, Y, B, 0, tt2.ptr, ty.ptr, ctypes.c_void_p(sh), None) != 0):
                    raise RuntimeError("three-call sequence failed")
            ms_f = median_ms(torch, stream, run_fused, reps)
            yf = y.clone()
            ms_3 = median_ms(torch, stream, run_three, reps)
            torch.cuda.synchronize()
            rel = (torch.linalg.vector_norm(yf - y) / torch.linalg.vector_norm(y)).item()
            for t in (tx, ty, tt1, tt2):
                t.close()
            fused.close()
            plain.close()
        alg = 16.0 * x.numel()
        res[f"shift_fft2_shift_{X}x{Y}_b{B}"] = {
            "fused_us": round(ms_f * 1e3, 2), "three_calls_us": round(ms_3 * 1e3, 2),
            "speedup": round(ms_3 / ms_f, 2), "fused_alg_GBps": round(alg / ms_f / 1e6, 1),
            "fused_launches": 2, "three_call_launches": 4, "rel_l2_fused_vs_three": float(f"{rel:.3e}")}
    return res


def bench_istft(torch, dev, stream, reps):
    """ISTFT form: asdFftExecIstft on a complex64 (channel, nfft, nframes) spectrum (center,
    two-sided, complex output), event-timed through the C ABI, beside torch.istft (rocFFT) on
    the same data.  Algorithmic bytes = spectrum read + window read + signal written."""
    from sip_amd import _native as N
    from sip_amd.fft import Istft
    L = N.lib()
    res = {}
    for channel, nfft, nframes, hop in ((16, 1024, 1000, 256), (64, 400, 500, 160)):
        g = torch.Generator(device=dev)
        g.manual_seed(nfft + hop)
        spec = torch.complex(torch.rand(channel, nfft, nframes, generator=g, device=dev) * 2 - 1,
                             torch.rand(channel, nfft, nframes, generator=g, device=dev) * 2 - 1).contiguous()
        k = torch.arange(nfft, device=dev, dtype=torch.float64)
        win = (0.6 + 0.4 * torch.sin(2 * math.pi * k / nfft + 0.7)).float()
        with torch.cuda.stream(stream):
            p = Istft(channel, nfft, nframes, hop)
            out = torch.empty((channel, p.out_len), dtype=torch.complex64, device=dev)
            p.bind_workspace(dev)
            p.bind_stream(stream.cuda_stream)
            ts = [N.AclTensor(t.shape, N.ACL_FLOAT if t.dtype == torch.float32 else N.ACL_COMPLEX64, t.data_ptr())
                  for t in (spec, win, out)]

            def run():
                st = L.asdFftExecIstft(p.handle, *[t.ptr for t in ts])
                if st != 0:
                    raise RuntimeError(f"asdFftExecIstft returned {st}")
            ms = median_ms(torch, stream, run, reps)
            ms_torch = median_ms(torch, stream, lambda: torch.istft(
                spec, n_fft=nfft, hop_length=hop, win_length=nfft, window=win, center=True, normalized=False,
                onesided=False, return_complex=True), reps)
            ref = torch.istft(spec, n_fft=nfft, hop_length=hop, win_length=nfft, window=win, center=True,
                              normalized=False, onesided=False, return_complex=True)
            torch.cuda.synchronize()
            rel = (torch.linalg.vector_norm(out - ref) / torch.linalg.vector_norm(ref)).item()
            for t in ts:
                t.close()
            p.close()
        alg = 8.0 * spec.numel() + 4.0 * nfft + 8.0 * out.numel()
        res[f"istft_c{channel}_n{nfft}_t{nframes}_h{hop}"] = {
            "us": round(ms * 1e3, 2), "alg_GBps": round(alg / ms / 1e6, 1),
            "hbm_frac": round(alg / ms / 1e6 / HBM_PEAK_GBS, 4), "launches": 2,
            "torch_istft_us": round(ms_torch * 1e3, 2), "speedup_vs_torch_istft": round(ms_torch / ms, 2),
            "rel_l2_vs_torch_istft": float(f"{rel:.3e}")}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--samples", type=int, default=300,
                    help="per-sample event timings per workload and per launch (median / mean reported)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--only", default=None, help="run one workload entry by name (profiling)")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--no-ragged", action="store_true", help="skip the ragged-batch form")
    ap.add_argument("--ragged-only", action="store_true", help="run only the ragged-batch form")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving form")
    ap.add_argument("--no-precision", action="store_true", help="skip the precision form")
    ap.add_argument("--no-separated", action="store_true", help="skip the separated-planes form")
    ap.add_argument("--separated-only", action="store_true", help="run only the separated-planes form")
    ap.add_argument("--no-istft", action="store_true", help="skip the ISTFT form")
    ap.add_argument("--no-shift", action="store_true", help="skip the folded-shift form")
    ap.add_argument("--shift-only", action="store_true", help="run only the folded-shift form")
    ap.add_argument("--istft-only", action="store_true", help="run only the ISTFT form")
    ap.add_argument("--no-vendor", action="store_true", help="skip the rocFFT (torch.fft) comparator form")
    ap.add_argument("--vendor-only", action="store_true", help="run only the rocFFT (torch.fft) comparator form")
    ap.add_argument("--vendor-no-torch", action="store_true",
                    help="comparator form without the torch.fft path (direct hipFFT only; for kernel traces)")
    ap.add_argument("--serve-only", action="store_true", help="run only the serving form")
    ap.add_argument("--precision-only", action="store_true", help="run only the precision form")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 256 MiB device-to-device tensor copy (reference dispatch for HBM counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only (--gpus 1); multi-GPU sharding is a design note, "
              "not a measured path", file=sys.stderr)
        sys.exit(2)

    import torch

    from sip_amd import _native as N
    from sip_amd.fft import Plan
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    if a.ragged_only:
        print(json.dumps({"forms": {"ragged": bench_ragged(torch, dev, stream, max(5, a.samples), 3)}}))
        return
    if a.serve_only:
        print(json.dumps({"forms": {"serving": bench_serve(torch, dev)}}))
        return
    if a.precision_only:
        print(json.dumps({"forms": {"precision": bench_precision(torch, dev, stream, max(5, a.samples))}}))
        return
    if a.separated_only:
        print(json.dumps({"forms": {"separated": bench_separated(torch, dev, stream, max(5, a.samples))}}))
        return
    if a.shift_only:
        print(json.dumps({"forms": {"shift": bench_shift(torch, dev, stream, max(5, a.samples))}}))
        return
    if a.istft_only:
        print(json.dumps({"forms": {"istft": bench_istft(torch, dev, stream, max(5, a.samples))}}))
        return
    if a.vendor_only:
        print(json.dumps({"forms": {"vendor": bench_vendor(torch, dev, stream, max(5, a.samples),
                                                           not a.vendor_no_torch)}}))
        return
    L = N.lib()
    entries = [w for w in WORKLOAD if a.only is None or w[0] == a.only or
               (a.only == "c2r_n8192_b1024" and w[0] == "r2c_n8192_b1024")]

    jobs = []
    with torch.cuda.stream(stream):
        prev_out = {}
        for name, kind, dims, batch, inv, seed in entries:
            plan = Plan(kind, dims, batch, inv)
            if kind == "c2r":
                x = prev_out["r2c_n8192_b1024"]
            else:
                x = make_inputs(torch, dev, kind, dims, batch, seed)
            out_dtype = torch.float32 if kind == "c2r" else torch.complex64
            y = torch.empty(plan.out_shape(), dtype=out_dtype, device=dev)
            plan.bind_workspace(dev)
            plan.bind_stream(stream.cuda_stream)
            tin = N.AclTensor(x.shape, N.ACL_FLOAT if x.dtype == torch.float32 else N.ACL_COMPLEX64, x.data_ptr())
            tout = N.AclTensor(y.shape, N.ACL_FLOAT if y.dtype == torch.float32 else N.ACL_COMPLEX64, y.data_ptr())
            fn = {"c2c": L.asdFftExecC2C, "r2c": L.asdFftExecR2C, "c2r": L.asdFftExecC2R}[kind]
            jobs.append(dict(name=name, kind=kind, dims=dims, batch=batch, inv=inv, plan=plan, x=x, y=y, tin=tin,
                             tout=tout, fn=fn, flop=flops(dims, batch), bytes=alg_bytes(kind, dims, batch),
                             steps=plan.step_bytes(), describe=plan.describe()))
            prev_out[name] = y
    torch.cuda.synchronize()

    def run_job(j):
        st = j["fn"](j["plan"].handle, j["tin"].ptr, j["tout"].ptr)
        if st != 0:
            raise RuntimeError(f"{j['name']}: exec returned {st}")

    def run_step():
        for j in jobs:
            run_job(j)

    with torch.cuda.stream(stream):
        for _ in range(a.warmup):
            run_step()
        torch.cuda.synchronize()
        t_wall0 = time.perf_counter()
        ms_per_step = event_time_ms(torch, stream, run_step, a.steps)
        wall_ms = (time.perf_counter() - t_wall0) * 1e3 / a.steps

        # per-workload and per-launch timings (same stream, same kernels).  Each repetition
        # takes the next (input, output) pair of a ring whose inputs and whose outputs each span
        # at least RING_BYTES, so the 256 MB device cache (MALL) cannot serve a repetition from
        # an earlier one's data, in-place steps included.
        breakdown = []
        launches = []
        for j in jobs:
            copies = ring_copies(j["x"], j["y"])
            ring = [(j["tin"], j["tout"])]
            keep = []
            for _ in range(copies - 1):
                xr = j["x"].clone()
                yr = torch.empty_like(j["y"])
                keep += [xr, yr]
                ring.append((N.AclTensor(xr.shape, N.ACL_FLOAT if xr.dtype == torch.float32 else N.ACL_COMPLEX64,
                                         xr.data_ptr()),
                             N.AclTensor(yr.shape, N.ACL_FLOAT if yr.dtype == torch.float32 else N.ACL_COMPLEX64,
                                         yr.data_ptr())))
            torch.cuda.synchronize()
            cursor = [0]

            def next_pair():
                cursor[0] = (cursor[0] + 1) % len(ring)
                return ring[cursor[0]]

            def run_ring():
                ti, to = next_pair()
                st = j["fn"](j["plan"].handle, ti.ptr, to.ptr)
                if st != 0:
                    raise RuntimeError(f"{j['name']}: exec returned {st}")
            dt = device_times_us(torch, stream, run_ring, a.samples)
            ms = dt["median"] * 1e-3
            ms_b2b = event_time_ms(torch, stream, run_ring, a.samples)
            # copy ceiling: a device-to-device copy of the same bytes (read input-size, write
            # output-size, the larger side split so both sides move) over the same ring
            cpy = copy_ceiling_us(torch, stream, j["x"], j["y"], a.samples)
            entry = {"name": j["name"], "us": round(ms * 1e3, 2), "us_p10": round(dt["p10"], 2),
                     "us_p90": round(dt["p90"], 2), "samples": dt["n"],
                     "us_back_to_back": round(ms_b2b * 1e3, 2),
                     "gflops": round(j["flop"] / ms / 1e6, 1),
                     "alg_GBps": round(j["bytes"] / ms / 1e6, 1),
                     "hbm_frac": round(j["bytes"] / ms / 1e6 / HBM_PEAK_GBS, 3),
                     "copy_us": round(cpy, 2), "frac_of_copy": round(cpy / (ms * 1e3), 3)}
            if j["kind"] in ("r2c", "c2r"):
                entry["gflops_conventional_2.5NlogN"] = round(0.5 * j["flop"] / ms / 1e6, 1)
            breakdown.append(entry)
            for si, sb in enumerate(j["steps"]):
                def one(si=si):
                    ti, to = next_pair()
                    st = L.asdFftExtExecStep(j["plan"].handle, ti.ptr, to.ptr, si)
                    if st != 0:
                        raise RuntimeError(f"{j['name']} step {si}: {st}")
                # the launch's average duration: HIP events around `samples` back-to-back launches
                # on its stream (the profiler's kernel averages compare with this); the
                # per-sample median adds each sample's event overhead (about 2-3 us)
                kms = event_time_ms(torch, stream, one, a.samples)
                kt = device_times_us(torch, stream, one, a.samples)
                launches.append({"workload": j["name"], "step": si, "us": kms * 1e3, "us_sample_median": kt["median"],
                                 "bytes": sb, "GBps": sb / kms / 1e6, "ring_pairs": len(ring)})
            for ti, to in ring[1:]:
                ti.close()
                to.close()
            del keep
    # per-configuration error against a float64 FFT, outside the timed region
    for j, entry in zip(jobs, breakdown):
        with torch.cuda.stream(stream):
            run_job(j)
        torch.cuda.synchronize()
        entry["rel_l2"] = sampled_error(torch, j["kind"], j["dims"], j["batch"], j["inv"], j["x"], j["y"])
    total_flop = sum(j["flop"] for j in jobs)
    value = total_flop / ms_per_step / 1e6  # GFLOP/s

    dom = max(launches, key=lambda r: r["us"])
    label = f"{dom['workload']} launch {dom['step']}"
    traffic = measured_traffic().get(label)
    roofline = {"bound": "hbm", "achieved": round(dom["GBps"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(dom["GBps"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": label, "us": round(dom["us"], 2), "algorithmic_bytes": dom["bytes"]}
    if traffic is not None:
        roofline["traffic_source"] = os.path.relpath(TRAFFIC_PROFILE, ROOT) + " (FETCH_SIZE x2, WRITE_SIZE x1 per copy calibration)"

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain-fp32-coverage: C2C 1D N={1000,2187,6000,4096} b256; R2C then C2R N=8192 "
                               "b1024; C2C 2D 2048x2048 b4; C2C 3D 256^3" if a.only is None else a.only,
                   "parallelism": "single-gpu",
                   "alg_bytes_per_step": sum(j["bytes"] for j in jobs),
                   "step_hbm_frac": round(sum(j["bytes"] for j in jobs) / ms_per_step / 1e6 / HBM_PEAK_GBS, 4),
                   "wall_ms_per_step": round(wall_ms, 4)},
        "breakdown": breakdown,
        "roofline": roofline,
    }
    if a.verbose:
        line["launches"] = [{k: (round(v, 2) if isinstance(v, float) else v) for k, v in r.items()} for r in launches]
        for j in jobs:
            print(j["describe"], file=sys.stderr)
    if a.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        line["calibration_copy_bytes"] = {"read": src.numel() * 4, "write": dst.numel() * 4}
        del src, dst
    forms = {}
    if not a.no_ragged:
        forms["ragged"] = bench_ragged(torch, dev, stream, max(5, a.samples), 3)
    if not a.no_serve:
        forms["serving"] = bench_serve(torch, dev)
    if not a.no_precision:
        forms["precision"] = bench_precision(torch, dev, stream, max(5, a.samples))
    if not a.no_separated:
        forms["separated"] = bench_separated(torch, dev, stream, max(5, a.samples))
    if not a.no_istft:
        forms["istft"] = bench_istft(torch, dev, stream, max(5, a.samples))
    if not a.no_shift:
        forms["shift"] = bench_shift(torch, dev, stream, max(5, a.samples))
    if not a.no_vendor:
        forms["vendor"] = bench_vendor(torch, dev, stream, max(5, a.samples))
    if forms:
        line["forms"] = forms
    if not a.no_cpu:
        numpy_cpu = cpu_numpy()
        line["cpu_baseline"] = numpy_cpu.pop("baseline")
        forms["cpu_numpy"] = numpy_cpu
        forms["cpu_baseline_oracle"] = cpu_oracle()
        line["forms"] = forms
    for j in jobs:
        j["tin"].close()
        j["tout"].close()
        j["plan"].close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
