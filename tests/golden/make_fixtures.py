"""Writes the golden fixtures under tests/golden/ (run in the build container only).

The reference computes its goldens at test time with numpy's FFT (pocketfft):
  C2C  tests/ut/unittest/fft/1d/c2c/c2c_data/gen_data.py:42-47
  R2C  tests/ut/unittest/fft/1d/r2c/r2c_data/gen_data.py:42-48
  C2R  tests/ut/unittest/fft/1d/c2r/c2r_data/gen_data.py:42-48
  2D   tests/ut/unittest/fft/2d/c2c/c2c_2d_data/gen_data_2d.py:64-67
  3D   tests/ut/unittest/fft/3d/c2c/c2c_3d_data/gen_data_3d.py:37-48
and feeds them the unit-test LCG inputs (seed 2026, U[-300,300),
tests/ut/framework/test_util/float_util.h:23-36, util.cpp:324-338).  No golden
files are committed in the reference, so this script regenerates the same
goldens for the same shapes (tests/ut/unittest/fft/*/test_fft_*.cpp case
lists) in complex128 / float64 and stores them with their inputs' recipe.

It also stores small slices of the benchmark workloads (seeded
numpy.random.default_rng inputs, uniform[-1,1]) with their complex128 goldens.

Output: ut_cases.npz, workload_cases.npz, manifest.json.  numpy is the
golden's own library; the reference tree is not imported or executed.
"""
from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


# --- independent restatement of the unit-test LCG (float32 arithmetic) ---
def lcg_uniform(n, state=2026, lower=-300.0, upper=300.0):
    out = np.empty(n, dtype=np.float32)
    s = state
    inv = np.float32(1.0 / 4294967296.0)
    lo = np.float32(lower)
    span = np.float32(upper) - np.float32(lower)
    for i in range(n):
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
        out[i] = lo + (np.float32(s) * inv) * span
    return out


def ut_complex(shape):
    n = int(np.prod(shape))
    return lcg_uniform(2 * n).view(np.complex64).reshape(shape)


def ut_real(shape):
    return lcg_uniform(int(np.prod(shape))).reshape(shape)


# (name, kind, dims, batch, inverse); shapes follow the reference test lists
C2C_1D = [17, 19, 49, 64, 81, 121, 125, 143, 169, 210, 289, 323, 343, 360, 361, 625, 729, 1024, 2310, 4096, 4620]
R2C_1D = [(1, 16), (1, 27), (1, 25), (1, 49), (1, 121), (1, 169), (1, 289), (1, 361), (1, 210), (2, 16), (4, 27),
          (8, 210), (1, 2048), (1, 2187), (1, 3125), (1, 2401), (2, 1573), (1, 2160), (1, 2520)]
C2R_1D = [(1, 16), (2, 32), (1, 64), (1, 27), (1, 25), (1, 49), (1, 121), (1, 289), (1, 361), (1, 30), (1, 210),
          (2, 30), (1, 2048), (1, 2187), (1, 3125), (2, 1573), (1, 2160), (1, 2520), (2, 1056)]
C2C_2D = [((64, 64), 1, False), ((64, 64), 1, True), ((32, 48), 2, False), ((112, 104), 1, False),
          ((500, 50), 2, False), ((50, 20), 2, True)]
C2C_3D = [((8, 8, 8), 1, False), ((16, 16, 16), 2, True), ((48, 40, 24), 1, False)]
REAL_2D = [("r2c", (64, 96), 2, False), ("r2c", (30, 50), 1, True), ("c2r", (64, 96), 2, True),
           ("c2r", (30, 50), 1, False), ("c2r", (21, 35), 1, True)]
REAL_3D = [("r2c", (8, 12, 16), 2, False), ("c2r", (8, 12, 16), 1, True), ("c2r", (6, 10, 9), 1, False)]


def golden_c2c(x, axes, inverse):
    x = x.astype(np.complex128)
    if inverse:
        return np.fft.ifftn(x, axes=axes, norm="forward")
    return np.fft.fftn(x, axes=axes)


def golden_r2c(x, ndim, inverse):
    y = np.fft.rfft(x.astype(np.float64), axis=-1)
    if inverse:
        y = np.conj(y)
    axes = tuple(range(-ndim, -1))
    if axes:
        y = np.fft.ifftn(y, axes=axes, norm="forward") if inverse else np.fft.fftn(y, axes=axes)
    return y


def golden_c2r(x, dims, inverse):
    ndim = len(dims)
    z = x.astype(np.complex128)
    axes = tuple(range(-ndim, -1))
    if axes:
        z = np.fft.ifftn(z, axes=axes, norm="forward") if inverse else np.fft.fftn(z, axes=axes)
    if inverse:
        return np.fft.irfft(z, n=dims[-1], axis=-1, norm="forward")
    return np.fft.hfft(z, n=dims[-1], axis=-1)


def main():
    ut = {}
    manifest = {"ut": [], "workload": []}

    def add(name, kind, dims, batch, inverse, golden):
        ut[name] = golden
        manifest["ut"].append({"name": name, "kind": kind, "dims": list(dims), "batch": batch,
                               "inverse": bool(inverse), "input": "lcg2026"})

    for n in C2C_1D:
        for inv in (False, True):
            x = ut_complex((1, n))
            add(f"c2c1d_n{n}_{'inv' if inv else 'fwd'}", "c2c", (n,), 1, inv, golden_c2c(x, (-1,), inv))
    for b, n in ((2, 1024), (2, 210)):
        x = ut_complex((b, n))
        add(f"c2c1d_n{n}_b{b}_fwd", "c2c", (n,), b, False, golden_c2c(x, (-1,), False))
    for b, n in R2C_1D:
        for inv in (False, True):
            x = ut_real((b, n))
            add(f"r2c1d_n{n}_b{b}_{'inv' if inv else 'fwd'}", "r2c", (n,), b, inv, golden_r2c(x, 1, inv))
    for b, n in C2R_1D:
        for inv in (False, True):
            x = ut_complex((b, n // 2 + 1))
            add(f"c2r1d_n{n}_b{b}_{'inv' if inv else 'fwd'}", "c2r", (n,), b, inv, golden_c2r(x, (n,), inv))
    for dims, b, inv in C2C_2D:
        x = ut_complex((b,) + dims)
        add(f"c2c2d_{dims[0]}x{dims[1]}_b{b}_{'inv' if inv else 'fwd'}", "c2c", dims, b, inv,
            golden_c2c(x, (-2, -1), inv))
    for dims, b, inv in C2C_3D:
        x = ut_complex((b,) + dims)
        add(f"c2c3d_{'x'.join(map(str, dims))}_b{b}_{'inv' if inv else 'fwd'}", "c2c", dims, b, inv,
            golden_c2c(x, (-3, -2, -1), inv))
    for kind, dims, b, inv in REAL_2D + REAL_3D:
        tag = "x".join(map(str, dims))
        if kind == "r2c":
            x = ut_real((b,) + dims)
            g = golden_r2c(x, len(dims), inv)
        else:
            x = ut_complex((b,) + dims[:-1] + (dims[-1] // 2 + 1,))
            g = golden_c2r(x, dims, inv)
        add(f"{kind}{len(dims)}d_{tag}_b{b}_{'inv' if inv else 'fwd'}", kind, dims, b, inv, g)

    # first LCG values, pinned independently of the C restatement
    manifest["lcg2026_first8"] = [float(v) for v in lcg_uniform(8)]

    # --- benchmark-workload slices (numpy default_rng, uniform[-1,1]) ---
    wl = {}

    def uni_c(rng, shape):
        re = rng.uniform(-1.0, 1.0, shape).astype(np.float32)
        im = rng.uniform(-1.0, 1.0, shape).astype(np.float32)
        return (re + 1j * im).astype(np.complex64)

    def add_wl(name, kind, dims, rows, inverse, x, g):
        wl[name + "_in"] = x
        wl[name + "_gold"] = g
        manifest["workload"].append({"name": name, "kind": kind, "dims": list(dims), "batch": rows,
                                     "inverse": bool(inverse)})

    rng = np.random.default_rng(0)
    x = uni_c(rng, (2, 4096))
    add_wl("c2c_n4096", "c2c", (4096,), 2, False, x, golden_c2c(x, (-1,), False))
    for seed, n in ((1, 1000), (2, 2187), (3, 6000), (4, 4096)):
        rng = np.random.default_rng(seed)
        x = uni_c(rng, (2, n))
        add_wl(f"c2c_n{n}_s{seed}", "c2c", (n,), 2, False, x, golden_c2c(x, (-1,), False))
    rng = np.random.default_rng(5)
    n = 8192
    t = np.arange(n)
    xr = rng.standard_normal((2, n)).astype(np.float64)
    for r in range(2):
        for _ in range(3):
            k = rng.integers(1, n // 2)
            xr[r] += rng.uniform(0.5, 4.0) * np.cos(2 * np.pi * k * t / n + rng.uniform(0, 2 * np.pi))
    xr = xr.astype(np.float32)
    add_wl("r2c_n8192", "r2c", (n,), 2, False, xr, golden_r2c(xr, 1, False))
    xc = np.fft.rfft(xr.astype(np.float64)).astype(np.complex64)
    add_wl("c2r_n8192", "c2r", (n,), 2, True, xc, golden_c2r(xc, (n,), True))
    rng = np.random.default_rng(8)
    x = uni_c(rng, (1, 65536)) * np.float32(10.0) ** rng.integers(-3, 4)
    add_wl("c2c_n65536", "c2c", (65536,), 1, False, x.astype(np.complex64),
           golden_c2c(x.astype(np.complex64), (-1,), False))
    rng = np.random.default_rng(6)
    x = uni_c(rng, (1, 2048, 16))
    add_wl("c2c2d_2048x16", "c2c", (2048, 16), 1, False, x, golden_c2c(x, (-2, -1), False))
    rng = np.random.default_rng(7)
    x = uni_c(rng, (1, 256, 8, 8))
    add_wl("c2c3d_256x8x8", "c2c", (256, 8, 8), 1, True, x, golden_c2c(x, (-3, -2, -1), True))

    np.savez_compressed(os.path.join(HERE, "ut_cases.npz"), **ut)
    np.savez_compressed(os.path.join(HERE, "workload_cases.npz"), **wl)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    print(len(ut), "ut cases,", len(manifest["workload"]), "workload cases")


if __name__ == "__main__":
    main()
