"""numpy front end of the CPU FP64 oracle (oracle/fft_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker.  The shipped library
never imports it.

Semantics (reference citations in fft_oracle.c):
  c2c(x, inverse)            fft / unnormalised ifft over the last axis
  r2c(x, inverse)            rfft / conj(rfft)
  c2r(x, n, inverse)         hfft(n) / unnormalised irfft(n)
  c2c_nd(x, ndim, inverse)   fftn over the trailing ndim axes
  r2c_nd(x, ndim, inverse)   R2C on the last axis, then C2C on the other FFT axes
  c2r_nd(x, dims, inverse)   C2C on the leading FFT axes, then C2R on the last
  shift2d(...)               fftshift / ifftshift(+scale) of 2D slices
  istft(spec, window, hop)   ISTFT: 1/nfft inverse DFT per frame, window, overlap-add,
                             squared-window envelope, centre crop
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build() -> str:
    """Compile the oracle with gcc (make) and return the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        lp = ctypes.POINTER(ctypes.c_long)
        lib.oracle_dft_direct.argtypes = [dp, dp, ctypes.c_long, ctypes.c_int]
        lib.oracle_fft.argtypes = [dp, dp, ctypes.c_long, ctypes.c_int]
        for name in ("oracle_c2c_nd", "oracle_r2c_nd", "oracle_c2r_nd"):
            getattr(lib, name).argtypes = [fp, dp, ctypes.c_long, lp, ctypes.c_int, ctypes.c_int]
        lib.oracle_lcg_uniform.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.c_float, ctypes.c_float,
                                           fp, ctypes.c_long]
        lib.oracle_shift2d.argtypes = [fp, fp, ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_long,
                                       ctypes.c_int, ctypes.c_double, ctypes.c_int]
        lib.oracle_istft.argtypes = [fp, fp, dp, ctypes.c_long, ctypes.c_long, ctypes.c_long, ctypes.c_long]
        lib.oracle_rel_l2.argtypes = [fp, dp, ctypes.c_long]
        lib.oracle_rel_l2.restype = ctypes.c_double
        _lib = lib
    return _lib


def _ptr(a, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _dims(dims):
    arr = (ctypes.c_long * len(dims))(*[int(d) for d in dims])
    return arr


def dft_direct(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """O(N^2) long-double DFT of one complex vector (sign -1 forward, +1 inverse, no scaling)."""
    x = np.ascontiguousarray(x, dtype=np.complex128).ravel()
    out = np.empty_like(x)
    _load().oracle_dft_direct(_ptr(x.view(np.float64), ctypes.c_double), _ptr(out.view(np.float64), ctypes.c_double),
                              x.size, 1 if inverse else -1)
    return out


def fft_1d(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """Mixed-radix FP64 FFT of one complex vector (sign -1 forward, +1 inverse, no scaling)."""
    x = np.ascontiguousarray(x, dtype=np.complex128).ravel()
    out = np.empty_like(x)
    _load().oracle_fft(_ptr(x.view(np.float64), ctypes.c_double), _ptr(out.view(np.float64), ctypes.c_double),
                       x.size, 1 if inverse else -1)
    return out


def c2c_nd(x: np.ndarray, ndim: int, inverse: bool = False) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.complex64)
    dims = x.shape[x.ndim - ndim:]
    batch = int(np.prod(x.shape[:x.ndim - ndim], dtype=np.int64))
    out = np.empty(x.shape, dtype=np.complex128)
    _load().oracle_c2c_nd(_ptr(x.view(np.float32), ctypes.c_float), _ptr(out.view(np.float64), ctypes.c_double),
                          batch, _dims(dims), ndim, int(inverse))
    return out


def r2c_nd(x: np.ndarray, ndim: int, inverse: bool = False) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    dims = x.shape[x.ndim - ndim:]
    batch = int(np.prod(x.shape[:x.ndim - ndim], dtype=np.int64))
    out_shape = x.shape[:-1] + (x.shape[-1] // 2 + 1,)
    out = np.empty(out_shape, dtype=np.complex128)
    _load().oracle_r2c_nd(_ptr(x, ctypes.c_float), _ptr(out.view(np.float64), ctypes.c_double),
                          batch, _dims(dims), ndim, int(inverse))
    return out


def c2r_nd(x: np.ndarray, dims, inverse: bool = False) -> np.ndarray:
    """x: complex [..., dims[:-1]..., dims[-1]//2+1]; returns float64 [..., dims...]."""
    dims = tuple(int(d) for d in dims)
    ndim = len(dims)
    x = np.ascontiguousarray(x, dtype=np.complex64)
    if x.shape[x.ndim - ndim:] != dims[:-1] + (dims[-1] // 2 + 1,):
        raise ValueError(f"c2r_nd: input shape {x.shape} does not match dims {dims}")
    batch = int(np.prod(x.shape[:x.ndim - ndim], dtype=np.int64))
    out = np.empty(x.shape[:x.ndim - ndim] + dims, dtype=np.float64)
    _load().oracle_c2r_nd(_ptr(x.view(np.float32), ctypes.c_float), _ptr(out, ctypes.c_double),
                          batch, _dims(dims), ndim, int(inverse))
    return out


def c2c(x, inverse=False):
    return c2c_nd(x, 1, inverse)


def r2c(x, inverse=False):
    return r2c_nd(x, 1, inverse)


def c2r(x, n, inverse=False):
    return c2r_nd(x, (n,), inverse)


def shift2d(x: np.ndarray, row: int, col: int, batch: int, dist: int = 0, inverse: bool = False,
            scale: float = 1.0) -> np.ndarray:
    """asdFftShift (inverse=False) / asdIfftShift (inverse=True) on complex64 slices of a flat
    input (slice s at element s*dist; dist <= 0 means row*col).  Returns complex64 (batch,row,col)."""
    dist = row * col if dist <= 0 else dist
    src = np.ascontiguousarray(x, dtype=np.complex64).reshape(-1)
    out = np.empty((batch, row, col), dtype=np.complex64)
    apply = 1 if (inverse and scale != 1.0) else 0
    _load().oracle_shift2d(_ptr(src.view(np.float32), ctypes.c_float), _ptr(out.view(np.float32), ctypes.c_float),
                           row, col, batch, dist, int(inverse), float(scale), apply)
    return out


def istft_out_len(nfft: int, hop: int, nframes: int) -> int:
    """Output samples per channel of the ISTFT (center = true, length = 0), istft.cpp:318-325."""
    return nfft + hop * (nframes - 1) - 2 * (nfft // 2)


def istft(spec: np.ndarray, window: np.ndarray, hop: int) -> np.ndarray:
    """ISTFT of a complex64 (channel, nfft, nframes) spectrum with a float32 window of nfft
    samples (center, not normalised, two-sided); returns complex128 (channel, out_len)."""
    spec = np.ascontiguousarray(spec, dtype=np.complex64)
    window = np.ascontiguousarray(window, dtype=np.float32)
    channel, nfft, nframes = spec.shape
    if window.shape != (nfft,):
        raise ValueError("istft: window must have nfft samples")
    out = np.empty((channel, istft_out_len(nfft, hop, nframes)), dtype=np.complex128)
    _load().oracle_istft(_ptr(spec.view(np.float32), ctypes.c_float), _ptr(window, ctypes.c_float),
                         _ptr(out.view(np.float64), ctypes.c_double), channel, nfft, nframes, hop)
    return out


def lcg_uniform(n: int, state: int = 2026, lower: float = -300.0, upper: float = 300.0):
    """The unit-test LCG stream (float_util.h:23-36); returns (values float32[n], next_state)."""
    st = ctypes.c_uint32(state)
    out = np.empty(n, dtype=np.float32)
    _load().oracle_lcg_uniform(ctypes.byref(st), lower, upper, _ptr(out, ctypes.c_float), n)
    return out, st.value


def ut_complex_input(shape, state: int = 2026):
    """complex64 tensor filled like OpTestCreateHostRandTensor (util.cpp:324-338): re, im per element."""
    n = int(np.prod(shape))
    vals, _ = lcg_uniform(2 * n, state)
    return vals.view(np.complex64).reshape(shape)


def ut_real_input(shape, state: int = 2026):
    n = int(np.prod(shape))
    vals, _ = lcg_uniform(n, state)
    return vals.reshape(shape)


def rel_l2(got: np.ndarray, golden: np.ndarray) -> float:
    """||got - golden||_2 / ||golden||_2 (run.py:503-506), computed in long double."""
    g = np.ascontiguousarray(got)
    if np.iscomplexobj(g):
        g = g.astype(np.complex64).view(np.float32)
    else:
        g = g.astype(np.float32)
    gold = np.ascontiguousarray(golden)
    gold = gold.astype(np.complex128).view(np.float64) if np.iscomplexobj(gold) else gold.astype(np.float64)
    if g.size != gold.size:
        raise ValueError("rel_l2: size mismatch")
    return float(_load().oracle_rel_l2(_ptr(g.ravel(), ctypes.c_float), _ptr(gold.ravel(), ctypes.c_double), g.size))


def rel_l2_per_row(got: np.ndarray, golden: np.ndarray, ndim: int = 1) -> np.ndarray:
    """Relative L2 error of every transform of a batch (leading axes), as float64."""
    got = np.asarray(got)
    golden = np.asarray(golden)
    lead = got.shape[:got.ndim - ndim]
    g2 = got.reshape(int(np.prod(lead, dtype=np.int64)), -1).astype(np.complex128)
    o2 = golden.reshape(g2.shape[0], -1).astype(np.complex128)
    num = np.linalg.norm(g2 - o2, axis=1)
    den = np.linalg.norm(o2, axis=1)
    return np.where(den > 0, num / np.where(den > 0, den, 1.0), num)
