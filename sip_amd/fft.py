"""Python host side over the C ABI (libasdfft.so).

`Plan` wraps one asdFft handle (create -> MakePlan -> workspace -> exec ->
destroy, the same sequence as the reference's tests, e.g.
tests/ut/unittest/fft/1d/c2c/test_fft_c2c_1d.cpp:103-136).

`asd_fft_c2c / asd_fft_r2c / asd_fft_c2r` mirror the reference's torch plugin
operators (sip_pta/torch_sip/__init__.py:427-447, sip_pta/csrc/fft/asd_fft_c2c.cpp:20-55,
asd_fft_r2c.cpp, asd_fft_c2r.cpp): a rank-2/3/4 tensor becomes a 1D/2D/3D plan
whose batch is the leading dimension; plans are kept in a small LRU cache
(sip_pta/csrc/FFTMixCache.cpp:18-40, capacity 10) and the workspace comes from
torch's caching allocator.  PyTorch is only used for device memory and streams;
every transform runs in the HIP kernels of libasdfft.so.
"""
from __future__ import annotations

import collections
import ctypes
import threading

from . import _native as N

_TYPE = {"c2c": N.ASCEND_FFT_C2C, "r2c": N.ASCEND_FFT_R2C, "c2r": N.ASCEND_FFT_C2R}


class Plan:
    """One plan handle.  dims: FFT sizes (1-3 of them); kind: 'c2c' | 'r2c' | 'c2r'."""

    def __init__(self, kind: str, dims, batch: int, inverse: bool = False, vertical: bool = False):
        self.L = N.lib()
        self.kind = kind
        self.dims = tuple(int(d) for d in dims)
        self.batch = int(batch)
        self.inverse = bool(inverse)
        self.vertical = bool(vertical)
        h = ctypes.c_void_p()
        N.check("asdFftCreate", self.L.asdFftCreate(ctypes.byref(h)))
        self.handle = h
        direction = N.ASCEND_FFT_INVERSE if inverse else N.ASCEND_FFT_FORWARD
        t = _TYPE[kind]
        try:
            if len(self.dims) == 1:
                st = self.L.asdFftMakePlan1D(h, self.dims[0], t, direction, self.batch,
                                             N.ASCEND_FFT_VERTICAL if vertical else N.ASCEND_FFT_HORIZONTAL)
            elif len(self.dims) == 2:
                st = self.L.asdFftMakePlan2D(h, self.dims[0], self.dims[1], t, direction, self.batch)
            elif len(self.dims) == 3:
                st = self.L.asdFftMakePlan3D(h, self.dims[0], self.dims[1], self.dims[2], t, direction, self.batch)
            else:
                raise ValueError("1 to 3 FFT dimensions")
            N.check("asdFftMakePlan", st)
            size = ctypes.c_size_t(0)
            N.check("asdFftGetWorkspaceSize", self.L.asdFftGetWorkspaceSize(h, ctypes.byref(size)))
            self.workspace_size = size.value
        except Exception:
            self.L.asdFftDestroy(h)
            self.handle = None
            raise
        self._workspace = None
        self._stream = None

    # --- shapes -------------------------------------------------------------------------
    def in_shape(self):
        d = self.dims
        if self.vertical:
            return (d[0], self.batch)
        if self.kind == "c2r":
            return (self.batch,) + d[:-1] + (d[-1] // 2 + 1,)
        return (self.batch,) + d

    def out_shape(self):
        d = self.dims
        if self.vertical:
            return (d[0], self.batch)
        if self.kind == "r2c":
            return (self.batch,) + d[:-1] + (d[-1] // 2 + 1,)
        return (self.batch,) + d

    def describe(self) -> str:
        need = ctypes.c_size_t(0)
        self.L.asdFftExtDescribePlan(self.handle, None, 0, ctypes.byref(need))
        buf = ctypes.create_string_buffer(need.value + 1)
        self.L.asdFftExtDescribePlan(self.handle, buf, len(buf), None)
        return buf.value.decode()

    def step_bytes(self):
        n = ctypes.c_int32(0)
        self.L.asdFftExtPlanSteps(self.handle, ctypes.byref(n), None)
        arr = (ctypes.c_double * max(1, n.value))()
        self.L.asdFftExtPlanSteps(self.handle, ctypes.byref(n), arr)
        return [arr[i] for i in range(n.value)]

    # --- execution (torch device tensors) ---------------------------------------------------
    def bind_workspace(self, device):
        import torch
        if self.workspace_size and (self._workspace is None or self._workspace.device != device):
            self._workspace = torch.empty(self.workspace_size, dtype=torch.uint8, device=device)
        if self.workspace_size:
            N.check("asdFftSetWorkspace", self.L.asdFftSetWorkspace(self.handle, self._workspace.data_ptr()))

    def bind_stream(self, stream_handle: int):
        if stream_handle and stream_handle != self._stream:
            N.check("asdFftSetStream", self.L.asdFftSetStream(self.handle, ctypes.c_void_p(stream_handle)))
            self._stream = stream_handle

    def execute(self, x, out=None):
        """Runs the plan on torch device tensor x (complex64 or float32); returns out."""
        import torch
        if out is None:
            dt = torch.float32 if self.kind == "c2r" else torch.complex64
            out = torch.empty(self.out_shape(), dtype=dt, device=x.device)
        self.bind_workspace(x.device)
        self.bind_stream(torch.cuda.current_stream(x.device).cuda_stream)
        xin = N.AclTensor(x.shape, N.ACL_FLOAT if x.dtype == torch.float32 else N.ACL_COMPLEX64, x.data_ptr())
        xout = N.AclTensor(out.shape, N.ACL_FLOAT if out.dtype == torch.float32 else N.ACL_COMPLEX64,
                           out.data_ptr())
        fn = {"c2c": self.L.asdFftExecC2C, "r2c": self.L.asdFftExecR2C, "c2r": self.L.asdFftExecC2R}[self.kind]
        try:
            N.check(f"asdFftExec{self.kind.upper()}", fn(self.handle, xin.ptr, xout.ptr))
        finally:
            xin.close()
            xout.close()
        return out

    def synchronize(self):
        N.check("asdFftSynchronize", self.L.asdFftSynchronize(self.handle))

    def close(self):
        if self.handle is not None:
            self.L.asdFftDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PlanCache:
    """LRU of plans keyed by (kind, dims, batch, inverse); the reference keeps 10."""

    def __init__(self, capacity: int = 10):
        self.capacity = capacity
        self._d = collections.OrderedDict()
        self._mu = threading.Lock()

    def get(self, kind, dims, batch, inverse):
        key = (kind, tuple(dims), int(batch), bool(inverse))
        with self._mu:
            p = self._d.get(key)
            if p is not None:
                self._d.move_to_end(key)
                return p
        p = Plan(kind, dims, batch, inverse)
        with self._mu:
            self._d[key] = p
            while len(self._d) > self.capacity:
                _, old = self._d.popitem(last=False)
                old.close()
        return p


_cache = PlanCache()


def _split(shape, kind):
    if len(shape) < 2 or len(shape) > 4:
        raise ValueError(f"input must have 2, 3 or 4 dimensions, got {len(shape)}")
    return int(shape[0]), tuple(int(s) for s in shape[1:])


def asd_fft_c2c(x, is_forward: bool = True):
    """complex64 (batch, N) / (batch, X, Y) / (batch, X, Y, Z) -> same shape."""
    import torch
    if x.dtype != torch.complex64:
        raise TypeError(f"asd_fft_c2c: input must be complex64, got {x.dtype}")
    x = x.contiguous()
    batch, dims = _split(x.shape, "c2c")
    return _cache.get("c2c", dims, batch, not is_forward).execute(x)


def asd_fft_r2c(x, is_forward: bool = True):
    """float32 (batch, ..., N) -> complex64 (batch, ..., N/2+1)."""
    import torch
    if x.dtype != torch.float32:
        raise TypeError("asd_fft_r2c: input tensor must be float32")
    x = x.contiguous()
    batch, dims = _split(x.shape, "r2c")
    return _cache.get("r2c", dims, batch, not is_forward).execute(x)


def asd_fft_c2r(x, is_forward: bool = True):
    """complex64 (batch, ..., M) -> float32 (batch, ..., 2(M-1)), as the reference plugin."""
    import torch
    if x.dtype != torch.complex64:
        raise TypeError("asd_fft_c2r: input tensor must be complex64")
    x = x.contiguous()
    batch, dims = _split(x.shape, "c2r")
    dims = dims[:-1] + (2 * (dims[-1] - 1),)
    return _cache.get("c2r", dims, batch, not is_forward).execute(x)
