"""Python host side over the C ABI (libasdfft.so).

`Plan` wraps one asdFft handle (create -> MakePlan -> workspace -> exec ->
destroy, the same sequence as the reference's tests, e.g.
tests/ut/unittest/fft/1d/c2c/test_fft_c2c_1d.cpp:103-136).

`asd_fft_c2c / asd_fft_c2c_sep / asd_fft_r2c / asd_fft_c2r / fft_shift / ifft_shift`
are the reference's torch plugin functions (sip_pta/torch_sip/__init__.py:427-540):
each calls the torch operator torch.ops.torch_sip.* of libasdfft_torch.so
(sip_amd/csrc/torch_ops.cpp), which keeps plans in a 10-entry LRU, takes the
workspace from torch's caching allocator and launches on the current stream.
PyTorch is only used for device memory and streams; every transform runs in the
HIP kernels of libasdfft.so.
"""
from __future__ import annotations

import ctypes

from . import _native as N

_TYPE = {"c2c": N.ASCEND_FFT_C2C, "r2c": N.ASCEND_FFT_R2C, "c2r": N.ASCEND_FFT_C2R,
         "c2c_sep": N.ASCEND_FFT_C2C_SEP}


class Plan:
    """One plan handle.  dims: FFT sizes (1-3 of them); kind: 'c2c' | 'r2c' | 'c2r' | 'c2c_sep'
    (C2C on separated real / imaginary float32 planes, run by execute_separated).  in_shift /
    out_shift ('fft' | 'ifft') and scale fold asdFftShift / asdIfftShift permutations and an
    output multiplier into a 1D or 2D C2C plan (asdFftExtSetShift)."""

    def __init__(self, kind: str, dims, batch: int, inverse: bool = False, vertical: bool = False,
                 precision: str = "fp32", in_shift: str | None = None, out_shift: str | None = None,
                 scale: float = 1.0):
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
            if precision != "fp32":
                mode = {"split16": N.ASDFFT_PRECISION_SPLIT16}[precision]
                N.check("asdFftExtSetPrecision", self.L.asdFftExtSetPrecision(h, mode))
            if in_shift or out_shift or scale != 1.0:
                modes = {None: N.ASDFFT_SHIFT_NONE, "fft": N.ASDFFT_SHIFT_FFT, "ifft": N.ASDFFT_SHIFT_IFFT}
                N.check("asdFftExtSetShift", self.L.asdFftExtSetShift(h, modes[in_shift], modes[out_shift], scale))
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
        # torch's default stream is the null stream, refused by asdFftSetStream; hipStreamLegacy
        # ((hipStream_t)1) names that same stream
        stream_handle = stream_handle or 1
        if stream_handle != self._stream:
            N.check("asdFftSetStream", self.L.asdFftSetStream(self.handle, ctypes.c_void_p(stream_handle)))
            self._stream = stream_handle

    def execute(self, x, out=None):
        """Runs the plan on torch device tensor x (complex64 or float32); returns out."""
        import torch
        if self.kind == "c2c_sep":
            raise TypeError("a c2c_sep plan runs through execute_separated")
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

    def execute_separated(self, re, im, out_re=None, out_im=None):
        """asdFftExecC2CSeparated on float32 device planes re, im (one shape); returns
        (out_re, out_im).  out_re / out_im may be re / im (in place)."""
        import torch
        if out_re is None:
            out_re = torch.empty_like(re)
        if out_im is None:
            out_im = torch.empty_like(im)
        self.bind_workspace(re.device)
        self.bind_stream(torch.cuda.current_stream(re.device).cuda_stream)
        ts = [N.AclTensor(t.shape, N.ACL_FLOAT if t.dtype == torch.float32 else N.ACL_COMPLEX64, t.data_ptr())
              for t in (re, im, out_re, out_im)]
        try:
            N.check("asdFftExecC2CSeparated", self.L.asdFftExecC2CSeparated(self.handle, *[t.ptr for t in ts]))
        finally:
            for t in ts:
                t.close()
        return out_re, out_im

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


class Istft:
    """One ISTFT plan (asdFftIstftMakePlan / asdFftExecIstft) for complex64 spectra of shape
    (channel, nfft, nframes): center, not normalised, two-sided, complex output of shape
    (channel, nfft + hop*(nframes-1) - 2*(nfft//2)).  Follows the reference's calling sequence
    (tests/ut/unittest/fft/istft/test_fft_istft.cpp:118-150)."""

    def __init__(self, channel: int, nfft: int, nframes: int, hop: int):
        self.L = N.lib()
        self.channel, self.nfft, self.nframes, self.hop = int(channel), int(nfft), int(nframes), int(hop)
        self.out_len = self.nfft + self.hop * (self.nframes - 1) - 2 * (self.nfft // 2)
        h = ctypes.c_void_p()
        N.check("asdFftCreate", self.L.asdFftCreate(ctypes.byref(h)))
        self.handle = h
        shape = N.AclTensor((self.channel, self.nfft, self.nframes), N.ACL_COMPLEX64, 0)
        try:
            N.check("asdFftIstftMakePlan", self.L.asdFftIstftMakePlan(h, shape.ptr, self.nfft, self.hop, self.nfft,
                                                                      True, False, False, 0, True))
            size = ctypes.c_size_t(0)
            N.check("asdFftGetWorkspaceSize", self.L.asdFftGetWorkspaceSize(h, ctypes.byref(size)))
            self.workspace_size = size.value
        except Exception:
            self.L.asdFftDestroy(h)
            self.handle = None
            raise
        finally:
            shape.close()
        self._workspace = None
        self._stream = None

    def describe(self) -> str:
        need = ctypes.c_size_t(0)
        self.L.asdFftExtDescribePlan(self.handle, None, 0, ctypes.byref(need))
        buf = ctypes.create_string_buffer(need.value + 1)
        self.L.asdFftExtDescribePlan(self.handle, buf, len(buf), None)
        return buf.value.decode()

    bind_workspace = Plan.bind_workspace
    bind_stream = Plan.bind_stream

    def execute(self, spec, window, out=None):
        """spec: complex64 (channel, nfft, nframes) device tensor; window: float32 (nfft,)."""
        import torch
        if out is None:
            out = torch.empty((self.channel, self.out_len), dtype=torch.complex64, device=spec.device)
        self.bind_workspace(spec.device)
        self.bind_stream(torch.cuda.current_stream(spec.device).cuda_stream)
        ts = [N.AclTensor(t.shape, N.ACL_FLOAT if t.dtype == torch.float32 else N.ACL_COMPLEX64, t.data_ptr())
              for t in (spec, window, out)]
        try:
            N.check("asdFftExecIstft", self.L.asdFftExecIstft(self.handle, *[t.ptr for t in ts]))
        finally:
            for t in ts:
                t.close()
        return out

    def close(self):
        if self.handle is not None:
            self.L.asdFftDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _side_bytes(kind: str, n: int):
    """(input bytes, output bytes) of one transform of this kind and length."""
    h = n // 2 + 1
    return {"c2c": (8 * n, 8 * n), "r2c": (4 * n, 8 * h), "c2r": (8 * h, 4 * n)}[kind]


class RaggedPlan:
    """A ragged batch (asdFftExtMakePlanRagged): entries of different kinds and lengths
    transformed by one exec over a packed input buffer and a packed output buffer.

    entries: (kind, n, count, inverse) tuples, optionally followed by explicit byte
    offsets (in_offset, out_offset); without them the entries are packed back to back in
    order, each starting on an 8-byte boundary.
    """

    def __init__(self, entries, graph: bool = False, autotune: bool = False, per_budget: bool = False):
        self.L = N.lib()
        rows = []
        ioff = ooff = 0
        for e in entries:
            kind, n, count, inverse = e[0], int(e[1]), int(e[2]), bool(e[3])
            ib, ob = _side_bytes(kind, n)
            if len(e) >= 6:
                i0, o0 = int(e[4]), int(e[5])
            else:
                i0, o0 = ioff, ooff
                ioff = (ioff + ib * count + 7) // 8 * 8
                ooff = (ooff + ob * count + 7) // 8 * 8
            rows.append((kind, n, count, inverse, i0, o0))
        self.entries = rows
        arr = (N.RaggedEntry * len(rows))()
        for i, (kind, n, count, inverse, i0, o0) in enumerate(rows):
            arr[i].type = _TYPE[kind]
            arr[i].direction = N.ASCEND_FFT_INVERSE if inverse else N.ASCEND_FFT_FORWARD
            arr[i].n = n
            arr[i].count = count
            arr[i].in_offset = i0
            arr[i].out_offset = o0
        h = ctypes.c_void_p()
        N.check("asdFftCreate", self.L.asdFftCreate(ctypes.byref(h)))
        self.handle = h
        flags = (N.ASDFFT_RAGGED_GRAPH if graph else 0) | (N.ASDFFT_RAGGED_AUTOTUNE if autotune else 0) | \
            (N.ASDFFT_RAGGED_PER_BUDGET if per_budget else 0)
        try:
            N.check("asdFftExtMakePlanRagged", self.L.asdFftExtMakePlanRagged(h, arr, len(rows), flags))
        except Exception:
            self.L.asdFftDestroy(h)
            self.handle = None
            raise
        ie, oe = ctypes.c_int64(0), ctypes.c_int64(0)
        by, fl = ctypes.c_double(0), ctypes.c_double(0)
        N.check("asdFftExtRaggedInfo", self.L.asdFftExtRaggedInfo(h, ctypes.byref(ie), ctypes.byref(oe),
                                                                   ctypes.byref(by), ctypes.byref(fl)))
        self.in_extent, self.out_extent = ie.value, oe.value
        self.algorithmic_bytes, self.flops = by.value, fl.value
        self._stream = None

    def describe(self) -> str:
        need = ctypes.c_size_t(0)
        self.L.asdFftExtDescribePlan(self.handle, None, 0, ctypes.byref(need))
        buf = ctypes.create_string_buffer(need.value + 1)
        self.L.asdFftExtDescribePlan(self.handle, buf, len(buf), None)
        return buf.value.decode()

    def views(self, buf, side: str):
        """Per-entry tensor views into a packed uint8 buffer: side 'in' or 'out'."""
        import torch
        out = []
        for kind, n, count, _, i0, o0 in self.entries:
            h = n // 2 + 1
            real = (kind == "r2c" and side == "in") or (kind == "c2r" and side == "out")
            width = n if (kind == "c2c" or real) else h
            dt = torch.float32 if real else torch.complex64
            esz = 4 if real else 8
            off = i0 if side == "in" else o0
            out.append(buf[off:off + count * width * esz].view(dt).view(count, width))
        return out

    def execute(self, in_buf, out_buf, stream_handle: int | None = None):
        """Enqueues the batch; in_buf / out_buf are device tensors covering the extents."""
        import torch
        if in_buf.numel() * in_buf.element_size() < self.in_extent:
            raise ValueError("input buffer smaller than the plan's input extent")
        if out_buf.numel() * out_buf.element_size() < self.out_extent:
            raise ValueError("output buffer smaller than the plan's output extent")
        s = stream_handle if stream_handle is not None else torch.cuda.current_stream(in_buf.device).cuda_stream
        if s and s != self._stream:
            N.check("asdFftSetStream", self.L.asdFftSetStream(self.handle, ctypes.c_void_p(s)))
            self._stream = s
        N.check("asdFftExtExecRagged", self.L.asdFftExtExecRagged(self.handle, ctypes.c_void_p(in_buf.data_ptr()),
                                                                   ctypes.c_void_p(out_buf.data_ptr())))
        return out_buf

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


class Server:
    """Serving loop (asdFftExtServe*): continuous batching of FFT requests under an HBM budget.

    Requests are (kind, dims, batch, inverse, input tensor, output tensor); dims has one
    (1D) or two (2D) sizes, the real length last for r2c / c2r.
    """

    def __init__(self, hbm_budget: int = 8 << 30, slots: int = 3, max_batch: int = 256, batching: bool = True):
        self.L = N.lib()
        cfg = N.ServeConfig(hbm_budget, slots, max_batch, 1 if batching else 0, 0)
        h = ctypes.c_void_p()
        N.check("asdFftExtServeCreate", self.L.asdFftExtServeCreate(ctypes.byref(h), ctypes.byref(cfg)))
        self.handle = h

    @staticmethod
    def request(kind, dims, batch, inverse, x_ptr: int, y_ptr: int):
        dims = tuple(int(d) for d in dims)
        r = N.ServeRequest()
        r.type = _TYPE[kind]
        r.direction = N.ASCEND_FFT_INVERSE if inverse else N.ASCEND_FFT_FORWARD
        r.rank = len(dims)
        r.dims[0] = dims[0]
        r.dims[1] = dims[1] if len(dims) > 1 else 0
        r.batch = int(batch)
        r.input = x_ptr
        r.output = y_ptr
        return r

    def submit(self, kind, dims, batch, inverse, x, y) -> int:
        r = self.request(kind, dims, batch, inverse, x.data_ptr(), y.data_ptr())
        t = ctypes.c_uint64(0)
        N.check("asdFftExtServeSubmit", self.L.asdFftExtServeSubmit(self.handle, ctypes.byref(r), ctypes.byref(t)))
        return t.value

    def wait(self, ticket: int, timeout_us: int = -1) -> float:
        lat = ctypes.c_double(0)
        N.check("asdFftExtServeWait", self.L.asdFftExtServeWait(self.handle, ticket, timeout_us, ctypes.byref(lat)))
        return lat.value

    def replay(self, requests, arrival_us):
        """requests: list of N.ServeRequest; arrival_us: offsets from the call.  Returns latencies (us)."""
        n = len(requests)
        arr = (N.ServeRequest * n)(*requests)
        t = (ctypes.c_double * n)(*[float(a) for a in arrival_us])
        lat = (ctypes.c_double * n)()
        N.check("asdFftExtServeReplay", self.L.asdFftExtServeReplay(self.handle, arr, t, n, lat))
        return [lat[i] for i in range(n)]

    def stats(self) -> dict:
        c = N.ServeCounters()
        N.check("asdFftExtServeStats", self.L.asdFftExtServeStats(self.handle, ctypes.byref(c)))
        return {name: getattr(c, name) for name, _ in N.ServeCounters._fields_}

    def close(self):
        if self.handle is not None:
            self.L.asdFftExtServeDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ops():
    return N.torch_ops()


def asd_fft_c2c(x, is_forward: bool = True):
    """complex64 (batch, N) / (batch, X, Y) / (batch, X, Y, Z) -> same shape."""
    return _ops().asd_fft_c2c(x, is_forward)


def asd_fft_c2c_sep(in_real, in_imag, is_forward: bool = True):
    """float32 planes (batch, N) -> 1D or (batch, X, Y, Z) -> 3D; returns (out_real, out_imag)."""
    return _ops().asd_fft_c2c_sep(in_real, in_imag, is_forward)


def asd_fft_r2c(x, is_forward: bool = True):
    """float32 (batch, ..., N) -> complex64 (batch, ..., N/2+1)."""
    return _ops().asd_fft_r2c(x, is_forward)


def asd_fft_c2r(x, is_forward: bool = True):
    """complex64 (batch, ..., M) -> float32 (batch, ..., 2(M-1)), as the reference plugin."""
    return _ops().asd_fft_c2r(x, is_forward)


def fft_shift(x, row: int | None = None, col: int | None = None, batch: int | None = None, dist: int = 0):
    """Zero frequency to the centre of each (row, col) slice: out[s][(r + ceil(row/2)) % row]
    [(c + ceil(col/2)) % col] = in[s][r][c].  A 3D (batch, row, col) tensor may omit
    row/col/batch; a flat layout names them (and dist, the slice distance, <= 0: contiguous).
    Mirrors sip_pta/torch_sip/__init__.py fft_shift."""
    if row is None or col is None or batch is None:
        if x.dim() != 3:
            raise ValueError(f"fft_shift: auto-infer requires a 3D [batch, row, col] tensor, got {x.dim()}D shape "
                             f"{tuple(x.shape)}. Explicitly pass row/col/batch for flattened input.")
        return _ops().asd_fft_shift_auto(x)
    return _ops().asd_fft_shift(x, row, col, batch, dist)


def ifft_shift(x, row: int | None = None, col: int | None = None, batch: int | None = None, dist: int = 0,
               scale: float = 1.0):
    """Inverse of fft_shift (floor instead of ceil) times float32(scale); same layouts.
    Mirrors sip_pta/torch_sip/__init__.py ifft_shift."""
    if row is None or col is None or batch is None:
        if x.dim() != 3:
            raise ValueError(f"ifft_shift: auto-infer requires a 3D [batch, row, col] tensor, got {x.dim()}D shape "
                             f"{tuple(x.shape)}. Explicitly pass row/col/batch for flattened input.")
        if dist != 0 or scale != 1.0:
            return _ops().asd_ifft_shift(x, x.shape[1], x.shape[2], x.shape[0], dist, scale)
        return _ops().asd_ifft_shift_auto(x)
    return _ops().asd_ifft_shift(x, row, col, batch, dist, scale)
