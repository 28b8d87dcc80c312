"""asdFftExtExecNoAlloc: the exec entry point that allocates nothing (its workspace is the
library's per-GPU static device array) against the ordinary exec with a caller-bound
workspace, through the C ABI, on the same plan and the same input.

The kernels and their order are the same, only the workspace address differs, so the outputs
must be identical bit for bit.  Two calls in a row share the one scratch array: the second
(a smaller plan) must not be disturbed by what the first left in it, and the first plan must
still agree when run again afterwards.
"""
import ctypes

import pytest

from sip_amd import _native as N

pytestmark = pytest.mark.gpu

C2C, C2R, R2C = N.ASCEND_FFT_C2C, N.ASCEND_FFT_C2R, N.ASCEND_FFT_R2C
FWD, INV = N.ASCEND_FFT_FORWARD, N.ASCEND_FFT_INVERSE
SCRATCH = 64 << 20


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    N.lib()
    return torch.device("cuda:0")


class Case:
    """One plan, its input, and outputs from both exec entry points."""

    def __init__(self, L, dev, kind, n, batch, direction, seed):
        import torch
        self.L, self.kind, self.n, self.batch = L, kind, n, batch
        self.h = ctypes.c_void_p()
        assert L.asdFftCreate(ctypes.byref(self.h)) == 0
        assert L.asdFftMakePlan1D(self.h, n, {"c2c": C2C, "r2c": R2C, "c2r": C2R}[kind], direction, batch,
                                  N.ASCEND_FFT_HORIZONTAL) == 0
        size = ctypes.c_size_t()
        assert L.asdFftGetWorkspaceSize(self.h, ctypes.byref(size)) == 0
        self.ws_bytes = size.value
        self.ws = torch.empty(max(size.value, 1), dtype=torch.uint8, device=dev)
        assert L.asdFftSetWorkspace(self.h, ctypes.c_void_p(self.ws.data_ptr())) == 0
        self.stream = torch.cuda.Stream(device=dev)  # a stream of its own (SetStream refuses NULL)
        assert L.asdFftSetStream(self.h, ctypes.c_void_p(self.stream.cuda_stream)) == 0
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        h = n // 2 + 1
        if kind == "c2c":
            self.x = torch.complex(torch.rand(batch, n, generator=g, device=dev) * 2 - 1,
                                   torch.rand(batch, n, generator=g, device=dev) * 2 - 1)
            out_shape, out_dt = (batch, n), torch.complex64
        elif kind == "r2c":
            self.x = torch.rand(batch, n, generator=g, device=dev) * 2 - 1
            out_shape, out_dt = (batch, h), torch.complex64
        else:
            self.x = torch.complex(torch.rand(batch, h, generator=g, device=dev) * 2 - 1,
                                   torch.rand(batch, h, generator=g, device=dev) * 2 - 1)
            out_shape, out_dt = (batch, n), torch.float32
        self.y_ref = torch.empty(out_shape, dtype=out_dt, device=dev)
        self.y = torch.full(out_shape, float("nan"), dtype=out_dt, device=dev) if out_dt == torch.float32 else \
            torch.full(out_shape, complex(float("nan"), 0.0), dtype=out_dt, device=dev)
        self.x_keep = self.x.clone()

    def tensors(self, y):
        import torch
        dt = lambda t: N.ACL_FLOAT if t.dtype == torch.float32 else N.ACL_COMPLEX64
        return N.AclTensor(self.x.shape, dt(self.x), self.x.data_ptr()), N.AclTensor(y.shape, dt(y), y.data_ptr())

    def run(self, no_alloc):
        import torch
        torch.cuda.synchronize()  # inputs and outputs were written on torch's stream
        y = self.y if no_alloc else self.y_ref
        ti, to = self.tensors(y)
        try:
            if no_alloc:
                st = self.L.asdFftExtExecNoAlloc(self.h, ti.ptr, to.ptr)
            else:
                fn = {"c2c": self.L.asdFftExecC2C, "r2c": self.L.asdFftExecR2C, "c2r": self.L.asdFftExecC2R}
                st = fn[self.kind](self.h, ti.ptr, to.ptr)
        finally:
            ti.close()
            to.close()
        torch.cuda.synchronize()
        assert torch.equal(self.x, self.x_keep), "the input was modified"
        return st

    def close(self):
        self.L.asdFftDestroy(self.h)


@pytest.mark.parametrize("kind,n,big,small,direction", [
    ("c2c", 65536, 64, 8, FWD),     # four-step: 32 MiB of workspace, then 4 MiB
    ("c2c", 65536, 64, 8, INV),
    ("r2c", 131072, 16, 3, FWD),    # even real beyond one workgroup: workspace of two complex buffers
    ("c2r", 131072, 16, 3, INV),
    ("c2c", 4096, 512, 7, FWD),     # no workspace at all
])
def test_no_alloc_matches_bound_workspace(torch_dev, kind, n, big, small, direction):
    import torch
    L = N.lib()
    first = Case(L, torch_dev, kind, n, big, direction, 100 + big)
    second = Case(L, torch_dev, kind, n, small, direction, 200 + small)
    try:
        assert first.ws_bytes <= SCRATCH and second.ws_bytes <= SCRATCH
        for case in (first, second):
            print(f"{kind} n={n} batch={case.batch} workspace={case.ws_bytes} B", flush=True)
            assert case.run(no_alloc=False) == 0
            assert case.run(no_alloc=True) == 0
            assert torch.equal(case.y.view(torch.float32), case.y_ref.view(torch.float32)), (kind, n, case.batch)
        # the first plan again, after the smaller one used the same scratch
        first.y.fill_(0)
        assert first.run(no_alloc=True) == 0
        assert torch.equal(first.y.view(torch.float32), first.y_ref.view(torch.float32))
    finally:
        first.close()
        second.close()


def test_no_alloc_refuses_plans_beyond_the_scratch(torch_dev):
    """A 65536-point batch of 256 needs 128 MiB of workspace: refused before any launch."""
    L = N.lib()
    h = ctypes.c_void_p()
    assert L.asdFftCreate(ctypes.byref(h)) == 0
    try:
        assert L.asdFftMakePlan1D(h, 65536, C2C, FWD, 256, N.ASCEND_FFT_HORIZONTAL) == 0
        size = ctypes.c_size_t()
        assert L.asdFftGetWorkspaceSize(h, ctypes.byref(size)) == 0 and size.value > SCRATCH
        x = N.AclTensor((256, 65536), N.ACL_COMPLEX64, 0x10000000)
        y = N.AclTensor((256, 65536), N.ACL_COMPLEX64, 0x90000000)
        assert L.asdFftExtExecNoAlloc(h, x.ptr, y.ptr) == N.ACL_ERROR_INVALID_PARAM
        x.close()
        y.close()
    finally:
        L.asdFftDestroy(h)
